// Wavefront primitives for CDNA (wave64): lane index, tiles, convergence.
//
// Public interface of trove/warp.h:34-93 (WARP_SIZE, thread_tile<N> with
// shfl/size/log_size/mask/id, warp_converged<N>), built on the hardware:
//  * the lane index comes from mbcnt (the wavefront's own lane numbering);
//  * thread_tile<N>::shfl is one ds_bpermute_b32 per dword, addressing the
//    source lane inside the caller's aligned N-lane group;
//  * warp_converged<N> reads the 64-bit EXEC mask through a ballot and tests
//    the caller's aligned N-lane group.
// WARP_SIZE is 64: a default tile is a whole wavefront.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace trove {

enum { WARP_SIZE = 64 };

__host__ __device__ constexpr std::size_t log2(std::size_t n) { return n < 2 ? 0 : 1 + log2(n / 2); }

// Lane index of the calling work-item inside its wavefront, 0..63.
__device__ __forceinline__ int lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Mask of the lanes executing this instruction (EXEC).
__device__ __forceinline__ uint64_t active_lanes() { return __builtin_amdgcn_ballot_w64(true); }

namespace detail {

// Cross-lane pull: every lane receives `v` from the lane whose byte address
// (4 * lane) it passes. Values wider than a dword move as independent dwords.
template <typename T>
__device__ __forceinline__ T bpermute(const T& v, int byte_addr) {
    constexpr int w = (sizeof(T) + 3) / 4;
    uint32_t a[w] = {};
    __builtin_memcpy(a, &v, sizeof(T));
#pragma unroll
    for (int i = 0; i < w; ++i) a[i] = (uint32_t)__builtin_amdgcn_ds_bpermute(byte_addr, (int)a[i]);
    T r;
    __builtin_memcpy(&r, a, sizeof(T));
    return r;
}

// Cross-lane push: every lane sends `v` to the lane whose byte address it
// passes. The destinations of the participating lanes must be distinct.
__device__ __forceinline__ uint32_t permute_push(uint32_t v, int byte_addr) {
    return (uint32_t)__builtin_amdgcn_ds_permute(byte_addr, (int)v);
}

// Broadcast of lane 0's value to the whole wavefront (scalar register).
template <typename T>
__device__ __forceinline__ T broadcast_first(const T& v) {
    constexpr int w = (sizeof(T) + 3) / 4;
    uint32_t a[w] = {};
    __builtin_memcpy(a, &v, sizeof(T));
#pragma unroll
    for (int i = 0; i < w; ++i) a[i] = (uint32_t)__builtin_amdgcn_readfirstlane((int)a[i]);
    T r;
    __builtin_memcpy(&r, a, sizeof(T));
    return r;
}

// True when `pred` holds on every lane of the caller's aligned N-lane group.
// Lanes of the group must all be active.
template <int N>
__device__ __forceinline__ bool tile_all(bool pred) {
    const uint64_t b = __builtin_amdgcn_ballot_w64(pred);
    if constexpr (N == 64) {
        return b == ~0ull;
    } else {
        const int shift = lane_id() & ~(N - 1);
        const uint64_t grp = ((1ull << N) - 1) << shift;
        return (b & grp) == grp;
    }
}

}  // namespace detail

// An aligned group of `threads` consecutive lanes of one wavefront.
template <std::size_t threads>
struct thread_tile {
    static_assert(threads >= 1 && threads <= 64 && (threads & (threads - 1)) == 0,
                  "a tile is a power-of-two group of at most 64 lanes");
    __device__ thread_tile() {}

    template <typename T>
    __device__ T shfl(const T& t, const int& i) const {
        return detail::bpermute(t, (first_lane() + (i & (int)mask())) << 2);
    }

    __host__ __device__ static constexpr int size() { return (int)threads; }
    __host__ __device__ static constexpr int log_size() { return (int)trove::log2(threads); }
    __host__ __device__ static constexpr int mask() { return (int)threads - 1; }
    // Rank of the calling lane inside its tile.
    __device__ int id() const { return lane_id() & mask(); }
    // Wavefront lane index of the tile's rank-0 lane.
    __device__ int first_lane() const { return lane_id() & ~mask(); }
};

// True when every lane of the caller's aligned warp_size-lane group is active.
template <std::size_t warp_size = WARP_SIZE>
__device__ inline bool warp_converged() {
    static_assert(warp_size >= 1 && warp_size <= 64 && (warp_size & (warp_size - 1)) == 0,
                  "warp_converged takes a power-of-two group of at most 64 lanes");
    const uint64_t exec = active_lanes();
    if constexpr (warp_size == 64) {
        return exec == ~0ull;
    } else {
        const int shift = lane_id() & ~(int)(warp_size - 1);
        const uint64_t grp = ((1ull << warp_size) - 1) << shift;
        return (exec & grp) == grp;
    }
}

}  // namespace trove
