// Indexed AoS movement: gather dst[i] = src[idx[i]], scatter dst[idx[i]] = src[i].
//  SHFL   - the reference benchmark's coalesced_ptr kernels
//           (tests/benchmark.cu:90-110): per-lane index read, struct read and
//           struct write through coalesced_ptr (tile cascade, address
//           broadcast, R2C / C2R, contiguity fast path on the contiguous side);
//           the scatter assigns one coalesced_ref to another (ptr.h:47-72).
//  TILE   - the same tile-indexed word accesses with the transposition
//           elided: the gathered words already sit in the column-side layout
//           a coalesced tile store takes (and scatter mirrors it).
//  DIRECT - the naive per-lane kernels (tests/benchmark.cu:112-128).
#include "gather_body.h"

namespace trove_amd {

// Each wavefront handles UNROLL 64-struct groups (one-shot grid; a capped
// grid strides).
template <typename T, int UNROLL>
__global__ __launch_bounds__(kBlock) void k_gather_tile(const T* __restrict__ src, const int32_t* __restrict__ idx,
                                                        T* __restrict__ dst, int64_t n) {
    const int64_t groups = (n + 63) >> 6;
    for (int64_t g0 = wave_index() * UNROLL; g0 < groups; g0 += wave_count() * UNROLL)
        gather_range_tile<T, UNROLL>(src, idx, dst, n, g0, g0 + UNROLL < groups ? g0 + UNROLL : groups);
}

// One 64-struct group per wavefront and no loop: inside a loop, LICM would
// hoist the per-lane schedule tables of all six cascade levels out of it and
// keep them live together (258 VGPRs at 60-byte structs). The trailing
// partial group runs the same code with its missing lanes inactive, so the
// coalesced_ptr cascade picks the widest converged sub-tile for it.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gather_shfl(const T* __restrict__ src, const int32_t* __restrict__ idx,
                                                        T* __restrict__ dst, int64_t n) {
    const int64_t g = wave_index();
    if (g < ((n + 63) >> 6)) gather_group_shfl<T>(src, idx, dst, n, g);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_gather_direct(const T* __restrict__ src, const int32_t* __restrict__ idx,
                                                          T* __restrict__ dst, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) dst[i] = src[idx[i]];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_scatter_shfl(const T* __restrict__ src, const int32_t* __restrict__ idx,
                                                         T* __restrict__ dst, int64_t n) {
    const int64_t i = (wave_index() << 6) + trove::lane_id();
    if (i < n) {
        trove::coalesced_ptr<T> s(const_cast<T*>(src));
        trove::coalesced_ptr<T> r(dst);
        r[idx[i]] = s[i];  // coalesced_ref = coalesced_ref
    }
}

template <typename T, int UNROLL>
__global__ __launch_bounds__(kBlock) void k_scatter_tile(const T* __restrict__ src, const int32_t* __restrict__ idx,
                                                         T* __restrict__ dst, int64_t n) {
    typedef trove::detail::word_layout<T> L;
    const int lane = trove::lane_id();
    const int64_t groups = n >> 6;
    const int64_t step = wave_count() * UNROLL;
    int64_t g = wave_index() * UNROLL;
    for (; g + UNROLL <= groups; g += step) {
        int32_t ix[UNROLL];
        uint32_t v[UNROLL][L::m * L::W];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            ix[u] = trove::detail::load_word<trove::streaming>(idx + ((g + u) << 6) + lane);
            trove::detail::tile_load_words<64, T, trove::streaming>(src + ((g + u) << 6), lane, v[u]);
        }
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) trove::detail::tile_scatter_indexed<64>(dst, ix[u], lane, 0, v[u]);
    }
    for (; g < groups; ++g) {
        uint32_t v[L::m * L::W];
        trove::detail::tile_load_words<64>(src + (g << 6), lane, v);
        trove::detail::tile_scatter_indexed<64>(dst, idx[(g << 6) + lane], lane, 0, v);
    }
    if (wave_index() == 0 && (n & 63)) {
        const int64_t i = (groups << 6) + lane;
        if (i < n) trove::detail::divergent_store(trove::detail::divergent_load(src + i), dst + idx[i]);
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_scatter_direct(const T* __restrict__ src, const int32_t* __restrict__ idx,
                                                           T* __restrict__ dst, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) dst[idx[i]] = src[i];
}

template <int B>
int gather_impl(const void* src_, const int32_t* idx, void* dst_, int64_t n, int variant, hipStream_t s) {
    typedef blob_t<B> T;
    const T* src = static_cast<const T*>(src_);
    T* dst = static_cast<T*>(dst_);
    constexpr bool tiled = trove::detail::use_shfl<T, 64>::value;
    constexpr int U = gather_groups_in_flight<T>();
    const int64_t units = ((n + 63) / 64 + U - 1) / U;
    if (variant == TROVE_VARIANT_AUTO) variant = tiled ? TROVE_VARIANT_TILE : TROVE_VARIANT_DIRECT;
    if (variant == TROVE_VARIANT_TILE && !tiled) variant = TROVE_VARIANT_DIRECT;
    if (variant == TROVE_VARIANT_SHFL) {
        hipLaunchKernelGGL((k_gather_shfl<T>), dim3((unsigned)grid_for((n + 63) / 64)), dim3(kBlock), 0, s, src,
                           idx, dst, n);
        return launch_status();
    }
    if (variant == TROVE_VARIANT_TILE) {
        if constexpr (tiled) {
            hipLaunchKernelGGL((k_gather_tile<T, U>), dim3((unsigned)grid_for(units > 0 ? units : 1)), dim3(kBlock),
                               0, s, src, idx, dst, n);
            return launch_status();
        }
    }
    if (variant == TROVE_VARIANT_DIRECT) {
        hipLaunchKernelGGL((k_gather_direct<T>), dim3((unsigned)grid_for((n + 63) / 64)), dim3(kBlock), 0, s, src,
                           idx, dst, n);
        return launch_status();
    }
    return TROVE_ERR_VARIANT;
}

template <int B>
int scatter_impl(const void* src_, const int32_t* idx, void* dst_, int64_t n, int variant, hipStream_t s) {
    typedef blob_t<B> T;
    const T* src = static_cast<const T*>(src_);
    T* dst = static_cast<T*>(dst_);
    constexpr bool tiled = trove::detail::use_shfl<T, 64>::value;
    constexpr int U = gather_groups_in_flight<T>();
    const int64_t units = ((n >> 6) + U - 1) / U;
    if (variant == TROVE_VARIANT_AUTO) variant = tiled ? TROVE_VARIANT_TILE : TROVE_VARIANT_DIRECT;
    if (variant == TROVE_VARIANT_TILE && !tiled) variant = TROVE_VARIANT_DIRECT;
    if (variant == TROVE_VARIANT_SHFL) {
        hipLaunchKernelGGL((k_scatter_shfl<T>), dim3((unsigned)grid_for((n + 63) / 64)), dim3(kBlock), 0, s, src,
                           idx, dst, n);
        return launch_status();
    }
    if (variant == TROVE_VARIANT_TILE) {
        if constexpr (tiled) {
            hipLaunchKernelGGL((k_scatter_tile<T, U>), dim3((unsigned)grid_for(units > 0 ? units : 1)),
                               dim3(kBlock), 0, s, src, idx, dst, n);
            return launch_status();
        }
    }
    if (variant == TROVE_VARIANT_DIRECT) {
        hipLaunchKernelGGL((k_scatter_direct<T>), dim3((unsigned)grid_for((n + 63) / 64)), dim3(kBlock), 0, s, src,
                           idx, dst, n);
        return launch_status();
    }
    return TROVE_ERR_VARIANT;
}

#define TROVE_INST_GATHER(B)                                                                          \
    template int gather_impl<B>(const void*, const int32_t*, void*, int64_t, int, hipStream_t);       \
    template int scatter_impl<B>(const void*, const int32_t*, void*, int64_t, int, hipStream_t);
TROVE_FOR_EACH_SIZE_IN_SHARD(TROVE_INST_GATHER)

}  // namespace trove_amd
