// Internal helpers shared by the kernel translation units of libtrove_amd.
#pragma once
#include <hip/hip_runtime.h>
#include <trove/aos.h>
#include <trove/block.h>
#include <trove/lds.h>
#include <trove/ptr.h>
#include <trove_amd.h>

#include <cstdint>
#include <type_traits>

namespace trove_amd {

// Work-group size of every kernel: four wavefronts.
constexpr int kBlock = 256;
// Largest struct of the LDS-staged variant: its stage is 64 * sizeof(T)
// bytes per wavefront (32 KiB at 512 bytes, 5 wavefronts per CU), and above
// 512 bytes the struct no longer fits the registers beside its chunks.
constexpr int kMaxLdsStructBytes = 512;

// TROVE_VARIANT_AUTO per struct size, from interleaved A/B timings on MI355X
// (median of 9-15 launches, 2^26 structs up to 64 bytes, 2^24-2^25 up to
// 256, 2^22-2^23 above) confirmed by a second, independent bench run: a path
// other than the default only where it was faster in both. Copies: the
// LDS-staged path for structs of at most 16 bytes (a full dwordx4 per lane
// instead of narrower per-lane words) and the 16-byte-word structs of 32 and
// 96-160 bytes (no two-rotation schedule); gathers: LDS-staged at 16, 48,
// 64, 192 and 384 bytes (+6-9% over the tile path at 384), coalesced_ptr at
// 224 bytes (+6-7%); scatters: LDS-staged at 8 and 16 bytes, coalesced_ptr
// at 64 bytes (+4%: 3.69 against 3.54 TB/s, interleaved). 4-byte scatters
// keep the naive kernel (0.93 TB/s; the streaming one-word tile scatter
// reaches 0.84). Everything else takes the transposition (copy) or the tile
// path (gather, scatter).
constexpr int auto_copy_variant(int b) {
    return (b == 4 || b == 8 || b == 12 || b == 16 || b == 32 || b == 96 || b == 128 || b == 160)
               ? TROVE_VARIANT_LDS
               : TROVE_VARIANT_SHFL;
}
constexpr int auto_gather_variant(int b) {
    return (b == 16 || b == 48 || b == 64 || b == 192 || b == 384) ? TROVE_VARIANT_LDS
           : b == 224                                                ? TROVE_VARIANT_SHFL
                                                                     : TROVE_VARIANT_TILE;
}
constexpr int auto_scatter_variant(int b) {
    return b == 4 ? TROVE_VARIANT_DIRECT
           : (b == 8 || b == 16) ? TROVE_VARIANT_LDS
           : b == 64 ? TROVE_VARIANT_SHFL
                     : TROVE_VARIANT_TILE;
}
// trove_scatter_permutation: struct sizes where inverting the permutation
// (one random 4-byte write per row) and gathering through the inverse beat
// scattering the rows. Random row writes that cover part of a 64-byte
// memory granule run far below whole-granule ones on MI355X (32-byte rows
// 0.73 TB/s, 60-byte rows 0.87 TB/s, 64-byte rows 2.86 TB/s of written
// bytes), while random row reads cost whole 128-byte line fills at ~5.4
// TB/s; the sweep at 2^26 rows predicts 1.14-1.53x from 12 to 60 bytes and a
// loss at 4, 8 and 64 bytes.
constexpr bool auto_scatter_by_inverse(int b) { return b % 4 == 0 && b >= 12 && b <= 60; }
constexpr int kWavesPerBlock = kBlock / 64;

// Cache policy P for a per-lane access (each lane moving its own struct word
// by word): a struct of several words per lane touches each line from
// several instructions, which only the cache turns into whole-line traffic
// (nontemporal per-lane words re-fetch or partially write the line for each
// word: 1.6 TB/s against 4.5 for ordinary accesses at 60-byte copies on
// MI355X), so P applies to one-word structs only.
template <typename T, typename P>
using per_lane_policy =
    typename std::conditional<trove::detail::word_layout<T>::m == 1, P, trove::cached>::type;

// The struct type moved for a given byte size: k 32-bit words when the size
// is a multiple of 4 (the reference benchmark's trove::array<int, k>),
// otherwise raw bytes.
template <int B>
using blob_t = typename std::conditional<B % 4 == 0, trove::array<uint32_t, (B % 4 == 0 ? B / 4 : 1)>,
                                         trove::array<uint8_t, B> >::type;

// Struct sizes compiled into the C ABI: the 4..64-byte sweep, larger sizes
// that reach the composite schedule (96, 160, 192, 224, 384, 768 B: 6, 10,
// 12, 14, 24, 48 16-byte words), power-of-two sizes up to the whole-wavefront
// limit m = 64 (1 KiB) and two sizes that are not a multiple of 4. They are
// split into shards so that the kernel translation units can be compiled in
// parallel (one shard per size above 256 B, whose kernels hold up to 256
// words per lane and compile slowest): a unit built with -DTROVE_SHARD=k
// instantiates shard k only (without it, all shards).
#define TROVE_SIZES_SHARD0(X) X(4) X(8) X(12) X(16) X(20) X(224)
#define TROVE_SIZES_SHARD1(X) X(24) X(28) X(32) X(36) X(256)
#define TROVE_SIZES_SHARD2(X) X(40) X(44) X(48) X(2) X(192)
#define TROVE_SIZES_SHARD3(X) X(52) X(56) X(6) X(160)
#define TROVE_SIZES_SHARD4(X) X(60) X(64) X(128)
#define TROVE_SIZES_SHARD5(X) X(96)
#define TROVE_SIZES_SHARD6(X) X(384)
#define TROVE_SIZES_SHARD7(X) X(512)
#define TROVE_SIZES_SHARD8(X) X(768)
#define TROVE_SIZES_SHARD9(X) X(1024)
#define TROVE_SIZE_SHARDS 10
#define TROVE_FOR_EACH_SIZE(X)                                                                                 \
    TROVE_SIZES_SHARD0(X) TROVE_SIZES_SHARD1(X) TROVE_SIZES_SHARD2(X) TROVE_SIZES_SHARD3(X) TROVE_SIZES_SHARD4(X) \
        TROVE_SIZES_SHARD5(X) TROVE_SIZES_SHARD6(X) TROVE_SIZES_SHARD7(X) TROVE_SIZES_SHARD8(X) TROVE_SIZES_SHARD9(X)

// The store / load-sum benchmark kernels of the reference harness are
// compiled up to this struct size.
constexpr int kMaxBenchStructBytes = 256;

#if !defined(TROVE_SHARD)
#define TROVE_FOR_EACH_SIZE_IN_SHARD(X) TROVE_FOR_EACH_SIZE(X)
#elif TROVE_SHARD == 0
#define TROVE_FOR_EACH_SIZE_IN_SHARD(X) TROVE_SIZES_SHARD0(X)
#elif TROVE_SHARD == 1
#define TROVE_FOR_EACH_SIZE_IN_SHARD(X) TROVE_SIZES_SHARD1(X)
#elif TROVE_SHARD == 2
#define TROVE_FOR_EACH_SIZE_IN_SHARD(X) TROVE_SIZES_SHARD2(X)
#elif TROVE_SHARD == 3
#define TROVE_FOR_EACH_SIZE_IN_SHARD(X) TROVE_SIZES_SHARD3(X)
#elif TROVE_SHARD == 4
#define TROVE_FOR_EACH_SIZE_IN_SHARD(X) TROVE_SIZES_SHARD4(X)
#elif TROVE_SHARD == 5
#define TROVE_FOR_EACH_SIZE_IN_SHARD(X) TROVE_SIZES_SHARD5(X)
#elif TROVE_SHARD == 6
#define TROVE_FOR_EACH_SIZE_IN_SHARD(X) TROVE_SIZES_SHARD6(X)
#elif TROVE_SHARD == 7
#define TROVE_FOR_EACH_SIZE_IN_SHARD(X) TROVE_SIZES_SHARD7(X)
#elif TROVE_SHARD == 8
#define TROVE_FOR_EACH_SIZE_IN_SHARD(X) TROVE_SIZES_SHARD8(X)
#elif TROVE_SHARD == 9
#define TROVE_FOR_EACH_SIZE_IN_SHARD(X) TROVE_SIZES_SHARD9(X)
#endif

// Global wavefront index and count for B-thread work-groups (scalar: the
// same on every lane).
template <int B = kBlock>
__device__ __forceinline__ int64_t wave_index() {
    return (int64_t)blockIdx.x * (B / 64) + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}
template <int B = kBlock>
__device__ __forceinline__ int64_t wave_count() {
    return (int64_t)gridDim.x * (B / 64);
}

// XCD-contiguous work mapping. MI355X dispatches consecutive work-groups to
// its 8 XCDs round-robin, so in dispatch order each XCD's L2 sees every 8th
// unit. unit_for_xcd maps work-group b of a one-shot grid of `total` units to
// a unit such that XCD x (= b % 8) walks a contiguous range of about
// total / 8 units. It is a bijection on [0, total) for any total.
constexpr int kXcds = 8;
__device__ __forceinline__ int64_t unit_for_xcd(int64_t b, int64_t total) {
    const int64_t x = b % kXcds, j = b / kXcds;
    const int64_t base = total / kXcds, extra = total % kXcds;
    return x * base + (x < extra ? x : extra) + j;
}

__host__ __device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Host side: launch bookkeeping and error reporting.
int hip_status(hipError_t e);         // maps to TROVE_OK / TROVE_ERR_HIP and records e
int launch_status();                  // checks hipGetLastError after a launch
int device_cu_count();                // compute units of the current device
// Work-groups of `waves_per_block` wavefronts for `waves` wavefront work
// units: one unit per wavefront (a one-shot grid; on MI355X this streams
// 15-25% faster than a resident-sized grid-stride loop). Kernels still
// stride for the (unreachable in practice) case of more than 2^31 groups.
int64_t grid_for(int64_t waves, int waves_per_block = kWavesPerBlock);

// Dynamic LDS to reserve per work-group so that at most about
// kStreamWindowBytes / bytes_per_block bytes of streaming work sit on one CU
// at a time (work-groups per CU = LDS per CU / reservation). Measured on
// MI355X: a resident window of ~32-40 KiB per CU streams best (4-KiB-per-
// wavefront copies: 10 wavefronts per CU reach 6.5-6.7 TB/s against
// 5.9-6.1 TB/s at full occupancy; 1-KiB wavefronts want no cap). Returns 0
// when no cap is needed.
constexpr int64_t kStreamWindowBytes = 40 * 1024;
size_t stream_lds_reservation(int64_t bytes_per_block);

// Dynamic LDS to reserve per work-group so that exactly blocks_per_cu
// work-groups of a kernel without static LDS fit on one CU (0 for
// blocks_per_cu <= 0: no cap). gfx950 allocates LDS in 1280-byte units, 128
// per CU: counted on MI355X with a per-CU residency counter, a reservation of
// r bytes admits floor(128 / ceil(r / 1280)) work-groups (16384 bytes admit
// 9, not 10; 13312 admit 11), which the occupancy calculator does not model.
constexpr int kLdsAllocBytes = 1280;
size_t cap_lds_reservation(int blocks_per_cu);

inline bool aligned(const void* p, int bytes) { return ((uintptr_t)p % (uintptr_t)bytes) == 0; }

template <int B>
constexpr int word_bytes() {
    return B % 16 == 0 ? 16 : B % 8 == 0 ? 8 : B % 4 == 0 ? 4 : 1;
}

inline int word_bytes_of(int b) { return b % 16 == 0 ? 16 : b % 8 == 0 ? 8 : b % 4 == 0 ? 4 : 1; }

// Per-size entry points, defined in the kernel translation units.
template <int B>
int copy_impl(const void* src, void* dst, int64_t n, int variant, int tile, hipStream_t s);
template <int B>
int gather_impl(const void* src, const int32_t* idx, void* dst, int64_t n, int variant, int tile, hipStream_t s);
template <int B>
int scatter_impl(const void* src, const int32_t* idx, void* dst, int64_t n, int variant, int tile, hipStream_t s);
template <int B>
int counting_impl(void* dst, int64_t n, int variant, int tile, hipStream_t s);
template <int B>
int loadsum_impl(const void* src, uint32_t* out, int64_t n, int variant, int tile, hipStream_t s);

}  // namespace trove_amd
