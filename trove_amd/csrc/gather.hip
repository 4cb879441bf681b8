// Indexed AoS movement: gather dst[i] = src[idx[i]], scatter dst[idx[i]] = src[i].
//  SHFL   - the reference benchmark's coalesced_ptr kernels
//           (tests/benchmark.cu:90-110): per-lane index read, struct read and
//           struct write through coalesced_ptr (tile cascade, address
//           broadcast, R2C / C2R, contiguity fast path on the contiguous side);
//           the scatter assigns one coalesced_ref to another (ptr.h:47-72).
//  TILE   - the same tile-indexed word accesses with the transposition
//           elided: the gathered words already sit in the column-side layout
//           a coalesced tile store takes (and scatter mirrors it).
//  LDS    - the coalesced_ptr kernels' LDS-staged counterpart: tile-indexed
//           word accesses, one struct per lane through an LDS stage, chunked
//           (dwordx4) accesses on the contiguous side (trove/lds.h).
//  DIRECT - the naive per-lane kernels (tests/benchmark.cu:112-128).
#include "gather_body.h"

namespace trove_amd {

// Launch shape of the standalone tile gather per struct size: 64-struct
// groups in flight per wavefront and work-group size. From interleaved
// timings of random-permutation gathers on MI355X (2^26 rows up to 64 bytes,
// 2^24 above), consistent over two runs: four groups at 60 and 128 bytes
// (+3% and +4.5% over two), four groups in 64-thread work-groups at 64
// bytes (+2.8%), 64-thread work-groups at 96 bytes (+1.5%); elsewhere
// gather_groups_in_flight in 256-thread work-groups. (The batch
// kernel keeps gather_groups_in_flight: its registers are the maximum over
// every size's branch.)
template <typename T>
struct gather_tile_shape {
    static constexpr int unroll =
        (sizeof(T) == 60 || sizeof(T) == 64 || sizeof(T) == 128) ? 4 : gather_groups_in_flight<T>();
    static constexpr int block = (sizeof(T) == 64 || sizeof(T) == 96) ? 64 : kBlock;
};

// Each wavefront handles UNROLL 64-struct groups (one-shot grid; a capped
// grid strides).
template <typename T, int UNROLL, int B>
__global__ __launch_bounds__(B) void k_gather_tile(const T* __restrict__ src, const int32_t* __restrict__ idx,
                                                   T* __restrict__ dst, int64_t n) {
    const int64_t groups = (n + 63) >> 6;
    for (int64_t g0 = wave_index<B>() * UNROLL; g0 < groups; g0 += wave_count<B>() * UNROLL)
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

// coalesced_ptr gather / scatter at a fixed tile width below 64 (the C ABI's
// `tile` argument): one group per wavefront, each TILE-lane tile on its own.
template <typename T, int TILE>
__global__ __launch_bounds__(kBlock) void k_gather_shfl_tile(const T* __restrict__ src,
                                                             const int32_t* __restrict__ idx, T* __restrict__ dst,
                                                             int64_t n) {
    const int64_t g = wave_index();
    if (g < ((n + 63) >> 6)) gather_group_shfl_tile<T, TILE>(src, idx, dst, n, g);
}

template <typename T, int TILE>
__global__ __launch_bounds__(kBlock) void k_scatter_shfl_tile(const T* __restrict__ src,
                                                              const int32_t* __restrict__ idx, T* __restrict__ dst,
                                                              int64_t n) {
    const int64_t i = (wave_index() << 6) + trove::lane_id();
    if (i < n) move_at_tile<TILE, trove::cached, trove::cached>(src + i, dst + idx[i]);
}

// Launches the fixed-tile gather (SCATTER false) or scatter (true).
template <typename T, bool SCATTER>
static int launch_shfl_tile(const T* src, const int32_t* idx, T* dst, int64_t n, int tile, hipStream_t s) {
    const unsigned blocks = (unsigned)grid_for((n + 63) / 64);
#define TROVE_TILE_CASE(W)                                                                                 \
    case W:                                                                                                \
        if constexpr (SCATTER) {                                                                           \
            hipLaunchKernelGGL((k_scatter_shfl_tile<T, W>), dim3(blocks), dim3(kBlock), 0, s, src, idx, dst, n); \
        } else {                                                                                           \
            hipLaunchKernelGGL((k_gather_shfl_tile<T, W>), dim3(blocks), dim3(kBlock), 0, s, src, idx, dst, n);  \
        }                                                                                                  \
        break;
    switch (tile) {
        TROVE_TILE_CASE(32)
        TROVE_TILE_CASE(16)
        TROVE_TILE_CASE(8)
        TROVE_TILE_CASE(4)
        TROVE_TILE_CASE(2)
        default: return TROVE_ERR_VARIANT;
    }
#undef TROVE_TILE_CASE
    return launch_status();
}

// LDS-staged gather: one group per one-wavefront work-group (64 * sizeof(T)
// bytes of dynamic LDS, up to 64 KiB at 1 KiB structs); the trailing partial
// group per lane.
template <typename T, int CW>
__global__ __launch_bounds__(64) void k_gather_lds(const T* __restrict__ src, const int32_t* __restrict__ idx,
                                                   T* __restrict__ dst, int64_t n) {
    extern __shared__ __align__(16) uint32_t lds_stage[];
    const int64_t g = blockIdx.x;
    if (g < (n >> 6)) {
        gather_group_lds<T, CW>(src, idx, dst, g, lds_stage);
    } else {
        const int64_t i = (g << 6) + trove::lane_id();
        if (i < n) trove::detail::divergent_store(trove::detail::divergent_load(src + idx[i]), dst + i);
    }
}

// LDS-staged scatter: CW-dword chunk loads of the group through the stage to
// one struct per lane (load_warp_contiguous_lds), then back through the stage
// to column-side words for the tile-indexed stores (m = 1: per-lane stores).
template <typename T, int CW>
__global__ __launch_bounds__(64) void k_scatter_lds(const T* __restrict__ src, const int32_t* __restrict__ idx,
                                                    T* __restrict__ dst, int64_t n) {
    extern __shared__ __align__(16) uint32_t lds_stage[];
    namespace d = trove::detail;
    typedef d::word_layout<T> L;
    const int64_t g = blockIdx.x;
    const int lane = trove::lane_id();
    const int64_t i = (g << 6) + lane;
    if (g >= (n >> 6)) {
        if (i < n) d::divergent_store(d::divergent_load(src + i), dst + idx[i]);
        return;
    }
    const int32_t ix = d::load_word<trove::streaming>(idx + i);
    d::lds_chunks<T, CW> t;
    d::lds_global_load<CW, trove::streaming>(src + (g << 6), lane, t);
    d::lds_put_chunks(lds_stage, lane, t);
    d::wave_lds_sync();
    const T v = d::lds_get_struct<T>(lds_stage, lane);
    if constexpr (L::m == 1) {
        d::divergent_store(v, dst + ix);
    } else {
        d::wave_lds_sync();
        d::lds_put_struct(lds_stage, lane, v);
        d::wave_lds_sync();
        uint32_t w[L::m * L::W];
        d::lds_get_columns<T>(lds_stage, lane, w);
        d::tile_scatter_indexed<64>(dst, ix, lane, 0, w);
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_gather_direct(const T* __restrict__ src, const int32_t* __restrict__ idx,
                                                          T* __restrict__ dst, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) dst[i] = src[idx[i]];
}

// The naive gather and scatter in the tile path's launch (same grid, work-
// group size and groups in flight; the tile path's cache policies where a
// per-lane access allows them), each lane moving its own struct word by
// word: what the tile path adds over a per-lane access on equal launch terms.
template <typename T, int UNROLL, int B>
__global__ __launch_bounds__(B) void k_gather_direct_launch(const T* __restrict__ src, const int32_t* __restrict__ idx,
                                                            T* __restrict__ dst, int64_t n) {
    namespace d = trove::detail;
    typedef trove::streaming S;
    typedef per_lane_policy<T, row_read_policy<T>> RP;
    typedef per_lane_policy<T, S> WP;
    const int lane = trove::lane_id();
    const int64_t groups = (n + 63) >> 6;
    for (int64_t g0 = wave_index<B>() * UNROLL; g0 < groups; g0 += wave_count<B>() * UNROLL) {
        if (((g0 + UNROLL) << 6) <= n) {
            int32_t ix[UNROLL];
            T v[UNROLL];
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) ix[u] = d::load_word<S>(idx + ((g0 + u) << 6) + lane);
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) v[u] = d::divergent_load<RP>(src + ix[u]);
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) d::divergent_store<WP>(v[u], dst + ((g0 + u) << 6) + lane);
        } else {
            for (int64_t i = (g0 << 6) + lane; i < n; i += 64)
                d::divergent_store<WP>(d::divergent_load<RP>(src + idx[i]), dst + i);
        }
    }
}

template <typename T, int UNROLL>
__global__ __launch_bounds__(kBlock) void k_scatter_direct_launch(const T* __restrict__ src,
                                                                  const int32_t* __restrict__ idx,
                                                                  T* __restrict__ dst, int64_t n) {
    namespace d = trove::detail;
    typedef trove::streaming S;
    const int lane = trove::lane_id();
    const int64_t groups = (n + 63) >> 6;
    for (int64_t g0 = wave_index() * UNROLL; g0 < groups; g0 += wave_count() * UNROLL) {
        if (((g0 + UNROLL) << 6) <= n) {
            int32_t ix[UNROLL];
            T v[UNROLL];
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) {
                ix[u] = d::load_word<S>(idx + ((g0 + u) << 6) + lane);
                v[u] = d::divergent_load<per_lane_policy<T, S>>(src + ((g0 + u) << 6) + lane);
            }
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) d::divergent_store(v[u], dst + ix[u]);
        } else {
            for (int64_t i = (g0 << 6) + lane; i < n; i += 64) d::divergent_store(d::divergent_load(src + i), dst + idx[i]);
        }
    }
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
    if constexpr (L::m == 1) {
        // One-word structs: streaming contiguous loads, per-lane row stores.
        typedef typename L::U U;
        const U* s = reinterpret_cast<const U*>(src);
        U* d = reinterpret_cast<U*>(dst);
        for (; g + UNROLL <= groups; g += step) {
            int32_t ix[UNROLL];
            U v[UNROLL];
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) {
                ix[u] = trove::detail::load_word<trove::streaming>(idx + ((g + u) << 6) + lane);
                v[u] = trove::detail::load_word<trove::streaming>(s + ((g + u) << 6) + lane);
            }
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) trove::detail::store_word<trove::cached>(d + ix[u], v[u]);
        }
        for (; g < groups; ++g) d[idx[(g << 6) + lane]] = s[(g << 6) + lane];
        if (wave_index() == 0 && (n & 63)) {
            const int64_t i = (groups << 6) + lane;
            if (i < n) d[idx[i]] = s[i];
        }
        return;
    }
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

// Structs the TILE variant covers: the shuffle-eligible sizes and the
// one-word structs of 4, 8 and 16 bytes (per-lane row words).
template <typename T>
constexpr bool tile_path() {
    return trove::detail::use_shfl<T, 64>::value ||
           (sizeof(T) % 4 == 0 && trove::detail::word_layout<T>::m == 1);
}

template <int B>
int gather_impl(const void* src_, const int32_t* idx, void* dst_, int64_t n, int variant, int tile, hipStream_t s) {
    typedef blob_t<B> T;
    const T* src = static_cast<const T*>(src_);
    T* dst = static_cast<T*>(dst_);
    constexpr bool tiled = tile_path<T>();
    typedef gather_tile_shape<T> G;
    const int64_t units = ((n + 63) / 64 + G::unroll - 1) / G::unroll;
    if (tile != 0 && tile != 64) {  // a sub-wavefront tile asks for coalesced_ptr at that width
        if (variant == TROVE_VARIANT_AUTO || variant == TROVE_VARIANT_SHFL)
            return launch_shfl_tile<T, false>(src, idx, dst, n, tile, s);
        if (variant != TROVE_VARIANT_DIRECT && variant != TROVE_VARIANT_DIRECT_LAUNCH) return TROVE_ERR_VARIANT;
    }
    if (variant == TROVE_VARIANT_AUTO) variant = auto_gather_variant(B);
    if (variant == TROVE_VARIANT_TILE && !tiled) variant = TROVE_VARIANT_DIRECT;
    if (variant == TROVE_VARIANT_LDS) {
        if constexpr (B % 4 == 0 && B <= kMaxLdsStructBytes) {
            const unsigned blocks = (unsigned)grid_for((n + 63) / 64, 1);
            const size_t lds = trove::lds_stage_bytes<T>();
            if (aligned(dst, 16)) {
                hipLaunchKernelGGL((k_gather_lds<T, 4>), dim3(blocks), dim3(64), lds, s, src, idx, dst, n);
            } else {
                hipLaunchKernelGGL((k_gather_lds<T, 1>), dim3(blocks), dim3(64), lds, s, src, idx, dst, n);
            }
            return launch_status();
        } else {
            return TROVE_ERR_VARIANT;
        }
    }
    if (variant == TROVE_VARIANT_SHFL) {
        hipLaunchKernelGGL((k_gather_shfl<T>), dim3((unsigned)grid_for((n + 63) / 64)), dim3(kBlock), 0, s, src,
                           idx, dst, n);
        return launch_status();
    }
    if (variant == TROVE_VARIANT_TILE) {
        if constexpr (tiled) {
            hipLaunchKernelGGL((k_gather_tile<T, G::unroll, G::block>),
                               dim3((unsigned)grid_for(units > 0 ? units : 1, G::block / 64)), dim3(G::block), 0, s,
                               src, idx, dst, n);
            return launch_status();
        }
    }
    if (variant == TROVE_VARIANT_DIRECT) {
        hipLaunchKernelGGL((k_gather_direct<T>), dim3((unsigned)grid_for((n + 63) / 64)), dim3(kBlock), 0, s, src,
                           idx, dst, n);
        return launch_status();
    }
    if (variant == TROVE_VARIANT_DIRECT_LAUNCH) {
        hipLaunchKernelGGL((k_gather_direct_launch<T, G::unroll, G::block>),
                           dim3((unsigned)grid_for(units > 0 ? units : 1, G::block / 64)), dim3(G::block), 0, s, src,
                           idx, dst, n);
        return launch_status();
    }
    return TROVE_ERR_VARIANT;
}

template <int B>
int scatter_impl(const void* src_, const int32_t* idx, void* dst_, int64_t n, int variant, int tile, hipStream_t s) {
    typedef blob_t<B> T;
    const T* src = static_cast<const T*>(src_);
    T* dst = static_cast<T*>(dst_);
    constexpr bool tiled = tile_path<T>();
    if (tile != 0 && tile != 64) {  // a sub-wavefront tile asks for coalesced_ptr at that width
        if (variant == TROVE_VARIANT_AUTO || variant == TROVE_VARIANT_SHFL)
            return launch_shfl_tile<T, true>(src, idx, dst, n, tile, s);
        if (variant != TROVE_VARIANT_DIRECT && variant != TROVE_VARIANT_DIRECT_LAUNCH) return TROVE_ERR_VARIANT;
    }
    // 4-byte rows: one group per wavefront. With four, the streaming loads
    // of the next groups slow the random 4-byte stores to half the rate
    // (0.51 vs 0.98 TB/s on MI355X); 8- and 16-byte rows gain from four.
    constexpr int U = sizeof(T) == 4 ? 1 : gather_groups_in_flight<T>();
    const int64_t units = ((n >> 6) + U - 1) / U;
    if (variant == TROVE_VARIANT_AUTO) variant = auto_scatter_variant(B);
    if (variant == TROVE_VARIANT_TILE && !tiled) variant = TROVE_VARIANT_DIRECT;
    if (variant == TROVE_VARIANT_LDS) {
        if constexpr (B % 4 == 0 && B <= kMaxLdsStructBytes) {
            const unsigned blocks = (unsigned)grid_for((n + 63) / 64, 1);
            const size_t lds = trove::lds_stage_bytes<T>();
            if (aligned(src, 16)) {
                hipLaunchKernelGGL((k_scatter_lds<T, 4>), dim3(blocks), dim3(64), lds, s, src, idx, dst, n);
            } else {
                hipLaunchKernelGGL((k_scatter_lds<T, 1>), dim3(blocks), dim3(64), lds, s, src, idx, dst, n);
            }
            return launch_status();
        } else {
            return TROVE_ERR_VARIANT;
        }
    }
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
    if (variant == TROVE_VARIANT_DIRECT_LAUNCH) {
        const int64_t all_units = (((n + 63) >> 6) + U - 1) / U;
        hipLaunchKernelGGL((k_scatter_direct_launch<T, U>), dim3((unsigned)grid_for(all_units > 0 ? all_units : 1)),
                           dim3(kBlock), 0, s, src, idx, dst, n);
        return launch_status();
    }
    return TROVE_ERR_VARIANT;
}

#define TROVE_INST_GATHER(B)                                                                          \
    template int gather_impl<B>(const void*, const int32_t*, void*, int64_t, int, int, hipStream_t);  \
    template int scatter_impl<B>(const void*, const int32_t*, void*, int64_t, int, int, hipStream_t);
TROVE_FOR_EACH_SIZE_IN_SHARD(TROVE_INST_GATHER)

}  // namespace trove_amd
