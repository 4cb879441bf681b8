// Contiguous AoS copy: dst[i] = src[i].
//  SHFL   - one struct per lane through load_warp_contiguous /
//           store_warp_contiguous<TILE> (the Trove low-level API), with the
//           streaming (nontemporal) cache policy.
//  LDS    - the LDS-staged transposition (load/store_warp_contiguous_lds):
//           dwordx4 tile traffic through a per-wavefront LDS stage.
//  DIRECT - the naive per-lane struct copy the compiler generates.
//  DIRECT_LAUNCH - the naive per-lane struct access (one struct per lane,
//           its words loaded and stored by the lane) in SHFL's launch: the
//           same grid, XCD order and residency cap, and SHFL's streaming
//           policy where a per-lane access allows it (one-word structs), so
//           that it differs from SHFL only by the transposition.
// Also the reference benchmark's store-only (counting structs) and load-only
// (per-struct word sum) kernels, tests/benchmark.cu:52-88.
//
// Launch shape (measured on MI355X, 2^26 structs): a one-shot grid, one
// work-group of B threads per wavefront-sized unit, each wavefront moving
// UNROLL 64-struct groups. Structs of more than one word per lane (the
// transposition path) stream best at UNROLL = 1 with 64-thread groups; one-
// word structs need 1 KiB per wavefront (UNROLL = 16 / words) to keep
// enough bytes in flight.
#include "common.h"

namespace trove_amd {

// XCD-contiguous mapping (unit_for_xcd) for structs of 20 bytes and more:
// +7-9% at 36-60 bytes, but -13% at 12 bytes, where dispatch order wins.
// The launch also caps resident work-groups per CU (stream_lds_reservation).
template <typename T>
struct copy_shape {
    static constexpr int words = sizeof(T) / 4 > 0 ? (int)(sizeof(T) / 4) : 1;
    static constexpr int m = trove::detail::word_layout<T>::m;
    static constexpr int unroll = (m == 1 && words < 4) ? 4 / words : 1;
    static constexpr int block = 64;
    static constexpr bool xcd_map = m > 1 && sizeof(T) >= 20;
};

// One 64-struct group at p (the group's first struct): through the
// warp-contiguous pair (NAIVE false) or per lane (NAIVE true; streaming for
// one-word structs only, see per_lane_policy).
template <typename T, int TILE, bool NAIVE>
__device__ __forceinline__ T copy_group_load(const T* p, int lane) {
    if constexpr (NAIVE) {
        return trove::detail::divergent_load<per_lane_policy<T, trove::streaming>>(p + lane);
    } else {
        return trove::load_warp_contiguous<TILE, trove::streaming>(p + lane);
    }
}
template <typename T, int TILE, bool NAIVE>
__device__ __forceinline__ void copy_group_store(const T& v, T* p, int lane) {
    if constexpr (NAIVE) {
        trove::detail::divergent_store<per_lane_policy<T, trove::streaming>>(v, p + lane);
    } else {
        trove::store_warp_contiguous<TILE, trove::streaming>(v, p + lane);
    }
}

template <typename T, int TILE, int UNROLL, int B, bool XCD, bool NAIVE = false>
__global__ __launch_bounds__(B) void k_copy_shfl(const T* __restrict__ src, T* __restrict__ dst, int64_t n) {
    const int lane = trove::lane_id();
    const int64_t groups = n >> 6;
    if constexpr (XCD) {
        // One unit (UNROLL groups) per wavefront; the grid has exactly
        // ceil(groups / UNROLL) wavefronts.
        const int64_t units = (groups + UNROLL - 1) / UNROLL;
        const int64_t w = wave_index<B>();
        if (w < units) {
            const int64_t g0 = unit_for_xcd(w, units) * UNROLL;
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) {
                if (g0 + u < groups) {
                    const T v = copy_group_load<T, TILE, NAIVE>(src + ((g0 + u) << 6), lane);
                    copy_group_store<T, TILE, NAIVE>(v, dst + ((g0 + u) << 6), lane);
                }
            }
        }
        return;
    }
    const int64_t step = wave_count<B>() * UNROLL;
    int64_t g = wave_index<B>() * UNROLL;
    for (; g + UNROLL <= groups; g += step) {
        T v[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) v[u] = copy_group_load<T, TILE, NAIVE>(src + ((g + u) << 6), lane);
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) copy_group_store<T, TILE, NAIVE>(v[u], dst + ((g + u) << 6), lane);
    }
    for (; g < groups; ++g) {
        const T v = copy_group_load<T, TILE, NAIVE>(src + (g << 6), lane);
        copy_group_store<T, TILE, NAIVE>(v, dst + (g << 6), lane);
    }
}

// LDS-staged variant (trove/lds.h): the same one-shot grid of one-wavefront
// work-groups, each staging its UNROLL groups (UNROLL * 64 * sizeof(T)
// contiguous bytes) through as many bytes of dynamic LDS. The unit moves as
// ONE chunk array, so that small structs (UNROLL > 1) still give every lane a
// full chunk: at 4 bytes one dwordx4 per lane instead of four 16-lane ones.
// CW is the chunk width (4 dwords when src and dst are 16-byte aligned, else
// 1). Struct l of each group passes through lane l's registers between the
// two stagings, as with load/store_warp_contiguous_lds.
template <int W>
struct unit_words {
    uint32_t w[W];
};

template <typename T, int UNROLL, int CW>
__global__ __launch_bounds__(64) void k_copy_lds(const T* __restrict__ src, T* __restrict__ dst, int64_t n) {
    extern __shared__ __align__(16) uint32_t lds_stage[];
    namespace d = trove::detail;
    constexpr int k = (int)(sizeof(T) / 4);
    typedef unit_words<UNROLL * k> Unit;  // 64 units = the wavefront's UNROLL groups
    // The chunks take Unit's stage layout and the structs T's: they agree
    // when the unit is one group, or when neither is swizzled (multi-group
    // units only occur for 4- and 8-byte structs).
    static_assert(UNROLL == 1 || (!d::lds_swizzle<T>::on && !d::lds_swizzle<Unit>::on), "stage layouts differ");
    const int lane = trove::lane_id();
    const int64_t groups = n >> 6;
    const int64_t g0 = (int64_t)blockIdx.x * UNROLL;
    if (g0 + UNROLL <= groups) {
        d::lds_chunks<Unit, CW> t;
        d::lds_global_load<CW, trove::streaming>(reinterpret_cast<const Unit*>(src + (g0 << 6)), lane, t);
        d::lds_put_chunks(lds_stage, lane, t);
        d::wave_lds_sync();
        T v[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) v[u] = d::lds_get_struct<T>(lds_stage + u * 64 * k, lane);
        d::wave_lds_sync();
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) d::lds_put_struct(lds_stage + u * 64 * k, lane, v[u]);
        d::wave_lds_sync();
        d::lds_get_chunks(lds_stage, lane, t);
        d::lds_global_store<CW, trove::streaming>(reinterpret_cast<Unit*>(dst + (g0 << 6)), lane, t);
    } else {
        for (int64_t g = g0; g < groups; ++g) {
            const T v = trove::load_warp_contiguous_lds<trove::streaming>(src + (g << 6) + lane, lds_stage);
            trove::store_warp_contiguous_lds<trove::streaming>(v, dst + (g << 6) + lane, lds_stage);
        }
    }
}

// Partial trailing group (n % 64 structs), one wavefront: the active lanes go
// through the tile cascade. A kernel of its own, so that the streaming kernel
// does not carry the cascade's registers (up to 512 per lane at 1 KiB).
template <typename T, bool NAIVE = false>
__global__ __launch_bounds__(64) void k_copy_tail(const T* __restrict__ src, T* __restrict__ dst, int64_t n) {
    const int64_t i = ((n >> 6) << 6) + trove::lane_id();
    if (i < n) {
        if constexpr (NAIVE) {
            typedef per_lane_policy<T, trove::streaming> P;
            trove::detail::divergent_store<P>(trove::detail::divergent_load<P>(src + i), dst + i);
        } else {
            trove::store(trove::load(src + i), dst + i);
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_copy_direct(const T* __restrict__ src, T* __restrict__ dst, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

template <typename T, int TILE, bool NAIVE = false>
static int launch_copy_shfl(const T* src, T* dst, int64_t n, hipStream_t s) {
    typedef copy_shape<T> C;
    const int64_t units = ((n >> 6) + C::unroll - 1) / C::unroll;
    const int64_t blocks = grid_for(units, C::block / 64);
    const size_t lds = stream_lds_reservation((int64_t)C::unroll * 64 * (int64_t)sizeof(T) * (C::block / 64));
    if (n >= 64) {
        hipLaunchKernelGGL((k_copy_shfl<T, TILE, C::unroll, C::block, C::xcd_map, NAIVE>), dim3((unsigned)blocks),
                           dim3(C::block), lds, s, src, dst, n);
        const int rc = launch_status();
        if (rc != TROVE_OK) return rc;
    }
    if (n & 63) {
        hipLaunchKernelGGL((k_copy_tail<T, NAIVE>), dim3(1), dim3(64), 0, s, src, dst, n);
        return launch_status();
    }
    return TROVE_OK;
}

template <typename T>
static int launch_copy_lds(const T* src, T* dst, int64_t n, hipStream_t s) {
    constexpr int words = (int)(sizeof(T) / 4);
    constexpr int unroll = words < 4 ? 4 / words : 1;
    const int64_t groups = n >> 6;
    if (groups > 0) {
        const int64_t blocks = grid_for((groups + unroll - 1) / unroll, 1);
        const size_t stage = (size_t)unroll * 64 * sizeof(T);
        const size_t window = stream_lds_reservation((int64_t)stage);
        const size_t lds = window > stage ? window : stage;
        if (aligned(src, 16) && aligned(dst, 16)) {
            hipLaunchKernelGGL((k_copy_lds<T, unroll, 4>), dim3((unsigned)blocks), dim3(64), lds, s, src, dst, n);
        } else {
            hipLaunchKernelGGL((k_copy_lds<T, unroll, 1>), dim3((unsigned)blocks), dim3(64), lds, s, src, dst, n);
        }
        const int rc = launch_status();
        if (rc != TROVE_OK) return rc;
    }
    if (n & 63) {
        hipLaunchKernelGGL((k_copy_tail<T>), dim3(1), dim3(64), 0, s, src, dst, n);
        return launch_status();
    }
    return TROVE_OK;
}

template <int B>
int copy_impl(const void* src_, void* dst_, int64_t n, int variant, int tile, hipStream_t s) {
    typedef blob_t<B> T;
    const T* src = static_cast<const T*>(src_);
    T* dst = static_cast<T*>(dst_);
    if (variant == TROVE_VARIANT_AUTO)  // a sub-wavefront tile asks for the transposition
        variant = (tile == 0 || tile == 64) ? auto_copy_variant(B) : TROVE_VARIANT_SHFL;
    if (variant == TROVE_VARIANT_LDS) {
        if constexpr (B % 4 == 0 && B <= kMaxLdsStructBytes) {
            return launch_copy_lds<T>(src, dst, n, s);
        } else {
            return TROVE_ERR_VARIANT;
        }
    }
    if (variant == TROVE_VARIANT_SHFL) {
        switch (tile == 0 ? 64 : tile) {
            case 64: return launch_copy_shfl<T, 64>(src, dst, n, s);
            case 32: return launch_copy_shfl<T, 32>(src, dst, n, s);
            case 16: return launch_copy_shfl<T, 16>(src, dst, n, s);
            case 8: return launch_copy_shfl<T, 8>(src, dst, n, s);
            case 4: return launch_copy_shfl<T, 4>(src, dst, n, s);
            case 2: return launch_copy_shfl<T, 2>(src, dst, n, s);
            default: return TROVE_ERR_VARIANT;
        }
    }
    if (variant == TROVE_VARIANT_DIRECT) {
        const int64_t blocks = grid_for((n + 63) / 64);
        hipLaunchKernelGGL((k_copy_direct<T>), dim3((unsigned)blocks), dim3(kBlock), 0, s, src, dst, n);
        return launch_status();
    }
    if (variant == TROVE_VARIANT_DIRECT_LAUNCH) return launch_copy_shfl<T, 64, true>(src, dst, n, s);
    return TROVE_ERR_VARIANT;
}

// ---- store / load benchmarks of the reference (32-bit-word structs) ----

template <typename T>
__device__ __forceinline__ T counting_struct(int64_t i) {
    constexpr int k = sizeof(T) / 4;
    return trove::counting_array<T>::impl((uint32_t)(i * k));
}

// Wavefront w of the one-shot grid (XCD: work-groups in XCD-contiguous order).
template <int B, bool XCD>
__device__ __forceinline__ int64_t bench_wave() {
    if constexpr (XCD) {
        return unit_for_xcd(blockIdx.x, (int64_t)gridDim.x) * (B / 64) +
               __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    } else {
        return wave_index<B>();
    }
}

// TILE: the store_warp_contiguous / load_warp_contiguous tile width (the
// reference benchmark's TILE_SIZE, tests/Makefile:4; 64 = a whole wavefront).
template <typename T, int TILE, int UNROLL, int B, bool XCD>
__global__ __launch_bounds__(B) void k_counting_shfl(T* __restrict__ dst, int64_t n) {
    const int lane = trove::lane_id();
    const int64_t groups = n >> 6;
    const int64_t step = wave_count<B>() * UNROLL;
    int64_t g = bench_wave<B, XCD>() * UNROLL;
    for (; g + UNROLL <= groups; g += step) {
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            const int64_t i = ((g + u) << 6) + lane;
            trove::store_warp_contiguous<TILE, trove::streaming>(counting_struct<T>(i), dst + i);
        }
    }
    for (; g < groups; ++g) {
        const int64_t i = (g << 6) + lane;
        trove::store_warp_contiguous<TILE, trove::streaming>(counting_struct<T>(i), dst + i);
    }
    if (wave_index<B>() == 0 && (n & 63)) {
        const int64_t i = (groups << 6) + lane;
        if (i < n) trove::store(counting_struct<T>(i), dst + i);
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_counting_direct(T* __restrict__ dst, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) dst[i] = counting_struct<T>(i);
}

template <typename T, int TILE, int UNROLL, int B, bool XCD>
__global__ __launch_bounds__(B) void k_loadsum_shfl(const T* __restrict__ src, uint32_t* __restrict__ out,
                                                    int64_t n) {
    const int lane = trove::lane_id();
    const int64_t groups = n >> 6;
    const int64_t step = wave_count<B>() * UNROLL;
    int64_t g = bench_wave<B, XCD>() * UNROLL;
    for (; g + UNROLL <= groups; g += step) {
        T v[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u)
            v[u] = trove::load_warp_contiguous<TILE, trove::streaming>(src + ((g + u) << 6) + lane);
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) out[((g + u) << 6) + lane] = trove::sum(v[u]);
    }
    for (; g < groups; ++g) {
        const int64_t i = (g << 6) + lane;
        out[i] = trove::sum(trove::load_warp_contiguous<TILE, trove::streaming>(src + i));
    }
    if (wave_index<B>() == 0 && (n & 63)) {
        const int64_t i = (groups << 6) + lane;
        if (i < n) out[i] = trove::sum(trove::load(src + i));
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_loadsum_direct(const T* __restrict__ src, uint32_t* __restrict__ out,
                                                           int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) out[i] = trove::sum(src[i]);
}

// Launch shape of the store-only and load-only kernels per struct size, in
// 4-byte words: work-group size, XCD-contiguous work-group order, and a cap
// on resident work-groups per CU (cap_lds_reservation; 0 = none). Chosen on
// MI355X from sweeps at 2^24 and 2^26 structs (single launches and ten back
// to back), then checked A/B against the previous shapes with the reference
// benchmark program at 16,384,000 structs and at 2^26:
// - A store-only wavefront of small structs finishes so quickly that
//   dispatching 64-thread work-groups bounds the kernel: 256-thread
//   work-groups store 12-20-byte structs at 5.7-7.1 TB/s against 3.1-5.9
//   with 64 threads.
// - XCD order lifts stores from 12 bytes and loads from 20 bytes by 2-14%;
//   it costs the 16-byte store 4-8% at 2^24 structs.
// - Stores of 20 and 32-64 bytes stream best with a cap of 2-11 work-groups
//   (5-16 wavefronts, about 10-25 KiB of stores per CU): 5.9-7.1 TB/s in
//   the reference program at its size against 5.3-6.5 uncapped in dispatch
//   order (0.93-1.3x of it at 2^26). Loads of 44-52 and 60-64 bytes take a
//   looser cap of 16-21 work-groups.
// One-word structs of 4, 8 and 16 bytes keep 256-thread work-groups in
// dispatch order; above 64 bytes only the XCD order applies (no cap measured
// there).
template <typename T>
struct bench_shape {
    static constexpr int words = (int)(sizeof(T) / 4);
    static constexpr int store_block =
        (words <= 5 || words == 8 || words == 10 || words == 11 || words == 13) ? 256 : 64;
    static constexpr int store_cap = words == 5                                   ? 4
                                     : words == 8 || words == 10 || words == 11 || words == 13 ? 2
                                     : words == 6                                 ? 11
                                     : words == 9                                 ? 8
                                     : words == 12 || (words >= 14 && words <= 16) ? 5
                                                                                   : 0;
    static constexpr bool store_xcd = words == 3 || words >= 5;
    static constexpr int load_block = (words == 3 || words == 4) ? 256 : 64;
    static constexpr int load_cap = words >= 11 && words <= 13 ? 21 : words >= 15 && words <= 16 ? 16 : 0;
    static constexpr bool load_xcd = words >= 5;
};

// Launches the store-only / load-only kernel at tile width TILE with the
// launch shape measured for the whole-wavefront tile.
template <typename T, int TILE>
static int launch_counting(T* dst, int64_t n, hipStream_t s) {
    typedef copy_shape<T> C;
    typedef bench_shape<T> H;
    constexpr int WB = H::store_block;
    const int64_t units = ((n >> 6) + C::unroll - 1) / C::unroll;
    const int64_t blocks = grid_for(units, WB / 64);
    hipLaunchKernelGGL((k_counting_shfl<T, TILE, C::unroll, WB, H::store_xcd>), dim3((unsigned)blocks), dim3(WB),
                       cap_lds_reservation(H::store_cap), s, dst, n);
    return launch_status();
}

template <typename T, int TILE>
static int launch_loadsum(const T* src, uint32_t* out, int64_t n, hipStream_t s) {
    typedef copy_shape<T> C;
    typedef bench_shape<T> H;
    constexpr int WB = H::load_block;
    const int64_t units = ((n >> 6) + C::unroll - 1) / C::unroll;
    const int64_t blocks = grid_for(units, WB / 64);
    hipLaunchKernelGGL((k_loadsum_shfl<T, TILE, C::unroll, WB, H::load_xcd>), dim3((unsigned)blocks), dim3(WB),
                       cap_lds_reservation(H::load_cap), s, src, out, n);
    return launch_status();
}

template <int B>
int counting_impl(void* dst_, int64_t n, int variant, int tile, hipStream_t s) {
    if constexpr (B % 4 != 0 || B > kMaxBenchStructBytes) {
        return TROVE_ERR_SIZE;
    } else {
        typedef blob_t<B> T;
        T* dst = static_cast<T*>(dst_);
        if (variant == TROVE_VARIANT_AUTO || variant == TROVE_VARIANT_SHFL) {
            switch (tile == 0 ? 64 : tile) {
                case 64: return launch_counting<T, 64>(dst, n, s);
                case 32: return launch_counting<T, 32>(dst, n, s);
                case 16: return launch_counting<T, 16>(dst, n, s);
                case 8: return launch_counting<T, 8>(dst, n, s);
                case 4: return launch_counting<T, 4>(dst, n, s);
                case 2: return launch_counting<T, 2>(dst, n, s);
                default: return TROVE_ERR_VARIANT;
            }
        }
        if (variant == TROVE_VARIANT_DIRECT) {
            const int64_t blocks = grid_for((n + 63) / 64);
            hipLaunchKernelGGL((k_counting_direct<T>), dim3((unsigned)blocks), dim3(kBlock), 0, s, dst, n);
            return launch_status();
        }
        return TROVE_ERR_VARIANT;
    }
}

template <int B>
int loadsum_impl(const void* src_, uint32_t* out, int64_t n, int variant, int tile, hipStream_t s) {
    if constexpr (B % 4 != 0 || B > kMaxBenchStructBytes) {
        return TROVE_ERR_SIZE;
    } else {
        typedef blob_t<B> T;
        const T* src = static_cast<const T*>(src_);
        if (variant == TROVE_VARIANT_AUTO || variant == TROVE_VARIANT_SHFL) {
            switch (tile == 0 ? 64 : tile) {
                case 64: return launch_loadsum<T, 64>(src, out, n, s);
                case 32: return launch_loadsum<T, 32>(src, out, n, s);
                case 16: return launch_loadsum<T, 16>(src, out, n, s);
                case 8: return launch_loadsum<T, 8>(src, out, n, s);
                case 4: return launch_loadsum<T, 4>(src, out, n, s);
                case 2: return launch_loadsum<T, 2>(src, out, n, s);
                default: return TROVE_ERR_VARIANT;
            }
        }
        if (variant == TROVE_VARIANT_DIRECT) {
            const int64_t blocks = grid_for((n + 63) / 64);
            hipLaunchKernelGGL((k_loadsum_direct<T>), dim3((unsigned)blocks), dim3(kBlock), 0, s, src, out, n);
            return launch_status();
        }
        return TROVE_ERR_VARIANT;
    }
}

#define TROVE_INST_COPY(B)                                                          \
    template int copy_impl<B>(const void*, void*, int64_t, int, int, hipStream_t); \
    template int counting_impl<B>(void*, int64_t, int, int, hipStream_t);          \
    template int loadsum_impl<B>(const void*, uint32_t*, int64_t, int, int, hipStream_t);
TROVE_FOR_EACH_SIZE_IN_SHARD(TROVE_INST_COPY)

}  // namespace trove_amd
