// Contiguous AoS copy: dst[i] = src[i].
//  SHFL   - one struct per lane through load_warp_contiguous /
//           store_warp_contiguous<TILE> (the Trove low-level API).
//  DIRECT - the naive per-lane struct copy the compiler generates.
// Also the reference benchmark's store-only (counting structs) and load-only
// (per-struct word sum) kernels, tests/benchmark.cu:52-88.
#include "common.h"

namespace trove_amd {

// 64-struct groups a wavefront keeps in flight per loop iteration.
template <typename T>
constexpr int groups_in_flight() {
    return sizeof(T) <= 16 ? 8 : sizeof(T) <= 32 ? 4 : sizeof(T) <= 64 ? 2 : 1;
}

template <typename T, int TILE, int UNROLL>
__global__ __launch_bounds__(kBlock) void k_copy_shfl(const T* __restrict__ src, T* __restrict__ dst, int64_t n) {
    const int lane = trove::lane_id();
    const int64_t groups = n >> 6;
    const int64_t step = wave_count() * UNROLL;
    int64_t g = wave_index() * UNROLL;
    for (; g + UNROLL <= groups; g += step) {
        T v[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) v[u] = trove::load_warp_contiguous<TILE>(src + ((g + u) << 6) + lane);
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) trove::store_warp_contiguous<TILE>(v[u], dst + ((g + u) << 6) + lane);
    }
    for (; g < groups; ++g) {
        const T v = trove::load_warp_contiguous<TILE>(src + (g << 6) + lane);
        trove::store_warp_contiguous<TILE>(v, dst + (g << 6) + lane);
    }
    // Partial trailing group: the active lanes go through the tile cascade.
    if (wave_index() == 0 && (n & 63)) {
        const int64_t i = (groups << 6) + lane;
        if (i < n) trove::store(trove::load(src + i), dst + i);
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_copy_direct(const T* __restrict__ src, T* __restrict__ dst, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

template <typename T, int TILE>
static int launch_copy_shfl(const T* src, T* dst, int64_t n, hipStream_t s) {
    constexpr int U = groups_in_flight<T>();
    const int64_t units = ((n >> 6) + U - 1) / U;
    const int64_t blocks = grid_for(units > 0 ? units : 1);
    hipLaunchKernelGGL((k_copy_shfl<T, TILE, U>), dim3((unsigned)blocks), dim3(kBlock), 0, s, src, dst, n);
    return launch_status();
}

template <int B>
int copy_impl(const void* src_, void* dst_, int64_t n, int variant, int tile, hipStream_t s) {
    typedef blob_t<B> T;
    const T* src = static_cast<const T*>(src_);
    T* dst = static_cast<T*>(dst_);
    if (variant == TROVE_VARIANT_AUTO || variant == TROVE_VARIANT_SHFL) {
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
    return TROVE_ERR_VARIANT;
}

// ---- store / load benchmarks of the reference (32-bit-word structs) ----

template <typename T>
__device__ __forceinline__ T counting_struct(int64_t i) {
    constexpr int k = sizeof(T) / 4;
    return trove::counting_array<T>::impl((uint32_t)(i * k));
}

template <typename T, int UNROLL>
__global__ __launch_bounds__(kBlock) void k_counting_shfl(T* __restrict__ dst, int64_t n) {
    const int lane = trove::lane_id();
    const int64_t groups = n >> 6;
    const int64_t step = wave_count() * UNROLL;
    int64_t g = wave_index() * UNROLL;
    for (; g + UNROLL <= groups; g += step) {
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            const int64_t i = ((g + u) << 6) + lane;
            trove::store_warp_contiguous(counting_struct<T>(i), dst + i);
        }
    }
    for (; g < groups; ++g) {
        const int64_t i = (g << 6) + lane;
        trove::store_warp_contiguous(counting_struct<T>(i), dst + i);
    }
    if (wave_index() == 0 && (n & 63)) {
        const int64_t i = (groups << 6) + lane;
        if (i < n) trove::store(counting_struct<T>(i), dst + i);
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_counting_direct(T* __restrict__ dst, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) dst[i] = counting_struct<T>(i);
}

template <typename T, int UNROLL>
__global__ __launch_bounds__(kBlock) void k_loadsum_shfl(const T* __restrict__ src, uint32_t* __restrict__ out,
                                                         int64_t n) {
    const int lane = trove::lane_id();
    const int64_t groups = n >> 6;
    const int64_t step = wave_count() * UNROLL;
    int64_t g = wave_index() * UNROLL;
    for (; g + UNROLL <= groups; g += step) {
        T v[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) v[u] = trove::load_warp_contiguous(src + ((g + u) << 6) + lane);
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) out[((g + u) << 6) + lane] = trove::sum(v[u]);
    }
    for (; g < groups; ++g) {
        const int64_t i = (g << 6) + lane;
        out[i] = trove::sum(trove::load_warp_contiguous(src + i));
    }
    if (wave_index() == 0 && (n & 63)) {
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

template <int B>
int counting_impl(void* dst_, int64_t n, int variant, hipStream_t s) {
    if constexpr (B % 4 != 0) {
        return TROVE_ERR_SIZE;
    } else {
        typedef blob_t<B> T;
        T* dst = static_cast<T*>(dst_);
        if (variant == TROVE_VARIANT_AUTO || variant == TROVE_VARIANT_SHFL) {
            constexpr int U = groups_in_flight<T>();
            const int64_t units = ((n >> 6) + U - 1) / U;
            const int64_t blocks = grid_for(units > 0 ? units : 1);
            hipLaunchKernelGGL((k_counting_shfl<T, U>), dim3((unsigned)blocks), dim3(kBlock), 0, s, dst, n);
            return launch_status();
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
int loadsum_impl(const void* src_, uint32_t* out, int64_t n, int variant, hipStream_t s) {
    if constexpr (B % 4 != 0) {
        return TROVE_ERR_SIZE;
    } else {
        typedef blob_t<B> T;
        const T* src = static_cast<const T*>(src_);
        if (variant == TROVE_VARIANT_AUTO || variant == TROVE_VARIANT_SHFL) {
            constexpr int U = groups_in_flight<T>();
            const int64_t units = ((n >> 6) + U - 1) / U;
            const int64_t blocks = grid_for(units > 0 ? units : 1);
            hipLaunchKernelGGL((k_loadsum_shfl<T, U>), dim3((unsigned)blocks), dim3(kBlock), 0, s, src, out, n);
            return launch_status();
        }
        if (variant == TROVE_VARIANT_DIRECT) {
            const int64_t blocks = grid_for((n + 63) / 64);
            hipLaunchKernelGGL((k_loadsum_direct<T>), dim3((unsigned)blocks), dim3(kBlock), 0, s, src, out, n);
            return launch_status();
        }
        return TROVE_ERR_VARIANT;
    }
}

#define TROVE_INST_COPY(B)                                                                    \
    template int copy_impl<B>(const void*, void*, int64_t, int, int, hipStream_t);           \
    template int counting_impl<B>(void*, int64_t, int, hipStream_t);                         \
    template int loadsum_impl<B>(const void*, uint32_t*, int64_t, int, hipStream_t);
TROVE_FOR_EACH_SIZE_IN_SHARD(TROVE_INST_COPY)

}  // namespace trove_amd
