// Block interface kernels (tests/block.cu:13-54) and plain device helpers.
#include "common.h"

namespace trove_amd {

// Lanes of one 64-array group that a seeded hash sends down the fallback
// path before the convergence test: one group in eight is affected, and in
// it each lane with probability 1/2.
__device__ __forceinline__ bool divergent_lane(uint64_t seed, int64_t index) {
    const uint64_t h = splitmix64(seed ^ (uint64_t)(index >> 6));
    if ((h & 7) != 0) return false;
    return (splitmix64(h + (uint64_t)(index & 63)) & 1) != 0;
}

template <int S>
__global__ __launch_bounds__(kBlock) void k_block_copy(const uint32_t* __restrict__ x, uint32_t* __restrict__ r,
                                                       int64_t l, int64_t e, int divergent, uint64_t seed) {
    typedef trove::array<uint32_t, S> s_ary;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    const int64_t gi = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    for (int64_t index = gi; index < l; index += stride) {
        if (divergent && divergent_lane(seed, index)) {
            const s_ary d = trove::load_array<S>(x, index);
            trove::store_array(r, index, d);
        } else if (trove::warp_converged()) {
            const s_ary d = trove::load_array_warp_contiguous<S>(x, index);
            trove::store_array_warp_contiguous(r, index, d);
        } else {
            const s_ary d = trove::load_array<S>(x, index);
            trove::store_array(r, index, d);
        }
    }
    for (int64_t j = l * S + gi; j < e; j += stride) r[j] = x[j];
}

template <int S>
__global__ __launch_bounds__(kBlock) void k_block_write(uint32_t* __restrict__ r, int64_t l) {
    typedef trove::array<uint32_t, S> s_ary;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t index = (int64_t)blockIdx.x * kBlock + threadIdx.x; index < l; index += stride) {
        const s_ary d = trove::counting_array<s_ary>::impl((uint32_t)(S * index));
        if (trove::warp_converged()) {
            trove::store_array_warp_contiguous<S>(r, index, d);
        } else {
            trove::store_array(r, index, d);
        }
    }
}

template <int S>
static int launch_block_copy(const uint32_t* x, uint32_t* r, int64_t e, int divergent, uint64_t seed,
                             hipStream_t s) {
    const int64_t l = e / S;
    const int64_t blocks = grid_for((l + 63) / 64 > 0 ? (l + 63) / 64 : 1);
    hipLaunchKernelGGL((k_block_copy<S>), dim3((unsigned)blocks), dim3(kBlock), 0, s, x, r, l, e, divergent, seed);
    return launch_status();
}

template <int S>
static int launch_block_write(uint32_t* r, int64_t l, hipStream_t s) {
    const int64_t blocks = grid_for((l + 63) / 64 > 0 ? (l + 63) / 64 : 1);
    hipLaunchKernelGGL((k_block_write<S>), dim3((unsigned)blocks), dim3(kBlock), 0, s, r, l);
    return launch_status();
}

#define TROVE_FOR_EACH_S(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16)

int block_copy_dispatch(const uint32_t* x, uint32_t* r, int64_t e, int s, int divergent, uint64_t seed,
                        hipStream_t st) {
    switch (s) {
#define TROVE_CASE_BC(S) \
    case S: return launch_block_copy<S>(x, r, e, divergent, seed, st);
        TROVE_FOR_EACH_S(TROVE_CASE_BC)
#undef TROVE_CASE_BC
        default: return TROVE_ERR_INVALID;
    }
}

int block_write_dispatch(uint32_t* r, int64_t l, int s, hipStream_t st) {
    switch (s) {
#define TROVE_CASE_BW(S) \
    case S: return launch_block_write<S>(r, l, st);
        TROVE_FOR_EACH_S(TROVE_CASE_BW)
#undef TROVE_CASE_BW
        default: return TROVE_ERR_INVALID;
    }
}

// ---- plain device helpers ----

// Plain copy in the fastest shape measured on MI355X: one-shot grid, 16
// bytes per lane, 4 KiB per work-group, nontemporal loads and stores.
__global__ __launch_bounds__(kBlock) void k_memcpy16(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                     int64_t n16) {
    typedef uint32_t vec __attribute__((ext_vector_type(4)));
    const vec* s = reinterpret_cast<const vec*>(src);
    vec* d = reinterpret_cast<vec*>(dst);
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n16; i += stride)
        __builtin_nontemporal_store(__builtin_nontemporal_load(s + i), d + i);
}

int memcpy_ceiling_launch(const void* src, void* dst, int64_t bytes, hipStream_t s) {
    const int64_t n16 = bytes / 16;
    const int64_t blocks = grid_for((n16 + 63) / 64);
    hipLaunchKernelGGL(k_memcpy16, dim3((unsigned)blocks), dim3(kBlock), 0, s, static_cast<const uint4*>(src),
                       static_cast<uint4*>(dst), n16);
    return launch_status();
}

__host__ __device__ __forceinline__ uint32_t word_hash(uint64_t seed, uint64_t i) {
    return (uint32_t)(splitmix64(seed * 0x632BE59BD9B4E019ull + i) >> 32);
}

__global__ __launch_bounds__(kBlock) void k_fill_words(uint32_t* __restrict__ out, int64_t n, uint64_t seed) {
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) out[i] = word_hash(seed, i);
}

int fill_words_launch(uint32_t* out, int64_t n, uint64_t seed, hipStream_t s) {
    const int64_t blocks = grid_for((n + 63) / 64 > 0 ? (n + 63) / 64 : 1);
    hipLaunchKernelGGL(k_fill_words, dim3((unsigned)blocks), dim3(kBlock), 0, s, out, n, seed);
    return launch_status();
}

uint32_t host_word_hash(uint64_t seed, uint64_t i) { return word_hash(seed, i); }

}  // namespace trove_amd
