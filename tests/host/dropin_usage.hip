// Compile-only check that kernels written against Trove's documented API
// (README.md "High-level", "Block" and "Low-level" interfaces;
// tests/benchmark.cu; tests/block.cu) build unchanged against
// include/trove for gfx950.
#include <trove/aos.h>
#include <trove/block.h>
#include <trove/lds.h>
#include <trove/ptr.h>

template <typename T>
__global__ void trove_gather(const int length, const int* indices, trove::coalesced_ptr<T> src,
                             trove::coalesced_ptr<T> dst) {
    int global_index = threadIdx.x + blockDim.x * blockIdx.x;
    if (global_index < length) {
        int index = indices[global_index];
        T data = src[index];
        dst[global_index] = data;
    }
}

template <typename T, int s>
__global__ void test_block_copy(const T* x, T* r, int l) {
    typedef trove::array<T, s> s_ary;
    int global_index = threadIdx.x + blockIdx.x * blockDim.x;
    for (int index = global_index; index < l; index += gridDim.x * blockDim.x) {
        if (trove::warp_converged()) {
            s_ary d = trove::load_array_warp_contiguous<s>(x, index);
            trove::store_array_warp_contiguous(r, index, d);
        } else {
            s_ary d = trove::load_array<s>(x, index);
            trove::store_array(r, index, d);
        }
    }
}

template <typename T>
__global__ void low_level(T* s, typename T::value_type* r) {
    int global_index = threadIdx.x + blockDim.x * blockIdx.x;
    T data = trove::load_warp_contiguous<32>(s + global_index);
    r[global_index] = trove::sum(data);
    trove::store_warp_contiguous(data, s + global_index);
}

template <typename T>
__global__ void ref_to_ref(const int* indices, T* raw_s, T* raw_r) {
    int global_index = threadIdx.x + blockDim.x * blockIdx.x;
    trove::coalesced_ptr<T> s(raw_s);
    trove::coalesced_ptr<T> r(raw_r);
    r[global_index] = s[indices[global_index]];
}

template <typename T>
__global__ void transpose_api(T* p) {
    trove::thread_tile<64> tile;
    trove::array<int, 5> idx;
    int rot;
    trove::array<T, 5> v = trove::make_array<T>(p[0], p[1], p[2], p[3], p[4]);
    trove::r2c_compute_indices(idx, rot, tile);
    trove::r2c_warp_transpose(v, idx, rot, tile);
    trove::c2r_compute_indices(idx, rot, tile);
    trove::c2r_warp_transpose(v, idx, rot, tile);
    trove::c2r_warp_transpose(v, tile);
    trove::r2c_warp_transpose(v, tile);
    p[threadIdx.x] = trove::get<2>(trove::rotate(v, threadIdx.x & 3));
}

// The LDS-staged addition (trove/lds.h) in a user kernel: four wavefronts
// per work-group, each with its own stage.
template <typename T>
__global__ __launch_bounds__(256) void lds_staged_copy(const T* src, T* dst) {
    __shared__ __align__(16) uint32_t stage[4 * trove::lds_stage_bytes<T>() / 4];
    uint32_t* mine = stage + (threadIdx.x / 64) * (trove::lds_stage_bytes<T>() / 4);
    const int i = blockIdx.x * 256 + threadIdx.x;
    T v = trove::load_warp_contiguous_lds<trove::streaming>(src + i, mine);
    trove::store_warp_contiguous_lds(v, dst + i, mine);
}

struct odd_struct {
    char c[6];
};

template __global__ void trove_gather<trove::array<int, 5> >(const int, const int*,
                                                             trove::coalesced_ptr<trove::array<int, 5> >,
                                                             trove::coalesced_ptr<trove::array<int, 5> >);
template __global__ void trove_gather<trove::array<float, 12> >(const int, const int*,
                                                                trove::coalesced_ptr<trove::array<float, 12> >,
                                                                trove::coalesced_ptr<trove::array<float, 12> >);
template __global__ void trove_gather<odd_struct>(const int, const int*, trove::coalesced_ptr<odd_struct>,
                                                  trove::coalesced_ptr<odd_struct>);
template __global__ void test_block_copy<int, 5>(const int*, int*, int);
template __global__ void test_block_copy<float, 7>(const float*, float*, int);
template __global__ void low_level<trove::array<int, 9> >(trove::array<int, 9>*, int*);
template __global__ void ref_to_ref<trove::array<int, 6> >(const int*, trove::array<int, 6>*,
                                                           trove::array<int, 6>*);
template __global__ void transpose_api<int>(int*);
template __global__ void transpose_api<double>(double*);
template __global__ void lds_staged_copy<trove::array<int, 5> >(const trove::array<int, 5>*, trove::array<int, 5>*);
template __global__ void lds_staged_copy<trove::array<float, 8> >(const trove::array<float, 8>*,
                                                                 trove::array<float, 8>*);
