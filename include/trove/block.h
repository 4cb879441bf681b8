// Block interface: s consecutive elements per lane as one trove::array.
//
// Public interface of trove/block.h:35-65. Lane work-item `idx` owns
// elements [idx*s, idx*s + s) of the T array. The *_warp_contiguous forms
// need a converged tile with consecutive idx across its lanes and use the
// coalesced transposition path; load_array / store_array are per-lane
// compiler-generated accesses for the non-converged case.
#pragma once
#include <trove/aos.h>
#include <trove/warp.h>

namespace trove {

template <int s, std::size_t tile_size = WARP_SIZE, typename Policy = cached, typename T, typename I>
__device__ __forceinline__ trove::array<T, s> load_array_warp_contiguous(const T* src, const I& idx) {
    typedef trove::array<T, s> array_type;
    const array_type* src_ptr = (const array_type*)(src) + idx;
    return load_warp_contiguous<tile_size, Policy>(src_ptr);
}

template <int s, typename T, typename I>
__device__ __forceinline__ trove::array<T, s> load_array(const T* src, const I& idx) {
    typedef trove::array<T, s> array_type;
    const array_type* src_ptr = (const array_type*)(src) + idx;
    return *src_ptr;
}

template <int s, std::size_t tile_size = WARP_SIZE, typename Policy = cached, typename T, typename I>
__device__ __forceinline__ void store_array_warp_contiguous(T* dest, const I& idx, const trove::array<T, s>& src) {
    typedef trove::array<T, s> array_type;
    array_type* dest_ptr = (array_type*)(dest) + idx;
    store_warp_contiguous<tile_size, Policy>(src, dest_ptr);
}

template <int s, typename T, typename I>
__device__ __forceinline__ void store_array(T* dest, const I& idx, const trove::array<T, s>& src) {
    typedef trove::array<T, s> array_type;
    array_type* dest_ptr = (array_type*)(dest) + idx;
    *dest_ptr = src;
}

}  // namespace trove
