// Strided per-lane memory access used by the warp-wide paths.
//
// Public interface of trove/memory.h:136-168: warp_load / warp_store move
// element k of the array at ptr[offset + k * stride], and uncoalesced_store
// writes element k to ptr[k]: its stride argument is accepted and ignored,
// as in the reference (memory.h:59-98 advances the offset by 1). When every
// lane of a tile passes offset = its rank and stride = the tile width, each
// element access is one coalesced wavefront instruction (global_load_dword,
// _dwordx2 or _dwordx4 depending on the element width).
#pragma once
#include <trove/array.h>

namespace trove {

template <typename Array>
__host__ __device__ __forceinline__ void warp_store(const Array& t, typename Array::head_type* ptr, int offset, int stride) {
    ptr[offset] = t.head;
    if constexpr (Array::size > 1) warp_store(t.tail, ptr, offset + stride, stride);
}

template <typename Array>
__host__ __device__ __forceinline__ Array warp_load(const typename Array::head_type* ptr, int offset, int stride) {
    if constexpr (Array::size == 1) {
        return Array(ptr[offset]);
    } else {
        return Array(ptr[offset], warp_load<typename Array::tail_type>(ptr, offset + stride, stride));
    }
}

template <typename Array>
__host__ __device__ __forceinline__ Array warp_load(const volatile typename Array::head_type* ptr, int offset, int stride) {
    if constexpr (Array::size == 1) {
        return Array(ptr[offset]);
    } else {
        return Array(ptr[offset], warp_load<typename Array::tail_type>(ptr, offset + stride, stride));
    }
}

template <typename Array>
__host__ __device__ __forceinline__ void uncoalesced_store(const Array& t, typename Array::head_type* ptr, int stride = 1) {
    ptr[0] = t.head;
    if constexpr (Array::size > 1) uncoalesced_store(t.tail, ptr + 1, stride);
}

template <typename Array>
__host__ __device__ __forceinline__ void uncoalesced_store(const Array& t, volatile typename Array::head_type* ptr,
                                           int stride = 1) {
    ptr[0] = t.head;
    if constexpr (Array::size > 1) uncoalesced_store(t.tail, ptr + 1, stride);
}

}  // namespace trove
