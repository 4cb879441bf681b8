// Per-lane access used when a tile is not converged or a struct is not
// eligible for the transposition.
//
// Contract of trove/detail/fallback.h:67-129: when sizeof(T) is a multiple of
// 4 the struct moves as its m words of type U (dismember_type), each lane on
// its own; otherwise it is an ordinary compiler-generated access.
#pragma once
#include <trove/detail/dismember.h>

namespace trove {
namespace detail {

template <typename T>
struct use_divergent {
    static const bool value = (sizeof(T) % 4) == 0;
};

template <typename T>
__device__ __forceinline__ typename enable_if<use_divergent<T>::value, T>::type divergent_load(const T* src) {
    typedef typename dismember_type<T>::type U;
    constexpr int m = aliased_size<T, U>::value;
    const U* p = reinterpret_cast<const U*>(src);
    U w[m];
#pragma unroll
    for (int j = 0; j < m; ++j) w[j] = p[j];
    T out;
    __builtin_memcpy(&out, w, sizeof(T));
    return out;
}

template <typename T>
__device__ __forceinline__ typename enable_if<!use_divergent<T>::value, T>::type divergent_load(const T* src) {
    return *src;
}

template <typename T>
__device__ __forceinline__ typename enable_if<use_divergent<T>::value>::type divergent_store(const T& data, T* dest) {
    typedef typename dismember_type<T>::type U;
    constexpr int m = aliased_size<T, U>::value;
    U w[m];
    __builtin_memcpy(w, &data, sizeof(T));
    U* p = reinterpret_cast<U*>(dest);
#pragma unroll
    for (int j = 0; j < m; ++j) p[j] = w[j];
}

template <typename T>
__device__ __forceinline__ typename enable_if<!use_divergent<T>::value>::type divergent_store(const T& data, T* dest) {
    *dest = data;
}

}  // namespace detail
}  // namespace trove
