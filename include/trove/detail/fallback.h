// Per-lane access used when a tile is not converged or a struct is not
// eligible for the transposition.
//
// Contract of trove/detail/fallback.h:67-129: when sizeof(T) is a multiple of
// 4 the struct moves as its m words of type U (dismember_type), each lane on
// its own, with the cache policy given (cached by default); otherwise it is
// an ordinary compiler-generated access.
#pragma once
#include <trove/detail/dismember.h>

namespace trove {
namespace detail {

template <typename T>
struct use_divergent {
    static const bool value = (sizeof(T) % 4) == 0;
};

// Each word is copied into (out of) the struct on its own: one memcpy of a
// whole struct of several hundred bytes is not split into registers and
// lands in scratch memory.
template <typename Policy = cached, typename T>
__device__ __forceinline__ typename enable_if<use_divergent<T>::value, T>::type divergent_load(const T* src) {
    typedef typename dismember_type<T>::type U;
    constexpr int m = aliased_size<T, U>::value;
    const U* p = reinterpret_cast<const U*>(src);
    T out;
#pragma unroll
    for (int j = 0; j < m; ++j) {
        const U u = load_word<Policy>(p + j);
        __builtin_memcpy(reinterpret_cast<char*>(&out) + j * sizeof(U), &u, sizeof(U));
    }
    return out;
}

template <typename Policy = cached, typename T>
__device__ __forceinline__ typename enable_if<!use_divergent<T>::value, T>::type divergent_load(const T* src) {
    return *src;
}

template <typename Policy = cached, typename T>
__device__ __forceinline__ typename enable_if<use_divergent<T>::value>::type divergent_store(const T& data, T* dest) {
    typedef typename dismember_type<T>::type U;
    constexpr int m = aliased_size<T, U>::value;
    U* p = reinterpret_cast<U*>(dest);
#pragma unroll
    for (int j = 0; j < m; ++j) {
        U u;
        __builtin_memcpy(&u, reinterpret_cast<const char*>(&data) + j * sizeof(U), sizeof(U));
        store_word<Policy>(p + j, u);
    }
}

template <typename Policy = cached, typename T>
__device__ __forceinline__ typename enable_if<!use_divergent<T>::value>::type divergent_store(const T& data, T* dest) {
    *dest = data;
}

}  // namespace detail
}  // namespace trove
