// Splitting a struct into machine words and joining it back.
//
// Same rule as trove/detail/dismember.h:37-66: the word type U is the widest
// of 16, 8 or 4 bytes that divides sizeof(T) (else 1 byte), and a T is
// m = sizeof(T) / sizeof(U) words. On CDNA a 16-byte word is one
// global_load_dwordx4 per lane; the transposition moves it as 4 dwords.
// lyse / fuse (dismember.h:117-129) copy through memcpy, which the compiler
// lowers to register renaming.
#pragma once
#include <trove/array.h>
#include <trove/utility.h>
#include <cstdint>
#include <type_traits>

namespace trove {
namespace detail {

struct alignas(8) word2 {
    uint32_t x, y;
};
struct alignas(16) word4 {
    uint32_t x, y, z, w;
};

template <typename T>
struct dismember_type {
    typedef typename std::conditional<
        sizeof(T) % 16 == 0, word4,
        typename std::conditional<sizeof(T) % 8 == 0, word2,
                                  typename std::conditional<sizeof(T) % 4 == 0, uint32_t, char>::type>::type>::type
        type;
};

template <typename T, typename U>
struct aliased_size {
    static_assert(sizeof(T) % sizeof(U) == 0, "struct size must be a multiple of the word size");
    static const int value = sizeof(T) / sizeof(U);
};

}  // namespace detail

// Cache policy of the global accesses of the warp-wide paths.
//  cached    - ordinary loads and stores (the reference's behaviour);
//  streaming - nontemporal loads and stores (`nt`): data touched once
//              bypasses cache retention, which measurably raises copy
//              bandwidth on MI355X.
struct cached {};
struct streaming {};

namespace detail {

// One word access. Words of 8 and 16 bytes move as one 2- / 4-dword vector
// (aligned to its size, like word2 / word4), never as separate fields: a
// struct-typed access is split into field accesses early on, and when the
// compiler then merges the code tails of two paths, the fields get separate
// addresses and are no longer re-joined into one dwordx4 access.
template <typename Policy, typename U>
__device__ __forceinline__ U load_word(const U* p) {
    if constexpr (sizeof(U) % 4 == 0) {
        constexpr int W = sizeof(U) / 4;
        typedef typename std::conditional<W == 1, uint32_t, uint32_t __attribute__((ext_vector_type(W)))>::type vec;
        const vec* q = reinterpret_cast<const vec*>(p);
        vec t;
        if constexpr (std::is_same<Policy, streaming>::value) {
            t = __builtin_nontemporal_load(q);
        } else {
            t = *q;
        }
        U u;
        __builtin_memcpy(&u, &t, sizeof(U));
        return u;
    } else {
        return *p;
    }
}

template <typename Policy, typename U>
__device__ __forceinline__ void store_word(U* p, const U& u) {
    if constexpr (sizeof(U) % 4 == 0) {
        constexpr int W = sizeof(U) / 4;
        typedef typename std::conditional<W == 1, uint32_t, uint32_t __attribute__((ext_vector_type(W)))>::type vec;
        vec t;
        __builtin_memcpy(&t, &u, sizeof(U));
        if constexpr (std::is_same<Policy, streaming>::value) {
            __builtin_nontemporal_store(t, reinterpret_cast<vec*>(p));
        } else {
            *reinterpret_cast<vec*>(p) = t;
        }
    } else {
        *p = u;
    }
}

// Struct <-> words, one memcpy per word: a single memcpy of a struct of
// several hundred bytes is not split into registers and lands in scratch
// memory, while word-sized copies stay in VGPRs.
template <typename U, typename T>
__host__ __device__ __forceinline__ void struct_to_words(const T& in, U* w) {
    static_assert(sizeof(T) % sizeof(U) == 0, "struct is not a whole number of words");
#pragma unroll
    for (int j = 0; j < (int)(sizeof(T) / sizeof(U)); ++j)
        __builtin_memcpy(&w[j], reinterpret_cast<const char*>(&in) + j * sizeof(U), sizeof(U));
}

template <typename U, typename T>
__host__ __device__ __forceinline__ void words_to_struct(const U* w, T& out) {
    static_assert(sizeof(T) % sizeof(U) == 0, "struct is not a whole number of words");
#pragma unroll
    for (int j = 0; j < (int)(sizeof(T) / sizeof(U)); ++j)
        __builtin_memcpy(reinterpret_cast<char*>(&out) + j * sizeof(U), &w[j], sizeof(U));
}

template <typename U, typename T>
__host__ __device__ array<U, aliased_size<T, U>::value> lyse(const T& in) {
    array<U, aliased_size<T, U>::value> out;
    static_assert(sizeof(out) == sizeof(T), "word array must alias the struct");
    U w[aliased_size<T, U>::value];
    struct_to_words(in, w);
    words_to_struct(w, out);
    return out;
}

template <typename T, typename U>
__host__ __device__ T fuse(const array<U, aliased_size<T, U>::value>& in) {
    U w[aliased_size<T, U>::value];
    struct_to_words(in, w);
    T out;
    words_to_struct(w, out);
    return out;
}

}  // namespace detail
}  // namespace trove
