// trove::array<T, m>: a statically indexed, value-semantics register array.
//
// Public interface of trove/array.h:32-200: the head/tail cons-list layout
// (so sizeof(array<T, m>) == m * sizeof(T) and an array<T, s>* may alias a
// T* run of s elements, which the block interface relies on), get<i>,
// make_array (variadic and from a C array) and make_carray. Indexing is by
// compile-time position only, so the compiler keeps every element in a VGPR.
// Copies are the implicit member-wise ones, so an array of trivially copyable
// elements is itself trivially copyable and may be moved with memcpy.
#pragma once
#include <hip/hip_runtime.h>

namespace trove {

template <typename T, int m>
struct array {
    typedef T value_type;
    typedef T head_type;
    typedef array<T, m - 1> tail_type;
    static const int size = m;
    head_type head;
    tail_type tail;
    __host__ __device__ array(head_type h, const tail_type& t) : head(h), tail(t) {}
    __host__ __device__ array() : head(), tail() {}
    array(const array& other) = default;
    array& operator=(const array& other) = default;
    __host__ __device__ bool operator==(const array& other) const {
        return head == other.head && tail == other.tail;
    }
    __host__ __device__ bool operator!=(const array& other) const { return !(*this == other); }
};

template <typename T>
struct array<T, 1> {
    typedef T value_type;
    typedef T head_type;
    static const int size = 1;
    head_type head;
    __host__ __device__ array(head_type h) : head(h) {}
    __host__ __device__ array() : head() {}
    array(const array& other) = default;
    array& operator=(const array& other) = default;
    __host__ __device__ bool operator==(const array& other) const { return head == other.head; }
    __host__ __device__ bool operator!=(const array& other) const { return !(*this == other); }
};

template <typename T>
struct array<T, 0> {};

template <int i, typename T, int m>
__host__ __device__ T& get(array<T, m>& src) {
    static_assert(i >= 0 && i < m, "trove::get index out of range");
    if constexpr (i == 0) {
        return src.head;
    } else {
        return get<i - 1>(src.tail);
    }
}

template <int i, typename T, int m>
__host__ __device__ T get(const array<T, m>& src) {
    static_assert(i >= 0 && i < m, "trove::get index out of range");
    if constexpr (i == 0) {
        return src.head;
    } else {
        return get<i - 1>(src.tail);
    }
}

template <typename T>
__host__ __device__ auto make_array() {
    return array<T, 0>();
}

template <typename T>
__host__ __device__ auto make_array(T a0) {
    return array<T, 1>(a0);
}

template <typename T, typename... Args>
__host__ __device__ auto make_array(T a0, Args... args) {
    return array<T, 1 + sizeof...(Args)>(a0, make_array<T>(args...));
}

namespace detail {
template <typename T, int s>
__host__ __device__ array<T, s> array_from_c(const T* c) {
    if constexpr (s == 1) {
        return array<T, 1>(c[0]);
    } else {
        return array<T, s>(c[0], array_from_c<T, s - 1>(c + 1));
    }
}

template <typename T, int s>
__host__ __device__ void array_to_c(const array<T, s>& a, T* c) {
    c[0] = a.head;
    if constexpr (s > 1) array_to_c<T, s - 1>(a.tail, c + 1);
}
}  // namespace detail

// Build an array from a C array of s elements (s given explicitly).
template <typename T, int s>
__host__ __device__ array<T, s> make_array(T cary[s]) {
    return detail::array_from_c<T, s>(cary);
}

// Write an array out to a C array of s elements.
template <typename T, int s>
__host__ __device__ void make_carray(const array<T, s>& ary, T result[s]) {
    detail::array_to_c<T, s>(ary, result);
}

}  // namespace trove
