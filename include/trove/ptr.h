// coalesced_ptr<T>: a pointer whose element reads and writes are performed
// cooperatively by the converged lanes around the caller.
//
// Public interface of trove/ptr.h:34-90. Reading a coalesced_ref calls
// trove::load, assigning a T calls trove::store. Assigning one coalesced_ref
// to another (dst[i] = src[j]) moves the structs with the same tile cascade
// but without the two transpositions a load followed by a store would run
// (detail::copy_cascade); the reference's version of that operator omits the
// tile argument of its dispatch calls and cannot be instantiated
// (ptr.h:47-50).
#pragma once
#include <trove/aos.h>

namespace trove {
namespace detail {

template <typename T, typename Policy = cached>
struct coalesced_ref {
    T* m_ptr;
    __device__ explicit coalesced_ref(T* ptr) : m_ptr(ptr) {}

    __device__ __forceinline__ operator T() { return trove::load<T, Policy>(m_ptr); }

    __device__ __forceinline__ coalesced_ref& operator=(const T& data) {
        trove::store<T, Policy>(data, m_ptr);
        return *this;
    }

    __device__ __forceinline__ coalesced_ref& operator=(const coalesced_ref& other) {
        trove::detail::copy_cascade<Policy, Policy>(other.m_ptr, m_ptr);
        return *this;
    }

    // From a reference of the other policy: each side keeps its own.
    template <typename P, typename = typename std::enable_if<!std::is_same<P, Policy>::value>::type>
    __device__ __forceinline__ coalesced_ref& operator=(const coalesced_ref<T, P>& other) {
        trove::detail::copy_cascade<P, Policy>(other.m_ptr, m_ptr);
        return *this;
    }
};

}  // namespace detail

// Policy (an addition, defaulted to the reference's behaviour) is the cache
// policy of the element accesses: trove::cached, or trove::streaming for
// arrays touched once (nontemporal loads and stores).
template <typename T, typename Policy = cached>
struct coalesced_ptr {
    T* m_ptr;
    __device__ coalesced_ptr(T* ptr) : m_ptr(ptr) {}
    __device__ trove::detail::coalesced_ref<T, Policy> operator*() { return trove::detail::coalesced_ref<T, Policy>(m_ptr); }
    template <typename I>
    __device__ trove::detail::coalesced_ref<T, Policy> operator[](const I& idx) {
        return trove::detail::coalesced_ref<T, Policy>(m_ptr + idx);
    }
    __device__ operator T*() { return m_ptr; }
};

}  // namespace trove
