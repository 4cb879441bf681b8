// Compile-time helpers and the test-data helpers of the public interface.
//
// Names and meanings follow trove/utility.h:33-159: counting_array and sum
// (the benchmark's iota structs and checksum), static_log (floor log2),
// is_power_of_two, is_odd, value_if, value_identity, inverse (preimage of a
// compile-time map), null_type, cons_c, static_range, enable_if and
// size_multiple_power_of_two. The implementations are constexpr functions
// rather than recursive template chains.
#pragma once
#include <trove/array.h>
#include <cstddef>

namespace trove {

// counting_array<array<T, s>>::impl(v, i) = {v, v + i, v + 2i, ...}.
template <typename T>
struct counting_array {};

template <typename T, int s>
struct counting_array<array<T, s> > {
    typedef array<T, s> Array;
    __host__ __device__ static Array impl(T v = 0, T i = 1) {
        if constexpr (s == 1) {
            return Array(v);
        } else {
            return Array(v, counting_array<array<T, s - 1> >::impl(v + i, i));
        }
    }
};

// Sum of the elements of an array (wrapping arithmetic for integers).
template <typename T, int s>
__host__ __device__ T sum(const array<T, s>& a) {
    if constexpr (s == 1) {
        return a.head;
    } else {
        return a.head + sum(a.tail);
    }
}

namespace detail {
constexpr int constexpr_floor_log2(int m) {
    int r = 0;
    while (m > 1) {
        m >>= 1;
        ++r;
    }
    return r;
}
}  // namespace detail

template <int m>
struct static_log {
    static_assert(m > 0, "static_log of a non-positive value");
    static constexpr int value = detail::constexpr_floor_log2(m);
};

template <int m>
struct is_power_of_two {
    static constexpr bool value = (m & (m - 1)) == 0;
};

template <int m>
struct is_odd {
    static constexpr bool value = (m & 1) == 1;
};

template <bool cond, typename T, typename Then, typename Else>
struct value_if {
    static constexpr T value = Then::value;
};

template <typename T, typename Then, typename Else>
struct value_if<false, T, Then, Else> {
    static constexpr T value = Else::value;
};

template <typename T, T x>
struct value_identity {
    static constexpr T value = x;
};

// inverse<T, Fn, x>::value: the smallest p >= p0 with Fn<p>::value == x.
template <typename T, template <T> class Fn, T x, T p = 0>
struct inverse {
    static constexpr T value =
        value_if<Fn<p>::value == x, T, value_identity<T, p>, inverse<T, Fn, x, p + 1> >::value;
};

struct null_type {};

template <typename T, T i, typename Tail = null_type>
struct cons_c {
    static const T head = i;
    typedef Tail tail;
};

template <int k, int l>
struct static_range {
    static const int head = k;
    typedef static_range<k + 1, l> tail;
};

template <int f>
struct static_range<f, f> {
    static const int head = f;
    typedef null_type tail;
};

template <bool b, typename T = void>
struct enable_if {
    typedef T type;
};

template <typename T>
struct enable_if<false, T> {};

// True when sizeof(T) is a multiple of 2^p.
template <typename T, int p>
struct size_multiple_power_of_two {
    static constexpr bool value = (sizeof(T) & ((std::size_t(1) << p) - 1)) == 0;
};

}  // namespace trove
