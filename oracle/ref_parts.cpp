// C ABI over the parts of the reference that compile unmodified on this
// image: its compile-time number theory (trove/static_gcd.h,
// trove/static_mod_inverse.h, trove/utility.h static_log), its register
// array and runtime rotation (trove/array.h, trove/rotate.h) and its test
// helpers (counting_array, sum; trove/utility.h:33-78). Built by
// oracle/Makefile (target `ref`) from the reference sources in place, host
// code only, into oracle/_ref/libref_parts.so. TEST INFRASTRUCTURE ONLY:
// tests/test_oracle.py compares the oracle against these tables.
#include <trove/static_gcd.h>
#include <trove/static_mod_inverse.h>
#include <trove/utility.h>
#include <trove/rotate.h>

#include <utility>

namespace {

constexpr int kN = 33;  // tables over 1..32

template <int U, int... V>
void gcd_row(int* out, std::integer_sequence<int, V...>) {
    ((out[U * kN + V + 1] = trove::static_gcd<U, V + 1>::value), ...);
}
template <int... U>
void gcd_table(int* out, std::integer_sequence<int, U...>) {
    (gcd_row<U + 1>(out, std::make_integer_sequence<int, kN - 1>{}), ...);
}

template <int A, int... M>
void inv_row(int* out, std::integer_sequence<int, M...>) {
    ((out[A * kN + M + 1] = trove::static_mod_inverse<A, M + 1>::value), ...);
}
template <int... A>
void inv_table(int* out, std::integer_sequence<int, A...>) {
    (inv_row<A>(out, std::make_integer_sequence<int, kN - 1>{}), ...);
}

template <int... M>
void log_table(int* out, std::integer_sequence<int, M...>) {
    ((out[M + 1] = trove::static_log<M + 1>::value), ...);
}

template <int M>
void rotate_one(const int* t, int a, int* r) {
    trove::array<int, M> x = trove::make_array<int, M>(const_cast<int*>(t));
    trove::make_carray<int, M>(trove::rotate(x, a), r);
}

template <int M>
int counting_sum(int v, int step, int* words) {
    trove::array<int, M> x = trove::counting_array<trove::array<int, M> >::impl(v, step);
    trove::make_carray<int, M>(x, words);
    return trove::sum(x);
}

}  // namespace

extern "C" {

// out[u * 33 + v] = static_gcd<u, v>::value, 1 <= u, v <= 32.
void ref_gcd_table(int* out) { gcd_table(out, std::make_integer_sequence<int, kN - 1>{}); }
// out[a * 33 + m] = static_mod_inverse<a, m>::value, 0 <= a <= 32, 1 <= m <= 32.
void ref_mod_inverse_table(int* out) { inv_table(out, std::make_integer_sequence<int, kN>{}); }
// out[m] = static_log<m>::value, 1 <= m <= 32.
void ref_log_table(int* out) { log_table(out, std::make_integer_sequence<int, kN - 1>{}); }

// r = rotate(t, a) for an array of m ints, 1 <= m <= 16; returns -1 otherwise.
int ref_rotate(int m, const int* t, int a, int* r) {
    switch (m) {
#define RC(M) \
    case M: rotate_one<M>(t, a, r); return 0;
        RC(1) RC(2) RC(3) RC(4) RC(5) RC(6) RC(7) RC(8) RC(9) RC(10) RC(11) RC(12) RC(13) RC(14) RC(15) RC(16)
#undef RC
        default: return -1;
    }
}

// words = counting_array<array<int, m>>::impl(v, step); returns sum(words).
int ref_counting_sum(int m, int v, int step, int* words) {
    switch (m) {
#define CC(M) \
    case M: return counting_sum<M>(v, step, words);
        CC(1) CC(2) CC(3) CC(4) CC(5) CC(6) CC(7) CC(8) CC(9) CC(10) CC(11) CC(12) CC(13) CC(14) CC(15) CC(16)
#undef CC
        default: return 0;
    }
}

}  // extern "C"
