// C ABI over the parts of the reference that compile unmodified on this
// image: its compile-time number theory (trove/static_gcd.h,
// trove/static_mod_inverse.h, trove/utility.h static_log, is_power_of_two,
// is_odd, inverse), its register array, strided per-lane memory access and
// runtime rotation (trove/array.h, trove/memory.h, trove/rotate.h) and its
// test helpers (counting_array, sum; trove/utility.h:33-78). Built by
// oracle/Makefile (target `ref`) from the reference sources in place, host
// code only, into oracle/_ref/libref_parts.so. TEST INFRASTRUCTURE ONLY:
// tests/test_oracle.py compares the oracle against these tables.
#include <trove/static_gcd.h>
#include <trove/static_mod_inverse.h>
#include <trove/utility.h>
#include <trove/rotate.h>
#include <trove/memory.h>

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

// static_mod_inverse<a, m, r> for r <= m (the reference's recursion only
// stops at r == m); -2 marks r > m.
template <int R, int A, int... M>
void inv_r_row(int* out, std::integer_sequence<int, M...>) {
    ((out[((R - 1) * kN + A) * 32 + M] = (R <= M + 1 ? trove::static_mod_inverse<A, M + 1, (R <= M + 1 ? R : M + 1)>::value : -2)), ...);
}
template <int R, int... A>
void inv_r_block(int* out, std::integer_sequence<int, A...>) {
    (inv_r_row<R, A>(out, std::make_integer_sequence<int, 32>{}), ...);
}

template <int... M>
void bits_table(int* out, std::integer_sequence<int, M...>) {
    ((out[M] = int(trove::is_power_of_two<M>::value) + 2 * int(trove::is_odd<M>::value)), ...);
}

template <int p>
struct affine64 {
    static const int value = (5 * p + 3) % 64;
};
template <int P0, int... X>
void inverse_row(int* out, std::integer_sequence<int, X...>) {
    ((out[X] = trove::inverse<int, affine64, X, P0>::value), ...);
}

template <int S>
void memory_one(const int* src, int offset, int stride, int* loaded, int* stored, int* unc) {
    typedef trove::array<int, S> A;
    const A a = trove::warp_load<A>(src, offset, stride);
    trove::make_carray<int, S>(a, loaded);
    trove::warp_store(a, stored, offset, stride);
    trove::uncoalesced_store(a, unc, 3);
}

}  // namespace

extern "C" {

// out[((r - 1) * 33 + a) * 32 + m - 1] = static_mod_inverse<a, m, r>::value
// for 1 <= r <= 6, 0 <= a <= 32, 1 <= m <= 32 (-2 where r > m).
void ref_mod_inverse_r_table(int* out) {
    inv_r_block<1>(out, std::make_integer_sequence<int, kN>{});
    inv_r_block<2>(out, std::make_integer_sequence<int, kN>{});
    inv_r_block<3>(out, std::make_integer_sequence<int, kN>{});
    inv_r_block<4>(out, std::make_integer_sequence<int, kN>{});
    inv_r_block<5>(out, std::make_integer_sequence<int, kN>{});
    inv_r_block<6>(out, std::make_integer_sequence<int, kN>{});
}
// out[m] = is_power_of_two<m> + 2 * is_odd<m>, 0 <= m <= 256.
void ref_bits_table(int* out) { bits_table(out, std::make_integer_sequence<int, 257>{}); }
// out[x] = inverse<int, (5p + 3) % 64, x>::value, out[64 + x] the same
// search from p = 10, 0 <= x < 64.
void ref_inverse_table(int* out) {
    inverse_row<0>(out, std::make_integer_sequence<int, 64>{});
    inverse_row<10>(out + 64, std::make_integer_sequence<int, 64>{});
}
// One lane of trove/memory.h on int arrays of s elements (s in 1, 2, 5, 16):
// loaded = warp_load(src, offset, stride), warp_store of it into stored, and
// uncoalesced_store of it to unc with stride 3; returns -1 for another s.
int ref_warp_memory(int s, const int* src, int offset, int stride, int* loaded, int* stored, int* unc) {
    switch (s) {
        case 1: memory_one<1>(src, offset, stride, loaded, stored, unc); return 0;
        case 2: memory_one<2>(src, offset, stride, loaded, stored, unc); return 0;
        case 5: memory_one<5>(src, offset, stride, loaded, stored, unc); return 0;
        case 16: memory_one<16>(src, offset, stride, loaded, stored, unc); return 0;
        default: return -1;
    }
}

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
