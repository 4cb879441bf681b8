// Compile-time greatest common divisor.
//
// Mirrors the public name and meaning of trove/static_gcd.h:59-75
// (static_gcd<u, v>::value). The reference evaluates it by binary-gcd
// template recursion and declares the zero-argument specialisations as
// `bool` (static_gcd.h:67-75), so static_gcd<0, v>::value collapses to 1;
// here gcd(0, v) = v and gcd(u, 0) = u, as in the mathematics.
#pragma once

namespace trove {

namespace detail {
constexpr int constexpr_gcd(int u, int v) {
    while (v != 0) {
        int t = u % v;
        u = v;
        v = t;
    }
    return u < 0 ? -u : u;
}
}  // namespace detail

template <int u, int v>
struct static_gcd {
    static constexpr int value = detail::constexpr_gcd(u, v);
};

}  // namespace trove
