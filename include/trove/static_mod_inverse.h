// Compile-time modular inverse.
//
// Same contract as trove/static_mod_inverse.h:32-52: value is the smallest
// r' in [r, m) with (r' * a) % m == 1, or -1 when there is none (including
// m <= r). The search starts at r = 1 by default. Evaluated by a constexpr
// search instead of template recursion.
#pragma once

namespace trove {

namespace detail {
constexpr int constexpr_mod_inverse(int a, int m, int r = 1) {
    for (; r < m; ++r) {
        if ((r * a) % m == 1) return r;
    }
    return -1;
}
}  // namespace detail

template <int a, int m, int r = 1>
struct static_mod_inverse {
    static constexpr int value = detail::constexpr_mod_inverse(a, m, r);
};

}  // namespace trove
