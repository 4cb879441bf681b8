// Compile-time modular inverse.
//
// Same contract as trove/static_mod_inverse.h:32-52: value is the smallest
// r in [1, m) with (r * a) % m == 1, or -1 when no such r exists (including
// m <= 1). Evaluated by a constexpr search instead of template recursion.
#pragma once

namespace trove {

namespace detail {
constexpr int constexpr_mod_inverse(int a, int m) {
    for (int r = 1; r < m; ++r) {
        if ((r * a) % m == 1) return r;
    }
    return -1;
}
}  // namespace detail

template <int a, int m, int r = 1>
struct static_mod_inverse {
    static_assert(r == 1, "the search start is fixed at 1");
    static constexpr int value = detail::constexpr_mod_inverse(a, m);
};

}  // namespace trove
