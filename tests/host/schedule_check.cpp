// Host check of the device library's transposition schedule
// (include/trove/detail/schedule.h) and of trove::rotate.
//
// For every tile width n in {2, ..., 64} and every m in [1, n], the R2C step
// sequence of trove/transpose.h, driven by tx_schedule<n, m>, must map the
// column-side layout (lane a, slot b = a + b*n) to the row-side layout
// (lane c, slot d = c*m + d), and the C2R sequence must invert it, with the
// C2R pushes landing on distinct (lane, slot) pairs. The dynamic path's word
// walk (trove/aos.h: word j of rank id belongs to struct (id + j*n) / m at
// word (id + j*n) % m) must equal that integer division for every rank, and
// the LDS stage swizzle must be a conflict-free permutation (swizzle_checks).
// Exits non-zero on the first failure.
#include <trove/aos.h>
#include <trove/array.h>
#include <trove/detail/schedule.h>
#include <trove/lds.h>
#include <trove/rotate.h>
#include <trove/utility.h>

#include <cstdio>
#include <vector>

using trove::detail::tx_schedule;

static void rotl(std::vector<int>& x, int m, int amt) {
    std::vector<int> t(m);
    for (int j = 0; j < m; ++j) t[j] = x[(j + amt) % m];
    x = t;
}

template <int n, int m>
static bool check() {
    using S = tx_schedule<n, m>;
    std::vector<std::vector<int>> v(n, std::vector<int>(m));
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < m; ++b) v[a][b] = a + b * n;
    // R2C: static permutation P, rotate right by the lane, pull, rotate by h.
    for (int L = 0; L < n; ++L) {
        std::vector<int> t(m);
        for (int b = 0; b < m; ++b) t[S::P(b)] = v[L][b];
        v[L] = t;
        rotl(v[L], m, S::unrotate_lane(L));
    }
    auto x = v;
    for (int L = 0; L < n; ++L)
        for (int r = 0; r < m; ++r) x[L][r] = v[S::partner(L, r)][r];
    for (int L = 0; L < n; ++L)
        if (S::has_row_rotation) rotl(x[L], m, S::h(L));
    for (int c = 0; c < n; ++c)
        for (int d = 0; d < m; ++d)
            if (x[c][d] != c * m + d) {
                std::printf("R2C n=%d m=%d lane %d slot %d holds %d\n", n, m, c, d, x[c][d]);
                return false;
            }
    // C2R: rotate right by h, push, rotate left by the lane, inverse P.
    auto y = x;
    for (int L = 0; L < n; ++L)
        if (S::has_row_rotation) rotl(y[L], m, S::unrotate_row(L));
    auto z = y;
    std::vector<int> seen(n * m, 0);
    for (int L = 0; L < n; ++L)
        for (int r = 0; r < m; ++r) {
            const int dst = S::partner(L, r);
            if (seen[dst * m + r]++) {
                std::printf("C2R n=%d m=%d: two pushes to lane %d slot %d\n", n, m, dst, r);
                return false;
            }
            z[dst][r] = y[L][r];
        }
    for (int L = 0; L < n; ++L) {
        rotl(z[L], m, L % m);
        std::vector<int> t(m);
        for (int b = 0; b < m; ++b) t[b] = z[L][S::P(b)];
        z[L] = t;
    }
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < m; ++b)
            if (z[a][b] != a + b * n) {
                std::printf("C2R n=%d m=%d lane %d slot %d holds %d\n", n, m, a, b, z[a][b]);
                return false;
            }
    for (int b = 0; b < m; ++b)
        if (S::P(S::Pinv(b)) != b) return false;
    return true;
}

template <int n, int m>
static int all_m() {
    if constexpr (m > n) {
        return 0;
    } else {
        return (check<n, m>() ? 0 : 1) + all_m<n, m + 1>();
    }
}

template <int n, int m>
static int walk_checks() {
    if constexpr (m > n) {
        return 0;
    } else {
        int fails = 0;
        for (int id = 0; id < n; ++id) {
            int owner[m], word[m];
            trove::detail::word_owners<n, m>(id, owner, word);
            for (int j = 0; j < m; ++j)
                if (owner[j] != (id + j * n) / m || word[j] != (id + j * n) % m) {
                    if (!fails) std::printf("word walk n=%d m=%d id %d j %d\n", n, m, id, j);
                    ++fails;
                }
        }
        return fails + walk_checks<n, m + 1>();
    }
}

template <int m>
static int rotate_checks() {
    int fails = 0;
    auto a = trove::counting_array<trove::array<int, m> >::impl(100);
    int limit = 1;
    while (limit < m) limit <<= 1;
    for (int amt = 0; amt < limit; ++amt) {
        auto r = trove::rotate(a, amt);
        int c[m], base[m];
        trove::make_carray<int, m>(r, c);
        trove::make_carray<int, m>(a, base);
        for (int j = 0; j < m; ++j)
            if (c[j] != base[(j + amt) % m]) ++fails;
    }
    return fails;
}

// LDS stage swizzle (trove/lds.h) for structs of m 16-byte words: the map
// (struct l, word j) -> slot is a permutation of [0, 64 m) that keeps every
// struct in its own m slots (the word-index form agrees), and for every word
// j the 16 structs of each LDS pass land in 16 different 16-byte bank groups
// (slot mod 16), where the unswizzled layout gives 16 / gcd(m, 16) of them.
template <int m>
static int swizzle_checks() {
    typedef trove::array<uint32_t, 4 * m> T;
    using S = trove::detail::lds_swizzle<T>;
    static_assert(S::on, "16-byte words");
    int f = 0;
    std::vector<int> seen(64 * m, 0);
    for (int l = 0; l < 64; ++l)
        for (int j = 0; j < m; ++j) {
            const int s = S::of(l, j);
            if (s < l * m || s >= l * m + m || S::of(l * m + j) != s) ++f;
            if (s >= 0 && s < 64 * m) seen[s]++;
        }
    for (int x : seen) f += x != 1;
    for (int j = 0; j < m; ++j)
        for (int p = 0; p < 4; ++p) {
            int groups = 0;
            for (int l = 16 * p; l < 16 * p + 16; ++l) groups |= 1 << (S::of(l, j) & 15);
            f += groups != 0xFFFF;
        }
    if (f) std::printf("lds swizzle m=%d: %d failures\n", m, f);
    return f;
}

int main() {
    int f = all_m<2, 1>() + all_m<4, 1>() + all_m<8, 1>() + all_m<16, 1>() + all_m<32, 1>() + all_m<64, 1>();
    f += rotate_checks<2>() + rotate_checks<3>() + rotate_checks<5>() + rotate_checks<7>() + rotate_checks<12>() +
         rotate_checks<15>() + rotate_checks<16>();
    f += walk_checks<2, 1>() + walk_checks<4, 1>() + walk_checks<8, 1>() + walk_checks<16, 1>() +
         walk_checks<32, 1>() + walk_checks<64, 1>();
    f += swizzle_checks<2>() + swizzle_checks<3>() + swizzle_checks<4>() + swizzle_checks<6>() +
         swizzle_checks<8>() + swizzle_checks<10>() + swizzle_checks<12>() + swizzle_checks<14>() +
         swizzle_checks<16>() + swizzle_checks<24>() + swizzle_checks<32>();
    std::printf("schedule/rotate/word-walk/swizzle failures: %d\n", f);
    return f == 0 ? 0 : 1;
}
