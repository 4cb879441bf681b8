// Lane/slot schedule of the in-register C2R / R2C transposition.
//
// A tile of n lanes (n a power of two) holds n*m elements, m per lane.
//   column side ("C"): lane a, slot b holds element a + b*n   (what coalesced
//                      warp-wide loads produce: memory.h warp_load layout)
//   row side    ("R"): lane c, slot d holds element c*m + d   (lane c owns
//                      struct c, the per-lane view of load_warp_contiguous)
// R2C goes column side -> row side (the load direction, trove/transpose.h
// r2c_warp_transpose, transpose.h:672-696); C2R is its inverse (the store
// direction, transpose.h:638-662).
//
// A cross-lane exchange instruction moves one register slot of every lane,
// so every element must pass through an intermediate slot ("colour") r that
// is a bijection over the elements a lane sends and over those it receives.
// With g = gcd(m, n), m' = m/g, n' = n/g the colour used here is
//     r(w) = (w mod m + floor(w / (m' n))) mod m,
// one formula for odd m (g = 1: r = destination slot), power-of-two m
// (m' = 1) and composite m. On the column side it factors as
//     r = (a + P(b)) mod m,  P(b) = (t*n + q*(m'*n + 1)) mod m,  b = q*m' + t,
// a lane-independent slot permutation followed by a rotation by the lane
// index; on the row side as r = (d + h(c)) mod m with h(c) = floor(c / n').
// Hence, per lane L:
//   R2C: v <- P-permute(v); v <- rotate_right(v, L); v[r] <- pull(v[r] of
//        lane partner(L, r)) for each r; v <- rotate_left(v, h(L))
//   C2R: v <- rotate_right(v, h(L)); push v[r] to lane partner(L, r) for each
//        r; v <- rotate_left(v, L); v <- inverse P-permute(v)
// with partner(c, r) = (c*m + (r - h(c)) mod m) mod n for both directions
// (R2C pulls from it, C2R pushes to it). Odd m has h == 0, so R2C needs one
// runtime rotation and C2R one; even m needs two each, as in the reference.
//
// This is a re-derivation of the rotate / shuffle / static-permute
// decomposition the reference implements in three schemas (transpose.h:
// 38-90 constants, 92-182 permutations, 185-410 offsets, 477-625 steps);
// it follows the same decomposition for in-place transposition (Catanzaro,
// Keller, Garland, PPoPP 2014) but with its own colouring.
#pragma once
#include <hip/hip_runtime.h>
#include <trove/static_gcd.h>

namespace trove {
namespace detail {

template <int n, int m>
struct tx_schedule {
    static_assert(n >= 1 && (n & (n - 1)) == 0, "tile width must be a power of two");
    static_assert(m >= 1 && m <= n, "the transposition needs m <= tile width");
    static constexpr int g = detail::constexpr_gcd(m, n);
    static constexpr int mp = m / g;  // odd part of m
    static constexpr int np = n / g;  // lanes per row-rotation step
    static constexpr bool has_row_rotation = g > 1;

    // Column-side static slot permutation: slot b moves to slot P(b).
    __host__ __device__ static constexpr int P(int b) {
        return ((b % mp) * n + (b / mp) * (mp * n + 1)) % m;
    }
    // Its inverse: P(Pinv(y)) == y.
    __host__ __device__ static constexpr int Pinv(int y) {
        for (int b = 0; b < m; ++b)
            if (P(b) == y) return b;
        return -1;
    }
    // Row-side rotation amount of lane c (always < g <= m).
    __host__ __device__ static constexpr int h(int c) { return c / np; }
    // Lane a tile lane exchanges slot r with (R2C source, C2R destination).
    __host__ __device__ static constexpr int partner(int c, int r) {
        return (c * m + (r - h(c) + m) % m) & (n - 1);
    }
    // Left-rotation amount equivalent to rotating right by the lane index.
    __host__ __device__ static constexpr int unrotate_lane(int c) { return (m - c % m) % m; }
    // Left-rotation amount equivalent to rotating right by h(c).
    __host__ __device__ static constexpr int unrotate_row(int c) { return (m - h(c)) % m; }
};

}  // namespace detail
}  // namespace trove
