// In-register C2R / R2C transposition across the lanes of a tile.
//
// Public interface of trove/transpose.h:629-696 (r2c/c2r_compute_indices,
// r2c/c2r_warp_transpose with and without precomputed indices). The schedule
// is the one of trove/detail/schedule.h; the exchange is lowered to
//   R2C: ds_bpermute_b32 (pull from the partner lane), one per dword,
//   C2R: ds_permute_b32  (push to the partner lane),   one per dword,
// so both directions share one partner table. Static slot permutations are
// register renaming; runtime rotations are select chains (rotate.h).
//
// The `indices` / `rotation` pair of the public functions holds, for the
// calling lane, the tile-relative partner lane of every slot and the lane
// rotation of the direction (R2C: left-rotation applied before the exchange;
// C2R: left-rotation applied after it). Even m also needs the row rotation,
// which the transpose recomputes from the tile rank, as the reference's
// power-of-two and composite schemas do (transpose.h:492-523, 594-625).
#pragma once
#include <trove/array.h>
#include <trove/rotate.h>
#include <trove/warp.h>
#include <trove/detail/schedule.h>
#include <cstdint>

namespace trove {
namespace detail {

// R2C on m elements of W dwords each (flat in v), tile width n.
// `id` is the rank in the tile, `first` the wavefront lane of rank 0.
template <int n, int m, int W>
__device__ __forceinline__ void r2c_words(uint32_t (&v)[m * W], int id, int first, const int* partner,
                                          int lane_rot) {
    typedef tx_schedule<n, m> S;
    uint32_t t[m * W];
#pragma unroll
    for (int b = 0; b < m; ++b)
#pragma unroll
        for (int x = 0; x < W; ++x) t[S::P(b) * W + x] = v[b * W + x];
    rotate_words_left<m, W>(t, lane_rot);
#pragma unroll
    for (int r = 0; r < m; ++r) {
        const int addr = (first + partner[r]) << 2;
#pragma unroll
        for (int x = 0; x < W; ++x) t[r * W + x] = (uint32_t)__builtin_amdgcn_ds_bpermute(addr, (int)t[r * W + x]);
    }
    if constexpr (S::has_row_rotation) rotate_words_left<m, W>(t, S::h(id));
#pragma unroll
    for (int i = 0; i < m * W; ++i) v[i] = t[i];
}

// C2R on m elements of W dwords each: the inverse of r2c_words.
template <int n, int m, int W>
__device__ __forceinline__ void c2r_words(uint32_t (&v)[m * W], int id, int first, const int* partner,
                                          int lane_rot) {
    typedef tx_schedule<n, m> S;
    if constexpr (S::has_row_rotation) rotate_words_left<m, W>(v, S::unrotate_row(id));
#pragma unroll
    for (int r = 0; r < m; ++r) {
        const int addr = (first + partner[r]) << 2;
#pragma unroll
        for (int x = 0; x < W; ++x) v[r * W + x] = (uint32_t)__builtin_amdgcn_ds_permute(addr, (int)v[r * W + x]);
    }
    rotate_words_left<m, W>(v, lane_rot);
    uint32_t t[m * W];
#pragma unroll
    for (int b = 0; b < m; ++b)
#pragma unroll
        for (int x = 0; x < W; ++x) t[b * W + x] = v[S::P(b) * W + x];
#pragma unroll
    for (int i = 0; i < m * W; ++i) v[i] = t[i];
}

template <int n, int m>
__device__ __forceinline__ void partner_table(int id, int (&partner)[m]) {
#pragma unroll
    for (int r = 0; r < m; ++r) partner[r] = tx_schedule<n, m>::partner(id, r);
}

// Self-contained forms: derive the per-lane tables from the tile rank.
template <int n, int m, int W>
__device__ __forceinline__ void r2c_words(uint32_t (&v)[m * W], int id, int first) {
    int partner[m];
    partner_table<n, m>(id, partner);
    r2c_words<n, m, W>(v, id, first, partner, tx_schedule<n, m>::unrotate_lane(id));
}

template <int n, int m, int W>
__device__ __forceinline__ void c2r_words(uint32_t (&v)[m * W], int id, int first) {
    int partner[m];
    partner_table<n, m>(id, partner);
    c2r_words<n, m, W>(v, id, first, partner, id % m);
}

template <typename T, int i>
struct array_words {
    static_assert(sizeof(T) % 4 == 0, "transposed elements must be whole dwords");
    static constexpr int W = sizeof(T) / 4;
};

}  // namespace detail

template <typename Tile, int i>
__device__ __forceinline__ void c2r_compute_indices(array<int, i>& indices, int& rotation, const Tile& tile) {
    const int id = tile.id();
    int p[i];
    detail::partner_table<Tile::size(), i>(id, p);
    indices = make_array<int, i>(p);
    rotation = id % i;
}

template <typename Tile, int i>
__device__ __forceinline__ void r2c_compute_indices(array<int, i>& indices, int& rotation, const Tile& tile) {
    const int id = tile.id();
    int p[i];
    detail::partner_table<Tile::size(), i>(id, p);
    indices = make_array<int, i>(p);
    rotation = detail::tx_schedule<Tile::size(), i>::unrotate_lane(id);
}

template <typename Tile, typename T, int i>
__device__ __forceinline__ void c2r_warp_transpose(array<T, i>& src, const array<int, i>& indices, int rotation, const Tile& tile) {
    constexpr int W = detail::array_words<T, i>::W;
    uint32_t v[i * W];
    int p[i];
    __builtin_memcpy(v, &src, sizeof(v));
    make_carray<int, i>(indices, p);
    detail::c2r_words<Tile::size(), i, W>(v, tile.id(), tile.first_lane(), p, rotation);
    __builtin_memcpy(&src, v, sizeof(v));
}

template <typename Tile, typename T, int i>
__device__ __forceinline__ void c2r_warp_transpose(array<T, i>& src, const Tile& tile) {
    constexpr int W = detail::array_words<T, i>::W;
    uint32_t v[i * W];
    __builtin_memcpy(v, &src, sizeof(v));
    detail::c2r_words<Tile::size(), i, W>(v, tile.id(), tile.first_lane());
    __builtin_memcpy(&src, v, sizeof(v));
}

template <typename Tile, typename T, int i>
__device__ __forceinline__ void r2c_warp_transpose(array<T, i>& src, const array<int, i>& indices, int rotation, const Tile& tile) {
    constexpr int W = detail::array_words<T, i>::W;
    uint32_t v[i * W];
    int p[i];
    __builtin_memcpy(v, &src, sizeof(v));
    make_carray<int, i>(indices, p);
    detail::r2c_words<Tile::size(), i, W>(v, tile.id(), tile.first_lane(), p, rotation);
    __builtin_memcpy(&src, v, sizeof(v));
}

template <typename Tile, typename T, int i>
__device__ __forceinline__ void r2c_warp_transpose(array<T, i>& src, const Tile& tile) {
    constexpr int W = detail::array_words<T, i>::W;
    uint32_t v[i * W];
    __builtin_memcpy(v, &src, sizeof(v));
    detail::r2c_words<Tile::size(), i, W>(v, tile.id(), tile.first_lane());
    __builtin_memcpy(&src, v, sizeof(v));
}

}  // namespace trove
