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
// Every step acts on the W dwords of an element alike, so with ByPlane the
// transposition runs one dword plane (dword x of every element) at a time:
// a barrel-shifter stage then keeps m temporaries alive instead of m * W,
// at the price of less independent work between the exchanges.
template <int n, int m, int W, bool ByPlane = false>
__device__ __forceinline__ void r2c_words(uint32_t (&v)[m * W], int id, int first, const int* partner,
                                          int lane_rot) {
    typedef tx_schedule<n, m> S;
    constexpr int PW = ByPlane ? 1 : W;  // dwords of an element per pass
#pragma unroll
    for (int x0 = 0; x0 < W; x0 += PW) {
        uint32_t t[m * PW];
#pragma unroll
        for (int b = 0; b < m; ++b)
#pragma unroll
            for (int x = 0; x < PW; ++x) t[S::P(b) * PW + x] = v[b * W + x0 + x];
        rotate_words_left<m, PW>(t, lane_rot);
#pragma unroll
        for (int r = 0; r < m; ++r) {
            const int addr = (first + partner[r]) << 2;
#pragma unroll
            for (int x = 0; x < PW; ++x)
                t[r * PW + x] = (uint32_t)__builtin_amdgcn_ds_bpermute(addr, (int)t[r * PW + x]);
        }
        if constexpr (S::has_row_rotation) rotate_words_left<m, PW>(t, S::h(id));
#pragma unroll
        for (int r = 0; r < m; ++r)
#pragma unroll
            for (int x = 0; x < PW; ++x) v[r * W + x0 + x] = t[r * PW + x];
    }
}

// C2R on m elements of W dwords each: the inverse of r2c_words.
template <int n, int m, int W, bool ByPlane = false>
__device__ __forceinline__ void c2r_words(uint32_t (&v)[m * W], int id, int first, const int* partner,
                                          int lane_rot) {
    typedef tx_schedule<n, m> S;
    constexpr int PW = ByPlane ? 1 : W;
#pragma unroll
    for (int x0 = 0; x0 < W; x0 += PW) {
        uint32_t t[m * PW];
#pragma unroll
        for (int r = 0; r < m; ++r)
#pragma unroll
            for (int x = 0; x < PW; ++x) t[r * PW + x] = v[r * W + x0 + x];
        if constexpr (S::has_row_rotation) rotate_words_left<m, PW>(t, S::unrotate_row(id));
#pragma unroll
        for (int r = 0; r < m; ++r) {
            const int addr = (first + partner[r]) << 2;
#pragma unroll
            for (int x = 0; x < PW; ++x)
                t[r * PW + x] = (uint32_t)__builtin_amdgcn_ds_permute(addr, (int)t[r * PW + x]);
        }
        rotate_words_left<m, PW>(t, lane_rot);
#pragma unroll
        for (int b = 0; b < m; ++b)
#pragma unroll
            for (int x = 0; x < PW; ++x) v[b * W + x0 + x] = t[S::P(b) * PW + x];
    }
}

// Whether the coalesced_ptr path (load_dispatch / store_dispatch) transposes
// one dword plane at a time: for structs of 16-byte words holding 84 to 112
// dwords (336 to 448 bytes), where the full-width form's second copy of the
// struct pushes the kernel past the register budget of two wavefronts per
// SIMD and one plane does not. Measured on MI355X at 384 bytes (2^22
// structs): coalesced_ptr gather 3.83 -> 5.41 TB/s (274 -> 255 registers,
// one -> two wavefronts per SIMD), scatter 4.23 -> 4.90. At 96-256 and 512
// bytes the planes changed the gathers by -4% to +1%, at 768 bytes by -7%,
// and the warp-contiguous copies by -3% to +2%, so they keep the full width.
template <int m, int W>
__host__ __device__ constexpr bool tx_by_plane() {
    return W == 4 && m * W > 80 && m * W <= 112;
}

// Square tiles (m == n, a power of two): the transposition is the plain
// n x n transpose of (lane, slot) -- column side lane a slot b holds word
// a + n*b, which the row side wants in lane b slot a -- and is its own
// inverse. It runs as log2(n) butterfly stages without LDS or runtime
// rotations: stage d (= 1, 2, 4, ... n/2) exchanges, for every slot pair
// (b, b + d) with bit d of b clear, the slot b of lanes with bit d set against
// the slot b + d of their partner lane l ^ d. Lane distance 32 is one
// v_permlane32_swap per slot pair and 16 one v_permlane16_swap; 8 and 4 are
// one DPP move per slot (row_ror, writing only the lanes of the DPP banks --
// groups of 4 lanes -- that take the partner's value); 2 and 1 fall inside a
// bank and take a quad_perm DPP move plus a select.
template <int d>
__device__ __forceinline__ void butterfly(uint32_t& a, uint32_t& b, bool bit) {
    // After: a(l) = bit ? b(l ^ d) : a(l);  b(l) = bit ? b(l) : a(l ^ d).
    if constexpr (d == 32) {
        auto r = __builtin_amdgcn_permlane32_swap(a, b, false, false);
        a = r[0];
        b = r[1];
    } else if constexpr (d == 16) {
        auto r = __builtin_amdgcn_permlane16_swap(a, b, false, false);
        a = r[0];
        b = r[1];
    } else if constexpr (d == 8) {
        // row_ror:8 reads lane l ^ 8 of the 16-lane row; banks 2-3 (lanes
        // 8-15 of a row) have bit 3 set, banks 0-1 clear.
        const uint32_t na = (uint32_t)__builtin_amdgcn_update_dpp((int)a, (int)b, 0x128, 0xF, 0xC, false);
        const uint32_t nb = (uint32_t)__builtin_amdgcn_update_dpp((int)b, (int)a, 0x128, 0xF, 0x3, false);
        a = na;
        b = nb;
    } else if constexpr (d == 4) {
        // Banks 1 and 3 (bit 2 set) read lane l - 4 (row_ror:4), banks 0 and
        // 2 lane l + 4 (row_ror:12); neither wraps around its row.
        const uint32_t na = (uint32_t)__builtin_amdgcn_update_dpp((int)a, (int)b, 0x124, 0xF, 0xA, false);
        const uint32_t nb = (uint32_t)__builtin_amdgcn_update_dpp((int)b, (int)a, 0x12C, 0xF, 0x5, false);
        a = na;
        b = nb;
    } else {
        static_assert(d == 1 || d == 2, "lane distances are powers of two below 64");
        constexpr int quad = d == 1 ? 0xB1 : 0x4E;  // quad_perm [1,0,3,2] / [2,3,0,1]: lane l ^ d
        const uint32_t b_from = (uint32_t)__builtin_amdgcn_mov_dpp((int)b, quad, 0xF, 0xF, false);
        const uint32_t a_from = (uint32_t)__builtin_amdgcn_mov_dpp((int)a, quad, 0xF, 0xF, false);
        a = bit ? b_from : a;
        b = bit ? b : a_from;
    }
}

template <int n, int W, int d = 1>
__device__ __forceinline__ void square_transpose(uint32_t (&v)[n * W], int lane) {
    if constexpr (d < n) {
        const bool bit = (lane & d) != 0;
#pragma unroll
        for (int b = 0; b < n; ++b) {
            if (b & d) continue;
#pragma unroll
            for (int x = 0; x < W; ++x) butterfly<d>(v[b * W + x], v[(b + d) * W + x], bit);
        }
        square_transpose<n, W, 2 * d>(v, lane);
    }
}

template <int n, int m>
__device__ __forceinline__ void partner_table(int id, int (&partner)[m]) {
#pragma unroll
    for (int r = 0; r < m; ++r) partner[r] = tx_schedule<n, m>::partner(id, r);
}

// Self-contained forms: derive the per-lane tables from the tile rank.
// Square tiles take the butterfly transpose; every other shape the
// partner schedule.
template <int n, int m, int W, bool ByPlane = false>
__device__ __forceinline__ void r2c_words(uint32_t (&v)[m * W], int id, int first) {
    if constexpr (m == n) {
        square_transpose<n, W>(v, first + id);
    } else {
        int partner[m];
        partner_table<n, m>(id, partner);
        r2c_words<n, m, W, ByPlane>(v, id, first, partner, tx_schedule<n, m>::unrotate_lane(id));
    }
}

template <int n, int m, int W, bool ByPlane = false>
__device__ __forceinline__ void c2r_words(uint32_t (&v)[m * W], int id, int first) {
    if constexpr (m == n) {
        square_transpose<n, W>(v, first + id);
    } else {
        int partner[m];
        partner_table<n, m>(id, partner);
        c2r_words<n, m, W, ByPlane>(v, id, first, partner, id % m);
    }
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
