// Runtime rotation of a register array by a per-lane amount.
//
// rotate(t, a)[j] = t[(j + a) mod m], the contract of trove/rotate.h:111-115.
// The amount differs between lanes, so registers cannot be renamed: the
// rotation is a barrel shifter of ceil(log2 m) stages, each stage a set of
// per-lane selects (v_cndmask_b32) taken when the matching bit of `a` is set.
// `a` must lie in [0, 2^ceil(log2 m)); the result is the rotation by a mod m.
#pragma once
#include <trove/array.h>
#include <cstdint>

namespace trove {
namespace detail {

// All-ones when bit `bit` of `a` is set, else zero, opaque to the optimizer.
// A stage of the barrel shifter is then a bitwise select (one v_bfi_b32 per
// dword). Written as `c ? x : y` on array elements instead, LLVM folds the
// selects into loads at a selected index before the loops are unrolled, the
// register array becomes a dynamically indexed vector, and every access turns
// into a compare-and-select chain over all m positions (about 60x more VALU
// work at m = 15).
__host__ __device__ __forceinline__ uint32_t stage_mask(int a, int bit) {
    uint32_t msk = (a & bit) ? ~0u : 0u;
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+v"(msk));
#endif
    return msk;
}

// Rotate m elements of W dwords each, stored flat in v, left by amt.
template <int m, int W>
__host__ __device__ __forceinline__ void rotate_words_left(uint32_t (&v)[m * W], int amt) {
#pragma unroll
    for (int bit = 1; bit < m; bit <<= 1) {
        const uint32_t take = stage_mask(amt, bit);
        uint32_t t[m * W];
#pragma unroll
        for (int j = 0; j < m; ++j) {
            const int src = (j + bit) % m;
#pragma unroll
            for (int x = 0; x < W; ++x) t[j * W + x] = (v[src * W + x] & take) | (v[j * W + x] & ~take);
        }
#pragma unroll
        for (int i = 0; i < m * W; ++i) v[i] = t[i];
    }
}

}  // namespace detail

template <typename T, int i>
__host__ __device__ __forceinline__ array<T, i> rotate(const array<T, i>& t, int a) {
    static_assert(sizeof(array<T, i>) == i * sizeof(T), "array must be densely packed");
    if constexpr (i == 1 || sizeof(T) % 4 != 0) {
        // Generic element type: select element-wise.
        T c[i], r[i];
        make_carray<T, i>(t, c);
        for (int j = 0; j < i; ++j) r[j] = c[(j + a) % i];
        return make_array<T, i>(r);
    } else {
        constexpr int W = sizeof(T) / 4;
        uint32_t v[i * W];
        __builtin_memcpy(v, &t, sizeof(v));
        detail::rotate_words_left<i, W>(v, a);
        array<T, i> r;
        __builtin_memcpy(&r, v, sizeof(v));
        return r;
    }
}

}  // namespace trove
