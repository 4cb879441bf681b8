// LDS-staged warp-contiguous access: the alternative to the in-register
// transposition of load_warp_contiguous / store_warp_contiguous (aos.h).
//
// A wavefront's 64 structs are 64 * sizeof(T) contiguous bytes. Here they
// move between global memory and a wavefront-private LDS stage in CW-dword
// chunks (CW = 4, one global_load/store_dwordx4 = 1 KiB per wave instruction,
// when the tile base is 16-byte aligned; CW = 1 otherwise), and each lane
// reads or writes its own struct in the stage (dwords [lane * k, lane * k + k)
// with k = sizeof(T) / 4, as 4-, 8- or 16-byte words like the register path).
// No cross-lane shuffles and no rotations; the price is the stage
// (lds_stage_bytes<T>() = 64 * sizeof(T) bytes of LDS per wavefront) and two
// LDS round trips per direction. For occupancy, gfx950 allocates a
// work-group's LDS in 1280-byte units, 128 per CU (counted on MI355X), so a
// work-group's stages cost their total rounded up to a multiple of 1280.
//
// Structs of 16-byte words are stored swizzled: the m words of struct l are
// rotated by rot(l) = (l mod 16) / (16 / gcd(m, 16)) slots, so that the 16
// lanes of one LDS pass read or write 16 different 16-byte bank groups of
// gfx950's 64 four-byte banks (unswizzled, the lanes of a pass all hit the
// same group at 128- and 256-byte strides). The stage's contents are
// therefore only meaningful through these functions.
//
// The reference's transposition is register-only (transpose.h); these
// functions have no counterpart there. Preconditions are those of
// load_warp_contiguous<64>: all 64 lanes active, lane l passes tile base + l,
// sizeof(T) % 4 == 0, the struct's words aligned to their size; the stage is
// 16-byte aligned LDS that no other wavefront uses meanwhile.
#pragma once
#include <trove/aos.h>
#include <trove/static_gcd.h>

#include <cstddef>
#include <cstdint>
#include <type_traits>

namespace trove {

template <typename T>
constexpr std::size_t lds_stage_bytes() {
    return (std::size_t)WARP_SIZE * sizeof(T);
}

namespace detail {

// Orders a wavefront's LDS accesses across its lanes. A wavefront's LDS
// instructions execute in issue order, so this only has to keep the compiler
// from moving accesses across it (the fence pair rocprim's wave_barrier uses).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int W>
using dword_vec = typename std::conditional<W == 1, uint32_t, uint32_t __attribute__((ext_vector_type(W)))>::type;

// LDS view of a stage pointer (an address-space cast, so that the accesses
// are ds_* instructions whatever the caller's pointer type says).
template <typename V>
__device__ __forceinline__ __attribute__((address_space(3))) V* lds_view(uint32_t* p) {
    return (__attribute__((address_space(3))) V*)p;
}
template <typename V>
__device__ __forceinline__ const __attribute__((address_space(3))) V* lds_view(const uint32_t* p) {
    return (const __attribute__((address_space(3))) V*)p;
}

// LDS slot of the tile's U-words. Identity for 4- and 8-byte words; for
// 16-byte words (W = 4, m > 1) word j of struct l sits at l * m +
// (j + rot(l)) mod m, rot(l) = (l mod 16) >> log2(16 / gcd(m, 16)) < m.
template <typename T>
struct lds_swizzle {
    typedef word_layout<T> L;
    static constexpr bool on = L::W == 4 && L::m > 1;
    static constexpr int g = on ? static_gcd<L::m, 16>::value : 1;
    static constexpr int shift = g >= 16 ? 0 : g >= 8 ? 1 : g >= 4 ? 2 : g >= 2 ? 3 : 4;
    static __host__ __device__ __forceinline__ int rot(int l) { return (l & 15) >> shift; }
    // Slot of word j (0 <= j < m) of struct l.
    static __host__ __device__ __forceinline__ int of(int l, int j) {
        if constexpr (!on) {
            return l * L::m + j;
        } else {
            const int w = j + rot(l);
            return l * L::m + (w < L::m ? w : w - L::m);
        }
    }
    // Slot of tile U-word u (struct u / m, word u % m).
    static __host__ __device__ __forceinline__ int of(int u) {
        if constexpr (!on) {
            return u;
        } else {
            return of(u / L::m, u % L::m);
        }
    }
};

// A tile's bytes as CW-dword chunks in registers: chunk c = lane + 64 r
// (r < rounds) is tile dwords [CW c, CW c + CW). The last round is partial
// when 64 * k / CW is not a multiple of 64.
template <typename T, int CW>
struct lds_chunks {
    static_assert(sizeof(T) % 4 == 0, "LDS staging needs a whole number of dwords per struct");
    static_assert(CW == 1 || CW == 2 || CW == 4, "chunk width is 1, 2 or 4 dwords");
    static constexpr int dwords = WARP_SIZE * (int)(sizeof(T) / 4);
    static constexpr int count = dwords / CW;
    static constexpr int rounds = (count + WARP_SIZE - 1) / WARP_SIZE;
    typedef dword_vec<CW> vec;
    vec c[rounds];
    static __device__ __forceinline__ bool has(int lane, int r) {
        return count % WARP_SIZE == 0 || lane + WARP_SIZE * r < count;
    }
};

template <int CW, typename Policy, typename T>
__device__ __forceinline__ void lds_global_load(const T* tile_base, int lane, lds_chunks<T, CW>& t) {
    typedef lds_chunks<T, CW> C;
    const typename C::vec* p = reinterpret_cast<const typename C::vec*>(tile_base);
#pragma unroll
    for (int r = 0; r < C::rounds; ++r)
        if (C::has(lane, r)) t.c[r] = load_word<Policy>(p + lane + WARP_SIZE * r);
}

template <int CW, typename Policy, typename T>
__device__ __forceinline__ void lds_global_store(T* tile_base, int lane, const lds_chunks<T, CW>& t) {
    typedef lds_chunks<T, CW> C;
    typename C::vec* p = reinterpret_cast<typename C::vec*>(tile_base);
#pragma unroll
    for (int r = 0; r < C::rounds; ++r)
        if (C::has(lane, r)) store_word<Policy>(p + lane + WARP_SIZE * r, t.c[r]);
}

// Stage position of chunk c in units of the chunk: a 4-dword chunk of a
// 16-byte-word tile is one U-word, a dword chunk a quarter of one.
template <typename S, int CW>
__device__ __forceinline__ int lds_chunk_pos(int c) {
    if constexpr (!S::on) {
        return c;
    } else if constexpr (CW == 4) {
        return S::of(c);
    } else {
        static_assert(CW == 1, "16-byte-word tiles move in 4- or 1-dword chunks");
        return 4 * S::of(c >> 2) + (c & 3);
    }
}

template <int CW, typename T>
__device__ __forceinline__ void lds_put_chunks(uint32_t* stage, int lane, const lds_chunks<T, CW>& t) {
    typedef lds_chunks<T, CW> C;
    auto s = lds_view<typename C::vec>(stage);
#pragma unroll
    for (int r = 0; r < C::rounds; ++r)
        if (C::has(lane, r)) s[lds_chunk_pos<lds_swizzle<T>, CW>(lane + WARP_SIZE * r)] = t.c[r];
}

template <int CW, typename T>
__device__ __forceinline__ void lds_get_chunks(const uint32_t* stage, int lane, lds_chunks<T, CW>& t) {
    typedef lds_chunks<T, CW> C;
    auto s = lds_view<typename C::vec>(stage);
#pragma unroll
    for (int r = 0; r < C::rounds; ++r)
        if (C::has(lane, r)) t.c[r] = s[lds_chunk_pos<lds_swizzle<T>, CW>(lane + WARP_SIZE * r)];
}

// The lane's own struct in the stage, moved as the struct's words.
template <typename T>
__device__ __forceinline__ T lds_get_struct(const uint32_t* stage, int lane) {
    typedef word_layout<T> L;
    typedef dword_vec<L::W> V;
    static_assert(sizeof(V) == sizeof(typename L::U), "word size");
    auto s = lds_view<V>(stage);
    V w[L::m];
#pragma unroll
    for (int j = 0; j < L::m; ++j) w[j] = s[lds_swizzle<T>::of(lane, j)];
    T out;
    words_to_struct(w, out);
    return out;
}

template <typename T>
__device__ __forceinline__ void lds_put_struct(uint32_t* stage, int lane, const T& v) {
    typedef word_layout<T> L;
    typedef dword_vec<L::W> V;
    V w[L::m];
    struct_to_words(v, w);
    auto s = lds_view<V>(stage);
#pragma unroll
    for (int j = 0; j < L::m; ++j) s[lds_swizzle<T>::of(lane, j)] = w[j];
}

// Column-side words (word j of lane l = tile word l + 64 j, as the indexed
// tile accesses produce and take) to and from the stage.
template <typename T>
__device__ __forceinline__ void lds_put_columns(uint32_t* stage, int lane,
                                                const uint32_t (&v)[word_layout<T>::m * word_layout<T>::W]) {
    typedef word_layout<T> L;
    typedef dword_vec<L::W> V;
    auto s = lds_view<V>(stage);
#pragma unroll
    for (int j = 0; j < L::m; ++j) {
        V w;
        __builtin_memcpy(&w, &v[j * L::W], sizeof(V));
        s[lds_swizzle<T>::of(lane + WARP_SIZE * j)] = w;
    }
}

template <typename T>
__device__ __forceinline__ void lds_get_columns(const uint32_t* stage, int lane,
                                                uint32_t (&v)[word_layout<T>::m * word_layout<T>::W]) {
    typedef word_layout<T> L;
    typedef dword_vec<L::W> V;
    auto s = lds_view<V>(stage);
#pragma unroll
    for (int j = 0; j < L::m; ++j) {
        const V w = s[lds_swizzle<T>::of(lane + WARP_SIZE * j)];
        __builtin_memcpy(&v[j * L::W], &w, sizeof(V));
    }
}

// Whether a tile base takes dwordx4 chunks (wave-uniform).
template <typename T>
__device__ __forceinline__ bool lds_wide(const T* tile_base) {
    return (__builtin_amdgcn_readfirstlane((int)(reinterpret_cast<uintptr_t>(tile_base) & 15)) == 0);
}

}  // namespace detail

// Lane l gets struct l of the wavefront's tile (src = tile base + l).
template <typename Policy = cached, typename T>
__device__ __forceinline__ T load_warp_contiguous_lds(const T* src, uint32_t* stage) {
    const int lane = lane_id();
    const T* base = src - lane;
    detail::wave_lds_sync();
    if (detail::lds_wide(base)) {
        detail::lds_chunks<T, 4> t;
        detail::lds_global_load<4, Policy>(base, lane, t);
        detail::lds_put_chunks(stage, lane, t);
    } else {
        detail::lds_chunks<T, 1> t;
        detail::lds_global_load<1, Policy>(base, lane, t);
        detail::lds_put_chunks(stage, lane, t);
    }
    detail::wave_lds_sync();
    return detail::lds_get_struct<T>(stage, lane);
}

// Lane l's struct goes to tile base + l (dest = tile base + l).
template <typename Policy = cached, typename T>
__device__ __forceinline__ void store_warp_contiguous_lds(const T& data, T* dest, uint32_t* stage) {
    const int lane = lane_id();
    T* base = dest - lane;
    detail::wave_lds_sync();
    detail::lds_put_struct(stage, lane, data);
    detail::wave_lds_sync();
    if (detail::lds_wide(base)) {
        detail::lds_chunks<T, 4> t;
        detail::lds_get_chunks(stage, lane, t);
        detail::lds_global_store<4, Policy>(base, lane, t);
    } else {
        detail::lds_chunks<T, 1> t;
        detail::lds_get_chunks(stage, lane, t);
        detail::lds_global_store<1, Policy>(base, lane, t);
    }
}

}  // namespace trove
