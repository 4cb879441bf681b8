// Array-of-structures access: the low-level warp-contiguous path and the
// dynamic per-lane-address path with its tile cascade.
//
// Public interface of trove/aos.h:66-105 and 258-290:
//   load_warp_contiguous<tile>(p) / store_warp_contiguous<tile>(v, p):
//     every lane of a converged tile passes p = tile base + its rank. The tile
//     moves its tile*sizeof(T) bytes with m coalesced word accesses per lane
//     (word j of lane l is tile word l + j*tile), then R2C / C2R puts struct
//     l in lane l.
//   load(p) / store(v, p): arbitrary per-lane addresses. The widest converged
//     tile (64, 32, ..., 2 lanes) is used; lanes in no converged pair fall
//     back to per-lane word accesses.
// Eligibility is the reference's (aos.h:40-61): sizeof(T) % 4 == 0 and
// 1 < m <= tile; everything else is a per-lane access.
//
// Dynamic path, per tile (aos.h:110-219 for the reference's version):
//   1. p0 = the address of the tile's rank-0 lane (readfirstlane for a whole
//      wavefront, one 64-bit bpermute for a smaller tile).
//   2. If every lane holds p0 + rank, the access is warp-contiguous and takes
//      the coalesced path (the reference has this check commented out,
//      aos.h:205-207, 234-236).
//   3. Otherwise word j of lane l belongs to tile struct s = (l + j*tile) / m
//      at word (l + j*tile) % m; its address comes from lane s. When every
//      lane's offset from p0 fits in 32 bits (the usual case: one array) the
//      offset is broadcast with one ds_bpermute instead of a 64-bit pointer
//      (two). The broadcasts of a chunk of words are issued before their
//      accesses, and broadcast pointers keep the caller's pointer provenance
//      (global memory accesses when the compiler knows the array is global).
#pragma once
#include <trove/array.h>
#include <trove/memory.h>
#include <trove/transpose.h>
#include <trove/utility.h>
#include <trove/warp.h>
#include <trove/detail/dismember.h>
#include <trove/detail/fallback.h>

namespace trove {

namespace detail {

template <typename T, std::size_t tile_size>
struct size_in_range {
    typedef typename dismember_type<T>::type U;
    static const int size = aliased_size<T, U>::value;
    static const bool value = (size > 1) && (size <= (int)tile_size);
};

template <typename T, std::size_t tile_size>
struct use_shfl {
    static const bool value = size_multiple_power_of_two<T, 2>::value && size_in_range<T, tile_size>::value;
};

template <typename T, std::size_t tile_size>
struct use_direct {
    static const bool value = !use_shfl<T, tile_size>::value;
};

template <typename T>
struct word_layout {
    typedef typename dismember_type<T>::type U;
    static constexpr int m = aliased_size<T, U>::value;
    static constexpr int W = sizeof(U) / 4;
};

// Coalesced loads of a tile's words: word j of rank id is base[id + j*n].
template <int n, typename T, typename Policy = cached>
__device__ __forceinline__ void tile_load_words(const T* tile_base, int id,
                                                uint32_t (&v)[word_layout<T>::m * word_layout<T>::W]) {
    typedef word_layout<T> L;
    const typename L::U* base = reinterpret_cast<const typename L::U*>(tile_base);
#pragma unroll
    for (int j = 0; j < L::m; ++j) {
        const typename L::U u = load_word<Policy>(base + id + j * n);
        __builtin_memcpy(&v[j * L::W], &u, sizeof(u));
    }
}

template <int n, typename T, typename Policy = cached>
__device__ __forceinline__ void tile_store_words(T* tile_base, int id,
                                                 const uint32_t (&v)[word_layout<T>::m * word_layout<T>::W]) {
    typedef word_layout<T> L;
    typename L::U* base = reinterpret_cast<typename L::U*>(tile_base);
#pragma unroll
    for (int j = 0; j < L::m; ++j) {
        typename L::U u;
        __builtin_memcpy(&u, &v[j * L::W], sizeof(u));
        store_word<Policy>(base + id + j * n, u);
    }
}

// A pointer received from another lane, expressed as the caller's own
// pointer plus the byte distance to it. The address is the same; the
// pointer keeps `own`'s provenance, so when the compiler knows `own` to be
// global memory (a kernel argument) the broadcast pointer's accesses are
// global_load / global_store rather than flat accesses, whose completion
// every later LDS wait (lgkmcnt) would also have to cover.
template <typename P>
__device__ __forceinline__ P* rebase(P* own, uintptr_t target) {
    typedef typename std::conditional<std::is_const<P>::value, const char, char>::type byte;
    return reinterpret_cast<P*>(reinterpret_cast<byte*>(own) + (target - reinterpret_cast<uintptr_t>(own)));
}

// Address of the tile's rank-0 lane, seen by every lane of the tile.
template <int n, typename P>
__device__ __forceinline__ P* tile_first_pointer(P* p, int first) {
    const uintptr_t a = reinterpret_cast<uintptr_t>(p);
    if constexpr (n == 64) {
        return rebase(p, broadcast_first(a));
    } else {
        return rebase(p, bpermute(a, first << 2));
    }
}

// The dynamic path moves a tile's words in chunks: first every word's owner
// offset is broadcast (ds_bpermute), then every word's access is issued, so
// that the broadcasts' latencies overlap instead of each gating one access.
// A chunk keeps at most kWordChunk broadcasts in flight (the LDS counter
// tracks 15 outstanding operations).
constexpr int kWordChunk = 8;

// Word j of rank id in a gathered tile: struct s = (id + j*n) / m, word u.
// Walks (s, u) incrementally over j.
template <int n, int m>
struct word_walk {
    int s, u;
    __host__ __device__ __forceinline__ explicit word_walk(int id) : s(id / m), u(id % m) {}
    __host__ __device__ __forceinline__ void next() {
        u += n % m;
        if (u >= m) {
            u -= m;
            s += 1;
        }
        s += n / m;
    }
};

// Loads the words of a tile whose structs sit at arbitrary per-lane
// addresses, in the column-side layout (word j of rank id = tile word
// id + j*n of the concatenated structs).
// Word j of rank id belongs to the struct of rank owner[j] (tile-relative)
// at word word[j]: the word_walk sequence as arrays.
template <int n, int m>
__host__ __device__ __forceinline__ void word_owners(int id, int (&owner)[m], int (&word)[m]) {
    word_walk<n, m> w(id);
#pragma unroll
    for (int j = 0; j < m; ++j) {
        owner[j] = w.s;
        word[j] = w.u;
        w.next();
    }
}

// Word access j goes to base + off[j] (in structs), word word[j]; off[j] is
// the 32-bit value `key` of lane first + owner[j]. Broadcasts first, then
// accesses, chunk by chunk.
template <int n, typename T, bool kStore, typename Policy = cached, typename V>
__device__ __forceinline__ void tile_words_by_key(T* base, int key, int id, int first, V* v) {
    typedef word_layout<T> L;
    typedef typename std::conditional<std::is_const<T>::value, const typename L::U, typename L::U>::type U;
    int owner[L::m], word[L::m];
    word_owners<n, L::m>(id, owner, word);
#pragma unroll
    for (int j0 = 0; j0 < L::m; j0 += kWordChunk) {
        constexpr int c = L::m < kWordChunk ? L::m : kWordChunk;
        int off[c];
#pragma unroll
        for (int j = 0; j < c; ++j)
            if (j0 + j < L::m) off[j] = __builtin_amdgcn_ds_bpermute((first + owner[j0 + j]) << 2, key);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < c; ++j) {
            if (j0 + j >= L::m) continue;
            U* p = reinterpret_cast<U*>(base + off[j]) + word[j0 + j];
            if constexpr (kStore) {
                typename L::U u;
                __builtin_memcpy(&u, &v[(j0 + j) * L::W], sizeof(u));
                store_word<Policy>(p, u);
            } else {
                const typename L::U u = load_word<Policy>(p);
                __builtin_memcpy(&v[(j0 + j) * L::W], &u, sizeof(u));
            }
        }
    }
}

// Same with full per-lane pointers (owner offsets beyond 32 bits): each word
// receives its owner's pointer.
template <int n, typename T, bool kStore, typename Policy = cached, typename V>
__device__ __forceinline__ void tile_words_by_pointer(T* own, int id, int first, V* v) {
    typedef word_layout<T> L;
    typedef typename std::conditional<std::is_const<T>::value, const typename L::U, typename L::U>::type U;
    int owner[L::m], word[L::m];
    word_owners<n, L::m>(id, owner, word);
    const uintptr_t a = reinterpret_cast<uintptr_t>(own);
#pragma unroll
    for (int j0 = 0; j0 < L::m; j0 += kWordChunk / 2) {
        constexpr int c = L::m < kWordChunk / 2 ? L::m : kWordChunk / 2;
        uintptr_t q[c];
#pragma unroll
        for (int j = 0; j < c; ++j)
            if (j0 + j < L::m) q[j] = bpermute(a, (first + owner[j0 + j]) << 2);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < c; ++j) {
            if (j0 + j >= L::m) continue;
            U* p = reinterpret_cast<U*>(rebase(own, q[j])) + word[j0 + j];
            if constexpr (kStore) {
                typename L::U u;
                __builtin_memcpy(&u, &v[(j0 + j) * L::W], sizeof(u));
                store_word<Policy>(p, u);
            } else {
                const typename L::U u = load_word<Policy>(p);
                __builtin_memcpy(&v[(j0 + j) * L::W], &u, sizeof(u));
            }
        }
    }
}

// The lane's address relative to its tile's rank-0 address, in structs.
// `exact` is false when the byte distance is not a whole number of structs
// (lanes pointing into different arrays); such a tile never takes the
// struct-offset shortcuts. Integer arithmetic on purpose: a C++ pointer
// difference is an exact division, which lets the compiler assume the
// distance is always whole.
struct tile_offset {
    long long delta;
    bool exact;
};

template <typename T>
__device__ __forceinline__ tile_offset offset_from(const T* p, const T* p0) {
    const long long bytes = (long long)(reinterpret_cast<uintptr_t>(p) - reinterpret_cast<uintptr_t>(p0));
    const long long delta = bytes / (long long)sizeof(T);
    return {delta, delta * (long long)sizeof(T) == bytes};
}

// Loads the words of a tile whose structs sit at arbitrary per-lane
// addresses, in the column-side layout (word j of rank id = tile word
// id + j*n of the concatenated structs).
template <int n, typename Policy = cached, typename T>
__device__ __forceinline__ void tile_gather_words(const T* src, const T* p0, tile_offset o, int id, int first,
                                                  uint32_t (&v)[word_layout<T>::m * word_layout<T>::W]) {
    if (tile_all<n>(o.exact && o.delta == (long long)(int)o.delta)) {
        tile_words_by_key<n, const T, false, Policy>(p0, (int)o.delta, id, first, v);
    } else {
        tile_words_by_pointer<n, const T, false, Policy>(src, id, first, v);
    }
}

template <int n, typename Policy = cached, typename T>
__device__ __forceinline__ void tile_scatter_words(T* dst, T* p0, tile_offset o, int id, int first,
                                                   const uint32_t (&v)[word_layout<T>::m * word_layout<T>::W]) {
    if (tile_all<n>(o.exact && o.delta == (long long)(int)o.delta)) {
        tile_words_by_key<n, T, true, Policy>(p0, (int)o.delta, id, first, v);
    } else {
        tile_words_by_pointer<n, T, true, Policy>(dst, id, first, v);
    }
}

// Tile gather from one array: rank id wants base[index]. Produces the
// column-side layout of the gathered structs (what a coalesced tile store
// expects), broadcasting the 32-bit index of each word's owner.
template <int n, typename Policy = cached, typename T>
__device__ __forceinline__ void tile_gather_indexed(const T* base, int index, int id, int first,
                                                    uint32_t (&v)[word_layout<T>::m * word_layout<T>::W]) {
    tile_words_by_key<n, const T, false, Policy>(base, index, id, first, v);
}

// Tile scatter into one array: the structs held in column-side layout go to
// base[index of their owner rank].
template <int n, typename Policy = cached, typename T>
__device__ __forceinline__ void tile_scatter_indexed(T* base, int index, int id, int first,
                                                     const uint32_t (&v)[word_layout<T>::m * word_layout<T>::W]) {
    tile_words_by_key<n, T, true, Policy>(base, index, id, first, v);
}

}  // namespace detail

// Policy is trove::cached (default: ordinary accesses, as in the reference)
// or trove::streaming (nontemporal accesses for data touched once).
template <std::size_t tile_size = WARP_SIZE, typename Policy = cached, typename T>
__device__ __forceinline__ typename enable_if<detail::use_shfl<T, tile_size>::value, T>::type load_warp_contiguous(const T* src) {
    constexpr int n = (int)tile_size;
    typedef detail::word_layout<T> L;
    const int lane = lane_id();
    const int id = lane & (n - 1);
    uint32_t v[L::m * L::W];
    detail::tile_load_words<n, T, Policy>(src - id, id, v);
    detail::r2c_words<n, L::m, L::W>(v, id, lane & ~(n - 1));
    T out;
    detail::words_to_struct(v, out);
    return out;
}

template <std::size_t tile_size = WARP_SIZE, typename Policy = cached, typename T>
__device__ __forceinline__ typename enable_if<detail::use_direct<T, tile_size>::value, T>::type load_warp_contiguous(const T* src) {
    if constexpr (detail::use_divergent<T>::value) {
        typedef detail::word_layout<T> L;
        const typename L::U* p = reinterpret_cast<const typename L::U*>(src);
        typename L::U w[L::m];
#pragma unroll
        for (int j = 0; j < L::m; ++j) w[j] = detail::load_word<Policy>(p + j);
        T out;
        detail::words_to_struct(w, out);
        return out;
    } else {
        return detail::divergent_load(src);
    }
}

template <std::size_t tile_size = WARP_SIZE, typename Policy = cached, typename T>
__device__ __forceinline__ typename enable_if<detail::use_shfl<T, tile_size>::value>::type store_warp_contiguous(const T& data,
                                                                                                 T* dest) {
    constexpr int n = (int)tile_size;
    typedef detail::word_layout<T> L;
    const int lane = lane_id();
    const int id = lane & (n - 1);
    uint32_t v[L::m * L::W];
    detail::struct_to_words(data, v);
    detail::c2r_words<n, L::m, L::W>(v, id, lane & ~(n - 1));
    detail::tile_store_words<n, T, Policy>(dest - id, id, v);
}

template <std::size_t tile_size = WARP_SIZE, typename Policy = cached, typename T>
__device__ __forceinline__ typename enable_if<detail::use_direct<T, tile_size>::value>::type store_warp_contiguous(const T& data,
                                                                                                   T* dest) {
    if constexpr (detail::use_divergent<T>::value) {
        typedef detail::word_layout<T> L;
        typename L::U w[L::m];
        detail::struct_to_words(data, w);
        typename L::U* p = reinterpret_cast<typename L::U*>(dest);
#pragma unroll
        for (int j = 0; j < L::m; ++j) detail::store_word<Policy>(p + j, w[j]);
    } else {
        detail::divergent_store(data, dest);
    }
}

namespace detail {

template <typename Policy = cached, typename T, typename Tile>
__device__ __forceinline__ typename enable_if<use_shfl<T, Tile::size()>::value, T>::type load_dispatch(const T* src,
                                                                                       const Tile& tile) {
    constexpr int n = Tile::size();
    typedef word_layout<T> L;
    const int id = tile.id();
    const int first = tile.first_lane();
    const T* p0 = tile_first_pointer<n>(src, first);
    const tile_offset o = offset_from(src, p0);
    uint32_t v[L::m * L::W];
    if (tile_all<n>(o.exact && o.delta == id)) {
        tile_load_words<n, T, Policy>(p0, id, v);
    } else {
        tile_gather_words<n, Policy>(src, p0, o, id, first, v);
    }
    r2c_words<n, L::m, L::W, tx_by_plane<L::m, L::W>()>(v, id, first);
    T out;
    words_to_struct(v, out);
    return out;
}

template <typename Policy = cached, typename T, typename Tile>
__device__ __forceinline__ typename enable_if<use_direct<T, Tile::size()>::value, T>::type load_dispatch(const T* src, const Tile&) {
    return divergent_load<Policy>(src);
}

template <typename Policy = cached, typename T, typename Tile>
__device__ __forceinline__ typename enable_if<use_shfl<T, Tile::size()>::value>::type store_dispatch(const T& data, T* dest,
                                                                                     const Tile& tile) {
    constexpr int n = Tile::size();
    typedef word_layout<T> L;
    const int id = tile.id();
    const int first = tile.first_lane();
    T* p0 = tile_first_pointer<n>(dest, first);
    const tile_offset o = offset_from(dest, p0);
    uint32_t v[L::m * L::W];
    struct_to_words(data, v);
    c2r_words<n, L::m, L::W, tx_by_plane<L::m, L::W>()>(v, id, first);
    if (tile_all<n>(o.exact && o.delta == id)) {
        tile_store_words<n, T, Policy>(p0, id, v);
    } else {
        tile_scatter_words<n, Policy>(dest, p0, o, id, first, v);
    }
}

template <typename Policy = cached, typename T, typename Tile>
__device__ __forceinline__ typename enable_if<use_direct<T, Tile::size()>::value>::type store_dispatch(const T& data, T* dest,
                                                                                       const Tile&) {
    divergent_store<Policy>(data, dest);
}

// *dst = *src for the structs of one converged tile, without the struct ever
// sitting in one lane: the tile's source words are loaded in the column-side
// layout (word j of rank id = word (id + j*n) mod m of the struct of rank
// (id + j*n) / m) and stored in the same layout to the destination structs
// of the same ranks, so neither R2C nor C2R runs. Every load of the tile
// completes before its first store, as when the struct passes through
// registers, so overlapping source and destination read the old values.
template <typename LoadP, typename StoreP, typename T, typename Tile>
__device__ __forceinline__ typename enable_if<use_shfl<T, Tile::size()>::value>::type copy_dispatch(const T* src, T* dst,
                                                                                    const Tile& tile) {
    constexpr int n = Tile::size();
    typedef word_layout<T> L;
    const int id = tile.id();
    const int first = tile.first_lane();
    uint32_t v[L::m * L::W];
    const T* s0 = tile_first_pointer<n>(src, first);
    const tile_offset os = offset_from(src, s0);
    if (tile_all<n>(os.exact && os.delta == id)) {
        tile_load_words<n, T, LoadP>(s0, id, v);
    } else {
        tile_gather_words<n, LoadP>(src, s0, os, id, first, v);
    }
    // A lane stores words other lanes of the tile loaded from the same
    // addresses when source and destination overlap: wait for every load.
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    T* d0 = tile_first_pointer<n>(dst, first);
    const tile_offset od = offset_from(dst, d0);
    if (tile_all<n>(od.exact && od.delta == id)) {
        tile_store_words<n, T, StoreP>(d0, id, v);
    } else {
        tile_scatter_words<n, StoreP>(dst, d0, od, id, first, v);
    }
}

template <typename LoadP, typename StoreP, typename T, typename Tile>
__device__ __forceinline__ typename enable_if<use_direct<T, Tile::size()>::value>::type copy_dispatch(const T* src, T* dst,
                                                                                      const Tile&) {
    divergent_store<StoreP>(divergent_load<LoadP>(src), dst);
}

// *dst = *src with the widest converged tile around the lane (the tile
// cascade of load / store, decided once for both sides): what assigning one
// coalesced_ref to another does.
template <typename LoadP, typename StoreP, typename T>
__device__ __forceinline__ void copy_cascade(const T* src, T* dst) {
    if (warp_converged<WARP_SIZE>()) {
        copy_dispatch<LoadP, StoreP>(src, dst, thread_tile<WARP_SIZE>());
    } else if (warp_converged<WARP_SIZE / 2>()) {
        copy_dispatch<LoadP, StoreP>(src, dst, thread_tile<WARP_SIZE / 2>());
    } else if (warp_converged<WARP_SIZE / 4>()) {
        copy_dispatch<LoadP, StoreP>(src, dst, thread_tile<WARP_SIZE / 4>());
    } else if (warp_converged<WARP_SIZE / 8>()) {
        copy_dispatch<LoadP, StoreP>(src, dst, thread_tile<WARP_SIZE / 8>());
    } else if (warp_converged<WARP_SIZE / 16>()) {
        copy_dispatch<LoadP, StoreP>(src, dst, thread_tile<WARP_SIZE / 16>());
    } else if (warp_converged<WARP_SIZE / 32>()) {
        copy_dispatch<LoadP, StoreP>(src, dst, thread_tile<WARP_SIZE / 32>());
    } else {
        divergent_store<StoreP>(divergent_load<LoadP>(src), dst);
    }
}

}  // namespace detail

// Load *src cooperatively with the widest converged tile around the lane.
// Policy (an addition, defaulted to the reference's behaviour) is the cache
// policy of the global accesses: trove::cached or trove::streaming.
template <typename T, typename Policy = cached>
__device__ __forceinline__ T load(const T* src) {
    if (warp_converged<WARP_SIZE>()) {
        return detail::load_dispatch<Policy>(src, thread_tile<WARP_SIZE>());
    } else if (warp_converged<WARP_SIZE / 2>()) {
        return detail::load_dispatch<Policy>(src, thread_tile<WARP_SIZE / 2>());
    } else if (warp_converged<WARP_SIZE / 4>()) {
        return detail::load_dispatch<Policy>(src, thread_tile<WARP_SIZE / 4>());
    } else if (warp_converged<WARP_SIZE / 8>()) {
        return detail::load_dispatch<Policy>(src, thread_tile<WARP_SIZE / 8>());
    } else if (warp_converged<WARP_SIZE / 16>()) {
        return detail::load_dispatch<Policy>(src, thread_tile<WARP_SIZE / 16>());
    } else if (warp_converged<WARP_SIZE / 32>()) {
        return detail::load_dispatch<Policy>(src, thread_tile<WARP_SIZE / 32>());
    } else {
        return detail::divergent_load<Policy>(src);
    }
}

// Store data to *dest cooperatively with the widest converged tile.
template <typename T, typename Policy = cached>
__device__ __forceinline__ void store(const T& data, T* dest) {
    if (warp_converged<WARP_SIZE>()) {
        detail::store_dispatch<Policy>(data, dest, thread_tile<WARP_SIZE>());
    } else if (warp_converged<WARP_SIZE / 2>()) {
        detail::store_dispatch<Policy>(data, dest, thread_tile<WARP_SIZE / 2>());
    } else if (warp_converged<WARP_SIZE / 4>()) {
        detail::store_dispatch<Policy>(data, dest, thread_tile<WARP_SIZE / 4>());
    } else if (warp_converged<WARP_SIZE / 8>()) {
        detail::store_dispatch<Policy>(data, dest, thread_tile<WARP_SIZE / 8>());
    } else if (warp_converged<WARP_SIZE / 16>()) {
        detail::store_dispatch<Policy>(data, dest, thread_tile<WARP_SIZE / 16>());
    } else if (warp_converged<WARP_SIZE / 32>()) {
        detail::store_dispatch<Policy>(data, dest, thread_tile<WARP_SIZE / 32>());
    } else {
        detail::divergent_store<Policy>(data, dest);
    }
}

}  // namespace trove
