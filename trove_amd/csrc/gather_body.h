// Per-wavefront gather bodies over a range of 64-struct groups, shared by the
// standalone gather kernels and the batch kernels.
//
// A wavefront calls one of these with the same [g0, g1) on every lane;
// group g covers structs [64 g, 64 g + 64) of a job of n structs, and the last
// group may be partial (n % 64 lanes).
#pragma once
#include "common.h"

namespace trove_amd {

// Index guards: every gather body passes each index it reads through one
// before using it as a row, and calls report() once at its end. no_guard
// trusts the caller (the C ABI's gathers). row_guard maps an index outside
// [0, last] to row 0 with a select (no branch between an index load and the
// next one: a branch there would wait for each index before loading the
// next, one group at a time) and remembers it per lane; report() then sets
// *flag (when given), so that a job with a bad index never reads outside
// its source. The flag is written with a system-scope store so that it may
// live in pinned host memory. A body's guard is its own copy.
struct no_guard {
    __device__ __forceinline__ int32_t operator()(int32_t ix) { return ix; }
    __device__ __forceinline__ void report() const {}
};

struct row_guard {
    uint32_t last;      // largest valid row (0xffffffff: every index valid)
    int32_t* flag;      // nullable
    bool bad = false;   // this lane met an index above last
    __device__ __forceinline__ int32_t operator()(int32_t ix) {
        const bool ok = (uint32_t)ix <= last;
        bad = bad || !ok;
        return ok ? ix : 0;
    }
    __device__ __forceinline__ void report() const {
        if (bad && flag) __hip_atomic_store(flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
};

// Cache policy of the random row reads of every gather path: ordinary loads
// up to 48 bytes, nontemporal loads above. Measured on MI355X (2^26 random
// rows, interleaved runs): nontemporal row loads cost 10-20% at 8-24 bytes,
// are even from 28 to 48 bytes, and gain 1-4% from 52 to 64 bytes and 2-17%
// from 96 to 256 bytes, on the tile path and through coalesced_ptr alike.
template <typename T>
using row_read_policy = typename std::conditional<(sizeof(T) > 48), trove::streaming, trove::cached>::type;

// 64-struct groups a wavefront keeps in flight on the tile path. Random rows
// cost whole 128-byte line fills, and the HBM serves random lines faster the
// more of them are outstanding (measured on MI355X: 3.8 TB/s of line reads
// at 8 lines per wavefront, 5.5 at 256); two groups per wavefront beat one by
// 3-9% from 36 to 256 bytes and match three or four from 20 to 32 bytes.
// Above 256 bytes two groups' words no longer fit the registers.
template <typename T>
constexpr int gather_groups_in_flight() {
    return sizeof(T) <= 16 ? 4 : sizeof(T) <= 256 ? 2 : 1;
}

// Tile path: indexed word loads in column-side layout (each word's owner
// index broadcast by ds_bpermute), then coalesced tile stores; no
// transposition. UNROLL groups are in flight per iteration. Structs of 4, 8
// or 16 bytes take gather_range_word.
// The index stream and the destination are touched once (streaming policy);
// the gathered rows follow row_read_policy.
// One-word structs (m = 1: 4, 8 and 16 bytes) have nothing to broadcast or
// transpose: each lane loads its row as one word. They keep the tile path's
// policies (streaming index loads and stores, UNROLL groups in flight).
template <typename T, int UNROLL, typename G = no_guard>
__device__ __forceinline__ void gather_range_word(const T* __restrict__ src, const int32_t* __restrict__ idx,
                                                  T* __restrict__ dst, int64_t n, int64_t g0, int64_t g1,
                                                  G guard = G()) {
    typedef typename trove::detail::word_layout<T>::U U;
    typedef trove::streaming S;
    const U* s = reinterpret_cast<const U*>(src);
    U* d = reinterpret_cast<U*>(dst);
    const int lane = trove::lane_id();
    const int64_t full = n >> 6;
    const int64_t full_end = g1 < full ? g1 : full;
    int64_t g = g0;
    for (; g + UNROLL <= full_end; g += UNROLL) {
        int32_t ix[UNROLL];
        U v[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) ix[u] = guard(trove::detail::load_word<S>(idx + ((g + u) << 6) + lane));
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) v[u] = trove::detail::load_word<trove::cached>(s + ix[u]);
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) trove::detail::store_word<S>(d + ((g + u) << 6) + lane, v[u]);
    }
    for (; g < full_end; ++g) {
        const int64_t i = (g << 6) + lane;
        trove::detail::store_word<S>(d + i, trove::detail::load_word<trove::cached>(s + guard(trove::detail::load_word<S>(idx + i))));
    }
    if (g1 > full && (n & 63)) {
        const int64_t i = (full << 6) + lane;
        if (i < n) d[i] = s[guard(idx[i])];
    }
    guard.report();
}

template <typename T, int UNROLL, typename G = no_guard>
__device__ __forceinline__ void gather_range_tile(const T* __restrict__ src, const int32_t* __restrict__ idx,
                                                  T* __restrict__ dst, int64_t n, int64_t g0, int64_t g1,
                                                  G guard = G()) {
    typedef trove::detail::word_layout<T> L;
    typedef trove::streaming S;
    if constexpr (L::m == 1) {
        gather_range_word<T, UNROLL>(src, idx, dst, n, g0, g1, guard);
        return;
    }
    const int lane = trove::lane_id();
    const int64_t full = n >> 6;
    const int64_t full_end = g1 < full ? g1 : full;
    int64_t g = g0;
    for (; g + UNROLL <= full_end; g += UNROLL) {
        int32_t ix[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) ix[u] = guard(trove::detail::load_word<S>(idx + ((g + u) << 6) + lane));
        uint32_t v[UNROLL][L::m * L::W];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) trove::detail::tile_gather_indexed<64, row_read_policy<T>>(src, ix[u], lane, 0, v[u]);
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) trove::detail::tile_store_words<64, T, S>(dst + ((g + u) << 6), lane, v[u]);
    }
    for (; g < full_end; ++g) {
        uint32_t v[L::m * L::W];
        trove::detail::tile_gather_indexed<64, row_read_policy<T>>(src, guard(trove::detail::load_word<S>(idx + (g << 6) + lane)),
                                                                   lane, 0, v);
        trove::detail::tile_store_words<64, T, S>(dst + (g << 6), lane, v);
    }
    if (g1 > full && (n & 63)) {
        const int64_t i = (full << 6) + lane;
        if (i < n) trove::detail::divergent_store(trove::detail::divergent_load(src + guard(idx[i])), dst + i);
    }
    guard.report();
}

// coalesced_ptr path (tests/benchmark.cu:90-99) for ONE group g: a partial
// last group runs the tile cascade with its missing lanes inactive. The
// destination is written once, so it is a streaming coalesced_ptr
// (nontemporal stores: +3-8% over cached ones on MI355X at 12-64 bytes); the
// rows are read with row_read_policy. There is deliberately no loop here:
// around the cascade, LICM would keep the tables of all six tile levels live
// at once.
template <typename T, typename G = no_guard>
__device__ __forceinline__ void gather_group_shfl(const T* __restrict__ src, const int32_t* __restrict__ idx,
                                                  T* __restrict__ dst, int64_t n, int64_t g, G guard = G()) {
    const int64_t i = (g << 6) + trove::lane_id();
    if (i < n) {
        trove::coalesced_ptr<T, row_read_policy<T>> s(const_cast<T*>(src));
        trove::coalesced_ptr<T, trove::streaming> r(dst);
        const T v = s[guard(idx[i])];
        r[i] = v;
    }
    guard.report();
}

// coalesced_ptr path at a fixed tile width TILE < 64: each converged
// TILE-lane tile of the wavefront moves its own structs (what the
// reference's coalesced_ptr does on a 32-lane warp); a tile with inactive
// lanes (in a partial last group) moves its structs per lane. The load and
// the store see the same active lanes, so one convergence check serves both,
// and each branch keeps its struct in registers (a struct returned through a
// branch landed in scratch memory at 512 bytes and above). LoadP / StoreP:
// the cache policies of the two sides.
template <int TILE, typename LoadP, typename StoreP, typename T>
__device__ __forceinline__ void move_at_tile(const T* from, T* to) {
    namespace d = trove::detail;
    if (trove::warp_converged<TILE>()) {
        const trove::thread_tile<TILE> tile;
        const T v = d::load_dispatch<LoadP>(from, tile);
        d::store_dispatch<StoreP>(v, to, tile);
    } else {
        d::divergent_store<StoreP>(d::divergent_load<LoadP>(from), to);
    }
}

template <typename T, int TILE>
__device__ __forceinline__ void gather_group_shfl_tile(const T* __restrict__ src, const int32_t* __restrict__ idx,
                                                       T* __restrict__ dst, int64_t n, int64_t g) {
    const int64_t i = (g << 6) + trove::lane_id();
    if (i < n) move_at_tile<TILE, row_read_policy<T>, trove::streaming>(src + idx[i], dst + i);
}

// LDS-staged path (trove/lds.h) for ONE full group g, the LDS counterpart of
// the coalesced_ptr path: the tile path's indexed loads (column-side words)
// go through the stage so that lane l holds struct l, as a coalesced_ptr read
// delivers it, then back through the stage to CW-dword chunk stores. stage:
// 64 * sizeof(T) bytes of LDS private to the wavefront. One-word structs
// (m = 1) load their row per lane and only stage the store.
template <typename T, int CW, typename G = no_guard>
__device__ __forceinline__ void gather_group_lds(const T* __restrict__ src, const int32_t* __restrict__ idx,
                                                 T* __restrict__ dst, int64_t g, uint32_t* stage, G guard = G()) {
    namespace d = trove::detail;
    typedef d::word_layout<T> L;
    const int lane = trove::lane_id();
    const int32_t ix = guard(d::load_word<trove::streaming>(idx + (g << 6) + lane));
    T v;
    if constexpr (L::m == 1) {
        v = d::divergent_load(src + ix);
    } else {
        uint32_t w[L::m * L::W];
        d::tile_gather_indexed<64, row_read_policy<T>>(src, ix, lane, 0, w);
        d::wave_lds_sync();
        d::lds_put_columns<T>(stage, lane, w);
        d::wave_lds_sync();
        v = d::lds_get_struct<T>(stage, lane);
    }
    d::wave_lds_sync();
    d::lds_put_struct(stage, lane, v);
    d::wave_lds_sync();
    d::lds_chunks<T, CW> t;
    d::lds_get_chunks(stage, lane, t);
    d::lds_global_store<CW, trove::streaming>(dst + (g << 6), lane, t);
    guard.report();
}

// LDS-staged gather over groups [g0, g1); the partial last group per lane.
template <typename T, typename G = no_guard>
__device__ __forceinline__ void gather_range_lds(const T* __restrict__ src, const int32_t* __restrict__ idx,
                                                 T* __restrict__ dst, int64_t n, int64_t g0, int64_t g1,
                                                 uint32_t* stage, G guard = G()) {
    const int64_t full = n >> 6;
    const int64_t full_end = g1 < full ? g1 : full;
    if (trove::detail::lds_wide(dst)) {
        for (int64_t g = g0; g < full_end; ++g) gather_group_lds<T, 4>(src, idx, dst, g, stage, guard);
    } else {
        for (int64_t g = g0; g < full_end; ++g) gather_group_lds<T, 1>(src, idx, dst, g, stage, guard);
    }
    if (g1 > full && (n & 63)) {
        const int64_t i = (full << 6) + trove::lane_id();
        if (i < n) trove::detail::divergent_store(trove::detail::divergent_load(src + guard(idx[i])), dst + i);
    }
    guard.report();
}

// Naive per-lane gather (tests/benchmark.cu:112-119).
template <typename T, typename G = no_guard>
__device__ __forceinline__ void gather_range_direct(const T* __restrict__ src, const int32_t* __restrict__ idx,
                                                    T* __restrict__ dst, int64_t n, int64_t g0, int64_t g1,
                                                    G guard = G()) {
    const int lane = trove::lane_id();
    for (int64_t g = g0; g < g1; ++g) {
        const int64_t i = (g << 6) + lane;
        if (i < n) dst[i] = src[guard(idx[i])];
    }
    guard.report();
}

}  // namespace trove_amd
