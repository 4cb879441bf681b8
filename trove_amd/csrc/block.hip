// Block interface kernels (tests/block.cu:13-54) and plain device helpers.
#include "common.h"

#include <type_traits>

namespace trove_amd {

// Lanes of one 64-array group that a seeded hash sends down the fallback
// path before the convergence test: one group in eight is affected, and in
// it each lane with probability 1/2.
__device__ __forceinline__ bool divergent_group(uint64_t seed, int64_t group) {
    return (splitmix64(seed ^ (uint64_t)group) & 7) == 0;
}

__device__ __forceinline__ bool divergent_lane(uint64_t seed, int64_t index) {
    const uint64_t h = splitmix64(seed ^ (uint64_t)(index >> 6));
    if ((h & 7) != 0) return false;
    return (splitmix64(h + (uint64_t)(index & 63)) & 1) != 0;
}

// UNROLL 64-array groups per 64-thread work-group (one-shot grid; 4 / S
// groups for arrays under 12 bytes so a wavefront moves about 1 KiB),
// XCD-contiguous unit order from 20-byte arrays up, streaming accesses. The
// partial last group has inactive lanes, so warp_converged() is false there
// and it takes the fallback, as in the reference's test.
template <int S>
struct block_shape {
    static constexpr int unroll = S == 1 ? 4 : S == 2 ? 2 : 1;
};

template <int S>
__global__ __launch_bounds__(64) void k_block_copy(const uint32_t* __restrict__ x, uint32_t* __restrict__ r, int64_t l,
                                                   int64_t e, int divergent, uint64_t seed) {
    typedef trove::array<uint32_t, S> s_ary;
    constexpr int U = block_shape<S>::unroll;
    const int64_t unit = S >= 5 ? unit_for_xcd(blockIdx.x, (int64_t)gridDim.x) : (int64_t)blockIdx.x;
    // Every group of the unit whole, undisturbed and converged: all loads
    // before all stores, so that U loads are in flight per wavefront, and no
    // per-lane bounds branch around the warp-contiguous accesses (one-group
    // units included).
    bool whole = ((unit * U + U) << 6) <= l;
#pragma unroll
    for (int u = 0; u < U; ++u) whole = whole && !(divergent && divergent_group(seed, unit * U + u));
    if (whole && trove::warp_converged()) {
        s_ary d[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
            d[u] = trove::load_array_warp_contiguous<S, trove::WARP_SIZE, trove::streaming>(
                x, ((unit * U + u) << 6) + threadIdx.x);
#pragma unroll
        for (int u = 0; u < U; ++u)
            trove::store_array_warp_contiguous<S, trove::WARP_SIZE, trove::streaming>(
                r, ((unit * U + u) << 6) + threadIdx.x, d[u]);
    } else {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t index = ((unit * U + u) << 6) + threadIdx.x;
            if (index < l) {
                if (divergent && divergent_lane(seed, index)) {
                    const s_ary d = trove::load_array<S>(x, index);
                    trove::store_array(r, index, d);
                } else if (trove::warp_converged()) {
                    const s_ary d = trove::load_array_warp_contiguous<S, trove::WARP_SIZE, trove::streaming>(x, index);
                    trove::store_array_warp_contiguous<S, trove::WARP_SIZE, trove::streaming>(r, index, d);
                } else {
                    const s_ary d = trove::load_array<S>(x, index);
                    trove::store_array(r, index, d);
                }
            }
        }
    }
    // Trailing e - l*S elements, one per lane of the first work-group.
    const int64_t j = l * S + (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (j < e) r[j] = x[j];
}

template <int S>
__global__ __launch_bounds__(kBlock) void k_block_write(uint32_t* __restrict__ r, int64_t l) {
    typedef trove::array<uint32_t, S> s_ary;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t index = (int64_t)blockIdx.x * kBlock + threadIdx.x; index < l; index += stride) {
        const s_ary d = trove::counting_array<s_ary>::impl((uint32_t)(S * index));
        if (trove::warp_converged()) {
            trove::store_array_warp_contiguous<S>(r, index, d);
        } else {
            trove::store_array(r, index, d);
        }
    }
}

template <int S>
static int launch_block_copy(const uint32_t* x, uint32_t* r, int64_t e, int divergent, uint64_t seed,
                             hipStream_t s) {
    const int64_t l = e / S;
    // One work-group per 64-array group; at least enough lanes for the
    // e - l*S < S <= 64 trailing elements.
    constexpr int U = block_shape<S>::unroll;
    const int64_t groups = (l + 63) / 64;
    const int64_t blocks = grid_for(groups > 0 ? (groups + U - 1) / U : 1, 1);
    const size_t lds = stream_lds_reservation((int64_t)U * 64 * 4 * S);
    hipLaunchKernelGGL((k_block_copy<S>), dim3((unsigned)blocks), dim3(64), lds, s, x, r, l, e, divergent, seed);
    return launch_status();
}

template <int S>
static int launch_block_write(uint32_t* r, int64_t l, hipStream_t s) {
    const int64_t blocks = grid_for((l + 63) / 64 > 0 ? (l + 63) / 64 : 1);
    hipLaunchKernelGGL((k_block_write<S>), dim3((unsigned)blocks), dim3(kBlock), 0, s, r, l);
    return launch_status();
}

#define TROVE_FOR_EACH_S(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16)

int block_copy_dispatch(const uint32_t* x, uint32_t* r, int64_t e, int s, int divergent, uint64_t seed,
                        hipStream_t st) {
    switch (s) {
#define TROVE_CASE_BC(S) \
    case S: return launch_block_copy<S>(x, r, e, divergent, seed, st);
        TROVE_FOR_EACH_S(TROVE_CASE_BC)
#undef TROVE_CASE_BC
        default: return TROVE_ERR_INVALID;
    }
}

int block_write_dispatch(uint32_t* r, int64_t l, int s, hipStream_t st) {
    switch (s) {
#define TROVE_CASE_BW(S) \
    case S: return launch_block_write<S>(r, l, st);
        TROVE_FOR_EACH_S(TROVE_CASE_BW)
#undef TROVE_CASE_BW
        default: return TROVE_ERR_INVALID;
    }
}

// ---- plain device helpers ----

// Plain copy in the fastest shape measured on MI355X: one-shot grid, 16
// bytes per lane, 4 KiB per work-group, nontemporal loads and stores.
__global__ __launch_bounds__(kBlock) void k_memcpy16(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                     int64_t n16) {
    typedef uint32_t vec __attribute__((ext_vector_type(4)));
    const vec* s = reinterpret_cast<const vec*>(src);
    vec* d = reinterpret_cast<vec*>(dst);
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n16; i += stride)
        __builtin_nontemporal_store(__builtin_nontemporal_load(s + i), d + i);
}

// Random-line read ceiling: the 8-lane group g of a wavefront reads the
// whole 128-byte line min(line_idx[r], n_lines - 1) of buf (16 bytes per lane) and
// out[r] receives the XOR of its 32 words. Every wavefront issues all of
// its kLineRounds rounds of loads (8 lines each) before using any: 64 line
// fills in flight per wavefront at full occupancy (measured on MI355X: 8
// rounds 5.34 TB/s, 16 rounds 5.32, 32 rounds 5.0 at a quarter of the
// occupancy). The rate at which HBM serves random 128-byte lines bounds
// every random gather.
constexpr int kLineRounds = 8;

__global__ __launch_bounds__(kBlock) void k_random_lines(const uint4* __restrict__ buf, uint32_t n_lines,
                                                         const uint32_t* __restrict__ line_idx,
                                                         uint32_t* __restrict__ out, int64_t n) {
    typedef uint32_t vec __attribute__((ext_vector_type(4)));
    const vec* b = reinterpret_cast<const vec*>(buf);
    const int lane = trove::lane_id();
    const int64_t r0 = wave_index() * (int64_t)(kLineRounds * 8) + (lane >> 3);
    // Branch-free: rounds past n re-read the last line (their results are
    // not written), so that all index loads, then all line loads, issue
    // without a wait between them (a load under a branch is followed by a
    // wait for every load in flight).
    uint32_t li[kLineRounds];
#pragma unroll
    for (int k = 0; k < kLineRounds; ++k) li[k] = line_idx[r0 + 8 * k < n ? r0 + 8 * k : n - 1];
    vec v[kLineRounds];
#pragma unroll
    for (int k = 0; k < kLineRounds; ++k) v[k] = b[(int64_t)(li[k] < n_lines ? li[k] : n_lines - 1) * 8 + (lane & 7)];
#pragma unroll
    for (int k = 0; k < kLineRounds; ++k) {
        uint32_t x = v[k].x ^ v[k].y ^ v[k].z ^ v[k].w;
        x ^= (uint32_t)__shfl_xor((int)x, 1);
        x ^= (uint32_t)__shfl_xor((int)x, 2);
        x ^= (uint32_t)__shfl_xor((int)x, 4);
        const int64_t r = r0 + 8 * k;
        if ((lane & 7) == 0 && r < n) out[r] = x;
    }
}

int random_lines_launch(const void* buf, int64_t n_lines, const uint32_t* line_idx, uint32_t* out, int64_t n,
                        hipStream_t s) {
    const int64_t waves = (n + kLineRounds * 8 - 1) / (kLineRounds * 8);
    const int64_t blocks = (waves + kWavesPerBlock - 1) / kWavesPerBlock;
    hipLaunchKernelGGL(k_random_lines, dim3((unsigned)blocks), dim3(kBlock), 0, s, static_cast<const uint4*>(buf),
                       (uint32_t)n_lines, line_idx, out, n);
    return launch_status();
}

// Random-row write ceiling: write r (r < n) stores ROW bytes, the words
// w_j = r * (ROW / 4) + j, at row min(row_idx[r], n_rows - 1) of buf. A row
// is G lanes of one word each (16-byte words when ROW is a multiple of 16,
// else 4-byte words: 60-byte rows are 15 lanes), 64 / G rows per wavefront
// per round, kRowRounds rounds per wavefront with every index load issued
// before any store. The rate at which HBM absorbs random row writes of that
// size bounds every scatter of rows of that size.
constexpr int kRowRounds = 8;

template <int ROW, bool NT>
__global__ __launch_bounds__(kBlock) void k_random_rows_write(uint8_t* __restrict__ buf, uint32_t n_rows,
                                                              const int32_t* __restrict__ row_idx, int64_t n) {
    constexpr int W = ROW % 16 == 0 ? 16 : 4;  // bytes per lane
    constexpr int G = ROW / W;                  // lanes per row
    constexpr int P = 64 / G;                   // rows per wavefront round
    typedef uint32_t vec4 __attribute__((ext_vector_type(4)));
    typedef typename std::conditional<W == 16, vec4, uint32_t>::type word;
    const int lane = trove::lane_id();
    const int slot = lane / G, part = lane % G;
    if (slot >= P) return;  // lanes past the last whole row (60-byte rows: lanes 60..63)
    const int64_t r0 = wave_index() * (int64_t)(kRowRounds * P) + slot;
    int64_t row[kRowRounds];
#pragma unroll
    for (int k = 0; k < kRowRounds; ++k) {
        const int64_t r = r0 + (int64_t)P * k;
        const uint32_t ix = (uint32_t)row_idx[r < n ? r : n - 1];
        row[k] = ix < n_rows ? ix : n_rows - 1;
    }
#pragma unroll
    for (int k = 0; k < kRowRounds; ++k) {
        const int64_t r = r0 + (int64_t)P * k;
        if (r < n) {
            const uint32_t w0 = (uint32_t)(r * (ROW / 4)) + (uint32_t)(part * (W / 4));
            word v;
            if constexpr (W == 16) {
                v = vec4{w0, w0 + 1, w0 + 2, w0 + 3};
            } else {
                v = w0;
            }
            word* p = reinterpret_cast<word*>(buf + row[k] * ROW + part * W);
            if constexpr (NT) {
                __builtin_nontemporal_store(v, p);
            } else {
                *p = v;
            }
        }
    }
}

template <int ROW>
static void launch_rows_write(bool nt, unsigned blocks, uint8_t* b, uint32_t n_rows, const int32_t* row_idx,
                              int64_t n, hipStream_t s) {
    if (nt) {
        hipLaunchKernelGGL((k_random_rows_write<ROW, true>), dim3(blocks), dim3(kBlock), 0, s, b, n_rows, row_idx, n);
    } else {
        hipLaunchKernelGGL((k_random_rows_write<ROW, false>), dim3(blocks), dim3(kBlock), 0, s, b, n_rows, row_idx, n);
    }
}

int random_rows_write_launch(void* buf, int64_t n_rows, int row_bytes, const int32_t* row_idx, int64_t n,
                             bool nontemporal, hipStream_t s) {
    int rows_per_wave = 0;
    switch (row_bytes) {
        case 32: rows_per_wave = 64 / 2; break;
        case 60: rows_per_wave = 64 / 15; break;
        case 64: rows_per_wave = 64 / 4; break;
        case 128: rows_per_wave = 64 / 8; break;
        default: return TROVE_ERR_SIZE;
    }
    const int64_t waves = (n + kRowRounds * rows_per_wave - 1) / (kRowRounds * rows_per_wave);
    const unsigned blocks = (unsigned)((waves + kWavesPerBlock - 1) / kWavesPerBlock);
    uint8_t* b = static_cast<uint8_t*>(buf);
    const uint32_t nr = (uint32_t)n_rows;
    switch (row_bytes) {
        case 32: launch_rows_write<32>(nontemporal, blocks, b, nr, row_idx, n, s); break;
        case 60: launch_rows_write<60>(nontemporal, blocks, b, nr, row_idx, n, s); break;
        case 64: launch_rows_write<64>(nontemporal, blocks, b, nr, row_idx, n, s); break;
        default: launch_rows_write<128>(nontemporal, blocks, b, nr, row_idx, n, s); break;
    }
    return launch_status();
}

// inv[perm[i]] = i for a permutation perm of [0, n): one 4-byte random write
// per element, as ordinary (cached) stores. The naive 4-byte scatter, the
// same random writes, moves 2^26 elements in 0.86 ms on MI355X (78 G
// writes/s), several times the rate of random 32-byte rows. An entry outside
// [0, n) is skipped, so that no input writes outside inv.
__global__ __launch_bounds__(kBlock) void k_invert_permutation(const int32_t* __restrict__ perm,
                                                               int32_t* __restrict__ inv, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        const int32_t p = __builtin_nontemporal_load(perm + i);
        if (p >= 0 && (int64_t)p < n) inv[p] = (int32_t)i;
    }
}

__global__ __launch_bounds__(kBlock) void k_zero_words(int32_t* __restrict__ out, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) out[i] = 0;
}

int invert_permutation_launch(const int32_t* perm, int32_t* inv, int64_t n, bool zero_first, hipStream_t s) {
    const int64_t blocks = grid_for((n + 63) / 64);
    if (zero_first) {
        // A fill kernel rather than hipMemsetAsync: callers may capture this
        // sequence into a hipGraph.
        hipLaunchKernelGGL(k_zero_words, dim3((unsigned)blocks), dim3(kBlock), 0, s, inv, n);
        const int rc = launch_status();
        if (rc != TROVE_OK) return rc;
    }
    hipLaunchKernelGGL(k_invert_permutation, dim3((unsigned)blocks), dim3(kBlock), 0, s, perm, inv, n);
    return launch_status();
}

int memcpy_ceiling_launch(const void* src, void* dst, int64_t bytes, hipStream_t s) {
    const int64_t n16 = bytes / 16;
    const int64_t blocks = grid_for((n16 + 63) / 64);
    hipLaunchKernelGGL(k_memcpy16, dim3((unsigned)blocks), dim3(kBlock), 0, s, static_cast<const uint4*>(src),
                       static_cast<uint4*>(dst), n16);
    return launch_status();
}

__host__ __device__ __forceinline__ uint32_t word_hash(uint64_t seed, uint64_t i) {
    return (uint32_t)(splitmix64(seed * 0x632BE59BD9B4E019ull + i) >> 32);
}

__global__ __launch_bounds__(kBlock) void k_fill_words(uint32_t* __restrict__ out, int64_t n, uint64_t seed) {
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) out[i] = word_hash(seed, i);
}

int fill_words_launch(uint32_t* out, int64_t n, uint64_t seed, hipStream_t s) {
    const int64_t blocks = grid_for((n + 63) / 64 > 0 ? (n + 63) / 64 : 1);
    hipLaunchKernelGGL(k_fill_words, dim3((unsigned)blocks), dim3(kBlock), 0, s, out, n, seed);
    return launch_status();
}

uint32_t host_word_hash(uint64_t seed, uint64_t i) { return word_hash(seed, i); }

}  // namespace trove_amd
