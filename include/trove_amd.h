/*
 * trove_amd.h - C ABI of the MI355X (gfx950) array-of-structures load/store
 * library.
 *
 * The reference (bryancatanzaro/trove) is a header-only device library with
 * no host API: its only host entry points are the benchmark and test
 * programs, which launch kernels over trove::coalesced_ptr,
 * load_warp_contiguous / store_warp_contiguous and the block interface.
 * Each function below is the launch of one of those kernels (or of its
 * naive per-lane twin), with the reference program it replaces cited. The
 * device-side drop-in is the header library under include/trove/.
 *
 * Conventions
 *  - Buffers are device pointers owned by the caller; `n` counts structs of
 *    `struct_bytes` bytes. Pointers must be aligned to the struct's word size
 *    (16, 8 or 4 bytes: the widest of those dividing struct_bytes).
 *  - `stream` is a hipStream_t (NULL = the default stream). Every call only
 *    enqueues work; nothing synchronises unless stated.
 *  - Return value: TROVE_OK (0) or a negative TROVE_ERR_* code; a HIP
 *    failure is reported as TROVE_ERR_HIP and trove_last_hip_error() gives
 *    the hipError_t. No function throws.
 *  - Indices (gather / scatter) must lie in [0, length of the indexed array)
 *    and, for scatter, be distinct; they are not range-checked here.
 */
#ifndef TROVE_AMD_H
#define TROVE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TROVE_OK 0
#define TROVE_ERR_INVALID (-1)  /* null pointer, negative count, bad parameter */
#define TROVE_ERR_SIZE (-2)     /* struct size not compiled into the library */
#define TROVE_ERR_ALIGN (-3)    /* pointer not aligned to the struct's word size */
#define TROVE_ERR_VARIANT (-4)  /* unknown variant or tile width */
#define TROVE_ERR_HIP (-5)      /* HIP runtime failure */
#define TROVE_ERR_NOMEM (-6)    /* device or host allocation failed */
#define TROVE_ERR_BUDGET (-7)   /* serving: request larger than the whole budget */
#define TROVE_ERR_FULL (-8)     /* serving: no free request slot */
#define TROVE_ERR_NOTFOUND (-9) /* unknown handle or request id */
#define TROVE_ERR_AGAIN (-10)   /* serving: request not finished yet */
#define TROVE_ERR_RANGE (-11)   /* index out of range (where indices are checked) */

/* Kernel variants. */
#define TROVE_VARIANT_AUTO 0   /* the library's choice (tuned table) */
#define TROVE_VARIANT_SHFL 1   /* Trove API path: warp-contiguous transposition / coalesced_ptr */
#define TROVE_VARIANT_DIRECT 2 /* naive per-lane compiler-generated struct access */
#define TROVE_VARIANT_TILE 3   /* gather/scatter: tile-indexed word access, transposition elided */

/* Message for a return code. */
const char* trove_strerror(int code);
/* hipError_t of the last TROVE_ERR_HIP on the calling thread. */
int trove_last_hip_error(void);
/* Library version, major * 10000 + minor * 100 + patch. */
int trove_version(void);
/* 1 when copy / gather / scatter are compiled for this struct size. */
int trove_struct_size_supported(int struct_bytes);
/* Word size (16, 8, 4 or 1 bytes) the struct is moved in. */
int trove_struct_word_bytes(int struct_bytes);

/* dst[i] = src[i], i < n: one struct per lane through load_warp_contiguous /
 * store_warp_contiguous<tile> (variant SHFL; tile in {2,4,...,64}, 0 = 64) or
 * the naive per-lane copy (DIRECT). Replaces the launches of
 * tests/benchmark.cu:74-88 (contiguous load) and :52-72 (contiguous store)
 * combined into the copy the configs measure; device side: aos.h:66-105. */
int trove_copy(const void* src, void* dst, int64_t n, int struct_bytes, int variant, int tile, void* stream);

/* dst[i] = src[idx[i]], i < n. SHFL: `T d = coalesced_ptr(src)[idx[i]];
 * coalesced_ptr(dst)[i] = d;` (tests/benchmark.cu:90-99); DIRECT: the naive
 * gather (tests/benchmark.cu:112-119); TILE: tile-indexed loads + coalesced
 * stores with the transposition elided. */
int trove_gather(const void* src, const int32_t* idx, void* dst, int64_t n, int struct_bytes, int variant,
                 void* stream);

/* dst[idx[i]] = src[i], i < n, idx distinct. SHFL: tests/benchmark.cu:101-110;
 * DIRECT: tests/benchmark.cu:121-128; TILE: coalesced loads + tile-indexed
 * stores with the transposition elided. */
int trove_scatter(const void* src, const int32_t* idx, void* dst, int64_t n, int struct_bytes, int variant,
                  void* stream);

/* dst[i] = counting_array(i * k) (k = struct_bytes / 4 words): the store
 * benchmark of tests/benchmark.cu:52-72 (SHFL: store_warp_contiguous;
 * DIRECT: plain stores). struct_bytes must be a multiple of 4. */
int trove_store_counting(void* dst, int64_t n, int struct_bytes, int variant, void* stream);

/* out[i] = sum of the 32-bit words of src[i] (wrapping): the load benchmark
 * of tests/benchmark.cu:74-88. struct_bytes must be a multiple of 4. */
int trove_load_sum(const void* src, uint32_t* out, int64_t n, int struct_bytes, int variant, void* stream);

/* Block interface copy over n_elems 32-bit elements, s per lane
 * (s in 1..16): a grid-stride loop over n_elems / s arrays that calls
 * load_array_warp_contiguous / store_array_warp_contiguous when the
 * wavefront is converged and load_array / store_array otherwise
 * (tests/block.cu:34-54); the n_elems % s trailing elements are copied per
 * lane. divergent != 0 masks lanes (seeded hash) in one of eight 64-array
 * groups so that those wavefronts take the fallback. */
int trove_block_copy(const void* src, void* dst, int64_t n_elems, int s, int divergent, uint64_t seed,
                     void* stream);

/* dst[j] = j for the first n_arrays * s 32-bit elements through
 * store_array_warp_contiguous (tests/block.cu:13-32), s in 1..16. */
int trove_block_write(void* dst, int64_t n_arrays, int s, void* stream);

/* Plain 16-byte-per-lane device copy of `bytes` bytes (a multiple of 16):
 * the practical bandwidth ceiling the AoS paths are compared against. */
int trove_memcpy_ceiling(const void* src, void* dst, int64_t bytes, void* stream);

/* Deterministic inputs (host memory; no GPU needed).
 * trove_gen_words: out[i] = 32-bit hash of (seed, i).
 * trove_gen_permutation: uniform random permutation of [0, n) (Fisher-Yates,
 *   xoshiro256** seeded from seed). */
int trove_gen_words(uint32_t* out, int64_t n, uint64_t seed);
int trove_gen_permutation(int32_t* out, int64_t n, uint64_t seed);
/* Device fill with the same values as trove_gen_words (stream-ordered). */
int trove_fill_words_device(uint32_t* out, int64_t n, uint64_t seed, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TROVE_AMD_H */
