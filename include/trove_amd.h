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
#define TROVE_VARIANT_AUTO 0   /* the library's per-size choice (measured on MI355X; see each entry point) */
#define TROVE_VARIANT_SHFL 1   /* Trove API path: warp-contiguous transposition / coalesced_ptr */
#define TROVE_VARIANT_DIRECT 2 /* naive per-lane compiler-generated struct access */
#define TROVE_VARIANT_TILE 3   /* gather/scatter: tile-indexed word access, transposition elided */
#define TROVE_VARIANT_LDS 4    /* LDS-staged transposition (trove/lds.h): dwordx4 tile traffic, per-lane LDS
                                  reads; struct sizes that are multiples of 4 up to 512 bytes */
#define TROVE_VARIANT_DIRECT_LAUNCH 5 /* the naive per-lane struct access launched as SHFL (copy) or TILE
                                         (gather, scatter) are: the same grid, groups in flight, XCD order
                                         and cache policies, so that it differs from them only by the
                                         transposition (a launch-equal baseline; copy, gather, scatter) */

/* Operations for trove_auto_variant. */
#define TROVE_OP_COPY 0
#define TROVE_OP_GATHER 1
#define TROVE_OP_SCATTER 2

/* The variant TROVE_VARIANT_AUTO resolves to for trove_copy (tile 64),
 * trove_gather or trove_scatter at this struct size; TROVE_ERR_SIZE for a
 * size the library does not compile, TROVE_ERR_INVALID for an unknown op.
 * No reference counterpart (the reference has no host API). */
int trove_auto_variant(int op, int struct_bytes);

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
 * store_warp_contiguous<tile> (variant SHFL; tile in {2,4,...,64}, 0 = 64),
 * through load/store_warp_contiguous_lds (LDS; tile 0 or 64) or the naive
 * per-lane copy (DIRECT). AUTO: LDS for 4, 8, 12, 16, 32, 96, 128 and 160
 * bytes, else SHFL (SHFL whenever tile is below 64). Replaces the launches
 * of tests/benchmark.cu:74-88 (contiguous load) and :52-72 (contiguous
 * store) combined into one copy; device side: aos.h:66-105. */
int trove_copy(const void* src, void* dst, int64_t n, int struct_bytes, int variant, int tile, void* stream);

/* dst[i] = src[idx[i]], i < n. SHFL: `T d = coalesced_ptr(src)[idx[i]];
 * coalesced_ptr(dst)[i] = d;` (tests/benchmark.cu:90-99); DIRECT: the naive
 * gather (tests/benchmark.cu:112-119); TILE: tile-indexed loads + coalesced
 * stores with the transposition elided (4-, 8- and 16-byte structs: per-lane
 * row words); LDS: the tile-indexed loads staged through LDS to one struct
 * per lane and back out with dwordx4 stores (the LDS-staged counterpart of
 * SHFL). AUTO: LDS for 16, 48, 64, 192 and 384 bytes, SHFL for 224, else TILE.
 * tile: 0 or 64 = whole wavefronts; a power of two 2..32 runs SHFL (also for
 * AUTO) with each converged tile of that many lanes on its own, as the
 * reference's coalesced_ptr runs on a 32-lane warp; TILE and LDS need whole
 * wavefronts (TROVE_ERR_VARIANT otherwise), DIRECT ignores it. */
int trove_gather(const void* src, const int32_t* idx, void* dst, int64_t n, int struct_bytes, int variant, int tile,
                 void* stream);

/* dst[idx[i]] = src[i], i < n, idx distinct. SHFL: tests/benchmark.cu:101-110;
 * DIRECT: tests/benchmark.cu:121-128; TILE: coalesced loads + tile-indexed
 * stores with the transposition elided; LDS: dwordx4 loads staged through
 * LDS to one struct per lane, then tile-indexed stores. AUTO: DIRECT for 4
 * bytes, LDS for 8 and 16, SHFL for 64, else TILE. tile as for trove_gather. */
int trove_scatter(const void* src, const int32_t* idx, void* dst, int64_t n, int struct_bytes, int variant, int tile,
                  void* stream);

/* inv[perm[i]] = i, i < n, for a permutation perm of [0, n) (not checked;
 * an entry outside [0, n) is skipped, and entries of inv that no perm[i]
 * names keep their contents). No reference counterpart: with it, a scatter
 * through one permutation can be run as gathers through the inverse (see
 * trove_scatter_permutation). */
int trove_invert_permutation(const int32_t* perm, int32_t* inv, int64_t n, void* stream);

/* dst[perm[i]] = src[i], i < n, for a permutation perm of [0, n) (dst holds
 * n structs; not checked): the result of trove_scatter, computed where that
 * is faster as trove_invert_permutation into workspace (n int32, device
 * memory, 4-byte aligned) followed by trove_gather(src, workspace, dst)
 * with the library's per-size path. That is the case for struct sizes of 12
 * to 60 bytes (multiples of 4): random row writes that cover part of a
 * 64-byte memory granule are several times slower than random row reads.
 * Other sizes, or workspace == NULL, run trove_scatter(AUTO). The
 * workspace is zeroed before the inversion, so that for a perm that is not a
 * permutation the result is unspecified but every access stays inside src,
 * dst and workspace. The reference's random scatter benchmark scatters
 * through a permutation (tests/benchmark.cu:101-110, 244-258). */
int trove_scatter_permutation(const void* src, const int32_t* perm, void* dst, int64_t n, int struct_bytes,
                              int32_t* workspace, void* stream);
/* 1 when trove_scatter_permutation with a workspace inverts and gathers at
 * this struct size, 0 when it scatters, TROVE_ERR_SIZE for a size the
 * library does not compile. */
int trove_scatter_permutation_inverts(int struct_bytes);

/* dst[i] = counting_array(i * k) (k = struct_bytes / 4 words): the store
 * benchmark of tests/benchmark.cu:52-72 (SHFL: store_warp_contiguous<tile>;
 * DIRECT: plain stores). tile is the reference program's TILE_SIZE
 * (tests/Makefile:4, tests/benchmark.cu:47-50; its default there is 32):
 * 0 = 64, else a power of two 2..64 (TROVE_ERR_VARIANT otherwise); DIRECT
 * ignores it. struct_bytes must be a multiple of 4 and at most 256 (else
 * TROVE_ERR_SIZE). */
int trove_store_counting(void* dst, int64_t n, int struct_bytes, int variant, int tile, void* stream);

/* out[i] = sum of the 32-bit words of src[i] (wrapping): the load benchmark
 * of tests/benchmark.cu:74-88 (SHFL: load_warp_contiguous<tile>, tile as for
 * trove_store_counting). struct_bytes must be a multiple of 4 and at most
 * 256 (else TROVE_ERR_SIZE). */
int trove_load_sum(const void* src, uint32_t* out, int64_t n, int struct_bytes, int variant, int tile,
                   void* stream);

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

/* Random 128-byte line reads: out[r] = XOR of the 32 words of line
 * min(line_idx[r], n_lines - 1) of buf (128-byte aligned, n_lines lines), with
 * 64 lines in flight per wavefront: the rate at which the device
 * serves random line fills, the ceiling of random gathers. No reference
 * counterpart (benchmark plumbing like trove_memcpy_ceiling). */
int trove_random_lines_ceiling(const void* buf, int64_t n_lines, const uint32_t* line_idx, uint32_t* out,
                               int64_t n_reads, void* stream);

/* Random row writes: for r < n_writes, row min(row_idx[r], n_rows - 1) of
 * buf (rows of row_bytes = 32, 60, 64 or 128 bytes; TROVE_ERR_SIZE
 * otherwise) receives the 32-bit words r * (row_bytes / 4) + j, j <
 * row_bytes / 4, written as 16-byte lane words (4-byte ones for 60-byte
 * rows) with many rows in flight per wavefront, as ordinary stores or, with
 * nontemporal != 0, nontemporal ones: the rate at which the device absorbs
 * random row writes, the ceiling of scatters of that row size. Rows should
 * be distinct (else which write lands is unspecified). No reference
 * counterpart (benchmark plumbing like trove_memcpy_ceiling). */
int trove_random_rows_write_ceiling(void* buf, int64_t n_rows, int row_bytes, const int32_t* row_idx,
                                    int64_t n_writes, int nontemporal, void* stream);

/* ---- Batched jobs: many independent gathers / copies in one launch ----
 * No reference counterpart (the reference launches one kernel per
 * benchmark case, tests/benchmark.cu:143-145); a job with indices is the
 * gather of trove_gather, a job without (idx == NULL) the copy of
 * trove_copy. */
typedef struct trove_job {
    const void* src;      /* device: source structs */
    const int32_t* idx;   /* device: n indices into src, or NULL for a copy */
    void* dst;            /* device: n output structs */
    int64_t n;            /* structs to gather / copy */
    int32_t struct_bytes; /* 4..64, a multiple of 4 */
    int32_t variant;      /* gather: TILE / SHFL / LDS / DIRECT; copy: SHFL (also for TILE) / LDS / DIRECT;
                             AUTO: the per-size choice of trove_gather / trove_copy */
    int32_t chunk_groups; /* 64-struct groups per work item (0 = 16) */
    int32_t reserved;     /* must be 0 */
} trove_job;

/* Opaque plan. Runs of one plan may be issued from several host threads
 * (the graph of TROVE_BATCH_GRAPH is captured once, under a lock); they
 * write the same outputs, so runs that overlap on the device race. */
typedef struct trove_batch trove_batch;

#define TROVE_BATCH_LAUNCH 0   /* one launch, one wavefront per work item of every job */
#define TROVE_BATCH_PERSISTENT TROVE_BATCH_LAUNCH /* the same mode's earlier name */
#define TROVE_BATCH_GRAPH 1    /* the TROVE_BATCH_LAUNCH launch captured once as a hipGraph, then replayed */
#define TROVE_BATCH_LOOPED 2   /* one trove_gather / trove_copy launch per job (the comparison) */
#define TROVE_BATCH_RESIDENT 3 /* one launch of only the wavefronts resident at once, each striding over the
                                  work items (a persistent kernel) */

/* Validates the jobs and uploads the job table (allocates device memory). */
int trove_batch_create(const trove_job* jobs, int32_t n_jobs, trove_batch** out);
/* Enqueues the whole batch on `stream` in the given mode. */
int trove_batch_run(trove_batch* b, int mode, void* stream);
/* Runs the jobs as trove_batch_create, trove_batch_run(TROVE_BATCH_LAUNCH)
 * and trove_batch_destroy would, but allocates nothing: the job table goes
 * into a static device array of the library (one per device, held until the
 * process ends; TROVE_BATCH_STATIC_MAX_JOBS descriptors of 56 bytes, 224
 * KiB), filled by small kernel launches that carry the descriptors in their
 * arguments. No reference counterpart: for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Jobs are validated as in
 * trove_batch_create; more than TROVE_BATCH_STATIC_MAX_JOBS jobs return
 * TROVE_ERR_FULL. Every call on one device shares that device's table:
 * calls in order on one stream are safe, calls that overlap on one device
 * (other streams, other host threads) race. */
#define TROVE_BATCH_STATIC_MAX_JOBS 4096
int trove_batch_run_static(const trove_job* jobs, int32_t n_jobs, void* stream);
/* Work items and algorithmic bytes ((2 * struct_bytes + 4) * n per gather,
 * 2 * struct_bytes * n per copy, summed). */
int trove_batch_stats(const trove_batch* b, int64_t* work_items, int64_t* algorithmic_bytes);
int trove_batch_destroy(trove_batch* b);

/* ---- Serving: continuously batched gathers against a device table ----
 * No reference counterpart. Requests "gather rows idx[0..n) of the table"
 * are admitted in FIFO order while a request slot and `n` rows of the
 * output budget are free; every step gathers the next slice (at most
 * slice_indices rows) of every admitted request in one launch, so requests
 * that arrive while others run join the next step. A finished request's
 * rows stay in the device output arena until trove_serve_release. One host
 * thread drives a server; calls on one server must not overlap. */
typedef struct trove_serve trove_serve; /* opaque server */

/* table: device rows of row_bytes bytes (4..64, multiple of 4), n_rows of
 * them. budget_bytes bounds the output rows of admitted requests (the index
 * arena is budget_bytes / row_bytes * 4 bytes more). variant: the gather
 * path of every step (TILE, SHFL, LDS, DIRECT, or AUTO: trove_gather's
 * per-size choice). Work runs on
 * `stream`; host index uploads run on a private stream. */
int trove_serve_create(const void* table, int64_t n_rows, int32_t row_bytes, int64_t budget_bytes,
                       int32_t max_slots, int64_t slice_indices, int32_t variant, void* stream,
                       trove_serve** out);
/* Queues a request. Host indices are range-checked (TROVE_ERR_RANGE) and
 * uploaded to a device staging pool of the server (as many entries as the
 * output budget has rows, at most 2^28; allocated at the first host-index
 * submission) before the
 * call returns, or copied to host memory when that pool is full, so the
 * caller's array may change or be freed right after. Device indices are read in place and must stay valid
 * until the request is done; the step kernels check each of them against
 * the table as they gather (an index outside [0, n_rows) reads row 0 and is
 * never dereferenced past the table), and such a request completes as
 * rejected: trove_serve_poll, _output and _fetch then return
 * TROVE_ERR_RANGE. TROVE_ERR_BUDGET when n rows exceed the whole budget. */
int trove_serve_submit(trove_serve* s, const int32_t* idx, int64_t n, int idx_on_device, uint64_t* request_id);
/* Retires finished steps, admits queued requests and launches one step.
 * *in_flight = requests queued or running plus steps outstanding. */
int trove_serve_step(trove_serve* s, int32_t* in_flight);
/* *done = 1 once the request's rows are written; latency from submission.
 * Returns TROVE_ERR_RANGE (with *done = 1) for a completed request that
 * held a device index outside the table. */
int trove_serve_poll(trove_serve* s, uint64_t request_id, int32_t* done, double* latency_s);
/* Device address of a finished request's n output rows (until released). */
int trove_serve_output(trove_serve* s, uint64_t request_id, void** dev_out);
/* *n = the request's row count, in any state (the size of the buffer
 * trove_serve_fetch fills is *n times row_bytes); TROVE_ERR_NOTFOUND for an
 * unknown or released request. */
int trove_serve_rows(const trove_serve* s, uint64_t request_id, int64_t* n);
/* Copies a finished request's n rows to `out` (host or device memory, n *
 * row_bytes bytes), synchronously. */
int trove_serve_fetch(trove_serve* s, uint64_t request_id, void* out);
/* Frees a finished request's rows (the caller's reads must be complete). */
int trove_serve_release(trove_serve* s, uint64_t request_id);
/* Steps until every submitted request is done (blocking). */
int trove_serve_drain(trove_serve* s);
int trove_serve_stats(const trove_serve* s, int64_t* completed, int64_t* indices_done, int64_t* steps,
                      int32_t* queued, int32_t* active);
/* Replays an arrival trace in real time: request i (idx[i], n[i] rows) is
 * submitted at arrival_s[i] seconds after the start (ascending); returns
 * when all are done, with latency_s[i] = completion - scheduled arrival
 * (-1 for a request rejected for a device index outside the table) and,
 * when request_ids is not NULL, request_ids[i] = request i's id (for
 * trove_serve_output / _fetch / _release when its rows are kept).
 * release_on_done frees each request's rows as soon as it completes.
 * Request ids are opaque: no order or starting value is promised. */
int trove_serve_replay(trove_serve* s, int32_t n_req, const double* arrival_s, const int32_t* const* idx,
                       const int64_t* n, int idx_on_device, int release_on_done, double* latency_s,
                       uint64_t* request_ids);
int trove_serve_destroy(trove_serve* s);

/* Deterministic inputs (host memory; no GPU needed).
 * trove_gen_words: out[i] = 32-bit hash of (seed, i).
 * trove_gen_permutation: uniform random permutation of [0, n) (Fisher-Yates,
 *   xoshiro256** seeded from seed). */
int trove_gen_words(uint32_t* out, int64_t n, uint64_t seed);
int trove_gen_permutation(int32_t* out, int64_t n, uint64_t seed);
/* Device fill with the same values as trove_gen_words (stream-ordered). */
int trove_fill_words_device(uint32_t* out, int64_t n, uint64_t seed, void* stream);

#define TROVE_PATTERN_IDENTITY 1 /* out[i] = i % n_src */
#define TROVE_PATTERN_RANDOM 2   /* a permutation of [0, n) when n == n_src, else uniform draws */
#define TROVE_PATTERN_RUNS 3     /* ascending runs of geometric length (mean 32) at uniform starts */
/* n indices into [0, n_src) following `pattern`. */
int trove_gen_indices(int32_t* out, int64_t n, int64_t n_src, int pattern, uint64_t seed);
/* n draws from Zipf(s) over ranks 1..n_items, returned as rows rank - 1
 * (rejection-inversion, Hoermann & Derflinger 1996). */
int trove_gen_zipf(int32_t* out, int64_t n, int64_t n_items, double s, uint64_t seed);
/* n integers log-uniform on [lo, hi]: floor(exp(U(log lo, log(hi + 1)))). */
int trove_gen_loguniform(int64_t* out, int64_t n, int64_t lo, int64_t hi, uint64_t seed);

#ifdef __cplusplus
}
#endif

#endif /* TROVE_AMD_H */
