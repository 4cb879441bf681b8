// extern "C" entry points of libtrove_amd: argument validation, dispatch by
// struct size to the kernel translation units, error mapping.
#include "common.h"

#include <cstdlib>
#include <cstring>
#include <random>

namespace trove_amd {

int block_copy_dispatch(const uint32_t* x, uint32_t* r, int64_t e, int s, int divergent, uint64_t seed,
                        hipStream_t st);
int block_write_dispatch(uint32_t* r, int64_t l, int s, hipStream_t st);
int memcpy_ceiling_launch(const void* src, void* dst, int64_t bytes, hipStream_t s);
int random_rows_write_launch(void* buf, int64_t n_rows, int row_bytes, const int32_t* row_idx, int64_t n,
                             bool nontemporal, hipStream_t s);
int invert_permutation_launch(const int32_t* perm, int32_t* inv, int64_t n, bool zero_first, hipStream_t s);
int random_lines_launch(const void* buf, int64_t n_lines, const uint32_t* line_idx, uint32_t* out, int64_t n,
                        hipStream_t s);
int fill_words_launch(uint32_t* out, int64_t n, uint64_t seed, hipStream_t s);
uint32_t host_word_hash(uint64_t seed, uint64_t i);

static thread_local int g_last_hip_error = 0;

int hip_status(hipError_t e) {
    if (e == hipSuccess) return TROVE_OK;
    g_last_hip_error = (int)e;
    return TROVE_ERR_HIP;
}

int launch_status() { return hip_status(hipGetLastError()); }

int device_cu_count() {
    static thread_local int cached_dev = -1;
    static thread_local int cached_cus = 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (dev != cached_dev) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        cached_dev = dev;
        cached_cus = cus;
    }
    return cached_cus;
}

static int lds_per_cu() {
    static thread_local int cached_dev = -1;
    static thread_local int lds = 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 160 * 1024;
    if (dev != cached_dev) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess ||
            v <= 0)
            v = 160 * 1024;
        cached_dev = dev;
        lds = v;
    }
    return lds;
}

size_t cap_lds_reservation(int blocks_per_cu) {
    if (blocks_per_cu <= 0) return 0;
    const int units = lds_per_cu() / kLdsAllocBytes;
    const int per = units / blocks_per_cu;
    return per > 0 ? (size_t)per * kLdsAllocBytes : 0;
}

size_t stream_lds_reservation(int64_t bytes_per_block) {
    if (bytes_per_block <= 0) return 0;
    int64_t groups = kStreamWindowBytes / bytes_per_block;
    if (groups >= 32) return 0;  // a CU holds at most 32 wavefronts anyway
    if (groups < 4) groups = 4;
    const int64_t per = (int64_t)lds_per_cu() / groups;
    return (size_t)(per / 512 * 512);
}

int64_t grid_for(int64_t waves, int waves_per_block) {
    if (waves < 1) waves = 1;
    const int64_t blocks = (waves + waves_per_block - 1) / waves_per_block;
    const int64_t max_grid = 0x7fffffff;
    return blocks > max_grid ? max_grid : blocks;
}

typedef int (*copy_fn)(const void*, void*, int64_t, int, int, hipStream_t);
typedef int (*index_fn)(const void*, const int32_t*, void*, int64_t, int, int, hipStream_t);
typedef int (*counting_fn)(void*, int64_t, int, int, hipStream_t);
typedef int (*loadsum_fn)(const void*, uint32_t*, int64_t, int, int, hipStream_t);

struct size_entry {
    int bytes;
    copy_fn copy;
    index_fn gather;
    index_fn scatter;
    counting_fn counting;
    loadsum_fn loadsum;
};

static const size_entry* find_size(int bytes) {
    static const size_entry table[] = {
#define TROVE_ENTRY(B) {B, &copy_impl<B>, &gather_impl<B>, &scatter_impl<B>, &counting_impl<B>, &loadsum_impl<B>},
        TROVE_FOR_EACH_SIZE(TROVE_ENTRY)
#undef TROVE_ENTRY
    };
    for (const size_entry& e : table)
        if (e.bytes == bytes) return &e;
    return nullptr;
}

static bool valid_variant(int v) { return v >= TROVE_VARIANT_AUTO && v <= TROVE_VARIANT_DIRECT_LAUNCH; }
// Tile widths of the warp-contiguous API: 0 (= 64) or a power of two 2..64.
static bool valid_tile(int tile) { return tile == 0 || (tile >= 2 && tile <= 64 && (tile & (tile - 1)) == 0); }

// Largest n the one-unit-per-wavefront kernels cover with a grid of at most
// 2^31 - 1 single-wavefront work-groups (2^37 structs, far beyond any HBM).
constexpr int64_t kMaxStructs = (int64_t)0x7fffffff * 64;

}  // namespace trove_amd

using namespace trove_amd;

extern "C" {

const char* trove_strerror(int code) {
    switch (code) {
        case TROVE_OK: return "success";
        case TROVE_ERR_INVALID: return "invalid argument";
        case TROVE_ERR_SIZE: return "struct size not supported";
        case TROVE_ERR_ALIGN: return "pointer not aligned to the struct's word size";
        case TROVE_ERR_VARIANT: return "unknown variant or tile width";
        case TROVE_ERR_HIP: return "HIP runtime error";
        case TROVE_ERR_NOMEM: return "out of memory";
        case TROVE_ERR_BUDGET: return "request exceeds the memory budget";
        case TROVE_ERR_FULL: return "no free request slot";
        case TROVE_ERR_NOTFOUND: return "unknown handle or request";
        case TROVE_ERR_AGAIN: return "not finished";
        case TROVE_ERR_RANGE: return "index out of range";
        default: return "unknown error";
    }
}

int trove_last_hip_error(void) { return g_last_hip_error; }

int trove_version(void) { return 10000; }

int trove_struct_size_supported(int struct_bytes) { return find_size(struct_bytes) != nullptr; }

int trove_struct_word_bytes(int struct_bytes) { return struct_bytes > 0 ? word_bytes_of(struct_bytes) : 0; }

int trove_auto_variant(int op, int struct_bytes) {
    if (op < TROVE_OP_COPY || op > TROVE_OP_SCATTER) return TROVE_ERR_INVALID;
    if (!find_size(struct_bytes)) return TROVE_ERR_SIZE;
    if (op == TROVE_OP_COPY) return auto_copy_variant(struct_bytes);
    // The tile path covers every size that is a whole number of dwords.
    const int v = op == TROVE_OP_GATHER ? auto_gather_variant(struct_bytes) : auto_scatter_variant(struct_bytes);
    return v == TROVE_VARIANT_TILE && struct_bytes % 4 != 0 ? TROVE_VARIANT_DIRECT : v;
}

int trove_copy(const void* src, void* dst, int64_t n, int struct_bytes, int variant, int tile, void* stream) {
    const size_entry* e = find_size(struct_bytes);
    if (!e) return TROVE_ERR_SIZE;
    if (n < 0 || n > kMaxStructs) return TROVE_ERR_INVALID;
    if (!valid_variant(variant) || variant == TROVE_VARIANT_TILE) return TROVE_ERR_VARIANT;
    if (!valid_tile(tile)) return TROVE_ERR_VARIANT;
    if (variant == TROVE_VARIANT_LDS && (tile != 0 && tile != 64)) return TROVE_ERR_VARIANT;  // whole wavefronts
    if (n == 0) return TROVE_OK;
    if (!src || !dst) return TROVE_ERR_INVALID;
    const int w = word_bytes_of(struct_bytes);
    if (!aligned(src, w) || !aligned(dst, w)) return TROVE_ERR_ALIGN;
    return e->copy(src, dst, n, variant, tile, static_cast<hipStream_t>(stream));
}

// Argument checks shared by gather and scatter. *launch is false when there
// is nothing to do (n == 0).
static int index_common(const size_entry* e, const void* src, const int32_t* idx, void* dst, int64_t n,
                        int struct_bytes, int variant, bool* launch) {
    *launch = false;
    if (!e) return TROVE_ERR_SIZE;
    if (n < 0 || n > kMaxStructs) return TROVE_ERR_INVALID;
    if (!valid_variant(variant)) return TROVE_ERR_VARIANT;
    if (n == 0) return TROVE_OK;
    if (!src || !dst || !idx) return TROVE_ERR_INVALID;
    const int w = word_bytes_of(struct_bytes);
    if (!aligned(src, w) || !aligned(dst, w) || !aligned(idx, 4)) return TROVE_ERR_ALIGN;
    *launch = true;
    return TROVE_OK;
}

int trove_gather(const void* src, const int32_t* idx, void* dst, int64_t n, int struct_bytes, int variant, int tile,
                 void* stream) {
    const size_entry* e = find_size(struct_bytes);
    bool launch = false;
    if (!valid_tile(tile)) return TROVE_ERR_VARIANT;
    const int rc = index_common(e, src, idx, dst, n, struct_bytes, variant, &launch);
    if (rc != TROVE_OK || !launch) return rc;
    return e->gather(src, idx, dst, n, variant, tile, static_cast<hipStream_t>(stream));
}

int trove_scatter(const void* src, const int32_t* idx, void* dst, int64_t n, int struct_bytes, int variant, int tile,
                  void* stream) {
    const size_entry* e = find_size(struct_bytes);
    bool launch = false;
    if (!valid_tile(tile)) return TROVE_ERR_VARIANT;
    const int rc = index_common(e, src, idx, dst, n, struct_bytes, variant, &launch);
    if (rc != TROVE_OK || !launch) return rc;
    return e->scatter(src, idx, dst, n, variant, tile, static_cast<hipStream_t>(stream));
}

int trove_invert_permutation(const int32_t* perm, int32_t* inv, int64_t n, void* stream) {
    if (n < 0 || n > kMaxStructs) return TROVE_ERR_INVALID;
    if (n == 0) return TROVE_OK;
    if (!perm || !inv) return TROVE_ERR_INVALID;
    if (!aligned(perm, 4) || !aligned(inv, 4)) return TROVE_ERR_ALIGN;
    return invert_permutation_launch(perm, inv, n, false, static_cast<hipStream_t>(stream));
}

int trove_scatter_permutation_inverts(int struct_bytes) {
    if (!find_size(struct_bytes)) return TROVE_ERR_SIZE;
    return auto_scatter_by_inverse(struct_bytes) ? 1 : 0;
}

int trove_scatter_permutation(const void* src, const int32_t* perm, void* dst, int64_t n, int struct_bytes,
                              int32_t* workspace, void* stream) {
    const size_entry* e = find_size(struct_bytes);
    bool launch = false;
    const int rc = index_common(e, src, perm, dst, n, struct_bytes, TROVE_VARIANT_AUTO, &launch);
    if (rc != TROVE_OK || !launch) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (!workspace || !auto_scatter_by_inverse(struct_bytes))
        return e->scatter(src, perm, dst, n, TROVE_VARIANT_AUTO, 0, s);
    if (!aligned(workspace, 4)) return TROVE_ERR_ALIGN;
    // Zeroed first: should perm not be a permutation, the rows nothing maps
    // to gather row 0 and every access stays inside src, dst and workspace.
    const int irc = invert_permutation_launch(perm, workspace, n, true, s);
    if (irc != TROVE_OK) return irc;
    return e->gather(src, workspace, dst, n, TROVE_VARIANT_AUTO, 0, s);
}

int trove_store_counting(void* dst, int64_t n, int struct_bytes, int variant, int tile, void* stream) {
    const size_entry* e = find_size(struct_bytes);
    if (!e || struct_bytes % 4 != 0) return TROVE_ERR_SIZE;
    if (n < 0) return TROVE_ERR_INVALID;
    if (!valid_variant(variant) || variant == TROVE_VARIANT_TILE || variant == TROVE_VARIANT_LDS ||
        variant == TROVE_VARIANT_DIRECT_LAUNCH || !valid_tile(tile))
        return TROVE_ERR_VARIANT;
    if (n == 0) return TROVE_OK;
    if (!dst) return TROVE_ERR_INVALID;
    if (!aligned(dst, word_bytes_of(struct_bytes))) return TROVE_ERR_ALIGN;
    return e->counting(dst, n, variant, tile, static_cast<hipStream_t>(stream));
}

int trove_load_sum(const void* src, uint32_t* out, int64_t n, int struct_bytes, int variant, int tile,
                   void* stream) {
    const size_entry* e = find_size(struct_bytes);
    if (!e || struct_bytes % 4 != 0) return TROVE_ERR_SIZE;
    if (n < 0) return TROVE_ERR_INVALID;
    if (!valid_variant(variant) || variant == TROVE_VARIANT_TILE || variant == TROVE_VARIANT_LDS ||
        variant == TROVE_VARIANT_DIRECT_LAUNCH || !valid_tile(tile))
        return TROVE_ERR_VARIANT;
    if (n == 0) return TROVE_OK;
    if (!src || !out) return TROVE_ERR_INVALID;
    if (!aligned(src, word_bytes_of(struct_bytes)) || !aligned(out, 4)) return TROVE_ERR_ALIGN;
    return e->loadsum(src, out, n, variant, tile, static_cast<hipStream_t>(stream));
}

int trove_block_copy(const void* src, void* dst, int64_t n_elems, int s, int divergent, uint64_t seed,
                     void* stream) {
    if (s < 1 || s > 16 || n_elems < 0 || n_elems / s > kMaxStructs) return TROVE_ERR_INVALID;
    if (n_elems == 0) return TROVE_OK;
    if (!src || !dst) return TROVE_ERR_INVALID;
    const int w = word_bytes_of(4 * s);
    if (!aligned(src, w) || !aligned(dst, w)) return TROVE_ERR_ALIGN;
    return block_copy_dispatch(static_cast<const uint32_t*>(src), static_cast<uint32_t*>(dst), n_elems, s,
                               divergent, seed, static_cast<hipStream_t>(stream));
}

int trove_block_write(void* dst, int64_t n_arrays, int s, void* stream) {
    if (s < 1 || s > 16 || n_arrays < 0 || n_arrays > kMaxStructs) return TROVE_ERR_INVALID;
    if (n_arrays == 0) return TROVE_OK;
    if (!dst) return TROVE_ERR_INVALID;
    if (!aligned(dst, word_bytes_of(4 * s))) return TROVE_ERR_ALIGN;
    return block_write_dispatch(static_cast<uint32_t*>(dst), n_arrays, s, static_cast<hipStream_t>(stream));
}

int trove_memcpy_ceiling(const void* src, void* dst, int64_t bytes, void* stream) {
    if (bytes < 0 || bytes % 16 != 0) return TROVE_ERR_INVALID;
    if (bytes == 0) return TROVE_OK;
    if (!src || !dst) return TROVE_ERR_INVALID;
    if (!aligned(src, 16) || !aligned(dst, 16)) return TROVE_ERR_ALIGN;
    return memcpy_ceiling_launch(src, dst, bytes, static_cast<hipStream_t>(stream));
}

int trove_random_lines_ceiling(const void* buf, int64_t n_lines, const uint32_t* line_idx, uint32_t* out,
                               int64_t n_reads, void* stream) {
    if (n_lines <= 0 || n_lines > 0xffffffffLL || n_reads < 0 || n_reads > kMaxStructs) return TROVE_ERR_INVALID;
    if (n_reads == 0) return TROVE_OK;
    if (!buf || !line_idx || !out) return TROVE_ERR_INVALID;
    if (!aligned(buf, 128) || !aligned(line_idx, 4) || !aligned(out, 4)) return TROVE_ERR_ALIGN;
    return random_lines_launch(buf, n_lines, line_idx, out, n_reads, static_cast<hipStream_t>(stream));
}

int trove_random_rows_write_ceiling(void* buf, int64_t n_rows, int row_bytes, const int32_t* row_idx,
                                    int64_t n_writes, int nontemporal, void* stream) {
    if (row_bytes != 32 && row_bytes != 60 && row_bytes != 64 && row_bytes != 128) return TROVE_ERR_SIZE;
    if (n_rows <= 0 || n_rows > 0xffffffffLL || n_writes < 0 || n_writes > kMaxStructs) return TROVE_ERR_INVALID;
    if (n_writes == 0) return TROVE_OK;
    if (!buf || !row_idx) return TROVE_ERR_INVALID;
    if (!aligned(buf, row_bytes % 16 == 0 ? 16 : 4) || !aligned(row_idx, 4)) return TROVE_ERR_ALIGN;
    return random_rows_write_launch(buf, n_rows, row_bytes, row_idx, n_writes, nontemporal != 0,
                                    static_cast<hipStream_t>(stream));
}

int trove_gen_words(uint32_t* out, int64_t n, uint64_t seed) {
    if (n < 0 || (n > 0 && !out)) return TROVE_ERR_INVALID;
    for (int64_t i = 0; i < n; ++i) out[i] = host_word_hash(seed, (uint64_t)i);
    return TROVE_OK;
}

int trove_fill_words_device(uint32_t* out, int64_t n, uint64_t seed, void* stream) {
    if (n < 0) return TROVE_ERR_INVALID;
    if (n == 0) return TROVE_OK;
    if (!out) return TROVE_ERR_INVALID;
    if (!aligned(out, 4)) return TROVE_ERR_ALIGN;
    return fill_words_launch(out, n, seed, static_cast<hipStream_t>(stream));
}

}  // extern "C"
