// Batched jobs: many independent gathers dst_j[i] = src_j[idx_j[i]] and
// copies dst_j[i] = src_j[i] (idx_j == nullptr) of different struct sizes and
// lengths in ONE launch.
//
// Work is cut into items of `chunk_groups` 64-struct groups of one job; item
// t belongs to the job j with first_item[j] <= t < first_item[j + 1]. Every
// wavefront walks the item space with a grid stride, finds its job by a
// binary search over the (wave-uniform, scalar-loaded) job table and runs the
// job's variant for its struct size. The launch can be captured once into a
// hipGraph and replayed. The looped mode launches the standalone gather or
// copy kernel once per job on the same stream, for comparison at equal
// output.
#include "batch.h"
#include "gather_body.h"

#include <algorithm>
#include <mutex>
#include <new>
#include <vector>

namespace trove_amd {

#define TROVE_BATCH_SIZES(X) \
    X(4) X(8) X(12) X(16) X(20) X(24) X(28) X(32) X(36) X(40) X(44) X(48) X(52) X(56) X(60) X(64)

// Copy job items: whole 64-struct groups through load/store_warp_contiguous
// (SHFL) or per lane (DIRECT); the partial last group per lane, so that the
// batch kernel carries no tile cascade.
template <typename T>
__device__ __forceinline__ void copy_range(const T* __restrict__ src, T* __restrict__ dst, int64_t n, int64_t g0,
                                           int64_t g1, bool shfl) {
    const int lane = trove::lane_id();
    const int64_t full = n >> 6;
    for (int64_t g = g0; g < g1; ++g) {
        const int64_t i = (g << 6) + lane;
        if (g < full && shfl) {
            const T v = trove::load_warp_contiguous<64, trove::streaming>(src + i);
            trove::store_warp_contiguous<64, trove::streaming>(v, dst + i);
        } else if (i < n) {
            trove::detail::divergent_store(trove::detail::divergent_load(src + i), dst + i);
        }
    }
}

// Job pointers come from the job table in memory, so the compiler cannot
// tell their address space and would use flat accesses (whose completion
// every LDS wait of the transposition and the index broadcast also covers).
// The C ABI takes device memory only, so each is rebuilt from its address as
// a global-memory pointer (an integer-to-pointer conversion into address
// space 1; a pointer-to-pointer round trip would be folded away).
template <typename P>
__device__ __forceinline__ P* assume_global(P* p) {
    typedef __attribute__((address_space(1))) P global_p;
    return (P*)(global_p*)reinterpret_cast<uintptr_t>(p);
}

// LDS-staged copy items (load/store_warp_contiguous_lds); the partial last
// group per lane.
template <typename T>
__device__ __forceinline__ void copy_range_lds(const T* __restrict__ src, T* __restrict__ dst, int64_t n, int64_t g0,
                                               int64_t g1, uint32_t* stage) {
    const int lane = trove::lane_id();
    const int64_t full = n >> 6;
    for (int64_t g = g0; g < g1; ++g) {
        const int64_t i = (g << 6) + lane;
        if (g < full) {
            const T v = trove::load_warp_contiguous_lds<trove::streaming>(src + i, stage);
            trove::store_warp_contiguous_lds<trove::streaming>(v, dst + i, stage);
        } else if (i < n) {
            trove::detail::divergent_store(trove::detail::divergent_load(src + i), dst + i);
        }
    }
}

// Per-wavefront LDS stage of the batch kernels: 64 structs of the largest
// batch struct size (64 bytes), allocated only for tables with LDS jobs.
constexpr int kBatchStageBytes = 64 * 64;

template <int B>
__device__ __forceinline__ void batch_item(const job_desc& J0, int64_t local, uint32_t* stage, int32_t* flag) {
    row_guard guard{J0.last_row, flag};
    typedef blob_t<B> T;
    job_desc J = J0;
    J.src = assume_global(J.src);
    J.dst = assume_global(J.dst);
    if (J.idx) J.idx = assume_global(J.idx);
    const T* src = static_cast<const T*>(J.src);
    T* dst = static_cast<T*>(J.dst);
    const int64_t groups = (J.n + 63) >> 6;
    const int64_t g0 = local * J.chunk_groups;
    const int64_t g1 = g0 + J.chunk_groups < groups ? g0 + J.chunk_groups : groups;
    if (J.idx == nullptr) {
        if (J.variant == TROVE_VARIANT_LDS) {
            copy_range_lds<T>(src, dst, J.n, g0, g1, stage);
        } else {
            copy_range<T>(src, dst, J.n, g0, g1, J.variant != TROVE_VARIANT_DIRECT);
        }
        return;
    }
    if (J.variant == TROVE_VARIANT_LDS) {
        gather_range_lds<T>(src, J.idx, dst, J.n, g0, g1, stage, guard);
        return;
    }
    if (J.variant == TROVE_VARIANT_TILE) {  // every batch size is a multiple of 4 bytes
        gather_range_tile<T, gather_groups_in_flight<T>()>(src, J.idx, dst, J.n, g0, g1, guard);
        return;
    }
    if constexpr (trove::detail::use_shfl<T, 64>::value) {
        if (J.variant == TROVE_VARIANT_SHFL) {  // items of one group (trove_batch_create)
            gather_group_shfl<T>(src, J.idx, dst, J.n, g0, guard);
            return;
        }
    }
    gather_range_direct<T>(src, J.idx, dst, J.n, g0, g1, guard);
}

// Work item `it` of a job table (device memory or kernel arguments): its job
// by binary search on first_item (wave-uniform, scalar loads), then the
// job's path for its struct size.
template <typename Table>
__device__ __forceinline__ void run_item(const Table& jobs, int n_jobs, int64_t it, uint32_t* stage,
                                         int32_t* flags) {
    int lo = 0, hi = n_jobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].first_item <= it) {
            lo = mid;
        } else {
            hi = mid - 1;
        }
    }
    const job_desc J = jobs[lo];
    const int64_t local = it - J.first_item;
    int32_t* flag = flags ? flags + lo : nullptr;
    switch (J.bytes) {
#define TROVE_BATCH_CASE(B)                   \
    case B:                                   \
        batch_item<B>(J, local, stage, flag); \
        break;
        TROVE_BATCH_SIZES(TROVE_BATCH_CASE)
#undef TROVE_BATCH_CASE
        default:
            break;
    }
}

// One work item per wavefront, no loop over items: in a loop, LICM hoists
// the per-lane tables of every (size, variant) branch and keeps them all
// live (275 VGPRs). The grid has exactly total_items wavefronts.
// The wavefront's slice of the dynamic LDS (present when the launch has LDS
// jobs; otherwise no item touches it).
__device__ __forceinline__ uint32_t* batch_stage() {
    extern __shared__ __align__(16) uint32_t lds_stage[];
    return lds_stage + (threadIdx.x >> 6) * (kBatchStageBytes / 4);
}

__global__ __launch_bounds__(kBlock) void k_batch_gather(const job_desc* __restrict__ jobs, int n_jobs,
                                                         int64_t total_items, int32_t* flags) {
    const int64_t it = wave_index();
    if (it < total_items) run_item(jobs, n_jobs, it, batch_stage(), flags);
}

__global__ __launch_bounds__(kBlock) void k_batch_gather_pack(const job_pack p) {
    const int64_t it = wave_index();
    if (it < p.total_items) run_item(p.jobs, p.n_jobs, it, batch_stage(), p.flags);
}

// Resident-grid variant: a grid of only as many wavefronts as fit on the
// device at once, each striding over the work items (t, t + W, t + 2W, ...,
// W = wavefronts in the grid). The loop lets LICM keep every size's and
// variant's per-lane tables live at once (162 VGPRs, 3 wavefronts per SIMD,
// against 94 and 5 for one item per wavefront).
__global__ __launch_bounds__(kBlock) void k_batch_resident(const job_desc* __restrict__ jobs, int n_jobs,
                                                           int64_t total_items, int32_t* flags) {
    for (int64_t it = wave_index(); it < total_items; it += wave_count()) run_item(jobs, n_jobs, it, batch_stage(), flags);
}

int launch_job_table_resident(const job_desc* d_table, int n_jobs, int64_t total_items, bool lds, int32_t* flags,
                              hipStream_t s) {
    const size_t stage = lds ? (size_t)kBatchStageBytes * kWavesPerBlock : 0;
    int per_cu = 0;
    int rc = hip_status(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_resident, kBlock, stage));
    if (rc != TROVE_OK) return rc;
    const int64_t resident = (int64_t)(per_cu > 0 ? per_cu : 1) * device_cu_count();
    const int64_t needed = (total_items + kWavesPerBlock - 1) / kWavesPerBlock;
    const int64_t blocks = needed < resident ? needed : resident;
    hipLaunchKernelGGL(k_batch_resident, dim3((unsigned)blocks), dim3(kBlock), stage, s, d_table, n_jobs,
                       total_items, flags);
    return launch_status();
}

int launch_job_table(const job_desc* d_table, int n_jobs, int64_t total_items, bool lds, int32_t* flags,
                     hipStream_t s) {
    // One wavefront per item: never capped.
    const int64_t blocks = (total_items + kWavesPerBlock - 1) / kWavesPerBlock;
    const size_t stage = lds ? (size_t)kBatchStageBytes * kWavesPerBlock : 0;
    hipLaunchKernelGGL(k_batch_gather, dim3((unsigned)blocks), dim3(kBlock), stage, s, d_table, n_jobs,
                       total_items, flags);
    return launch_status();
}

int launch_job_pack(const job_pack& p, bool lds, hipStream_t s) {
    const int64_t blocks = (p.total_items + kWavesPerBlock - 1) / kWavesPerBlock;
    const size_t stage = lds ? (size_t)kBatchStageBytes * kWavesPerBlock : 0;
    hipLaunchKernelGGL(k_batch_gather_pack, dim3((unsigned)blocks), dim3(kBlock), stage, s, p);
    return launch_status();
}

}  // namespace trove_amd

using namespace trove_amd;

struct trove_batch {
    std::vector<trove_job> jobs;
    std::vector<job_desc> table;
    job_desc* d_table = nullptr;
    int64_t total_items = 0;
    int64_t algorithmic_bytes = 0;
    bool lds = false;  // some job takes the LDS-staged path
    int device = 0;
    hipStream_t capture_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::mutex capture_mutex;  // the lazy capture of TROVE_BATCH_GRAPH, once per handle
};

static int batch_launch(trove_batch* b, hipStream_t s) {
    return launch_job_table(b->d_table, (int)b->table.size(), b->total_items, b->lds, nullptr, s);
}

// Validates one job and builds its descriptor: the variant AUTO resolves to,
// its work items (starting at first_item) and algorithmic bytes.
static int describe_job(const trove_job& J, int64_t first_item, job_desc* d, int64_t* items, int64_t* bytes_out) {
    const int bytes = J.struct_bytes;
    if (bytes < 4 || bytes > 64 || bytes % 4 != 0) return TROVE_ERR_SIZE;
    if (J.n < 0 || (J.n > 0 && (!J.src || !J.dst)) || J.reserved != 0) return TROVE_ERR_INVALID;
    if (J.variant < TROVE_VARIANT_AUTO || J.variant > TROVE_VARIANT_LDS || J.chunk_groups < 0)
        return TROVE_ERR_VARIANT;
    const int w = word_bytes_of(bytes);
    if (!aligned(J.src, w) || !aligned(J.dst, w) || !aligned(J.idx, 4)) return TROVE_ERR_ALIGN;
    *d = job_desc{};
    d->src = J.src;
    d->idx = J.idx;
    d->dst = J.dst;
    d->n = J.n;
    d->bytes = bytes;
    d->last_row = 0xffffffffu;  // batch indices are the caller's to keep in range
    const bool copy = J.idx == nullptr;
    if (copy) {
        // A copy job: SHFL (load/store_warp_contiguous), LDS
        // (load/store_warp_contiguous_lds) or DIRECT; AUTO takes the
        // library's per-size choice, TILE means SHFL.
        const int v = J.variant == TROVE_VARIANT_AUTO ? auto_copy_variant(bytes) : J.variant;
        d->variant = v == TROVE_VARIANT_TILE ? TROVE_VARIANT_SHFL : v;
    } else {
        d->variant = J.variant == TROVE_VARIANT_AUTO ? auto_gather_variant(bytes) : J.variant;
    }
    d->chunk_groups = J.chunk_groups > 0 ? J.chunk_groups : 16;
    // coalesced_ptr gather items are one group: a loop around the tile
    // cascade would keep every level's tables live.
    if (!copy && d->variant == TROVE_VARIANT_SHFL) d->chunk_groups = 1;
    d->first_item = first_item;
    const int64_t groups = (J.n + 63) / 64;
    *items = (groups + d->chunk_groups - 1) / d->chunk_groups;
    *bytes_out = (2 * (int64_t)bytes + (copy ? 0 : 4)) * J.n;
    return TROVE_OK;
}

// ---- trove_batch_run_static: the job table in a static device array ----
//
// `scratch` holds the job table of trove_batch_run_static: one array per
// device, kept for the life of the process. Host code never passes its name
// to a kernel or a copy (there it is only a host-side stand-in); it uses the
// device address hipGetSymbolAddress gives, looked up once per device with
// that device current.
static __device__ __align__(256) job_desc scratch[TROVE_BATCH_STATIC_MAX_JOBS];

constexpr int kMaxDevices = 64;
static job_desc* g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_scratch_mutex;

static int scratch_for_current_device(job_desc** out) {
    int dev = 0;
    int rc = hip_status(hipGetDevice(&dev));
    if (rc != TROVE_OK) return rc;
    if (dev < 0 || dev >= kMaxDevices) return TROVE_ERR_INVALID;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        rc = hip_status(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (rc != TROVE_OK) return rc;
        g_scratch_ptr[dev] = static_cast<job_desc*>(p);
    }
    *out = g_scratch_ptr[dev];
    return TROVE_OK;
}

// Writes up to kArgJobs descriptors, passed in the kernel arguments, to
// table[first .. first + p.n_jobs).
__global__ __launch_bounds__(64) void k_store_jobs(const job_pack p, int first, job_desc* table) {
    const int j = (int)threadIdx.x;
    if (j < p.n_jobs) table[first + j] = p.jobs[j];
}

// Captures the one-launch mode into b->graph / b->graph_exec once. On any
// failure the capture stream and graph are destroyed again, so that a later
// call starts clean.
static int capture_graph(trove_batch* b) {
    int rc = hip_status(hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking));
    if (rc == TROVE_OK) {
        rc = hip_status(hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal));
        if (rc == TROVE_OK) {
            const int lrc = batch_launch(b, b->capture_stream);
            rc = hip_status(hipStreamEndCapture(b->capture_stream, &b->graph));
            if (rc == TROVE_OK) rc = lrc;
        }
    }
    if (rc == TROVE_OK) rc = hip_status(hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0));
    if (rc != TROVE_OK) {
        if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        b->graph_exec = nullptr;
        b->graph = nullptr;
    }
    if (b->capture_stream) (void)hipStreamDestroy(b->capture_stream);
    b->capture_stream = nullptr;
    return rc;
}

extern "C" {

int trove_batch_create(const trove_job* jobs, int32_t n_jobs, trove_batch** out) {
    if (!out || n_jobs <= 0 || !jobs) return TROVE_ERR_INVALID;
    *out = nullptr;
    trove_batch* b = new (std::nothrow) trove_batch();
    if (!b) return TROVE_ERR_NOMEM;
    b->jobs.assign(jobs, jobs + n_jobs);
    int64_t item = 0;
    for (int j = 0; j < n_jobs; ++j) {
        job_desc d{};
        int64_t items = 0, bytes = 0;
        const int rc = describe_job(jobs[j], item, &d, &items, &bytes);
        if (rc != TROVE_OK) {
            delete b;
            return rc;
        }
        b->lds = b->lds || d.variant == TROVE_VARIANT_LDS;
        item += items;
        b->algorithmic_bytes += bytes;
        b->table.push_back(d);
    }
    b->total_items = item;
    int rc = hip_status(hipGetDevice(&b->device));
    if (rc == TROVE_OK) rc = hip_status(hipMalloc(&b->d_table, sizeof(job_desc) * b->table.size()));
    if (rc == TROVE_OK)
        rc = hip_status(
            hipMemcpy(b->d_table, b->table.data(), sizeof(job_desc) * b->table.size(), hipMemcpyHostToDevice));
    if (rc != TROVE_OK) {
        if (b->d_table) (void)hipFree(b->d_table);
        delete b;
        return rc;
    }
    *out = b;
    return TROVE_OK;
}

int trove_batch_run(trove_batch* b, int mode, void* stream) {
    if (!b) return TROVE_ERR_INVALID;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (b->total_items == 0) return TROVE_OK;
    if (mode == TROVE_BATCH_LAUNCH) return batch_launch(b, s);
    if (mode == TROVE_BATCH_RESIDENT)
        return launch_job_table_resident(b->d_table, (int)b->table.size(), b->total_items, b->lds, nullptr, s);
    if (mode == TROVE_BATCH_GRAPH) {
        hipGraphExec_t exec = nullptr;
        {
            std::lock_guard<std::mutex> lock(b->capture_mutex);
            if (!b->graph_exec) {
                const int rc = capture_graph(b);
                if (rc != TROVE_OK) return rc;
            }
            exec = b->graph_exec;
        }
        return hip_status(hipGraphLaunch(exec, s));
    }
    if (mode == TROVE_BATCH_LOOPED) {
        for (const job_desc& d : b->table) {
            if (d.n == 0) continue;
            const int rc = d.idx ? trove_gather(d.src, d.idx, d.dst, d.n, d.bytes, d.variant, 64, stream)
                                 : trove_copy(d.src, d.dst, d.n, d.bytes, d.variant, 64, stream);
            if (rc != TROVE_OK) return rc;
        }
        return TROVE_OK;
    }
    return TROVE_ERR_INVALID;
}

int trove_batch_run_static(const trove_job* jobs, int32_t n_jobs, void* stream) {
    if (!jobs || n_jobs <= 0) return TROVE_ERR_INVALID;
    if (n_jobs > TROVE_BATCH_STATIC_MAX_JOBS) return TROVE_ERR_FULL;
    // Validate every job before anything is written to the device.
    int64_t total_items = 0;
    bool lds = false;
    for (int j = 0; j < n_jobs; ++j) {
        job_desc d;
        int64_t items = 0, bytes = 0;
        const int rc = describe_job(jobs[j], total_items, &d, &items, &bytes);
        if (rc != TROVE_OK) return rc;
        total_items += items;
        lds = lds || d.variant == TROVE_VARIANT_LDS;
    }
    job_desc* table = nullptr;
    int rc = scratch_for_current_device(&table);
    if (rc != TROVE_OK) return rc;
    if (total_items == 0) return TROVE_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    // Descriptors travel to the table in the kernel arguments of one small
    // launch per 64 jobs (stream-ordered before the batch launch).
    job_pack p;
    p.pad = 0;
    p.total_items = 0;
    p.flags = nullptr;
    int64_t item = 0;
    for (int first = 0; first < n_jobs; first += kArgJobs) {
        const int cnt = n_jobs - first < kArgJobs ? n_jobs - first : kArgJobs;
        for (int j = 0; j < cnt; ++j) {
            int64_t items = 0, bytes = 0;
            (void)describe_job(jobs[first + j], item, &p.jobs[j], &items, &bytes);
            item += items;
        }
        p.n_jobs = cnt;
        hipLaunchKernelGGL(k_store_jobs, dim3(1), dim3(64), 0, s, p, first, table);
        rc = launch_status();
        if (rc != TROVE_OK) return rc;
    }
    return launch_job_table(table, n_jobs, total_items, lds, nullptr, s);
}

int trove_batch_stats(const trove_batch* b, int64_t* work_items, int64_t* algorithmic_bytes) {
    if (!b) return TROVE_ERR_INVALID;
    if (work_items) *work_items = b->total_items;
    if (algorithmic_bytes) *algorithmic_bytes = b->algorithmic_bytes;
    return TROVE_OK;
}

int trove_batch_destroy(trove_batch* b) {
    if (!b) return TROVE_ERR_INVALID;
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->d_table) (void)hipFree(b->d_table);
    delete b;
    return TROVE_OK;
}

}  // extern "C"
