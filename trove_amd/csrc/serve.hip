// Continuously batched gather server over a device-resident struct table.
// (Device-resident request indices are read in place: the caller keeps them
// valid until the request is done. Host indices are range-checked and copied
// at submission, so the caller's array may change right after.)
//
// No reference counterpart: the reference is a device library. This engine
// serves requests "gather these rows of the table" with:
//  * admission by an HBM budget: every admitted request owns a contiguous
//    range of rows in a device arena (its output rows and, at the same row
//    offsets in a second arena, its indices); requests wait in FIFO order
//    until a request slot and a range that fits are free;
//  * continuous batching: each step gathers the next slice (at most
//    `slice_indices`) of every active request in ONE launch of the
//    batch-gather kernel, so requests admitted while others are
//    still on the GPU join the next step and their work overlaps on the
//    device;
//  * a step of up to 64 requests passes its job table in the kernel
//    arguments; larger steps upload it from pinned memory first;
//  * pipelining: up to kInflight steps are queued on the compute stream;
//    host index uploads run on a separate copy stream and the compute stream
//    waits on them by event;
//  * completion: a request is done when the step holding its last slice has
//    finished (event query); its output stays in the arena until released;
//  * index safety: every step's gather checks each index against the table
//    (an index outside it reads row 0 instead) and sets the job's flag in
//    pinned host memory; a request with a flagged slice completes as
//    rejected (TROVE_ERR_RANGE) and its rows are never handed out.
// The host side is single-threaded: the caller drives it with step() (or
// replay()), and every call is non-blocking except drain().
#include "batch.h"

#include <chrono>
#include <deque>
#include <map>
#include <memory>
#include <new>
#include <vector>

namespace trove_amd {

namespace {

// Steps queued on the compute stream at most. Fewer steps in flight make
// each step take more of the requests that arrived meanwhile instead of
// queueing behind small steps already launched. Measured on MI355X (2^24 x
// 32-B table, Zipf requests, interleaved runs): against four, two steps
// lower p50 at 1 / 1.5 M requests/s from 0.054-0.060 / 0.083-0.098 ms to
// 0.033-0.035 / 0.046-0.057 ms with saturation unchanged (1.99-2.08 M
// requests/s); one step costs 2-3% of saturation.
constexpr int kInflight = 2;

double now_s() {
    using clk = std::chrono::steady_clock;
    return std::chrono::duration<double>(clk::now().time_since_epoch()).count();
}

// First-fit allocator of row ranges in [0, capacity).
class range_allocator {
  public:
    explicit range_allocator(int64_t capacity) { free_[0] = capacity; }
    int64_t alloc(int64_t len) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= len) {
                const int64_t start = it->first, rest = it->second - len;
                free_.erase(it);
                if (rest > 0) free_[start + len] = rest;
                return start;
            }
        }
        return -1;
    }
    void release(int64_t start, int64_t len) {
        auto next = free_.lower_bound(start);
        if (next != free_.end() && start + len == next->first) {
            len += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == start) {
                prev->second += len;
                return;
            }
        }
        free_[start] = len;
    }

  private:
    std::map<int64_t, int64_t> free_;
};

enum req_state { QUEUED = 0, ACTIVE = 1, DONE = 2 };

struct request {
    uint64_t id = 0;
    const int32_t* idx = nullptr;  // caller's device indices (host indices: host_idx)
    std::unique_ptr<int32_t[]> host_idx;  // copy of host indices, taken at submission
    bool idx_on_device = false;
    bool live = false;       // the slot holds a request (request_table)
    bool bad_index = false;  // a step flagged an index outside the table
    int64_t n = 0;
    int64_t rows = 0;       // arena rows reserved (>= 1)
    int64_t row0 = -1;      // arena offset
    int64_t scheduled = 0;  // indices handed to steps so far
    int64_t stage0 = -1;    // host indices uploaded at submission: their offset in the staging pool
    int state = QUEUED;
    double t_submit = 0, t_done = 0;
    const int32_t* src_idx = nullptr;  // indices the kernel reads (caller's or arena copy)
};

// Requests live in the slots of one vector, and a released request's slot is
// reused. A request id is (generation << 32) | slot: ids stay unique across
// reuse (until one slot has been reused 2^32 times), and a lookup is an index
// plus an id check, with no allocation per request once the vector has grown
// to the number of live requests. (Measured on MI355X's host: with a hash map
// of requests, submitting and completing a request cost about 1 us of host
// time, enough to saturate the host loop near 1 M requests/s on a slow core.)
class request_table {
  public:
    request* find(uint64_t id) {
        const uint64_t k = id & 0xffffffffu;
        return k < slots_.size() && slots_[k].live && slots_[k].id == id ? &slots_[k] : nullptr;
    }
    const request* find(uint64_t id) const { return const_cast<request_table*>(this)->find(id); }
    // The request of an id known to be live.
    request& operator[](uint64_t id) { return slots_[id & 0xffffffffu]; }
    // A fresh request in a free slot. May move the slots: references to
    // requests do not survive it.
    request& create() {
        uint32_t k;
        if (!free_.empty()) {
            k = free_.back();
            free_.pop_back();
        } else {
            k = (uint32_t)slots_.size();
            slots_.emplace_back();
        }
        request& r = slots_[k];
        uint32_t gen = (uint32_t)(r.id >> 32) + 1;
        if (gen == 0) gen = 1;
        r = request();
        r.id = ((uint64_t)gen << 32) | k;
        r.live = true;
        return r;
    }
    void erase(request& r) {
        r.live = false;
        r.host_idx.reset();
        free_.push_back((uint32_t)(r.id & 0xffffffffu));
    }
    // Room for n more requests without moving the slots.
    void reserve_more(size_t n) {
        const size_t want = slots_.size() - free_.size() + n;
        if (want > slots_.capacity()) slots_.reserve(want);
        free_.reserve(slots_.capacity());
    }
    // Upper bound (exclusive) of the slot numbers of ids created before the
    // slots next grow beyond their reserved capacity.
    size_t slot_bound() const { return slots_.capacity(); }

  private:
    std::vector<request> slots_;
    std::vector<uint32_t> free_;
};

struct step_rec {
    hipEvent_t done = nullptr;
    std::vector<uint64_t> finishing;
    std::vector<uint64_t> job_req;  // request id of each job of the step
    bool busy = false;
};

}  // namespace

}  // namespace trove_amd

using namespace trove_amd;

struct trove_serve {
    const void* table = nullptr;
    int64_t n_rows = 0;
    int32_t row_bytes = 0;
    int64_t capacity_rows = 0;
    int32_t max_slots = 0;
    int64_t slice = 0;
    int32_t variant = TROVE_VARIANT_TILE;
    hipStream_t stream = nullptr;       // compute (caller's stream)
    hipStream_t copy_stream = nullptr;  // index uploads
    uint8_t* out_arena = nullptr;
    int32_t* idx_arena = nullptr;
    range_allocator* arena = nullptr;
    hipEvent_t upload_done = nullptr;  // last host-index upload pass
    // Device staging pool for host indices uploaded at submission (see
    // ensure_dstage); a request that finds it full keeps a host copy and
    // uploads it at admission.
    int32_t* dstage = nullptr;
    range_allocator* dstage_alloc = nullptr;
    bool dstage_tried = false;
    bool upload_pending = false;  // submission uploads not yet ordered before a step
    std::deque<uint64_t> queue;   // QUEUED, FIFO
    std::vector<uint64_t> active;  // ACTIVE, admission order
    std::vector<uint64_t> still_active;  // launch_step's next `active` (kept for its capacity)
    request_table reqs;
    // step ring
    step_rec steps[kInflight];
    job_desc* h_jobs[kInflight] = {};  // pinned
    job_desc* d_jobs[kInflight] = {};
    int32_t* h_flags[kInflight] = {};  // pinned, coherent: per-job bad-index flags the step kernel sets
    int32_t* d_flags[kInflight] = {};  // the same memory as the device addresses it
    int max_jobs = 0;
    job_pack pack = {};  // kernel-argument job table of small steps
    uint64_t step_seq = 0;  // next step slot = step_seq % kInflight
    uint64_t retired_seq = 0;
    std::vector<uint64_t> newly_done;  // completions not yet seen by replay (recorded while it runs)
    bool replaying = false;
    // statistics
    int64_t completed = 0, indices_done = 0, steps_launched = 0;
};

static void free_serve(trove_serve* s) {
    for (int k = 0; k < kInflight; ++k) {
        if (s->steps[k].done) (void)hipEventDestroy(s->steps[k].done);
        if (s->h_jobs[k]) (void)hipHostFree(s->h_jobs[k]);
        if (s->h_flags[k]) (void)hipHostFree(s->h_flags[k]);
        if (s->d_jobs[k]) (void)hipFree(s->d_jobs[k]);
    }
    if (s->upload_done) (void)hipEventDestroy(s->upload_done);
    if (s->copy_stream) (void)hipStreamDestroy(s->copy_stream);
    if (s->out_arena) (void)hipFree(s->out_arena);
    if (s->idx_arena) (void)hipFree(s->idx_arena);
    if (s->dstage) (void)hipFree(s->dstage);
    delete s->dstage_alloc;
    delete s->arena;
    delete s;
}

// Marks the requests of every finished step done, oldest first.
static int retire(trove_serve* s) {
    while (s->retired_seq < s->step_seq) {
        step_rec& st = s->steps[s->retired_seq % kInflight];
        const hipError_t q = hipEventQuery(st.done);
        if (q == hipErrorNotReady) break;
        if (q != hipSuccess) return hip_status(q);
        const double t = now_s();
        const volatile int32_t* flags = s->h_flags[s->retired_seq % kInflight];
        for (size_t j = 0; j < st.job_req.size(); ++j)
            if (flags[j]) s->reqs[st.job_req[j]].bad_index = true;
        st.job_req.clear();
        for (uint64_t id : st.finishing) {
            request& r = s->reqs[id];
            r.state = DONE;
            r.t_done = t;
            r.host_idx.reset();  // uploaded and gathered: no longer needed
            if (r.stage0 >= 0) {
                s->dstage_alloc->release(r.stage0, r.n);
                r.stage0 = -1;
            }
            s->completed++;
            s->indices_done += r.n;
            if (s->replaying) s->newly_done.push_back(id);
        }
        st.finishing.clear();
        st.busy = false;
        s->retired_seq++;
    }
    return TROVE_OK;
}

// Admits queued requests in FIFO order while a slot and an arena range fit.
// Device-resident indices are read in place; host indices are uploaded into
// the index arena on the copy stream, and one event after the pass orders
// them before the next step's kernel.
static int admit(trove_serve* s) {
    bool uploaded = false;
    while (!s->queue.empty() && (int)s->active.size() < s->max_slots) {
        request& r = s->reqs[s->queue.front()];
        const int64_t row0 = s->arena->alloc(r.rows);
        if (row0 < 0) break;
        r.row0 = row0;
        if (r.idx_on_device) {
            r.src_idx = r.idx;
        } else if (r.stage0 >= 0) {
            r.src_idx = s->dstage + r.stage0;  // uploaded at submission
        } else {
            int32_t* dst = s->idx_arena + row0;
            const int rc = hip_status(hipMemcpyAsync(dst, r.host_idx.get(), sizeof(int32_t) * r.n,
                                                     hipMemcpyHostToDevice, s->copy_stream));
            if (rc != TROVE_OK) return rc;
            r.src_idx = dst;
            uploaded = true;
        }
        r.state = ACTIVE;
        s->active.push_back(r.id);
        s->queue.pop_front();
    }
    if (uploaded || s->upload_pending) {
        // One event after every upload issued so far (at submission and here)
        // orders them before the next step's kernel.
        s->upload_pending = false;
        int rc = hip_status(hipEventRecord(s->upload_done, s->copy_stream));
        if (rc == TROVE_OK) rc = hip_status(hipStreamWaitEvent(s->stream, s->upload_done, 0));
        if (rc != TROVE_OK) return rc;
    }
    return TROVE_OK;
}

// Builds and launches one step: the next slice of every active request.
static int launch_step(trove_serve* s) {
    if (s->step_seq - s->retired_seq >= (uint64_t)kInflight) return TROVE_OK;  // ring full
    const int k = (int)(s->step_seq % kInflight);
    step_rec& st = s->steps[k];
    job_desc* jobs = s->h_jobs[k];
    const int32_t last_row = (int32_t)(s->n_rows - 1);
    int nj = 0;
    int64_t items = 0;
    std::vector<uint64_t>& still_active = s->still_active;
    still_active.clear();
    for (uint64_t id : s->active) {
        request& r = s->reqs[id];
        const int64_t left = r.n - r.scheduled;
        if (nj >= s->max_jobs) {
            still_active.push_back(id);
            continue;
        }
        const int64_t cnt = left < s->slice ? left : s->slice;
        job_desc& d = jobs[nj++];
        d.src = s->table;
        d.idx = r.src_idx + r.scheduled;
        d.dst = s->out_arena + (r.row0 + r.scheduled) * (int64_t)s->row_bytes;
        d.n = cnt;
        d.bytes = s->row_bytes;
        d.variant = s->variant;
        // Work items of two 64-row groups: a step lasts a few hundred
        // microseconds, and small items shorten its tail. Measured on MI355X
        // (2^24 x 32-B table, Zipf requests): against eight groups per item,
        // saturation +3% and p50 / p99 at 80-95% load 20-30% lower; one group
        // per item costs 7% of saturation throughput.
        d.chunk_groups = s->variant == TROVE_VARIANT_SHFL ? 1 : 2;
        d.last_row = (uint32_t)last_row;
        d.first_item = items;
        st.job_req.push_back(id);
        items += ((cnt + 63) / 64 + d.chunk_groups - 1) / d.chunk_groups;
        r.scheduled += cnt;
        if (r.scheduled == r.n)
            st.finishing.push_back(id);
        else
            still_active.push_back(id);
    }
    if (nj == 0) return TROVE_OK;
    s->active.swap(still_active);
    // The slot's previous step has retired: its flags are free to clear.
    for (int j = 0; j < nj; ++j) s->h_flags[k][j] = 0;
    int rc = TROVE_OK;
    if (nj <= kArgJobs) {
        // Small steps carry their jobs in the kernel arguments: no upload.
        job_pack& p = s->pack;
        for (int j = 0; j < nj; ++j) p.jobs[j] = jobs[j];
        p.n_jobs = nj;
        p.pad = 0;
        p.total_items = items;
        p.flags = s->d_flags[k];
        rc = launch_job_pack(p, s->variant == TROVE_VARIANT_LDS, s->stream);
    } else {
        rc = hip_status(hipMemcpyAsync(s->d_jobs[k], jobs, sizeof(job_desc) * nj, hipMemcpyHostToDevice, s->stream));
        if (rc == TROVE_OK)
            rc = launch_job_table(s->d_jobs[k], nj, items, s->variant == TROVE_VARIANT_LDS, s->d_flags[k], s->stream);
    }
    if (rc == TROVE_OK) rc = hip_status(hipEventRecord(st.done, s->stream));
    if (rc != TROVE_OK) return rc;
    st.busy = true;
    s->step_seq++;
    s->steps_launched++;
    return TROVE_OK;
}

static int serve_step(trove_serve* s) {
    int rc = retire(s);
    if (rc == TROVE_OK) rc = admit(s);
    if (rc == TROVE_OK) rc = launch_step(s);
    return rc;
}

// Whether every src[i], i < n, lies in [0, n_rows): one read-only pass with
// no early exit, so the compiler vectorises it (an unsigned compare also
// catches negative indices).
static bool all_in_range(const int32_t* __restrict__ src, int64_t n, int64_t n_rows) {
    const uint32_t lim = (uint32_t)n_rows;  // n_rows <= INT32_MAX (trove_serve_create)
    uint32_t bad = 0;
    for (int64_t i = 0; i < n; ++i) bad |= (uint32_t)((uint32_t)src[i] >= lim);
    return bad == 0;
}

// The device staging pool, allocated at the first host-index submission
// (servers fed device indices never hold it): as many entries as the output
// budget has rows, at most kMaxStageInts. Without it (allocation refused)
// every host-index request takes the host-copy path.
constexpr int64_t kMaxStageInts = (int64_t)1 << 28;  // 1 GiB of device memory
static void ensure_dstage(trove_serve* s) {
    if (s->dstage_tried) return;
    s->dstage_tried = true;
    const int64_t ints = s->capacity_rows < kMaxStageInts ? s->capacity_rows : kMaxStageInts;
    if (hipMalloc(&s->dstage, (size_t)ints * sizeof(int32_t)) != hipSuccess) {
        s->dstage = nullptr;
        (void)hipGetLastError();
        return;
    }
    s->dstage_alloc = new (std::nothrow) range_allocator(ints);
    if (!s->dstage_alloc) {
        (void)hipFree(s->dstage);
        s->dstage = nullptr;
    }
}

// dst[i] = src[i] for i < n; false when some index lies outside [0, n_rows).
// One pass with no early exit, so the compiler vectorises it (an unsigned
// compare also catches negative indices); the caller discards dst on false.
static bool copy_checked(int32_t* __restrict__ dst, const int32_t* __restrict__ src, int64_t n, int64_t n_rows) {
    const uint32_t lim = (uint32_t)n_rows;  // n_rows <= INT32_MAX (trove_serve_create)
    uint32_t bad = 0;
    for (int64_t i = 0; i < n; ++i) {
        const int32_t v = src[i];
        dst[i] = v;
        bad |= (uint32_t)((uint32_t)v >= lim);
    }
    return bad == 0;
}

// Zero-length requests are done at submission.
static void finish_empty(trove_serve* s, request& r) {
    r.state = DONE;
    r.t_done = now_s();
    s->completed++;
}

extern "C" {

int trove_serve_create(const void* table, int64_t n_rows, int32_t row_bytes, int64_t budget_bytes,
                       int32_t max_slots, int64_t slice_indices, int32_t variant, void* stream,
                       trove_serve** out) {
    if (!out) return TROVE_ERR_INVALID;
    *out = nullptr;
    if (!table || n_rows <= 0 || n_rows > INT32_MAX) return TROVE_ERR_INVALID;
    if (row_bytes < 4 || row_bytes > 64 || row_bytes % 4 != 0) return TROVE_ERR_SIZE;
    if (budget_bytes < row_bytes || max_slots <= 0 || slice_indices < 64) return TROVE_ERR_INVALID;
    if (variant < TROVE_VARIANT_AUTO || variant > TROVE_VARIANT_LDS) return TROVE_ERR_VARIANT;
    if (!aligned(table, word_bytes_of(row_bytes))) return TROVE_ERR_ALIGN;
    trove_serve* s = new (std::nothrow) trove_serve();
    if (!s) return TROVE_ERR_NOMEM;
    s->table = table;
    s->n_rows = n_rows;
    s->row_bytes = row_bytes;
    s->capacity_rows = budget_bytes / row_bytes;
    s->max_slots = max_slots;
    s->slice = slice_indices;
    s->variant = variant == TROVE_VARIANT_AUTO ? auto_gather_variant(row_bytes) : variant;
    s->stream = static_cast<hipStream_t>(stream);
    s->max_jobs = max_slots;
    s->arena = new (std::nothrow) range_allocator(s->capacity_rows);
    int rc = s->arena ? TROVE_OK : TROVE_ERR_NOMEM;
    if (rc == TROVE_OK) rc = hip_status(hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking));
    if (rc == TROVE_OK) rc = hip_status(hipEventCreateWithFlags(&s->upload_done, hipEventDisableTiming));
    if (rc == TROVE_OK) rc = hip_status(hipMalloc(&s->out_arena, (size_t)(s->capacity_rows * row_bytes)));
    if (rc == TROVE_OK) rc = hip_status(hipMalloc(&s->idx_arena, (size_t)(s->capacity_rows * sizeof(int32_t))));
    for (int k = 0; k < kInflight && rc == TROVE_OK; ++k) {
        rc = hip_status(hipEventCreateWithFlags(&s->steps[k].done, hipEventDisableTiming));
        if (rc == TROVE_OK) rc = hip_status(hipHostMalloc((void**)&s->h_jobs[k], sizeof(job_desc) * s->max_jobs));
        if (rc == TROVE_OK) rc = hip_status(hipMalloc(&s->d_jobs[k], sizeof(job_desc) * s->max_jobs));
        if (rc == TROVE_OK)
            rc = hip_status(hipHostMalloc((void**)&s->h_flags[k], sizeof(int32_t) * s->max_jobs,
                                          hipHostMallocMapped | hipHostMallocCoherent));
        if (rc == TROVE_OK) rc = hip_status(hipHostGetDevicePointer((void**)&s->d_flags[k], s->h_flags[k], 0));
    }
    if (rc != TROVE_OK) {
        free_serve(s);
        return rc == TROVE_ERR_HIP && trove_last_hip_error() == (int)hipErrorOutOfMemory ? TROVE_ERR_NOMEM : rc;
    }
    *out = s;
    return TROVE_OK;
}

int trove_serve_submit(trove_serve* s, const int32_t* idx, int64_t n, int idx_on_device, uint64_t* request_id) {
    if (!s || !request_id || n < 0 || (n > 0 && !idx)) return TROVE_ERR_INVALID;
    if (n > s->capacity_rows) return TROVE_ERR_BUDGET;
    std::unique_ptr<int32_t[]> copy;
    int64_t stage0 = -1;
    if (!idx_on_device) {
        // Host indices are checked, then uploaded to the staging pool right
        // away (an upload from pageable memory has read the caller's array
        // when it returns), or, with the pool full, checked and copied to the
        // host in one pass; either way later changes to the caller's array
        // (or its release) cannot reach the gather.
        if (n > 0) {
            ensure_dstage(s);
            if (s->dstage_alloc && (stage0 = s->dstage_alloc->alloc(n)) >= 0) {
                if (!all_in_range(idx, n, s->n_rows)) {
                    s->dstage_alloc->release(stage0, n);
                    return TROVE_ERR_RANGE;
                }
                const int rc = hip_status(hipMemcpyAsync(s->dstage + stage0, idx, sizeof(int32_t) * n,
                                                         hipMemcpyHostToDevice, s->copy_stream));
                if (rc != TROVE_OK) {
                    s->dstage_alloc->release(stage0, n);
                    return rc;
                }
                s->upload_pending = true;
            } else {
                copy.reset(new (std::nothrow) int32_t[n]);
                if (!copy) return TROVE_ERR_NOMEM;
                if (!copy_checked(copy.get(), idx, n, s->n_rows)) return TROVE_ERR_RANGE;
            }
        }
    } else if (!aligned(idx, 4)) {
        return TROVE_ERR_ALIGN;
    }
    request& r = s->reqs.create();
    r.idx_on_device = idx_on_device != 0;
    if (r.idx_on_device) {
        r.idx = idx;
    } else {
        r.host_idx = std::move(copy);
        r.stage0 = stage0;
    }
    r.n = n;
    r.rows = n > 0 ? n : 1;
    r.t_submit = now_s();
    *request_id = r.id;
    if (n == 0) {
        finish_empty(s, r);
        return TROVE_OK;
    }
    s->queue.push_back(r.id);
    return TROVE_OK;
}

int trove_serve_step(trove_serve* s, int32_t* in_flight) {
    if (!s) return TROVE_ERR_INVALID;
    const int rc = serve_step(s);
    if (in_flight) *in_flight = (int32_t)(s->queue.size() + s->active.size() + (s->step_seq - s->retired_seq));
    return rc;
}

int trove_serve_poll(trove_serve* s, uint64_t request_id, int32_t* done, double* latency_s) {
    if (!s || !done) return TROVE_ERR_INVALID;
    const request* r = s->reqs.find(request_id);
    if (!r) return TROVE_ERR_NOTFOUND;
    *done = r->state == DONE;
    if (latency_s) *latency_s = r->state == DONE ? r->t_done - r->t_submit : -1.0;
    return r->state == DONE && r->bad_index ? TROVE_ERR_RANGE : TROVE_OK;
}

int trove_serve_output(trove_serve* s, uint64_t request_id, void** dev_out) {
    if (!s || !dev_out) return TROVE_ERR_INVALID;
    const request* r = s->reqs.find(request_id);
    if (!r) return TROVE_ERR_NOTFOUND;
    if (r->state != DONE) return TROVE_ERR_AGAIN;
    if (r->bad_index) return TROVE_ERR_RANGE;
    *dev_out = r->row0 >= 0 ? s->out_arena + r->row0 * (int64_t)s->row_bytes : nullptr;
    return TROVE_OK;
}

int trove_serve_rows(const trove_serve* s, uint64_t request_id, int64_t* n) {
    if (!s || !n) return TROVE_ERR_INVALID;
    const request* r = s->reqs.find(request_id);
    if (!r) return TROVE_ERR_NOTFOUND;
    *n = r->n;
    return TROVE_OK;
}

int trove_serve_fetch(trove_serve* s, uint64_t request_id, void* out) {
    if (!s || !out) return TROVE_ERR_INVALID;
    const request* r = s->reqs.find(request_id);
    if (!r) return TROVE_ERR_NOTFOUND;
    if (r->state != DONE) return TROVE_ERR_AGAIN;
    if (r->bad_index) return TROVE_ERR_RANGE;
    if (r->n == 0) return TROVE_OK;
    const void* src = s->out_arena + r->row0 * (int64_t)s->row_bytes;
    return hip_status(hipMemcpy(out, src, (size_t)(r->n * s->row_bytes), hipMemcpyDefault));
}

int trove_serve_release(trove_serve* s, uint64_t request_id) {
    if (!s) return TROVE_ERR_INVALID;
    request* r = s->reqs.find(request_id);
    if (!r) return TROVE_ERR_NOTFOUND;
    if (r->state != DONE) return TROVE_ERR_AGAIN;
    if (r->row0 >= 0) s->arena->release(r->row0, r->rows);
    s->reqs.erase(*r);
    return TROVE_OK;
}

int trove_serve_drain(trove_serve* s) {
    if (!s) return TROVE_ERR_INVALID;
    for (;;) {
        int rc = serve_step(s);
        if (rc != TROVE_OK) return rc;
        if (s->queue.empty() && s->active.empty() && s->retired_seq == s->step_seq) return TROVE_OK;
        if (!s->queue.empty() && s->active.empty() && s->retired_seq == s->step_seq) {
            // Nothing running and the head does not fit: only releases free space.
            return TROVE_ERR_FULL;
        }
        if (s->retired_seq < s->step_seq) {
            rc = hip_status(hipEventSynchronize(s->steps[s->retired_seq % kInflight].done));
            if (rc != TROVE_OK) return rc;
        }
    }
}

int trove_serve_stats(const trove_serve* s, int64_t* completed, int64_t* indices_done, int64_t* steps,
                      int32_t* queued, int32_t* active) {
    if (!s) return TROVE_ERR_INVALID;
    if (completed) *completed = s->completed;
    if (indices_done) *indices_done = s->indices_done;
    if (steps) *steps = s->steps_launched;
    if (queued) *queued = (int32_t)s->queue.size();
    if (active) *active = (int32_t)s->active.size();
    return TROVE_OK;
}

// The replay loop proper; trove_serve_replay brackets it with the recording
// of completions in newly_done.
static int replay_loop(trove_serve* s, int32_t n_req, const double* arrival_s, const int32_t* const* idx,
                       const int64_t* n, int idx_on_device, int release_on_done, double* latency_s,
                       uint64_t* request_ids) {
    std::vector<uint64_t> own_ids(request_ids ? 0 : n_req, 0);
    uint64_t* ids = request_ids ? request_ids : own_ids.data();
    // Arrival index of the request in each table slot (-1: none of this
    // replay's); the reservation keeps every slot number below slot_bound().
    s->reqs.reserve_more((size_t)n_req);
    std::vector<int32_t> arrival_of(s->reqs.slot_bound(), -1);
    int32_t next = 0, done = 0;
    // A rejected request (an index outside the table) reports latency -1.
    auto latency_of = [](const request& r) { return r.bad_index ? -1.0 : r.t_done - r.t_submit; };
    const double t0 = now_s();
    while (done < n_req) {
        const double t = now_s() - t0;
        // At most max_slots submissions between steps: admission cannot take
        // more at once, and a burst of due arrivals (a saturation run) then
        // reaches the GPU after one slot's worth of host work instead of all
        // of it. Latency still counts from each request's scheduled arrival.
        int32_t burst = s->max_slots;
        while (next < n_req && arrival_s[next] <= t && burst-- > 0) {
            const int rc = trove_serve_submit(s, idx[next], n[next], idx_on_device, &ids[next]);
            if (rc != TROVE_OK) return rc;
            // Latency counts from the scheduled arrival, not from submission.
            request& r = s->reqs[ids[next]];
            r.t_submit = t0 + arrival_s[next];
            if (r.state == DONE) {
                latency_s[next] = latency_of(r);
                ++done;
                if (release_on_done) trove_serve_release(s, ids[next]);
            } else {
                arrival_of[ids[next] & 0xffffffffu] = next;
            }
            ++next;
        }
        const int rc = serve_step(s);
        if (rc != TROVE_OK) return rc;
        if (!s->queue.empty() && s->active.empty() && s->retired_seq == s->step_seq && s->newly_done.empty()) {
            // Nothing runs and the queue head does not fit: only a release can
            // free space, and outputs are being kept.
            return TROVE_ERR_FULL;
        }
        for (uint64_t id : s->newly_done) {
            int32_t& a = arrival_of[id & 0xffffffffu];
            if (a < 0 || ids[a] != id) continue;  // submitted before this replay
            latency_s[a] = latency_of(s->reqs[id]);
            ++done;
            a = -1;
            if (release_on_done) trove_serve_release(s, id);
        }
        s->newly_done.clear();
    }
    return TROVE_OK;
}

int trove_serve_replay(trove_serve* s, int32_t n_req, const double* arrival_s, const int32_t* const* idx,
                       const int64_t* n, int idx_on_device, int release_on_done, double* latency_s,
                       uint64_t* request_ids) {
    if (!s || n_req < 0 || (n_req > 0 && (!arrival_s || !idx || !n || !latency_s))) return TROVE_ERR_INVALID;
    s->newly_done.clear();
    s->replaying = true;
    const int rc = replay_loop(s, n_req, arrival_s, idx, n, idx_on_device, release_on_done, latency_s, request_ids);
    s->replaying = false;
    s->newly_done.clear();
    return rc;
}

int trove_serve_destroy(trove_serve* s) {
    if (!s) return TROVE_ERR_INVALID;
    (void)hipStreamSynchronize(s->stream);
    (void)hipStreamSynchronize(s->copy_stream);
    free_serve(s);
    return TROVE_OK;
}

}  // extern "C"
