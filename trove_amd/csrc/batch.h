// Job table of the batch-gather kernels (batch.hip), shared with
// the serving engine, which launches one table per step.
#pragma once
#include "common.h"

namespace trove_amd {

// One gather job (a copy job when idx is null) as the kernel sees it. Work
// item t of the launch belongs to the job with first_item <= t < next job's
// first_item and covers chunk_groups 64-struct groups of it.
struct job_desc {
    const void* src;
    const int32_t* idx;
    void* dst;
    int64_t n;
    int64_t first_item;
    int32_t bytes;  // 4..64, a multiple of 4
    int32_t variant;
    int32_t chunk_groups;
    uint32_t last_row;  // indices above it (as uint32) are rejected; 0xffffffff: unchecked
};
static_assert(sizeof(job_desc) == 56, "job_desc layout");

// Launches the batch gather over a device job table (sorted by
// first_item, total_items work items in all). lds: some job is LDS-staged
// (the launch then carries a 4 KiB LDS stage per wavefront). flags
// (nullable): flags[j] is set to 1 when job j meets an index above its
// last_row (that index reads row 0 instead).
int launch_job_table(const job_desc* d_table, int n_jobs, int64_t total_items, bool lds, int32_t* flags,
                     hipStream_t s);

// Up to kArgJobs jobs travel in the kernel arguments instead (56 * 64 + 16
// bytes, inside the 4 KiB argument segment): no table upload on the launch
// path.
constexpr int kArgJobs = 64;
struct job_pack {
    job_desc jobs[kArgJobs];
    int32_t n_jobs;
    int32_t pad;
    int64_t total_items;
    int32_t* flags;  // as for launch_job_table
};
static_assert(sizeof(job_pack) <= 4096 - 64, "kernel argument segment");
int launch_job_pack(const job_pack& p, bool lds, hipStream_t s);

}  // namespace trove_amd
