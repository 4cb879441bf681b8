// Costs of the primitives a persistent device-side request queue would be
// built from, measured on the device the tests run on (built into
// tests/host/libtrove_devtests.so, driven by tests/test_gpu_queue_primitives.py):
//  * work claims: agent-scope fetch-add on one counter, and on eight
//    counters 256 bytes apart, from 4,096 wavefronts (200 each);
//  * polling: a dependent chain of system-scope acquire loads of pinned,
//    coherent host memory, from one lane;
//  * publishing results: 8,192 wavefronts each writing 4 KiB (nontemporal)
//    and then, per write, nothing, an agent-scope or a system-scope release
//    fence followed by a counter fetch-add (a queue's "this item is done").
// Every loop runs a fixed number of iterations and waits on nothing.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

constexpr int kClaimBlocks = 1024, kClaimIters = 200;  // 4,096 wavefronts
constexpr int kCounterStride = 32;                     // 256 bytes between counters
constexpr int kPollIters = 2000;
constexpr int kFenceBlocks = 2048, kFenceIters = 64;  // 8,192 wavefronts x 4 KiB

__global__ __launch_bounds__(256) void k_claims(unsigned long long* ctr, int n_counters, unsigned long long* sink) {
    const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    unsigned long long* c = ctr + (blockIdx.x % n_counters) * kCounterStride;
    if ((threadIdx.x & 63) == 0) {
        unsigned long long acc = 0;
        for (int i = 0; i < kClaimIters; ++i)
            acc += __hip_atomic_fetch_add(c, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        sink[wave] = acc;
    }
}

__global__ void k_poll(const unsigned long long* host, unsigned long long* sink) {
    if (threadIdx.x == 0) {
        unsigned long long acc = 0, next = 0;
        for (int i = 0; i < kPollIters; ++i) {
            const unsigned long long v =
                __hip_atomic_load(host + (next & 7), __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
            acc += v;
            next = v + (unsigned long long)i;  // each address depends on the previous load
        }
        sink[0] = acc;
    }
}

// fence: 0 none, 1 agent-scope release, 2 system-scope release.
__global__ __launch_bounds__(256) void k_publish(uint32_t* buf, int fence, unsigned long long* done) {
    const int lane = (int)(threadIdx.x & 63);
    const int64_t wave = (int64_t)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    uint32_t* w = buf + wave * 1024;  // 4 KiB per wavefront
    for (int i = 0; i < kFenceIters; ++i) {
#pragma unroll
        for (int j = 0; j < 16; ++j) __builtin_nontemporal_store((uint32_t)(i + j), w + j * 64 + lane);
        if (fence == 1) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        if (fence == 2) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
        if (fence != 0 && lane == 0) __hip_atomic_fetch_add(done, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

float elapsed_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, a, b);
    return ms;
}

}  // namespace

extern "C" {

// out[0..7]: M claims/s on one counter, M claims/s over eight counters,
// 1 if every counter ends at its exact count, us per dependent host-memory
// load, us per publishing iteration with no fence / an agent-scope / a
// system-scope release fence, 1 if the done counter ends at its exact count.
// Returns 0, or -2 on a HIP error.
int devtest_queue_primitives(double* out) {
    hipEvent_t a = nullptr, b = nullptr;
    unsigned long long *ctr = nullptr, *sink = nullptr, *host = nullptr, *host_dev = nullptr;
    uint32_t* buf = nullptr;
    const size_t ctr_words = 8 * kCounterStride;
    const int claim_waves = kClaimBlocks * 4, fence_waves = kFenceBlocks * 4;
    bool ok = hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess &&
              hipMalloc(&ctr, ctr_words * sizeof(unsigned long long)) == hipSuccess &&
              hipMalloc(&sink, (size_t)claim_waves * sizeof(unsigned long long)) == hipSuccess &&
              hipHostMalloc((void**)&host, 8 * sizeof(unsigned long long), hipHostMallocCoherent | hipHostMallocMapped) ==
                  hipSuccess &&
              hipHostGetDevicePointer((void**)&host_dev, host, 0) == hipSuccess &&
              hipMalloc(&buf, (size_t)fence_waves * 4096) == hipSuccess;
    bool counts_ok = true;
    for (int k = 0; ok && k < 2; ++k) {
        const int n_counters = k == 0 ? 1 : 8;
        for (int rep = 0; ok && rep < 2; ++rep) {  // the first launch warms up
            ok = hipMemset(ctr, 0, ctr_words * sizeof(unsigned long long)) == hipSuccess &&
                 hipEventRecord(a) == hipSuccess;
            hipLaunchKernelGGL(k_claims, dim3(kClaimBlocks), dim3(256), 0, 0, ctr, n_counters, sink);
            ok = ok && hipGetLastError() == hipSuccess && hipEventRecord(b) == hipSuccess &&
                 hipEventSynchronize(b) == hipSuccess;
        }
        if (!ok) break;
        out[k] = (double)claim_waves * kClaimIters / (elapsed_ms(a, b) * 1e3);
        unsigned long long got[8 * kCounterStride];
        ok = hipMemcpy(got, ctr, sizeof(got), hipMemcpyDeviceToHost) == hipSuccess;
        for (int q = 0; ok && q < n_counters; ++q) {
            const unsigned long long blocks = (kClaimBlocks - q + n_counters - 1) / n_counters;
            counts_ok = counts_ok && got[q * kCounterStride] == blocks * 4 * kClaimIters;
        }
    }
    out[2] = counts_ok ? 1.0 : 0.0;
    if (ok) {
        for (int i = 0; i < 8; ++i) host[i] = 0;
        for (int rep = 0; ok && rep < 2; ++rep) {
            ok = hipEventRecord(a) == hipSuccess;
            hipLaunchKernelGGL(k_poll, dim3(1), dim3(64), 0, 0, host_dev, sink);
            ok = ok && hipGetLastError() == hipSuccess && hipEventRecord(b) == hipSuccess &&
                 hipEventSynchronize(b) == hipSuccess;
        }
        if (ok) out[3] = elapsed_ms(a, b) * 1e3 / kPollIters;
    }
    bool done_ok = true;
    for (int fence = 0; ok && fence < 3; ++fence) {
        for (int rep = 0; ok && rep < 2; ++rep) {
            ok = hipMemset(ctr, 0, sizeof(unsigned long long)) == hipSuccess && hipEventRecord(a) == hipSuccess;
            hipLaunchKernelGGL(k_publish, dim3(kFenceBlocks), dim3(256), 0, 0, buf, fence, ctr);
            ok = ok && hipGetLastError() == hipSuccess && hipEventRecord(b) == hipSuccess &&
                 hipEventSynchronize(b) == hipSuccess;
        }
        if (!ok) break;
        out[4 + fence] = elapsed_ms(a, b) * 1e3 / kFenceIters;
        unsigned long long done = 0;
        ok = hipMemcpy(&done, ctr, sizeof(done), hipMemcpyDeviceToHost) == hipSuccess;
        done_ok = done_ok && done == (fence ? (unsigned long long)fence_waves * kFenceIters : 0ull);
    }
    out[7] = done_ok ? 1.0 : 0.0;
    if (buf) (void)hipFree(buf);
    if (host) (void)hipHostFree(host);
    if (sink) (void)hipFree(sink);
    if (ctr) (void)hipFree(ctr);
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    return ok ? 0 : -2;
}

}  // extern "C"
