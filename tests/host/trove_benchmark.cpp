// Host driver with the shape of the reference's benchmark program
// (tests/benchmark.cu:130-372): for structs of k = lower..upper 32-bit words
// it runs the contiguous store and load benchmarks and the random scatter and
// gather benchmarks, each as the Trove kernel ("SHFL") and the naive kernel
// ("Direct"), times `iterations` launches with HIP events, checks the Trove
// kernel's output against the naive kernel's ("gold", as the reference does
// with thrust::equal) and prints one line per benchmark:
//
//     <name>, <k>, <GB/s>, Results passed | INCORRECT
//
// GB/s follows the reference's accounting: store S, load S + 4, scatter and
// gather 2 S + 4 bytes per struct and launch. The tile-path and LDS-staged
// gather and scatter (no reference counterpart) are printed after the
// reference's lines.
// Everything goes through the C ABI (include/trove_amd.h).
//
// Usage: trove_benchmark [--n N] [--lower L] [--upper U] [--iterations I] [--seed S] [--tile T]
//                        [--contiguous-only 1]
// (defaults: the reference's N = 80*8*100*256 structs, k = 1..16, 10 launches;
// tile 64, a whole wavefront). --tile is the reference's TILE_SIZE
// (tests/Makefile:4, tests/benchmark.cu:47-50, 32 there by default): the tile
// width of store_warp_contiguous / load_warp_contiguous in the contiguous
// store and load benchmarks.
// Exit status 0 when every result passed, 1 otherwise, 2 on an error.
#include <hip/hip_runtime.h>
#include <trove_amd.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

void hip_ok(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
        std::exit(2);
    }
}

void trove_ok(int rc, const char* what) {
    if (rc != TROVE_OK) {
        std::fprintf(stderr, "%s: %s\n", what, trove_strerror(rc));
        std::exit(2);
    }
}

struct device_buffer {
    void* p = nullptr;
    size_t bytes = 0;
    explicit device_buffer(size_t b) : bytes(b) { hip_ok(hipMalloc(&p, b > 0 ? b : 1), "hipMalloc"); }
    ~device_buffer() { (void)hipFree(p); }
    device_buffer(const device_buffer&) = delete;
    device_buffer& operator=(const device_buffer&) = delete;
};

// Byte equality of two device buffers, compared on the host in chunks.
bool equal(const device_buffer& a, const device_buffer& b) {
    if (a.bytes != b.bytes) return false;
    const size_t chunk = size_t(64) << 20;
    std::vector<unsigned char> ha(chunk), hb(chunk);
    for (size_t off = 0; off < a.bytes; off += chunk) {
        const size_t len = a.bytes - off < chunk ? a.bytes - off : chunk;
        hip_ok(hipMemcpy(ha.data(), static_cast<const char*>(a.p) + off, len, hipMemcpyDeviceToHost), "hipMemcpy");
        hip_ok(hipMemcpy(hb.data(), static_cast<const char*>(b.p) + off, len, hipMemcpyDeviceToHost), "hipMemcpy");
        if (std::memcmp(ha.data(), hb.data(), len) != 0) return false;
    }
    return true;
}

template <typename F>
float time_ms(F launch, int iterations) {
    hipEvent_t start, stop;
    hip_ok(hipEventCreate(&start), "hipEventCreate");
    hip_ok(hipEventCreate(&stop), "hipEventCreate");
    launch();  // untimed: loads the kernel's code object
    hip_ok(hipEventRecord(start, nullptr), "hipEventRecord");
    for (int j = 0; j < iterations; ++j) launch();
    hip_ok(hipEventRecord(stop, nullptr), "hipEventRecord");
    hip_ok(hipEventSynchronize(stop), "hipEventSynchronize");
    float ms = 0.f;
    hip_ok(hipEventElapsedTime(&ms, start, stop), "hipEventElapsedTime");
    (void)hipEventDestroy(start);
    (void)hipEventDestroy(stop);
    return ms;
}

struct options {
    int64_t n = int64_t(80) * 8 * 100 * 256;
    int lower = 1, upper = 16, iterations = 10, tile = 64;
    bool contiguous_only = false;  // --contiguous-only 1: the store and load benchmarks alone
    uint64_t seed = 1;
};

bool all_passed = true;

void report(const char* name, int k, double bytes_per_launch, int iterations, float ms, bool correct) {
    const double gbs = bytes_per_launch * iterations / (ms * 1e6);
    std::printf("%s, %d, %.1f, %s\n", name, k, gbs, correct ? "Results passed" : "INCORRECT");
    std::fflush(stdout);
    all_passed = all_passed && correct;
}

// tests/benchmark.cu:130-161: counting structs stored, gold = the Direct store.
void contiguous_store(const options& o, const char* name, int k, int variant) {
    const int s = 4 * k;
    device_buffer r(size_t(o.n) * s);
    const float ms = time_ms([&] { trove_ok(trove_store_counting(r.p, o.n, s, variant, o.tile, nullptr), name); },
                             o.iterations);
    bool correct = true;
    if (variant != TROVE_VARIANT_DIRECT) {
        device_buffer g(r.bytes);
        trove_ok(trove_store_counting(g.p, o.n, s, TROVE_VARIANT_DIRECT, 0, nullptr), "gold store");
        correct = equal(r, g);
    }
    report(name, k, double(s) * o.n, o.iterations, ms, correct);
}

// tests/benchmark.cu:183-226: input = 0, 1, 2, ... as 32-bit words (the
// Direct counting store writes exactly that), output = per-struct word sums.
void contiguous_load(const options& o, const char* name, int k, int variant) {
    const int s = 4 * k;
    device_buffer src(size_t(o.n) * s);
    trove_ok(trove_store_counting(src.p, o.n, s, TROVE_VARIANT_DIRECT, 0, nullptr), "fill");
    device_buffer r(size_t(o.n) * 4);
    const float ms = time_ms(
        [&] { trove_ok(trove_load_sum(src.p, static_cast<uint32_t*>(r.p), o.n, s, variant, o.tile, nullptr), name); },
        o.iterations);
    bool correct = true;
    if (variant != TROVE_VARIANT_DIRECT) {
        device_buffer g(r.bytes);
        trove_ok(trove_load_sum(src.p, static_cast<uint32_t*>(g.p), o.n, s, TROVE_VARIANT_DIRECT, 0, nullptr),
                 "gold load");
        correct = equal(r, g);
    }
    report(name, k, double(s + 4) * o.n, o.iterations, ms, correct);
}

// tests/benchmark.cu:260-297: scatter r[perm[i]] = s[i] / gather r[i] = s[perm[i]].
void random_access(const options& o, const device_buffer& perm, const char* name, int k, bool gather,
                   int variant) {
    const int s = 4 * k;
    device_buffer src(size_t(o.n) * s);
    trove_ok(trove_store_counting(src.p, o.n, s, TROVE_VARIANT_DIRECT, 0, nullptr), "fill");
    device_buffer r(src.bytes);
    const int32_t* idx = static_cast<const int32_t*>(perm.p);
    auto run = [&](void* out, int v) {
        // The reference's gathers and scatters go through coalesced_ptr, which
        // has no tile width: whole wavefronts here.
        return gather ? trove_gather(src.p, idx, out, o.n, s, v, 0, nullptr)
                      : trove_scatter(src.p, idx, out, o.n, s, v, 0, nullptr);
    };
    const float ms = time_ms([&] { trove_ok(run(r.p, variant), name); }, o.iterations);
    bool correct = true;
    if (variant != TROVE_VARIANT_DIRECT) {
        device_buffer g(r.bytes);
        trove_ok(run(g.p, TROVE_VARIANT_DIRECT), "gold");
        correct = equal(r, g);
    }
    report(name, k, double(2 * s + 4) * o.n, o.iterations, ms, correct);
}

options parse(int argc, char** argv) {
    options o;
    for (int a = 1; a < argc; ++a) {
        const std::string arg = argv[a];
        if (a + 1 >= argc) {
            std::fprintf(stderr, "missing value for %s\n", arg.c_str());
            std::exit(2);
        }
        const char* v = argv[++a];
        if (arg == "--n") o.n = std::atoll(v);
        else if (arg == "--lower") o.lower = std::atoi(v);
        else if (arg == "--upper") o.upper = std::atoi(v);
        else if (arg == "--iterations") o.iterations = std::atoi(v);
        else if (arg == "--seed") o.seed = std::strtoull(v, nullptr, 10);
        else if (arg == "--tile") o.tile = std::atoi(v);
        else if (arg == "--contiguous-only") o.contiguous_only = std::atoi(v) != 0;
        else {
            std::fprintf(stderr, "unknown option %s\n", arg.c_str());
            std::exit(2);
        }
    }
    const bool tile_ok = o.tile >= 2 && o.tile <= 64 && (o.tile & (o.tile - 1)) == 0;
    if (o.n <= 0 || o.n > INT32_MAX || o.lower < 1 || o.upper > 16 || o.lower > o.upper || o.iterations < 1 ||
        !tile_ok) {
        std::fprintf(stderr,
                     "need 0 < n <= 2^31-1, 1 <= lower <= upper <= 16, iterations >= 1, tile a power of two 2..64\n");
        std::exit(2);
    }
    return o;
}

}  // namespace

int main(int argc, char** argv) {
    const options o = parse(argc, argv);
    for (int k = o.lower; k <= o.upper; ++k) contiguous_store(o, "Contiguous SHFL Store", k, TROVE_VARIANT_SHFL);
    for (int k = o.lower; k <= o.upper; ++k) contiguous_store(o, "Contiguous Direct Store", k, TROVE_VARIANT_DIRECT);
    for (int k = o.lower; k <= o.upper; ++k) contiguous_load(o, "Contiguous SHFL Load", k, TROVE_VARIANT_SHFL);
    for (int k = o.lower; k <= o.upper; ++k) contiguous_load(o, "Contiguous Direct Load", k, TROVE_VARIANT_DIRECT);
    if (o.contiguous_only) return all_passed ? 0 : 1;

    // A uniform random permutation of [0, n) (the reference sorts rand() keys;
    // this is Fisher-Yates from a seeded generator).
    std::vector<int32_t> h_perm(static_cast<size_t>(o.n));
    trove_ok(trove_gen_permutation(h_perm.data(), o.n, o.seed), "permutation");
    device_buffer perm(size_t(o.n) * 4);
    hip_ok(hipMemcpy(perm.p, h_perm.data(), perm.bytes, hipMemcpyHostToDevice), "hipMemcpy");

    for (int k = o.lower; k <= o.upper; ++k) random_access(o, perm, "SHFL Scatter", k, false, TROVE_VARIANT_SHFL);
    for (int k = o.lower; k <= o.upper; ++k) random_access(o, perm, "Direct Scatter", k, false, TROVE_VARIANT_DIRECT);
    for (int k = o.lower; k <= o.upper; ++k) random_access(o, perm, "SHFL Gather", k, true, TROVE_VARIANT_SHFL);
    for (int k = o.lower; k <= o.upper; ++k) random_access(o, perm, "Direct Gather", k, true, TROVE_VARIANT_DIRECT);
    for (int k = o.lower; k <= o.upper; ++k) random_access(o, perm, "Tile Scatter", k, false, TROVE_VARIANT_TILE);
    for (int k = o.lower; k <= o.upper; ++k) random_access(o, perm, "Tile Gather", k, true, TROVE_VARIANT_TILE);
    for (int k = o.lower; k <= o.upper; ++k) random_access(o, perm, "LDS Scatter", k, false, TROVE_VARIANT_LDS);
    for (int k = o.lower; k <= o.upper; ++k) random_access(o, perm, "LDS Gather", k, true, TROVE_VARIANT_LDS);
    return all_passed ? 0 : 1;
}
