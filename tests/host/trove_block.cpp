// Host driver with the shape of the reference's block-interface test
// (tests/block.cu:56-75): writes l arrays of s counting ints through the
// block interface, checks them against 0, 1, 2, ... (test_block_write), then
// copies that sequence through the block interface and checks the copy
// (test_block_copy); prints the reference's two result lines. Calls the
// library only through its C ABI (include/trove_amd.h).
//
// Usage: trove_block [--l L] [--s S] [--divergent 0|1]
// (defaults: the reference's l = 100000*256 + 17 arrays of s = 5 ints;
// --divergent 1 masks lanes in one of eight wavefront groups of the copy so
// that those take the non-converged path). Exit status 0 when both pass.
#include <hip/hip_runtime.h>
#include <trove_amd.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
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

bool is_counting(const void* dev, int64_t n) {
    std::vector<int32_t> h(static_cast<size_t>(n));
    hip_ok(hipMemcpy(h.data(), dev, size_t(n) * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    for (int64_t j = 0; j < n; ++j)
        if (h[size_t(j)] != int32_t(j)) return false;
    return true;
}

}  // namespace

int main(int argc, char** argv) {
    int64_t l = int64_t(100000) * 256 + 17;
    int s = 5, divergent = 0;
    for (int a = 1; a + 1 < argc; a += 2) {
        const std::string arg = argv[a];
        if (arg == "--l") l = std::atoll(argv[a + 1]);
        else if (arg == "--s") s = std::atoi(argv[a + 1]);
        else if (arg == "--divergent") divergent = std::atoi(argv[a + 1]);
        else {
            std::fprintf(stderr, "unknown option %s\n", arg.c_str());
            return 2;
        }
    }
    if (argc % 2 == 0 || l <= 0 || s < 1 || s > 16 || l * s > INT32_MAX) {
        std::fprintf(stderr, "usage: trove_block [--l L] [--s 1..16] [--divergent 0|1]\n");
        return 2;
    }
    const int64_t int_length = l * s;
    void *d, *g, *e;
    hip_ok(hipMalloc(&d, size_t(int_length) * 4), "hipMalloc");
    hip_ok(hipMalloc(&g, size_t(int_length) * 4), "hipMalloc");
    hip_ok(hipMalloc(&e, size_t(int_length) * 4), "hipMalloc");

    trove_ok(trove_block_write(d, l, s, nullptr), "trove_block_write");
    hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize");
    const bool write_ok = is_counting(d, int_length);
    std::printf("test_block_write results pass: %s\n", write_ok ? "true" : "false");

    std::vector<int32_t> h(static_cast<size_t>(int_length));
    for (int64_t j = 0; j < int_length; ++j) h[size_t(j)] = int32_t(j);
    hip_ok(hipMemcpy(g, h.data(), size_t(int_length) * 4, hipMemcpyHostToDevice), "hipMemcpy");
    trove_ok(trove_block_copy(g, e, int_length, s, divergent, 7, nullptr), "trove_block_copy");
    hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize");
    const bool copy_ok = is_counting(e, int_length);
    std::printf("test_block_copy results pass: %s\n", copy_ok ? "true" : "false");

    (void)hipFree(d);
    (void)hipFree(g);
    (void)hipFree(e);
    return write_ok && copy_ok ? 0 : 1;
}
