// Device-side unit checks of the public device API, built into
// tests/host/libtrove_devtests.so and driven by tests/test_gpu_device_api.py.
// Each entry point runs one 64-lane work-group and writes what the lanes saw;
// the host test computes the expected values.
#include <hip/hip_runtime.h>
#include <trove/aos.h>
#include <trove/lds.h>
#include <trove/ptr.h>

#include <cstdint>

namespace {

struct w16 {
    uint32_t a, b, c, d;
};

template <typename T>
__device__ T encode(uint32_t v) {
    T t;
    uint32_t w[sizeof(T) / 4];
    for (int x = 0; x < (int)(sizeof(T) / 4); ++x) w[x] = v * 4u + (uint32_t)x;
    __builtin_memcpy(&t, w, sizeof(T));
    return t;
}

template <typename T>
__device__ uint32_t decode(const T& t) {
    uint32_t w[sizeof(T) / 4];
    __builtin_memcpy(w, &t, sizeof(T));
    // Every word must agree with the first; a mismatch is reported as ~0.
    for (int x = 1; x < (int)(sizeof(T) / 4); ++x)
        if (w[x] != w[0] + (uint32_t)x) return ~0u;
    return w[0] / 4u;
}

// Lane id of tile rank r: column-side value r + j*N goes in, the public
// explicit-index R2C must produce r*M + j; then the self-contained C2R must
// restore the input.
template <int N, int M, typename T>
__global__ __launch_bounds__(64) void k_transpose(uint32_t* r2c_out, uint32_t* c2r_out) {
    trove::thread_tile<N> tile;
    const int r = tile.id();
    const int lane = threadIdx.x;
    T in[M];
    for (int j = 0; j < M; ++j) in[j] = encode<T>((uint32_t)(r + j * N));
    trove::array<T, M> v = trove::make_array<T, M>(in);
    trove::array<int, M> idx;
    int rot;
    trove::r2c_compute_indices(idx, rot, tile);
    trove::r2c_warp_transpose(v, idx, rot, tile);
    T out[M];
    trove::make_carray<T, M>(v, out);
    for (int j = 0; j < M; ++j) r2c_out[lane * M + j] = decode(out[j]);
    trove::c2r_compute_indices(idx, rot, tile);
    trove::c2r_warp_transpose(v, idx, rot, tile);
    trove::make_carray<T, M>(v, out);
    for (int j = 0; j < M; ++j) c2r_out[lane * M + j] = decode(out[j]);
}

// The self-contained forms r2c_warp_transpose(v, tile) / c2r_warp_transpose(v,
// tile): square tiles (m == n) take the butterfly transpose there.
template <int N, int M, typename T>
__global__ __launch_bounds__(64) void k_transpose_self(uint32_t* r2c_out, uint32_t* c2r_out) {
    trove::thread_tile<N> tile;
    const int r = tile.id();
    const int lane = threadIdx.x;
    T in[M];
    for (int j = 0; j < M; ++j) in[j] = encode<T>((uint32_t)(r + j * N));
    trove::array<T, M> v = trove::make_array<T, M>(in);
    trove::r2c_warp_transpose(v, tile);
    T out[M];
    trove::make_carray<T, M>(v, out);
    for (int j = 0; j < M; ++j) r2c_out[lane * M + j] = decode(out[j]);
    trove::c2r_warp_transpose(v, tile);
    trove::make_carray<T, M>(v, out);
    for (int j = 0; j < M; ++j) c2r_out[lane * M + j] = decode(out[j]);
}

template <int M>
__global__ __launch_bounds__(64) void k_rotate(uint32_t* out) {
    const int lane = threadIdx.x;
    int lim = 1;
    while (lim < M) lim <<= 1;
    const int a = lane % lim;
    uint32_t t[M];
    for (int j = 0; j < M; ++j) t[j] = 100u * (uint32_t)lane + (uint32_t)j;
    trove::array<uint32_t, M> v = trove::make_array<uint32_t, M>(t);
    v = trove::rotate(v, a);
    trove::make_carray<uint32_t, M>(v, t);
    for (int j = 0; j < M; ++j) out[lane * M + j] = t[j];
}

__global__ __launch_bounds__(64) void k_converged(uint64_t mask, int32_t* out) {
    const int lane = threadIdx.x;
    for (int k = 0; k < 6; ++k) out[lane * 6 + k] = -1;
    if ((mask >> lane) & 1ull) {
        out[lane * 6 + 0] = trove::warp_converged<64>();
        out[lane * 6 + 1] = trove::warp_converged<32>();
        out[lane * 6 + 2] = trove::warp_converged<16>();
        out[lane * 6 + 3] = trove::warp_converged<8>();
        out[lane * 6 + 4] = trove::warp_converged<4>();
        out[lane * 6 + 5] = trove::warp_converged<2>();
    }
}

template <int N>
__global__ __launch_bounds__(64) void k_shfl(const int* src_rank, uint32_t* out) {
    trove::thread_tile<N> tile;
    const int lane = threadIdx.x;
    const uint64_t v = (uint64_t)lane * 0x100000001ull + 7;
    const uint64_t got = tile.shfl(v, src_rank[lane]);
    out[2 * lane] = (uint32_t)got;
    out[2 * lane + 1] = (uint32_t)(got >> 32);
}

// Active lanes (mask) copy src[perm[lane]] to dst[lane] (gather) or
// src[lane] to dst[perm[lane]] (scatter) through trove::load / trove::store.
template <int S>
__global__ __launch_bounds__(64) void k_cascade(uint64_t mask, int scatter, const trove::array<uint32_t, S / 4>* src, const int* perm,
                          trove::array<uint32_t, S / 4>* dst) {
    const int lane = threadIdx.x;
    if ((mask >> lane) & 1ull) {
        if (scatter) {
            trove::store(trove::load(src + lane), dst + perm[lane]);
        } else {
            trove::store(trove::load(src + perm[lane]), dst + lane);
        }
    }
}

// The same through the streaming cache policy: mode 0 calls
// trove::load / trove::store<T, streaming>, mode 1 assigns between two
// streaming coalesced_ptrs, mode 2 assigns a cached coalesced_ptr's element
// to a streaming one (the cross-policy ref = ref), mode 3 assigns between two
// cached coalesced_ptrs.
template <int S>
__global__ __launch_bounds__(64) void k_cascade_policy(uint64_t mask, int scatter, int mode,
                                                       const trove::array<uint32_t, S / 4>* src, const int* perm,
                                                       trove::array<uint32_t, S / 4>* dst) {
    typedef trove::array<uint32_t, S / 4> T;
    typedef trove::streaming NT;
    const int lane = threadIdx.x;
    if ((mask >> lane) & 1ull) {
        const int from = scatter ? lane : perm[lane], to = scatter ? perm[lane] : lane;
        if (mode == 0) {
            trove::store<T, NT>(trove::load<T, NT>(src + from), dst + to);
        } else if (mode == 1) {
            trove::coalesced_ptr<T, NT> s(const_cast<T*>(src));
            trove::coalesced_ptr<T, NT> d(dst);
            d[to] = s[from];
        } else if (mode == 2) {
            trove::coalesced_ptr<T> s(const_cast<T*>(src));
            trove::coalesced_ptr<T, NT> d(dst);
            d[to] = s[from];
        } else {
            trove::coalesced_ptr<T> s(const_cast<T*>(src));
            trove::coalesced_ptr<T> d(dst);
            d[to] = s[from];
        }
    }
}

// ref = ref within one array: the wavefront permutes the 64 structs of buf
// in place (scatter: buf[perm[lane]] = buf[lane]; gather: buf[lane] =
// buf[perm[lane]]); every lane must read the structs as they were before.
template <int S>
__global__ __launch_bounds__(64) void k_inplace(int scatter, trove::array<uint32_t, S / 4>* buf, const int* perm) {
    typedef trove::array<uint32_t, S / 4> T;
    const int lane = threadIdx.x;
    trove::coalesced_ptr<T> p(buf);
    if (scatter) {
        p[perm[lane]] = p[lane];
    } else {
        p[lane] = p[perm[lane]];
    }
}

// coalesced_ptr over a __shared__ array: active lanes (mask) read struct
// perm[lane] of 64 structs staged in LDS, through the dynamic path (pointers
// into LDS are flat pointers to the LDS aperture).
template <int S>
__global__ __launch_bounds__(64) void k_shared_gather(uint64_t mask, const trove::array<uint32_t, S / 4>* src,
                                                      const int* perm, trove::array<uint32_t, S / 4>* dst) {
    typedef trove::array<uint32_t, S / 4> T;
    __shared__ uint32_t words[64 * (S / 4)];  // raw storage: array<> has a constructor
    T* staged = reinterpret_cast<T*>(words);
    const int lane = threadIdx.x;
    staged[lane] = src[lane];
    __syncthreads();
    if ((mask >> lane) & 1ull) {
        trove::coalesced_ptr<T> p(staged);
        dst[lane] = p[perm[lane]];
    }
}

// Per-lane pointers into two views of one buffer that start `shift` bytes
// apart (a word, not a whole struct): lane l loads the struct at byte
// base + shift_l + perm[l] * S through trove::load and stores it to
// dst[l] through trove::store (the dynamic path on both sides).
template <int S>
__global__ __launch_bounds__(64) void k_mixed(uint64_t mask, const unsigned char* base, const int* perm,
                                              const int* shift, trove::array<uint32_t, S / 4>* dst) {
    typedef trove::array<uint32_t, S / 4> T;
    const int lane = threadIdx.x;
    if ((mask >> lane) & 1ull) {
        const T* p = reinterpret_cast<const T*>(base + shift[lane] + (long)perm[lane] * S);
        trove::store(trove::load(p), dst + lane);
    }
}

// LDS-staged warp-contiguous API (trove/lds.h) in a work-group of WAVES
// wavefronts, each with its own stage: wavefront w of block b handles tile
// t = b * WAVES + w at byte offset `off` of src / dst. Lane l stores the
// struct load_warp_contiguous_lds gave it to got[t * 64 + l] with a plain
// per-lane store, then writes it back through store_warp_contiguous_lds.
template <int S, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_lds_roundtrip(const unsigned char* src, int off,
                                                              trove::array<uint32_t, S / 4>* got, unsigned char* dst) {
    typedef trove::array<uint32_t, S / 4> T;
    __shared__ __align__(16) uint32_t stage[WAVES * 64 * (S / 4)];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long t = (long)blockIdx.x * WAVES + w;
    const T* s = reinterpret_cast<const T*>(src + off) + t * 64;
    T* d = reinterpret_cast<T*>(dst + off) + t * 64;
    uint32_t* my = stage + w * 64 * (S / 4);
    const T v = trove::load_warp_contiguous_lds(s + lane, my);
    got[t * 64 + lane] = v;
    trove::store_warp_contiguous_lds(v, d + lane, my);
}

// A user struct of mixed field types (8-byte aligned, 24 bytes = three
// 8-byte words): a coalesced_ptr gather written as ref = ref, the register
// transposition and the LDS-staged pair, one wavefront.
struct particle {
    float x, y, z;
    int id;
    double mass;
};
static_assert(sizeof(particle) == 24 && alignof(particle) == 8, "particle layout");

__global__ __launch_bounds__(64) void k_user_struct(const particle* src, const int* perm, particle* gathered,
                                                    particle* copied, particle* lds_copied) {
    __shared__ __align__(16) uint32_t stage[64 * sizeof(particle) / 4];
    const int lane = threadIdx.x;
    trove::coalesced_ptr<particle> s(const_cast<particle*>(src));
    trove::coalesced_ptr<particle> g(gathered);
    g[lane] = s[perm[lane]];
    const particle v = trove::load_warp_contiguous(src + lane);
    trove::store_warp_contiguous(v, copied + lane);
    const particle w = trove::load_warp_contiguous_lds(src + lane, stage);
    trove::store_warp_contiguous_lds(w, lds_copied + lane, stage);
}

}  // namespace

extern "C" {

// 64 particles (24 bytes each) in src; outputs of the same size.
int devtest_user_struct(const void* src, const int* perm, void* gathered, void* copied, void* lds_copied) {
    hipLaunchKernelGGL(k_user_struct, dim3(1), dim3(64), 0, 0, (const particle*)src, perm, (particle*)gathered,
                       (particle*)copied, (particle*)lds_copied);
    return hipDeviceSynchronize() == hipSuccess ? 0 : -2;
}


// Returns 0 and fills r2c_out / c2r_out (64 * m words each), -1 for an
// unsupported combination, -2 on a launch error.
int devtest_transpose(int n, int m, int word_bytes, uint32_t* r2c_out, uint32_t* c2r_out) {
#define TT(N, M, T)                                                                \
    if (n == N && m == M && word_bytes == (int)sizeof(T)) {                        \
        hipLaunchKernelGGL((k_transpose<N, M, T>), dim3(1), dim3(64), 0, 0, r2c_out, c2r_out); \
        return hipDeviceSynchronize() == hipSuccess ? 0 : -2;                      \
    }
#define TT_M(N, T) TT(N, 2, T) TT(N, 3, T) TT(N, 4, T) TT(N, 5, T) TT(N, 6, T) TT(N, 7, T) TT(N, 8, T) \
    TT(N, 9, T) TT(N, 10, T) TT(N, 12, T) TT(N, 14, T) TT(N, 15, T) TT(N, 16, T)
    TT_M(64, uint32_t)
    TT_M(32, uint32_t)
    TT_M(16, uint32_t)
    TT(8, 2, uint32_t) TT(8, 3, uint32_t) TT(8, 5, uint32_t) TT(8, 6, uint32_t) TT(8, 8, uint32_t)
    TT(4, 2, uint32_t) TT(4, 3, uint32_t) TT(4, 4, uint32_t)
    TT(2, 2, uint32_t)
    TT(64, 2, uint64_t) TT(64, 3, uint64_t) TT(64, 6, uint64_t) TT(64, 12, uint64_t)
    TT(64, 2, w16) TT(64, 5, w16) TT(64, 6, w16) TT(64, 16, w16) TT(32, 6, w16)
    TT(64, 24, uint32_t) TT(64, 33, uint32_t) TT(64, 64, uint32_t)
    TT(64, 24, w16) TT(64, 32, w16) TT(32, 32, w16) TT(64, 48, w16) TT(64, 64, w16)
#undef TT_M
#undef TT
    return -1;
}

// Same cases and contract as devtest_transpose, through the self-contained forms.
int devtest_transpose_self(int n, int m, int word_bytes, uint32_t* r2c_out, uint32_t* c2r_out) {
#define TS(N, M, T)                                                                \
    if (n == N && m == M && word_bytes == (int)sizeof(T)) {                        \
        hipLaunchKernelGGL((k_transpose_self<N, M, T>), dim3(1), dim3(64), 0, 0, r2c_out, c2r_out); \
        return hipDeviceSynchronize() == hipSuccess ? 0 : -2;                      \
    }
    TS(64, 64, uint32_t) TS(64, 64, w16) TS(32, 32, uint32_t) TS(32, 32, w16) TS(16, 16, uint32_t)
    TS(16, 16, uint64_t) TS(8, 8, uint32_t) TS(4, 4, uint32_t) TS(2, 2, uint32_t) TS(2, 2, w16)
    TS(64, 5, uint32_t) TS(64, 6, w16) TS(32, 12, uint32_t) TS(64, 48, w16)
#undef TS
    return -1;
}

int devtest_rotate(int m, uint32_t* out) {
#define RT(M)                                                                    \
    if (m == M) {                                                                \
        hipLaunchKernelGGL((k_rotate<M>), dim3(1), dim3(64), 0, 0, out);         \
        return hipDeviceSynchronize() == hipSuccess ? 0 : -2;                    \
    }
    RT(1) RT(2) RT(3) RT(4) RT(5) RT(6) RT(7) RT(8) RT(9) RT(11) RT(13) RT(15) RT(16)
#undef RT
    return -1;
}

int devtest_converged(uint64_t mask, int32_t* out) {
    hipLaunchKernelGGL(k_converged, dim3(1), dim3(64), 0, 0, mask, out);
    return hipDeviceSynchronize() == hipSuccess ? 0 : -2;
}

int devtest_shfl(int n, const int* src_rank, uint32_t* out) {
#define SH(N)                                                                    \
    if (n == N) {                                                                \
        hipLaunchKernelGGL((k_shfl<N>), dim3(1), dim3(64), 0, 0, src_rank, out); \
        return hipDeviceSynchronize() == hipSuccess ? 0 : -2;                    \
    }
    SH(64) SH(32) SH(16) SH(8) SH(4) SH(2)
#undef SH
    return -1;
}

int devtest_cascade(int s, uint64_t mask, int scatter, const void* src, const int* perm, void* dst) {
#define CS(S)                                                                                        \
    if (s == S) {                                                                                    \
        hipLaunchKernelGGL((k_cascade<S>), dim3(1), dim3(64), 0, 0, mask, scatter,                  \
                           (const trove::array<uint32_t, S / 4>*)src, perm, (trove::array<uint32_t, S / 4>*)dst); \
        return hipDeviceSynchronize() == hipSuccess ? 0 : -2;                                        \
    }
    CS(8) CS(12) CS(20) CS(32) CS(48) CS(60) CS(96) CS(384) CS(512) CS(1024)
#undef CS
    return -1;
}

int devtest_cascade_policy(int s, uint64_t mask, int scatter, int mode, const void* src, const int* perm, void* dst) {
#define CP(S)                                                                                        \
    if (s == S) {                                                                                    \
        hipLaunchKernelGGL((k_cascade_policy<S>), dim3(1), dim3(64), 0, 0, mask, scatter, mode,     \
                           (const trove::array<uint32_t, S / 4>*)src, perm, (trove::array<uint32_t, S / 4>*)dst); \
        return hipDeviceSynchronize() == hipSuccess ? 0 : -2;                                        \
    }
    CP(8) CP(20) CP(32) CP(60) CP(64) CP(96) CP(256)
#undef CP
    return -1;
}

int devtest_inplace(int s, int scatter, void* buf, const int* perm) {
#define IP(S)                                                                                        \
    if (s == S) {                                                                                    \
        hipLaunchKernelGGL((k_inplace<S>), dim3(1), dim3(64), 0, 0, scatter,                         \
                           (trove::array<uint32_t, S / 4>*)buf, perm);                                \
        return hipDeviceSynchronize() == hipSuccess ? 0 : -2;                                        \
    }
    IP(8) IP(20) IP(32) IP(60) IP(64) IP(96) IP(256)
#undef IP
    return -1;
}

int devtest_shared_gather(int s, uint64_t mask, const void* src, const int* perm, void* dst) {
#define SG(S)                                                                                            \
    if (s == S) {                                                                                        \
        hipLaunchKernelGGL((k_shared_gather<S>), dim3(1), dim3(64), 0, 0, mask,                          \
                           (const trove::array<uint32_t, S / 4>*)src, perm, (trove::array<uint32_t, S / 4>*)dst); \
        return hipDeviceSynchronize() == hipSuccess ? 0 : -2;                                            \
    }
    SG(8) SG(20) SG(32) SG(60) SG(64)
#undef SG
    return -1;
}

// tiles must be a multiple of the work-group's wavefront count (4 up to 64
// bytes, 1 above).
int devtest_lds(int s, int off, int tiles, const void* src, void* got, void* dst) {
#define LR(S, W)                                                                                          \
    if (s == S) {                                                                                         \
        if (tiles % W) return -1;                                                                         \
        hipLaunchKernelGGL((k_lds_roundtrip<S, W>), dim3(tiles / W), dim3(64 * W), 0, 0,                  \
                           (const unsigned char*)src, off, (trove::array<uint32_t, S / 4>*)got, (unsigned char*)dst); \
        return hipDeviceSynchronize() == hipSuccess ? 0 : -2;                                             \
    }
    LR(4, 4) LR(8, 4) LR(12, 4) LR(20, 4) LR(32, 4) LR(60, 4) LR(64, 4) LR(96, 1) LR(256, 1)
#undef LR
    return -1;
}

int devtest_mixed(int s, uint64_t mask, const void* base, const int* perm, const int* shift, void* dst) {
#define MX(S)                                                                                              \
    if (s == S) {                                                                                          \
        hipLaunchKernelGGL((k_mixed<S>), dim3(1), dim3(64), 0, 0, mask, (const unsigned char*)base, perm, shift, \
                           (trove::array<uint32_t, S / 4>*)dst);                                           \
        return hipDeviceSynchronize() == hipSuccess ? 0 : -2;                                              \
    }
    MX(12) MX(20) MX(32) MX(48) MX(96)
#undef MX
    return -1;
}

}  // extern "C"

// ---- trove/memory.h and the compile-time number theory ----
#include <trove/memory.h>
#include <trove/static_gcd.h>
#include <trove/static_mod_inverse.h>
#include <trove/utility.h>

#include <utility>

namespace {

// Lane l: a = warp_load<array<T, S>>(src, offset[l], stride[l]) (element k
// from src[offset[l] + k * stride[l]]), written out as loaded[l * S + k];
// warp_store(a, stored, offset[l], stride[l]); uncoalesced_store(a,
// unc + l * S, 3) (element k to unc[l * S + k]; the stride is ignored). VOL
// takes the volatile-pointer overloads.
template <typename T, int S, bool VOL>
__global__ __launch_bounds__(64) void k_memory(const T* src, const int* offset, const int* stride, T* loaded,
                                               T* stored, T* unc) {
    typedef trove::array<T, S> A;
    const int lane = threadIdx.x;
    A a;
    if constexpr (VOL) {
        a = trove::warp_load<A>(static_cast<const volatile T*>(src), offset[lane], stride[lane]);
    } else {
        a = trove::warp_load<A>(src, offset[lane], stride[lane]);
    }
    T out[S];
    trove::make_carray<T, S>(a, out);
    for (int k = 0; k < S; ++k) loaded[lane * S + k] = out[k];
    trove::warp_store(a, stored, offset[lane], stride[lane]);
    if constexpr (VOL) {
        trove::uncoalesced_store(a, static_cast<volatile T*>(unc + lane * S), 3);
    } else {
        trove::uncoalesced_store(a, unc + lane * S, 3);
    }
}

// Tables of the public compile-time helpers, evaluated by the product
// templates at compile time and copied out by a kernel.
template <int N>
struct itable {
    int v[N];
};

constexpr int kNT = 65;  // arguments 0..64

template <size_t... I>
constexpr itable<sizeof...(I)> mod_inverse_tab(std::index_sequence<I...>) {
    // entry a * 64 + (m - 1) = static_mod_inverse<a, m>, a in [0, 64], m in [1, 64]
    return {{trove::static_mod_inverse<int(I / 64), int(I % 64) + 1>::value...}};
}
template <size_t... I>
constexpr itable<sizeof...(I)> mod_inverse_r_tab(std::index_sequence<I...>) {
    // entry (r - 1) * 33 * 32 + a * 32 + (m - 1) = static_mod_inverse<a, m, r>,
    // r in [1, 6], a in [0, 32], m in [1, 32]
    return {{trove::static_mod_inverse<int((I / 32) % 33), int(I % 32) + 1, int(I / (33 * 32)) + 1>::value...}};
}
template <size_t... I>
constexpr itable<sizeof...(I)> gcd_tab(std::index_sequence<I...>) {
    // entry u * 65 + v = static_gcd<u, v>, u, v in [0, 64]
    return {{trove::static_gcd<int(I / kNT), int(I % kNT)>::value...}};
}
template <size_t... I>
constexpr itable<sizeof...(I)> log_tab(std::index_sequence<I...>) {
    // entry m - 1 = static_log<m>, m in [1, 1024]
    return {{trove::static_log<int(I) + 1>::value...}};
}
template <size_t... I>
constexpr itable<sizeof...(I)> bits_tab(std::index_sequence<I...>) {
    // entry m = is_power_of_two<m> + 2 * is_odd<m>, m in [0, 1024]
    return {{(int(trove::is_power_of_two<int(I)>::value) + 2 * int(trove::is_odd<int(I)>::value))...}};
}
// A compile-time map on [0, 64) for inverse<>: p -> (5 p + 3) mod 64 (a
// bijection, so every x has a preimage in every run of 64 arguments).
template <int p>
struct affine64 {
    static constexpr int value = (5 * p + 3) % 64;
};
template <size_t... I>
constexpr itable<sizeof...(I)> inverse_tab(std::index_sequence<I...>) {
    // entry x = inverse<int, affine64, x>, then entry 64 + x with the search
    // starting at p = 10
    return {{trove::inverse<int, affine64, int(I % 64), (I < 64 ? 0 : 10)>::value...}};
}

constexpr int kInvN = 65 * 64, kInvRN = 6 * 33 * 32, kGcdN = kNT * kNT, kLogN = 1024, kBitsN = 1025,
              kInverseN = 128;
__constant__ itable<kInvN> c_inv = mod_inverse_tab(std::make_index_sequence<kInvN>{});
__constant__ itable<kInvRN> c_inv_r = mod_inverse_r_tab(std::make_index_sequence<kInvRN>{});
__constant__ itable<kGcdN> c_gcd = gcd_tab(std::make_index_sequence<kGcdN>{});
__constant__ itable<kLogN> c_log = log_tab(std::make_index_sequence<kLogN>{});
__constant__ itable<kBitsN> c_bits = bits_tab(std::make_index_sequence<kBitsN>{});
__constant__ itable<kInverseN> c_inverse = inverse_tab(std::make_index_sequence<kInverseN>{});

// out: [inv | inv_r | gcd | log | bits | inverse], in that order.
__global__ __launch_bounds__(256) void k_static_tables(int* out) {
    for (int i = threadIdx.x; i < kInvN; i += 256) out[i] = c_inv.v[i];
    out += kInvN;
    for (int i = threadIdx.x; i < kInvRN; i += 256) out[i] = c_inv_r.v[i];
    out += kInvRN;
    for (int i = threadIdx.x; i < kGcdN; i += 256) out[i] = c_gcd.v[i];
    out += kGcdN;
    for (int i = threadIdx.x; i < kLogN; i += 256) out[i] = c_log.v[i];
    out += kLogN;
    for (int i = threadIdx.x; i < kBitsN; i += 256) out[i] = c_bits.v[i];
    out += kBitsN;
    for (int i = threadIdx.x; i < kInverseN; i += 256) out[i] = c_inverse.v[i];
}

}  // namespace

extern "C" {

// Element bytes eb (4, 8 or 16) and array length s; src holds the words
// warp_load reads, offset / stride one int per lane; loaded and unc are
// 64 * s elements, stored must cover every offset + k * stride. vol != 0:
// the volatile overloads (4- and 8-byte elements). Returns -1 for an
// unsupported combination, -2 on a launch error.
int devtest_memory(int eb, int s, int vol, const void* src, const int* offset, const int* stride, void* loaded,
                   void* stored, void* unc) {
#define MEM(E, T, S, V)                                                                                       \
    if (eb == E && s == S && (vol != 0) == V) {                                                              \
        hipLaunchKernelGGL((k_memory<T, S, V>), dim3(1), dim3(64), 0, 0, (const T*)src, offset, stride,      \
                           (T*)loaded, (T*)stored, (T*)unc);                                                  \
        return hipDeviceSynchronize() == hipSuccess ? 0 : -2;                                                \
    }
    MEM(4, uint32_t, 1, false) MEM(4, uint32_t, 2, false) MEM(4, uint32_t, 5, false) MEM(4, uint32_t, 16, false)
    MEM(8, uint64_t, 3, false) MEM(16, w16, 2, false) MEM(16, w16, 5, false)
    MEM(4, uint32_t, 5, true) MEM(8, uint64_t, 3, true)
#undef MEM
    return -1;
}

// Sizes of the sections devtest_static_tables writes, in order.
void devtest_static_table_sizes(int* sizes) {
    sizes[0] = kInvN;
    sizes[1] = kInvRN;
    sizes[2] = kGcdN;
    sizes[3] = kLogN;
    sizes[4] = kBitsN;
    sizes[5] = kInverseN;
}

int devtest_static_tables(int* out) {
    hipLaunchKernelGGL(k_static_tables, dim3(1), dim3(256), 0, 0, out);
    return hipDeviceSynchronize() == hipSuccess ? 0 : -2;
}

}  // extern "C"
