/*
 * CPU oracle for the Trove array-of-structures load/store path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker that the HIP library is
 * compared against (and the CPU baseline of the benchmark). Nothing in the
 * product path may call it: only tests/, the smoke check and the benchmark's
 * CPU-baseline leg load liboracle.so.
 *
 * Contents, each restating the reference (bryancatanzaro/trove) at the cited
 * lines:
 *   1. semantics of the access paths as scalar loops (copy, gather, scatter,
 *      the benchmark's counting store and summing load, the block interface);
 *   2. the reference's compile-time number theory (gcd, modular inverse,
 *      floor log2, power-of-two / odd tests, inverse of a map), its strided
 *      per-lane memory access (warp_load / warp_store / uncoalesced_store)
 *      and its runtime rotation;
 *   3. the reference's R2C / C2R lane schedule (three schemas: odd,
 *      power-of-two, composite) simulated for a whole tile of n lanes;
 *   4. OpenMP versions of copy and gather (the CPU baseline).
 * How each part is pinned (golden vectors, library routines, closed forms)
 * is in tests/test_oracle.py.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- 1. semantics --------------------------------------------------- */

/* load_warp_contiguous(p) == *p and store_warp_contiguous(v, p) == (*p = v)
 * (aos.h:66-105): a contiguous copy moves every struct unchanged. Byte loop,
 * one struct at a time. */
void oracle_copy(const uint8_t* src, uint8_t* dst, int64_t n, int s) {
    for (int64_t i = 0; i < n; ++i)
        for (int b = 0; b < s; ++b) dst[i * s + b] = src[i * s + b];
}

/* benchmark_shfl_gather / benchmark_direct_gather (tests/benchmark.cu:90-99,
 * 112-119): r[g] = s[indices[g]]. */
void oracle_gather(const uint8_t* src, const int32_t* idx, uint8_t* dst, int64_t n, int s) {
    for (int64_t i = 0; i < n; ++i) {
        const int64_t j = idx[i];
        for (int b = 0; b < s; ++b) dst[i * s + b] = src[j * s + b];
    }
}

/* benchmark_shfl_scatter / benchmark_direct_scatter (tests/benchmark.cu:
 * 101-110, 121-128): r[indices[g]] = s[g]. */
void oracle_scatter(const uint8_t* src, const int32_t* idx, uint8_t* dst, int64_t n, int s) {
    for (int64_t i = 0; i < n; ++i) {
        const int64_t j = idx[i];
        for (int b = 0; b < s; ++b) dst[j * s + b] = src[i * s + b];
    }
}

/* benchmark_contiguous_*_store (tests/benchmark.cu:52-72): struct g of k
 * 32-bit words is counting_array(g * k) = {g*k, g*k + 1, ...}
 * (utility.h:33-53). */
void oracle_counting_store(uint32_t* dst, int64_t n, int k) {
    for (int64_t g = 0; g < n; ++g) {
        uint32_t v = (uint32_t)(g * k);
        for (int j = 0; j < k; ++j) dst[g * k + j] = v + (uint32_t)j;
    }
}

/* benchmark_contiguous_*_load (tests/benchmark.cu:74-88): r[g] = sum(s[g]),
 * the wrapping sum of the struct's k words (utility.h:55-78). */
void oracle_load_sum(const uint32_t* src, uint32_t* out, int64_t n, int k) {
    for (int64_t g = 0; g < n; ++g) {
        uint32_t acc = 0;
        for (int j = 0; j < k; ++j) acc += src[g * k + j];
        out[g] = acc;
    }
}

/* test_block_copy (tests/block.cu:34-54): arrays of s elements at index
 * 0..l-1 copied (the converged and fallback branches move the same values);
 * l = e / s, and the e - l*s trailing elements are copied as well. */
void oracle_block_copy(const uint32_t* x, uint32_t* r, int64_t e, int s) {
    const int64_t l = e / s;
    for (int64_t index = 0; index < l; ++index)
        for (int j = 0; j < s; ++j) r[index * s + j] = x[index * s + j];
    for (int64_t j = l * s; j < e; ++j) r[j] = x[j];
}

/* test_block_write (tests/block.cu:13-32): array `index` is
 * counting_array(s * index), so element j of the output is j. */
void oracle_block_write(uint32_t* r, int64_t l, int s) {
    for (int64_t index = 0; index < l; ++index)
        for (int j = 0; j < s; ++j) r[index * s + j] = (uint32_t)(s * index + j);
}

/* ---- 2. number theory and rotation ----------------------------------- */

/* static_gcd (static_gcd.h:36-65): Stein's binary gcd, as the reference's
 * template recursion (both even: 2*gcd(u/2, v/2); one even: halve it; both
 * odd: gcd(|u-v|/2, min(u, v))). Positive arguments; the reference's zero
 * specialisations are `bool` (static_gcd.h:67-75), restated here as the
 * plain values 0 -> v and u -> 0 return. */
int oracle_gcd(int u, int v) {
    if (u == v) return u;
    if (u == 0) return v;
    if (v == 0) return u;
    const int uo = u & 1, vo = v & 1;
    if (!uo && !vo) return oracle_gcd(u >> 1, v >> 1) << 1;
    if (!uo && vo) return oracle_gcd(u >> 1, v);
    if (uo && !vo) return oracle_gcd(u, v >> 1);
    return u > v ? oracle_gcd((u - v) >> 1, v) : oracle_gcd((v - u) >> 1, u);
}

/* static_mod_inverse (static_mod_inverse.h:32-52): the first r = 1, 2, ...,
 * m-1 with (r * a) % m == 1, else -1. */
int oracle_mod_inverse(int a, int m) {
    for (int r = 1; r < m; ++r)
        if ((r * a) % m == 1) return r;
    return -1;
}

/* static_mod_inverse<a, m, r> with an explicit start (static_mod_inverse.h:
 * 32-52 recurses r, r + 1, ... and stops with -1 at r == m): the first
 * r' = r, ..., m - 1 with (r' * a) % m == 1, else -1. */
int oracle_mod_inverse_from(int a, int m, int r) {
    for (; r < m; ++r)
        if ((r * a) % m == 1) return r;
    return -1;
}

/* is_power_of_two / is_odd (utility.h:96-104): (m & (m - 1)) == 0 and
 * (m & 1) == 1, returned as bit 0 and bit 1 of one value. */
int oracle_pow2_odd_bits(int m) { return ((m & (m - 1)) == 0) + 2 * ((m & 1) == 1); }

/* inverse<T, Fn, x, p> (utility.h:121-126): the first p' = p, p + 1, ...
 * with Fn(p') == x, for Fn given as a table fn[0..n_fn); -1 when none lies
 * in the table. */
int oracle_inverse(const int32_t* fn, int n_fn, int x, int p) {
    for (; p < n_fn; ++p)
        if (fn[p] == x) return p;
    return -1;
}

/* warp_load / warp_store / uncoalesced_store (memory.h:101-168) for one
 * lane's array of s elements of eb bytes: element k of the array is
 * ptr[offset + k * stride] for warp_load / warp_store, and ptr[k] for
 * uncoalesced_store (memory.h:59-98 advances its offset by 1 whatever the
 * stride). */
void oracle_warp_load(const uint8_t* ptr, int eb, int s, int offset, int stride, uint8_t* out) {
    for (int k = 0; k < s; ++k) memcpy(out + (int64_t)k * eb, ptr + ((int64_t)offset + (int64_t)k * stride) * eb, (size_t)eb);
}

void oracle_warp_store(const uint8_t* in, int eb, int s, uint8_t* ptr, int offset, int stride) {
    for (int k = 0; k < s; ++k) memcpy(ptr + ((int64_t)offset + (int64_t)k * stride) * eb, in + (int64_t)k * eb, (size_t)eb);
}

void oracle_uncoalesced_store(const uint8_t* in, int eb, int s, uint8_t* ptr, int stride) {
    (void)stride;
    for (int k = 0; k < s; ++k) memcpy(ptr + (int64_t)k * eb, in + (int64_t)k * eb, (size_t)eb);
}

/* static_log (utility.h:80-94): floor(log2 m) for m >= 1. */
int oracle_floor_log2(int m) {
    int r = 0;
    while (m > 1) {
        m >>= 1;
        ++r;
    }
    return r;
}

/* rotate (rotate.h:34-115): stage i = 1, 2, 4, ... < m moves element
 * (j + i) % m into slot j when bit i of a is set, applied in increasing i. */
void oracle_rotate(const int32_t* t, int32_t* r, int m, int a) {
    int32_t cur[1024], nxt[1024];
    memcpy(cur, t, sizeof(int32_t) * m);
    for (int i = 1; i < m; i <<= 1) {
        for (int j = 0; j < m; ++j) nxt[j] = (a & i) ? cur[(j + i) % m] : cur[j];
        memcpy(cur, nxt, sizeof(int32_t) * m);
    }
    memcpy(r, cur, sizeof(int32_t) * m);
}

/* ---- 3. the reference's R2C / C2R schedule --------------------------- */

/* Tile state: lane l, slot j at d[l * m + j]. */

enum { SCHEMA_ODD = 0, SCHEMA_PO2 = 1, SCHEMA_COMPOSITE = 2 };

/* tx_algorithm (transpose.h:38-55). */
int oracle_schema(int m) {
    if ((m & (m - 1)) == 0) return SCHEMA_PO2;
    if (m & 1) return SCHEMA_ODD;
    return SCHEMA_COMPOSITE;
}

static void rotate_lane(int32_t* d, int m, int a) { /* rotate.h:111-115 in place */
    int32_t tmp[1024];
    oracle_rotate(d, tmp, m, a);
    memcpy(d, tmp, sizeof(int32_t) * m);
}

/* tx_permute_impl (transpose.h:92-118): out[x] = in[P(x)]. */
static void permute_lane(int32_t* d, int m, const int* P) {
    int32_t tmp[1024];
    for (int x = 0; x < m; ++x) tmp[x] = d[P[x]];
    memcpy(d, tmp, sizeof(int32_t) * m);
}

/* warp_shuffle (transpose.h:413-437): slot j of lane l takes slot j of lane
 * idx[l][j]. */
static void shuffle_tile(int32_t* d, int n, int m, const int* idx) {
    int32_t* tmp = (int32_t*)malloc(sizeof(int32_t) * n * m);
    for (int l = 0; l < n; ++l)
        for (int j = 0; j < m; ++j) tmp[l * m + j] = d[idx[l * m + j] * m + j];
    memcpy(d, tmp, sizeof(int32_t) * n * m);
    free(tmp);
}

/* composite_c2r_permute_fn (transpose.h:129-138): F(x) = (x*o - x/p) % m with
 * o = n % m, p = m / gcd(m, n). */
static int composite_f(int n, int m, int x) {
    const int o = n % m, p = m / oracle_gcd(m, n);
    return (x * o - x / p) % m;
}

/* r2c_warp_transpose on a whole tile (transpose.h:525-625, 664-696).
 * d holds n lanes of m slots; returns 0, or -1 for unsupported (m > n). */
int oracle_r2c_tile(int32_t* d, int n, int m) {
    if (m < 2 || m > n || n > 1024 || m > 1024) return -1;
    int* idx = (int*)malloc(sizeof(int) * n * m);
    int P[1024];
    const int schema = oracle_schema(m);
    const int mask = n - 1;
    if (schema == SCHEMA_ODD) {
        /* r2c_offset_constants<odd>::permute = inv(n, m) (transpose.h:82-85);
         * rotation = warp_id * (m - permute) % m (transpose.h:528-540);
         * offsets: (warp_id * m + (n % m * j) % m) & mask (transpose.h:242-250,
         * 320-342); steps rotate, shuffle, permute (transpose.h:581-592). */
        const int permute = oracle_mod_inverse(n, m);
        for (int x = 0; x < m; ++x) P[x] = (permute * x) % m;
        for (int l = 0; l < n; ++l) {
            const int init = (l * m) & mask;
            for (int j = 0; j < m; ++j) idx[l * m + j] = (init + ((n % m) * j) % m) & mask;
            rotate_lane(d + l * m, m, (l * (m - permute)) % m);
        }
        shuffle_tile(d, n, m, idx);
        for (int l = 0; l < n; ++l) permute_lane(d + l * m, m, P);
    } else if (schema == SCHEMA_PO2) {
        /* rotation = warp_id % m; initial offset ((warp_id << log m) +
         * (warp_id >> log(n/m))) & mask, lb = initial & ~(m-1), next = (o == lb)
         * ? o + m - 1 : o - 1 (transpose.h:345-363, 542-561); steps rotate,
         * shuffle, permute by (m-1)x, rotate by (m - warp_id/(n/m)) % m
         * (transpose.h:594-608). */
        const int lm = oracle_floor_log2(m), lnm = oracle_floor_log2(n / m);
        for (int x = 0; x < m; ++x) P[x] = ((m - 1) * x) % m;
        for (int l = 0; l < n; ++l) {
            int o = ((l << lm) + (l >> lnm)) & mask;
            const int lb = o & ~(m - 1);
            for (int j = 0; j < m; ++j) {
                idx[l * m + j] = o;
                o = (o == lb) ? o + m - 1 : o - 1;
            }
            rotate_lane(d + l * m, m, l % m);
        }
        shuffle_tile(d, n, m, idx);
        for (int l = 0; l < n; ++l) {
            permute_lane(d + l * m, m, P);
            rotate_lane(d + l * m, m, (m - l / (n / m)) % m);
        }
    } else {
        /* composite: permute by F^-1, rotate by m - warp_id % m, shuffle with
         * offsets lb = m*warp_id & mask, ub = lb + m, o0 = lb + warp_id/(n/c),
         * o -> o+1 wrapping ub -> lb, then rotate by m - warp_id/(n/c)
         * (transpose.h:129-147, 366-400, 563-576, 610-625). */
        const int c = oracle_gcd(m, n);
        for (int x = 0; x < m; ++x) {
            /* composite_r2c_permute_fn: the preimage of x under F (utility.h
             * inverse, transpose.h:140-147). */
            int y = 0;
            while (composite_f(n, m, y) != x) ++y;
            P[x] = y;
        }
        for (int l = 0; l < n; ++l) {
            permute_lane(d + l * m, m, P);
            rotate_lane(d + l * m, m, m - (l % m));
            const int lb = (m * l) & mask, ub = lb + m;
            int o = lb + l / (n / c);
            for (int j = 0; j < m; ++j) {
                idx[l * m + j] = o & mask;
                o = (o + 1 == ub) ? lb : o + 1;
            }
        }
        shuffle_tile(d, n, m, idx);
        for (int l = 0; l < n; ++l) rotate_lane(d + l * m, m, m - l / (n / c));
    }
    free(idx);
    return 0;
}

/* c2r_warp_transpose on a whole tile (transpose.h:440-523, 629-662). */
int oracle_c2r_tile(int32_t* d, int n, int m) {
    if (m < 2 || m > n || n > 1024 || m > 1024) return -1;
    int* idx = (int*)malloc(sizeof(int) * n * m);
    int P[1024];
    const int schema = oracle_schema(m);
    const int mask = n - 1;
    if (schema == SCHEMA_ODD || schema == SCHEMA_PO2) {
        int offset, rot_const = 0, permute;
        if (schema == SCHEMA_ODD) {
            /* c2r_offset_constants<odd> (transpose.h:60-65). */
            offset = n - oracle_mod_inverse(m, n);
            rot_const = oracle_mod_inverse(n, m);
            permute = oracle_mod_inverse(rot_const, m);
        } else {
            /* c2r_offset_constants<power_of_two> (transpose.h:67-71). */
            offset = n - n / m;
            permute = m - 1;
        }
        for (int x = 0; x < m; ++x) P[x] = (permute * x) % m;
        const int lm = oracle_floor_log2(m), ln = oracle_floor_log2(n);
        for (int l = 0; l < n; ++l) {
            /* initial offsets (transpose.h:253-261 odd, 228-237 po2); the
             * sequence adds `offset` and reduces once mod n (185-226). */
            int o = schema == SCHEMA_ODD ? ((n - l) * offset) & mask : ((l * (n + 1)) >> lm) & mask;
            for (int j = 0; j < m; ++j) {
                if (o >= n) o -= n;
                idx[l * m + j] = o;
                o += offset;
            }
            /* po2 pre-rotation by warp_id >> (log n - log m)
             * (transpose.h:492-507). */
            if (schema == SCHEMA_PO2) rotate_lane(d + l * m, m, l >> (ln - lm));
        }
        shuffle_tile(d, n, m, idx);
        for (int l = 0; l < n; ++l) {
            /* rotation: warp_id * r % m (odd, transpose.h:443-453) or
             * (m - warp_id) & (m - 1) (po2, transpose.h:455-464); steps
             * permute then rotate (transpose.h:480-490). */
            const int rot = schema == SCHEMA_ODD ? (l * rot_const) % m : (m - l) & (m - 1);
            permute_lane(d + l * m, m, P);
            rotate_lane(d + l * m, m, rot);
        }
    } else {
        /* composite (transpose.h:73-77, 263-309, 466-475, 509-523). */
        const int c = oracle_gcd(m, n);
        const int k = oracle_mod_inverse(m / c, n / c);
        const int lc = oracle_floor_log2(c), lndc = oracle_floor_log2(n / c);
        for (int x = 0; x < m; ++x) P[x] = composite_f(n, m, x);
        for (int l = 0; l < n; ++l) {
            int i = l;
            for (int p = 0; p < m; ++p) {
                idx[l * m + p] = ((((i >> lc) * k) & (n / c - 1)) + ((i & (c - 1)) << lndc)) & mask;
                int ni = i + n - 1;
                if (p == m - c + (l & (c - 1))) ni += m;
                i = ni;
            }
            rotate_lane(d + l * m, m, l >> lndc);
        }
        shuffle_tile(d, n, m, idx);
        for (int l = 0; l < n; ++l) {
            rotate_lane(d + l * m, m, l % m);
            permute_lane(d + l * m, m, P);
        }
    }
    free(idx);
    return 0;
}

/* ---- 4. OpenMP CPU baseline ------------------------------------------ */

int oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void oracle_copy_omp(const uint8_t* src, uint8_t* dst, int64_t n, int s) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) memcpy(dst + i * s, src + i * s, (size_t)s);
}

void oracle_gather_omp(const uint8_t* src, const int32_t* idx, uint8_t* dst, int64_t n, int s) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) memcpy(dst + i * s, src + (int64_t)idx[i] * s, (size_t)s);
}
