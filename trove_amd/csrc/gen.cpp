// Deterministic host-side input generators (seeded; no GPU needed).
#include <trove_amd.h>

#include <cmath>
#include <cstdint>

namespace {

inline uint64_t splitmix_next(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

// xoshiro256** (Blackman & Vigna), state seeded from splitmix64.
struct xoshiro256ss {
    uint64_t s[4];
    explicit xoshiro256ss(uint64_t seed) {
        for (uint64_t& w : s) w = splitmix_next(seed);
    }
    uint64_t next() {
        const uint64_t result = rotl64(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl64(s[3], 45);
        return result;
    }
    // Uniform integer in [0, bound), bound > 0 (Lemire's multiply-and-reject).
    uint64_t below(uint64_t bound) {
        unsigned __int128 m = (unsigned __int128)next() * bound;
        uint64_t lo = (uint64_t)m;
        if (lo < bound) {
            const uint64_t thresh = (0 - bound) % bound;
            while (lo < thresh) {
                m = (unsigned __int128)next() * bound;
                lo = (uint64_t)m;
            }
        }
        return (uint64_t)(m >> 64);
    }
};

}  // namespace

extern "C" int trove_gen_permutation(int32_t* out, int64_t n, uint64_t seed) {
    if (n < 0 || n > INT32_MAX || (n > 0 && !out)) return TROVE_ERR_INVALID;
    for (int64_t i = 0; i < n; ++i) out[i] = (int32_t)i;
    xoshiro256ss rng(seed);
    for (int64_t i = n - 1; i > 0; --i) {
        const int64_t j = (int64_t)rng.below((uint64_t)i + 1);
        const int32_t t = out[i];
        out[i] = out[j];
        out[j] = t;
    }
    return TROVE_OK;
}

namespace {

inline double uniform01(xoshiro256ss& rng) { return (double)(rng.next() >> 11) * 0x1.0p-53; }

// Zipf(s) over 1..N by rejection-inversion (W. Hoermann, G. Derflinger,
// "Rejection-inversion to generate variates from monotone discrete
// distributions", ACM TOMACS 6(3), 1996): invert the integral H of the hat
// h(x) = x^-s over [1.5 - 1, N + 0.5] and accept k = round(x) when x is
// within the squeeze or u lies under H(k + 0.5) - h(k).
struct zipf_sampler {
    double s, hx1, hn, squeeze;
    int64_t n;
    static double expm1_over_x(double x) { return std::fabs(x) > 1e-8 ? std::expm1(x) / x : 1 + x * 0.5 * (1 + x / 3 * (1 + 0.25 * x)); }
    static double log1p_over_x(double x) { return std::fabs(x) > 1e-8 ? std::log1p(x) / x : 1 - x * (0.5 - x * (1.0 / 3 - 0.25 * x)); }
    double h(double x) const { return std::exp(-s * std::log(x)); }
    // H(x) = (x^(1-s) - 1) / (1 - s), continuous at s = 1 (log x).
    double H(double x) const {
        const double lx = std::log(x);
        return expm1_over_x((1 - s) * lx) * lx;
    }
    double Hinv(double y) const {
        double t = y * (1 - s);
        if (t < -1) t = -1;
        return std::exp(log1p_over_x(t) * y);
    }
    zipf_sampler(int64_t n_, double s_) : s(s_), n(n_) {
        hx1 = H(1.5) - 1;
        hn = H((double)n + 0.5);
        squeeze = 2 - Hinv(H(2.5) - h(2));
    }
    int64_t sample(xoshiro256ss& rng) const {
        for (;;) {
            const double u = hn + uniform01(rng) * (hx1 - hn);
            const double x = Hinv(u);
            int64_t k = (int64_t)(x + 0.5);
            if (k < 1) k = 1;
            if (k > n) k = n;
            if ((double)k - x <= squeeze || u >= H((double)k + 0.5) - h((double)k)) return k;
        }
    }
};

}  // namespace

extern "C" int trove_gen_indices(int32_t* out, int64_t n, int64_t n_src, int pattern, uint64_t seed) {
    if (n < 0 || n_src <= 0 || n_src > INT32_MAX || (n > 0 && !out)) return TROVE_ERR_INVALID;
    xoshiro256ss rng(seed);
    switch (pattern) {
        case TROVE_PATTERN_IDENTITY:
            for (int64_t i = 0; i < n; ++i) out[i] = (int32_t)(i % n_src);
            return TROVE_OK;
        case TROVE_PATTERN_RANDOM:
            if (n == n_src) return trove_gen_permutation(out, n, seed);
            for (int64_t i = 0; i < n; ++i) out[i] = (int32_t)rng.below((uint64_t)n_src);
            return TROVE_OK;
        case TROVE_PATTERN_RUNS: {
            // Geometric run lengths with mean 32: P(len = k) = (1-p)^(k-1) p, p = 1/32.
            const double log_q = std::log(1.0 - 1.0 / 32.0);
            int64_t i = 0;
            while (i < n) {
                const double u = 1.0 - uniform01(rng);  // (0, 1]
                int64_t len = 1 + (int64_t)std::floor(std::log(u) / log_q);
                if (len > n_src) len = n_src;
                const int64_t start = (int64_t)rng.below((uint64_t)(n_src - len + 1));
                for (int64_t k = 0; k < len && i < n; ++k, ++i) out[i] = (int32_t)(start + k);
            }
            return TROVE_OK;
        }
        default:
            return TROVE_ERR_INVALID;
    }
}

extern "C" int trove_gen_zipf(int32_t* out, int64_t n, int64_t n_items, double s, uint64_t seed) {
    if (n < 0 || n_items <= 0 || n_items > INT32_MAX || !(s > 0) || (n > 0 && !out)) return TROVE_ERR_INVALID;
    xoshiro256ss rng(seed);
    const zipf_sampler z(n_items, s);
    for (int64_t i = 0; i < n; ++i) out[i] = (int32_t)(z.sample(rng) - 1);
    return TROVE_OK;
}

extern "C" int trove_gen_loguniform(int64_t* out, int64_t n, int64_t lo, int64_t hi, uint64_t seed) {
    if (n < 0 || lo < 1 || hi < lo || (n > 0 && !out)) return TROVE_ERR_INVALID;
    xoshiro256ss rng(seed);
    const double a = std::log((double)lo), b = std::log((double)hi + 1.0);
    for (int64_t i = 0; i < n; ++i) {
        int64_t v = (int64_t)std::floor(std::exp(a + (b - a) * uniform01(rng)));
        if (v < lo) v = lo;
        if (v > hi) v = hi;
        out[i] = v;
    }
    return TROVE_OK;
}
