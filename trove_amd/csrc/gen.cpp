// Deterministic host-side input generators (seeded; no GPU needed).
#include <trove_amd.h>

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
