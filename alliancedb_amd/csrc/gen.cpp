// gen.cpp -- seeded input generators (host, multithreaded, thread-count
// independent output).
//
// The reference generates one window per run: unique keys by a Knuth/Sattolo
// shuffle of 1..n (generator.cpp:85-93, 285-338) and Zipf keys from a CDF
// look-up table with binary search (genzipf.cpp:100-207), from glibc rand().
// These generators give the same distributions from a counter-based stream so
// that any element can be produced independently (parallel, reproducible):
//
//   stream value i = splitmix64 output number i of the seed
//   unique keys    = 1 + P(i), P a seeded 4-round Feistel permutation of
//                    [0, 2^(2h)) restricted to [0, n) by cycle walking
//   uniform keys   = lo + floor(u * (hi - lo + 1)) with 64-bit multiply-shift
//   Zipf keys      = rank by rejection-inversion (Hoermann & Derflinger,
//                    ACM TOMACS 6(3), 1996), then 1 + P(rank - 1)
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/alliancedb/aj.h"

namespace {

constexpr uint64_t GOLDEN = 0x9E3779B97F4A7C15ull;

inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// output number i (0-based) of splitmix64 started at `seed`
inline uint64_t stream_at(uint64_t seed, uint64_t i) { return mix64(seed + (i + 1) * GOLDEN); }
inline double unit(uint64_t x) { return (double)(x >> 11) * (1.0 / 9007199254740992.0); }

struct Feistel {
    uint64_t n;
    int half;  // bits per half
    uint64_t mask;
    uint64_t keys[4];
    Feistel(uint64_t n_, uint64_t seed) : n(n_) {
        half = 1;
        while ((1ull << (2 * half)) < n) ++half;
        mask = (1ull << half) - 1;
        for (int r = 0; r < 4; ++r) keys[r] = stream_at(seed ^ 0xA5A5A5A5DEADBEEFull, r);
    }
    uint64_t round_trip(uint64_t x) const {
        uint64_t l = x >> half, r = x & mask;
        for (int k = 0; k < 4; ++k) {
            uint64_t f = mix64(r ^ keys[k]) & mask;
            uint64_t nl = r;
            r = l ^ f;
            l = nl;
        }
        return (l << half) | r;
    }
    uint64_t operator()(uint64_t i) const {
        uint64_t x = i;
        do x = round_trip(x);
        while (x >= n);
        return x;
    }
};

template <class F> void parallel_for(uint64_t n, F f) {
    unsigned T = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    if (n < (1u << 16)) T = 1;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; ++t) {
        uint64_t b = n * t / T, e = n * (t + 1) / T;
        th.emplace_back([=] {
            for (uint64_t i = b; i < e; ++i) f(i);
        });
    }
    for (auto &x : th) x.join();
}

// Rejection-inversion sampler for P(k) ~ k^-q on {1..N}, q > 0.
struct Zipf {
    double q, hx1, hn, s;
    uint64_t N;
    static double helper1(double x) { return std::fabs(x) > 1e-8 ? std::log1p(x) / x : 1 - x * (0.5 - x / 3.0); }
    static double helper2(double x) { return std::fabs(x) > 1e-8 ? std::expm1(x) / x : 1 + x * 0.5 * (1 + x / 3.0); }
    double h(double x) const { return std::exp(-q * std::log(x)); }
    double H(double x) const {
        const double lx = std::log(x);
        return helper2((1 - q) * lx) * lx;
    }
    double Hinv(double y) const {
        double t = y * (1 - q);
        if (t < -1) t = -1;
        return std::exp(helper1(t) * y);
    }
    Zipf(uint64_t n, double q_) : q(q_), N(n) {
        hx1 = H(1.5) - 1.0;
        hn = H((double)n + 0.5);
        s = 2 - Hinv(H(2.5) - h(2));
    }
    uint64_t sample(uint64_t seed, uint64_t i) const {
        double x = 1;
        uint64_t k = 1;
        for (uint64_t j = 0; j < 64; ++j) {
            const double u = hn + unit(stream_at(seed, i * 64 + j)) * (hx1 - hn);
            x = Hinv(u);
            double kd = std::floor(x + 0.5);
            if (kd < 1) kd = 1;
            if (kd > (double)N) kd = (double)N;
            k = (uint64_t)kd;
            if (kd - x <= s || u >= H(kd + 0.5) - h(kd)) return k;
        }
        return k;
    }
};

}  // namespace

extern "C" {

int aj_gen_unique(tuple_t *out, uint64_t n, uint64_t seed) {
    if (!out && n) return AJ_ERR_ARG;
    if (n >= (1ull << 31)) return AJ_ERR_ARG;
    Feistel P(std::max<uint64_t>(n, 1), seed);
    parallel_for(n, [&](uint64_t i) {
        out[i].key = (int32_t)(1 + P(i));
        out[i].payloadID = (int32_t)i;
    });
    return AJ_OK;
}

int aj_gen_uniform(tuple_t *out, uint64_t n, uint32_t lo, uint32_t hi, uint64_t seed) {
    if ((!out && n) || hi < lo) return AJ_ERR_ARG;
    const uint64_t range = (uint64_t)hi - lo + 1;
    parallel_for(n, [&](uint64_t i) {
        const uint64_t r = (uint64_t)(((__uint128_t)stream_at(seed, i) * range) >> 64);
        out[i].key = (int32_t)(lo + r);
        out[i].payloadID = (int32_t)i;
    });
    return AJ_OK;
}

int aj_gen_zipf(tuple_t *out, uint64_t n, uint32_t nkeys, double theta, uint64_t seed) {
    if ((!out && n) || nkeys == 0 || !(theta >= 0) || theta > 10) return AJ_ERR_ARG;
    Feistel P(nkeys, seed ^ 0x5bd1e995ull);
    if (theta == 0.0) {
        parallel_for(n, [&](uint64_t i) {
            const uint64_t r = (uint64_t)(((__uint128_t)stream_at(seed, i) * nkeys) >> 64);
            out[i].key = (int32_t)(1 + P(r));
            out[i].payloadID = (int32_t)i;
        });
        return AJ_OK;
    }
    Zipf z(nkeys, theta);
    parallel_for(n, [&](uint64_t i) {
        const uint64_t rank = z.sample(seed, i);
        out[i].key = (int32_t)(1 + P(rank - 1));
        out[i].payloadID = (int32_t)i;
    });
    return AJ_OK;
}

}  // extern "C"
