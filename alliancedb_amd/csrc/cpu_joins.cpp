// cpu_joins.cpp -- PRO_CPU / NPO_CPU: the multithreaded CPU joins with the
// drop-in signature, the baseline a driver selects by name next to the GPU
// joins (the reference's algos[] registry, hashing/main.cpp:306-338).  They run
// on host threads only; the GPU entry points never call them.
//
// PRO_CPU follows the reference's PRO phase structure
// (parallel_radix_join.cpp:981-1473): pass 1 partitions each relation by the
// low radix bits (per-thread histograms, offsets partition-major /
// thread-minor, parallel_radix_partition :739-806), pass 2 clusters every
// pass-1 partition by the next bits (radix_cluster :573-614, here out of place
// and taken from a shared counter instead of the task queue), and each final
// partition pair is joined by bucket chaining (bucket_chaining_join :291-354:
// N = next power of two >= |R_p|, bucket = (key >> radix bits) & (N - 1)).
// Default radix configuration: the reference's NUM_RADIX_BITS = 8 in
// NUM_PASSES = 2 (prj_params.h:16, 21); the options "radix_bits" and "passes"
// override it.  Unlike the reference it never writes into its inputs.
//
// NPO_CPU follows no_partitioning_join.cpp:156-476: one table over all of R
// with next_pow2(|R| / 2) buckets indexed by key & (buckets - 1)
// (allocate_hashtable, common_functions.cpp:74-102), built by all threads,
// then probed by all threads.  A bucket here is the head of a chain through a
// per-tuple next array, pushed with an atomic exchange (the reference latches
// a bucket and appends to 2-tuple overflow buckets, common_functions.cpp:
// 350-400); the matches are the same.
//
// Both return count and checksum (sum of R.payloadID << 32 | S.payloadID over
// matches, mod 2^64) in an aj_result_t; total_ms is the wall time of the join
// including its thread start-up and temporaries, device_ms is 0.  Threads:
// nthreads, or the host's hardware threads when nthreads <= 0.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "aj_internal.h"

namespace aj {

namespace {

inline uint32_t key32(uint64_t t) { return (uint32_t)(t >> 32); }
inline uint64_t pairv(uint64_t r, uint64_t s) { return (r << 32) | (uint32_t)s; }

template <class F> void run_threads(int T, F &&f) {
    std::vector<std::thread> th;
    th.reserve(T);
    for (int t = 0; t < T; ++t) th.emplace_back(f, t);
    for (auto &x : th) x.join();
}

uint64_t pow2_at_least(uint64_t v) {
    uint64_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

// Pass 1 of one relation: thread t takes rows [t n / T, (t+1) n / T), digit
// key & (F - 1); out is partition-major, thread-minor.  off[d] = start of
// partition d (F + 1 entries).
void partition_pass1(const uint64_t *in, uint64_t n, int T, int bits, uint64_t *out,
                     std::vector<uint64_t> &off) {
    const uint32_t F = 1u << bits, mask = F - 1;
    std::vector<uint64_t> hist((size_t)T * F, 0);
    run_threads(T, [&](int t) {
        const uint64_t b = n * t / T, e = n * (t + 1) / T;
        uint64_t *h = &hist[(size_t)t * F];
        for (uint64_t i = b; i < e; ++i) ++h[key32(in[i]) & mask];
    });
    off.assign(F + 1, 0);
    std::vector<uint64_t> base((size_t)T * F);
    uint64_t run = 0;
    for (uint32_t d = 0; d < F; ++d) {
        off[d] = run;
        for (int t = 0; t < T; ++t) {
            base[(size_t)t * F + d] = run;
            run += hist[(size_t)t * F + d];
        }
    }
    off[F] = run;
    run_threads(T, [&](int t) {
        const uint64_t b = n * t / T, e = n * (t + 1) / T;
        uint64_t *cur = &base[(size_t)t * F];
        for (uint64_t i = b; i < e; ++i) out[cur[key32(in[i]) & mask]++] = in[i];
    });
}

// Pass 2: every pass-1 partition of both relations (2 * F1 jobs from a shared
// counter) is clustered by (key >> b1) & (F2 - 1) into out at the same range.
// off2[rel][d1 * F2 + d2] = start of final partition (F1 * F2 + 1 entries).
void partition_pass2(const uint64_t *in[2], uint64_t *out[2], const std::vector<uint64_t> off1[2],
                     int T, int b1, int b2, std::vector<uint64_t> off2[2]) {
    const uint32_t F1 = 1u << b1, F2 = 1u << b2, mask = F2 - 1;
    for (int r = 0; r < 2; ++r) off2[r].assign((size_t)F1 * F2 + 1, 0);
    std::atomic<uint32_t> next{0};
    run_threads(T, [&](int) {
        std::vector<uint64_t> cur(F2);
        for (uint32_t job; (job = next.fetch_add(1)) < 2 * F1;) {
            const int r = job >= F1;
            const uint32_t d1 = job - (r ? F1 : 0);
            const uint64_t b = off1[r][d1], e = off1[r][d1 + 1];
            std::fill(cur.begin(), cur.end(), 0);
            for (uint64_t i = b; i < e; ++i) ++cur[(key32(in[r][i]) >> b1) & mask];
            uint64_t run = b;
            for (uint32_t d2 = 0; d2 < F2; ++d2) {
                off2[r][(size_t)d1 * F2 + d2] = run;
                const uint64_t c = cur[d2];
                cur[d2] = run;
                run += c;
            }
            for (uint64_t i = b; i < e; ++i) {
                const uint64_t v = in[r][i];
                out[r][cur[(key32(v) >> b1) & mask]++] = v;
            }
        }
    });
    for (int r = 0; r < 2; ++r) off2[r][(size_t)F1 * F2] = off1[r][F1];
}

struct Acc {
    uint64_t count = 0, sum = 0;
    char pad[48];
};

result_t *make_result(uint64_t count, uint64_t sum, int nthreads, int algo, float ms) {
    aj_result_t *res = (aj_result_t *)calloc(1, sizeof(aj_result_t));
    if (!res) {
        set_error(AJ_ERR_NOMEM, "host allocation of the result failed");
        return nullptr;
    }
    res->base.totalresults = (int64_t)count;
    res->base.resultlist = nullptr;
    res->base.nthreads = nthreads;
    res->checksum = sum;
    res->status = AJ_OK;
    res->total_ms = ms;
    res->algo = algo;
    return &res->base;
}

bool check(const relation_t *r, const char *name) {
    if (!r || (r->num_tuples && !r->tuples)) {
        set_error(AJ_ERR_ARG, "relation %s is NULL or has no tuple array", name);
        return false;
    }
    if (r->num_tuples >= (1ull << 31)) {
        set_error(AJ_ERR_ARG, "relation %s too large (limit 2^31-1 tuples)", name);
        return false;
    }
    return true;
}

int thread_count(int nthreads) {
    if (nthreads > 0) return std::min(nthreads, 1024);
    const unsigned h = std::thread::hardware_concurrency();
    return h ? (int)h : 1;
}

}  // namespace

result_t *pro_cpu(relation_t *relR, relation_t *relS, int nthreads) {
    if (!check(relR, "R") || !check(relS, "S")) return nullptr;
    const auto t0 = std::chrono::steady_clock::now();
    const int T = thread_count(nthreads);
    const uint64_t nR = relR->num_tuples, nS = relS->num_tuples;
    int bits = (int)aj_get_option("radix_bits");
    int passes = (int)aj_get_option("passes");
    if (bits <= 0) bits = 8;
    if (passes <= 0) passes = 2;
    if (bits > 24) bits = 24;
    if (passes == 2 && bits < 2) passes = 1;
    const int b1 = passes == 2 ? bits / 2 : bits, b2 = bits - b1;
    uint64_t count = 0, sum = 0;
    if (nR && nS) {
        const uint64_t *in[2] = {reinterpret_cast<const uint64_t *>(relR->tuples),
                                 reinterpret_cast<const uint64_t *>(relS->tuples)};
        const uint64_t n[2] = {nR, nS};
        std::vector<uint64_t> p1[2], p2[2];
        std::vector<uint64_t> off1[2], off2[2];
        try {
            for (int r = 0; r < 2; ++r) {
                p1[r].resize(n[r]);
                if (b2) p2[r].resize(n[r]);
            }
        } catch (const std::bad_alloc &) {
            set_error(AJ_ERR_NOMEM, "host allocation of the partition buffers failed");
            return nullptr;
        }
        for (int r = 0; r < 2; ++r) partition_pass1(in[r], n[r], T, b1, p1[r].data(), off1[r]);
        const uint64_t *jin[2] = {p1[0].data(), p1[1].data()};
        const std::vector<uint64_t> *off = off1;
        if (b2) {
            const uint64_t *pin[2] = {p1[0].data(), p1[1].data()};
            uint64_t *pout[2] = {p2[0].data(), p2[1].data()};
            partition_pass2(pin, pout, off1, T, b1, b2, off2);
            jin[0] = p2[0].data();
            jin[1] = p2[1].data();
            off = off2;
        }
        const uint64_t P = off[0].size() - 1;
        std::atomic<uint64_t> next{0};
        std::vector<Acc> acc(T);
        run_threads(T, [&](int t) {
            std::vector<uint32_t> bucket, chain;
            uint64_t c = 0, x = 0;
            for (uint64_t p; (p = next.fetch_add(1)) < P;) {
                const uint64_t rb = off[0][p], nr = off[0][p + 1] - rb;
                const uint64_t sb = off[1][p], ns = off[1][p + 1] - sb;
                if (!nr || !ns) continue;
                const uint64_t N = pow2_at_least(nr);
                bucket.assign(N, 0);
                chain.resize(nr);
                const uint64_t *Rp = jin[0] + rb, *Sp = jin[1] + sb;
                for (uint64_t i = 0; i < nr; ++i) {
                    const uint32_t h = (key32(Rp[i]) >> bits) & (uint32_t)(N - 1);
                    chain[i] = bucket[h];
                    bucket[h] = (uint32_t)(i + 1);
                }
                for (uint64_t j = 0; j < ns; ++j) {
                    const uint64_t s = Sp[j];
                    const uint32_t k = key32(s);
                    for (uint32_t i = bucket[(k >> bits) & (uint32_t)(N - 1)]; i; i = chain[i - 1])
                        if (key32(Rp[i - 1]) == k) {
                            ++c;
                            x += pairv(Rp[i - 1], s);
                        }
                }
            }
            acc[t].count = c;
            acc[t].sum = x;
        });
        for (const Acc &a : acc) {
            count += a.count;
            sum += a.sum;
        }
    }
    const float ms =
        std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return make_result(count, sum, nthreads, AJ_ALGO_PRO, ms);
}

result_t *npo_cpu(relation_t *relR, relation_t *relS, int nthreads) {
    if (!check(relR, "R") || !check(relS, "S")) return nullptr;
    const auto t0 = std::chrono::steady_clock::now();
    const int T = thread_count(nthreads);
    const uint64_t nR = relR->num_tuples, nS = relS->num_tuples;
    uint64_t count = 0, sum = 0;
    if (nR && nS) {
        const uint64_t *R = reinterpret_cast<const uint64_t *>(relR->tuples);
        const uint64_t *S = reinterpret_cast<const uint64_t *>(relS->tuples);
        const uint64_t nb = pow2_at_least(std::max<uint64_t>(nR / 2, 1));
        const uint32_t mask = (uint32_t)(nb - 1);
        std::unique_ptr<std::atomic<uint32_t>[]> head;
        std::vector<uint32_t> chain;
        try {
            head.reset(new std::atomic<uint32_t>[nb]);
            chain.resize(nR);
        } catch (const std::bad_alloc &) {
            set_error(AJ_ERR_NOMEM, "host allocation of the hash table failed");
            return nullptr;
        }
        run_threads(T, [&](int t) {
            const uint64_t b = nb * t / T, e = nb * (t + 1) / T;
            for (uint64_t i = b; i < e; ++i) head[i].store(0, std::memory_order_relaxed);
        });
        run_threads(T, [&](int t) {
            const uint64_t b = nR * t / T, e = nR * (t + 1) / T;
            for (uint64_t i = b; i < e; ++i)
                chain[i] = head[key32(R[i]) & mask].exchange((uint32_t)(i + 1),
                                                             std::memory_order_relaxed);
        });
        std::vector<Acc> acc(T);
        run_threads(T, [&](int t) {
            const uint64_t b = nS * t / T, e = nS * (t + 1) / T;
            uint64_t c = 0, x = 0;
            for (uint64_t j = b; j < e; ++j) {
                const uint32_t k = key32(S[j]);
                for (uint32_t i = head[k & mask].load(std::memory_order_relaxed); i; i = chain[i - 1])
                    if (key32(R[i - 1]) == k) {
                        ++c;
                        x += pairv(R[i - 1], S[j]);
                    }
            }
            acc[t].count = c;
            acc[t].sum = x;
        });
        for (const Acc &a : acc) {
            count += a.count;
            sum += a.sum;
        }
    }
    const float ms =
        std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return make_result(count, sum, nthreads, AJ_ALGO_NPO, ms);
}

}  // namespace aj

extern "C" {

result_t *PRO_CPU(relation_t *relR, relation_t *relS, int nthreads) {
    return aj::pro_cpu(relR, relS, nthreads);
}

result_t *NPO_CPU(relation_t *relR, relation_t *relS, int nthreads) {
    return aj::npo_cpu(relR, relS, nthreads);
}

}  // extern "C"
