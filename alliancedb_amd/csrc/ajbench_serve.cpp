// ajbench_serve.cpp -- served window-join load generator of the benchmark
// driver (ajbench -a SERVE).  No counterpart in the reference, whose driver
// joins one window per run.
//
// Offered load: `streams` stream pairs, each emitting window-join requests
// with exponential inter-arrival times; per-side window sizes lognormal
// (median, sigma), capped; R keys uniform over `domain`, S keys uniform for
// the first half of the streams and Zipf(theta) for the second half.  The
// total offered rate is load x peak tuples/s (R+S).  Client threads walk the
// merged schedule: admit the window (aj_server_window_alloc, which waits while
// the HBM budget is taken), generate its tuples on the device, and submit it
// at its scheduled close time (or as soon as possible when behind).  Latency =
// result on the host - scheduled close time.  Afterwards a sample of windows
// is regenerated and joined alone with a plain one-window launch; the served
// results must be identical.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../include/alliancedb/aj.h"
#include "ajbench_serve.h"

namespace {

struct Req {
    double close_us;
    int stream;
    uint64_t nR, nS, seed;
    double theta;
};

struct Done {
    uint64_t count, checksum;
    double lat_us;
    int batch;
    int status;
};

double pct(std::vector<double> v, double p) {
    if (v.empty()) return 0;
    std::sort(v.begin(), v.end());
    const size_t i = std::min(v.size() - 1, (size_t)std::ceil(p * v.size()) - (p > 0 ? 1 : 0));
    return v[i];
}

uint64_t draw_size(std::lognormal_distribution<double> &size, std::mt19937_64 &rng,
                   uint64_t cap) {
    return std::min<uint64_t>(cap, std::max<uint64_t>(1, (uint64_t)size(rng)));
}

double host_us() {
    return std::chrono::duration<double, std::micro>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// The two reference rates of the served workload, on P.calibrate windows of the
// same distribution generated into the server's window memory ahead of time:
// plain one-window launches (a batch plan sized to the largest window, one
// window per aj_batch_run on one stream, back to back, one sync at the end)
// and the server's batched peak (every window submitted at once, until the
// last result is on the host).  The served results must equal the plain ones.
int run_calibration(const ServeParams &P, aj_ctx_t *ctx, aj_server_t *srv) {
    const int K = P.calibrate;
    std::mt19937_64 rng(P.seed ^ 0xCA11B7A7Eull);
    std::lognormal_distribution<double> size(std::log(P.median), P.sigma);
    std::vector<aj_window_t> win(K);
    uint64_t tuples = 0, maxw = 1;
    for (int i = 0; i < K; ++i) {
        const uint64_t nR = draw_size(size, rng, P.cap_tuples);
        const uint64_t nS = draw_size(size, rng, P.cap_tuples);
        const double theta = (i % P.streams) >= P.streams / 2 ? P.theta : 0.0;
        if (aj_server_window_alloc(srv, nR, nS, -1, &win[i]) != AJ_OK ||
            aj_server_window_fill(srv, &win[i], P.seed * 7919ull + (uint64_t)i, P.domain,
                                  theta) != AJ_OK) {
            fprintf(stderr, "[ERROR] calibration windows: %s\n", aj_last_error());
            return 1;
        }
        tuples += nR + nS;
        maxw = std::max(maxw, std::max(nR, nS));
    }
    if (aj_device_sync() != AJ_OK) {
        fprintf(stderr, "[ERROR] calibration fill: %s\n", aj_last_error());
        return 1;
    }
    aj_batch_opts_t bo;
    memset(&bo, 0, sizeof(bo));
    bo.no_graph = 1;
    aj_batch_plan_t *plan = nullptr;
    void *dres = nullptr, *dbounds = nullptr;
    const uint32_t bounds[2] = {0u, 1u};  // every generated tuple has ts 0
    std::vector<uint64_t> plain(2 * (size_t)K);
    double t_plain = 0;
    int rc = 0;
    if (aj_batch_plan_create(ctx, maxw, maxw, 1, maxw, &bo, &plan) != AJ_OK ||
        aj_device_alloc(&dres, 16 * (uint64_t)K) != AJ_OK ||
        aj_device_alloc(&dbounds, sizeof(bounds)) != AJ_OK ||
        aj_copy_h2d(dbounds, bounds, sizeof(bounds)) != AJ_OK) {
        rc = 1;
    } else {
        uint64_t *res = static_cast<uint64_t *>(dres);
        auto one = [&](int i) {
            return aj_batch_run(plan, win[i].dR, win[i].nR, win[i].dS, win[i].nS,
                                (const uint32_t *)dbounds, 1, res + 2 * i, res + 2 * i + 1,
                                nullptr);
        };
        for (int i = 0; i < std::min(K, 4) && rc == 0; ++i) rc = one(i) != AJ_OK;  // warm
        if (rc == 0) rc = aj_device_sync() != AJ_OK;
        const double t0 = host_us();
        for (int i = 0; i < K && rc == 0; ++i) rc = one(i) != AJ_OK;
        if (rc == 0) rc = aj_device_sync() != AJ_OK;
        t_plain = host_us() - t0;
        if (rc == 0) rc = aj_copy_d2h(plain.data(), dres, 16 * (uint64_t)K) != AJ_OK;
    }
    if (plan) aj_batch_plan_destroy(plan);
    if (dres) aj_device_free(dres);
    if (dbounds) aj_device_free(dbounds);
    if (rc) {
        fprintf(stderr, "[ERROR] calibration plain launches: %s\n", aj_last_error());
        return 1;
    }
    // the batched peak: all windows queued at once
    const double t1 = aj_server_now_us(srv);
    for (int i = 0; i < K; ++i)
        if (aj_server_submit(srv, &win[i], (uint64_t)i, -1) != AJ_OK) {
            fprintf(stderr, "[ERROR] calibration submit: %s\n", aj_last_error());
            return 1;
        }
    aj_server_drain(srv);
    std::vector<aj_server_result_t> got(K);
    int n = 0, mismatches = 0, bad = 0;
    double last = t1;
    while (n < K) {
        int m = 0;
        if (aj_server_poll(srv, got.data() + n, K - n, 1000000, &m) != AJ_OK || m == 0) break;
        n += m;
    }
    for (int i = 0; i < n; ++i) {
        const aj_server_result_t &x = got[i];
        last = std::max(last, x.done_us);
        bad += x.status != AJ_OK;
        if (x.tag >= (uint64_t)K || x.count != plain[2 * x.tag] ||
            x.checksum != plain[2 * x.tag + 1])
            ++mismatches;
    }
    aj_server_stats_t st;
    aj_server_stats(srv, &st);
    const double t_batch = last - t1;
    printf("{\"calibration\": true, \"windows\": %d, \"tuples\": %llu, "
           "\"largest_window\": %llu, \"plain_ms\": %.3f, \"plain_Mtps\": %.1f, "
           "\"batched_ms\": %.3f, \"batched_Mtps\": %.1f, \"batches\": %llu, "
           "\"results\": %d, \"bad_status\": %d, \"mismatches\": %d}\n",
           K, (unsigned long long)tuples, (unsigned long long)maxw, t_plain / 1e3,
           tuples / t_plain, t_batch / 1e3, tuples / t_batch, (unsigned long long)st.batches, n,
           bad, mismatches);
    fflush(stdout);
    return (n != K || bad || mismatches) ? 2 : 0;
}

}  // namespace

int run_serve(const ServeParams &P) {
    aj_ctx_t *ctx = nullptr;
    if (aj_ctx_create(0, &ctx) != AJ_OK) {
        fprintf(stderr, "[ERROR] context: %s\n", aj_last_error());
        return 1;
    }
    // ---- schedule
    const double rate_tps = P.load * P.peak_mtps * 1e6;
    const double mean_side = P.median * std::exp(P.sigma * P.sigma / 2);
    const double req_per_s = rate_tps / (2 * mean_side);
    std::vector<Req> sched;
    std::mt19937_64 rng(P.seed);
    for (int s = 0; s < P.streams; ++s) {
        std::exponential_distribution<double> gap(req_per_s / P.streams);
        std::lognormal_distribution<double> size(std::log(P.median), P.sigma);
        double t = 0;
        for (uint64_t k = 0;; ++k) {
            t += gap(rng) * 1e6;
            if (t > P.seconds * 1e6) break;
            Req r;
            r.close_us = t;
            r.stream = s;
            r.nR = std::min<uint64_t>(P.cap_tuples, std::max<uint64_t>(1, (uint64_t)size(rng)));
            r.nS = std::min<uint64_t>(P.cap_tuples, std::max<uint64_t>(1, (uint64_t)size(rng)));
            r.seed = P.seed * 1000003ull + (uint64_t)s * 1000000007ull + k;
            r.theta = s >= P.streams / 2 ? P.theta : 0.0;
            sched.push_back(r);
        }
    }
    std::sort(sched.begin(), sched.end(), [](const Req &a, const Req &b) { return a.close_us < b.close_us; });
    uint64_t offered = 0;
    for (const Req &r : sched) offered += r.nR + r.nS;

    aj_server_opts_t so;
    memset(&so, 0, sizeof(so));
    so.hbm_budget = (uint64_t)P.budget;
    so.slots = P.slots;
    so.max_batch_windows = P.batch_windows;
    so.max_batch_tuples = std::max<uint64_t>(P.batch_tuples, P.cap_tuples);
    so.max_window_tuples = P.cap_tuples;
    aj_server_t *srv = nullptr;
    if (aj_server_create(ctx, &so, &srv) != AJ_OK) {
        fprintf(stderr, "[ERROR] server: %s\n", aj_last_error());
        return 1;
    }
    if (P.calibrate > 0) {
        const int rc = run_calibration(P, ctx, srv);
        aj_server_destroy(srv);
        aj_ctx_destroy(ctx);
        return rc;
    }
    std::vector<Done> done(sched.size());
    std::atomic<uint64_t> n_done{0};
    std::atomic<int> failed{0};
    std::string err;
    std::mutex err_mu;
    // start the schedule a little after the server clock's origin
    const double t_start = aj_server_now_us(srv) + 20000;

    std::thread poller([&] {
        std::vector<aj_server_result_t> buf(1024);
        while (n_done.load() < sched.size() && !failed.load()) {
            int n = 0;
            if (aj_server_poll(srv, buf.data(), (int)buf.size(), 100000, &n) != AJ_OK) break;
            for (int i = 0; i < n; ++i) {
                const aj_server_result_t &x = buf[i];
                Done &d = done[x.tag];
                d.count = x.count;
                d.checksum = x.checksum;
                d.lat_us = x.done_us - x.close_us;
                d.batch = x.batch_windows;
                d.status = x.status;
            }
            n_done += n;
        }
    });
    std::vector<std::thread> clients;
    for (int c = 0; c < P.clients; ++c) {
        clients.emplace_back([&, c] {
            for (size_t i = c; i < sched.size() && !failed.load(); i += P.clients) {
                const Req &r = sched[i];
                aj_window_t w;
                int rc = aj_server_window_alloc(srv, r.nR, r.nS, -1, &w);
                if (rc == AJ_OK)
                    rc = aj_server_window_fill(srv, &w, r.seed, P.domain, r.theta);
                const double close = t_start + r.close_us;
                double now = aj_server_now_us(srv);
                if (now < close)
                    std::this_thread::sleep_for(std::chrono::microseconds((int64_t)(close - now)));
                if (rc == AJ_OK) rc = aj_server_submit(srv, &w, (uint64_t)i, close);
                if (rc != AJ_OK) {
                    std::lock_guard<std::mutex> lk(err_mu);
                    err = aj_last_error();
                    failed = 1;
                }
            }
        });
    }
    for (auto &t : clients) t.join();
    if (!failed.load()) aj_server_drain(srv);
    poller.join();
    aj_server_stats_t st;
    aj_server_stats(srv, &st);
    if (failed.load()) {
        fprintf(stderr, "[ERROR] serve: %s\n", err.c_str());
        aj_server_destroy(srv);
        aj_ctx_destroy(ctx);
        return 1;
    }
    // throughput: tuples over the span from the first scheduled close to the last result
    std::vector<double> lat;
    lat.reserve(sched.size());
    double last_done = 0;
    uint64_t bad_status = 0, batched = 0;
    for (size_t i = 0; i < sched.size(); ++i) {
        lat.push_back(done[i].lat_us);
        last_done = std::max(last_done, sched[i].close_us + done[i].lat_us);
        bad_status += done[i].status != AJ_OK;
        batched += done[i].batch;
    }
    const double span_us = last_done - sched.front().close_us;

    // ---- verification: a sample of windows joined alone with a plain launch
    int verified = 0, mismatches = 0;
    if (P.verify > 0 && !sched.empty()) {
        aj_batch_opts_t bo;
        memset(&bo, 0, sizeof(bo));
        bo.no_graph = 1;
        aj_batch_plan_t *plan = nullptr;
        void *dR = nullptr, *dS = nullptr, *dres = nullptr, *dbounds = nullptr;
        const uint32_t bounds[2] = {0u, 1u};  // every generated tuple has ts 0
        if (aj_batch_plan_create(ctx, P.cap_tuples, P.cap_tuples, 1, P.cap_tuples, &bo, &plan) == AJ_OK &&
            aj_device_alloc(&dR, 12 * P.cap_tuples) == AJ_OK &&
            aj_device_alloc(&dS, 12 * P.cap_tuples) == AJ_OK && aj_device_alloc(&dres, 16) == AJ_OK &&
            aj_device_alloc(&dbounds, 8) == AJ_OK &&
            aj_copy_h2d(dbounds, bounds, sizeof(bounds)) == AJ_OK) {
            const size_t step = std::max<size_t>(1, sched.size() / P.verify);
            for (size_t i = 0; i < sched.size(); i += step) {
                const Req &r = sched[i];
                uint64_t res[2] = {0, 0};
                if (aj_gen_stream_tuples(ctx, (aj_stream_tuple_t *)dR, r.nR, r.seed, P.domain, 0.0, 0, nullptr) != AJ_OK ||
                    aj_gen_stream_tuples(ctx, (aj_stream_tuple_t *)dS, r.nS, r.seed ^ 0x5DEECE66Dull, P.domain, r.theta, 0, nullptr) != AJ_OK ||
                    aj_batch_run(plan, (aj_stream_tuple_t *)dR, r.nR, (aj_stream_tuple_t *)dS, r.nS,
                                 (const uint32_t *)dbounds, 1, (uint64_t *)dres, (uint64_t *)dres + 1, nullptr) != AJ_OK ||
                    aj_device_sync() != AJ_OK || aj_copy_d2h(res, dres, 16) != AJ_OK) {
                    fprintf(stderr, "[ERROR] verify: %s\n", aj_last_error());
                    ++mismatches;
                    break;
                }
                ++verified;
                if (res[0] != done[i].count || res[1] != done[i].checksum) ++mismatches;
            }
        } else {
            fprintf(stderr, "[ERROR] verify setup: %s\n", aj_last_error());
            mismatches = -1;
        }
        if (plan) aj_batch_plan_destroy(plan);
        if (dR) aj_device_free(dR);
        if (dS) aj_device_free(dS);
        if (dres) aj_device_free(dres);
        if (dbounds) aj_device_free(dbounds);
    }
    printf("{\"load\": %.3f, \"streams\": %d, \"windows\": %zu, \"offered_Mtps\": %.1f, "
           "\"achieved_Mtps\": %.1f, \"tuples\": %llu, \"span_ms\": %.2f, "
           "\"p50_ms\": %.3f, \"p99_ms\": %.3f, \"p999_ms\": %.3f, \"max_ms\": %.3f, "
           "\"batches\": %llu, \"avg_batch_windows\": %.2f, \"alloc_waits\": %llu, "
           "\"window_GB\": %.1f, \"window_GB_peak\": %.2f, \"workspace_GB\": %.2f, "
           "\"bad_status\": %llu, \"verified\": %d, \"mismatches\": %d}\n",
           P.load, P.streams, sched.size(), offered / (P.seconds * 1e6) / 1.0,
           offered / span_us, (unsigned long long)offered, span_us / 1e3, pct(lat, 0.5) / 1e3,
           pct(lat, 0.99) / 1e3, pct(lat, 0.999) / 1e3, pct(lat, 1.0) / 1e3,
           (unsigned long long)st.batches,
           sched.empty() ? 0.0 : (double)batched / sched.size(), (unsigned long long)st.alloc_waits,
           st.window_bytes / 1e9, st.window_bytes_peak / 1e9, st.workspace_bytes / 1e9,
           (unsigned long long)bad_status, verified, mismatches);
    fflush(stdout);
    aj_server_destroy(srv);
    aj_ctx_destroy(ctx);
    return (mismatches || bad_status) ? 2 : 0;
}
