// tune.cpp -- autotuner of the PRO geometry (radix bits, passes, LDS round,
// partition work-group size) per size bucket, with a persisted table, and of
// a window batch's geometry (LDS round, partition target, work-group chunk).
//
// No counterpart in the reference, whose geometry is fixed at compile time
// (NUM_RADIX_BITS, NUM_PASSES: prj_params.h:16-21).  aj_tune_pro times a small
// grid of candidates around the size heuristic on the caller's device
// relations (every candidate must return the same count and checksum) and
// keeps the fastest for the bucket (floor log2 |R|, floor log2 |S|).  PRO calls
// whose options leave the geometry open use the bucket's entry.  The table is
// saved as JSON, one entry per line.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "aj_internal.h"

namespace aj {

namespace {

std::mutex g_tune_mu;
std::map<std::pair<int, int>, aj_tune_entry_t> g_table;

int log2_floor(uint64_t n) {
    int l = 0;
    while (n > 1) {
        n >>= 1;
        ++l;
    }
    return l;
}

bool valid_entry(const aj_tune_entry_t &e) {
    const bool cap_ok = e.cap == 2048 || e.cap == 4096 || e.cap == 8192 || e.cap == 8704 ||
                        e.cap == 12288;
    return e.log2_r >= 0 && e.log2_r < 32 && e.log2_s >= 0 && e.log2_s < 32 && e.radix_bits >= 0 &&
           e.radix_bits <= 24 && e.passes >= 0 && e.passes <= 2 && cap_ok &&
           (e.part_block == 256 || e.part_block == 512) && e.part_chunk % 8192 == 0 &&
           (e.one_pass_hist == 0 || e.one_pass_hist == 1);
}

}  // namespace

bool tune_lookup(uint64_t nR, uint64_t nS, aj_tune_entry_t *out) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_table.find({log2_floor(nR), log2_floor(nS)});
    if (it == g_table.end()) return false;
    *out = it->second;
    return true;
}

}  // namespace aj

using namespace aj;

extern "C" {

int aj_tune_put(const aj_tune_entry_t *e) {
    if (!e || !valid_entry(*e)) {
        set_error(AJ_ERR_ARG, "aj_tune_put: invalid entry");
        return AJ_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_table[{e->log2_r, e->log2_s}] = *e;
    return AJ_OK;
}

int aj_tune_get(uint64_t nR, uint64_t nS, aj_tune_entry_t *out) {
    if (!out) {
        set_error(AJ_ERR_ARG, "aj_tune_get: NULL argument");
        return AJ_ERR_ARG;
    }
    if (!tune_lookup(nR, nS, out)) {
        set_error(AJ_ERR_STATE, "no tuned entry for this size bucket");
        return AJ_ERR_STATE;
    }
    return AJ_OK;
}

int aj_tune_clear(void) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_table.clear();
    return AJ_OK;
}

int aj_tune_save(const char *path) {
    if (!path) {
        set_error(AJ_ERR_ARG, "aj_tune_save: NULL path");
        return AJ_ERR_ARG;
    }
    std::vector<aj_tune_entry_t> es;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        for (const auto &kv : g_table) es.push_back(kv.second);
    }
    FILE *f = fopen(path, "w");
    if (!f) {
        set_error(AJ_ERR_ARG, "cannot write %s", path);
        return AJ_ERR_ARG;
    }
    fprintf(f, "{\"format\": \"alliancedb_amd PRO tuning\", \"version\": 1, \"entries\": [\n");
    for (size_t i = 0; i < es.size(); ++i) {
        const aj_tune_entry_t &e = es[i];
        fprintf(f,
                "{\"log2_r\": %d, \"log2_s\": %d, \"radix_bits\": %d, \"passes\": %d, \"cap\": %d, "
                "\"part_block\": %d, \"part_chunk\": %u, \"device_ms\": %.4f, "
                "\"one_pass_hist\": %d}%s\n",
                e.log2_r, e.log2_s, e.radix_bits, e.passes, e.cap, e.part_block, e.part_chunk,
                e.device_ms, e.one_pass_hist, i + 1 < es.size() ? "," : "");
    }
    fprintf(f, "]}\n");
    const int bad = ferror(f);
    fclose(f);
    if (bad) {
        set_error(AJ_ERR_ARG, "error writing %s", path);
        return AJ_ERR_ARG;
    }
    return AJ_OK;
}

int aj_tune_load(const char *path) {
    if (!path) {
        set_error(AJ_ERR_ARG, "aj_tune_load: NULL path");
        return AJ_ERR_ARG;
    }
    FILE *f = fopen(path, "r");
    if (!f) {
        set_error(AJ_ERR_ARG, "cannot read %s", path);
        return AJ_ERR_ARG;
    }
    char line[512];
    std::vector<aj_tune_entry_t> es;
    bool header = false;
    while (fgets(line, sizeof(line), f)) {
        if (strstr(line, "\"format\": \"alliancedb_amd PRO tuning\"")) header = true;
        if (!strstr(line, "\"log2_r\"")) continue;
        aj_tune_entry_t e;
        memset(&e, 0, sizeof(e));
        // tables written before one_pass_hist existed end after device_ms (0)
        const int got = sscanf(line,
                               "{\"log2_r\": %d, \"log2_s\": %d, \"radix_bits\": %d, \"passes\": %d, "
                               "\"cap\": %d, \"part_block\": %d, \"part_chunk\": %u, \"device_ms\": %f, "
                               "\"one_pass_hist\": %d}",
                               &e.log2_r, &e.log2_s, &e.radix_bits, &e.passes, &e.cap,
                               &e.part_block, &e.part_chunk, &e.device_ms, &e.one_pass_hist);
        if (got < 8 || !valid_entry(e)) {
            fclose(f);
            set_error(AJ_ERR_ARG, "%s: malformed entry: %s", path, line);
            return AJ_ERR_ARG;
        }
        es.push_back(e);
    }
    fclose(f);
    if (!header) {
        set_error(AJ_ERR_ARG, "%s is not a tuning table", path);
        return AJ_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (const auto &e : es) g_table[{e.log2_r, e.log2_s}] = e;
    return AJ_OK;
}

int aj_tune_pro(aj_ctx_t *ctx, const tuple_t *dR, uint64_t nR, const tuple_t *dS, uint64_t nS,
                int reps, aj_tune_entry_t *best) {
    if (!ctx || !dR || !dS || !best || nR == 0 || nS == 0 || reps < 1) {
        set_error(AJ_ERR_ARG, "aj_tune_pro: bad arguments");
        return AJ_ERR_ARG;
    }
    // candidates around the size heuristic (with the table ignored)
    aj_join_opts_t o0;
    memset(&o0, 0, sizeof(o0));
    o0.cap = 8704;
    const ProConfig h = choose_pro_config(nR, nS, &o0);
    const int b0 = h.bits1 + h.bits2;
    struct Cand {
        int bits, passes, cap, block, hist;  // hist: one pass through the histogram pass
    };
    std::vector<Cand> cands;
    const bool staged_on = aj_get_option("staged") != 0;
    for (int bits = std::max(1, b0 - 1); bits <= std::min(24, b0 + 1); ++bits)
        for (int passes = 1; passes <= 2; ++passes) {
            if (passes == 1 && bits > 12) continue;  // one pass: runs of >= 2 tuples per 8K tile
            if (passes == 2 && bits < 4) continue;
            for (int cap : {8192, 8704, 12288}) {
                cands.push_back({bits, passes, cap, 512, 0});
                // one pass of <= 11 bits runs the staged pass unless told otherwise
                if (passes == 1 && bits <= WSTAGE_MAX_BITS && staged_on)
                    cands.push_back({bits, passes, cap, 512, 1});
            }
        }
    if (b0 == 0) cands.push_back({0, 0, 8704, 512, 0});
    uint64_t want_c = 0, want_x = 0;
    bool have = false;
    float best_ms = 1e30f;
    Cand win{};
    auto run = [&](const Cand &c, float *ms) -> int {
        aj_join_opts_t o;
        memset(&o, 0, sizeof(o));
        o.algo = AJ_ALGO_PRO;
        o.radix_bits = c.bits;
        o.passes = c.passes;
        o.cap = c.cap;
        const int64_t saved_block = aj_get_option("part_block");
        const int64_t saved_staged = aj_get_option("staged");
        aj_set_option("part_block", c.block);
        if (c.hist) aj_set_option("staged", 0);
        aj_join_out_t out;
        int rc = aj_join_dev(ctx, dR, nR, dS, nS, &o, &out, nullptr);  // warm-up
        std::vector<float> t;
        for (int r = 0; rc == AJ_OK && r < reps; ++r) {
            rc = aj_join_dev(ctx, dR, nR, dS, nS, &o, &out, nullptr);
            t.push_back(out.device_ms);
        }
        aj_set_option("part_block", saved_block);
        aj_set_option("staged", saved_staged);
        if (rc != AJ_OK) return rc;
        if (!have) {
            want_c = out.count;
            want_x = out.checksum;
            have = true;
        } else if (out.count != want_c || out.checksum != want_x) {
            set_error(AJ_ERR_STATE, "tuning candidate (%d bits, %d passes, cap %d) disagrees",
                      c.bits, c.passes, c.cap);
            return AJ_ERR_STATE;
        }
        std::sort(t.begin(), t.end());
        *ms = t[t.size() / 2];
        return AJ_OK;
    };
    for (const Cand &c : cands) {
        float ms = 0;
        AJ_TRY(run(c, &ms));
        if (ms < best_ms) {
            best_ms = ms;
            win = c;
        }
    }
    {  // second stage: partition work-group size for the winner
        Cand c = win;
        c.block = 256;
        float ms = 0;
        AJ_TRY(run(c, &ms));
        if (ms < best_ms) {
            best_ms = ms;
            win = c;
        }
    }
    aj_tune_entry_t e;
    memset(&e, 0, sizeof(e));
    e.log2_r = log2_floor(nR);
    e.log2_s = log2_floor(nS);
    e.radix_bits = win.bits;
    e.passes = win.passes;
    e.cap = win.cap;
    e.part_block = win.block;
    e.part_chunk = 0;
    e.device_ms = best_ms;
    e.one_pass_hist = win.hist;
    AJ_TRY(aj_tune_put(&e));
    *best = e;
    return AJ_OK;
}

int aj_tune_batch(aj_ctx_t *ctx, const aj_stream_tuple_t *dR, uint64_t nR,
                  const aj_stream_tuple_t *dS, uint64_t nS, const uint32_t *d_bounds, int nwin,
                  uint64_t max_window_tuples, int reps, aj_batch_opts_t *best, float *best_ms) {
    if (!ctx || !dR || !dS || !d_bounds || nwin < 1 || reps < 1 || !best) {
        set_error(AJ_ERR_ARG, "aj_tune_batch: bad arguments");
        return AJ_ERR_ARG;
    }
    uint64_t *dres = nullptr;
    AJ_TRY(aj_device_alloc(reinterpret_cast<void **>(&dres), 16 * (uint64_t)nwin));
    std::vector<uint64_t> want, got(2 * (size_t)nwin);
    float win_ms = 1e30f;
    aj_batch_opts_t win{};
    int rc = AJ_OK;
    // candidates: LDS round x partition target (the round's power of two, or
    // half of it: more, smaller partitions) x tuples per partition work-group
    // (derived from the sizes, or fixed multiples of the 16384-tuple tile)
    for (int cap : {8192, 8704, 12288}) {
        uint32_t t = 1;
        while (2 * t <= (uint32_t)cap) t *= 2;
        for (uint32_t target : {t, t / 2})
            for (uint32_t chunk : {0u, 65536u, 262144u}) {
                aj_batch_opts_t o{};
                o.cap = cap;
                o.target = target;
                o.chunk = chunk;
                aj_batch_plan_t *plan = nullptr;
                rc = aj_batch_plan_create(ctx, nR, nS, nwin, max_window_tuples, &o, &plan);
                double ms = 0;
                for (int r = -1; rc == AJ_OK && r < reps; ++r) {  // r = -1: warm-up (and capture)
                    const auto t0 = std::chrono::steady_clock::now();
                    rc = aj_batch_run(plan, dR, nR, dS, nS, d_bounds, nwin, dres, dres + nwin,
                                      nullptr);
                    if (rc == AJ_OK) rc = aj_device_sync();
                    const auto t1 = std::chrono::steady_clock::now();
                    if (r >= 0) ms += std::chrono::duration<double, std::milli>(t1 - t0).count();
                }
                if (rc == AJ_OK) rc = aj_copy_d2h(got.data(), dres, 16 * (uint64_t)nwin);
                if (plan) aj_batch_plan_destroy(plan);
                if (rc != AJ_OK) break;
                if (want.empty()) {
                    want = got;
                } else if (got != want) {
                    rc = set_error(AJ_ERR_STATE,
                                   "batch tuning candidate (cap %d, target %u, chunk %u) disagrees",
                                   cap, target, chunk);
                    break;
                }
                if (ms / reps < win_ms) {
                    win_ms = (float)(ms / reps);
                    win = o;
                }
            }
        if (rc != AJ_OK) break;
    }
    aj_device_free(dres);
    if (rc != AJ_OK) return rc;
    *best = win;
    if (best_ms) *best_ms = win_ms;
    return AJ_OK;
}

}  // extern "C"
