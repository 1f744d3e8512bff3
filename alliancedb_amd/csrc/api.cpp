// api.cpp -- the C ABI: errors, options, contexts, drop-in PRO/NPO, device API.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "aj_internal.h"

namespace aj {

int pro_device(Ctx *ctx, const uint64_t *dR, uint64_t nR, const uint64_t *dS, uint64_t nS,
               const ProConfig &cfg, uint64_t *pairs, uint64_t pairs_cap, int records,
               aj_join_out_t *out, hipStream_t st);
int npo_device(Ctx *ctx, const uint64_t *dR, uint64_t nR, const uint64_t *dS, uint64_t nS,
               uint64_t *pairs, uint64_t pairs_cap, int records, aj_join_out_t *out,
               hipStream_t st);

// ---------------------------------------------------------------- errors
static thread_local char g_err[512] = "";
static thread_local int g_status = AJ_OK;

int set_error(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    g_status = code;
    return code;
}
int last_status() { return g_status; }
static void clear_error() {
    g_err[0] = 0;
    g_status = AJ_OK;
}

// ---------------------------------------------------------------- memory
static std::atomic<uint64_t> g_dev_budget{0};  // 0 = unlimited
static std::atomic<uint64_t> g_dev_used{0};

int dev_malloc(void **p, size_t bytes) {
    *p = nullptr;
    const uint64_t budget = g_dev_budget.load();
    if (budget && g_dev_used.load() + bytes > budget) {
        set_error(AJ_ERR_NOMEM, "device budget exceeded: %zu bytes requested, %llu of %llu used",
                  bytes, (unsigned long long)g_dev_used.load(), (unsigned long long)budget);
        return AJ_ERR_NOMEM;
    }
    hipError_t e = hipMalloc(p, bytes ? bytes : 1);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        set_error(AJ_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
        return AJ_ERR_NOMEM;
    }
    g_dev_used += bytes;
    return AJ_OK;
}

void dev_free(void *p) {
    if (p) (void)hipFree(p);
}

int DevBuf::ensure(size_t bytes) {
    if (bytes <= cap && p) return AJ_OK;
    if (arena) {
        const size_t o = (arena->used + 255) / 256 * 256;
        if (o > arena->size || bytes > arena->size - o) {
            set_error(AJ_ERR_CAPACITY,
                      "the no-allocation scratch of %zu bytes does not hold this join's "
                      "workspace (%zu more bytes needed at offset %zu)",
                      arena->size, bytes, o);
            return AJ_ERR_CAPACITY;
        }
        p = arena->base + o;
        cap = bytes;
        arena->used = o + bytes;
        return AJ_OK;
    }
    if (p) {
        dev_free(p);
        g_dev_used -= cap;
        p = nullptr;
        cap = 0;
    }
    // grow with headroom so that slowly growing sizes do not reallocate each call
    size_t want = bytes + bytes / 8;
    int rc = dev_malloc(&p, want);
    if (rc != AJ_OK) {
        rc = dev_malloc(&p, bytes);
        if (rc != AJ_OK) return rc;
        want = bytes;
    }
    cap = want;
    return AJ_OK;
}

void DevBuf::release() {
    if (arena) {
        p = nullptr;
        cap = 0;
        return;
    }
    if (p) {
        dev_free(p);
        g_dev_used -= cap;
    }
    p = nullptr;
    cap = 0;
}

// ---------------------------------------------------------------- options
static std::atomic<int64_t> opt_materialize{0}, opt_radix_bits{0}, opt_passes{0}, opt_device{0},
    opt_cap{0}, opt_split{0}, opt_kernel_timing{0}, opt_part_block{0}, opt_part_chunk{0},
    opt_part_hist2d{1}, opt_tuned{1}, opt_timing{0}, opt_shj_batch{0}, opt_staged{1},
    opt_stage_sets{0}, opt_stage_chunk{0}, opt_stage_block{0}, opt_stage_cache{-1},
    opt_npo_region_min{1ll << 22}, opt_heavy{1};

static thread_local bool g_force_timing = false;  // phase timing for one drop-in call
bool kernel_timing_enabled() { return g_force_timing || opt_kernel_timing.load() != 0; }
bool staged_enabled() { return opt_staged.load() != 0; }
uint64_t npo_region_min() { return (uint64_t)opt_npo_region_min.load(); }

static int ctx_init(Ctx *c, int device) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        set_error(AJ_ERR_NODEV, "no HIP device available");
        return AJ_ERR_NODEV;
    }
    if (device < 0 || device >= n) {
        set_error(AJ_ERR_ARG, "device %d out of range (%d devices)", device, n);
        return AJ_ERR_ARG;
    }
    c->device = device;
    AJ_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    AJ_HIP(hipGetDeviceProperties(&prop, device));
    c->num_cu = prop.multiProcessorCount;
    AJ_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    AJ_HIP(hipEventCreate(&c->ev0));
    AJ_HIP(hipEventCreate(&c->ev1));
    AJ_HIP(hipHostMalloc((void **)&c->hres, HRES_SLOTS * sizeof(uint64_t), hipHostMallocDefault));
    return AJ_OK;
}

static void ctx_fini(Ctx *c) {
    (void)hipSetDevice(c->device);
    DevBuf *bufs[] = {&c->inR, &c->inS, &c->p1R, &c->p1S, &c->p2R, &c->p2S, &c->hist,
                      &c->scan_tmp, &c->plan, &c->seg, &c->tasks, &c->acc, &c->pairs,
                      &c->table, &c->misc, &c->hist2d, &c->steps, &c->stage,
                      &c->chain, &c->chain2, &c->heads, &c->heads2, &c->carry, &c->carry2,
                      &c->heavy, &c->light};
    for (DevBuf *b : bufs) b->release();
    if (c->pro1) {
        batch_plan_release(c->pro1);
        delete c->pro1;
        c->pro1 = nullptr;
    }
    for (hipEvent_t e : c->evs) (void)hipEventDestroy(e);
    c->evs.clear();
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    for (hipEvent_t &e : c->evk)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->hres) (void)hipHostFree(c->hres);
    c->hres = nullptr;
}

static std::mutex g_default_mu;
static Ctx *g_default = nullptr;

Ctx *default_ctx(int *status) {
    std::lock_guard<std::mutex> lk(g_default_mu);
    if (!g_default) {
        Ctx *c = new Ctx();
        int rc = ctx_init(c, (int)opt_device.load());
        if (rc != AJ_OK) {
            delete c;
            *status = rc;
            return nullptr;
        }
        g_default = c;
    }
    *status = AJ_OK;
    return g_default;
}

static ProConfig config_from(uint64_t nR, uint64_t nS, const aj_join_opts_t *o) {
    aj_join_opts_t eff;
    memset(&eff, 0, sizeof(eff));
    if (o) eff = *o;
    if (!eff.radix_bits) eff.radix_bits = (int)opt_radix_bits.load();
    if (!eff.passes) eff.passes = (int)opt_passes.load();
    if (!eff.cap) eff.cap = (int)opt_cap.load();
    if (!eff.split) eff.split = (uint32_t)opt_split.load();
    aj_tune_entry_t te;
    const bool use_table = opt_tuned.load() && !eff.radix_bits && !eff.passes && !eff.cap &&
                           tune_lookup(nR, nS, &te);
    if (use_table) {
        eff.radix_bits = te.radix_bits;
        eff.passes = te.passes;
        eff.cap = te.cap;
    }
    ProConfig c = choose_pro_config(nR, nS, &eff);
    if (use_table) {
        c.block = te.part_block;
        if (te.part_chunk) c.chunk = (int)te.part_chunk;
    }
    if (opt_part_block.load()) c.block = (int)opt_part_block.load();
    if (opt_part_chunk.load()) c.chunk = (int)opt_part_chunk.load();
    c.p2d = opt_part_hist2d.load() != 0;
    c.staged = opt_staged.load() != 0;
    // a tuned one-pass geometry that was fastest through the histogram pass
    if (use_table && te.one_pass_hist) c.staged = false;
    c.stage_sets = (int)opt_stage_sets.load();
    c.stage_chunk = (uint64_t)opt_stage_chunk.load();
    c.stage_block = (int)opt_stage_block.load();
    c.stage_cache = (int)opt_stage_cache.load();
    return c;
}

// ---------------------------------------------------------------- heavy keys
// A counting PRO whose S has keys that hold a large share of it (Zipf S at
// theta >= 1): a strided sample of S names the keys seen at least HEAVY_MIN
// times in HEAVY_SAMPLE draws (the hottest HEAVY_MAX); when they hold at least
// HEAVY_SHARE of the sample, those keys are joined by their aggregates
// (heavy.hip) and the partitioned join takes R and the rest of S.  The result
// is the same sum either way.  A first sample of HEAVY_SCREEN draws ends the
// check unless the keys drawn HEAVY_SCREEN_MIN times in it hold at least
// HEAVY_SCREEN_SHARE of it (below that, the partitioned join of what would be
// left costs about as much as the whole: measured at Zipf theta 1.0, where the
// keys the full sample finds hold ~1/3 of S).  Returns with
// *applied false (nothing joined) when the samples find too little.  Only
// probe sides of at least 2^22 tuples and 4 times the build side are sampled
// (the skewed foreign-key shape), so other joins pay nothing for the check.
constexpr uint32_t HEAVY_SAMPLE = 1u << 14, HEAVY_MIN = 8;
constexpr uint32_t HEAVY_SCREEN = 1u << 11, HEAVY_SCREEN_MIN = 4;
constexpr double HEAVY_SCREEN_SHARE = 0.4, HEAVY_SHARE = 0.15;

static bool heavy_applies(const Ctx *ctx, uint64_t nR, uint64_t nS, const uint64_t *pairs,
                          int records) {
    return opt_heavy.load() && !pairs && !records && !ctx->arena_only && nR > 0 &&
           nS >= (1ull << 22) && nS >= 4 * nR;
}

// Keys drawn at least `min` times in the sample, as (draws, key), hottest
// first, at most HEAVY_MAX (counted in an open-addressed table of 2m entries).
static std::vector<std::pair<uint32_t, uint32_t>> heavy_select(const std::vector<uint32_t> &sample,
                                                               uint32_t min) {
    const size_t cap = 2 * sample.size();  // a power of two, as the sample size
    std::vector<uint32_t> tkey(cap), tcnt(cap, 0);
    for (uint32_t k : sample) {
        size_t h = (size_t)(((uint64_t)k * 0x9E3779B97F4A7C15ull) >> 32) & (cap - 1);
        while (tcnt[h] && tkey[h] != k) h = (h + 1) & (cap - 1);
        tkey[h] = k;
        ++tcnt[h];
    }
    std::vector<std::pair<uint32_t, uint32_t>> hot;
    for (size_t i = 0; i < cap; ++i)
        if (tcnt[i] >= min) hot.push_back({tcnt[i], tkey[i]});
    std::sort(hot.begin(), hot.end(), [](const std::pair<uint32_t, uint32_t> &a,
                                         const std::pair<uint32_t, uint32_t> &b) {
        return a.first != b.first ? a.first > b.first : a.second < b.second;
    });
    if (hot.size() > (size_t)HEAVY_MAX) hot.resize(HEAVY_MAX);
    return hot;
}

static int pro_heavy(Ctx *ctx, const uint64_t *dR, uint64_t nR, const uint64_t *dS, uint64_t nS,
                     const ProConfig &cfg, aj_join_out_t *out, hipStream_t st, bool *applied) {
    *applied = false;
    const uint32_t m = HEAVY_SAMPLE;
    const size_t o_keys = 0, o_slots = 4 * (size_t)m, o_agg = o_slots + 32 * HEAVY_MAX,
                 o_n = o_agg + 32 * HEAVY_MAX, bytes = o_n + 64;
    AJ_TRY(ctx->heavy.ensure(bytes));
    uint8_t *hb = ctx->heavy.as<uint8_t>();
    uint32_t *d_sample = reinterpret_cast<uint32_t *>(hb + o_keys);
    std::vector<uint32_t> sample(HEAVY_SCREEN);
    AJ_TRY(heavy_sample(dS, nS, HEAVY_SCREEN, d_sample, st));
    AJ_HIP(hipMemcpyAsync(sample.data(), d_sample, 4 * (size_t)HEAVY_SCREEN,
                          hipMemcpyDeviceToHost, st));
    AJ_HIP(hipStreamSynchronize(st));
    uint64_t screened = 0;
    for (const auto &h : heavy_select(sample, HEAVY_SCREEN_MIN)) screened += h.first;
    if ((double)screened < HEAVY_SCREEN_SHARE * HEAVY_SCREEN) return AJ_OK;
    sample.resize(m);
    AJ_TRY(heavy_sample(dS, nS, m, d_sample, st));
    AJ_HIP(hipMemcpyAsync(sample.data(), d_sample, 4 * (size_t)m, hipMemcpyDeviceToHost, st));
    AJ_HIP(hipStreamSynchronize(st));
    const std::vector<std::pair<uint32_t, uint32_t>> hot = heavy_select(sample, HEAVY_MIN);
    uint64_t drawn = 0;
    for (const auto &h : hot) drawn += h.first;
    if (hot.empty() || (double)drawn < HEAVY_SHARE * m) return AJ_OK;
    const uint32_t nh = (uint32_t)hot.size();
    std::vector<uint32_t> keys(nh);
    std::vector<uint64_t> slots(4 * HEAVY_MAX);
    for (uint32_t i = 0; i < nh; ++i) keys[i] = hot[i].second;
    heavy_table(keys.data(), nh, slots.data());
    AJ_TRY(ctx->light.ensure(8 * nS + 256));
    HeavyDesc d;
    d.R = dR;
    d.S = dS;
    d.nR = nR;
    d.nS = nS;
    d.slots = reinterpret_cast<uint64_t *>(hb + o_slots);
    d.nheavy = nh;
    d.agg = reinterpret_cast<unsigned long long *>(hb + o_agg);
    d.light = ctx->light.as<uint64_t>();
    d.light_n = reinterpret_cast<unsigned long long *>(hb + o_n);
    d.num_cu = ctx->num_cu;
    AJ_HIP(hipMemcpyAsync((void *)d.slots, slots.data(), 8 * slots.size(), hipMemcpyHostToDevice,
                          st));
    AJ_TRY(fill_u64(reinterpret_cast<uint64_t *>(d.agg), 4 * HEAVY_MAX + 1, 0, st));
    AJ_HIP(record_event(ctx->ev0, st));
    AJ_TRY(heavy_split(d, st));
    AJ_HIP(record_event(ctx->ev1, st));
    std::vector<uint64_t> agg(4 * HEAVY_MAX + 1);
    AJ_HIP(hipMemcpyAsync(agg.data(), d.agg, 8 * agg.size(), hipMemcpyDeviceToHost, st));
    AJ_HIP(hipStreamSynchronize(st));
    const float split_ms = elapsed_ms(ctx->ev0, ctx->ev1);
    const uint64_t n_light = agg[4 * HEAVY_MAX];
    uint64_t count = 0, sum = 0;
    for (uint32_t i = 0; i < nh; ++i) {
        const uint64_t cR = agg[i], sR = agg[HEAVY_MAX + i], cS = agg[2 * HEAVY_MAX + i],
                       sS = agg[3 * HEAVY_MAX + i];
        count += cR * cS;
        sum += ((sR * cS) << 32) + cR * sS;
    }
    memset(out, 0, sizeof(*out));
    if (n_light) AJ_TRY(pro_device(ctx, dR, nR, d.light, n_light, cfg, nullptr, 0, 0, out, st));
    out->count += count;
    out->checksum += sum;
    out->device_ms += split_ms;
    out->heavy_ms = split_ms;
    out->light = n_light;
    out->flags |= AJ_OUT_HEAVY;
    *applied = true;
    return AJ_OK;
}

// Tuples per side per eager step: the caller's, else 256 steps of >= 65536.
static uint64_t shj_batch(uint64_t nR, uint64_t nS, const aj_join_opts_t *opts) {
    if (opts && opts->batch) return opts->batch;
    const uint64_t n = std::max(nR, nS);
    return std::max<uint64_t>(65536, (n + 255) / 256);
}

static int join_dev_locked(Ctx *ctx, const uint64_t *dR, uint64_t nR, const uint64_t *dS,
                           uint64_t nS, const aj_join_opts_t *opts, aj_join_out_t *out,
                           hipStream_t st) {
    memset(out, 0, sizeof(*out));
    if (nR == 0 || nS == 0) {
        // every step of an eager join over an empty side matches nothing
        if (opts && opts->algo == AJ_ALGO_SHJ && opts->progress) {
            const uint64_t n = shj_steps(nR, nS, shj_batch(nR, nS, opts));
            std::vector<uint64_t> z(n, 0);
            if (n) AJ_HIP(hipMemcpy(opts->progress, z.data(), 8 * n, hipMemcpyDefault));
        }
        return AJ_OK;
    }
    const int algo = opts ? opts->algo : AJ_ALGO_PRO;
    uint64_t *pairs = opts ? opts->pairs : nullptr;
    const uint64_t pcap = opts ? opts->pairs_capacity : 0;
    const int records = opts ? opts->records : 0;
    if (records < 0 || records > 2) {
        set_error(AJ_ERR_ARG, "records must be 0, 1 or 2");
        return AJ_ERR_ARG;
    }
    int rc;
    if (algo == AJ_ALGO_SHJ) {
        rc = shj_device(ctx, dR, nR, dS, nS, shj_batch(nR, nS, opts), opts ? opts->progress : nullptr,
                        pairs, pcap, records, out, st);
    } else if (algo == AJ_ALGO_NPO) {
        rc = npo_device(ctx, dR, nR, dS, nS, pairs, pcap, records, out, st);
    } else if (algo == AJ_ALGO_PRO) {
        const ProConfig cfg = config_from(nR, nS, opts);
        bool done = false;
        rc = AJ_OK;
        if (heavy_applies(ctx, nR, nS, pairs, records))
            rc = pro_heavy(ctx, dR, nR, dS, nS, cfg, out, st, &done);
        if (rc == AJ_OK && !done)
            rc = pro_device(ctx, dR, nR, dS, nS, cfg, pairs, pcap, records, out, st);
    } else {
        set_error(AJ_ERR_ARG, "unknown algorithm %d", algo);
        return AJ_ERR_ARG;
    }
    if (rc != AJ_OK) return rc;
    if (pairs && out->count > pcap) {
        set_error(AJ_ERR_CAPACITY, "pair buffer holds %llu pairs, %llu needed",
                  (unsigned long long)pcap, (unsigned long long)out->count);
        return AJ_ERR_CAPACITY;
    }
    return AJ_OK;
}

static int check_relation(const relation_t *r, const char *name) {
    if (!r) {
        set_error(AJ_ERR_ARG, "relation %s is NULL", name);
        return AJ_ERR_ARG;
    }
    if (r->num_tuples && !r->tuples) {
        set_error(AJ_ERR_ARG, "relation %s has %llu tuples but no tuple array", name,
                  (unsigned long long)r->num_tuples);
        return AJ_ERR_ARG;
    }
    if (r->num_tuples >= (1ull << 31)) {
        set_error(AJ_ERR_ARG, "relation %s too large (%llu tuples, limit 2^31-1)", name,
                  (unsigned long long)r->num_tuples);
        return AJ_ERR_ARG;
    }
    return AJ_OK;
}

// ------------------------------------------- reference-format result lists
// The reference materialises, per joining thread, one record per match: the
// matching S tuple <S.key, S.payloadID> (parallel_radix_join.cpp:333-338,
// common_functions.cpp:252-257), appended to a chainedtuplebuffer_t whose
// newest 1M-tuple buffer is at the head of the chain (tuple_buffer.h:31-47,
// cb_next_writepos :93-109).  materialize = 2 returns exactly that structure:
// entry 0 of result_t.resultlist holds every record, the other entries are
// empty chains (one buffer, writepos 0) as chainedtuplebuffer_init leaves them.

static void chain_free(aj_chainedtuplebuffer_t *cb) {
    if (!cb) return;
    aj_tuplebuffer_t *b = cb->buf;
    while (b) {
        aj_tuplebuffer_t *nx = b->next;
        free(b->tuples);
        free(b);
        b = nx;
    }
    free(cb);
}

// A chain holding the n records of dsrc (device memory; may be NULL when n = 0),
// record i in buffer i / AJ_CHAINEDBUFF_NUMTUPLESPERBUF counted from the tail.
static aj_chainedtuplebuffer_t *chain_from_device(const uint64_t *dsrc, uint64_t n, int *rc) {
    const uint64_t per = AJ_CHAINEDBUFF_NUMTUPLESPERBUF;
    const uint64_t nbuf = n ? (n + per - 1) / per : 1;
    auto *cb = (aj_chainedtuplebuffer_t *)calloc(1, sizeof(aj_chainedtuplebuffer_t));
    if (!cb) {
        *rc = set_error(AJ_ERR_NOMEM, "host allocation of a result chain failed");
        return nullptr;
    }
    for (uint64_t k = 0; k < nbuf; ++k) {  // oldest first; each new buffer becomes the head
        auto *b = (aj_tuplebuffer_t *)calloc(1, sizeof(aj_tuplebuffer_t));
        void *t = nullptr;
        if (!b || posix_memalign(&t, 64, per * sizeof(tuple_t))) {
            free(b);
            chain_free(cb);
            *rc = set_error(AJ_ERR_NOMEM, "host allocation of a result buffer failed");
            return nullptr;
        }
        b->tuples = (tuple_t *)t;
        b->next = cb->buf;
        cb->buf = b;
        const uint64_t len = std::min(per, n - std::min(n, k * per));
        if (len && hipMemcpy(b->tuples, dsrc + k * per, len * 8, hipMemcpyDeviceToHost)) {
            chain_free(cb);
            *rc = set_error(AJ_ERR_HIP, "device-to-host copy of result records failed");
            return nullptr;
        }
    }
    cb->readcursor = cb->writecursor = cb->buf;
    cb->writepos = (uint32_t)(n - (nbuf - 1) * per);
    cb->numbufs = (uint32_t)nbuf;
    return cb;
}

static void free_resultlist(threadresult_t *l, int n) {
    if (!l) return;
    for (int i = 0; i < n; ++i) chain_free((aj_chainedtuplebuffer_t *)l[i].results);
    free(l);
}

// resultlist of nt entries, entry 0 holding the n device records of dsrc.
static threadresult_t *build_resultlist(const uint64_t *dsrc, uint64_t n, int nt, int *rc) {
    auto *l = (threadresult_t *)calloc((size_t)nt, sizeof(threadresult_t));
    if (!l) {
        *rc = set_error(AJ_ERR_NOMEM, "host allocation of the result list failed");
        return nullptr;
    }
    for (int i = 0; i < nt; ++i) {
        const uint64_t m = i == 0 ? n : 0;
        l[i].threadid = (uint32_t)i;
        l[i].nresults = (int64_t)m;
        l[i].results = chain_from_device(dsrc, m, rc);
        if (!l[i].results) {
            free_resultlist(l, i);
            return nullptr;
        }
    }
    return l;
}

// Timed records: one per match, (device clock ticks since the join started << 32)
// | S.payloadID, from a join run with records = 2 after a clock stamp.
static int timed_run(Ctx *ctx, aj_join_opts_t o, uint64_t nR, uint64_t nS, uint64_t count,
                     std::vector<uint64_t> *timed, double *clock_hz, uint64_t *ntimed_r) {
    AJ_TRY(ctx->pairs.ensure((count + 1) * 8));
    uint64_t *stamp = ctx->pairs.as<uint64_t>() + count;
    AJ_TRY(device_clock_stamp(stamp, ctx->stream));
    o.pairs = ctx->pairs.as<uint64_t>();
    o.pairs_capacity = count;
    o.records = 2;
    aj_join_out_t out;
    AJ_TRY(join_dev_locked(ctx, ctx->inR.as<uint64_t>(), nR, ctx->inS.as<uint64_t>(), nS, &o,
                           &out, ctx->stream));
    if (out.count != count || out.npairs != count || out.nrecords_r > count)
        return set_error(AJ_ERR_STATE, "timed join disagrees with the counting join");
    *ntimed_r = out.nrecords_r;  // SHJ: the last ones are R tuples' probes
    timed->resize(count + 1);
    if (hipMemcpyAsync(timed->data(), ctx->pairs.p, (count + 1) * 8, hipMemcpyDeviceToHost,
                       ctx->stream) ||
        hipStreamSynchronize(ctx->stream))
        return set_error(AJ_ERR_HIP, "device-to-host copy of timed records failed");
    const uint32_t t0 = (uint32_t)(*timed)[count];
    timed->pop_back();
    for (uint64_t &v : *timed) v = ((uint64_t)((uint32_t)(v >> 32) - t0) << 32) | (uint32_t)v;
    int khz = 0;
    AJ_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device));
    *clock_hz = khz * 1e3;
    return AJ_OK;
}

static result_t *dropin(int algo, relation_t *relR, relation_t *relS, int nthreads) {
    clear_error();
    if (check_relation(relR, "R") || check_relation(relS, "S")) return nullptr;
    int st = AJ_OK;
    Ctx *ctx = default_ctx(&st);
    if (!ctx) return nullptr;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (hipSetDevice(ctx->device) != hipSuccess) {
        set_error(AJ_ERR_HIP, "hipSetDevice failed");
        return nullptr;
    }
    const uint64_t nR = relR->num_tuples, nS = relS->num_tuples;
    auto t0 = std::chrono::steady_clock::now();
    aj_join_out_t out;
    memset(&out, 0, sizeof(out));
    const int64_t mat = opt_materialize.load();
    const bool timing = opt_timing.load() != 0;
    const int nt = std::max(nthreads, 1);
    uint64_t npairs = 0;
    std::vector<uint64_t> host_pairs, timed;
    double clock_hz = 0;
    uint64_t ntimed_r = 0;
    threadresult_t *rlist = nullptr;
    int rc = AJ_OK;
    do {
        if (nR == 0 || nS == 0) break;
        if ((rc = ctx->inR.ensure(nR * 8))) break;
        if ((rc = ctx->inS.ensure(nS * 8))) break;
        if (hipMemcpyAsync(ctx->inR.p, relR->tuples, nR * 8, hipMemcpyHostToDevice, ctx->stream) ||
            hipMemcpyAsync(ctx->inS.p, relS->tuples, nS * 8, hipMemcpyHostToDevice, ctx->stream)) {
            set_error(AJ_ERR_HIP, "host-to-device copy failed");
            rc = AJ_ERR_HIP;
            break;
        }
        aj_join_opts_t o;
        memset(&o, 0, sizeof(o));
        o.algo = algo;
        if (algo == AJ_ALGO_SHJ) o.batch = (uint64_t)opt_shj_batch.load();  // 0 = default
        g_force_timing = timing;  // per-phase device times of the counting run
        rc = join_dev_locked(ctx, ctx->inR.as<uint64_t>(), nR, ctx->inS.as<uint64_t>(), nS, &o,
                             &out, ctx->stream);
        g_force_timing = false;
        if (rc || out.count == 0) break;
        if (mat) {
            if ((rc = ctx->pairs.ensure(out.count * 8))) break;
            aj_join_opts_t om = o;
            om.pairs = ctx->pairs.as<uint64_t>();
            om.pairs_capacity = out.count;
            om.records = mat == 2;
            aj_join_out_t out2;
            rc = join_dev_locked(ctx, ctx->inR.as<uint64_t>(), nR, ctx->inS.as<uint64_t>(), nS,
                                 &om, &out2, ctx->stream);
            if (rc) break;
            if (out2.count != out.count || out2.npairs != out.count) {
                rc = set_error(AJ_ERR_STATE, "materialising join disagrees with the counting join");
                break;
            }
            if (hipStreamSynchronize(ctx->stream)) {
                rc = set_error(AJ_ERR_HIP, "join stream failed");
                break;
            }
            if (mat == 2) {
                rlist = build_resultlist(ctx->pairs.as<uint64_t>(), out2.npairs, nt, &rc);
                if (!rlist) break;
            } else {
                npairs = out2.npairs;
                host_pairs.resize(npairs);
                if (hipMemcpy(host_pairs.data(), ctx->pairs.p, npairs * 8, hipMemcpyDeviceToHost)) {
                    rc = set_error(AJ_ERR_HIP, "device-to-host copy of pairs failed");
                    break;
                }
            }
        }
        if (timing) rc = timed_run(ctx, o, nR, nS, out.count, &timed, &clock_hz, &ntimed_r);
    } while (0);
    if (rc) {
        free_resultlist(rlist, nt);
        return nullptr;
    }
    if (mat == 2 && !rlist) {  // no matches (or an empty input): nt empty chains
        rlist = build_resultlist(nullptr, 0, nt, &rc);
        if (!rlist) return nullptr;
    }
    const uint64_t ntimed = timed.size();
    aj_result_t *res =
        (aj_result_t *)malloc(sizeof(aj_result_t) + (npairs + ntimed) * sizeof(uint64_t));
    if (!res) {
        free_resultlist(rlist, nt);
        set_error(AJ_ERR_NOMEM, "host allocation of the result failed");
        return nullptr;
    }
    memset(res, 0, sizeof(aj_result_t));
    res->base.totalresults = (int64_t)out.count;
    res->base.resultlist = rlist;
    res->base.nthreads = rlist ? nt : nthreads;
    res->checksum = out.checksum;
    res->npairs = npairs;
    res->pairs = npairs ? reinterpret_cast<uint64_t *>(res + 1) : nullptr;
    if (npairs) memcpy(res->pairs, host_pairs.data(), npairs * 8);
    res->status = AJ_OK;
    res->device_ms = out.device_ms;
    res->ntimed = ntimed;
    res->timed = ntimed ? reinterpret_cast<uint64_t *>(res + 1) + npairs : nullptr;
    if (ntimed) memcpy(res->timed, timed.data(), ntimed * 8);
    res->clock_hz = clock_hz;
    res->ntimed_r = ntimed_r;
    res->algo = algo;
    if (timing) memcpy(res->phase_ms, out.phase_ms, sizeof(res->phase_ms));
    res->total_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0)
                        .count();
    return &res->base;
}

}  // namespace aj

using namespace aj;

struct aj_ctx {
    Ctx c;
};

namespace aj {
Ctx *ctx_of(aj_ctx_t *c) { return &c->c; }
int join_dev(Ctx *ctx, const uint64_t *dR, uint64_t nR, const uint64_t *dS, uint64_t nS,
             const aj_join_opts_t *opts, aj_join_out_t *out, hipStream_t st) {
    return join_dev_locked(ctx, dR, nR, dS, nS, opts, out, st);
}
}  // namespace aj

extern "C" {

const char *aj_last_error(void) { return g_err; }
int aj_last_status(void) { return g_status; }

result_t *PRO(relation_t *relR, relation_t *relS, int nthreads) {
    return dropin(AJ_ALGO_PRO, relR, relS, nthreads);
}

result_t *NPO(relation_t *relR, relation_t *relS, int nthreads) {
    return dropin(AJ_ALGO_NPO, relR, relS, nthreads);
}

result_t *SHJ_JM_NP(relation_t *relR, relation_t *relS, int nthreads) {
    return dropin(AJ_ALGO_SHJ, relR, relS, nthreads);
}

result_t *SHJ_JBCR_NP(relation_t *relR, relation_t *relS, int nthreads) {
    return dropin(AJ_ALGO_SHJ, relR, relS, nthreads);
}

result_t *SHJ_st(relation_t *relR, relation_t *relS, int nthreads) {
    return dropin(AJ_ALGO_SHJ, relR, relS, nthreads);
}

const aj_result_t *aj_result_ext(const result_t *r) {
    return reinterpret_cast<const aj_result_t *>(r);
}

void aj_free_result(result_t *r) {
    if (!r) return;
    free_resultlist(r->resultlist, r->nthreads);
    free(r);
}

int aj_write_result_file(const result_t *r, const char *path) {
    clear_error();
    if (!r || !path) return set_error(AJ_ERR_ARG, "aj_write_result_file: NULL argument");
    if (r->totalresults > 0 && !r->resultlist)
        return set_error(AJ_ERR_STATE, "result holds no record lists (option materialize 2)");
    FILE *f = fopen(path, "w");
    if (!f) return set_error(AJ_ERR_ARG, "cannot write %s", path);
    std::vector<char> buf(1 << 20);
    size_t used = 0;
    bool bad = false;
    auto put_int = [&](int32_t v, char end) {
        char tmp[12];
        int k = 0;
        uint32_t u = v < 0 ? 0u - (uint32_t)v : (uint32_t)v;
        do {
            tmp[k++] = (char)('0' + u % 10);
            u /= 10;
        } while (u);
        if (v < 0) buf[used++] = '-';
        while (k) buf[used++] = tmp[--k];
        buf[used++] = end;
    };
    const int nt = r->resultlist ? r->nthreads : 0;
    for (int i = 0; i < nt && !bad; ++i) {
        const threadresult_t &tr = r->resultlist[i];
        const auto *cb = (const aj_chainedtuplebuffer_t *)tr.results;
        if (tr.nresults <= 0) continue;
        // newest record first, as cb_begin_backwards / cb_read_backwards read
        // (tuple_buffer.h:57-90), one "%d %d" line of key and payloadID each
        const aj_tuplebuffer_t *b = cb ? cb->buf : nullptr;
        uint64_t pos = cb && cb->writepos ? cb->writepos - 1 : 0;
        for (int64_t j = 0; j < tr.nresults; ++j) {
            if (!b || (cb->writepos == 0 && j == 0)) {
                bad = true;
                set_error(AJ_ERR_STATE, "result list %d holds fewer records than it counts", i);
                break;
            }
            const tuple_t t = b->tuples[pos];
            if (used + 32 > buf.size()) {
                bad |= fwrite(buf.data(), 1, used, f) != used;
                used = 0;
            }
            put_int(t.key, ' ');
            put_int(t.payloadID, '\n');
            if (pos == 0) {
                b = b->next;
                pos = AJ_CHAINEDBUFF_NUMTUPLESPERBUF - 1;
            } else {
                --pos;
            }
        }
    }
    if (!bad && used) bad = fwrite(buf.data(), 1, used, f) != used;
    bad |= ferror(f) != 0;
    fclose(f);
    if (bad) return g_status != AJ_OK ? g_status : set_error(AJ_ERR_ARG, "error writing %s", path);
    return AJ_OK;
}

int aj_set_option(const char *name, int64_t value) {
    clear_error();
    if (!name) {
        set_error(AJ_ERR_ARG, "option name is NULL");
        return AJ_ERR_ARG;
    }
    struct {
        const char *n;
        std::atomic<int64_t> *v;
        int64_t lo, hi;
    } tab[] = {{"materialize", &opt_materialize, 0, 2}, {"radix_bits", &opt_radix_bits, 0, 24},
               {"passes", &opt_passes, 0, 2},           {"device", &opt_device, 0, 1024},
               {"cap", &opt_cap, 0, 12288},             {"split", &opt_split, 0, 1ll << 30},
               {"kernel_timing", &opt_kernel_timing, 0, 1},
               {"part_block", &opt_part_block, 0, 512},
               {"part_chunk", &opt_part_chunk, 0, 1ll << 24},
               {"part_hist2d", &opt_part_hist2d, 0, 1},
               {"tuned", &opt_tuned, 0, 1},
               {"timing", &opt_timing, 0, 1},
               {"shj_batch", &opt_shj_batch, 0, 1ll << 31},
               {"staged", &opt_staged, 0, 1},
               {"stage_sets", &opt_stage_sets, 0, 64},
               {"stage_chunk", &opt_stage_chunk, 0, 1ll << 31},
               {"stage_block", &opt_stage_block, 0, 1024},
               {"stage_cache", &opt_stage_cache, -1, 15},
               {"npo_region_min", &opt_npo_region_min, 0, 1ll << 32},
               {"heavy", &opt_heavy, 0, 1}};
    for (auto &e : tab) {
        if (strcmp(e.n, name) == 0) {
            if (value < e.lo || value > e.hi) {
                set_error(AJ_ERR_ARG, "option %s out of range [%lld, %lld]", name,
                          (long long)e.lo, (long long)e.hi);
                return AJ_ERR_ARG;
            }
            if (e.v == &opt_part_block && value && value != 256 && value != 512) {
                set_error(AJ_ERR_ARG, "option part_block must be 0, 256 or 512");
                return AJ_ERR_ARG;
            }
            if (e.v == &opt_stage_block && value && value != 512 && value != 1024) {
                set_error(AJ_ERR_ARG, "option stage_block must be 0, 512 or 1024");
                return AJ_ERR_ARG;
            }
            if (e.v == &opt_part_chunk && value % 8192) {
                set_error(AJ_ERR_ARG, "option part_chunk must be a multiple of 8192");
                return AJ_ERR_ARG;
            }
            e.v->store(value);
            return AJ_OK;
        }
    }
    if (strcmp(name, "device_budget") == 0) {
        if (value < 0) {
            set_error(AJ_ERR_ARG, "device_budget must be >= 0");
            return AJ_ERR_ARG;
        }
        g_dev_budget.store((uint64_t)value);
        return AJ_OK;
    }
    set_error(AJ_ERR_ARG, "unknown option '%s'", name);
    return AJ_ERR_ARG;
}

int64_t aj_get_option(const char *name) {
    if (!name) return -1;
    if (!strcmp(name, "materialize")) return opt_materialize.load();
    if (!strcmp(name, "radix_bits")) return opt_radix_bits.load();
    if (!strcmp(name, "passes")) return opt_passes.load();
    if (!strcmp(name, "device")) return opt_device.load();
    if (!strcmp(name, "cap")) return opt_cap.load();
    if (!strcmp(name, "split")) return opt_split.load();
    if (!strcmp(name, "kernel_timing")) return opt_kernel_timing.load();
    if (!strcmp(name, "part_block")) return opt_part_block.load();
    if (!strcmp(name, "part_chunk")) return opt_part_chunk.load();
    if (!strcmp(name, "part_hist2d")) return opt_part_hist2d.load();
    if (!strcmp(name, "tuned")) return opt_tuned.load();
    if (!strcmp(name, "timing")) return opt_timing.load();
    if (!strcmp(name, "shj_batch")) return opt_shj_batch.load();
    if (!strcmp(name, "staged")) return opt_staged.load();
    if (!strcmp(name, "stage_sets")) return opt_stage_sets.load();
    if (!strcmp(name, "stage_chunk")) return opt_stage_chunk.load();
    if (!strcmp(name, "stage_block")) return opt_stage_block.load();
    if (!strcmp(name, "stage_cache")) return opt_stage_cache.load();
    if (!strcmp(name, "npo_region_min")) return opt_npo_region_min.load();
    if (!strcmp(name, "heavy")) return opt_heavy.load();
    if (!strcmp(name, "device_budget")) return (int64_t)g_dev_budget.load();
    if (!strcmp(name, "device_used")) return (int64_t)g_dev_used.load();
    return -1;
}

static int ensure_default_device() {
    int st = AJ_OK;
    Ctx *c = default_ctx(&st);
    if (!c) return st;
    AJ_HIP(hipSetDevice(c->device));
    return AJ_OK;
}

int aj_device_alloc(void **ptr, uint64_t bytes) {
    clear_error();
    if (!ptr) {
        set_error(AJ_ERR_ARG, "ptr is NULL");
        return AJ_ERR_ARG;
    }
    *ptr = nullptr;
    AJ_TRY(ensure_default_device());
    if (hipMalloc(ptr, bytes ? bytes : 1) != hipSuccess) {
        (void)hipGetLastError();
        *ptr = nullptr;
        set_error(AJ_ERR_NOMEM, "hipMalloc(%llu) failed", (unsigned long long)bytes);
        return AJ_ERR_NOMEM;
    }
    return AJ_OK;
}

int aj_device_free(void *ptr) {
    clear_error();
    if (!ptr) return AJ_OK;
    AJ_TRY(ensure_default_device());
    AJ_HIP(hipFree(ptr));
    return AJ_OK;
}

int aj_copy_h2d(void *dst, const void *src, uint64_t bytes) {
    clear_error();
    if (!bytes) return AJ_OK;
    if (!dst || !src) {
        set_error(AJ_ERR_ARG, "NULL pointer");
        return AJ_ERR_ARG;
    }
    AJ_TRY(ensure_default_device());
    AJ_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return AJ_OK;
}

int aj_copy_d2h(void *dst, const void *src, uint64_t bytes) {
    clear_error();
    if (!bytes) return AJ_OK;
    if (!dst || !src) {
        set_error(AJ_ERR_ARG, "NULL pointer");
        return AJ_ERR_ARG;
    }
    AJ_TRY(ensure_default_device());
    AJ_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return AJ_OK;
}

int aj_copy_d2d(void *dst, const void *src, uint64_t bytes) {
    clear_error();
    if (!bytes) return AJ_OK;
    if (!dst || !src) {
        set_error(AJ_ERR_ARG, "NULL pointer");
        return AJ_ERR_ARG;
    }
    int st = AJ_OK;
    Ctx *c = default_ctx(&st);
    if (!c) return st;
    AJ_HIP(hipSetDevice(c->device));
    AJ_TRY(copy_d2d(dst, src, bytes, c->num_cu, nullptr));
    AJ_HIP(hipDeviceSynchronize());
    return AJ_OK;
}

int aj_access_kernel(int kind, void *table, uint64_t slots, uint64_t n, float *ms) {
    clear_error();
    if (!table) {
        set_error(AJ_ERR_ARG, "NULL table");
        return AJ_ERR_ARG;
    }
    int st = AJ_OK;
    Ctx *c = default_ctx(&st);
    if (!c) return st;
    AJ_HIP(hipSetDevice(c->device));
    uint64_t *tab = static_cast<uint64_t *>(table);
    AJ_TRY(fill_u64(tab, 2 * slots, ~0ull, nullptr));
    hipEvent_t a = nullptr, b = nullptr;
    AJ_HIP(hipEventCreate(&a));
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return set_error(AJ_ERR_HIP, "hipEventCreate failed");
    }
    int rc = hipEventRecord(a, nullptr) == hipSuccess ? AJ_OK : AJ_ERR_HIP;
    if (rc == AJ_OK) rc = access_kernel(kind, tab, slots, n, c->num_cu, nullptr);
    if (rc == AJ_OK && hipEventRecord(b, nullptr) != hipSuccess) rc = AJ_ERR_HIP;
    if (rc == AJ_OK && hipEventSynchronize(b) != hipSuccess) rc = AJ_ERR_HIP;
    float t = 0;
    if (rc == AJ_OK && hipEventElapsedTime(&t, a, b) != hipSuccess) rc = AJ_ERR_HIP;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc == AJ_ERR_HIP) return set_error(AJ_ERR_HIP, "aj_access_kernel: HIP event failure");
    if (rc != AJ_OK) return rc;
    if (ms) *ms = t;
    return AJ_OK;
}

int aj_device_sync(void) {
    clear_error();
    AJ_TRY(ensure_default_device());
    AJ_HIP(hipDeviceSynchronize());
    return AJ_OK;
}

int aj_ctx_create(int device, aj_ctx_t **out) {
    clear_error();
    if (!out) {
        set_error(AJ_ERR_ARG, "out is NULL");
        return AJ_ERR_ARG;
    }
    *out = nullptr;
    std::unique_ptr<aj_ctx> c(new aj_ctx());
    int rc = ctx_init(&c->c, device);
    if (rc != AJ_OK) return rc;
    *out = c.release();
    return AJ_OK;
}

int aj_ctx_destroy(aj_ctx_t *ctx) {
    clear_error();
    if (!ctx) return AJ_OK;
    {
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        ctx_fini(&ctx->c);
    }
    delete ctx;
    return AJ_OK;
}

int aj_join_dev(aj_ctx_t *ctx, const tuple_t *dR, uint64_t nR, const tuple_t *dS, uint64_t nS,
                const aj_join_opts_t *opts, aj_join_out_t *out, void *stream) {
    clear_error();
    if (!ctx || !out || (nR && !dR) || (nS && !dS)) {
        set_error(AJ_ERR_ARG, "aj_join_dev: NULL argument");
        return AJ_ERR_ARG;
    }
    if (nR >= (1ull << 31) || nS >= (1ull << 31)) {
        set_error(AJ_ERR_ARG, "aj_join_dev: relation too large (limit 2^31-1 tuples)");
        return AJ_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    AJ_HIP(hipSetDevice(ctx->c.device));
    hipStream_t st = stream ? (hipStream_t)stream : ctx->c.stream;
    return join_dev_locked(&ctx->c, reinterpret_cast<const uint64_t *>(dR), nR,
                           reinterpret_cast<const uint64_t *>(dS), nS, opts, out, st);
}

}  // extern "C"
