// server.hip -- serving loop for window joins arriving from many streams, and
// the device generator of window tuples.
//
// No counterpart in the reference, which joins one window per process run
// (benchmark.cpp:164-257).  Structure:
//
//   window memory  one device buffer of 12-byte tuples, carved per window by
//                  a first-fit arena (arena.h); window_alloc waits on a
//                  condition variable while the budget is taken (admission)
//   ingest stream  fills/uploads of window tuples; every batch waits for the
//                  ingest work recorded before it (device-side event wait)
//   slots          `slots` batch workspaces, each with its own HIP stream,
//                  pinned range/result staging and a completion event
//   dispatcher     thread: when a slot is free and windows are queued, packs
//                  them FIFO into a batch (window, tuple limits), copies the
//                  ranges, launches the batched join (windows.hip,
//                  enqueue_ranges) and the result copy on the slot's stream
//   completer      thread: waits for the slots' events in launch order,
//                  publishes per-window results, returns window memory and
//                  the slot
//
// While one batch runs, the next is formed from windows that arrived
// meanwhile and launched on another stream (continuous batching).
#include <chrono>
#include <cmath>
#include <cstring>
#include <condition_variable>
#include <deque>
#include <memory>
#include <thread>
#include <unordered_map>

#include "aj_internal.h"
#include "arena.h"
#include "device_common.h"

namespace aj {

// ---------------------------------------------------------------- generator
namespace {

__host__ __device__ __forceinline__ uint64_t g_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__host__ __device__ __forceinline__ uint64_t g_stream_at(uint64_t seed, uint64_t i) {
    return g_mix64(seed + (i + 1) * 0x9E3779B97F4A7C15ull);
}

// Seeded permutation of [0, n) (4-round Feistel network on the smallest even
// bit width covering n, cycle-walking back into range).
struct Perm {
    uint64_t n, mask;
    int half;
    uint64_t keys[4];
};

Perm make_perm(uint64_t n, uint64_t seed) {
    Perm p;
    p.n = n;
    p.half = 1;
    while ((1ull << (2 * p.half)) < n) ++p.half;
    p.mask = (1ull << p.half) - 1;
    for (int r = 0; r < 4; ++r) p.keys[r] = g_stream_at(seed ^ 0xA5A5A5A5DEADBEEFull, r);
    return p;
}

__device__ __forceinline__ uint64_t perm_apply(const Perm &p, uint64_t i) {
    uint64_t x = i;
    do {
        uint64_t l = x >> p.half, r = x & p.mask;
        for (int k = 0; k < 4; ++k) {
            const uint64_t f = g_mix64(r ^ p.keys[k]) & p.mask;
            const uint64_t nl = r;
            r = l ^ f;
            l = nl;
        }
        x = (l << p.half) | r;
    } while (x >= p.n);
    return x;
}

// Rejection-inversion sampling of P(k) ~ k^-q on {1..N} (Hoermann &
// Derflinger, ACM TOMACS 6(3), 1996): constants from the host, draws on the
// device.
struct ZipfC {
    double q, hx1, hn, s;
    uint64_t N;
};

__host__ __device__ inline double z_helper1(double x) {
    return fabs(x) > 1e-8 ? log1p(x) / x : 1 - x * (0.5 - x / 3.0);
}
__host__ __device__ inline double z_helper2(double x) {
    return fabs(x) > 1e-8 ? expm1(x) / x : 1 + x * 0.5 * (1 + x / 3.0);
}
__host__ __device__ inline double z_h(double q, double x) { return exp(-q * log(x)); }
__host__ __device__ inline double z_H(double q, double x) {
    const double lx = log(x);
    return z_helper2((1 - q) * lx) * lx;
}
__host__ __device__ inline double z_Hinv(double q, double y) {
    double t = y * (1 - q);
    if (t < -1) t = -1;
    return exp(z_helper1(t) * y);
}

ZipfC make_zipf(uint64_t n, double q) {
    ZipfC z;
    z.q = q;
    z.N = n;
    z.hx1 = z_H(q, 1.5) - 1.0;
    z.hn = z_H(q, (double)n + 0.5);
    z.s = 2 - z_Hinv(q, z_H(q, 2.5) - z_h(q, 2));
    return z;
}

__device__ uint64_t zipf_rank(const ZipfC &z, uint64_t seed, uint64_t i) {
    uint64_t k = 1;
    for (uint64_t j = 0; j < 64; ++j) {
        const double u01 = (double)(g_stream_at(seed, i * 64 + j) >> 11) * (1.0 / 9007199254740992.0);
        const double u = z.hn + u01 * (z.hx1 - z.hn);
        const double x = z_Hinv(z.q, u);
        double kd = floor(x + 0.5);
        if (kd < 1) kd = 1;
        if (kd > (double)z.N) kd = (double)z.N;
        k = (uint64_t)kd;
        if (kd - x <= z.s || u >= z_H(z.q, kd + 0.5) - z_h(z.q, kd)) return k;
    }
    return k;
}

__global__ void k_gen_tuples(uint32_t *out, uint64_t n, uint64_t seed, uint32_t domain,
                             int zipf, ZipfC z, Perm p, uint32_t payload0) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t key;
        if (zipf) key = 1u + (uint32_t)perm_apply(p, zipf_rank(z, seed, i) - 1);
        else key = 1u + (uint32_t)__umul64hi(g_stream_at(seed, i), (uint64_t)domain);
        out[3 * i + 0] = key;
        out[3 * i + 1] = payload0 + (uint32_t)i;
        out[3 * i + 2] = 0u;
    }
}

// Many window fills in one launch: fill f writes n tuples at out as
// k_gen_tuples would.  The work is cut into chunks of FILL_CHUNK tuples;
// chunk0[f] is fill f's first chunk.  Fills of the same memory never share a
// launch (the server flushes first).
constexpr int FILL_MAX = 16;
constexpr uint64_t FILL_CHUNK = 4096;
struct FillDesc {
    uint32_t *out;
    uint64_t n, seed;
    uint32_t domain, payload0;
    int zipf;
    ZipfC z;
    Perm p;
};
struct FillBatch {
    FillDesc d[FILL_MAX];
    uint64_t chunk0[FILL_MAX + 1];
    int nd;
};

__global__ __launch_bounds__(256) void k_gen_fills(FillBatch b) {
    const uint64_t total = b.chunk0[b.nd];
    for (uint64_t c = blockIdx.x; c < total; c += gridDim.x) {
        int d = 0;
        while (d + 1 < b.nd && b.chunk0[d + 1] <= c) ++d;
        const FillDesc &f = b.d[d];
        const uint64_t i0 = (c - b.chunk0[d]) * FILL_CHUNK;
        const uint64_t i1 = min(i0 + FILL_CHUNK, f.n);
        for (uint64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
            uint32_t key;
            if (f.zipf) key = 1u + (uint32_t)perm_apply(f.p, zipf_rank(f.z, f.seed, i) - 1);
            else key = 1u + (uint32_t)__umul64hi(g_stream_at(f.seed, i), (uint64_t)f.domain);
            f.out[3 * i + 0] = key;
            f.out[3 * i + 1] = f.payload0 + (uint32_t)i;
            f.out[3 * i + 2] = 0u;
        }
    }
}

FillDesc make_fill(aj_stream_tuple_t *out, uint64_t n, uint64_t seed, uint32_t domain,
                   double theta, uint32_t payload0) {
    FillDesc f;
    f.out = (uint32_t *)out;
    f.n = n;
    f.seed = seed;
    f.domain = domain;
    f.payload0 = payload0;
    f.zipf = theta > 0;
    f.z = f.zipf ? make_zipf(domain, theta) : ZipfC{};
    f.p = make_perm(domain, seed ^ 0x5bd1e995ull);
    return f;
}

}  // namespace

int gen_stream_tuples(Ctx *c, aj_stream_tuple_t *d_out, uint64_t n, uint64_t seed,
                      uint32_t domain, double theta, uint32_t payload0, hipStream_t st) {
    if (!n) return AJ_OK;
    if (!d_out || domain == 0 || !(theta >= 0) || theta > 10) {
        set_error(AJ_ERR_ARG, "gen_stream_tuples: bad arguments");
        return AJ_ERR_ARG;
    }
    const int zipf = theta > 0;
    const ZipfC z = zipf ? make_zipf(domain, theta) : ZipfC{};
    const Perm p = make_perm(domain, seed ^ 0x5bd1e995ull);
    const unsigned grid = (unsigned)std::min<uint64_t>((n + 255) / 256, (uint64_t)c->num_cu * 16);
    hipLaunchKernelGGL(k_gen_tuples, dim3(grid), dim3(256), 0, st, (uint32_t *)d_out, n, seed,
                       domain, zipf, z, p, payload0);
    AJ_HIP(hipGetLastError());
    return AJ_OK;
}

// ---------------------------------------------------------------- server
struct SReq {
    uint64_t id = 0, tag = 0;
    uint64_t r_off = 0, nR = 0, s_off = 0, nS = 0;
    double close_us = 0, start_us = 0;
    bool submitted = false;  // queued or in a batch: no longer the client's to change
};

struct SSlot {
    BatchPlan plan;
    hipStream_t st = nullptr;
    hipEvent_t done = nullptr;
    uint64_t *h_ranges = nullptr;  // pinned: 4 * max_win
    uint64_t *h_res = nullptr;     // pinned: count[max_win] | checksum[max_win]
    DevBuf res;                    // device: count[max_win] | checksum[max_win]
    std::vector<SReq *> reqs;
    double start_us = 0;
    int status = AJ_OK;
};

struct Server {
    Ctx *ctx = nullptr;
    aj_server_opts_t o{};
    DevBuf wmem;  // window memory (12-byte tuples)
    aj_stream_tuple_t *base = nullptr;
    Admission adm;  // window memory, in tuples
    uint64_t workspace_bytes = 0;

    std::mutex mu;
    std::condition_variable cv;
    std::unordered_map<uint64_t, SReq *> live;  // allocated, not yet finished
    std::deque<SReq *> pending;                 // submitted, not in a batch
    std::deque<int> inflight;                   // slots in launch order
    std::vector<int> free_slots;
    std::deque<aj_server_result_t> done;
    std::vector<std::unique_ptr<SSlot>> slots;
    uint64_t next_id = 1;
    uint64_t submitted = 0, finished = 0;
    uint64_t n_batches = 0, tuples_done = 0;
    bool stop = false;
    bool dispatcher_done = false;  // the completer outlives the dispatcher

    std::mutex ingest_mu;
    hipStream_t ingest = nullptr;
    hipEvent_t ingest_ev = nullptr;
    // window fills not yet launched (under ingest_mu): launched together, at
    // the latest before the next ingest-stream operation or batch
    std::vector<FillDesc> fills;

    std::thread dispatcher, completer;
    std::chrono::steady_clock::time_point t0;

    double now_us() const {
        return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0)
            .count();
    }
};

// Launches the pending window fills on the ingest stream, FILL_MAX per
// launch (caller holds ingest_mu).  Everything enqueued on the ingest stream
// later runs after them, as if each fill had been launched when it was made.
static int flush_fills(Server *s) {
    size_t i = 0;
    while (i < s->fills.size()) {
        FillBatch b;
        b.nd = 0;
        uint64_t ch = 0;
        for (; i < s->fills.size() && b.nd < FILL_MAX; ++i) {
            b.d[b.nd] = s->fills[i];
            b.chunk0[b.nd] = ch;
            ch += (s->fills[i].n + FILL_CHUNK - 1) / FILL_CHUNK;
            ++b.nd;
        }
        b.chunk0[b.nd] = ch;
        if (ch) {
            const unsigned grid = (unsigned)std::min<uint64_t>(ch, (uint64_t)s->ctx->num_cu * 8);
            hipLaunchKernelGGL(k_gen_fills, dim3(grid), dim3(256), 0, s->ingest, b);
        }
    }
    s->fills.clear();
    AJ_HIP(hipGetLastError());
    return AJ_OK;
}

// Before window memory goes back to the arena without a batch having read it:
// a pending fill of it must reach the stream before any later window's.
static int flush_fills_locked(Server *s) {
    std::lock_guard<std::mutex> lk(s->ingest_mu);
    if (s->fills.empty()) return AJ_OK;
    AJ_HIP(hipSetDevice(s->ctx->device));
    return flush_fills(s);
}

static void publish(Server *s, SReq *r, uint64_t count, uint64_t sum, int status, int nwin,
                    double start_us, double done_us) {
    aj_server_result_t x;
    x.id = r->id;
    x.tag = r->tag;
    x.count = count;
    x.checksum = sum;
    x.status = status;
    x.batch_windows = nwin;
    x.close_us = r->close_us;
    x.start_us = start_us;
    x.done_us = done_us;
    s->done.push_back(x);
    s->adm.release(r->r_off, r->nR);
    s->adm.release(r->s_off, r->nS);
    s->live.erase(r->id);
    s->tuples_done += r->nR + r->nS;
    ++s->finished;
    delete r;
}

// Launches one batch on slot `sl` (called without the server lock).
static int launch_batch(Server *s, SSlot *sl) {
    Ctx *c = s->ctx;
    const int nwin = (int)sl->reqs.size();
    const uint64_t W = (uint64_t)sl->plan.max_win;
    uint64_t totR = 0, totS = 0;
    for (int i = 0; i < nwin; ++i) {
        const SReq *r = sl->reqs[i];
        sl->h_ranges[i] = r->r_off;
        sl->h_ranges[W + i] = r->r_off + r->nR;
        sl->h_ranges[2 * W + i] = r->s_off;
        sl->h_ranges[3 * W + i] = r->s_off + r->nS;
        totR += r->nR;
        totS += r->nS;
    }
    {
        std::lock_guard<std::mutex> lk(s->ingest_mu);
        AJ_TRY(flush_fills(s));
        AJ_HIP(hipEventRecord(s->ingest_ev, s->ingest));
        AJ_HIP(hipStreamWaitEvent(sl->st, s->ingest_ev, 0));
    }
    AJ_HIP(hipMemcpyAsync(batch_ranges(&sl->plan), sl->h_ranges, 32 * W, hipMemcpyHostToDevice,
                          sl->st));
    uint64_t *cnt = sl->res.as<uint64_t>();
    {
        std::lock_guard<std::mutex> lk(c->mu);
        AJ_TRY(enqueue_ranges(&sl->plan, s->base, totR, s->base, totS, nwin, cnt, cnt + W, sl->st));
    }
    AJ_HIP(hipMemcpyAsync(sl->h_res, cnt, sizeof(uint64_t) * nwin, hipMemcpyDeviceToHost, sl->st));
    AJ_HIP(hipMemcpyAsync(sl->h_res + W, cnt + W, sizeof(uint64_t) * nwin, hipMemcpyDeviceToHost,
                          sl->st));
    AJ_HIP(hipEventRecord(sl->done, sl->st));
    return AJ_OK;
}

static void dispatcher_main(Server *s) {
    (void)hipSetDevice(s->ctx->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv.wait(lk, [&] { return s->stop || (!s->pending.empty() && !s->free_slots.empty()); });
        if (s->stop) {
            s->dispatcher_done = true;
            s->cv.notify_all();
            break;
        }
        const int si = s->free_slots.back();
        s->free_slots.pop_back();
        SSlot *sl = s->slots[si].get();
        sl->reqs.clear();
        uint64_t totR = 0, totS = 0;
        const double now = s->now_us();
        while (!s->pending.empty() && (int)sl->reqs.size() < s->o.max_batch_windows) {
            SReq *r = s->pending.front();
            if (r->nR == 0 || r->nS == 0) {  // nothing to join
                s->pending.pop_front();
                (void)flush_fills_locked(s);
                publish(s, r, 0, 0, AJ_OK, 1, now, now);
                continue;
            }
            if (!sl->reqs.empty() && (totR + r->nR > s->o.max_batch_tuples ||
                                      totS + r->nS > s->o.max_batch_tuples))
                break;
            s->pending.pop_front();
            totR += r->nR;
            totS += r->nS;
            sl->reqs.push_back(r);
        }
        if (sl->reqs.empty()) {
            s->free_slots.push_back(si);
            s->cv.notify_all();
            continue;
        }
        sl->start_us = now;
        lk.unlock();
        sl->status = launch_batch(s, sl);
        lk.lock();
        ++s->n_batches;
        s->inflight.push_back(si);
        s->cv.notify_all();
    }
}

static void completer_main(Server *s) {
    (void)hipSetDevice(s->ctx->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv.wait(lk, [&] { return !s->inflight.empty() || s->dispatcher_done; });
        if (s->inflight.empty()) break;  // the dispatcher has stopped and nothing is in flight
        const int si = s->inflight.front();
        SSlot *sl = s->slots[si].get();
        lk.unlock();
        int status = sl->status;
        if (status == AJ_OK) {
            const hipError_t e = hipEventSynchronize(sl->done);
            if (e != hipSuccess) {
                set_error(AJ_ERR_HIP, "batch failed: %s", hipGetErrorString(e));
                status = AJ_ERR_HIP;
            }
        } else {
            // a launch that failed part-way may have queued the ranges copy and
            // some kernels: they still read window memory and h_ranges, so wait
            // for them before the windows' memory and the slot are reused
            (void)hipStreamSynchronize(sl->st);
            (void)hipGetLastError();
        }
        const double t = s->now_us();
        lk.lock();
        s->inflight.pop_front();
        const uint64_t W = (uint64_t)sl->plan.max_win;
        const int nwin = (int)sl->reqs.size();
        for (int i = 0; i < nwin; ++i)
            publish(s, sl->reqs[i], status == AJ_OK ? sl->h_res[i] : 0,
                    status == AJ_OK ? sl->h_res[W + i] : 0, status, nwin, sl->start_us, t);
        sl->reqs.clear();
        s->free_slots.push_back(si);
        s->cv.notify_all();
    }
}

static void server_free(Server *s) {
    (void)hipSetDevice(s->ctx->device);
    for (auto &sl : s->slots) {
        batch_plan_release(&sl->plan);
        sl->res.release();
        if (sl->h_ranges) (void)hipHostFree(sl->h_ranges);
        if (sl->h_res) (void)hipHostFree(sl->h_res);
        if (sl->done) (void)hipEventDestroy(sl->done);
        if (sl->st) (void)hipStreamDestroy(sl->st);
    }
    s->slots.clear();
    if (s->ingest_ev) (void)hipEventDestroy(s->ingest_ev);
    if (s->ingest) (void)hipStreamDestroy(s->ingest);
    s->wmem.release();
    for (auto &kv : s->live) delete kv.second;
    s->live.clear();
}

static int server_init(Server *s) {
    Ctx *c = s->ctx;
    AJ_HIP(hipSetDevice(c->device));
    AJ_HIP(hipStreamCreateWithFlags(&s->ingest, hipStreamNonBlocking));
    AJ_HIP(hipEventCreateWithFlags(&s->ingest_ev, hipEventDisableTiming));
    aj_batch_opts_t bo;
    memset(&bo, 0, sizeof(bo));
    bo.cap = s->o.cap;
    bo.no_graph = 1;
    // 64K-tuple chunks: a served window (median 256K per side) spreads over
    // several work-groups, so a batch of a few windows still fills the GPU (p99
    // at half load 2.9 -> 1.6 ms against 256K chunks, same saturated rate)
    bo.chunk = 65536;
    const uint64_t W = (uint64_t)s->o.max_batch_windows;
    for (int i = 0; i < s->o.slots; ++i) {
        std::unique_ptr<SSlot> sl(new SSlot());
        {
            std::lock_guard<std::mutex> lk(c->mu);
            AJ_TRY(batch_plan_setup(&sl->plan, c, s->o.max_batch_tuples, s->o.max_batch_tuples,
                                    s->o.max_batch_windows, s->o.max_window_tuples, &bo));
        }
        AJ_TRY(sl->res.ensure(16 * W));
        AJ_HIP(hipStreamCreateWithFlags(&sl->st, hipStreamNonBlocking));
        AJ_HIP(hipEventCreateWithFlags(&sl->done, hipEventDisableTiming));
        AJ_HIP(hipHostMalloc((void **)&sl->h_ranges, 32 * W, hipHostMallocDefault));
        AJ_HIP(hipHostMalloc((void **)&sl->h_res, 16 * W, hipHostMallocDefault));
        s->workspace_bytes += sl->plan.work.cap + sl->plan.hist.cap + sl->plan.scan_tmp.cap +
                              sl->plan.outR.cap + sl->plan.outS.cap + sl->res.cap;
        s->free_slots.push_back(i);
        s->slots.push_back(std::move(sl));
    }
    if (s->workspace_bytes >= s->o.hbm_budget) {
        set_error(AJ_ERR_CAPACITY,
                  "HBM budget of %llu bytes does not cover the batch workspaces (%llu bytes)",
                  (unsigned long long)s->o.hbm_budget, (unsigned long long)s->workspace_bytes);
        return AJ_ERR_CAPACITY;
    }
    const uint64_t wbytes = (s->o.hbm_budget - s->workspace_bytes) / 12 * 12;
    if (wbytes < 12) {
        set_error(AJ_ERR_CAPACITY, "no window memory left in the HBM budget");
        return AJ_ERR_CAPACITY;
    }
    AJ_TRY(s->wmem.ensure(wbytes));
    s->base = s->wmem.as<aj_stream_tuple_t>();
    s->adm.reset(wbytes / 12);
    s->t0 = std::chrono::steady_clock::now();
    s->dispatcher = std::thread(dispatcher_main, s);
    s->completer = std::thread(completer_main, s);
    return AJ_OK;
}

static int window_check(Server *s, const aj_window_t *w, SReq **out) {
    auto it = s->live.find(w->id);
    if (it == s->live.end()) {
        set_error(AJ_ERR_ARG, "unknown window %llu", (unsigned long long)w->id);
        return AJ_ERR_ARG;
    }
    *out = it->second;
    return AJ_OK;
}

}  // namespace aj

using namespace aj;

struct aj_server {
    Server s;
};

extern "C" {

int aj_gen_stream_tuples(aj_ctx_t *ctx, aj_stream_tuple_t *d_out, uint64_t n, uint64_t seed,
                         uint32_t key_domain, double theta, uint32_t payload0, void *stream) {
    if (!ctx) {
        set_error(AJ_ERR_ARG, "aj_gen_stream_tuples: NULL context");
        return AJ_ERR_ARG;
    }
    Ctx *c = ctx_of(ctx);
    AJ_HIP(hipSetDevice(c->device));
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    AJ_TRY(gen_stream_tuples(c, d_out, n, seed, key_domain, theta, payload0, st));
    if (!stream) AJ_HIP(hipStreamSynchronize(st));
    return AJ_OK;
}

int aj_server_create(aj_ctx_t *ctx, const aj_server_opts_t *opts, aj_server_t **out) {
    if (!ctx || !out) {
        set_error(AJ_ERR_ARG, "aj_server_create: NULL argument");
        return AJ_ERR_ARG;
    }
    *out = nullptr;
    std::unique_ptr<aj_server> S(new aj_server());
    Server &s = S->s;
    s.ctx = ctx_of(ctx);
    if (opts) s.o = *opts;
    if (!s.o.hbm_budget) s.o.hbm_budget = 200000000000ull;
    if (!s.o.slots) s.o.slots = 3;
    if (!s.o.max_batch_windows) s.o.max_batch_windows = 256;
    if (!s.o.max_batch_tuples) s.o.max_batch_tuples = 1ull << 27;
    if (!s.o.max_window_tuples) s.o.max_window_tuples = 1ull << 24;
    if (!s.o.cap) s.o.cap = 8704;
    if (s.o.slots < 1 || s.o.slots > 16 || s.o.max_batch_windows < 1 ||
        s.o.max_batch_windows > 65536 || s.o.max_window_tuples > s.o.max_batch_tuples ||
        2 * s.o.max_batch_tuples >= (1ull << 32)) {
        set_error(AJ_ERR_ARG, "aj_server_create: slots 1..16, windows per batch 1..65536, "
                              "max_window_tuples <= max_batch_tuples < 2^31");
        return AJ_ERR_ARG;
    }
    const int rc = server_init(&s);
    if (rc != AJ_OK) {
        server_free(&s);
        return rc;
    }
    *out = S.release();
    return AJ_OK;
}

int aj_server_destroy(aj_server_t *srv) {
    if (!srv) return AJ_OK;
    Server &s = srv->s;
    {
        std::lock_guard<std::mutex> lk(s.mu);
        s.stop = true;
        s.cv.notify_all();
    }
    s.adm.close();
    if (s.dispatcher.joinable()) s.dispatcher.join();
    if (s.completer.joinable()) s.completer.join();
    server_free(&s);
    delete srv;
    return AJ_OK;
}

int aj_server_window_alloc(aj_server_t *srv, uint64_t nR, uint64_t nS, int64_t timeout_us,
                           aj_window_t *win) {
    if (!srv || !win) {
        set_error(AJ_ERR_ARG, "aj_server_window_alloc: NULL argument");
        return AJ_ERR_ARG;
    }
    Server &s = srv->s;
    if (nR > s.o.max_window_tuples || nS > s.o.max_window_tuples) {
        set_error(AJ_ERR_ARG, "window of %llu/%llu tuples exceeds max_window_tuples %llu",
                  (unsigned long long)nR, (unsigned long long)nS,
                  (unsigned long long)s.o.max_window_tuples);
        return AJ_ERR_ARG;
    }
    {
        std::lock_guard<std::mutex> lk(s.mu);
        if (s.stop) {
            set_error(AJ_ERR_STATE, "server is shutting down");
            return AJ_ERR_STATE;
        }
    }
    uint64_t ro = 0, so = 0;
    switch (s.adm.acquire(nR, nS, timeout_us, &ro, &so)) {
        case Admission::OK: break;
        case Admission::NEVER:
            set_error(AJ_ERR_CAPACITY, "window of %llu tuples exceeds the window memory (%llu)",
                      (unsigned long long)(nR + nS), (unsigned long long)s.adm.capacity());
            return AJ_ERR_CAPACITY;
        case Admission::CLOSED:
            set_error(AJ_ERR_STATE, "server is shutting down");
            return AJ_ERR_STATE;
        default:
            set_error(AJ_ERR_CAPACITY, timeout_us == 0 ? "window memory exhausted"
                                                       : "timed out waiting for window memory");
            return AJ_ERR_CAPACITY;
    }
    std::lock_guard<std::mutex> lk(s.mu);
    SReq *r = new SReq();
    r->id = s.next_id++;
    r->r_off = ro;
    r->nR = nR;
    r->s_off = so;
    r->nS = nS;
    s.live[r->id] = r;
    win->id = r->id;
    win->dR = s.base + ro;
    win->dS = s.base + so;
    win->nR = nR;
    win->nS = nS;
    return AJ_OK;
}

int aj_server_window_release(aj_server_t *srv, const aj_window_t *win) {
    if (!srv || !win) {
        set_error(AJ_ERR_ARG, "aj_server_window_release: NULL argument");
        return AJ_ERR_ARG;
    }
    Server &s = srv->s;
    std::lock_guard<std::mutex> lk(s.mu);
    SReq *r = nullptr;
    AJ_TRY(window_check(&s, win, &r));
    if (r->submitted) {
        set_error(AJ_ERR_STATE, "window %llu is already submitted", (unsigned long long)r->id);
        return AJ_ERR_STATE;
    }
    AJ_TRY(flush_fills_locked(&s));
    s.adm.release(r->r_off, r->nR);
    s.adm.release(r->s_off, r->nS);
    s.live.erase(r->id);
    delete r;
    return AJ_OK;
}

int aj_server_window_fill(aj_server_t *srv, const aj_window_t *win, uint64_t seed,
                          uint32_t key_domain, double theta_s) {
    if (!srv || !win) {
        set_error(AJ_ERR_ARG, "aj_server_window_fill: NULL argument");
        return AJ_ERR_ARG;
    }
    Server &s = srv->s;
    {
        std::lock_guard<std::mutex> lk(s.mu);
        SReq *r = nullptr;
        AJ_TRY(window_check(&s, win, &r));
        if (r->submitted) {
            set_error(AJ_ERR_STATE, "window %llu is already submitted", (unsigned long long)r->id);
            return AJ_ERR_STATE;
        }
    }
    if (key_domain == 0 || !(theta_s >= 0) || theta_s > 10) {
        set_error(AJ_ERR_ARG, "aj_server_window_fill: bad key domain or theta");
        return AJ_ERR_ARG;
    }
    AJ_HIP(hipSetDevice(s.ctx->device));
    std::lock_guard<std::mutex> lk(s.ingest_mu);
    // a second fill of the window: the first goes out first (one launch must
    // not write the same memory twice)
    for (const FillDesc &f : s.fills)
        if (f.out == (uint32_t *)win->dR || f.out == (uint32_t *)win->dS) {
            AJ_TRY(flush_fills(&s));
            break;
        }
    if (win->nR) s.fills.push_back(make_fill(win->dR, win->nR, seed, key_domain, 0.0, 0u));
    if (win->nS)
        s.fills.push_back(
            make_fill(win->dS, win->nS, seed ^ 0x5DEECE66Dull, key_domain, theta_s, 0u));
    if (s.fills.size() >= (size_t)FILL_MAX) AJ_TRY(flush_fills(&s));
    return AJ_OK;
}

int aj_server_window_upload(aj_server_t *srv, const aj_window_t *win,
                            const aj_stream_tuple_t *hostR, const aj_stream_tuple_t *hostS) {
    if (!srv || !win || (win->nR && !hostR) || (win->nS && !hostS)) {
        set_error(AJ_ERR_ARG, "aj_server_window_upload: NULL argument");
        return AJ_ERR_ARG;
    }
    Server &s = srv->s;
    {
        std::lock_guard<std::mutex> lk(s.mu);
        SReq *r = nullptr;
        AJ_TRY(window_check(&s, win, &r));
        if (r->submitted) {
            set_error(AJ_ERR_STATE, "window %llu is already submitted", (unsigned long long)r->id);
            return AJ_ERR_STATE;
        }
    }
    AJ_HIP(hipSetDevice(s.ctx->device));
    std::lock_guard<std::mutex> lk(s.ingest_mu);
    AJ_TRY(flush_fills(&s));
    if (win->nR)
        AJ_HIP(hipMemcpyAsync(win->dR, hostR, 12 * win->nR, hipMemcpyHostToDevice, s.ingest));
    if (win->nS)
        AJ_HIP(hipMemcpyAsync(win->dS, hostS, 12 * win->nS, hipMemcpyHostToDevice, s.ingest));
    AJ_HIP(hipStreamSynchronize(s.ingest));
    return AJ_OK;
}

int aj_server_submit(aj_server_t *srv, const aj_window_t *win, uint64_t tag, double close_us) {
    if (!srv || !win) {
        set_error(AJ_ERR_ARG, "aj_server_submit: NULL argument");
        return AJ_ERR_ARG;
    }
    Server &s = srv->s;
    std::lock_guard<std::mutex> lk(s.mu);
    SReq *r = nullptr;
    AJ_TRY(window_check(&s, win, &r));
    if (s.stop) {
        set_error(AJ_ERR_STATE, "server is shutting down");
        return AJ_ERR_STATE;
    }
    if (r->submitted) {
        set_error(AJ_ERR_STATE, "window %llu is already submitted", (unsigned long long)r->id);
        return AJ_ERR_STATE;
    }
    r->submitted = true;
    r->tag = tag;
    r->close_us = close_us < 0 ? s.now_us() : close_us;
    s.pending.push_back(r);
    ++s.submitted;
    s.cv.notify_all();
    return AJ_OK;
}

int aj_server_poll(aj_server_t *srv, aj_server_result_t *out, int max, int64_t timeout_us,
                   int *n) {
    if (!srv || !n || (max > 0 && !out)) {
        set_error(AJ_ERR_ARG, "aj_server_poll: NULL argument");
        return AJ_ERR_ARG;
    }
    Server &s = srv->s;
    std::unique_lock<std::mutex> lk(s.mu);
    *n = 0;
    if (s.done.empty() && timeout_us != 0) {
        auto ready = [&] { return !s.done.empty(); };
        if (timeout_us < 0) s.cv.wait(lk, ready);
        else s.cv.wait_for(lk, std::chrono::microseconds(timeout_us), ready);
    }
    while (*n < max && !s.done.empty()) {
        out[(*n)++] = s.done.front();
        s.done.pop_front();
    }
    return AJ_OK;
}

int aj_server_drain(aj_server_t *srv) {
    if (!srv) {
        set_error(AJ_ERR_ARG, "aj_server_drain: NULL argument");
        return AJ_ERR_ARG;
    }
    Server &s = srv->s;
    std::unique_lock<std::mutex> lk(s.mu);
    s.cv.wait(lk, [&] { return s.finished >= s.submitted && s.pending.empty() && s.inflight.empty(); });
    return AJ_OK;
}

double aj_server_now_us(aj_server_t *srv) { return srv ? srv->s.now_us() : 0.0; }

int aj_server_stats(aj_server_t *srv, aj_server_stats_t *out) {
    if (!srv || !out) {
        set_error(AJ_ERR_ARG, "aj_server_stats: NULL argument");
        return AJ_ERR_ARG;
    }
    Server &s = srv->s;
    std::lock_guard<std::mutex> lk(s.mu);
    out->windows_done = s.finished;
    out->batches = s.n_batches;
    out->tuples_done = s.tuples_done;
    out->window_bytes = 12 * s.adm.capacity();
    out->window_bytes_used = 12 * s.adm.used();
    out->window_bytes_peak = 12 * s.adm.peak();
    out->workspace_bytes = s.workspace_bytes;
    out->alloc_waits = s.adm.waits();
    out->pending = s.pending.size();
    out->in_flight = s.inflight.size();
    return AJ_OK;
}

}  // extern "C"
