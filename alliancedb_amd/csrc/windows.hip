// windows.hip -- batched intra-window joins (many tumbling windows per launch)
// and the device-side stream generator.
//
// Reference: one lazy join per process over one window whose size is
// window/interval*step tuples (benchmark.cpp:164-205), with timestamps
// assigned by add_ts (generator.cpp:133-180).  Here a batch holds many
// windows of two timestamp-ordered streams of 12-byte {key, payload, ts}
// tuples, and every window is joined in the same kernels:
//
//   k_window_bounds  window w = [bound[w], bound[w+1]) in ts; its tuple range
//                    in each stream by binary search (streams are ts-ordered)
//   k_window_bits    radix bits per window from its R size (about half an
//                    LDS round per partition)
//   partition pass   one segmented pass over both streams: reads the 12-byte
//                    tuples, writes 8-byte tuples window-major and
//                    partition-major (partition.hip)
//   join             the persistent LDS join (join.hip) with per-window 64-bit
//                    count/checksum accumulators
//
// The sequence is captured once into a hipGraph per plan and replayed; new
// input pointers or sizes update the captured graph in place.
//
// Bytes per stream tuple: 12 (histogram read) + 12 (scatter read) + 8 (write)
// + 8 (build/probe read) = 40.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>

#include "aj_internal.h"
#include "device_common.h"

namespace aj {

constexpr uint32_t WTILE = 16384;  // the window scatter's tile: chunks are whole tiles

__global__ void k_window_bounds(const uint32_t *R, uint64_t nR, const uint32_t *S, uint64_t nS,
                                const uint32_t *bounds, int nwin, uint64_t *rb, uint64_t *re,
                                uint64_t *sb, uint64_t *se) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w > nwin) return;
    const uint32_t t = bounds[w];
    // first index with ts >= t (ts is word 2 of each 12-byte tuple)
    uint64_t lo = 0, hi = nR;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (R[3 * mid + 2] < t) lo = mid + 1;
        else hi = mid;
    }
    const uint64_t ri = lo;
    lo = 0;
    hi = nS;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (S[3 * mid + 2] < t) lo = mid + 1;
        else hi = mid;
    }
    const uint64_t si = lo;
    if (w < nwin) {
        rb[w] = ri;
        sb[w] = si;
    }
    if (w > 0) {
        re[w - 1] = ri;
        se[w - 1] = si;
    }
}

__global__ void k_window_bits(const uint64_t *rb, const uint64_t *re, int nwin, uint32_t *bits,
                              uint32_t target, int max_bits) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nwin) return;
    const uint64_t n = re[w] - rb[w];
    int b = 0;
    while (b < max_bits && (n >> b) > target) ++b;
    bits[w] = (uint32_t)b;
}

static size_t lay(size_t &off, size_t bytes) {
    size_t o = off;
    off += (bytes + 255) / 256 * 256;
    return o;
}

int batch_plan_init(BatchPlan *p) {
    const uint64_t W = (uint64_t)p->max_win;
    int b = 0;
    while (b < 13 && (p->max_window >> b) > p->target) ++b;
    p->max_bits = b;
    p->max_chunks = (p->max_R + p->chunk - 1) / p->chunk + (p->max_S + p->chunk - 1) / p->chunk +
                    2 * W;
    p->max_hist = (1ull << b) * p->max_chunks;
    // partitions per window: 2^bits with (n >> (bits-1)) > target, so at most
    // 2n/target + 1 per window
    p->max_parts = std::min<uint64_t>(W << b, 2 * p->max_R / p->target + W);
    // the staged pass's S spill area: max_S (any skew of S fits), listed runs of
    // 8 tuples on average (a hot key past its partition's capacity spills a run
    // from every tile that holds it: at 11 bits, Zipf keys spill ~10^2 short runs
    // per tile)
    p->spill_cap = std::max<uint64_t>(p->max_S, 2 * WTILE);
    p->spill_max = (uint32_t)std::min<uint64_t>(p->spill_cap / 8 + 4096, 1u << 30);
    // join entries: each partition, and (staged) each partition's S extension
    const uint64_t jparts = 2 * p->max_parts;
    BatchLayout &L = p->L;
    size_t off = 0;
    // window ranges: rb | re | sb | se, W entries each, in one block
    L.rb = lay(off, 32 * W);
    L.re = L.rb + 8 * W;
    L.sb = L.rb + 16 * W;
    L.se = L.rb + 24 * W;
    L.bits = lay(off, 4 * W);
    for (int r = 0; r < 2; ++r) {
        L.cpre[r] = lay(off, 8 * (W + 1));
        L.hb[r] = lay(off, 8 * W);
        L.ob[r] = lay(off, 8 * (W + 1));
        L.poff[r] = lay(off, 8 * (p->max_parts + 2));
    }
    L.pb = lay(off, 4 * (W + 1));
    L.hdr = lay(off, 64);
    L.tcnt = lay(off, 4 * (jparts + 1));
    L.tbase = lay(off, 4 * (jparts + 1));
    L.tinfo = lay(off, 2 * sizeof(uint4) * join_max_tasks(jparts, p->max_S, p->split));
    L.tres = lay(off, sizeof(ulonglong2) * join_max_tasks(jparts, p->max_S, p->split));
    L.scal = lay(off, 64);
    L.wcap = lay(off, 2 * 4 * W);
    L.wbase = lay(off, 2 * 8 * W);
    for (int r = 0; r < 2; ++r) {
        L.wcur[r] = lay(off, 4 * (p->max_parts + 1));
        L.pstart[r] = lay(off, 8 * (jparts + 2));
        L.plen[r] = lay(off, 4 * (jparts + 2));
    }
    L.strad = lay(off, 4 * (p->max_parts + 1));
    L.extoff = lay(off, 4 * (p->max_parts + 1));
    L.splist = lay(off, sizeof(uint4) * (uint64_t)p->spill_max);
    L.bytes = off;
    p->scan_bytes = std::max(pass_scan_tmp_bytes(std::max<uint64_t>(p->max_hist, 1)),
                             join_scan_tmp_bytes(jparts));
    AJ_TRY(p->work.ensure(L.bytes));
    AJ_TRY(p->hist.ensure(sizeof(uint32_t) * 2 * p->max_hist + 256));
    AJ_TRY(p->scan_tmp.ensure(p->scan_bytes));
    // the staged pass needs room for its partitions' slack; windows of more
    // radix bits than its LDS holds, or layouts past 32-bit offsets, use the
    // histogram pass
    p->out_cap[0] = p->max_R;
    p->out_cap[1] = p->max_S;
    if (p->staged) {
        const uint64_t cR = window_stage_bound(p->max_R, p->max_parts);
        const uint64_t cS = window_stage_bound(p->max_S, p->max_parts);
        if (p->max_bits <= WSTAGE_MAX_BITS && cR < (1ull << 32) &&
            cS + p->spill_cap < (1ull << 32)) {
            p->out_cap[0] = cR;
            p->out_cap[1] = cS;
        } else {
            p->staged = false;
        }
    }
    AJ_TRY(p->outR.ensure(8 * p->out_cap[0] + 256));
    AJ_TRY(p->outS.ensure(8 * (p->out_cap[1] + (p->staged ? 2 * p->spill_cap : 0)) + 256));
    return AJ_OK;
}

template <class T> static T *W_(const BatchPlan *p, size_t off) {
    return reinterpret_cast<T *>(static_cast<char *>(p->work.p) + off);
}

// Enqueues the whole batch on `st` (no host synchronisation; capturable).
static int enqueue_batch(BatchPlan *p, const aj_stream_tuple_t *dR, uint64_t nR,
                         const aj_stream_tuple_t *dS, uint64_t nS, const uint32_t *bounds, int nwin,
                         uint64_t *cnt, uint64_t *sum, hipStream_t st) {
    const BatchLayout &L = p->L;
    const unsigned g = (unsigned)((nwin + 1 + 255) / 256);
    hipLaunchKernelGGL(k_window_bounds, dim3(g), dim3(256), 0, st, (const uint32_t *)dR, nR,
                       (const uint32_t *)dS, nS, bounds, nwin, W_<uint64_t>(p, L.rb),
                       W_<uint64_t>(p, L.re), W_<uint64_t>(p, L.sb), W_<uint64_t>(p, L.se));
    return enqueue_ranges(p, dR, nR, dS, nS, nwin, cnt, sum, st);
}

uint64_t *batch_ranges(BatchPlan *p) { return W_<uint64_t>(p, p->L.rb); }
const uint32_t *batch_fallback_flag(BatchPlan *p) {
    return reinterpret_cast<const uint32_t *>(W_<uint64_t>(p, p->L.scal) + 5);
}

// Enqueues the joins of nwin windows whose tuple ranges (in tuples of dR and
// dS) are already in the plan's range block.
int enqueue_ranges(BatchPlan *p, const void *dR, uint64_t nR, const void *dS, uint64_t nS,
                   int nwin, uint64_t *cnt, uint64_t *sum, hipStream_t st, int stride,
                   const Marks *mk, int m0) {
    const BatchLayout &L = p->L;
    AJ_TRY(fill_u64(cnt, nwin, 0, st));
    AJ_TRY(fill_u64(sum, nwin, 0, st));
    AJ_TRY(fill_u64(W_<uint64_t>(p, L.scal), 8, 0, st));
    // scalars: [2] pair cursor, [3] task counter, [5] fallback / layout flags,
    // [6] spill area cursor / spilled runs, [7] extension cursor
    uint32_t *flag = reinterpret_cast<uint32_t *>(W_<uint64_t>(p, L.scal) + 5);
    uint32_t *spill = reinterpret_cast<uint32_t *>(W_<uint64_t>(p, L.scal) + 6);
    if (p->staged)
        for (int r = 0; r < 2; ++r) AJ_TRY(fill_u32(W_<uint32_t>(p, L.wcur[r]), p->max_parts, 0, st));
    const unsigned g = (unsigned)((nwin + 255) / 256);
    hipLaunchKernelGGL(k_window_bits, dim3(g), dim3(256), 0, st, W_<uint64_t>(p, L.rb),
                       W_<uint64_t>(p, L.re), nwin, W_<uint32_t>(p, L.bits), p->target,
                       p->max_bits);
    PassDesc d;
    d.in[0] = dR;
    d.in[1] = dS;
    d.in_stride = stride;
    d.out[0] = p->outR.as<uint64_t>();
    d.out[1] = p->outS.as<uint64_t>();
    d.in_begin[0] = W_<uint64_t>(p, L.rb);
    d.in_end[0] = W_<uint64_t>(p, L.re);
    d.in_begin[1] = W_<uint64_t>(p, L.sb);
    d.in_end[1] = W_<uint64_t>(p, L.se);
    d.bits = W_<uint32_t>(p, L.bits);
    d.nseg = nwin;
    d.shift = 0;
    d.max_bits = p->max_bits;
    d.chunk = p->chunk;
    d.n_total[0] = nR;
    d.n_total[1] = nS;
    d.max_chunks = p->max_chunks;
    d.max_hist = p->max_hist;
    d.max_parts = p->max_parts;
    PassWork w;
    w.hist = p->hist.as<uint32_t>();
    for (int r = 0; r < 2; ++r) {
        w.seg_chunk_pre[r] = W_<uint64_t>(p, L.cpre[r]);
        w.seg_hist_base[r] = W_<uint64_t>(p, L.hb[r]);
        w.seg_out_base[r] = W_<uint64_t>(p, L.ob[r]);
        w.part_off[r] = W_<uint64_t>(p, L.poff[r]);
    }
    w.seg_part_base = W_<uint32_t>(p, L.pb);
    w.header = W_<uint64_t>(p, L.hdr);
    w.scan_tmp = p->scan_tmp.p;
    w.scan_tmp_bytes = p->scan_bytes;
    JoinDesc j;
    if (p->staged) {
        // the staged pass first; the histogram pass runs (and rewrites the
        // partitions exactly) only when the staged one set the overflow flag
        WStageDesc ws;
        ws.cap = W_<uint32_t>(p, L.wcap);
        ws.base = W_<uint64_t>(p, L.wbase);
        ws.flag = flag;
        ws.lflag = flag + 1;
        ws.W = p->max_win;
        ws.strad = W_<uint32_t>(p, L.strad);
        ws.spill_cur = spill;
        ws.spill_n = spill + 1;
        ws.ext_cur = spill + 2;
        ws.spill_list = W_<uint4>(p, L.splist);
        ws.spill_cap = p->spill_cap;
        ws.spill_max = p->spill_max;
        ws.ext_off = W_<uint32_t>(p, L.extoff);
        for (int r = 0; r < 2; ++r) {
            ws.cur[r] = W_<uint32_t>(p, L.wcur[r]);
            ws.out_cap[r] = p->out_cap[r];
            ws.pstart[r] = W_<uint64_t>(p, L.pstart[r]);
            ws.plen[r] = W_<uint32_t>(p, L.plen[r]);
        }
        AJ_TRY(run_partition_plan(d, w, st));
        if (mk)
            for (int i = 1; i < 4; ++i) mk->mark(m0 + i);
        AJ_TRY(run_window_stage(d, w, ws, st));
        d.guard = flag;
        d.skip_plan = true;
        AJ_TRY(run_partition_pass(d, w, st));
        AJ_TRY(run_window_parts(d, w, ws, st));
        if (mk)
            for (int i = 4; i < 9; ++i) mk->mark(m0 + i);
        j.roff = ws.pstart[0];
        j.soff = ws.pstart[1];
        j.rlen = ws.plen[0];
        j.slen = ws.plen[1];
        j.max_parts = 2 * p->max_parts;
        j.nreal = p->max_parts;
    } else {
        if (mk)
            for (int i = 1; i < 4; ++i) mk->mark(m0 + i);
        AJ_TRY(run_partition_pass(d, w, st));
        if (mk)
            for (int i = 4; i < 9; ++i) mk->mark(m0 + i);
        j.roff = w.part_off[0];
        j.soff = w.part_off[1];
        j.max_parts = p->max_parts;
    }
    j.R = p->outR.as<uint64_t>();
    j.S = p->outS.as<uint64_t>();
    j.nparts_dev = nullptr;
    j.part_seg_base = w.seg_part_base;
    j.nseg = nwin;
    j.hash_shift = 0;
    j.seg_bits = W_<uint32_t>(p, L.bits);  // one pass of bits[w] radix bits per window
    j.cap = p->cap;
    j.split = p->split;
    j.seg_count = cnt;
    j.seg_sum = sum;
    j.pairs = nullptr;
    j.pairs_cap = 0;
    j.pair_cursor = W_<uint64_t>(p, L.scal) + 2;
    j.task_cnt = W_<uint32_t>(p, L.tcnt);
    j.task_base = W_<uint32_t>(p, L.tbase);
    j.task_next = reinterpret_cast<uint32_t *>(W_<uint64_t>(p, L.scal) + 3);
    j.task_info = W_<uint4>(p, L.tinfo);
    j.task_res = W_<ulonglong2>(p, L.tres);
    j.max_tasks = join_max_tasks(j.max_parts, p->max_S, p->split);
    j.scan_tmp = p->scan_tmp.p;
    j.scan_tmp_bytes = p->scan_bytes;
    j.num_cu = p->ctx->num_cu;
    AJ_TRY(run_join(j, st, mk, m0 + 9));
    return AJ_OK;
}

static int check_batch_args(const BatchPlan *p, uint64_t nR, uint64_t nS, int nwin) {
    if (nR > p->max_R || nS > p->max_S || nwin > p->max_win || nwin < 1) {
        set_error(AJ_ERR_ARG, "batch exceeds its plan (nR %llu/%llu, nS %llu/%llu, windows %d/%d)",
                  (unsigned long long)nR, (unsigned long long)p->max_R, (unsigned long long)nS,
                  (unsigned long long)p->max_S, nwin, p->max_win);
        return AJ_ERR_ARG;
    }
    return AJ_OK;
}

int batch_plan_setup(BatchPlan *p, Ctx *c, uint64_t max_R, uint64_t max_S, int max_windows,
                     uint64_t max_window_tuples, const aj_batch_opts_t *opts) {
    p->ctx = c;
    p->max_R = std::max<uint64_t>(max_R, 1);
    p->max_S = std::max<uint64_t>(max_S, 1);
    p->max_win = max_windows;
    p->max_window = max_window_tuples ? max_window_tuples : p->max_R;
    p->cap = (opts && opts->cap) ? opts->cap : 8704;
    p->split = join_task_split(opts ? opts->split : 0u, p->cap);
    // default target: the largest power of two that fits one LDS round (the
    // rest of the round absorbs the spread of partition sizes)
    uint32_t t = 1;
    while (2 * t <= (uint32_t)p->cap) t *= 2;
    p->target = (opts && opts->target) ? opts->target : t;
    uint32_t chunk = (opts && opts->chunk) ? opts->chunk : 0;
    if (!chunk) {
        const uint64_t want = (max_R + max_S) / ((uint64_t)c->num_cu * 4);
        chunk = (uint32_t)std::min<uint64_t>(
            std::max<uint64_t>((want + WTILE - 1) / WTILE * WTILE, WTILE), 1u << 20);
    }
    p->chunk = (chunk + WTILE - 1) / WTILE * WTILE;
    p->use_graph = !(opts && opts->no_graph);
    p->staged = staged_enabled();
    if (p->cap != 4096 && p->cap != 8192 && p->cap != 8704 && p->cap != 12288) {
        set_error(AJ_ERR_ARG, "cap must be 4096, 8192, 8704 or 12288");
        return AJ_ERR_ARG;
    }
    return batch_plan_init(p);
}

void batch_plan_release(BatchPlan *p) {
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->cap_stream) (void)hipStreamDestroy(p->cap_stream);
    p->exec = nullptr;
    p->cap_stream = nullptr;
    p->work.release();
    p->hist.release();
    p->scan_tmp.release();
    p->outR.release();
    p->outS.release();
}

// ---------------------------------------------------------------- generator
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ uint64_t stream_at(uint64_t seed, uint64_t i) {
    return mix64(seed + (i + 1) * 0x9E3779B97F4A7C15ull);
}

// exponential inter-arrival gaps in nanoseconds for a rate of `rate` tuples/us
__global__ void k_gen_gaps(uint64_t *gaps, uint64_t n, uint64_t seed, double ns_per_tuple) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const double u = ((stream_at(seed, i) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        gaps[i] = (uint64_t)llrint(-log(u) * ns_per_tuple);
    }
}

// ts = arrival (us); key uniform on [1, max(1, nominal_w / dup)] of its window
__global__ void k_gen_fill(const uint64_t *arrival_ns, uint64_t n, uint64_t seed,
                           const uint32_t *bounds, const uint32_t *nominal, int nwin, uint32_t dup,
                           uint32_t *out) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t us = arrival_ns[i] / 1000;
        const uint32_t ts = us > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)us;
        // window of ts: last w with bounds[w] <= ts
        int lo = 0, hi = nwin + 1;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (bounds[mid] <= ts) lo = mid;
            else hi = mid;
        }
        const uint32_t w = (uint32_t)min(lo, nwin - 1);
        const uint32_t domain = max(1u, nominal[w] / max(dup, 1u));
        const uint32_t key =
            1u + (uint32_t)__umul64hi(stream_at(seed ^ 0x6A09E667F3BCC909ull, i), (uint64_t)domain);
        out[3 * i + 0] = key;
        out[3 * i + 1] = (uint32_t)i;
        out[3 * i + 2] = ts;
    }
}

}  // namespace aj

using namespace aj;

struct aj_batch_plan {
    BatchPlan p;
};

extern "C" {

int aj_batch_plan_create(aj_ctx_t *ctx, uint64_t max_R, uint64_t max_S, int max_windows,
                         uint64_t max_window_tuples, const aj_batch_opts_t *opts,
                         aj_batch_plan_t **out) {
    if (!ctx || !out || max_windows < 1 || max_R + max_S >= (1ull << 32)) {
        set_error(AJ_ERR_ARG, "aj_batch_plan_create: bad arguments");
        return AJ_ERR_ARG;
    }
    *out = nullptr;
    Ctx *c = ctx_of(ctx);
    std::lock_guard<std::mutex> lk(c->mu);
    AJ_HIP(hipSetDevice(c->device));
    std::unique_ptr<aj_batch_plan> P(new aj_batch_plan());
    BatchPlan &p = P->p;
    AJ_TRY(batch_plan_setup(&p, c, max_R, max_S, max_windows, max_window_tuples, opts));
    AJ_HIP(hipStreamCreateWithFlags(&p.cap_stream, hipStreamNonBlocking));
    *out = P.release();
    return AJ_OK;
}

int aj_batch_plan_destroy(aj_batch_plan_t *plan) {
    if (!plan) return AJ_OK;
    BatchPlan &p = plan->p;
    {
        std::lock_guard<std::mutex> lk(p.ctx->mu);
        (void)hipSetDevice(p.ctx->device);
        batch_plan_release(&p);
    }
    delete plan;
    return AJ_OK;
}

int aj_batch_run(aj_batch_plan_t *plan, const aj_stream_tuple_t *dR, uint64_t nR,
                 const aj_stream_tuple_t *dS, uint64_t nS, const uint32_t *d_bounds, int nwin,
                 uint64_t *d_count, uint64_t *d_checksum, void *stream) {
    if (!plan || !d_bounds || !d_count || !d_checksum || (nR && !dR) || (nS && !dS)) {
        set_error(AJ_ERR_ARG, "aj_batch_run: NULL argument");
        return AJ_ERR_ARG;
    }
    BatchPlan &p = plan->p;
    AJ_TRY(check_batch_args(&p, nR, nS, nwin));
    std::lock_guard<std::mutex> lk(p.ctx->mu);
    AJ_HIP(hipSetDevice(p.ctx->device));
    hipStream_t st = stream ? (hipStream_t)stream : p.ctx->stream;
    if (!p.use_graph) return enqueue_batch(&p, dR, nR, dS, nS, d_bounds, nwin, d_count, d_checksum, st);
    BatchPlan::Key k{dR, dS, d_bounds, nR, nS, nwin, d_count, d_checksum};
    const BatchPlan::Key &o = p.key;
    const bool same = p.have_key && k.R == o.R && k.S == o.S && k.bounds == o.bounds &&
                      k.nR == o.nR && k.nS == o.nS && k.nwin == o.nwin && k.cnt == o.cnt &&
                      k.sum == o.sum;
    if (!same) {
        if (!p.exec) {
            // one eager run first: sets kernel attributes outside of any capture
            AJ_TRY(enqueue_batch(&p, dR, nR, dS, nS, d_bounds, nwin, d_count, d_checksum, p.cap_stream));
            AJ_HIP(hipStreamSynchronize(p.cap_stream));
        }
        hipGraph_t g = nullptr;
        AJ_HIP(hipStreamBeginCapture(p.cap_stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_batch(&p, dR, nR, dS, nS, d_bounds, nwin, d_count, d_checksum, p.cap_stream);
        hipError_t e = hipStreamEndCapture(p.cap_stream, &g);
        if (rc != AJ_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) {
            set_error(AJ_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
            return AJ_ERR_HIP;
        }
        bool updated = false;
        if (p.exec) {
            hipGraphExecUpdateResult res;
            hipGraphNode_t err_node = nullptr;
            updated = hipGraphExecUpdate(p.exec, g, &err_node, &res) == hipSuccess &&
                      res == hipGraphExecUpdateSuccess;
            if (!updated) {
                (void)hipGetLastError();
                (void)hipGraphExecDestroy(p.exec);
                p.exec = nullptr;
            }
        }
        if (!updated) {
            hipError_t ie = hipGraphInstantiate(&p.exec, g, nullptr, nullptr, 0);
            if (ie != hipSuccess) {
                (void)hipGraphDestroy(g);
                set_error(AJ_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(ie));
                return AJ_ERR_HIP;
            }
        }
        (void)hipGraphDestroy(g);
        p.key = k;
        p.have_key = true;
    }
    AJ_HIP(hipGraphLaunch(p.exec, st));
    return AJ_OK;
}

int aj_gen_window_streams(aj_ctx_t *ctx, uint64_t seed, int nwin, const uint32_t *sizes,
                          double rate_per_us, uint32_t dup, aj_stream_tuple_t *d_out,
                          uint64_t capacity, uint32_t *d_bounds, uint64_t *n_out) {
    if (!ctx || nwin < 1 || !sizes || !d_out || !d_bounds || !n_out || !(rate_per_us > 0)) {
        set_error(AJ_ERR_ARG, "aj_gen_window_streams: bad arguments");
        return AJ_ERR_ARG;
    }
    Ctx *c = ctx_of(ctx);
    std::lock_guard<std::mutex> lk(c->mu);
    AJ_HIP(hipSetDevice(c->device));
    // boundaries: window w lasts sizes[w] / rate microseconds
    std::vector<uint32_t> b(nwin + 1);
    double t = 0, total = 0;
    for (int w = 0; w < nwin; ++w) {
        b[w] = (uint32_t)llround(t);
        t += sizes[w] / rate_per_us;
        total += sizes[w];
    }
    b[nwin] = (uint32_t)llround(t);
    if (t >= 4294967295.0) {
        set_error(AJ_ERR_ARG, "stream longer than 2^32 microseconds");
        return AJ_ERR_ARG;
    }
    uint64_t n = (uint64_t)(total + 8 * std::sqrt(total) + 1024);
    if (n > capacity) n = capacity;
    hipStream_t st = c->stream;
    AJ_TRY(c->misc.ensure(n * 16 + 512));
    AJ_TRY(c->seg.ensure(sizeof(uint32_t) * nwin + 256));
    AJ_HIP(hipMemcpyAsync(d_bounds, b.data(), sizeof(uint32_t) * (nwin + 1), hipMemcpyHostToDevice, st));
    AJ_HIP(hipMemcpyAsync(c->seg.p, sizes, sizeof(uint32_t) * nwin, hipMemcpyHostToDevice, st));
    uint64_t *gaps = c->misc.as<uint64_t>();
    uint64_t *arrival = gaps + ((n + 31) / 32 * 32);
    const unsigned grid = (unsigned)std::min<uint64_t>((n + 255) / 256, (uint64_t)c->num_cu * 32);
    hipLaunchKernelGGL(k_gen_gaps, dim3(grid), dim3(256), 0, st, gaps, n, seed, 1000.0 / rate_per_us);
    const size_t tmp = scan_tmp_bytes(n, sizeof(uint64_t));
    AJ_TRY(c->scan_tmp.ensure(tmp));
    AJ_TRY(inclusive_scan_u64(gaps, arrival, n, c->scan_tmp.p, tmp, st));
    hipLaunchKernelGGL(k_gen_fill, dim3(grid), dim3(256), 0, st, (const uint64_t *)arrival, n, seed,
                       (const uint32_t *)d_bounds, (const uint32_t *)c->seg.p, nwin, dup,
                       (uint32_t *)d_out);
    AJ_HIP(hipGetLastError());
    AJ_HIP(hipStreamSynchronize(st));
    *n_out = n;
    return AJ_OK;
}

}  // extern "C"
