// pipeline.hip -- device pipelines behind the C ABI: PRO (1 or 2 partition
// passes + LDS join) and NPO (global table), on device-resident 8-byte tuples.
#include <algorithm>
#include <cstring>

#include "aj_internal.h"
#include "device_common.h"

namespace aj {

constexpr uint32_t PART_TILE = 8192;  // must match the partition tile

// Bump layout of many small device arrays inside one DevBuf.
struct Layout {
    size_t off = 0;
    template <class T> size_t add(uint64_t n) {
        size_t o = off;
        off += ((n * sizeof(T) + 255) / 256) * 256;
        return o;
    }
};
template <class T> static T *at(const DevBuf &b, size_t off) {
    return reinterpret_cast<T *>(static_cast<char *>(b.p) + off);
}

// Writes the one-segment tables of pass 1 and zeroes the scalars.
__global__ void k_setup_single(uint64_t *ib0, uint64_t *ie0, uint64_t *ib1, uint64_t *ie1,
                               uint32_t *bits, uint64_t nR, uint64_t nS, uint32_t b1,
                               uint32_t *part_seg_base, uint32_t parts, uint64_t *scalars,
                               int nscalars) {
    if (threadIdx.x == 0) {
        ib0[0] = 0;
        ie0[0] = nR;
        ib1[0] = 0;
        ie1[0] = nS;
        bits[0] = b1;
        part_seg_base[0] = 0;
        part_seg_base[1] = parts;
    }
    for (int i = threadIdx.x; i < nscalars; i += blockDim.x) scalars[i] = 0;
}

// Offsets of the "no partitioning" case: one partition holding everything.
__global__ void k_setup_nopart(uint64_t *roff, uint64_t *soff, uint64_t nR, uint64_t nS,
                               uint32_t *part_seg_base, uint64_t *scalars, int nscalars) {
    if (threadIdx.x == 0) {
        roff[0] = 0;
        roff[1] = nR;
        soff[0] = 0;
        soff[1] = nS;
        part_seg_base[0] = 0;
        part_seg_base[1] = 1;
    }
    for (int i = threadIdx.x; i < nscalars; i += blockDim.x) scalars[i] = 0;
}

// Tuples per partition work-group: about two work-groups per CU (one wave of
// work-groups; measured best at 2^27 + 2^27 tuples).
static uint32_t pick_chunk(uint64_t n_total, int num_cu) {
    uint64_t want = n_total / ((uint64_t)num_cu * 2);
    want = (want + PART_TILE - 1) / PART_TILE * PART_TILE;
    want = std::max<uint64_t>(want, PART_TILE);
    want = std::min<uint64_t>(want, 1ull << 20);
    return (uint32_t)want;
}

ProConfig choose_pro_config(uint64_t nR, uint64_t nS, const aj_join_opts_t *o) {
    ProConfig c;
    if (o && o->cap) c.cap = o->cap;
    c.split = (int)join_task_split(o ? o->split : 0u, c.cap);
    int bits;
    if (o && o->radix_bits > 0) {
        bits = o->radix_bits;
    } else {
        // average R partition of the largest power of two that fits one LDS
        // round; a round's capacity above that power of two (8704 = 8192 +
        // 512) absorbs the spread of partition sizes, so uniform keys take
        // one round per partition
        bits = 0;
        uint64_t target = 1;
        while (2 * target <= (uint64_t)c.cap) target *= 2;
        while (bits < 24 && (nR >> bits) > target) ++bits;
    }
    // one pass only while the scatter keeps runs of >= 16 tuples per 8K tile
    int passes = (o && o->passes > 0) ? o->passes : (bits > 9 ? 2 : 1);
    if (bits == 0) passes = 0;
    if (passes == 1) {
        c.bits1 = bits;
        c.bits2 = 0;
    } else if (passes == 2) {
        // an odd bit goes to pass 2: pass 1 writes into every pass-1 partition
        // at once, pass 2 only into one pass-1 partition's range
        c.bits1 = bits / 2;
        c.bits2 = bits - c.bits1;
    }
    (void)nS;
    return c;
}

struct ProPlan {
    // pass geometry
    int b1, b2;
    uint32_t chunk1, chunk2;
    uint64_t max_chunks1, max_hist1, max_parts1;
    uint64_t max_chunks2, max_hist2, max_parts2;
    uint64_t parts_final;
    // layout offsets in ctx->plan
    size_t o_ib[2], o_ie[2], o_bits1, o_bits2;
    size_t o_cpre[2], o_hb[2], o_ob[2], o_pb, o_hdr, o_poff1[2];
    size_t o_cpre2[2], o_hb2[2], o_ob2[2], o_pb2, o_hdr2, o_poff2[2];
    size_t o_tcnt, o_tbase, o_tinfo, o_tres, o_psb, o_scal;
    bool p2d;                 // pass 2 from the 2-D pass-1 histogram
    size_t o_tot, o_ub;
    uint64_t max_units, unit, max_tasks;
    size_t o_uspan;
    size_t bytes;
    size_t scan_bytes;
};

static ProPlan plan_pro(uint64_t nR, uint64_t nS, const ProConfig &c, int num_cu) {
    ProPlan p;
    memset(&p, 0, sizeof(p));
    p.b1 = c.bits1;
    p.b2 = c.bits2;
    const uint64_t F1 = 1ull << p.b1, F2 = 1ull << p.b2;
    const bool part = (p.b1 + p.b2) > 0;
    p.chunk1 = c.chunk ? (uint32_t)c.chunk : pick_chunk(nR + nS, num_cu);
    p.max_chunks1 = (nR + p.chunk1 - 1) / p.chunk1 + (nS + p.chunk1 - 1) / p.chunk1;
    p.max_hist1 = F1 * p.max_chunks1;
    p.max_parts1 = F1;
    p.parts_final = part ? F1 : 1;
    if (p.b2) {
        p.chunk2 = p.chunk1;
        p.max_chunks2 = (nR + p.chunk2 - 1) / p.chunk2 + (nS + p.chunk2 - 1) / p.chunk2 + 2 * F1;
        p.max_hist2 = F2 * p.max_chunks2;
        p.max_parts2 = F1 * F2;
        p.parts_final = F1 * F2;
    }
    Layout L;
    for (int r = 0; r < 2; ++r) {
        p.o_ib[r] = L.add<uint64_t>(1);
        p.o_ie[r] = L.add<uint64_t>(1);
    }
    p.o_bits1 = L.add<uint32_t>(1);
    p.o_bits2 = L.add<uint32_t>(F1);
    const uint64_t nseg1 = 1, nseg2 = F1;
    for (int r = 0; r < 2; ++r) {
        p.o_cpre[r] = L.add<uint64_t>(nseg1 + 1);
        p.o_hb[r] = L.add<uint64_t>(nseg1);
        p.o_ob[r] = L.add<uint64_t>(nseg1 + 1);
        p.o_poff1[r] = L.add<uint64_t>(p.max_parts1 + 2);
        p.o_cpre2[r] = L.add<uint64_t>(nseg2 + 1);
        p.o_hb2[r] = L.add<uint64_t>(nseg2);
        p.o_ob2[r] = L.add<uint64_t>(nseg2 + 1);
        p.o_poff2[r] = L.add<uint64_t>(p.max_parts2 + 2);
    }
    p.o_pb = L.add<uint32_t>(nseg1 + 1);
    p.o_pb2 = L.add<uint32_t>(nseg2 + 1);
    p.o_hdr = L.add<uint64_t>(8);
    p.o_hdr2 = L.add<uint64_t>(8);
    const uint64_t jparts = p.parts_final;
    p.o_tcnt = L.add<uint32_t>(jparts + 1);
    p.o_tbase = L.add<uint32_t>(jparts + 1);
    p.max_tasks = join_max_tasks(jparts, nS, (uint32_t)c.split);
    p.o_tinfo = L.add<uint4>(2 * p.max_tasks);
    p.o_tres = L.add<ulonglong2>(p.max_tasks);
    p.o_psb = L.add<uint32_t>(2);
    p.o_scal = L.add<uint64_t>(8);
    p.p2d = p.b2 > 0 && p.b1 + p.b2 <= P2D_MAX_BITS && c.p2d;
    if (p.p2d) {
        // pass-2 units of about 40 per CU (at least one 16384-tuple tile of the
        // pass-2 scatter, at most a pass-1 chunk)
        p.unit = std::max<uint64_t>(2 * PART_TILE, std::min<uint64_t>(p.chunk1, (nR + nS) / ((uint64_t)num_cu * 40)));
        p.max_units = (nR + nS) / p.unit + 2 * F1 + 1;
        p.o_tot = L.add<uint32_t>(p.max_units * F2);
        p.o_ub = L.add<uint32_t>(2 * F1 + 2);
        p.o_uspan = L.add<uint64_t>(2 * p.max_units);
    }
    p.bytes = L.off;
    p.scan_bytes = std::max({pass_scan_tmp_bytes(std::max<uint64_t>(p.max_hist1, 1)),
                             pass_scan_tmp_bytes(std::max<uint64_t>(p.max_hist2, 1)),
                             join_scan_tmp_bytes(jparts)});
    return p;
}

static int ensure_phase_events(Ctx *c) {
    for (hipEvent_t &e : c->evk)
        if (!e) AJ_HIP(hipEventCreate(&e));
    return AJ_OK;
}

static void read_phases(Ctx *c, aj_join_out_t *out, int first, int last) {
    for (int i = first; i < last; ++i) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, c->evk[i], c->evk[i + 1]) != hipSuccess) {
            (void)hipGetLastError();
            ms = 0;
        }
        out->phase_ms[i] = ms;
    }
}

// scalars: [0]=count [1]=checksum [2]=pair cursor [3]=task counter (u32 in low half)
__global__ void k_clock_stamp(uint64_t *slot) {
    if (threadIdx.x == 0) *slot = __builtin_amdgcn_s_memrealtime();
}

int device_clock_stamp(uint64_t *slot, hipStream_t st) {
    hipLaunchKernelGGL(k_clock_stamp, dim3(1), dim3(64), 0, st, slot);
    AJ_HIP(hipGetLastError());
    return AJ_OK;
}

__global__ void k_staged_setup(uint32_t *psb, uint32_t parts, uint64_t *scalars, int nscalars) {
    if (threadIdx.x == 0) {
        psb[0] = 0;
        psb[1] = parts;
    }
    for (int i = threadIdx.x; i < nscalars; i += blockDim.x) scalars[i] = 0;
}

// Staged partitioning (staged.hip) is taken for two passes of at most 8 bits
// each on inputs of at least 2^20 tuples in total.
static bool staged_applies(uint64_t nR, uint64_t nS, const ProConfig &c) {
    return c.staged && c.bits1 >= 1 && c.bits2 >= 1 && c.bits1 <= 8 && c.bits2 <= 8 &&
           nR + nS >= (1ull << 20);
}

static float event_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
        (void)hipGetLastError();
        ms = 0;
    }
    return ms;
}

// PRO with staged partitioning.  *overflow is set (and out left unset) when a
// partition did not fit its capacity; the caller then joins with the
// histogram-based passes.
static int pro_staged(Ctx *ctx, const uint64_t *dR, uint64_t nR, const uint64_t *dS,
                      uint64_t nS, const ProConfig &cfg, uint64_t *pairs, uint64_t pairs_cap,
                      int records, aj_join_out_t *out, hipStream_t st, bool *overflow) {
    *overflow = false;
    StagedDesc sd;
    const int sets = cfg.stage_sets ? cfg.stage_sets : 4;
    // pass 1 in 512-lane work-groups (two per CU: one's tile work overlaps the
    // other's loads; 0.85 against 0.87 ms per 2^28 tuples), pass 2 in 1024-lane
    // ones (16K-tuple tiles: runs of 128 tuples per digit; with 8K tiles pass 2
    // took 1.22 against 1.02 ms); option stage_block sets both, and the other
    // cache modes keep 1024 lanes in both
    const int block = cfg.stage_block ? cfg.stage_block : 1024;
    const int block1 = cfg.stage_block ? cfg.stage_block
                       : (cfg.stage_cache < 0 || cfg.stage_cache == 9) ? 512 : 1024;
    AJ_TRY(staged_geometry(&sd.g, nR, nS, cfg.bits1, cfg.bits2, cfg.stage_chunk, sets, block,
                           ctx->num_cu, block1));
    if (cfg.stage_cache >= 0) sd.g.cache_mode = cfg.stage_cache;
    const StagedGeom &g = sd.g;
    const uint32_t F1 = 1u << g.b1, F = F1 << g.b2;
    const uint32_t split = (uint32_t)cfg.split;
    const uint64_t max_tasks = staged_max_tasks(g, split);
    Layout L;
    const size_t o_scur = L.add<uint32_t>(2ull * g.nset * F1);
    const size_t o_fcur = L.add<uint32_t>(2ull * F);
    const size_t o_tbase = L.add<uint32_t>((uint64_t)F + 1);
    const size_t o_tinfo = L.add<uint4>(2 * max_tasks);
    const size_t o_tres = L.add<ulonglong2>(max_tasks);
    const size_t o_psb = L.add<uint32_t>(2);
    const size_t o_scal = L.add<uint64_t>(8);
    AJ_TRY(ctx->plan.ensure(L.off));
    AJ_TRY(ctx->stage.ensure(g.stage_tuples * 8));
    AJ_TRY(ctx->p2R.ensure(g.out_tuples[0] * 8));
    AJ_TRY(ctx->p2S.ensure(g.out_tuples[1] * 8));
    const DevBuf &B = ctx->plan;
    uint64_t *scal = at<uint64_t>(B, o_scal);
    uint32_t *psb = at<uint32_t>(B, o_psb);
    sd.in[0] = dR;
    sd.in[1] = dS;
    sd.n[0] = nR;
    sd.n[1] = nS;
    sd.stage = ctx->stage.as<uint64_t>();
    sd.out[0] = ctx->p2R.as<uint64_t>();
    sd.out[1] = ctx->p2S.as<uint64_t>();
    sd.scur = at<uint32_t>(B, o_scur);
    sd.fcur = at<uint32_t>(B, o_fcur);
    sd.flag = reinterpret_cast<uint32_t *>(scal + 4);
    const bool timing = phase_timing(ctx);
    std::vector<hipEvent_t> *evs = nullptr;
    Marks mk;
    const Marks *mkp = nullptr;
    if (timing) {
        AJ_TRY(ensure_phase_events(ctx));
        const size_t need = (size_t)staged_events(sd);
        while (ctx->evs.size() < need) {
            hipEvent_t e;
            AJ_HIP(hipEventCreate(&e));
            ctx->evs.push_back(e);
        }
        evs = &ctx->evs;
        mk.ev = ctx->evk;
        mk.st = st;
        mkp = &mk;
    }
    AJ_HIP(record_event(ctx->ev0, st));
    mk.mark(0);
    hipLaunchKernelGGL(k_staged_setup, dim3(1), dim3(64), 0, st, psb, F, scal, 8);
    AJ_TRY(run_staged(sd, st, evs));
    for (int i = 1; i < 9; ++i) mk.mark(i);
    AJ_TRY(staged_tasks(sd, split, at<uint4>(B, o_tinfo), at<uint32_t>(B, o_tbase), F, max_tasks,
                        st));
    JoinDesc j;
    j.R = sd.out[0];
    j.S = sd.out[1];
    j.roff = j.soff = nullptr;
    j.nparts_dev = nullptr;
    j.max_parts = F;
    j.part_seg_base = psb;
    j.nseg = 1;
    j.hash_shift = g.b1 + g.b2;
    j.cap = cfg.cap;
    j.split = split;
    j.seg_count = scal + 0;
    j.seg_sum = scal + 1;
    j.pairs = pairs;
    j.pairs_cap = pairs_cap;
    j.records = records;
    j.pair_cursor = scal + 2;
    j.task_cnt = nullptr;
    j.task_base = at<uint32_t>(B, o_tbase);
    j.task_next = reinterpret_cast<uint32_t *>(scal + 3);
    j.task_info = at<uint4>(B, o_tinfo);
    j.task_res = at<ulonglong2>(B, o_tres);
    j.max_tasks = max_tasks;
    j.scan_tmp = nullptr;
    j.scan_tmp_bytes = 0;
    j.num_cu = ctx->num_cu;
    j.tasks_ready = true;
    AJ_TRY(run_join(j, st, mkp, 9));
    AJ_HIP(record_event(ctx->ev1, st));
    uint64_t local[HRES_SLOTS];
    uint64_t *host = result_slots(ctx, local);
    AJ_HIP(hipMemcpyAsync(host, scal, 5 * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    AJ_HIP(hipStreamSynchronize(st));
    const float ms = elapsed_ms(ctx->ev0, ctx->ev1);
    if ((uint32_t)host[4]) {
        *overflow = true;
        out->device_ms = ms;  // the attempt's time, added to the fallback's
        return AJ_OK;
    }
    out->count = host[0];
    out->checksum = host[1];
    out->npairs = pairs ? host[2] : 0;
    out->device_ms = ms;
    out->flags = AJ_OUT_STAGED;
    if (timing) {
        read_phases(ctx, out, AJ_PH_PLAN1, AJ_PH_JOIN + 1);
        const int ne = staged_events(sd);
        float s1 = 0, s2 = 0;
        for (int e = 0; e + 1 < ne; ++e) {
            const float t = event_ms(ctx->evs[e], ctx->evs[e + 1]);
            if (e & 1)
                s2 += t;
            else
                s1 += t;
        }
        out->phase_ms[AJ_PH_SCATTER1] = s1;
        out->phase_ms[AJ_PH_SCATTER2] = s2;
        out->phase_ms[AJ_PH_PLAN1] = std::max(0.0f, out->phase_ms[AJ_PH_PLAN1] - s1 - s2);
    }
    return AJ_OK;
}

// One pass of at most 11 radix bits without a histogram: the join is a window
// batch of one segment (windows.hip), whose staged pass reserves each tile
// run's room by atomics, spills S runs that overflow a partition into the
// partition's extension range, and reruns the pass from a histogram on the
// device only when an R partition overflows.  Counting joins on inputs of at
// least 2^20 tuples in total; the context's plan is kept for later calls.
static bool staged1_applies(const Ctx *ctx, uint64_t nR, uint64_t nS, const ProConfig &c,
                            const uint64_t *pairs) {
    return c.staged && staged_enabled() && !ctx->arena_only && !pairs && c.bits2 == 0 &&
           c.bits1 >= 1 && c.bits1 <= WSTAGE_MAX_BITS && nR + nS >= (1ull << 20) &&
           (c.cap == 4096 || c.cap == 8192 || c.cap == 8704 || c.cap == 12288);
}

__global__ void k_one_range(uint64_t *rb, uint64_t nR, uint64_t nS) {
    if (threadIdx.x == 0) {
        rb[0] = 0;   // rb
        rb[1] = nR;  // re
        rb[2] = 0;   // sb
        rb[3] = nS;  // se
    }
}

static int pro_staged1(Ctx *ctx, const uint64_t *dR, uint64_t nR, const uint64_t *dS,
                       uint64_t nS, const ProConfig &cfg, aj_join_out_t *out, hipStream_t st) {
    // the window plan sizes a window's bits by its R tuples over `target`
    const uint32_t target = (uint32_t)((nR + (1ull << cfg.bits1) - 1) >> cfg.bits1);
    BatchPlan *&p = ctx->pro1;
    if (!p || p->max_R < nR || p->max_S < nS || p->max_window != nR || p->target != target ||
        p->cap != cfg.cap || p->split != join_task_split((uint32_t)cfg.split, cfg.cap) ||
        p->staged != staged_enabled()) {
        if (p) {
            batch_plan_release(p);
            delete p;
            p = nullptr;
        }
        p = new BatchPlan();
        aj_batch_opts_t bo;
        memset(&bo, 0, sizeof(bo));
        bo.cap = cfg.cap;
        bo.split = (uint32_t)cfg.split;
        bo.target = std::max<uint32_t>(target, 1);
        bo.no_graph = 1;
        const int rc = batch_plan_setup(p, ctx, nR, nS, 1, nR, &bo);
        if (rc != AJ_OK) {
            batch_plan_release(p);
            delete p;
            p = nullptr;
            return rc;
        }
    }
    AJ_TRY(ctx->acc.ensure(64));
    uint64_t *res = ctx->acc.as<uint64_t>();
    Marks mk;
    const Marks *mkp = nullptr;
    if (phase_timing(ctx)) {
        AJ_TRY(ensure_phase_events(ctx));
        mk.ev = ctx->evk;
        mk.st = st;
        mkp = &mk;
    }
    AJ_HIP(record_event(ctx->ev0, st));
    mk.mark(0);
    hipLaunchKernelGGL(k_one_range, dim3(1), dim3(64), 0, st, batch_ranges(p), nR, nS);
    AJ_TRY(enqueue_ranges(p, dR, nR, dS, nS, 1, res, res + 1, st, 8, mkp, 0));
    AJ_HIP(record_event(ctx->ev1, st));
    uint64_t local[HRES_SLOTS];
    uint64_t *host = result_slots(ctx, local);
    host[2] = 0;
    AJ_HIP(hipMemcpyAsync(host, res, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    if (p->staged)
        AJ_HIP(hipMemcpyAsync(host + 2, batch_fallback_flag(p), sizeof(uint32_t),
                              hipMemcpyDeviceToHost, st));
    AJ_HIP(hipStreamSynchronize(st));
    const uint32_t flag = (uint32_t)host[2];
    out->count = host[0];
    out->checksum = host[1];
    out->npairs = 0;
    out->device_ms = elapsed_ms(ctx->ev0, ctx->ev1);
    out->flags = p->staged ? (flag ? AJ_OUT_FALLBACK : AJ_OUT_STAGED) : 0;
    if (mkp) read_phases(ctx, out, AJ_PH_PLAN1, AJ_PH_JOIN + 1);
    return AJ_OK;
}

int pro_device(Ctx *ctx, const uint64_t *dR, uint64_t nR, const uint64_t *dS, uint64_t nS,
               const ProConfig &cfg, uint64_t *pairs, uint64_t pairs_cap, int records,
               aj_join_out_t *out, hipStream_t st) {
    if (staged1_applies(ctx, nR, nS, cfg, pairs)) return pro_staged1(ctx, dR, nR, dS, nS, cfg, out, st);
    float attempt_ms = 0;
    if (staged_applies(nR, nS, cfg)) {
        bool overflow = false;
        AJ_TRY(pro_staged(ctx, dR, nR, dS, nS, cfg, pairs, pairs_cap, records, out, st,
                          &overflow));
        if (!overflow) return AJ_OK;
        attempt_ms = out->device_ms;
        memset(out, 0, sizeof(*out));
        out->flags = AJ_OUT_FALLBACK;
    }
    ProPlan p = plan_pro(nR, nS, cfg, ctx->num_cu);
    AJ_TRY(ctx->plan.ensure(p.bytes));
    AJ_TRY(ctx->scan_tmp.ensure(p.scan_bytes));
    AJ_TRY(ctx->hist.ensure(sizeof(uint32_t) * 2 * std::max(p.max_hist1, p.max_hist2) + 256));
    const bool part = (p.b1 + p.b2) > 0;
    if (part) {
        AJ_TRY(ctx->p1R.ensure(nR * 8 + 256));
        AJ_TRY(ctx->p1S.ensure(nS * 8 + 256));
        if (p.b2) {
            AJ_TRY(ctx->p2R.ensure(nR * 8 + 256));
            AJ_TRY(ctx->p2S.ensure(nS * 8 + 256));
        }
        if (p.p2d) AJ_TRY(ctx->hist2d.ensure(sizeof(uint32_t) * p.max_chunks1 << (p.b1 + p.b2)));
    }
    const DevBuf &B = ctx->plan;
    uint64_t *scal = at<uint64_t>(B, p.o_scal);
    uint32_t *psb = at<uint32_t>(B, p.o_psb);

    AJ_HIP(record_event(ctx->ev0, st));
    Marks mk;
    const Marks *mkp = nullptr;
    if (phase_timing(ctx)) {
        AJ_TRY(ensure_phase_events(ctx));
        mk.ev = ctx->evk;
        mk.st = st;
        mkp = &mk;
        mk.mark(0);
    }
    const uint64_t *jR = dR, *jS = dS, *roff, *soff;
    if (!part) {
        roff = at<uint64_t>(B, p.o_poff1[0]);
        soff = at<uint64_t>(B, p.o_poff1[1]);
        hipLaunchKernelGGL(k_setup_nopart, dim3(1), dim3(64), 0, st, (uint64_t *)roff,
                           (uint64_t *)soff, nR, nS, psb, scal, 8);
    } else {
        hipLaunchKernelGGL(k_setup_single, dim3(1), dim3(64), 0, st, at<uint64_t>(B, p.o_ib[0]),
                           at<uint64_t>(B, p.o_ie[0]), at<uint64_t>(B, p.o_ib[1]),
                           at<uint64_t>(B, p.o_ie[1]), at<uint32_t>(B, p.o_bits1), nR, nS,
                           (uint32_t)p.b1, psb, (uint32_t)p.parts_final, scal, 8);
        PassDesc d1;
        d1.in[0] = dR;
        d1.in[1] = dS;
        d1.in_stride = 8;
        d1.out[0] = ctx->p1R.as<uint64_t>();
        d1.out[1] = ctx->p1S.as<uint64_t>();
        for (int r = 0; r < 2; ++r) {
            d1.in_begin[r] = at<uint64_t>(B, p.o_ib[r]);
            d1.in_end[r] = at<uint64_t>(B, p.o_ie[r]);
        }
        d1.bits = at<uint32_t>(B, p.o_bits1);
        d1.nseg = 1;
        d1.shift = 0;
        d1.max_bits = p.b1;
        d1.chunk = p.chunk1;
        d1.block = cfg.block;
        if (p.p2d) {
            d1.hist2d = ctx->hist2d.as<uint32_t>();
            d1.bits2 = p.b2;
        }
        d1.n_total[0] = nR;
        d1.n_total[1] = nS;
        d1.exact_single = true;
        d1.max_chunks = p.max_chunks1;
        d1.max_hist = p.max_hist1;
        d1.max_parts = p.max_parts1;
        PassWork w1;
        w1.hist = ctx->hist.as<uint32_t>();
        for (int r = 0; r < 2; ++r) {
            w1.seg_chunk_pre[r] = at<uint64_t>(B, p.o_cpre[r]);
            w1.seg_hist_base[r] = at<uint64_t>(B, p.o_hb[r]);
            w1.seg_out_base[r] = at<uint64_t>(B, p.o_ob[r]);
            w1.part_off[r] = at<uint64_t>(B, p.o_poff1[r]);
        }
        w1.seg_part_base = at<uint32_t>(B, p.o_pb);
        w1.header = at<uint64_t>(B, p.o_hdr);
        w1.scan_tmp = ctx->scan_tmp.p;
        w1.scan_tmp_bytes = p.scan_bytes;
        AJ_TRY(run_partition_pass(d1, w1, st, mkp, 1));
        jR = ctx->p1R.as<uint64_t>();
        jS = ctx->p1S.as<uint64_t>();
        roff = w1.part_off[0];
        soff = w1.part_off[1];
        if (p.p2d) {
            P2Desc q;
            q.in[0] = jR;
            q.in[1] = jS;
            q.out[0] = ctx->p2R.as<uint64_t>();
            q.out[1] = ctx->p2S.as<uint64_t>();
            q.b1 = p.b1;
            q.b2 = p.b2;
            q.hist2d = ctx->hist2d.as<uint32_t>();
            q.unit_counts = at<uint32_t>(B, p.o_tot);
            q.unit_span = at<uint64_t>(B, p.o_uspan);
            for (int r = 0; r < 2; ++r) q.poff2[r] = at<uint64_t>(B, p.o_poff2[r]);
            q.ubase = at<uint32_t>(B, p.o_ub);
            q.w1 = &w1;
            q.hist_scan = w1.hist + p.max_hist1;
            q.unit = p.unit;
            q.max_units = p.max_units;
            q.block = cfg.block;
            AJ_TRY(run_pass2_2d(q, st, mkp, 5));
            jR = ctx->p2R.as<uint64_t>();
            jS = ctx->p2S.as<uint64_t>();
            roff = q.poff2[0];
            soff = q.poff2[1];
        } else if (p.b2) {
            AJ_TRY(fill_u32(at<uint32_t>(B, p.o_bits2), 1ull << p.b1, (uint32_t)p.b2, st));
            PassDesc d2 = d1;
            d2.in[0] = jR;
            d2.in[1] = jS;
            d2.out[0] = ctx->p2R.as<uint64_t>();
            d2.out[1] = ctx->p2S.as<uint64_t>();
            for (int r = 0; r < 2; ++r) {
                d2.in_begin[r] = w1.part_off[r];
                d2.in_end[r] = w1.part_off[r] + 1;
            }
            d2.bits = at<uint32_t>(B, p.o_bits2);
            d2.nseg = 1 << p.b1;
            d2.shift = p.b1;
            d2.max_bits = p.b2;
            d2.chunk = p.chunk2;
            d2.max_chunks = p.max_chunks2;
            d2.max_hist = p.max_hist2;
            d2.max_parts = p.max_parts2;
            PassWork w2 = w1;
            for (int r = 0; r < 2; ++r) {
                w2.seg_chunk_pre[r] = at<uint64_t>(B, p.o_cpre2[r]);
                w2.seg_hist_base[r] = at<uint64_t>(B, p.o_hb2[r]);
                w2.seg_out_base[r] = at<uint64_t>(B, p.o_ob2[r]);
                w2.part_off[r] = at<uint64_t>(B, p.o_poff2[r]);
            }
            w2.seg_part_base = at<uint32_t>(B, p.o_pb2);
            w2.header = at<uint64_t>(B, p.o_hdr2);
            AJ_TRY(run_partition_pass(d2, w2, st, mkp, 5));
            jR = ctx->p2R.as<uint64_t>();
            jS = ctx->p2S.as<uint64_t>();
            roff = w2.part_off[0];
            soff = w2.part_off[1];
        }
    }
    if (mkp) {  // phases that did not run take zero time
        for (int i = part ? (p.b2 ? 9 : 5) : 1; i < 9; ++i) mk.mark(i);
    }
    JoinDesc j;
    j.R = jR;
    j.S = jS;
    j.roff = roff;
    j.soff = soff;
    j.nparts_dev = nullptr;
    j.max_parts = p.parts_final;
    j.part_seg_base = psb;
    j.nseg = 1;
    j.hash_shift = p.b1 + p.b2;
    j.cap = cfg.cap;
    j.split = cfg.split;
    j.seg_count = scal + 0;
    j.seg_sum = scal + 1;
    j.pairs = pairs;
    j.pairs_cap = pairs_cap;
    j.records = records;
    j.pair_cursor = scal + 2;
    j.task_cnt = at<uint32_t>(B, p.o_tcnt);
    j.task_base = at<uint32_t>(B, p.o_tbase);
    j.task_next = reinterpret_cast<uint32_t *>(scal + 3);
    j.task_info = at<uint4>(B, p.o_tinfo);
    j.task_res = at<ulonglong2>(B, p.o_tres);
    j.max_tasks = p.max_tasks;
    j.scan_tmp = ctx->scan_tmp.p;
    j.scan_tmp_bytes = p.scan_bytes;
    j.num_cu = ctx->num_cu;
    AJ_TRY(run_join(j, st, mkp, 9));
    AJ_HIP(record_event(ctx->ev1, st));
    uint64_t local[HRES_SLOTS];
    uint64_t *host = result_slots(ctx, local);
    AJ_HIP(hipMemcpyAsync(host, scal, 3 * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    AJ_HIP(hipStreamSynchronize(st));
    const float ms = elapsed_ms(ctx->ev0, ctx->ev1);
    out->count = host[0];
    out->checksum = host[1];
    out->npairs = pairs ? host[2] : 0;
    out->device_ms = ms + attempt_ms;
    if (mkp) read_phases(ctx, out, AJ_PH_PLAN1, AJ_PH_JOIN + 1);
    return AJ_OK;
}

int npo_device(Ctx *ctx, const uint64_t *dR, uint64_t nR, const uint64_t *dS, uint64_t nS,
               uint64_t *pairs, uint64_t pairs_cap, int records, aj_join_out_t *out,
               hipStream_t st) {
    uint64_t cap = 2;  // load at most 1/2: one slot per distinct key
    int cap_log2 = 1;
    while (cap < 2 * nR) cap <<= 1, ++cap_log2;
    AJ_TRY(ctx->table.ensure(cap * 16));
    // Region build (npo.hip) for counting joins of at least npo_region_min R
    // tuples whose table splits into 4 to 2^16 regions of 4096 slots (8192 for
    // a table of 2^29 slots): R partitioned by hashed key with the staged passes.
    const int rs_log2 = cap_log2 <= 28 ? 12 : 13, rbits = cap_log2 - rs_log2;
    const uint64_t rmin = npo_region_min();
    const bool region = !pairs && rmin && nR >= rmin && rbits >= 2 && rbits <= 16;
    StagedDesc sd;
    if (region) {
        const int b1 = (rbits + 1) / 2, b2 = rbits / 2;
        AJ_TRY(staged_geometry(&sd.g, nR, nR, b1, b2, 0, 4, 1024, ctx->num_cu));
        sd.g.cache_mode = 9 | 16;  // keys hashed in pass 1
        const uint32_t F1 = 1u << b1, F = F1 << b2;
        Layout L;
        const size_t o_scur = L.add<uint32_t>(2ull * sd.g.nset * F1);
        const size_t o_fcur = L.add<uint32_t>(2ull * F);
        AJ_TRY(ctx->plan.ensure(L.off));
        AJ_TRY(ctx->stage.ensure(sd.g.stage_tuples * 8));
        AJ_TRY(ctx->p2R.ensure(sd.g.out_tuples[0] * 8));
        sd.in[0] = dR;
        sd.in[1] = nullptr;
        sd.n[0] = nR;
        sd.n[1] = 0;
        sd.stage = ctx->stage.as<uint64_t>();
        sd.out[0] = ctx->p2R.as<uint64_t>();
        sd.out[1] = nullptr;
        sd.scur = at<uint32_t>(ctx->plan, o_scur);
        sd.fcur = at<uint32_t>(ctx->plan, o_fcur);
    }
    if (pairs) {  // chains of the further tuples of a key, to write every match
        AJ_TRY(ctx->chain.ensure(nR * 4 + 4));
        AJ_TRY(ctx->heads.ensure(cap * 4));
    }
    AJ_TRY(ctx->acc.ensure(8 * npo_acc_words(ctx->num_cu)));
    uint64_t *acc = ctx->acc.as<uint64_t>();
    AJ_HIP(record_event(ctx->ev0, st));
    AJ_TRY(fill_u64(acc, 8, 0, st));
    NpoDesc d;
    d.R = dR;
    d.S = dS;
    d.nR = nR;
    d.nS = nS;
    d.table = ctx->table.as<uint64_t>();
    d.capacity = cap;
    d.head = pairs ? ctx->heads.as<uint32_t>() : nullptr;
    d.next = pairs ? ctx->chain.as<uint32_t>() : nullptr;
    d.acc = acc;
    if (region) {
        sd.flag = reinterpret_cast<uint32_t *>(acc + 4);
        d.stage = &sd;
        d.geom.b1 = (uint32_t)sd.g.b1;
        d.geom.b2 = (uint32_t)sd.g.b2;
        d.geom.rs_log2 = (uint32_t)rs_log2;
    }
    d.pairs = pairs;
    d.pairs_cap = pairs_cap;
    d.records = records;
    d.num_cu = ctx->num_cu;
    Marks mk;
    const Marks *mkp = nullptr;
    if (phase_timing(ctx)) {
        AJ_TRY(ensure_phase_events(ctx));
        mk.ev = ctx->evk;
        mk.st = st;
        mkp = &mk;
    }
    AJ_TRY(run_npo(d, st, mkp, AJ_PH_NPO_BUILD));
    AJ_HIP(record_event(ctx->ev1, st));
    uint64_t local[HRES_SLOTS];
    uint64_t *host = result_slots(ctx, local);
    AJ_HIP(hipMemcpyAsync(host, acc, 5 * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    AJ_HIP(hipStreamSynchronize(st));
    const float ms = elapsed_ms(ctx->ev0, ctx->ev1);
    out->count = host[0];
    out->checksum = host[1];
    out->npairs = pairs ? host[3] : 0;
    out->device_ms = ms;
    out->flags = region ? ((uint32_t)host[4] ? AJ_OUT_FALLBACK : AJ_OUT_STAGED) : 0u;
    if (mkp) read_phases(ctx, out, AJ_PH_NPO_BUILD, AJ_PH_NPO_PROBE + 1);
    return AJ_OK;
}

}  // namespace aj
