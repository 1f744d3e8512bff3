// shj.hip -- eager symmetric hash join (SHJ) on two HBM-resident hash tables.
//
// Reference: every arriving tuple is inserted into its own side's table and
// probed against the other side's, so each match is produced as soon as its
// later tuple arrives (SHJJoiner::join, helper/localjoiner.cpp:73-95; the
// single-thread driver shj alternates R[i], S[i] in index order, :33-64).  The
// multithreaded variants only change which thread sees which tuples: join
// matrix (every thread holds all of R and a slice of S, fetcher.h:146-190) or
// join biclique with content-sensitive groups and random S placement
// (ContRandShuffler, shuffler.cpp:117-166); the matches are the same.
//
// CDNA4 form: arrivals are taken in steps of `batch` tuples per side.  Step k
// runs two kernels: R[k*batch, (k+1)*batch) is inserted into TR and probed
// against TS, then S[k*batch, (k+1)*batch) is inserted into TS and probed
// against TR.  The kernel boundary is the only ordering needed: a pair (r_i,
// s_j) is found exactly once, in the step of its later tuple (r first on a
// tie), which for batch = 1 is the reference's alternating order.  Tables are
// the NPO layout (open addressing over whole 64-bit tuples, capacity the next
// power of two >= 2n, murmur3-finalised home slot, all-ones empty marker with
// all-ones tuples counted on the side).  Per-step cumulative match counts give
// the progressive result curve.
//
// Bytes per tuple: read 8 + own slot write 8 + probe slot reads (>= 8).
#include "aj_internal.h"
#include "device_common.h"

namespace aj {

constexpr int SBLOCK = 256;

// Insert v into tab (the all-ones tuple is counted in *own_sent instead) and
// count / checksum its matches in other (other_sent all-ones tuples there).
// Work-group b adds its matches to step_cnt[b] (this step's row) and its
// checksum to sum_part[b]: every launch has the same grid and launches are
// stream-ordered, so each slot has one writer at a time and no atomics contend.
__global__ __launch_bounds__(SBLOCK) void k_shj_step(const uint64_t *in, uint64_t lo,
                                                     uint64_t hi, uint64_t *tab, uint64_t mask,
                                                     uint64_t *own_sent, const uint64_t *other,
                                                     uint64_t omask, const uint64_t *other_sent,
                                                     int is_r, uint64_t *sum_part,
                                                     uint64_t *step_cnt) {
    __shared__ uint64_t red[SBLOCK / 64];
    const uint64_t nsent = *other_sent;  // not written during this kernel
    uint64_t cnt = 0, sum = 0, sent = 0;
    for (uint64_t i = lo + blockIdx.x * (uint64_t)SBLOCK + threadIdx.x; i < hi;
         i += (uint64_t)gridDim.x * SBLOCK) {
        const uint64_t v = in[i];
        const uint32_t k = key_of(v);
        // probe the other side first (its table does not change in this kernel)
        uint64_t h = fmix32(k) & omask;
        for (;;) {
            const uint64_t t = other[h];
            if (t == EMPTY_SLOT) break;
            if (key_of(t) == k) {
                ++cnt;
                sum += is_r ? pair_value(v, t) : pair_value(t, v);
            }
            h = (h + 1) & omask;
        }
        if (k == 0xFFFFFFFFu && nsent) {  // all-ones tuples of the other side
            cnt += nsent;
            const uint64_t pv = is_r ? pair_value(v, EMPTY_SLOT) : pair_value(EMPTY_SLOT, v);
            sum += nsent * pv;
        }
        if (v == EMPTY_SLOT) {
            ++sent;
            continue;
        }
        h = fmix32(k) & mask;
        // terminates: fewer tuples than slots
        while (atomicCAS((unsigned long long *)&tab[h], (unsigned long long)EMPTY_SLOT,
                         (unsigned long long)v) != EMPTY_SLOT)
            h = (h + 1) & mask;
    }
    sent = block_sum<SBLOCK>(sent, red);
    cnt = block_sum<SBLOCK>(cnt, red);
    sum = block_sum<SBLOCK>(sum, red);
    if (threadIdx.x == 0) {
        if (sent) atomicAdd((unsigned long long *)own_sent, (unsigned long long)sent);
        if (cnt) {
            step_cnt[blockIdx.x] += cnt;
            sum_part[blockIdx.x] += sum;
        }
    }
}

// The materialising form of k_shj_step: the same inserts, probes, per-block
// counts and checksums, and every match also written to out: mode 1 the pair
// (R payload << 32 | S payload), mode 2 the probing tuple <key, payloadID> (the
// record SHJJoiner::join's probe appends, common_functions.cpp:251-257), mode 3
// a timed record (device clock << 32 | the probing tuple's payloadID, the
// sample END_PROGRESSIVE_MEASURE takes there).  A wave counts its lanes'
// matches, reserves one range with a single atomic on *cursor, then walks its
// probe chains again to write; positions at or past cap are counted but not
// written (the caller reports the shortfall).  In mode 3 an R tuple's records
// fill out from the end backwards (its cursor counts them), so the caller can
// tell the two probe sides apart.  The
// loop runs a work-group-uniform number of rounds so the wave operations see
// every lane.
__global__ __launch_bounds__(SBLOCK) void k_shj_step_out(
    const uint64_t *in, uint64_t lo, uint64_t hi, uint64_t *tab, uint64_t mask,
    uint64_t *own_sent, const uint64_t *other, uint64_t omask, const uint64_t *other_sent,
    int is_r, uint64_t *sum_part, uint64_t *step_cnt, int mode, uint64_t *out, uint64_t cap,
    uint64_t *cursor) {
    __shared__ uint64_t red[SBLOCK / 64];
    const uint64_t nsent = *other_sent;  // not written during this kernel
    uint64_t cnt = 0, sum = 0, sent = 0;
    for (uint64_t base = lo + blockIdx.x * (uint64_t)SBLOCK; base < hi;
         base += (uint64_t)gridDim.x * SBLOCK) {
        const uint64_t i = base + threadIdx.x;
        const bool valid = i < hi;
        const uint64_t v = valid ? in[i] : 0;
        const uint32_t k = key_of(v);
        uint64_t c = 0;
        if (valid) {
            for (uint64_t h = fmix32(k) & omask;; h = (h + 1) & omask) {
                const uint64_t t = other[h];
                if (t == EMPTY_SLOT) break;
                c += key_of(t) == k;
            }
            if (k == 0xFFFFFFFFu) c += nsent;  // all-ones tuples of the other side
        }
        const uint64_t incl = wave_inclusive_scan(c);
        const uint64_t total = __shfl(incl, 63);
        uint64_t wbase = 0;
        if (lane_id() == 63 && total)
            wbase = atomicAdd((unsigned long long *)cursor, (unsigned long long)total);
        uint64_t pos = __shfl(wbase, 63) + incl - c;
        const uint32_t tnow = mode == 3 ? (uint32_t)__builtin_amdgcn_s_memrealtime() : 0u;
        if (c) {
            auto emit = [&](uint64_t t) {
                const uint64_t pv = is_r ? pair_value(v, t) : pair_value(t, v);
                sum += pv;
                if (pos < cap) {
                    if (mode == 3)
                        out[is_r ? cap - 1 - pos : pos] = ((uint64_t)tnow << 32) | payload_of(v);
                    else
                        out[pos] = mode == 2 ? v : pv;
                }
                ++pos;
            };
            for (uint64_t h = fmix32(k) & omask;; h = (h + 1) & omask) {
                const uint64_t t = other[h];
                if (t == EMPTY_SLOT) break;
                if (key_of(t) == k) emit(t);
            }
            if (k == 0xFFFFFFFFu)
                for (uint64_t n = 0; n < nsent; ++n) emit(EMPTY_SLOT);
            cnt += c;
        }
        if (!valid) continue;
        if (v == EMPTY_SLOT) {
            ++sent;
            continue;
        }
        uint64_t h = fmix32(k) & mask;
        // terminates: fewer tuples than slots
        while (atomicCAS((unsigned long long *)&tab[h], (unsigned long long)EMPTY_SLOT,
                         (unsigned long long)v) != EMPTY_SLOT)
            h = (h + 1) & mask;
    }
    sent = block_sum<SBLOCK>(sent, red);
    cnt = block_sum<SBLOCK>(cnt, red);
    sum = block_sum<SBLOCK>(sum, red);
    if (threadIdx.x == 0) {
        if (sent) atomicAdd((unsigned long long *)own_sent, (unsigned long long)sent);
        if (cnt) {
            step_cnt[blockIdx.x] += cnt;
            sum_part[blockIdx.x] += sum;
        }
    }
}

// Work-group k: rows[k] = matches of step k, the sum of the per-block row
// step_cnt[k * G + b] (coalesced reads across the work-group).
__global__ __launch_bounds__(SBLOCK) void k_shj_rows(const uint64_t *step_cnt, unsigned G,
                                                     uint64_t *rows) {
    __shared__ uint64_t red[SBLOCK / 64];
    const uint64_t k = blockIdx.x;
    uint64_t c = 0;
    for (unsigned b = threadIdx.x; b < G; b += SBLOCK) c += step_cnt[k * G + b];
    c = block_sum<SBLOCK>(c, red);
    if (threadIdx.x == 0) rows[k] = c;
}

// One work-group: prog[k] = matches of steps 0..k (in place over the step
// rows, a carried block scan 256 steps at a time); acc[0] = all matches,
// acc[1] = the checksum from the per-block partials.
__global__ __launch_bounds__(SBLOCK) void k_shj_finish(uint64_t steps, unsigned G,
                                                       const uint64_t *sum_part, uint64_t *prog,
                                                       uint64_t *acc) {
    __shared__ uint64_t red[SBLOCK / 64 + 1];
    uint64_t carry = 0;
    for (uint64_t base = 0; base < steps; base += SBLOCK) {
        const uint64_t k = base + threadIdx.x;
        const uint64_t c = k < steps ? prog[k] : 0;
        uint64_t total;
        const uint64_t pre = block_exclusive_scan<SBLOCK>(c, red, &total);
        if (k < steps) prog[k] = carry + pre + c;
        carry += total;
    }
    uint64_t x = 0;
    for (unsigned b = threadIdx.x; b < G; b += SBLOCK) x += sum_part[b];
    x = block_sum<SBLOCK>(x, red);
    if (threadIdx.x == 0) {
        acc[0] = carry;
        acc[1] = x;
    }
}

namespace {

unsigned shj_grid(uint64_t n, int num_cu) {
    uint64_t g = (n + SBLOCK - 1) / SBLOCK;
    const uint64_t cap = (uint64_t)num_cu * 8;
    if (g > cap) g = cap;
    return (unsigned)(g ? g : 1);
}

uint64_t table_cap(uint64_t n) {
    uint64_t c = 2;
    while (c <= 2 * n) c <<= 1;
    return c;
}

}  // namespace

uint64_t shj_steps(uint64_t nR, uint64_t nS, uint64_t batch) {
    const uint64_t n = nR > nS ? nR : nS;
    return batch ? (n + batch - 1) / batch : 0;
}

int shj_device(Ctx *ctx, const uint64_t *dR, uint64_t nR, const uint64_t *dS, uint64_t nS,
               uint64_t batch, uint64_t *progress, uint64_t *pairs, uint64_t pcap, int records,
               aj_join_out_t *out, hipStream_t st) {
    if (batch == 0) {
        set_error(AJ_ERR_ARG, "SHJ batch must be >= 1");
        return AJ_ERR_ARG;
    }
    // 1 pairs, 2 the reference's records, 3 timed records (R probes from the end)
    const int mode = pairs ? (records == 2 ? 3 : records ? 2 : 1) : 0;
    const uint64_t capR = table_cap(nR), capS = table_cap(nS);
    const uint64_t steps = shj_steps(nR, nS, batch);
    if (steps > (1ull << 20)) {
        set_error(AJ_ERR_ARG, "SHJ batch %llu gives %llu steps (limit 2^20)",
                  (unsigned long long)batch, (unsigned long long)steps);
        return AJ_ERR_ARG;
    }
    const unsigned G = shj_grid(batch, ctx->num_cu);  // one grid for every launch
    AJ_TRY(ctx->table.ensure(capR * 8));
    AJ_TRY(ctx->table2.ensure(capS * 8));
    AJ_TRY(ctx->acc.ensure(8 * 8));
    AJ_TRY(ctx->steps.ensure(8 * (steps * G + G + steps)));
    // [0] count [1] sum [2] R sentinels [3] S sentinels [4] output cursor
    // [5] R-probe output cursor (timed records)
    uint64_t *acc = ctx->acc.as<uint64_t>();
    uint64_t *tr = ctx->table.as<uint64_t>(), *ts = ctx->table2.as<uint64_t>();
    uint64_t *scnt = ctx->steps.as<uint64_t>();  // [steps][G] per-block step matches
    uint64_t *sum_part = scnt + steps * G;       // [G] per-block checksums
    uint64_t *prog = sum_part + G;               // [steps] cumulative matches
    AJ_HIP(record_event(ctx->ev0, st));
    AJ_TRY(fill_u64(acc, 8, 0, st));
    AJ_TRY(fill_u64(tr, capR, EMPTY_SLOT, st));
    AJ_TRY(fill_u64(ts, capS, EMPTY_SLOT, st));
    AJ_TRY(fill_u64(scnt, steps * G + G, 0, st));
    for (uint64_t k = 0; k < steps; ++k) {
        const uint64_t lo = k * batch;
        if (lo < nR) {
            const uint64_t hi = lo + batch < nR ? lo + batch : nR;
            if (mode)
                hipLaunchKernelGGL(k_shj_step_out, dim3(G), dim3(SBLOCK), 0, st, dR, lo, hi, tr,
                                   capR - 1, acc + 2, (const uint64_t *)ts, capS - 1,
                                   (const uint64_t *)(acc + 3), 1, sum_part, scnt + k * G, mode,
                                   pairs, pcap, mode == 3 ? acc + 5 : acc + 4);
            else
                hipLaunchKernelGGL(k_shj_step, dim3(G), dim3(SBLOCK), 0, st, dR, lo, hi, tr,
                                   capR - 1, acc + 2, (const uint64_t *)ts, capS - 1,
                                   (const uint64_t *)(acc + 3), 1, sum_part, scnt + k * G);
        }
        if (lo < nS) {
            const uint64_t hi = lo + batch < nS ? lo + batch : nS;
            if (mode)
                hipLaunchKernelGGL(k_shj_step_out, dim3(G), dim3(SBLOCK), 0, st, dS, lo, hi, ts,
                                   capS - 1, acc + 3, (const uint64_t *)tr, capR - 1,
                                   (const uint64_t *)(acc + 2), 0, sum_part, scnt + k * G, mode,
                                   pairs, pcap, acc + 4);
            else
                hipLaunchKernelGGL(k_shj_step, dim3(G), dim3(SBLOCK), 0, st, dS, lo, hi, ts,
                                   capS - 1, acc + 3, (const uint64_t *)tr, capR - 1,
                                   (const uint64_t *)(acc + 2), 0, sum_part, scnt + k * G);
        }
    }
    hipLaunchKernelGGL(k_shj_rows, dim3((unsigned)steps), dim3(SBLOCK), 0, st,
                       (const uint64_t *)scnt, G, prog);
    hipLaunchKernelGGL(k_shj_finish, dim3(1), dim3(SBLOCK), 0, st, steps, G,
                       (const uint64_t *)sum_part, prog, acc);
    AJ_HIP(hipGetLastError());
    AJ_HIP(record_event(ctx->ev1, st));
    uint64_t host[6];
    AJ_HIP(hipMemcpyAsync(host, acc, sizeof(host), hipMemcpyDeviceToHost, st));
    if (progress)
        AJ_HIP(hipMemcpyAsync(progress, prog, 8 * steps, hipMemcpyDefault, st));
    AJ_HIP(hipStreamSynchronize(st));
    const float ms = elapsed_ms(ctx->ev0, ctx->ev1);
    out->count = host[0];
    out->checksum = host[1];
    out->npairs = mode ? host[4] + host[5] : 0;
    out->nrecords_r = mode == 3 ? host[5] : 0;
    out->device_ms = ms;
    return AJ_OK;
}

}  // namespace aj
