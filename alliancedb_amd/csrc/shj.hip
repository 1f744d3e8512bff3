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
// the NPO layout (16-byte slots, one per distinct key, holding its first tuple
// and the count and payload sum of its further tuples; capacity the next power
// of two >= 2 max(|R|, |S|), murmur3-finalised home slot, all-ones empty marker
// with all-ones tuples counted on the side), the two sides' slots interleaved
// in 32-byte pairs when a step gives each lane at most one tuple (a tuple's
// probe and its insert are then one random line, not two), else two separate
// tables.  A counting probe takes its key's count and payload sum from the slot in
// O(1), the writing modes walk the key's chain, and an insert never walks past
// the copies of a key.  Per-step cumulative match counts give the progressive
// result curve.
//
// Each step is bound by random read-modify-writes: a probe read and a 64-bit
// CAS per tuple, issued together (they are independent), plus a 64-bit add per
// duplicate key.
//
// Bytes per tuple: read 8 + own slot write 16 + probe slot reads (>= 16).
#include "aj_internal.h"
#include "device_common.h"

namespace aj {

constexpr int SBLOCK = 256;

// One side's table: 16-byte slots as NPO's (word 0 the first tuple of a
// distinct key, all ones = empty; word 1 the aggregate of the key's further
// tuples, payload sum mod 2^32 << 32 | count), either interleaved with the
// other side's (slot h of R and slot h of S form one 32-byte pair, R's at
// tab + 4h, S's two words later: a tuple's probe of the other side and its
// claim in its own side touch the same line) or in a table of its own.  Also
// the carries out of that sum per
// slot (an S key's full payload sum is needed when R tuples probe it: the S
// payloads fill the low word of a pair, so their sum carries into the high
// word), the side's tuples, its all-ones tuple count, and -- in the writing
// modes -- the chains of further tuples (head per slot, side index + 1; links in
// next[]).
struct ShjSide {
    uint64_t *tab;
    uint64_t mask;
    uint32_t sh;     // slot h at tab + (h << sh) words: 2 interleaved, 1 separate tables
    uint32_t *carry; // per slot: carries out of the aggregate's 32-bit payload sum
    uint32_t *head;  // null: no chains (counting mode)
    uint32_t *next;
    const uint64_t *in;
    uint64_t *sent;
};

typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));

// k's slot in a side's table, or ~0 if k has none; first / agg receive the
// slot's words.  shj_find_from continues from slot h whose words w were
// already loaded.
__device__ __forceinline__ uint64_t shj_find_from(uint32_t k, const ShjSide &o, uint64_t h,
                                                  u64x2 w, uint64_t &first, uint64_t &agg) {
    for (;;) {
        if (w.x == EMPTY_SLOT) return ~0ull;
        if (key_of(w.x) == k) {
            first = w.x;
            agg = w.y;
            return h;
        }
        h = (h + 1) & o.mask;
        w = *reinterpret_cast<const u64x2 *>(o.tab + (h << o.sh));
    }
}

__device__ __forceinline__ uint64_t shj_find(uint32_t k, const ShjSide &o, uint64_t &first,
                                             uint64_t &agg) {
    const uint64_t h = fmix32(k) & o.mask;
    return shj_find_from(k, o, h, *reinterpret_cast<const u64x2 *>(o.tab + (h << o.sh)), first,
                         agg);
}

// Count and checksum of v's matches in the other side's table in O(1): the
// key's count is 1 + the further count, its payload sum the first payload plus
// the further sum (mod 2^32).
__device__ __forceinline__ void shj_count(uint64_t v, const ShjSide &o, int is_r, uint64_t h0,
                                          u64x2 w0, uint64_t &cnt, uint64_t &sum) {
    uint64_t first, agg;
    const uint64_t h = shj_find_from(key_of(v), o, h0, w0, first, agg);
    if (h == ~0ull) return;
    const uint32_t more = (uint32_t)agg;
    const uint64_t c = 1ull + more;
    cnt += c;
    if (is_r) {  // S payloads fill the low word of a pair: their full sum
        const uint64_t carries = more ? o.carry[h] : 0u;
        const uint64_t osum = payload_of(first) + ((carries << 32) | (agg >> 32));
        sum += c * ((uint64_t)payload_of(v) << 32) + osum;
    } else {  // R payloads fill the high word: their sum mod 2^32 suffices
        const uint32_t osum = payload_of(first) + (uint32_t)(agg >> 32);
        sum += ((uint64_t)osum << 32) + c * payload_of(v);
    }
}

// f(t) for every tuple t of the other side with v's key (the first, then the
// chain of further ones); the writing modes only.
template <class F>
__device__ __forceinline__ void shj_matches(uint32_t k, const ShjSide &o, F &&f) {
    uint64_t first, agg;
    const uint64_t h = shj_find(k, o, first, agg);
    if (h == ~0ull) return;
    f(first);
    for (uint32_t j = o.head[h]; j; j = o.next[j - 1]) f(o.in[j - 1]);
}

// Insert tuple i (value v, key k) into its side's table: claim k's slot with a
// 64-bit CAS, or add v to the slot's aggregate (one 64-bit add) and, in the
// writing modes, chain i behind it with one exchange.  shj_insert_from
// continues after a first CAS at slot h that returned old.
__device__ __forceinline__ void shj_insert_from(uint64_t i, uint64_t v, uint32_t k,
                                                const ShjSide &s, uint64_t h, uint64_t old) {
    // terminates: fewer distinct keys than slots
    for (;;) {
        if (old == EMPTY_SLOT) return;
        if (key_of(old) == k) {
            const uint32_t p = payload_of(v);
            const uint64_t prev = atomicAdd((unsigned long long *)&s.tab[(h << s.sh) + 1],
                                            (unsigned long long)(((uint64_t)p << 32) | 1u));
            // this add carried out of the 32-bit sum (the count half never does:
            // fewer than 2^32 tuples)
            if ((uint32_t)(prev >> 32) + p < p) atomicAdd(&s.carry[h], 1u);
            if (s.head) s.next[i] = atomicExch(&s.head[h], (uint32_t)(i + 1));
            return;
        }
        h = (h + 1) & s.mask;
        old = atomicCAS((unsigned long long *)&s.tab[h << s.sh], (unsigned long long)EMPTY_SLOT,
                        (unsigned long long)v);
    }
}

__device__ __forceinline__ uint64_t shj_claim(uint64_t v, const ShjSide &s, uint64_t h) {
    return atomicCAS((unsigned long long *)&s.tab[h << s.sh], (unsigned long long)EMPTY_SLOT,
                     (unsigned long long)v);
}

__device__ __forceinline__ void shj_insert(uint64_t i, uint64_t v, uint32_t k, const ShjSide &s) {
    const uint64_t h = fmix32(k) & s.mask;
    shj_insert_from(i, v, k, s, h, shj_claim(v, s, h));
}

// Insert tuple i of this side into its table (the all-ones tuple is counted in
// *own.sent instead) and count / checksum its matches in the other side's
// table (other.sent all-ones tuples there).  Work-group b adds its matches to
// step_cnt[b] (this step's row) and its checksum to sum_part[b]: every launch
// has the same grid and launches are stream-ordered, so each slot has one
// writer at a time and no atomics contend.
__global__ __launch_bounds__(SBLOCK) void k_shj_step(uint64_t lo, uint64_t hi, ShjSide own,
                                                     ShjSide other, int is_r, uint64_t *sum_part,
                                                     uint64_t *step_cnt) {
    __shared__ uint64_t red[SBLOCK / 64];
    const uint64_t nsent = *other.sent;  // not written during this kernel
    uint64_t cnt = 0, sum = 0, sent = 0;
    for (uint64_t i = lo + blockIdx.x * (uint64_t)SBLOCK + threadIdx.x; i < hi;
         i += (uint64_t)gridDim.x * SBLOCK) {
        const uint64_t v = own.in[i];
        const uint32_t k = key_of(v);
        const uint64_t h = fmix32(k) & own.mask;  // both sides have the same capacity
        // the probe's first slot (the other table does not change in this
        // kernel) and the insert's first claim are independent: both in flight
        // before either is looked at
        const u64x2 w0 = *reinterpret_cast<const u64x2 *>(other.tab + (h << other.sh));
        const bool ins = v != EMPTY_SLOT;
        const uint64_t old = ins ? shj_claim(v, own, h) : 0;
        shj_count(v, other, is_r, h, w0, cnt, sum);
        if (k == 0xFFFFFFFFu && nsent) {  // all-ones tuples of the other side
            cnt += nsent;
            const uint64_t pv = is_r ? pair_value(v, EMPTY_SLOT) : pair_value(EMPTY_SLOT, v);
            sum += nsent * pv;
        }
        if (!ins) {
            ++sent;
            continue;
        }
        shj_insert_from(i, v, k, own, h, old);
    }
    sent = block_sum<SBLOCK>(sent, red);
    cnt = block_sum<SBLOCK>(cnt, red);
    sum = block_sum<SBLOCK>(sum, red);
    if (threadIdx.x == 0) {
        if (sent) atomicAdd((unsigned long long *)own.sent, (unsigned long long)sent);
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
// matches, reserves one range with a single atomic on *cursor, then finds its
// matches again to write; positions at or past cap are counted but not written
// (the caller reports the shortfall).  In mode 3 an R tuple's records fill out
// from the end backwards (its cursor counts them), so the caller can tell the
// two probe sides apart.  The loop runs a work-group-uniform number of rounds
// so the wave operations see every lane.
__global__ __launch_bounds__(SBLOCK) void k_shj_step_out(uint64_t lo, uint64_t hi, ShjSide own,
                                                         ShjSide other, int is_r,
                                                         uint64_t *sum_part, uint64_t *step_cnt,
                                                         int mode, uint64_t *out, uint64_t cap,
                                                         uint64_t *cursor) {
    __shared__ uint64_t red[SBLOCK / 64];
    const uint64_t nsent = *other.sent;  // not written during this kernel
    uint64_t cnt = 0, sum = 0, sent = 0;
    for (uint64_t base = lo + blockIdx.x * (uint64_t)SBLOCK; base < hi;
         base += (uint64_t)gridDim.x * SBLOCK) {
        const uint64_t i = base + threadIdx.x;
        const bool valid = i < hi;
        const uint64_t v = valid ? own.in[i] : 0;
        const uint32_t k = key_of(v);
        uint64_t c = 0;
        if (valid) {
            uint64_t first, agg;
            if (shj_find(k, other, first, agg) != ~0ull) c = 1ull + (uint32_t)agg;
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
            shj_matches(k, other, emit);
            if (k == 0xFFFFFFFFu)
                for (uint64_t n = 0; n < nsent; ++n) emit(EMPTY_SLOT);
            cnt += c;
        }
        if (!valid) continue;
        if (v == EMPTY_SLOT) {
            ++sent;
            continue;
        }
        shj_insert(i, v, k, own);
    }
    sent = block_sum<SBLOCK>(sent, red);
    cnt = block_sum<SBLOCK>(cnt, red);
    sum = block_sum<SBLOCK>(sum, red);
    if (threadIdx.x == 0) {
        if (sent) atomicAdd((unsigned long long *)own.sent, (unsigned long long)sent);
        if (cnt) {
            step_cnt[blockIdx.x] += cnt;
            sum_part[blockIdx.x] += sum;
        }
    }
}

// Empty slots of a side's table: word 0 all ones, aggregate 0.
__global__ __launch_bounds__(SBLOCK) void k_shj_fill(uint4 *t, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        t[i] = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u);
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

uint64_t table_cap(uint64_t n) {  // load at most 1/2: one slot per distinct key
    uint64_t c = 2;
    while (c < 2 * n) c <<= 1;
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
    // one capacity for both sides: their slots are interleaved in pairs
    const uint64_t capR = table_cap(nR > nS ? nR : nS), capS = capR;
    const uint64_t steps = shj_steps(nR, nS, batch);
    if (steps > (1ull << 20)) {
        set_error(AJ_ERR_ARG, "SHJ batch %llu gives %llu steps (limit 2^20)",
                  (unsigned long long)batch, (unsigned long long)steps);
        return AJ_ERR_ARG;
    }
    const unsigned G = shj_grid(batch, ctx->num_cu);  // one grid for every launch
    // Interleaved slot pairs when a step gives every lane at most one tuple: a
    // probe and its claim then meet in the memory-side cache.  With several
    // tuples per lane, lines read by probes are the lines other lanes' atomics
    // modify, and separate tables run faster (measured on MI355X at 2^27 per
    // side: 9.9 against 8.5 G tuples/s at 2^19 per step, 9.0 against 10.9 at 2^23).
    const bool interleave = (uint64_t)G * SBLOCK >= batch;
    AJ_TRY(ctx->table.ensure(capR * 32));
    AJ_TRY(ctx->carry.ensure(capR * 4));
    AJ_TRY(ctx->carry2.ensure(capS * 4));
    if (mode) {  // chains of further tuples, to write every match
        AJ_TRY(ctx->chain.ensure(nR * 4 + 4));
        AJ_TRY(ctx->chain2.ensure(nS * 4 + 4));
        AJ_TRY(ctx->heads.ensure(capR * 4));
        AJ_TRY(ctx->heads2.ensure(capS * 4));
    }
    AJ_TRY(ctx->acc.ensure(8 * 8));
    AJ_TRY(ctx->steps.ensure(8 * (steps * G + G + steps)));
    // [0] count [1] sum [2] R sentinels [3] S sentinels [4] output cursor
    // [5] R-probe output cursor (timed records)
    uint64_t *acc = ctx->acc.as<uint64_t>();
    uint64_t *scnt = ctx->steps.as<uint64_t>();  // [steps][G] per-block step matches
    uint64_t *sum_part = scnt + steps * G;       // [G] per-block checksums
    uint64_t *prog = sum_part + G;               // [steps] cumulative matches
    const uint32_t sh = interleave ? 2 : 1;
    const ShjSide sideR{ctx->table.as<uint64_t>(), capR - 1, sh, ctx->carry.as<uint32_t>(),
                        mode ? ctx->heads.as<uint32_t>() : nullptr,
                        mode ? ctx->chain.as<uint32_t>() : nullptr, dR, acc + 2};
    const ShjSide sideS{ctx->table.as<uint64_t>() + (interleave ? 2 : 2 * capR), capS - 1, sh,
                        ctx->carry2.as<uint32_t>(),
                        mode ? ctx->heads2.as<uint32_t>() : nullptr,
                        mode ? ctx->chain2.as<uint32_t>() : nullptr, dS, acc + 3};
    AJ_HIP(record_event(ctx->ev0, st));
    AJ_TRY(fill_u64(acc, 8, 0, st));
    hipLaunchKernelGGL(k_shj_fill, dim3(shj_grid(2 * capR, 32)), dim3(SBLOCK), 0, st,
                       reinterpret_cast<uint4 *>(sideR.tab), 2 * capR);  // both sides' slots
    AJ_TRY(fill_u32(sideR.carry, capR, 0, st));
    AJ_TRY(fill_u32(sideS.carry, capS, 0, st));
    if (mode) {
        AJ_TRY(fill_u32(sideR.head, capR, 0, st));
        AJ_TRY(fill_u32(sideS.head, capS, 0, st));
    }
    AJ_TRY(fill_u64(scnt, steps * G + G, 0, st));
    for (uint64_t k = 0; k < steps; ++k) {
        const uint64_t lo = k * batch;
        if (lo < nR) {
            const uint64_t hi = lo + batch < nR ? lo + batch : nR;
            if (mode)
                hipLaunchKernelGGL(k_shj_step_out, dim3(G), dim3(SBLOCK), 0, st, lo, hi, sideR,
                                   sideS, 1, sum_part, scnt + k * G, mode, pairs, pcap,
                                   mode == 3 ? acc + 5 : acc + 4);
            else
                hipLaunchKernelGGL(k_shj_step, dim3(G), dim3(SBLOCK), 0, st, lo, hi, sideR, sideS,
                                   1, sum_part, scnt + k * G);
        }
        if (lo < nS) {
            const uint64_t hi = lo + batch < nS ? lo + batch : nS;
            if (mode)
                hipLaunchKernelGGL(k_shj_step_out, dim3(G), dim3(SBLOCK), 0, st, lo, hi, sideS,
                                   sideR, 0, sum_part, scnt + k * G, mode, pairs, pcap, acc + 4);
            else
                hipLaunchKernelGGL(k_shj_step, dim3(G), dim3(SBLOCK), 0, st, lo, hi, sideS, sideR,
                                   0, sum_part, scnt + k * G);
        }
    }
    hipLaunchKernelGGL(k_shj_rows, dim3((unsigned)steps), dim3(SBLOCK), 0, st,
                       (const uint64_t *)scnt, G, prog);
    hipLaunchKernelGGL(k_shj_finish, dim3(1), dim3(SBLOCK), 0, st, steps, G,
                       (const uint64_t *)sum_part, prog, acc);
    AJ_HIP(hipGetLastError());
    AJ_HIP(record_event(ctx->ev1, st));
    uint64_t local[HRES_SLOTS];
    uint64_t *host = result_slots(ctx, local);
    AJ_HIP(hipMemcpyAsync(host, acc, 6 * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
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
