// staged.hip -- two-pass PRO partitioning streamed through a staging buffer
// the size of the memory-side cache, with output space reserved by atomics.
//
// The reference partitions in two passes (parallel_radix_partition for pass 1,
// parallel_radix_join.cpp:739-806, then radix_cluster per pass-1 partition,
// :573-614), each pass preceded by a histogram of its input that fixes every
// thread's output offsets (:762-796).  Here both relations are cut into
// super-chunks of `chunk` tuples, and each super-chunk goes through both
// passes before the next one starts:
//
//   k_stage1  pass 1 of one super-chunk: work-groups take runs of consecutive
//             tiles, counting-sort each tile by the pass-1 digit in LDS, and
//             reserve room for each digit's run with one atomic add on that
//             digit's staging cursor; the runs land in the staging regions
//             (set, d1), where set = work-group mod nset spreads the atomics of
//             one digit over nset cursors.  The staging buffer is reused by
//             every super-chunk, so its lines stay in the memory-side cache
//             between this kernel's writes and k_stage2's reads.
//   k_stage2  pass 2 of the same super-chunk: each staging region is split
//             into units; a unit's tiles are counting-sorted by the pass-2
//             digit and each run reserves room in its final partition (d1, d2)
//             with one atomic add on that partition's cursor.
//   k_stage_tasks  join task records straight from the final cursors.
//
// No histogram is read: partitions are given fixed capacities (the expected
// size plus max(1/8, 8 sigma) of it plus 64) and the cursors say how much of
// each was filled.  A run that does not fit its region is written to a trash
// area instead and sets the overflow flag; the caller then joins the input
// again with the histogram-based passes, so the result never depends on the
// capacities (they only decide which path computes it).  The order of the
// runs within a partition depends on the order of the atomics; the join
// results (count, checksum, pair multiset) do not.
//
// HBM traffic per tuple: 8 read (pass 1) + 8 written (pass 2); the staging
// writes and reads go to the memory-side cache as far as it holds them.
#include <algorithm>
#include <cmath>

#include "aj_internal.h"
#include "device_common.h"
#include "stage_tiles.h"
#include "tiles.h"

namespace aj {


struct Stage1Args {
    const uint64_t *in;
    uint64_t begin, end;    // rows of this super-chunk
    uint64_t tiles_per_wg;  // consecutive tiles per work-group
    int b1, nset;
    uint64_t *stage;
    uint32_t scap;          // tuples per staging region
    uint32_t *scur;         // nset * F1 cursors of this super-chunk
    uint32_t *flag;
};

// MODE bits: 1 = input loads nontemporal, 2 = staging stores nontemporal,
// 16 = keys replaced by fmix32(key) as they are loaded
template <int PBLOCK, int MODE>
__global__ __launch_bounds__(PBLOCK, 4) void k_stage1(Stage1Args a) {
    constexpr int PTILE = PBLOCK * PITEMS;
    extern __shared__ uint64_t smem[];
    __shared__ uint32_t s_len;
    const uint32_t F = 1u << a.b1;
    const uint64_t b = a.begin + (uint64_t)blockIdx.x * a.tiles_per_wg * PTILE;
    if (b >= a.end) return;
    const uint64_t e = min(b + a.tiles_per_wg * PTILE, a.end);
    uint64_t *buf = smem;                                                // PTILE + NSPARE
    uint32_t *cnt = reinterpret_cast<uint32_t *>(buf + PTILE + NSPARE);  // F + NSPARE
    uint32_t *dst = cnt + F + NSPARE;                                    // F
    for (uint32_t d = threadIdx.x; d < F + NSPARE; d += PBLOCK) cnt[d] = 0;
    __syncthreads();
    const uint32_t set = blockIdx.x % (uint32_t)a.nset;
    struct {  // one segment: this work-group's run of tiles
        StageSeg s;
        __device__ bool get(uint32_t i, StageSeg &o) const {
            o = s;
            return i == 0;
        }
    } src;
    src.s.in = a.in;
    src.s.begin = b;
    src.s.end = e;
    src.s.r.cur = a.scur + set * F;
    src.s.r.base = set * F * a.scap;
    src.s.r.stride = a.scap;
    src.s.r.cap = a.scap;
    src.s.r.trash = (uint32_t)a.nset * F * a.scap;
    src.s.r.flag = a.flag;
    stage_tiles<PBLOCK, (MODE & 1) != 0, (MODE & 2) != 0, (MODE & 16) != 0>(
        src, a.stage, 0, F, buf, cnt, dst, &s_len);
}

struct Stage2Args {
    const uint64_t *stage;
    uint32_t scap;
    const uint32_t *scur;  // cursors k_stage1 filled
    uint32_t *scur_next;   // the other super-chunk's cursors: zeroed here
    int b1, b2, nset;
    uint32_t units;        // work items per staging region
    uint32_t part;         // tuples per work item (0: the region split evenly over units)
    uint32_t per_wg;       // consecutive work items per work-group
    uint64_t *out;         // final partitions of this relation
    uint32_t fcap;
    uint32_t *fcur;        // F1 * F2 cursors of this relation
    uint32_t *flag;
};

// Work item k of pass 2 is part units - 1 - (k mod units) of staging region
// k / units: a region's parts are read last first.  The staging regions are
// filled in time order, so their last parts were written last and are the
// likeliest to be still in the memory-side cache when a super-chunk's regions
// come close to its size (2^25 8-byte tuples: 268 MB of staging; reading the
// regions backwards took pass 2 from 1.05 to 1.02-1.03 ms per 2^28 tuples,
// forwards in the same tile-sized parts 1.06).  Parts are whole tiles when the
// regions hold at least one tile per work-group of a resident round; else each
// region is split evenly over units parts.  A work-group takes per_wg
// consecutive items and runs through them as one tile pipeline (regions are
// about 4 tiles each at the default geometry, so a work-group per region would
// load its first tile with nothing to overlap).
struct Stage2Src {
    const Stage2Args *a;
    uint32_t k0, k1;  // this work-group's items
    __device__ bool get(uint32_t i, StageSeg &o) const {
        const uint32_t k = k0 + i;
        if (k >= k1) return false;
        const uint32_t F1 = 1u << a->b1, F2 = 1u << a->b2;
        const uint32_t region = k / a->units, j = a->units - 1 - k % a->units;
        const uint32_t d1 = region & (F1 - 1);
        const uint32_t n = min(a->scur[region], a->scap);
        // an even split point keeps the 16-byte loads aligned
        const uint32_t per = a->part ? a->part : ((n + a->units - 1) / a->units + 1) & ~1u;
        o.in = a->stage + (uint64_t)region * a->scap;
        o.begin = min(j * per, n);
        o.end = min(j * per + per, n);
        o.r.cur = a->fcur + d1 * F2;
        o.r.base = d1 * F2 * a->fcap;
        o.r.stride = a->fcap;
        o.r.cap = a->fcap;
        o.r.trash = F1 * F2 * a->fcap;
        o.r.flag = a->flag;
        return true;
    }
};

// MODE bits: 4 = staging loads nontemporal, 8 = final stores nontemporal
template <int PBLOCK, int MODE>
__global__ __launch_bounds__(PBLOCK, 4) void k_stage2(Stage2Args a) {
    constexpr int PTILE = PBLOCK * PITEMS;
    extern __shared__ uint64_t smem[];
    __shared__ uint32_t s_len;
    const uint32_t F2 = 1u << a.b2;
    const uint32_t nreg = (uint32_t)a.nset << a.b1;
    if (blockIdx.x == 0)
        for (uint32_t i = threadIdx.x; i < nreg; i += PBLOCK) a.scur_next[i] = 0;
    const uint32_t items = nreg * a.units;
    const uint32_t k0 = blockIdx.x * a.per_wg;
    if (k0 >= items) return;
    uint64_t *buf = smem;
    uint32_t *cnt = reinterpret_cast<uint32_t *>(buf + PTILE + NSPARE);
    uint32_t *dst = cnt + F2 + NSPARE;
    for (uint32_t d = threadIdx.x; d < F2 + NSPARE; d += PBLOCK) cnt[d] = 0;
    __syncthreads();
    Stage2Src src;
    src.a = &a;
    src.k0 = k0;
    src.k1 = min(k0 + a.per_wg, items);
    stage_tiles<PBLOCK, (MODE & 4) != 0, (MODE & 8) != 0, false>(src, a.out, a.b1, F2, buf, cnt,
                                                                 dst, &s_len);
}

// Join tasks of the staged layout: T tasks per final partition p (S slices of
// `split`), task p * T + j = (R rows [p fcapR, + nR_p), S rows [p fcapS + j split,
// + its slice)); a task with nothing to join has nR = nS = 0 and its S start
// at the partition start (k_join then loads nothing it uses, and every address
// it touches lies in the partition's region or the trash area after it).
__global__ void k_stage_tasks(const uint32_t *fcur, uint32_t F, uint32_t fcapR, uint32_t fcapS,
                              uint32_t split, uint32_t T, uint32_t sh, uint4 *task_info,
                              uint32_t *task_base, uint64_t max_parts) {
    const uint64_t total = (uint64_t)F * T;
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < total;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t p = (uint32_t)(t / T), j = (uint32_t)(t % T);
        uint32_t nr = min(fcur[p], fcapR);
        const uint32_t ns_all = min(fcur[F + p], fcapS);
        const uint32_t s_off = j * split;
        uint32_t ns = ns_all > s_off ? min(split, ns_all - s_off) : 0u;
        if (!ns || !nr) ns = nr = 0;
        const uint64_t r0 = (uint64_t)p * fcapR;
        const uint64_t s0 = (uint64_t)p * fcapS + (ns ? s_off : 0u);
        task_info[2 * t] = make_uint4((uint32_t)r0, (uint32_t)(r0 >> 32), nr, 0u);
        task_info[2 * t + 1] = make_uint4((uint32_t)s0, (uint32_t)(s0 >> 32), ns, sh);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        task_base[0] = 0;
        task_base[max_parts] = (uint32_t)total;
    }
}

// Zeroes up to three counter ranges in one launch (the staging cursors, the
// final cursors and the overflow flag of a staged join).
__global__ void k_stage_zero(uint32_t *p0, uint64_t n0, uint32_t *p1, uint64_t n1, uint32_t *p2,
                             uint64_t n2) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t i0 = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    for (uint64_t i = i0; i < n0; i += step) p0[i] = 0;
    for (uint64_t i = i0; i < n1; i += step) p1[i] = 0;
    for (uint64_t i = i0; i < n2; i += step) p2[i] = 0;
}

// ---------------------------------------------------------------- host side
static uint32_t region_cap(double expect) {
    const double slack = std::max(expect / 8.0, 8.0 * std::sqrt(expect));
    const uint64_t c = (uint64_t)std::ceil(expect + slack) + 64;
    return (uint32_t)((c + 63) / 64 * 64);
}

int staged_geometry(StagedGeom *g, uint64_t nR, uint64_t nS, int b1, int b2, uint64_t chunk,
                    int nset, int block, int num_cu, int block1) {
    if (!block1) block1 = block;
    if (b1 < 1 || b2 < 1 || (1 << b1) > STAGE_MAXF || (1 << b2) > STAGE_MAXF ||
        (block != 512 && block != 1024) || (block1 != 512 && block1 != 1024) || nset < 1 ||
        nset > 64 || nR == 0 || nS == 0) {
        set_error(AJ_ERR_ARG,
                  "staged partitioning: bad geometry (%d + %d bits, blocks %d + %d, %d sets)", b1,
                  b2, block1, block, nset);
        return AJ_ERR_ARG;
    }
    // pass 1 tiles of ptile1 tuples, pass 2 tiles of ptile; both trash areas
    // hold the larger tile
    const uint64_t ptile1 = (uint64_t)block1 * PITEMS, ptile = (uint64_t)block * PITEMS;
    const uint64_t pmax = std::max(ptile1, ptile);
    g->b1 = b1;
    g->b2 = b2;
    g->nset = nset;
    g->block = block;
    g->block1 = block1;
    const uint64_t nmax = std::max(nR, nS);
    if (!chunk) chunk = 1ull << 25;
    chunk = std::max<uint64_t>(pmax, chunk / pmax * pmax);
    g->chunk = std::min<uint64_t>(chunk, (nmax + pmax - 1) / pmax * pmax);
    const uint64_t tiles = g->chunk / ptile1;
    const int per_cu1 = block1 == 1024 ? 1 : 2, per_cu = block == 1024 ? 1 : 2;
    const uint64_t g1 = std::min<uint64_t>(tiles, (uint64_t)num_cu * per_cu1);
    g->tiles_per_wg = (tiles + g1 - 1) / g1;
    g->grid1 = (uint32_t)((tiles + g->tiles_per_wg - 1) / g->tiles_per_wg);
    // the largest share of a super-chunk one cursor set can receive
    const uint64_t wg_per_set = (g->grid1 + nset - 1) / nset;
    const double set_tuples =
        (double)std::min<uint64_t>(g->chunk, wg_per_set * g->tiles_per_wg * ptile1);
    const uint32_t F1 = 1u << b1, F = F1 << b2;
    g->scap = region_cap(set_tuples / F1);
    g->fcap[0] = region_cap((double)nR / F);
    g->fcap[1] = region_cap((double)nS / F);
    const uint32_t nreg = (uint32_t)nset * F1;
    // pass 2: at least one work item per work-group of one resident round, and
    // each work-group takes its share of the items in one pipeline
    const uint64_t g2 = (uint64_t)num_cu * per_cu;
    const uint64_t spread = std::max<uint64_t>(1, (g2 + nreg - 1) / nreg);
    const uint64_t region_tiles = (g->scap + ptile - 1) / ptile;
    g->part = region_tiles >= spread ? (uint32_t)ptile : 0u;
    g->units = (uint32_t)std::max(spread, region_tiles);
    const uint64_t items = (uint64_t)nreg * g->units;
    g->per_wg = (uint32_t)((items + g2 - 1) / g2);
    g->grid2 = (uint32_t)((items + g->per_wg - 1) / g->per_wg);
    g->stage_tuples = (uint64_t)nreg * g->scap + pmax;
    for (int r = 0; r < 2; ++r) g->out_tuples[r] = (uint64_t)F * g->fcap[r] + pmax;
    if (g->stage_tuples >= (1ull << 32) || g->out_tuples[0] >= (1ull << 32) ||
        g->out_tuples[1] >= (1ull << 32)) {
        set_error(AJ_ERR_ARG, "staged partitioning: regions exceed 32-bit offsets");
        return AJ_ERR_ARG;
    }
    return AJ_OK;
}

template <int PBLOCK1, int PBLOCK, int MODE>
static int launch_staged(const StagedDesc &d, hipStream_t st, std::vector<hipEvent_t> *evs) {
    const StagedGeom &g = d.g;
    const uint32_t F1 = 1u << g.b1, F2 = 1u << g.b2;
    const size_t lds1 = stage_lds_bytes(PBLOCK1, F1), lds2 = stage_lds_bytes(PBLOCK, F2);
    constexpr int MODE2 = MODE & 15;  // pass 2 reads keys as pass 1 wrote them
    AJ_TRY(lds_limit_set((const void *)k_stage1<PBLOCK1, MODE>, lds1));
    AJ_TRY(lds_limit_set((const void *)k_stage2<PBLOCK, MODE2>, lds2));
    const uint32_t nreg = (uint32_t)g.nset * F1;
    int k = 0, ev = 0;
    for (int rel = 0; rel < 2; ++rel) {
        for (uint64_t b = 0; b < d.n[rel]; b += g.chunk, ++k) {
            uint32_t *cur = d.scur + (uint64_t)(k & 1) * nreg;
            uint32_t *next = d.scur + (uint64_t)((k + 1) & 1) * nreg;
            Stage1Args a1;
            a1.in = d.in[rel];
            a1.begin = b;
            a1.end = std::min(b + g.chunk, d.n[rel]);
            a1.tiles_per_wg = g.tiles_per_wg;
            a1.b1 = g.b1;
            a1.nset = g.nset;
            a1.stage = d.stage;
            a1.scap = g.scap;
            a1.scur = cur;
            a1.flag = d.flag;
            const uint64_t tiles =
                (a1.end - a1.begin + PBLOCK1 * PITEMS - 1) / (PBLOCK1 * PITEMS);
            const unsigned grid1 = (unsigned)((tiles + g.tiles_per_wg - 1) / g.tiles_per_wg);
            if (evs) (void)hipEventRecord((*evs)[ev++], st);
            hipLaunchKernelGGL((k_stage1<PBLOCK1, MODE>), dim3(grid1), dim3(PBLOCK1), lds1, st,
                               a1);
            if (evs) (void)hipEventRecord((*evs)[ev++], st);
            Stage2Args a2;
            a2.stage = d.stage;
            a2.scap = g.scap;
            a2.scur = cur;
            a2.scur_next = next;
            a2.b1 = g.b1;
            a2.b2 = g.b2;
            a2.nset = g.nset;
            a2.units = g.units;
            a2.part = g.part;
            a2.per_wg = g.per_wg;
            a2.out = d.out[rel];
            a2.fcap = g.fcap[rel];
            a2.fcur = d.fcur + (uint64_t)rel * (F1 * F2);
            a2.flag = d.flag;
            hipLaunchKernelGGL((k_stage2<PBLOCK, MODE2>), dim3(g.grid2), dim3(PBLOCK), lds2, st,
                               a2);
        }
    }
    if (evs) (void)hipEventRecord((*evs)[ev++], st);
    AJ_HIP(hipGetLastError());
    return AJ_OK;
}

int staged_events(const StagedDesc &d) {
    int k = 0;
    for (int rel = 0; rel < 2; ++rel) k += (int)((d.n[rel] + d.g.chunk - 1) / d.g.chunk);
    return 2 * k + 1;
}

int run_staged(const StagedDesc &d, hipStream_t st, std::vector<hipEvent_t> *evs) {
    const StagedGeom &g = d.g;
    const uint32_t F1 = 1u << g.b1, F = F1 << g.b2;
    if (evs && (int)evs->size() < staged_events(d)) {
        set_error(AJ_ERR_ARG, "staged partitioning: too few timing events");
        return AJ_ERR_ARG;
    }
    // both cursor sets of the staging regions, the final cursors and the flag
    const uint64_t nz = 2ull * g.nset * F1;
    hipLaunchKernelGGL(k_stage_zero, dim3(64), dim3(256), 0, st, d.scur, nz, d.fcur, 2ull * F,
                       d.flag, (uint64_t)1);
    switch ((g.block1 * 2 + g.block) * 32 + g.cache_mode) {
#define AJ_STAGED_CASE(B1, B, M) \
    case (B1 * 2 + B) * 32 + M: return launch_staged<B1, B, M>(d, st, evs);
        AJ_STAGED_CASE(1024, 1024, 0) AJ_STAGED_CASE(1024, 1024, 1) AJ_STAGED_CASE(1024, 1024, 5)
        AJ_STAGED_CASE(1024, 1024, 9) AJ_STAGED_CASE(1024, 1024, 13)
        AJ_STAGED_CASE(1024, 1024, 15) AJ_STAGED_CASE(512, 512, 1) AJ_STAGED_CASE(512, 512, 5)
        AJ_STAGED_CASE(512, 512, 9) AJ_STAGED_CASE(1024, 1024, 25) AJ_STAGED_CASE(512, 1024, 9)
#undef AJ_STAGED_CASE
        default:
            set_error(AJ_ERR_ARG,
                      "staged partitioning: no kernels for blocks %d + %d, cache mode %d",
                      g.block1, g.block, g.cache_mode);
            return AJ_ERR_ARG;
    }
}

int staged_tasks(const StagedDesc &d, uint32_t split, uint4 *task_info, uint32_t *task_base,
                 uint64_t max_parts, uint64_t max_tasks, hipStream_t st) {
    const StagedGeom &g = d.g;
    const uint32_t F = 1u << (g.b1 + g.b2);
    const uint32_t T = (g.fcap[1] + split - 1) / split;
    if ((uint64_t)F * T > max_tasks || max_parts != F) {
        set_error(AJ_ERR_ARG, "staged partitioning: task workspace too small");
        return AJ_ERR_ARG;
    }
    const unsigned grid = (unsigned)std::min<uint64_t>(((uint64_t)F * T + 255) / 256, 4096);
    hipLaunchKernelGGL(k_stage_tasks, dim3(grid), dim3(256), 0, st, d.fcur, F, g.fcap[0],
                       g.fcap[1], split, T, (uint32_t)(g.b1 + g.b2), task_info, task_base,
                       max_parts);
    AJ_HIP(hipGetLastError());
    return AJ_OK;
}

uint64_t staged_max_tasks(const StagedGeom &g, uint32_t split) {
    const uint64_t F = 1ull << (g.b1 + g.b2);
    return F * ((g.fcap[1] + split - 1) / split);
}

}  // namespace aj
