// partition.hip -- segmented radix partitioning (PRO partition passes).
//
// Restates the reference's pass structure (parallel_radix_partition,
// parallel_radix_join.cpp:739-806: per-thread histogram -> offsets
// partition-major / thread-minor -> scatter) for CDNA4: a "thread" of the
// reference becomes a chunk handled by one work-group, and one launch handles
// many independent segments (the windows of a batch, or the pass-1 partitions
// during pass 2).
//
//   k_plan      per segment: chunk counts, histogram bases, output bases,
//               partition bases (one work-group, block scans)
//   k_hist      one WG per chunk: LDS histogram of the digit (LDS atomics;
//               the LDS serialises same-address lanes of one instruction, so
//               a skewed digit costs at most ~1 cycle per tuple) written
//               column-major: hist[base + d*C + c]
//   scan        one flat exclusive scan over all histogram entries gives every
//               (segment, digit, chunk) its output offset -- the
//               partition-major, chunk-minor order of the reference
//   k_scatter   one WG per chunk, tiles of 8192 tuples: rank by LDS atomic,
//               LDS counting sort of the tile, then each digit's run is
//               written with consecutive lanes on consecutive addresses
//               (k_scatter_pipe: the same with fewer barriers and the next
//               tile's loads in flight during the write, for passes of at
//               most 10 bits)
//   k_part_off  partition start offsets for the join
//
// Bytes per tuple: k_hist reads 8, k_scatter reads 8 and writes 8.

#include <cmath>

#include "aj_internal.h"
#include "device_common.h"
#include "stage_tiles.h"
#include "tiles.h"

namespace aj {

constexpr int PTILE_MAX = 512 * PITEMS;   // largest tile (512-lane work-groups)
constexpr int MAX_PASS_BITS = 13;

struct PassArgs {
    const void *in0, *in1;
    int stride;
    uint64_t *out0, *out1;
    const uint64_t *in_begin0, *in_begin1, *in_end0, *in_end1;
    const uint32_t *bits;
    int nseg;
    int shift;
    uint32_t chunk;
    uint64_t n_total0;
    uint32_t *hist;        // counts, then (after the scan) offsets
    uint32_t *hist_scan;   // scanned offsets
    uint64_t *cpre0, *cpre1;
    uint64_t *hbase0, *hbase1;
    uint64_t *obase0, *obase1;
    uint32_t *pbase;
    uint64_t *header;
    uint64_t *poff0, *poff1;
    uint64_t max_hist;
    uint64_t max_parts;
    // nullable: the histogram, scatter and offset kernels return at once while
    // *guard is 0 (a pass that runs only when an earlier attempt overflowed)
    const uint32_t *guard;
};

__device__ __forceinline__ bool guard_off(const uint32_t *guard) { return guard && !*guard; }

// ---------------------------------------------------------------- k_plan
constexpr int PLAN_BLOCK = 1024;

__global__ __launch_bounds__(PLAN_BLOCK) void k_plan(PassArgs a) {
    __shared__ uint64_t tmp[PLAN_BLOCK / 64];
    uint64_t hist_run = 0;
    for (int rel = 0; rel < 2; ++rel) {
        const uint64_t *ib = rel ? a.in_begin1 : a.in_begin0;
        const uint64_t *ie = rel ? a.in_end1 : a.in_end0;
        uint64_t *cpre = rel ? a.cpre1 : a.cpre0;
        uint64_t *hb = rel ? a.hbase1 : a.hbase0;
        uint64_t *ob = rel ? a.obase1 : a.obase0;
        uint64_t chunk_run = 0, out_run = 0, part_run = 0;
        const uint64_t hist_rel_start = hist_run;
        for (int base = 0; base < a.nseg; base += PLAN_BLOCK) {
            const int k = base + threadIdx.x;
            uint64_t n = 0, c = 0, h = 0, f = 0;
            if (k < a.nseg) {
                n = ie[k] - ib[k];
                c = (n + a.chunk - 1) / a.chunk;
                f = 1ull << a.bits[k];
                h = f * c;
            }
            uint64_t tc, th, tn, tf;
            uint64_t pc = block_exclusive_scan<PLAN_BLOCK>(c, tmp, &tc);
            uint64_t ph = block_exclusive_scan<PLAN_BLOCK>(h, tmp, &th);
            uint64_t pn = block_exclusive_scan<PLAN_BLOCK>(n, tmp, &tn);
            uint64_t pf = block_exclusive_scan<PLAN_BLOCK>(f, tmp, &tf);
            if (k < a.nseg) {
                cpre[k] = chunk_run + pc;
                hb[k] = hist_run + ph;
                ob[k] = out_run + pn;
                if (rel == 0) a.pbase[k] = (uint32_t)(part_run + pf);
            }
            chunk_run += tc;
            hist_run += th;
            out_run += tn;
            part_run += tf;
        }
        if (threadIdx.x == 0) {
            cpre[a.nseg] = chunk_run;
            ob[a.nseg] = out_run;
            a.header[rel] = chunk_run;
            a.header[2 + rel] = hist_run - hist_rel_start;
            if (rel == 0) {
                a.pbase[a.nseg] = (uint32_t)part_run;
                a.header[4] = part_run;
            }
        }
        __syncthreads();
    }
}

// k_plan for one segment of known sizes nR, nS (the host computed the chunk
// counts c0, c1 and the digit count F): one lane writes the same plan.
__global__ void k_plan_single(PassArgs a, uint64_t nR, uint64_t nS, uint64_t c0, uint64_t c1,
                              uint32_t F) {
    if (threadIdx.x != 0) return;
    a.cpre0[0] = 0;
    a.cpre0[1] = c0;
    a.cpre1[0] = 0;
    a.cpre1[1] = c1;
    a.hbase0[0] = 0;
    a.hbase1[0] = (uint64_t)F * c0;
    a.obase0[0] = 0;
    a.obase0[1] = nR;
    a.obase1[0] = 0;
    a.obase1[1] = nS;
    a.pbase[0] = 0;
    a.pbase[1] = F;
    a.header[0] = c0;
    a.header[1] = c1;
    a.header[2] = (uint64_t)F * c0;
    a.header[3] = (uint64_t)F * c1;
    a.header[4] = F;
}

// Zero the histogram tail beyond the entries this launch uses (so the flat
// scan over the host-side bound stays correct).
__global__ void k_zero_tail(uint32_t *hist, const uint64_t *header, uint64_t max_hist,
                            const uint32_t *guard) {
    if (guard_off(guard)) return;
    const uint64_t used = header[2] + header[3];
    for (uint64_t i = used + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < max_hist;
         i += (uint64_t)gridDim.x * blockDim.x)
        hist[i] = 0;
}

struct ChunkLoc {
    int rel;
    int seg;
    uint64_t c;       // chunk index within the segment
    uint64_t nchunk;  // chunks of the segment
    uint64_t begin, end;
};

__device__ __forceinline__ bool locate_chunk(const PassArgs &a, uint64_t g, ChunkLoc &L) {
    const uint64_t tot0 = a.header[0], tot1 = a.header[1];
    if (g >= tot0 + tot1) return false;
    L.rel = g >= tot0;
    const uint64_t local = L.rel ? g - tot0 : g;
    const uint64_t *cpre = L.rel ? a.cpre1 : a.cpre0;
    L.seg = (int)last_le<uint64_t>(cpre, (uint64_t)a.nseg + 1, local);
    L.c = local - cpre[L.seg];
    L.nchunk = cpre[L.seg + 1] - cpre[L.seg];
    const uint64_t b = (L.rel ? a.in_begin1 : a.in_begin0)[L.seg];
    const uint64_t e = (L.rel ? a.in_end1 : a.in_end0)[L.seg];
    L.begin = b + L.c * a.chunk;
    L.end = min(L.begin + a.chunk, e);
    return true;
}

// ---------------------------------------------------------------- k_hist
template <int STRIDE, int PBLOCK>
__global__ __launch_bounds__(PBLOCK) void k_hist(PassArgs a) {
    if (guard_off(a.guard)) return;
    constexpr int PTILE = PBLOCK * PITEMS;
    extern __shared__ uint32_t cnt[];
    ChunkLoc L;
    if (!locate_chunk(a, blockIdx.x, L)) return;
    const int bits = (int)a.bits[L.seg];
    const uint32_t F = 1u << bits, mask = F - 1;
    const void *in = L.rel ? a.in1 : a.in0;
    for (uint32_t i = threadIdx.x; i < F; i += PBLOCK) cnt[i] = 0;
    __syncthreads();
    for (uint64_t base = tile_start(L.begin, STRIDE); base < L.end; base += PTILE) {
        uint64_t t[PITEMS];
        bool v[PITEMS];
        load_tile<PBLOCK>(in, STRIDE, base, L.begin, L.end, t, v);
#pragma unroll
        for (int j = 0; j < PITEMS; ++j) {
            const uint32_t d = (key_of(t[j]) >> a.shift) & mask;
            if (v[j]) atomicAdd(&cnt[d], 1u);
        }
    }
    __syncthreads();
    const uint64_t hb = (L.rel ? a.hbase1 : a.hbase0)[L.seg];
    for (uint32_t d = threadIdx.x; d < F; d += PBLOCK) a.hist[hb + d * L.nchunk + L.c] = cnt[d];
}

// ---------------------------------------------------------------- k_scatter
template <int STRIDE, int PBLOCK>
__global__ __launch_bounds__(PBLOCK) void k_scatter(PassArgs a) {
    if (guard_off(a.guard)) return;
    constexpr int PTILE = PBLOCK * PITEMS;
    extern __shared__ uint64_t smem[];
    __shared__ uint32_t scan_tmp[PBLOCK / 64];
    ChunkLoc L;
    if (!locate_chunk(a, blockIdx.x, L)) return;
    const int bits = (int)a.bits[L.seg];
    const uint32_t F = 1u << bits, mask = F - 1;
    uint64_t *buf = smem;                                      // PTILE tuples
    uint32_t *cursor = reinterpret_cast<uint32_t *>(buf + PTILE);  // F
    uint32_t *lofs = cursor + F;                               // F + 1
    const void *in = L.rel ? a.in1 : a.in0;
    uint64_t *out = L.rel ? a.out1 : a.out0;
    // S offsets follow every R tuple that lies in a segment
    const uint32_t rel_base = L.rel ? (uint32_t)a.obase0[a.nseg] : 0u;
    const uint64_t hb = (L.rel ? a.hbase1 : a.hbase0)[L.seg];
    for (uint32_t d = threadIdx.x; d < F; d += PBLOCK)
        cursor[d] = a.hist_scan[hb + d * L.nchunk + L.c] - rel_base;

    for (uint64_t base = tile_start(L.begin, STRIDE); base < L.end; base += PTILE) {
        for (uint32_t d = threadIdx.x; d <= F; d += PBLOCK) lofs[d] = 0;
        uint64_t t[PITEMS];
        bool v[PITEMS];
        load_tile<PBLOCK>(in, STRIDE, base, L.begin, L.end, t, v);
        __syncthreads();
        uint32_t rank[PITEMS];
#pragma unroll
        for (int j = 0; j < PITEMS; ++j) {
            const uint32_t d = (key_of(t[j]) >> a.shift) & mask;
            rank[j] = v[j] ? atomicAdd(&lofs[d], 1u) : 0u;
        }
        __syncthreads();
        const uint32_t tile_len = lds_exclusive_scan<PBLOCK>(lofs, F + 1, scan_tmp);
#pragma unroll
        for (int j = 0; j < PITEMS; ++j) {
            if (v[j]) {
                const uint32_t d = (key_of(t[j]) >> a.shift) & mask;
                buf[lofs[d] + rank[j]] = t[j];
            }
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < tile_len; i += PBLOCK) {
            const uint64_t u = buf[i];
            const uint32_t d = (key_of(u) >> a.shift) & mask;
            out[(uint64_t)cursor[d] + (i - lofs[d])] = u;
        }
        __syncthreads();
        for (uint32_t d = threadIdx.x; d < F; d += PBLOCK) cursor[d] += lofs[d + 1] - lofs[d];
        __syncthreads();
    }
}

// ---------------------------------------------------------------- k_scatter_pipe
// The scatter for passes of at most PIPE_MAX_BITS radix bits, with four
// barriers per tile:
//   ranks     LDS atomic per tuple on its digit's counter
//   scan      wave 0 alone scans the F counters (F/64 per lane) into the tile
//             starts, the per-digit write bases (global cursor - tile start)
//             and the advanced cursors
//   place     tuples into LDS at start + rank (digit order)
//   write     sorted tuple i goes to out[base[digit] + i]: consecutive lanes on
//             consecutive addresses within a digit's run; the next tile's loads
//             are issued before this step, so they are in flight meanwhile
constexpr int PIPE_MAX_BITS = 10;
// up to PIPE_WIDE_BITS with the block-wide scan (4 digits per lane)
constexpr int PIPE_WIDE_BITS = 11;
// LDS of the pipelined scatter: PTILE + NSPARE tuples, then F + NSPARE
// counters, F write bases and F cursors.
constexpr size_t pipe_lds_bytes(int pblock, uint32_t F) {
    return sizeof(uint64_t) * ((size_t)pblock * PITEMS + NSPARE) +
           sizeof(uint32_t) * (3 * (size_t)F + NSPARE);
}

// Tile loop of the pipelined scatter over input tuples [begin, end): digit
// (key >> shift) & (F-1); `cur` holds the F output cursors and `cnt` (F
// counters, then NSPARE spares) is zero (both visible to the whole
// work-group) on entry; `buf` holds PTILE tuples, then NSPARE spare slots.
// Every phase is branch-free over a lane's items: a tuple outside the range
// counts on its lane's spare counter and is placed in its lane's spare slot,
// so the LDS reads of a phase are issued in batches (WB items) and waited for
// once.  WIDE (more than 2^PIPE_MAX_BITS digits): every lane scans 4 digits
// and a block-wide scan orders them, instead of wave 0 scanning all of them.
template <int STRIDE, int PBLOCK, bool WIDE = false>
__device__ __forceinline__ void scatter_tiles(const void *in, uint64_t begin, uint64_t end,
                                              uint64_t *out, int shift, uint32_t F,
                                              uint64_t *buf, uint32_t *cnt, uint32_t *dst,
                                              uint32_t *cur, uint32_t *s_len) {
    constexpr int PTILE = PBLOCK * PITEMS;
    constexpr int MAXPER = WIDE ? 4 : (1 << PIPE_MAX_BITS) / 64;
    static_assert(!WIDE || MAXPER * PBLOCK >= (1 << PIPE_WIDE_BITS), "4 digits per lane cover F");
    constexpr int WB = 8;  // items per batch of LDS reads in the place / write phases
    static_assert(PITEMS % WB == 0, "a tile's items split into whole batches");
    const uint32_t mask = F - 1;
    uint64_t base = tile_start(begin, STRIDE);
    if (base >= end) return;
    uint64_t t[PITEMS];
    bool v[PITEMS];
    load_tile<PBLOCK>(in, STRIDE, base, begin, end, t, v);
    for (;;) {
        // ranks within the tile (< PTILE <= 2^13) packed two per register
        uint32_t rank2[PITEMS / 2];
#pragma unroll
        for (int j = 0; j < PITEMS; ++j) {
            const uint32_t d = (key_of(t[j]) >> shift) & mask;
            const uint32_t r =
                atomicAdd(&cnt[v[j] ? d : F + (threadIdx.x & (NSPARE - 1))], 1u) & 0xffffu;
            if (j & 1)
                rank2[j / 2] |= r << 16;
            else
                rank2[j / 2] = r;
        }
        __syncthreads();
        if constexpr (WIDE) {
            __shared__ uint32_t wscan[PBLOCK / 64];
            const uint32_t b = threadIdx.x * MAXPER;
            uint32_t c[MAXPER];
            uint32_t sum = 0;
#pragma unroll
            for (int k = 0; k < MAXPER; ++k) {
                c[k] = b + k < F ? cnt[b + k] : 0u;
                sum += c[k];
            }
            uint32_t total;
            uint32_t pre = block_exclusive_scan<PBLOCK>(sum, wscan, &total);
#pragma unroll
            for (int k = 0; k < MAXPER; ++k) {
                if (b + k < F) {
                    cnt[b + k] = pre;
                    dst[b + k] = cur[b + k] - pre;
                    cur[b + k] += c[k];
                    pre += c[k];
                }
            }
            if (threadIdx.x == 0) *s_len = total;
        } else if (threadIdx.x < 64) {
            const uint32_t lane = threadIdx.x;
            const uint32_t per = (F + 63) >> 6;
            const uint32_t b = lane * per;
            uint32_t c[MAXPER];
            uint32_t sum = 0;
#pragma unroll
            for (int k = 0; k < MAXPER; ++k) {
                c[k] = ((uint32_t)k < per && b + k < F) ? cnt[b + k] : 0u;
                sum += c[k];
            }
            const uint32_t inc = wave_inclusive_scan(sum);
            uint32_t pre = inc - sum;
#pragma unroll
            for (int k = 0; k < MAXPER; ++k) {
                if ((uint32_t)k < per && b + k < F) {
                    cnt[b + k] = pre;
                    dst[b + k] = cur[b + k] - pre;
                    cur[b + k] += c[k];
                    pre += c[k];
                }
            }
            if (lane == 63) *s_len = inc;
        }
        __syncthreads();
#pragma unroll
        for (int j0 = 0; j0 < PITEMS; j0 += WB) {
            uint32_t st[WB];
#pragma unroll
            for (int j = 0; j < WB; ++j)
                st[j] = cnt[(key_of(t[j0 + j]) >> shift) & mask];
#pragma unroll
            for (int j = 0; j < WB; ++j)
                buf[v[j0 + j] ? st[j] + ((rank2[(j0 + j) / 2] >> (16 * ((j0 + j) & 1))) & 0xffffu)
                              : (uint32_t)PTILE + (threadIdx.x & (NSPARE - 1))] = t[j0 + j];
        }
        __syncthreads();
        const uint32_t len = *s_len;
        const uint64_t nbase = base + PTILE;
        const bool more = nbase < end;
        if (more) load_tile<PBLOCK>(in, STRIDE, nbase, begin, end, t, v);
#pragma unroll
        for (int k0 = 0; k0 < PITEMS; k0 += WB) {
            uint64_t u[WB];
            uint32_t o[WB];
#pragma unroll
            for (int k = 0; k < WB; ++k) {
                const uint32_t i = threadIdx.x + (uint32_t)(k0 + k) * PBLOCK;
                u[k] = buf[i < len ? i : (uint32_t)PTILE];
            }
#pragma unroll
            for (int k = 0; k < WB; ++k) o[k] = dst[(key_of(u[k]) >> shift) & mask];
#pragma unroll
            for (int k = 0; k < WB; ++k) {
                const uint32_t i = threadIdx.x + (uint32_t)(k0 + k) * PBLOCK;
                if (i < len) out[(uint64_t)(uint32_t)(o[k] + i)] = u[k];
            }
        }
        for (uint32_t d = threadIdx.x; d < F; d += PBLOCK) cnt[d] = 0;
        __syncthreads();
        if (!more) break;
        base = nbase;
    }
}

template <int STRIDE, int PBLOCK, bool WIDE = false>
__global__ __launch_bounds__(PBLOCK, 4) void k_scatter_pipe(PassArgs a) {
    if (guard_off(a.guard)) return;
    constexpr int PTILE = PBLOCK * PITEMS;
    extern __shared__ uint64_t smem[];
    __shared__ uint32_t s_len;
    ChunkLoc L;
    if (!locate_chunk(a, blockIdx.x, L)) return;
    const uint32_t F = 1u << a.bits[L.seg];
    uint64_t *buf = smem;                                    // PTILE + NSPARE tuples
    uint32_t *cnt = reinterpret_cast<uint32_t *>(buf + PTILE + NSPARE);  // F + NSPARE counters
    uint32_t *dst = cnt + F + NSPARE;                        // F: cursor - tile start
    uint32_t *cur = dst + F;                                         // F: global write cursors
    // S offsets follow every R tuple that lies in a segment
    const uint32_t rel_base = L.rel ? (uint32_t)a.obase0[a.nseg] : 0u;
    const uint64_t hb = (L.rel ? a.hbase1 : a.hbase0)[L.seg];
    for (uint32_t d = threadIdx.x; d < F; d += PBLOCK) {
        cur[d] = a.hist_scan[hb + d * L.nchunk + L.c] - rel_base;
        cnt[d] = 0;
    }
    if (threadIdx.x < NSPARE) cnt[F + threadIdx.x] = 0;
    __syncthreads();
    scatter_tiles<STRIDE, PBLOCK, WIDE>(L.rel ? a.in1 : a.in0, L.begin, L.end,
                                        L.rel ? a.out1 : a.out0, a.shift, F, buf, cnt, dst, cur,
                                        &s_len);
}

// ---------------------------------------------------------------- pass 2 from a 2-D histogram
// For the one-segment PRO with two passes of b1 and b2 bits (F1*F2 <= 2^15):
//   k_hist2d     pass-1 histogram kernel that also counts (d1, d2) pairs of
//                its chunk in LDS and stores them in H, laid out
//                [relation][d1][chunk][d2]; the 1-D pass-1 histogram is the
//                row sums
//   k_p2_units   pass-2 work units: pass-1 partition d1 of a relation is
//                split at pass-1 chunk boundaries into about size/unit units
//   k_h2_usum    per unit and d2: the count over the unit's chunks (G);
//                per unit its input span in the pass-1 output
//   k_h2_base    one work-group per pass-1 partition: the (d1, d2) totals
//                from its units' counts, their exclusive scan from the
//                pass-1 partition start (the final partition offsets), and
//                each unit's start cursor per d2 (written over G)
//   k_scatter_p2 one work-group per unit: its pass-1 chunks' runs of
//                partition d1 are contiguous in the pass-1 output; the
//                cursors start where k_h2_base put them, so no pass-2
//                histogram is read
// This removes the pass-2 histogram read (8 bytes per tuple).
struct P2Args {
    const uint64_t *in0, *in1;   // pass-1 output
    uint64_t *out0, *out1;       // pass-2 output
    int b1, b2;
    uint32_t *H;                 // [rel][d1][chunk][d2] counts
    uint32_t *G;                 // [unit][d2] counts over the unit's chunks, then start cursors
    uint64_t *span;              // [unit] {input begin | rel << 63, input end}
    uint64_t *poff2_0, *poff2_1; // F1*F2+1 final offsets per relation
    uint32_t *ubase;             // 2*F1+1 unit bases
    const uint64_t *poff1_0, *poff1_1;  // pass-1 partition offsets
    const uint32_t *hist_scan;   // pass-1 scanned histogram (column-major d1*C + c)
    const uint64_t *hbase0, *hbase1;
    const uint64_t *header;      // pass-1 header: [0] chunks of R, [1] chunks of S
    const uint64_t *obase0;      // obase0[1] = R tuples (S offsets start there)
    uint64_t unit;               // tuples per unit (target)
};

// offset of H[rel][d1][c][0]
__device__ __forceinline__ uint64_t h_row(const P2Args &p, int rel, uint32_t d1, uint64_t c) {
    const uint64_t C0 = p.header[0], C = rel ? p.header[1] : C0;
    const uint64_t base = rel ? (C0 << p.b1) << p.b2 : 0;
    return base + (((uint64_t)d1 * C + c) << p.b2);
}

// unit u -> relation, partition d1, its index j among the partition's U units
// and its chunk range [c_lo, c_hi)
struct Unit {
    int rel;
    uint32_t d1, j, U, idx;
    uint64_t C, c_lo, c_hi;
};
__device__ __forceinline__ Unit unit_of(const P2Args &p, uint32_t u) {
    const uint32_t F1 = 1u << p.b1;
    Unit x;
    x.idx = (uint32_t)last_le<uint32_t>(p.ubase, 2ull * F1 + 1, u);
    x.rel = x.idx >= F1;
    x.d1 = x.idx - (x.rel ? F1 : 0);
    x.U = p.ubase[x.idx + 1] - p.ubase[x.idx];
    x.j = u - p.ubase[x.idx];
    x.C = x.rel ? p.header[1] : p.header[0];
    x.c_lo = x.j * x.C / x.U;
    x.c_hi = (x.j + 1) * x.C / x.U;
    return x;
}

template <int PBLOCK>
__global__ __launch_bounds__(PBLOCK) void k_hist2d(PassArgs a, P2Args p) {
    constexpr int HITEMS = 8;  // tuples per lane per tile (fewer registers than the scatter)
    constexpr int PTILE = PBLOCK * HITEMS;
    extern __shared__ uint32_t h[];
    ChunkLoc L;
    if (!locate_chunk(a, blockIdx.x, L)) return;
    const uint32_t F1 = 1u << p.b1, F2 = 1u << p.b2, FF = F1 * F2;
    const uint32_t m1 = F1 - 1, m2 = F2 - 1;
    for (uint32_t i = threadIdx.x; i < FF; i += PBLOCK) h[i] = 0;
    __syncthreads();
    const void *in = L.rel ? a.in1 : a.in0;
    for (uint64_t base = tile_start(L.begin, 8); base < L.end; base += PTILE) {
        uint64_t t[HITEMS];
        bool v[HITEMS];
        load_tile<PBLOCK, HITEMS>(in, 8, base, L.begin, L.end, t, v);
#pragma unroll
        for (int j = 0; j < HITEMS; ++j) {
            const uint32_t k = key_of(t[j]);
            if (v[j]) atomicAdd(&h[((k & m1) << p.b2) | ((k >> p.b1) & m2)], 1u);
        }
    }
    __syncthreads();
    // H[rel][d1][c][d2]: F1 runs of F2 counts
    const uint64_t C0 = a.header[0];
    const uint64_t hbase = L.rel ? (C0 << p.b1) << p.b2 : 0;
    for (uint32_t i = threadIdx.x; i < FF; i += PBLOCK) {
        const uint32_t d1 = i >> p.b2, d2 = i & m2;
        p.H[hbase + (((uint64_t)d1 * L.nchunk + L.c) << p.b2) + d2] = h[i];
    }
    const uint64_t hb = (L.rel ? a.hbase1 : a.hbase0)[L.seg];
    for (uint32_t d1 = threadIdx.x; d1 < F1; d1 += PBLOCK) {
        uint32_t sum = 0;
        for (uint32_t k = 0; k < F2; ++k) sum += h[(d1 << p.b2) | ((k + d1) & m2)];  // rotated: no bank conflicts
        a.hist[hb + d1 * L.nchunk + L.c] = sum;
    }
}

constexpr int H2B_BLOCK = 1024;
__global__ __launch_bounds__(H2B_BLOCK) void k_p2_units(P2Args p) {
    __shared__ uint32_t tmp[H2B_BLOCK / 64];
    const uint32_t F1 = 1u << p.b1;
    const uint64_t C[2] = {p.header[0], p.header[1]};
    uint32_t run = 0;
    for (uint32_t base = 0; base < 2 * F1; base += H2B_BLOCK) {
        const uint32_t i = base + threadIdx.x;
        uint32_t u = 0;
        if (i < 2 * F1) {
            const int rel = i >= F1;
            const uint32_t d1 = i - (rel ? F1 : 0);
            const uint64_t *po = rel ? p.poff1_1 : p.poff1_0;
            const uint64_t size = po[d1 + 1] - po[d1];
            if (size) {
                const uint64_t want = (size + p.unit - 1) / p.unit;
                u = (uint32_t)(want < C[rel] ? want : C[rel]);
                if (u == 0) u = 1;
            }
        }
        uint32_t total;
        const uint32_t pre = block_exclusive_scan<H2B_BLOCK>(u, tmp, &total);
        if (i < 2 * F1) p.ubase[i] = run + pre;
        run += total;
    }
    if (threadIdx.x == 0) p.ubase[2 * F1] = run;
}

// G[u][d2] = sum of H[rel][d1][c][d2] over the unit's chunks: one lane per d2
// (several when F2 exceeds the work-group), four independent loads in flight.
constexpr int USUM_BLOCK = 128;
__global__ __launch_bounds__(USUM_BLOCK) void k_h2_usum(P2Args p) {
    const uint32_t F1 = 1u << p.b1, F2 = 1u << p.b2;
    const uint32_t u = blockIdx.x;
    if (u >= p.ubase[2 * F1]) return;
    const Unit x = unit_of(p, u);
    if (threadIdx.x == 0) {
        // the runs of partition d1 written by pass-1 chunks [c_lo, c_hi)
        const uint64_t hb = (x.rel ? p.hbase1 : p.hbase0)[0];
        const uint64_t rel_base = x.rel ? p.obase0[1] : 0;
        const uint64_t *po1 = x.rel ? p.poff1_1 : p.poff1_0;
        const uint64_t begin = p.hist_scan[hb + x.d1 * x.C + x.c_lo] - rel_base;
        const uint64_t end =
            x.c_hi < x.C ? p.hist_scan[hb + x.d1 * x.C + x.c_hi] - rel_base : po1[x.d1 + 1];
        p.span[2 * (uint64_t)u] = begin | ((uint64_t)x.rel << 63);
        p.span[2 * (uint64_t)u + 1] = end;
    }
    const uint32_t *row0 = p.H + h_row(p, x.rel, x.d1, 0);
    for (uint32_t d2 = threadIdx.x; d2 < F2; d2 += USUM_BLOCK) {
        uint32_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
        uint64_t c = x.c_lo;
        for (; c + 4 <= x.c_hi; c += 4) {
            s0 += row0[(c << p.b2) + d2];
            s1 += row0[((c + 1) << p.b2) + d2];
            s2 += row0[((c + 2) << p.b2) + d2];
            s3 += row0[((c + 3) << p.b2) + d2];
        }
        for (; c < x.c_hi; ++c) s0 += row0[(c << p.b2) + d2];
        p.G[((uint64_t)u << p.b2) + d2] = s0 + s1 + s2 + s3;
    }
}

// One work-group per (relation, pass-1 partition d1), one lane per d2 (F2 <=
// H2B_LANES, or several d2 per lane): the (d1, d2) total is the sum of the
// counts of d1's units; the final offsets are the pass-1 partition start plus
// the exclusive scan of the totals over d2; unit j of d1 starts writing d2 at
// the final offset plus the counts of units 0..j-1, which replace the counts
// in G.  The last work-group of each relation also writes the end offset.
constexpr int H2B_LANES = 256;
__global__ __launch_bounds__(H2B_LANES) void k_h2_base(P2Args p) {
    __shared__ uint64_t tmp[H2B_LANES / 64];
    const uint32_t F1 = 1u << p.b1, F2 = 1u << p.b2;
    const uint32_t idx = blockIdx.x;  // rel * F1 + d1
    if (idx >= 2 * F1) return;
    const int rel = idx >= F1;
    const uint32_t d1 = idx - (rel ? F1 : 0);
    const uint64_t *po1 = rel ? p.poff1_1 : p.poff1_0;
    uint64_t *po = rel ? p.poff2_1 : p.poff2_0;
    const uint32_t u0 = p.ubase[idx], u1 = p.ubase[idx + 1];
    uint64_t run = po1[d1];
    for (uint32_t base = 0; base < F2; base += H2B_LANES) {
        const uint32_t d2 = base + threadIdx.x;
        const bool ok = d2 < F2;
        uint32_t total = 0;
        uint32_t k = u0;
        for (; k + 4 <= u1; k += 4) {  // four independent loads in flight
            const uint32_t *g = p.G + ((uint64_t)k << p.b2) + (ok ? d2 : 0);
            total += g[0] + g[F2] + g[2 * F2] + g[3 * F2];
        }
        for (; k < u1; ++k) total += p.G[((uint64_t)k << p.b2) + (ok ? d2 : 0)];
        if (!ok) total = 0;
        uint64_t sum;
        const uint64_t pre = block_exclusive_scan<H2B_LANES>((uint64_t)total, tmp, &sum);
        const uint64_t start = run + pre;
        if (ok) {
            po[((uint64_t)d1 << p.b2) + d2] = start;
            uint64_t c = start;
            for (uint32_t k2 = u0; k2 < u1; k2 += 8) {  // eight loads in flight
                uint32_t n[8];
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    n[q] = k2 + q < u1 ? p.G[((uint64_t)(k2 + q) << p.b2) + d2] : 0u;
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    if (k2 + q < u1) p.G[((uint64_t)(k2 + q) << p.b2) + d2] = (uint32_t)c;
                    c += n[q];
                }
            }
        }
        run += sum;
    }
    if (d1 == F1 - 1 && threadIdx.x == 0) po[(uint64_t)F1 << p.b2] = po1[F1];
}

template <int PBLOCK>
__global__ __launch_bounds__(PBLOCK, 4) void k_scatter_p2(P2Args p) {
    constexpr int PTILE = PBLOCK * PITEMS;
    extern __shared__ uint64_t smem[];
    __shared__ uint32_t s_len;
    const uint32_t F1 = 1u << p.b1, F2 = 1u << p.b2;
    const uint32_t u = blockIdx.x;
    if (u >= p.ubase[2 * F1]) return;
    // input span from the planning (one load instead of the unit search)
    const uint64_t s0 = p.span[2 * (uint64_t)u], end = p.span[2 * (uint64_t)u + 1];
    const int rel = (int)(s0 >> 63);
    const uint64_t begin = s0 & ~(1ull << 63);
    uint64_t *buf = smem;                                    // PTILE + NSPARE tuples
    uint32_t *cnt = reinterpret_cast<uint32_t *>(buf + PTILE + NSPARE);  // F2 + NSPARE counters
    uint32_t *dst = cnt + F2 + NSPARE;
    uint32_t *cur = dst + F2;
    for (uint32_t d = threadIdx.x; d < F2; d += PBLOCK) {
        cur[d] = p.G[((uint64_t)u << p.b2) + d];
        cnt[d] = 0;
    }
    if (threadIdx.x < NSPARE) cnt[F2 + threadIdx.x] = 0;
    __syncthreads();
    scatter_tiles<8, PBLOCK>(rel ? (const void *)p.in1 : (const void *)p.in0, begin, end,
                             rel ? p.out1 : p.out0, p.b1, F2, buf, cnt, dst, cur, &s_len);
}

// ---------------------------------------------------------------- k_part_off
__global__ void k_part_off(PassArgs a) {
    if (guard_off(a.guard)) return;
    const uint64_t parts = a.header[4];
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p <= a.max_parts;
         p += (uint64_t)gridDim.x * blockDim.x) {
        if (p > parts) {
            a.poff0[p] = a.obase0[a.nseg];
            a.poff1[p] = a.obase1[a.nseg];
            continue;
        }
        if (p == parts) {
            a.poff0[p] = a.obase0[a.nseg];
            a.poff1[p] = a.obase1[a.nseg];
            continue;
        }
        const int k = (int)last_le<uint32_t>(a.pbase, (uint64_t)a.nseg + 1, (uint32_t)p);
        const uint64_t d = p - a.pbase[k];
        for (int rel = 0; rel < 2; ++rel) {
            const uint64_t *cpre = rel ? a.cpre1 : a.cpre0;
            const uint64_t C = cpre[k + 1] - cpre[k];
            const uint64_t hb = (rel ? a.hbase1 : a.hbase0)[k];
            const uint64_t ob = (rel ? a.obase1 : a.obase0)[k];
            const uint64_t rb = rel ? a.obase0[a.nseg] : 0;
            uint64_t *po = rel ? a.poff1 : a.poff0;
            po[p] = C ? (uint64_t)a.hist_scan[hb + d * C] - rb : ob;
        }
    }
}

// ---------------------------------------------------------------- staged window pass
// (run_window_stage / run_window_parts)
struct WStageArgs {
    uint32_t *cap;
    uint64_t *base;
    uint32_t *cur0, *cur1;
    uint32_t *flag;       // the histogram pass must run (set before or during the scatter)
    uint32_t *lflag;      // the layout does not fit (set before the scatter starts)
    uint64_t lim0, lim1;  // layout limits: the output buffer less a trash tile
    int W;
    uint64_t *pstart0, *pstart1;
    uint32_t *plen0, *plen1;
    // S runs past their partition's capacity: spilled, then gathered into one
    // extension range per partition
    uint32_t *strad;
    uint32_t *sp_cur, *sp_n;
    uint4 *sp_list;
    uint32_t sp_base, sp_cap, sp_max;
    uint64_t ext_base;
    uint32_t *ext_cur, *ext_off;
};

constexpr int WPLAN_BLOCK = 1024;
constexpr int WSTAGE_MAXF = 1 << WSTAGE_MAX_BITS;

// A partition of e expected tuples: e + max(e/8, 8 sqrt(e)) + 64, in whole
// 64-tuple lines (as the staged PRO's regions).
__device__ __forceinline__ uint32_t wstage_cap(uint64_t n, uint32_t F) {
    const double e = (double)n / F;
    const double slack = fmax(e / 8.0, 8.0 * sqrt(e));
    const uint64_t c = (uint64_t)ceil(e + slack) + 64;
    return (uint32_t)((c + 63) / 64 * 64);
}

__global__ __launch_bounds__(WPLAN_BLOCK) void k_wstage_caps(PassArgs a, WStageArgs w) {
    __shared__ uint64_t tmp[WPLAN_BLOCK / 64];
    for (int rel = 0; rel < 2; ++rel) {
        const uint64_t *ib = rel ? a.in_begin1 : a.in_begin0;
        const uint64_t *ie = rel ? a.in_end1 : a.in_end0;
        uint64_t run = 0;
        for (int b = 0; b < a.nseg; b += WPLAN_BLOCK) {
            const int k = b + threadIdx.x;
            uint64_t size = 0;
            if (k < a.nseg) {
                const uint32_t F = 1u << a.bits[k];
                const uint32_t c = wstage_cap(ie[k] - ib[k], F);
                w.cap[rel * w.W + k] = c;
                size = (uint64_t)F * c;
            }
            uint64_t tot;
            const uint64_t pre = block_exclusive_scan<WPLAN_BLOCK>(size, tmp, &tot);
            if (k < a.nseg) w.base[rel * w.W + k] = run + pre;
            run += tot;
        }
        if (threadIdx.x == 0 && run > (rel ? w.lim1 : w.lim0)) {
            atomicOr(w.flag, 1u);
            atomicOr(w.lflag, 1u);
        }
        __syncthreads();
    }
}

// One work-group per chunk of a window: the chunk's tiles counting-sorted by
// digit in LDS, each run placed by one atomic reservation in its partition.
template <int PBLOCK, int STRIDE>
__global__ __launch_bounds__(PBLOCK, 4) void k_wscatter(PassArgs a, WStageArgs w) {
    constexpr int PTILE = PBLOCK * PITEMS;
    extern __shared__ uint64_t smem[];
    __shared__ uint32_t s_len;
    // the layout does not fit: the histogram pass runs instead.  (Only the flag
    // written before this launch is read here: the overflow flag that work-groups
    // of this launch set would end some waves of a work-group and not others.)
    if (*w.lflag) return;
    ChunkLoc L;
    if (!locate_chunk(a, blockIdx.x, L)) return;
    const uint32_t F = 1u << a.bits[L.seg];
    uint64_t *buf = smem;                                                // PTILE + NSPARE
    uint32_t *cnt = reinterpret_cast<uint32_t *>(buf + PTILE + NSPARE);  // F + NSPARE
    uint32_t *dst = cnt + F + NSPARE;                                    // F
    for (uint32_t d = threadIdx.x; d < F + NSPARE; d += PBLOCK) cnt[d] = 0;
    __syncthreads();
    struct {
        StageSeg s;
        __device__ bool get(uint32_t i, StageSeg &o) const {
            o = s;
            return i == 0;
        }
    } src;
    const uint32_t c = w.cap[L.rel * w.W + L.seg];
    src.s.in = L.rel ? a.in1 : a.in0;
    src.s.begin = L.begin;
    src.s.end = L.end;
    src.s.r.cur = (L.rel ? w.cur1 : w.cur0) + a.pbase[L.seg];
    src.s.r.base = (uint32_t)w.base[L.rel * w.W + L.seg];
    src.s.r.stride = c;
    src.s.r.cap = c;
    src.s.r.trash = (uint32_t)(L.rel ? w.lim1 : w.lim0);
    src.s.r.flag = w.flag;
    StageSpill sp{};  // S runs that do not fit are spilled; R runs set the flag
    if (L.rel) {
        sp.strad = w.strad;
        sp.cur = w.sp_cur;
        sp.n = w.sp_n;
        sp.list = w.sp_list;
        sp.base = w.sp_base;
        sp.cap = w.sp_cap;
        sp.max = w.sp_max;
        sp.part0 = a.pbase[L.seg];
    }
    stage_tiles<PBLOCK, true, false, false, decltype(src), STRIDE, WSTAGE_MAXF, true>(
        src, L.rel ? a.out1 : a.out0, a.shift, F, buf, cnt, dst, &s_len, &sp);
}

// Partition p's start and length per relation, from the staged layout, or from
// the exact offsets of the histogram pass after an overflow.  Entry p is the
// partition's main range; entry max_parts + p its S extension (with the same R
// range), non-empty only when S runs of the partition were spilled: those are
// given one range of the extension area here and gathered into it by
// k_wspill_gather.
__global__ void k_wstage_parts(PassArgs a, WStageArgs w) {
    const bool over = *w.flag != 0;
    const uint64_t parts = a.header[4];
    const uint64_t M = a.max_parts;
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < M;
         p += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t rs = 0, ss = 0, es = 0;
        uint32_t rl = 0, sl = 0, el = 0;
        if (p < parts) {
            if (over) {
                rs = a.poff0[p];
                rl = (uint32_t)(a.poff0[p + 1] - a.poff0[p]);
                ss = a.poff1[p];
                sl = (uint32_t)(a.poff1[p + 1] - a.poff1[p]);
            } else {
                const int k = (int)last_le<uint32_t>(a.pbase, (uint64_t)a.nseg + 1, (uint32_t)p);
                const uint64_t d = p - a.pbase[k];
                const uint32_t c0 = w.cap[k], c1 = w.cap[w.W + k];
                const uint32_t n0 = w.cur0[p], n1 = w.cur1[p];
                rs = w.base[k] + d * c0;
                rl = n0 < c0 ? n0 : c0;  // R never spills: n0 > c0 set the flag
                ss = w.base[w.W + k] + d * c1;
                sl = n1;
                if (n1 > c1) {
                    sl = w.strad[p];  // the runs from here on were spilled
                    el = n1 - sl;
                    const uint32_t off = atomicAdd(w.ext_cur, el);
                    w.ext_off[p] = off;
                    es = w.ext_base + off;
                }
            }
        }
        w.pstart0[p] = rs;
        w.plen0[p] = rl;
        w.pstart1[p] = ss;
        w.plen1[p] = sl;
        w.pstart0[M + p] = rs;
        w.plen0[M + p] = el ? rl : 0u;
        w.pstart1[M + p] = es;
        w.plen1[M + p] = el;
    }
}

// The spilled S runs, each moved to its place in its partition's extension
// range (the runs of a partition past its capacity cover [strad, cur) of the
// partition's positions, so a run's place follows from its start alone).
__global__ __launch_bounds__(256) void k_wspill_gather(PassArgs a, WStageArgs w) {
    if (*w.flag) return;
    const uint32_t n = min(*w.sp_n, w.sp_max);
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const uint4 e = w.sp_list[i];
        const uint64_t *src = a.out1 + w.sp_base + e.y;
        uint64_t *dst = a.out1 + w.ext_base + w.ext_off[e.x] + (e.z - w.strad[e.x]);
        for (uint32_t j = threadIdx.x; j < e.w; j += blockDim.x) dst[j] = src[j];
    }
}

// ---------------------------------------------------------------- device scan
// Prefix sums of n counters (u32: the histograms of a pass, the join's task
// counts; u64: the window generator's arrival times), modulo the type's width,
// in three launches: the sum of every tile of SCAN_TILE counters, one
// work-group's exclusive scan of those sums, and each tile's own scan from its
// start.  Loads and stores are coalesced (tiles staged through LDS); no
// work-group waits for another, so no launch order is assumed.
constexpr int SCAN_BLOCK = 256, SCAN_ITEMS = 16, SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;
constexpr int SUMS_BLOCK = 1024;

template <class T>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_tiles(const T *in, uint64_t n, T *sums,
                                                          const uint32_t *guard) {
    if (guard_off(guard)) return;
    __shared__ T red[SCAN_BLOCK / 64];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    T s = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        const uint64_t i = base + (uint64_t)j * SCAN_BLOCK + threadIdx.x;
        s += i < n ? in[i] : T(0);
    }
    s = block_sum<SCAN_BLOCK>(s, red);
    if (threadIdx.x == 0) sums[blockIdx.x] = s;
}

template <class T>
__global__ __launch_bounds__(SUMS_BLOCK) void k_scan_sums(T *sums, uint32_t tiles,
                                                         const uint32_t *guard) {
    if (guard_off(guard)) return;
    __shared__ T tmp[SUMS_BLOCK / 64];
    T carry = 0;
    for (uint32_t b = 0; b < tiles; b += SUMS_BLOCK) {
        const uint32_t i = b + threadIdx.x;
        const T v = i < tiles ? sums[i] : T(0);
        T total;
        const T pre = block_exclusive_scan<SUMS_BLOCK>(v, tmp, &total);
        if (i < tiles) sums[i] = carry + pre;
        carry += total;
    }
}

__device__ __forceinline__ uint32_t scan_slot(uint32_t k) { return k + (k >> 5); }

template <class T, bool INCLUSIVE>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_apply(const T *in, T *out, uint64_t n,
                                                         const T *sums, const uint32_t *guard) {
    if (guard_off(guard)) return;
    __shared__ T sh[SCAN_TILE + SCAN_TILE / 32];  // padded: lane t's run spans banks
    __shared__ T tmp[SCAN_BLOCK / 64];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        const uint32_t k = j * SCAN_BLOCK + threadIdx.x;
        const uint64_t i = base + k;
        sh[scan_slot(k)] = i < n ? in[i] : T(0);
    }
    __syncthreads();
    const uint32_t k0 = threadIdx.x * SCAN_ITEMS;
    T s = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) s += sh[scan_slot(k0 + j)];
    T pre = block_exclusive_scan<SCAN_BLOCK>(s, tmp, (T *)nullptr) + sums[blockIdx.x];
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        const T v = sh[scan_slot(k0 + j)];
        sh[scan_slot(k0 + j)] = INCLUSIVE ? pre + v : pre;
        pre += v;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        const uint32_t k = j * SCAN_BLOCK + threadIdx.x;
        const uint64_t i = base + k;
        if (i < n) out[i] = sh[scan_slot(k)];
    }
}

template <class T, bool INCLUSIVE>
static int device_scan(const T *in, T *out, uint64_t n, void *tmp, size_t tmp_bytes,
                       hipStream_t st, const uint32_t *guard = nullptr) {
    if (!n) return AJ_OK;
    const uint64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (tmp_bytes < sizeof(T) * (tiles + 1) || !tmp || tiles >= (1ull << 31) ||
        (const void *)in == (const void *)out) {
        set_error(AJ_ERR_ARG, "device scan: workspace of %zu bytes for %llu counters", tmp_bytes,
                  (unsigned long long)n);
        return AJ_ERR_ARG;
    }
    T *sums = static_cast<T *>(tmp);
    hipLaunchKernelGGL(k_scan_tiles<T>, dim3((unsigned)tiles), dim3(SCAN_BLOCK), 0, st, in, n,
                       sums, guard);
    hipLaunchKernelGGL(k_scan_sums<T>, dim3(1), dim3(SUMS_BLOCK), 0, st, sums, (uint32_t)tiles,
                       guard);
    hipLaunchKernelGGL((k_scan_apply<T, INCLUSIVE>), dim3((unsigned)tiles), dim3(SCAN_BLOCK), 0,
                       st, in, out, n, (const T *)sums, guard);
    AJ_HIP(hipGetLastError());
    return AJ_OK;
}

size_t scan_tmp_bytes(uint64_t n, size_t elem) {
    return elem * ((n + SCAN_TILE - 1) / SCAN_TILE + 1);
}

int exclusive_scan_u32(const uint32_t *in, uint32_t *out, uint64_t n, void *tmp,
                       size_t tmp_bytes, hipStream_t st, const uint32_t *guard) {
    return device_scan<uint32_t, false>(in, out, n, tmp, tmp_bytes, st, guard);
}

int inclusive_scan_u64(const uint64_t *in, uint64_t *out, uint64_t n, void *tmp,
                       size_t tmp_bytes, hipStream_t st) {
    return device_scan<uint64_t, true>(in, out, n, tmp, tmp_bytes, st);
}

// ---------------------------------------------------------------- host side
size_t pass_scan_tmp_bytes(uint64_t max_hist) { return scan_tmp_bytes(max_hist, 4); }

// Raises a kernel's dynamic-LDS limit on the current device to the largest size
// requested so far there; contexts on several devices or threads share the record.
int lds_limit_set(const void *fn, size_t bytes) {
    struct Rec {
        int dev;
        const void *fn;
        size_t bytes;
    };
    static std::mutex mu;
    static std::vector<Rec> recs;
    int dev = 0;
    AJ_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(mu);
    Rec *r = nullptr;
    for (Rec &x : recs)
        if (x.dev == dev && x.fn == fn) r = &x;
    if (r && bytes <= r->bytes) return AJ_OK;
    AJ_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    if (r)
        r->bytes = bytes;
    else
        recs.push_back({dev, fn, bytes});
    return AJ_OK;
}

template <int STRIDE, int PBLOCK>
static int launch_pass_kernels(const PassDesc &d, const PassWork &w, const PassArgs &a,
                               hipStream_t st, const Marks *mk, int m0) {
    const uint32_t F = 1u << d.max_bits;
    const size_t lds_hist = sizeof(uint32_t) * F;
    const size_t lds_scatter =
        sizeof(uint64_t) * PBLOCK * PITEMS + sizeof(uint32_t) * (2 * F + 1);
    AJ_TRY(lds_limit_set((const void *)k_hist<STRIDE, PBLOCK>, lds_hist));
    AJ_TRY(lds_limit_set((const void *)k_scatter<STRIDE, PBLOCK>, lds_scatter));
    const size_t lds_pipe = pipe_lds_bytes(PBLOCK, F);
    if (d.max_bits <= PIPE_MAX_BITS)
        AJ_TRY(lds_limit_set((const void *)k_scatter_pipe<STRIDE, PBLOCK>, lds_pipe));
    if (mk) mk->mark(m0);
    if (d.max_chunks && d.hist2d) {
        P2Args p{};
        p.b1 = d.max_bits;
        p.b2 = d.bits2;
        p.H = d.hist2d;
        const size_t lds2 = sizeof(uint32_t) << (d.max_bits + d.bits2);
        AJ_TRY(lds_limit_set((const void *)k_hist2d<PBLOCK>, lds2));
        hipLaunchKernelGGL((k_hist2d<PBLOCK>), dim3((unsigned)d.max_chunks), dim3(PBLOCK), lds2,
                           st, a, p);
    } else if (d.max_chunks) {
        hipLaunchKernelGGL((k_hist<STRIDE, PBLOCK>), dim3((unsigned)d.max_chunks), dim3(PBLOCK),
                           lds_hist, st, a);
    }
    if (mk) mk->mark(m0 + 1);
    if (d.max_hist)
        AJ_TRY(exclusive_scan_u32(a.hist, a.hist_scan, d.max_hist, w.scan_tmp, w.scan_tmp_bytes,
                                  st, a.guard));
    if (mk) mk->mark(m0 + 2);
    if (d.max_chunks) {
        if (PBLOCK == 512 && (d.hist2d || STRIDE == 12) && d.max_bits <= PIPE_MAX_BITS) {
            // pass 1 of a two-pass join, and window batches (12-byte tuples):
            // 1024-lane work-groups with tiles of 16384 tuples, whose runs per
            // digit are twice as long (the pass writes to every partition at
            // once; longer runs write faster)
            constexpr int BIG = 1024;
            const size_t lds_big = pipe_lds_bytes(BIG, F);
            AJ_TRY(lds_limit_set((const void *)k_scatter_pipe<STRIDE, BIG>, lds_big));
            hipLaunchKernelGGL((k_scatter_pipe<STRIDE, BIG>), dim3((unsigned)d.max_chunks), dim3(BIG),
                               lds_big, st, a);
        } else if (d.max_bits <= PIPE_MAX_BITS) {
            hipLaunchKernelGGL((k_scatter_pipe<STRIDE, PBLOCK>), dim3((unsigned)d.max_chunks),
                               dim3(PBLOCK), lds_pipe, st, a);
        } else if (d.max_bits <= PIPE_WIDE_BITS) {
            // 2^11 digits: 1024-lane work-groups with 16384-tuple tiles (runs
            // of ~8 tuples per digit), the digits scanned block-wide
            constexpr int BIG = 1024;
            const size_t lds_big = pipe_lds_bytes(BIG, F);
            AJ_TRY(lds_limit_set((const void *)k_scatter_pipe<STRIDE, BIG, true>, lds_big));
            hipLaunchKernelGGL((k_scatter_pipe<STRIDE, BIG, true>), dim3((unsigned)d.max_chunks),
                               dim3(BIG), lds_big, st, a);
        } else
            hipLaunchKernelGGL((k_scatter<STRIDE, PBLOCK>), dim3((unsigned)d.max_chunks),
                               dim3(PBLOCK), lds_scatter, st, a);
    }
    if (mk) mk->mark(m0 + 3);
    return AJ_OK;
}

static PassArgs pass_args(const PassDesc &d, const PassWork &w) {
    PassArgs a;
    a.in0 = d.in[0];
    a.in1 = d.in[1];
    a.stride = d.in_stride;
    a.out0 = d.out[0];
    a.out1 = d.out[1];
    a.in_begin0 = d.in_begin[0];
    a.in_begin1 = d.in_begin[1];
    a.in_end0 = d.in_end[0];
    a.in_end1 = d.in_end[1];
    a.bits = d.bits;
    a.nseg = d.nseg;
    a.shift = d.shift;
    a.chunk = d.chunk;
    a.n_total0 = d.n_total[0];
    a.hist = w.hist;
    a.hist_scan = w.hist + d.max_hist;
    a.cpre0 = w.seg_chunk_pre[0];
    a.cpre1 = w.seg_chunk_pre[1];
    a.hbase0 = w.seg_hist_base[0];
    a.hbase1 = w.seg_hist_base[1];
    a.obase0 = w.seg_out_base[0];
    a.obase1 = w.seg_out_base[1];
    a.pbase = w.seg_part_base;
    a.header = w.header;
    a.poff0 = w.part_off[0];
    a.poff1 = w.part_off[1];
    a.max_hist = d.max_hist;
    a.max_parts = d.max_parts;
    a.guard = d.guard;
    return a;
}

int run_partition_pass(const PassDesc &d, const PassWork &w, hipStream_t st, const Marks *mk,
                       int m0) {
    if (d.max_bits > MAX_PASS_BITS || d.max_bits < 0) {
        set_error(AJ_ERR_ARG, "radix bits per pass must be in [0, %d], got %d", MAX_PASS_BITS,
                  d.max_bits);
        return AJ_ERR_ARG;
    }
    if (d.block != 256 && d.block != 512) {
        set_error(AJ_ERR_ARG, "partition work-group size must be 256 or 512, got %d", d.block);
        return AJ_ERR_ARG;
    }
    if (d.chunk % PTILE_MAX) {
        set_error(AJ_ERR_ARG, "chunk must be a multiple of %d", PTILE_MAX);
        return AJ_ERR_ARG;
    }
    if (d.in_stride != 8 && d.in_stride != 12) {
        set_error(AJ_ERR_ARG, "tuple stride must be 8 or 12 bytes");
        return AJ_ERR_ARG;
    }
    if (d.hist2d && (d.in_stride != 8 || d.nseg != 1 || d.bits2 < 1 ||
                     d.max_bits + d.bits2 > P2D_MAX_BITS || d.max_bits > PIPE_MAX_BITS ||
                     d.bits2 > PIPE_MAX_BITS)) {
        set_error(AJ_ERR_ARG, "2-D pass-1 histogram needs one 8-byte segment and at most %d bits",
                  P2D_MAX_BITS);
        return AJ_ERR_ARG;
    }
    if (d.n_total[0] + d.n_total[1] >= (1ull << 32)) {
        set_error(AJ_ERR_ARG, "|R|+|S| per launch must be below 2^32");
        return AJ_ERR_ARG;
    }
    if (d.max_hist >= (1ull << 31)) {
        set_error(AJ_ERR_ARG, "histogram too large (%llu entries)", (unsigned long long)d.max_hist);
        return AJ_ERR_ARG;
    }
    const PassArgs a = pass_args(d, w);
    if (d.exact_single && d.nseg == 1) {
        const uint32_t F = 1u << d.max_bits;
        const uint64_t c0 = (d.n_total[0] + d.chunk - 1) / d.chunk;
        const uint64_t c1 = (d.n_total[1] + d.chunk - 1) / d.chunk;
        hipLaunchKernelGGL(k_plan_single, dim3(1), dim3(64), 0, st, a, d.n_total[0],
                           d.n_total[1], c0, c1, F);
        if ((c0 + c1) * F < d.max_hist)  // the scan covers max_hist entries
            hipLaunchKernelGGL(k_zero_tail, dim3(256), dim3(256), 0, st, a.hist, a.header,
                               d.max_hist, a.guard);
    } else {
        if (!d.skip_plan) hipLaunchKernelGGL(k_plan, dim3(1), dim3(PLAN_BLOCK), 0, st, a);
        hipLaunchKernelGGL(k_zero_tail, dim3(256), dim3(256), 0, st, a.hist, a.header,
                           d.max_hist, a.guard);
    }
    const bool s8 = d.in_stride == 8;
    int rc;
    if (d.block == 512)
        rc = s8 ? launch_pass_kernels<8, 512>(d, w, a, st, mk, m0)
                : launch_pass_kernels<12, 512>(d, w, a, st, mk, m0);
    else
        rc = s8 ? launch_pass_kernels<8, 256>(d, w, a, st, mk, m0)
                : launch_pass_kernels<12, 256>(d, w, a, st, mk, m0);
    AJ_TRY(rc);
    const unsigned pgrid = (unsigned)std::min<uint64_t>((d.max_parts + 256) / 256, 4096);
    hipLaunchKernelGGL(k_part_off, dim3(pgrid), dim3(256), 0, st, a);
    AJ_HIP(hipGetLastError());
    return AJ_OK;
}

int run_partition_plan(const PassDesc &d, const PassWork &w, hipStream_t st) {
    const PassArgs a = pass_args(d, w);
    hipLaunchKernelGGL(k_plan, dim3(1), dim3(PLAN_BLOCK), 0, st, a);
    AJ_HIP(hipGetLastError());
    return AJ_OK;
}



// ---------------------------------------------------------------- staged window pass, host
uint64_t window_stage_bound(uint64_t n, uint64_t parts) {
    // sum over windows of F * cap <= 1.125 n + 8 sqrt(F n) + 128 F, and the
    // sum of sqrt(F_w n_w) is at most sqrt(parts * n)
    const double nn = (double)n, pp = (double)std::max<uint64_t>(parts, 1);
    return (uint64_t)std::ceil(1.125 * nn + 8.0 * std::sqrt(pp * nn) + 128.0 * pp) +
           2ull * 1024 * PITEMS;
}

static WStageArgs wstage_args(const WStageDesc &ws) {
    WStageArgs w;
    w.cap = ws.cap;
    w.base = ws.base;
    w.cur0 = ws.cur[0];
    w.cur1 = ws.cur[1];
    w.flag = ws.flag;
    w.lflag = ws.lflag;
    // the last tile of each buffer is the trash area
    w.lim0 = ws.out_cap[0] - 1024 * PITEMS;
    w.lim1 = ws.out_cap[1] - 1024 * PITEMS;
    w.W = ws.W;
    w.pstart0 = ws.pstart[0];
    w.pstart1 = ws.pstart[1];
    w.plen0 = ws.plen[0];
    w.plen1 = ws.plen[1];
    w.strad = ws.strad;
    w.sp_cur = ws.spill_cur;
    w.sp_n = ws.spill_n;
    w.sp_list = ws.spill_list;
    w.sp_base = (uint32_t)ws.out_cap[1];  // the spill area follows the S layout
    w.sp_cap = (uint32_t)ws.spill_cap;
    w.sp_max = ws.spill_max;
    w.ext_base = ws.out_cap[1] + ws.spill_cap;  // then the extension ranges
    w.ext_cur = ws.ext_cur;
    w.ext_off = ws.ext_off;
    return w;
}

int run_window_stage(const PassDesc &d, const PassWork &wk, const WStageDesc &ws,
                     hipStream_t st) {
    if ((d.in_stride != 12 && d.in_stride != 8) || d.max_bits > WSTAGE_MAX_BITS || d.nseg > ws.W ||
        ws.out_cap[0] >= (1ull << 32) || ws.out_cap[1] + ws.spill_cap >= (1ull << 32) ||
        ws.out_cap[0] < 2 * 1024 * PITEMS || ws.out_cap[1] < 2 * 1024 * PITEMS) {
        set_error(AJ_ERR_ARG, "staged window pass: unsupported geometry");
        return AJ_ERR_ARG;
    }
    const PassArgs a = pass_args(d, wk);
    const WStageArgs w = wstage_args(ws);
    hipLaunchKernelGGL(k_wstage_caps, dim3(1), dim3(WPLAN_BLOCK), 0, st, a, w);
    constexpr int B = 1024;
    const size_t lds = stage_lds_bytes(B, 1u << d.max_bits);
    const void *fn = d.in_stride == 12 ? (const void *)k_wscatter<B, 12> : (const void *)k_wscatter<B, 8>;
    AJ_TRY(lds_limit_set(fn, lds));
    if (d.max_chunks && d.in_stride == 12)
        hipLaunchKernelGGL((k_wscatter<B, 12>), dim3((unsigned)d.max_chunks), dim3(B), lds, st, a, w);
    else if (d.max_chunks)
        hipLaunchKernelGGL((k_wscatter<B, 8>), dim3((unsigned)d.max_chunks), dim3(B), lds, st, a, w);
    AJ_HIP(hipGetLastError());
    return AJ_OK;
}

int run_window_parts(const PassDesc &d, const PassWork &wk, const WStageDesc &ws,
                     hipStream_t st) {
    const PassArgs a = pass_args(d, wk);
    const WStageArgs w = wstage_args(ws);
    const unsigned g = (unsigned)std::min<uint64_t>((d.max_parts + 255) / 256, 4096);
    hipLaunchKernelGGL(k_wstage_parts, dim3(g), dim3(256), 0, st, a, w);
    hipLaunchKernelGGL(k_wspill_gather, dim3(1024), dim3(256), 0, st, a, w);
    AJ_HIP(hipGetLastError());
    return AJ_OK;
}

template <int PBLOCK>
static int launch_scatter_p2(const P2Args &p, uint64_t max_units, hipStream_t st) {
    const size_t lds = pipe_lds_bytes(PBLOCK, 1u << p.b2);
    AJ_TRY(lds_limit_set((const void *)k_scatter_p2<PBLOCK>, lds));
    hipLaunchKernelGGL((k_scatter_p2<PBLOCK>), dim3((unsigned)max_units), dim3(PBLOCK), lds, st, p);
    return AJ_OK;
}

int run_pass2_2d(const P2Desc &d, hipStream_t st, const Marks *mk, int m0) {
    if (d.b1 < 1 || d.b2 < 1 || d.b1 + d.b2 > P2D_MAX_BITS || d.b2 > PIPE_MAX_BITS || !d.w1 ||
        d.unit == 0 || (d.block != 256 && d.block != 512)) {
        set_error(AJ_ERR_ARG, "run_pass2_2d: bad geometry");
        return AJ_ERR_ARG;
    }
    P2Args p{};
    p.in0 = d.in[0];
    p.in1 = d.in[1];
    p.out0 = d.out[0];
    p.out1 = d.out[1];
    p.b1 = d.b1;
    p.b2 = d.b2;
    p.H = d.hist2d;
    p.G = d.unit_counts;
    p.span = d.unit_span;
    p.poff2_0 = d.poff2[0];
    p.poff2_1 = d.poff2[1];
    p.ubase = d.ubase;
    p.poff1_0 = d.w1->part_off[0];
    p.poff1_1 = d.w1->part_off[1];
    p.hbase0 = d.w1->seg_hist_base[0];
    p.hbase1 = d.w1->seg_hist_base[1];
    p.header = d.w1->header;
    p.obase0 = d.w1->seg_out_base[0];
    p.unit = d.unit;
    p.hist_scan = d.hist_scan;
    if (!p.hist_scan || !p.H || !p.G || !p.span || !p.ubase) {
        set_error(AJ_ERR_ARG, "run_pass2_2d: missing workspace");
        return AJ_ERR_ARG;
    }
    if (mk) mk->mark(m0);
    hipLaunchKernelGGL(k_p2_units, dim3(1), dim3(H2B_BLOCK), 0, st, p);
    hipLaunchKernelGGL(k_h2_usum, dim3((unsigned)d.max_units), dim3(USUM_BLOCK), 0, st, p);
    hipLaunchKernelGGL(k_h2_base, dim3(2u << d.b1), dim3(H2B_LANES), 0, st, p);
    if (mk) {
        mk->mark(m0 + 1);
        mk->mark(m0 + 2);
    }
    // 1024 lanes: one work-group per CU at 128 VGPRs, 16384-tuple tiles
    AJ_TRY(d.block == 512 ? launch_scatter_p2<1024>(p, d.max_units, st)
                          : launch_scatter_p2<256>(p, d.max_units, st));
    if (mk) mk->mark(m0 + 3);
    AJ_HIP(hipGetLastError());
    return AJ_OK;
}

}  // namespace aj
