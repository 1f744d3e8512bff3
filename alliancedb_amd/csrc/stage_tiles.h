// stage_tiles.h -- the tile loop of the staged partitioning passes: tiles
// counting-sorted by digit in LDS, each digit's run written where one atomic
// reservation on its region's cursor puts it (no histogram pass).  Used by the
// staged two-pass PRO (staged.hip) and the staged window pass (windows.hip).
#pragma once

#include "device_common.h"
#include "tiles.h"

namespace aj {

constexpr int STAGE_MAXF = 256;  // digits per pass of the staged PRO (8 bits)

// Keys replaced by their murmur3 finalisation (a bijection, undone by
// fmix32_inv): the hashed partitioning of NPO's region build.
template <int N> __device__ __forceinline__ void hash_keys(uint64_t (&t)[N]) {
#pragma unroll
    for (int j = 0; j < N; ++j) t[j] = ((uint64_t)fmix32(key_of(t[j])) << 32) | (uint32_t)t[j];
}

// Where runs that do not fit their partition go, when the partitions may
// spill (the S side of window batches): a spill area filled in arrival order,
// and a list of the runs placed there.
struct StageSpill {
    uint32_t *strad;       // null: no spilling; else per partition, the start of
                           // the run that crossed its capacity (written by that run)
    uint32_t *cur, *n;     // tuples of the spill area taken; runs listed
    uint4 *list;           // per run: (partition, spill offset, start in the partition, length)
    uint32_t base, cap;    // out index and tuples of the spill area
    uint32_t max;          // list entries
    uint32_t part0;        // partition index of digit 0
};

// Where the runs of a tile go: region d starts at out index base + d * stride,
// holds cap tuples and has filled *cur[d] of them.
struct StageRes {
    uint32_t *cur;
    uint32_t base, stride, cap;
    uint32_t trash;  // out index of a tile-sized area for runs that do not fit
    uint32_t *flag;
};

// One contiguous input range of the tile loop and where its runs go.
struct StageSeg {
    const void *in;  // 8-byte or 12-byte tuples (the tile loop's STRIDE)
    uint64_t begin, end;
    StageRes r;
};

// The pipelined tile loop of the partition scatter (four barriers per tile,
// branch-free phases, the next tile's loads in flight during the write), with
// the write bases taken from atomic reservations instead of running cursors.
// Wave 0 scans the tile's F counters and issues one atomic add per non-empty
// digit; the returned offsets are needed only for the write phase, so they are
// turned into write bases after wave 0 has placed its own tuples.
// The input is a sequence of segments src.get(0), src.get(1), ... (each with
// its own destination regions); the loop runs through all of them as one
// pipeline, so the first tile of a segment is loaded while the last tile of
// the one before is written.  Every value src.get returns is wave-uniform.
// STRIDE: input tuples of 8 bytes (tuple_t) or 12 (window stream tuples, written
// out as 8-byte tuples); MAXF: the most digits a tile may have.  A run that does
// not fit its region goes to the trash area and sets *flag, or with SPILL (and
// sp.strad set) to the spill area, listed there; *flag then only when the spill
// area or its list is full.
template <int PBLOCK, bool NT_LOAD, bool NT_STORE, bool HASH, class Src, int STRIDE = 8,
          int MAXF = STAGE_MAXF, bool SPILL = false>
__device__ __forceinline__ void stage_tiles(const Src &src, uint64_t *out, int shift, uint32_t F,
                                            uint64_t *buf, uint32_t *cnt, uint32_t *dst,
                                            uint32_t *s_len, const StageSpill *spill = nullptr) {
    constexpr int PTILE = PBLOCK * PITEMS;
    // up to STAGE_MAXF digits wave 0 scans and reserves them all (MAXF / 64
    // per lane); beyond, every lane owns 4 digits and a block-wide scan orders
    // them, so no lane holds more than 4 reservations
    constexpr bool WIDE = MAXF > STAGE_MAXF;
    static_assert(!WIDE || MAXF <= 4 * PBLOCK, "at most 4 digits per lane");
    constexpr int MAXPER = WIDE ? 4 : MAXF / 64;
    constexpr int WB = 8;
    static_assert(PITEMS % WB == 0, "a tile's items split into whole batches");
    const uint32_t mask = F - 1;
    StageSeg sg;
    uint32_t si = 0;
    uint64_t base;
    for (;; ++si) {  // the first segment with a tuple
        if (!src.get(si, sg)) return;
        base = tile_start(sg.begin, STRIDE);
        if (sg.begin < sg.end) break;  // (base may be begin - 1 of an empty segment)
    }
    uint64_t t[PITEMS];
    bool v[PITEMS];
    load_tile<PBLOCK, PITEMS, NT_LOAD>(sg.in, STRIDE, base, sg.begin, sg.end, t, v);
    if constexpr (HASH) hash_keys<PITEMS>(t);
    for (;;) {
        const StageRes &r = sg.r;
        uint32_t rank2[PITEMS / 2];
        // A wave whose first tuple of the tile lies past the segment's end
        // holds none of its tuples (most waves of a segment's last tile, when
        // that holds only a few): it skips the count here and the placement
        // below.  Lane l's tuples are tb + 2 (j PBLOCK + l) + {0, 1} for
        // 8-byte tuples and tb + j PBLOCK + l for 12-byte ones.
        const uint64_t wfirst =
            base + (STRIDE == 8 ? 2ull : 1ull) * __builtin_amdgcn_readfirstlane(threadIdx.x);
        const bool busy = wfirst < sg.end;
        if (busy) {
#pragma unroll
            for (int j = 0; j < PITEMS; ++j) {
                const uint32_t d = (key_of(t[j]) >> shift) & mask;
                const uint32_t rk =
                    atomicAdd(&cnt[v[j] ? d : F + (threadIdx.x & (NSPARE - 1))], 1u) & 0xffffu;
                if (j & 1)
                    rank2[j / 2] |= rk << 16;
                else
                    rank2[j / 2] = rk;
            }
        }
        __syncthreads();
        uint32_t o[MAXPER], pk[MAXPER], ck[MAXPER];
        const uint32_t lane = threadIdx.x;
        const uint32_t per = WIDE ? (uint32_t)MAXPER : (F + 63) >> 6;
        const uint32_t b = lane * per;
        if constexpr (WIDE) {
            __shared__ uint32_t wscan[PBLOCK / 64];
            uint32_t sum = 0;
#pragma unroll
            for (int k = 0; k < MAXPER; ++k) {
                ck[k] = b + k < F ? cnt[b + k] : 0u;
                sum += ck[k];
            }
            uint32_t total;
            uint32_t pre = block_exclusive_scan<PBLOCK>(sum, wscan, &total);
#pragma unroll
            for (int k = 0; k < MAXPER; ++k) {
                pk[k] = pre;
                if (b + k < F) cnt[b + k] = pre;
                o[k] = ck[k] ? atomicAdd(&r.cur[b + k], ck[k]) : 0u;
                pre += ck[k];
            }
            if (lane == 0) *s_len = total;
        } else if (threadIdx.x < 64) {
            uint32_t sum = 0;
#pragma unroll
            for (int k = 0; k < MAXPER; ++k) {
                ck[k] = ((uint32_t)k < per && b + k < F) ? cnt[b + k] : 0u;
                sum += ck[k];
            }
            const uint32_t inc = wave_inclusive_scan(sum);
            uint32_t pre = inc - sum;
#pragma unroll
            for (int k = 0; k < MAXPER; ++k) {
                pk[k] = pre;
                if ((uint32_t)k < per && b + k < F) cnt[b + k] = pre;
                o[k] = ck[k] ? atomicAdd(&r.cur[b + k], ck[k]) : 0u;
                pre += ck[k];
            }
            if (lane == 63) *s_len = inc;
        }
        __syncthreads();
        if (busy) {
#pragma unroll
            for (int j0 = 0; j0 < PITEMS; j0 += WB) {
                uint32_t st[WB];
#pragma unroll
                for (int j = 0; j < WB; ++j) st[j] = cnt[(key_of(t[j0 + j]) >> shift) & mask];
#pragma unroll
                for (int j = 0; j < WB; ++j)
                    buf[v[j0 + j] ? st[j] + ((rank2[(j0 + j) / 2] >> (16 * ((j0 + j) & 1))) & 0xffffu)
                                  : (uint32_t)PTILE + (threadIdx.x & (NSPARE - 1))] = t[j0 + j];
            }
        }
        if (WIDE || threadIdx.x < 64) {
#pragma unroll
            for (int k = 0; k < MAXPER; ++k) {
                if (ck[k]) {
                    const bool fits = o[k] <= r.cap && ck[k] <= r.cap - o[k];
                    if constexpr (!SPILL) {
                        if (!fits) atomicOr(r.flag, 1u);
                        dst[b + k] = fits ? r.base + (b + k) * r.stride + o[k] - pk[k]
                                          : r.trash - pk[k];
                    } else {
                        uint32_t wb = r.base + (b + k) * r.stride + o[k] - pk[k];
                        if (!fits) {
                            bool placed = false;
                            wb = r.trash - pk[k];
                            const StageSpill &sp = *spill;
                            if (sp.strad) {
                                const uint32_t part = sp.part0 + b + k;
                                // the one run holding position cap of the partition
                                if (o[k] <= r.cap) sp.strad[part] = o[k];
                                const uint32_t off = atomicAdd(sp.cur, ck[k]);
                                const uint32_t idx = atomicAdd(sp.n, 1u);
                                if (idx < sp.max && off <= sp.cap && ck[k] <= sp.cap - off) {
                                    sp.list[idx] = make_uint4(part, off, o[k], ck[k]);
                                    wb = sp.base + off - pk[k];
                                    placed = true;
                                }
                            }
                            if (!placed) atomicOr(r.flag, 1u);
                        }
                        dst[b + k] = wb;
                    }
                }
            }
        }
        __syncthreads();
        const uint32_t len = *s_len;
        // the next tile: in this segment, else the first of a later one
        uint64_t nbase = base + PTILE;
        bool more = nbase < sg.end;
        StageSeg ns = sg;
        while (!more && src.get(++si, ns)) {
            nbase = tile_start(ns.begin, STRIDE);
            more = ns.begin < ns.end;
        }
        if (more) {
            load_tile<PBLOCK, PITEMS, NT_LOAD>(ns.in, STRIDE, nbase, ns.begin, ns.end, t, v);
            if constexpr (HASH) hash_keys<PITEMS>(t);
        }
        // the first lane of this wave: batches past the tile's end are skipped
        // (a wave-uniform branch; only a partly filled tile has them)
        const uint32_t wave0 = __builtin_amdgcn_readfirstlane(threadIdx.x);
#pragma unroll
        for (int k0 = 0; k0 < PITEMS; k0 += WB) {
            if (wave0 + (uint32_t)k0 * PBLOCK >= len) break;
            uint64_t u[WB];
            uint32_t w[WB];
#pragma unroll
            for (int k = 0; k < WB; ++k) {
                const uint32_t i = threadIdx.x + (uint32_t)(k0 + k) * PBLOCK;
                u[k] = buf[i < len ? i : (uint32_t)PTILE];
            }
#pragma unroll
            for (int k = 0; k < WB; ++k) w[k] = dst[(key_of(u[k]) >> shift) & mask];
#pragma unroll
            for (int k = 0; k < WB; ++k) {
                const uint32_t i = threadIdx.x + (uint32_t)(k0 + k) * PBLOCK;
                if (i < len) {
                    uint64_t *q = out + (uint64_t)(uint32_t)(w[k] + i);
                    if constexpr (NT_STORE)
                        __builtin_nontemporal_store(u[k], q);
                    else
                        *q = u[k];
                }
            }
        }
        for (uint32_t d = threadIdx.x; d < F; d += PBLOCK) cnt[d] = 0;
        __syncthreads();
        if (!more) break;
        base = nbase;
        sg = ns;
    }
}

constexpr size_t stage_lds_bytes(int pblock, uint32_t F) {
    return sizeof(uint64_t) * ((size_t)pblock * PITEMS + NSPARE) +
           sizeof(uint32_t) * (2 * (size_t)F + NSPARE);
}

}  // namespace aj
