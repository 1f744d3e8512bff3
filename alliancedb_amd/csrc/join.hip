// join.hip -- build/probe of partition pairs in LDS (PRO join phase).
//
// Reference: bucket_chaining_join, parallel_radix_join.cpp:291-354 (build a
// table over one R partition, stream the matching S partition, count every
// key-equal pair), driven by the join task queue (parallel_radix_join.cpp:
// 1409-1416) and, for large partitions, the skew split that gives S slices of
// one partition to several threads (parallel_radix_join.cpp:1299-1321).
//
// CDNA4 form:
//   k_task_count  tasks per partition = ceil(|S_p| / split) if |R_p|,|S_p| > 0
//                 (split: 8 LDS rounds of S by default)
//   scan          task bases (flat exclusive scan)
//   k_task_fill   task -> (R range, S slice, segment) records, one thread per
//                 task, so a worker finds its task with one load
//   k_join        persistent work-groups take tasks round-robin (t = wg,
//                 wg + grid, ...).  A task's first CAP tuples of R_p and its
//                 whole S slice arrive in registers before the task starts
//                 (loaded during the previous task); the build counting-sorts
//                 R by bucket into LDS (16-bit bucket starts, about two
//                 tuples per bucket) straight from the registers, the next
//                 task's R is requested, and every S tuple probes one bucket
//                 (six S tuples of a lane step through their buckets together).  R partitions larger than
//                 CAP take further rounds read from memory.  Each task stores
//                 its count and checksum; k_seg_reduce adds them per segment.
//                 With pairs enabled each round probes twice: count, reserve a
//                 range with one atomic, write.
//
// Bytes per tuple: R and S are each read once (R partitions above CAP: the
// extra rounds read their R tuples twice).

#include "aj_internal.h"
#include "device_common.h"

namespace aj {

constexpr int MAX_BUCKET_LG = 12;  // at most 4096 buckets per round

// Buckets of a round of m tuples: the power of two nb >= m/2 (about two
// tuples per bucket), at most 2^MAX_BUCKET_LG.
__host__ __device__ constexpr uint32_t max_buckets(uint32_t cap) {
    uint32_t p = 1;
    while (2 * p < cap && p < (1u << MAX_BUCKET_LG)) p *= 2;
    return p;
}
__device__ __forceinline__ int bucket_lg(uint32_t m) {
    int lg = 0;
    while ((2u << lg) < m && lg < MAX_BUCKET_LG) ++lg;
    return lg;
}

// LDS bytes of one round's table: the tuples in bucket order (8 bytes each)
// plus one spare slot, and the 16-bit bucket bounds (a leading 0, nb bucket
// ends, a spare cursor).
__host__ __device__ constexpr size_t join_lds_bytes(uint32_t cap) {
    return sizeof(uint64_t) * (cap + 1) + sizeof(uint16_t) * (max_buckets(cap) + 4);
}

// S tuples per lane probed together: their candidate tuples are read from LDS
// in one batch per step, so the LDS latency is paid once per step and group.
constexpr int PGROUP = 6;

struct JoinArgs {
    const uint64_t *R, *S, *roff, *soff;
    const uint32_t *rlen, *slen;  // nullable: partition p is [roff[p], + rlen[p]) (else up to roff[p+1])
    const uint32_t *part_seg_base;
    int nseg;
    uint64_t max_parts;
    uint64_t nreal;  // nonzero: entry p >= nreal belongs to partition p - nreal
    uint32_t split;
    uint64_t *seg_count, *seg_sum;
    uint64_t *pairs;
    uint64_t pairs_cap;
    uint64_t *pair_cursor;
    int records;  // 0: pairs; 1: the S tuple <S.key, S.payloadID>; 2: (clock << 32 | S.payloadID)
    int hash_shift;             // radix bits of every partition (when seg_bits is null)
    const uint32_t *seg_bits;   // nullable: radix bits per segment
    uint32_t *task_cnt, *task_base, *task_next;
    uint4 *task_info;
    ulonglong2 *task_res;  // per task {count, checksum}
};

// Per segment k: count/checksum = sums over its tasks, which are contiguous:
// [task_base[part_seg_base[k]], task_base[part_seg_base[k+1]]), and with nreal
// set also the tasks of its partitions' further S ranges, the same range of
// entries shifted by nreal.  One block per segment (grid-stride), plain stores:
// the accumulators need no zeroing.
constexpr int RBLOCK = 1024;
__global__ __launch_bounds__(RBLOCK) void k_seg_reduce(JoinArgs a) {
    __shared__ uint64_t red[2 * (RBLOCK / 64)];
    for (int k = blockIdx.x; k < a.nseg; k += gridDim.x) {
        uint64_t c = 0, x = 0;
        for (int h = 0; h < (a.nreal ? 2 : 1); ++h) {
            const uint64_t sh = h ? a.nreal : 0;
            const uint32_t tb = a.task_base[sh + a.part_seg_base[k]];
            const uint32_t te = a.task_base[sh + a.part_seg_base[k + 1]];
            uint32_t t = tb + threadIdx.x;
            for (; t + 3 * RBLOCK < te; t += 4 * RBLOCK) {  // four loads in flight
                const ulonglong2 r0 = a.task_res[t], r1 = a.task_res[t + RBLOCK];
                const ulonglong2 r2 = a.task_res[t + 2 * RBLOCK], r3 = a.task_res[t + 3 * RBLOCK];
                c += r0.x + r1.x + r2.x + r3.x;
                x += r0.y + r1.y + r2.y + r3.y;
            }
            for (; t < te; t += RBLOCK) {
                const ulonglong2 r = a.task_res[t];
                c += r.x;
                x += r.y;
            }
        }
        c = wave_sum(c);
        x = wave_sum(x);
        if (lane_id() == 0) {
            red[threadIdx.x / 64] = c;
            red[RBLOCK / 64 + threadIdx.x / 64] = x;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t cc = 0, xx = 0;
            for (int w = 0; w < RBLOCK / 64; ++w) {
                cc += red[w];
                xx += red[RBLOCK / 64 + w];
            }
            a.seg_count[k] = cc;
            a.seg_sum[k] = xx;
        }
        __syncthreads();
    }
}

__global__ void k_task_count(JoinArgs a) {
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p <= a.max_parts;
         p += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t n = 0;
        if (p < a.max_parts) {
            const uint64_t nr = a.rlen ? a.rlen[p] : a.roff[p + 1] - a.roff[p];
            const uint64_t ns = a.slen ? a.slen[p] : a.soff[p + 1] - a.soff[p];
            if (nr && ns) n = (uint32_t)((ns + a.split - 1) / a.split);
        }
        a.task_cnt[p] = n;
    }
}

// One thread per task: its partition is the last p with task_base[p] <= t.
__global__ void k_task_fill(JoinArgs a) {
    const uint32_t total = a.task_base[a.max_parts];
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < total;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t p = last_le<uint32_t>(a.task_base, a.max_parts + 1, (uint32_t)t);
        const uint32_t j = (uint32_t)t - a.task_base[p];
        const uint64_t pr = a.nreal && p >= a.nreal ? p - a.nreal : p;
        const uint32_t seg =
            a.nseg == 1 ? 0u
                        : (uint32_t)last_le<uint32_t>(a.part_seg_base, (uint64_t)a.nseg + 1,
                                                      (uint32_t)pr);
        const uint64_t r0 = a.roff[p], nr = a.rlen ? a.rlen[p] : a.roff[p + 1] - r0;
        const uint64_t s_end = a.slen ? a.soff[p] + a.slen[p] : a.soff[p + 1];
        const uint64_t s0 = a.soff[p] + (uint64_t)j * a.split;
        const uint64_t ns = min<uint64_t>(a.split, s_end - s0);
        const uint32_t sh = min(31u, a.seg_bits ? a.seg_bits[seg] : (uint32_t)a.hash_shift);
        a.task_info[2 * t] = make_uint4((uint32_t)r0, (uint32_t)(r0 >> 32), (uint32_t)nr, seg);
        a.task_info[2 * t + 1] = make_uint4((uint32_t)s0, (uint32_t)(s0 >> 32), (uint32_t)ns, sh);
    }
}

__device__ __forceinline__ uint32_t umin32(uint32_t x, uint32_t y) { return x < y ? x : y; }

// threadIdx.x behind an empty asm statement: the per-tuple indices derived from
// it are recomputed where they are used instead of being hoisted out of the
// task loop and kept live (or spilled) across every phase.
__device__ __forceinline__ uint32_t opaque_tid() {
    uint32_t t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

// Bucket of a key among 2^lg (hp = lg | sh << 8, lg <= 12, sh <= 31): every
// key of a partition has the same sh low bits (the partition bits), so the
// bucket is taken from the bits above them, as in bucket_chaining_join
// (HASH_BIT_MODULO, parallel_radix_join.cpp:68,309): the lg bits above the
// partition bits, xor the next lg bits, xor the lg bits after those (bit-field
// offsets capped at 31).  Keys that differ only in the first field never share
// a bucket (a dense key range fills every bucket evenly), and keys that agree
// in it are spread by the fields above.  Three bit-field extracts and one
// three-way xor: no multiply (a quarter-rate instruction).
__device__ __forceinline__ uint32_t bucket_of(uint32_t key, uint32_t hp) {
    const uint32_t lg = hp & 31u, sh = hp >> 8;
    const uint32_t o2 = min(sh + lg, 31u), o3 = min(sh + 2 * lg, 31u);
    return __builtin_amdgcn_ubfe(key, sh, lg) ^ __builtin_amdgcn_ubfe(key, o2, lg) ^
           __builtin_amdgcn_ubfe(key, o3, lg);
}

// Exclusive scan of the nb (a power of two, at most 8*JB) bucket counters
// `cnt` into 16-bit starts bst[0..nb]; bst[nb] = total.  Each lane takes
// nb/JB consecutive counters (at least one), read in one batch.  Ends with a
// barrier.
template <int JB>
__device__ __forceinline__ void bucket_scan(const uint32_t *cnt, uint16_t *bst, uint32_t nb,
                                            uint32_t *scan_tmp, uint32_t tid) {
    constexpr int MAXPER = 8;
    const uint32_t per = nb > JB ? nb / JB : 1u;
    const uint32_t b = tid * per;
    uint32_t v[MAXPER];
    uint32_t s = 0;
    unroll<MAXPER>([&](auto j) {
        const bool ok = j < per && b + j < nb;
        const uint32_t x = cnt[ok ? b + j : 0];
        v[j] = ok ? x : 0u;
        s += v[j];
    });
    uint32_t total;
    uint32_t pre = block_exclusive_scan<JB>(s, scan_tmp, &total);
    unroll<MAXPER>([&](auto j) {
        if (j < per && b + j < nb) bst[b + j] = (uint16_t)pre;
        pre += v[j];
    });
    if (tid == 0) bst[nb] = (uint16_t)total;
    __syncthreads();
}

// Counting-sort the m (<= CAP) tuples of R held in registers (element q of
// `rv` is tuple tid + q*JB, valid below m) by bucket into `tab`.  Count
// (counters in `tab`), scan into the 16-bit bounds `bst` (bst[0] = 0, bucket
// h's cursor at bst[h+1], starting at its first slot), then place each tuple
// at a position taken by an LDS atomic on its bucket's 16-bit cursor (two
// cursors share one 32-bit word; no carry crosses the halves since every
// cursor stays below 2^16).  After the build bucket h spans
// [bst[h], bst[h+1]).  Branch-free: lanes without a tuple count into the spare
// counter nb+1, advance the spare cursor bst[nb+2] and store into the spare
// slot tab[CAP].  `sh` is the partition's radix bits.  Returns the hash
// parameters hp = log2(nb) | sh << 8 for bucket_of.
template <int JB, int CAP>
__device__ __forceinline__ uint32_t build_regs(const uint64_t (&rv)[CAP / JB], uint32_t m,
                                               uint32_t sh, uint64_t *tab, uint16_t *bst,
                                               uint32_t *scan_tmp, uint32_t tid) {
    constexpr int ITEMS = CAP / JB;
    static_assert(max_buckets(CAP) <= 8u * JB, "bucket scan takes at most 8 counters per lane");
    const int lg = bucket_lg(m);
    const uint32_t hp = (uint32_t)lg | (sh << 8);
    const uint32_t nb = 1u << lg;
    uint32_t *cnt = reinterpret_cast<uint32_t *>(tab);
    for (uint32_t i = threadIdx.x; i < nb + 2; i += JB) cnt[i] = 0;
    if (threadIdx.x == 0) {
        bst[0] = 0;
        bst[nb + 2] = 0;
    }
    __syncthreads();
    // items no lane of this wave holds are skipped (a wave-uniform branch);
    // in a partly filled item, lanes without a tuple use the spare counter,
    // cursor and slot
    const uint32_t wave0 = __builtin_amdgcn_readfirstlane(tid);
    uint32_t hq[ITEMS];
    unroll<ITEMS>([&](auto q) {
        if (wave0 + q * JB < m) {
            hq[q] = tid + q * JB < m ? bucket_of(key_of(rv[q]), hp) : nb + 1;
            atomicAdd(&cnt[hq[q]], 1u);
        }
    });
    __syncthreads();
    bucket_scan<JB>(cnt, bst + 1, nb, scan_tmp, tid);
    uint32_t *cur = reinterpret_cast<uint32_t *>(bst);
    unroll<ITEMS>([&](auto q) {
        if (wave0 + q * JB < m) {
            const uint32_t c = hq[q] + 1;  // the bucket's cursor
            const uint32_t half = (c & 1u) * 16;
            const uint32_t old = atomicAdd(&cur[c >> 1], 1u << half);
            tab[tid + q * JB < m ? (old >> half) & 0xffffu : (uint32_t)CAP] = rv[q];
        }
    });
    __syncthreads();
    return hp;
}

// Probe S tuples held in registers against the round's table (bucket ends in
// `bst`): element j of `sv` is tuple tid + (Q0 + j)*JB of the task's slice
// (valid below ns).
// PGROUP tuples of a lane advance together through their buckets: each step
// reads one candidate per tuple (unconditionally, clamped inside the table)
// and then compares, so a step costs one LDS round trip.
template <int JB, int CAP, bool WRITE, int Q0, int N, int G0 = 0>
__device__ __forceinline__ void probe_part(const uint64_t (&sv)[N], uint32_t ns,
                                           const uint64_t *tab, const uint16_t *bst, uint32_t hp,
                                           uint64_t &cnt, uint64_t &sum, uint64_t *pairs,
                                           uint64_t &wpos, uint64_t cap, uint32_t tid,
                                           int records = 0, uint32_t tnow = 0) {
    if constexpr (G0 < N) {
        constexpr int G = N - G0 < PGROUP ? N - G0 : PGROUP;
        // groups that no lane of this wave holds are skipped (wave-uniform)
        if (__builtin_amdgcn_readfirstlane(tid) + (uint32_t)(Q0 + G0) * JB >= ns) return;
        uint32_t be[G];  // bucket cursor | end << 16
        unroll<G>([&](auto j) {
            const uint32_t h = bucket_of(key_of(sv[G0 + j]), hp);
            const uint32_t lo = bst[h], hi = bst[h + 1];
            be[j] = tid + (uint32_t)(Q0 + G0 + j) * JB < ns ? lo | (hi << 16) : 0u;
        });
        for (;;) {
            uint64_t r[G];
            unroll<G>([&](auto j) { r[j] = tab[umin32(be[j] & 0xffffu, CAP - 1)]; });
            bool more = false;
            unroll<G>([&](auto j) {
                const uint32_t x = be[j] & 0xffffu, e = be[j] >> 16;
                if (x < e) {
                    if (key_of(r[j]) == key_of(sv[G0 + j])) {
                        const uint64_t v = (r[j] << 32) | payload_of(sv[G0 + j]);
                        if (WRITE) {
                            if (wpos < cap)
                                pairs[wpos] = records == 2 ? ((uint64_t)tnow << 32) | payload_of(sv[G0 + j])
                                              : records ? sv[G0 + j] : v;
                            ++wpos;
                        } else {
                            ++cnt;
                            sum += v;
                        }
                    }
                    be[j] += 1;
                    more |= x + 1 < e;
                }
            });
            if (!more) break;
        }
        probe_part<JB, CAP, WRITE, Q0, N, G0 + PGROUP>(sv, ns, tab, bst, hp, cnt, sum, pairs,
                                                       wpos, cap, tid, records, tnow);
    }
}

// Counting probe (no pair output): PGROUP S tuples of a lane step through their
// buckets together, for as many steps as the longest bucket of any tuple in the
// wave (a wave-uniform trip count, so the loop control is scalar and the body is
// branch-free).  A tuple whose bucket [lo, hi) is shorter reads the `steps`
// slots from min(lo, m - steps): they cover its bucket and never leave the
// round's m tuples, and a slot outside the bucket never holds the probing key
// (every R tuple with that key is in the bucket), so each read needs neither a
// clamp nor a bound test, only the key compare.  Matches of tuple j are counted
// in mt[j]; the checksum is rebuilt from its two halves: the sum of R payloads
// (mod 2^32, the high word) and mt[j] * S payload.  Lanes without a tuple
// (past ns) are masked out of both.
template <int JB, int CAP, int Q0, int N, int G0 = 0>
__device__ __forceinline__ void probe_count(const uint64_t (&sv)[N], uint32_t ns, uint32_t m,
                                            const uint64_t *tab, const uint16_t *bst, uint32_t hp,
                                            uint64_t &cnt, uint64_t &sum, uint32_t tid) {
    if constexpr (G0 < N) {
        constexpr int G = N - G0 < PGROUP ? N - G0 : PGROUP;
        // groups that no lane of this wave holds are skipped (wave-uniform)
        if (__builtin_amdgcn_readfirstlane(tid) + (uint32_t)(Q0 + G0) * JB >= ns) return;
        // candidates are read as {payload, key} word pairs so that the key test
        // is one 32-bit compare
        const uint2 *tab2 = reinterpret_cast<const uint2 *>(tab);
        uint32_t lo[G], sk[G];
        bool ok[G];
        uint32_t mx = 0;
        unroll<G>([&](auto j) {
            uint32_t k = key_of(sv[G0 + j]);
            asm volatile("" : "+v"(k));  // a 32-bit value to the compiler: 32-bit key tests
            sk[j] = k;
            const uint32_t h = bucket_of(k, hp);
            const uint32_t b = bst[h], e = bst[h + 1];
            ok[j] = tid + (uint32_t)(Q0 + G0 + j) * JB < ns;
            lo[j] = b;
            const uint32_t len = ok[j] ? e - b : 0u;
            mx = mx > len ? mx : len;
        });
        const uint32_t steps = __builtin_amdgcn_readfirstlane(wave_max_u32_dpp(mx));
        const uint32_t lim = m - steps;  // steps <= the longest bucket <= m
        const uint2 *q[G];
        unroll<G>([&](auto j) { q[j] = tab2 + (lo[j] < lim ? lo[j] : lim); });
        uint32_t mt[G];
        uint32_t rsum = 0;
        unroll<G>([&](auto j) { mt[j] = 0; });
        uint32_t i = 0;
        for (; i + 2 <= steps; i += 2) {  // two slots per read pair
            uint2 r0[G], r1[G];
            unroll<G>([&](auto j) {
                r0[j] = q[j][0];
                r1[j] = q[j][1];
                q[j] += 2;
            });
            unroll<G>([&](auto j) {
                const bool h0 = (int)ok[j] & (int)(r0[j].y == sk[j]), h1 = (int)ok[j] & (int)(r1[j].y == sk[j]);
                mt[j] += (h0 ? 1u : 0u) + (h1 ? 1u : 0u);
                rsum += (h0 ? r0[j].x : 0u) + (h1 ? r1[j].x : 0u);
            });
        }
        if (i < steps) {
            uint2 r0[G];
            unroll<G>([&](auto j) { r0[j] = q[j][0]; });
            unroll<G>([&](auto j) {
                const bool h0 = (int)ok[j] & (int)(r0[j].y == sk[j]);
                mt[j] += h0 ? 1u : 0u;
                rsum += h0 ? r0[j].x : 0u;
            });
        }
        unroll<G>([&](auto j) {
            cnt += mt[j];
            sum += (uint64_t)mt[j] * payload_of(sv[G0 + j]);
        });
        sum += (uint64_t)rsum << 32;
        probe_count<JB, CAP, Q0, N, G0 + PGROUP>(sv, ns, m, tab, bst, hp, cnt, sum, tid);
    }
}

// Direct table of a round: when its m R tuples fall into distinct buckets among
// 2^DLG (the largest power of two <= CAP, so m <= 2^DLG), the table is the
// bucket array itself, one tuple per slot, and a probe reads one slot.  Keys
// that are dense above the partition bits (a permutation, as in the
// reference's unique-key relations) fill it exactly: the reference's
// bucket_chaining_join gets one tuple per bucket for them too, with
// next_pow2(|R_p|) buckets (parallel_radix_join.cpp:300-312).  An empty slot h
// holds a key whose bucket is h ^ 1, so no key probing slot h equals it; this
// needs the partition bits plus DLG to stay within 31 bits.
template <int CAP> __host__ __device__ constexpr int direct_lg() {
    int l = 0;
    while ((2 << l) <= CAP) ++l;
    return l;
}

__device__ __forceinline__ uint64_t direct_empty(uint32_t h, uint32_t sh) {
    return (uint64_t)((h ^ 1u) << sh) << 32;
}

// Places the round's tuples (element q of rv is tuple tid + q*JB, valid below
// m) into their slots with one LDS compare-and-swap each.  A tuple whose slot
// already holds its key (a duplicate) adds its payload to the slot's payload
// half and 1 to the slot's 8-bit count of further tuples (`ext`, 2^DLG bytes
// packed in words; read only when a duplicate was seen).  Returns, the same for
// every lane, 0 when two different keys met in a slot or a count passed 255
// (the caller builds the bucket table instead), 1 for a table of distinct
// keys, 2 for one with duplicates.
template <int JB, int CAP>
__device__ __forceinline__ uint32_t build_direct(const uint64_t (&rv)[CAP / JB], uint32_t m,
                                                 uint32_t sh, uint64_t *tab, uint32_t *ext,
                                                 uint32_t *s_flag, uint32_t tid) {
    constexpr int ITEMS = CAP / JB;
    constexpr int DLG = direct_lg<CAP>();
    constexpr uint32_t NSLOT = 1u << DLG;
    for (uint32_t i = tid; i < NSLOT; i += JB) tab[i] = direct_empty(i, sh);
    for (uint32_t i = tid; i < NSLOT / 4; i += JB) ext[i] = 0;
    if (tid == 0) *s_flag = 0;
    __syncthreads();
    const uint32_t hp = (uint32_t)DLG | (sh << 8);
    const uint32_t wave0 = __builtin_amdgcn_readfirstlane(tid);
    uint32_t fl = 0;  // 1: clash, 2: duplicate
    unroll<ITEMS>([&](auto q) {
        if (wave0 + q * JB < m && tid + q * JB < m) {
            const uint32_t k = key_of(rv[q]);
            const uint32_t h = bucket_of(k, hp);
            const unsigned long long e = direct_empty(h, sh);
            const unsigned long long old = atomicCAS(
                reinterpret_cast<unsigned long long *>(tab + h), e, (unsigned long long)rv[q]);
            if (old != e) {
                if ((uint32_t)(old >> 32) == k) {
                    uint32_t *slot = reinterpret_cast<uint32_t *>(tab + h);
                    atomicAdd(slot, payload_of(rv[q]));
                    const uint32_t sb = 8 * (h & 3u);
                    const uint32_t c = atomicAdd(ext + (h >> 2), 1u << sb);
                    fl |= ((c >> sb) & 0xffu) == 0xffu ? 1u : 2u;
                } else {
                    fl |= 1u;
                }
            }
        }
    });
    if (fl) atomicOr(s_flag, fl);
    __syncthreads();
    const uint32_t f = *s_flag;
    return (f & 1u) ? 0u : (f ? 2u : 1u);
}

// Counting probe of S tuples held in registers (element j of `sv` is tuple
// tid + (Q0 + j)*JB of the slice, valid below ns) against a direct table: one
// slot read and one 32-bit key test per tuple (with DUPS also the slot's count
// of further tuples); the checksum as in probe_count.
template <int JB, int CAP, bool DUPS, int Q0, int N>
__device__ __forceinline__ void probe_direct(const uint64_t (&sv)[N], uint32_t ns,
                                             const uint64_t *tab, const uint32_t *ext,
                                             uint32_t sh, uint64_t &cnt, uint64_t &sum,
                                             uint32_t tid) {
    constexpr int DLG = direct_lg<CAP>();
    if (__builtin_amdgcn_readfirstlane(tid) + (uint32_t)Q0 * JB >= ns) return;
    const uint32_t hp = (uint32_t)DLG | (sh << 8);
    const uint2 *tab2 = reinterpret_cast<const uint2 *>(tab);
    uint2 r[N];
    uint32_t sk[N], hs[N];
    unroll<N>([&](auto j) {
        uint32_t k = key_of(sv[j]);
        asm volatile("" : "+v"(k));  // a 32-bit value to the compiler: 32-bit key tests
        sk[j] = k;
        hs[j] = bucket_of(k, hp);
        r[j] = tab2[hs[j]];
    });
    uint32_t mt = 0, rsum = 0;
    uint64_t ssum = 0;
    unroll<N>([&](auto j) {
        const bool h = (int)(tid + (uint32_t)(Q0 + j) * JB < ns) & (int)(r[j].y == sk[j]);
        uint32_t c = 1;
        if constexpr (DUPS) c += (ext[hs[j] >> 2] >> (8 * (hs[j] & 3u))) & 0xffu;
        c = h ? c : 0u;
        mt += c;
        rsum += h ? r[j].x : 0u;
        ssum += (uint64_t)c * payload_of(sv[j]);
    });
    cnt += mt;
    sum += ssum + ((uint64_t)rsum << 32);
}

// Tuples tid + (Q0 + j)*JB, j < N, of `p` (n >= 1 tuples) into registers.
// The loads are unconditional so that all of them are in flight at once: an
// index at or past n reads tuple n-1 again, and callers ignore those elements.
template <int JB, int Q0, int N>
__device__ __forceinline__ void load_part(uint64_t (&v)[N], const uint64_t *p, uint32_t n,
                                          uint32_t tid) {
    unroll<N>([&](auto j) {
        const uint32_t i = tid + (uint32_t)(Q0 + j) * JB;
        v[j] = __builtin_nontemporal_load(p + umin32(i, n - 1));
    });
}

// A task record is two uint4: {r0 lo, r0 hi, |R_p|, segment}, {s0 lo, s0 hi,
// |S slice|, radix bits of the partition}; everything a worker needs in one
// independent load.
struct TaskRec {
    uint64_t r0, s0;
    uint32_t nr, ns, seg, sh;
};
__device__ __forceinline__ TaskRec decode_task(uint4 a, uint4 b) {
    TaskRec r;
    r.r0 = ((uint64_t)a.y << 32) | a.x;
    r.nr = a.z;
    r.seg = a.w;
    r.s0 = ((uint64_t)b.y << 32) | b.x;
    r.ns = b.z;
    r.sh = b.w;
    return r;
}

template <int JB, int CAP, bool PAIRS>
__global__ __launch_bounds__(JB, 4) void k_join(JoinArgs a) {
    constexpr int ITEMS = CAP / JB;
    static_assert(ITEMS * JB == CAP, "CAP must be a multiple of the work-group size");
    extern __shared__ uint64_t jsm[];
    uint64_t *tab = jsm;  // CAP tuples in bucket order + a spare slot
    uint16_t *bst = reinterpret_cast<uint16_t *>(jsm + CAP + 1);  // max_buckets(CAP) + 3 bounds
    // a direct table's 8-bit counts of further tuples per slot, in bst's place
    uint32_t *ext = reinterpret_cast<uint32_t *>(bst);
    static_assert((1u << direct_lg<CAP>()) <= 2 * (max_buckets(CAP) + 4),
                  "a direct table's counts fit the bucket bounds' space");
    __shared__ uint64_t red[2 * (JB / 64)];
    __shared__ uint32_t scan_tmp[JB / 64];
    __shared__ uint64_t s_base;
    __shared__ uint32_t s_dflag;
    constexpr int NW = JB / 64;
    constexpr int DLG = direct_lg<CAP>();
    // rounds left before the direct table is tried again after a clash (so
    // that keys that never fit it cost one attempt in 16 rounds)
    uint32_t dskip = 0;

    // Static round-robin over tasks (a shared device-wide task counter costs one
    // serialised atomic per task).  The S slice (at most CAP tuples) is read
    // in two halves: the first half of the next task's slice is loaded while
    // this task finishes, the second half after the build, in flight while the
    // first half probes.
    constexpr int SA = ITEMS / 2, SB = ITEMS - SA;
    const uint32_t total = a.task_base[a.max_parts];
    uint32_t t = blockIdx.x;
    if (t >= total) return;
    TaskRec cur = decode_task(a.task_info[2 * t], a.task_info[2 * t + 1]);
    uint64_t sa[SA];
    load_part<JB, 0, SA>(sa, a.S + cur.s0, umin32(cur.ns, CAP), opaque_tid());
    for (;;) {
        const uint32_t tn = t + gridDim.x;
        const bool has_next = tn < total;
        // the last task reads its own record again, so the loads below need
        // no condition
        const uint32_t tl = has_next ? tn : t;
        const uint4 n0 = a.task_info[2 * tl], n1 = a.task_info[2 * tl + 1];
        const uint32_t ns = cur.ns;
        uint64_t cnt = 0, sum = 0;
        for (uint64_t rb = cur.r0; rb < cur.r0 + cur.nr; rb += CAP) {
            // one LDS round of R: CAP tuples (more rounds only for partitions
            // larger than CAP)
            const uint32_t m = (uint32_t)min<uint64_t>(CAP, cur.r0 + cur.nr - rb);
            uint32_t hp = 0;
            uint32_t direct = 0;  // 1: direct table, 2: direct table with duplicates
            {
                uint64_t rv[ITEMS];
                load_part<JB, 0, ITEMS>(rv, a.R + rb, m, opaque_tid());
                if (!PAIRS && m <= (1u << DLG) && cur.sh + DLG <= 31) {
                    if (dskip == 0) {
                        direct = build_direct<JB, CAP>(rv, m, cur.sh, tab, ext, &s_dflag,
                                                       opaque_tid());
                        if (!direct) dskip = 16;
                    } else {
                        --dskip;
                    }
                }
                if (!direct) hp = build_regs<JB, CAP>(rv, m, cur.sh, tab, bst, scan_tmp, opaque_tid());
            }
            // the S slice in sub-slices of CAP (one, unless a heavy partition
            // gives the task more): the first half of the first sub-slice was
            // loaded ahead; it is read again only after a later sub-slice
            // replaced it
            for (uint32_t so = 0; so < ns; so += CAP) {
                const uint32_t sub = umin32(CAP, ns - so);
                if (ns > (uint32_t)CAP && (so > 0 || rb > cur.r0))
                    load_part<JB, 0, SA>(sa, a.S + cur.s0 + so, sub, opaque_tid());
                // second half: in flight while the first half probes
                uint64_t sb[SB];
                load_part<JB, SA, SB>(sb, a.S + cur.s0 + so, sub, opaque_tid());
                uint64_t c = 0, x = 0, wpos = 0;
                if (!PAIRS && direct == 1) {
                    probe_direct<JB, CAP, false, 0, SA>(sa, sub, tab, ext, cur.sh, c, x, opaque_tid());
                    probe_direct<JB, CAP, false, SA, SB>(sb, sub, tab, ext, cur.sh, c, x, opaque_tid());
                } else if (!PAIRS && direct == 2) {
                    probe_direct<JB, CAP, true, 0, SA>(sa, sub, tab, ext, cur.sh, c, x, opaque_tid());
                    probe_direct<JB, CAP, true, SA, SB>(sb, sub, tab, ext, cur.sh, c, x, opaque_tid());
                } else {
                    probe_count<JB, CAP, 0, SA>(sa, sub, m, tab, bst, hp, c, x, opaque_tid());
                    probe_count<JB, CAP, SA, SB>(sb, sub, m, tab, bst, hp, c, x, opaque_tid());
                }
                if (PAIRS) {
                    uint32_t total_c;
                    const uint32_t pre = block_exclusive_scan<JB>((uint32_t)c, scan_tmp, &total_c);
                    if (threadIdx.x == 0)
                        s_base = atomicAdd((unsigned long long *)a.pair_cursor,
                                           (unsigned long long)total_c);
                    __syncthreads();
                    wpos = s_base + pre;
                    uint64_t dc = 0, dx = 0;
                    // timed records: the device's constant-rate clock when this
                    // slice's matches are produced (low 32 bits)
                    const uint32_t tnow =
                        a.records == 2 ? (uint32_t)__builtin_amdgcn_s_memrealtime() : 0u;
                    probe_part<JB, CAP, true, 0, SA>(sa, sub, tab, bst, hp, dc, dx, a.pairs, wpos,
                                                    a.pairs_cap, opaque_tid(), a.records, tnow);
                    probe_part<JB, CAP, true, SA, SB>(sb, sub, tab, bst, hp, dc, dx, a.pairs,
                                                     wpos, a.pairs_cap, opaque_tid(), a.records,
                                                     tnow);
                }
                cnt += c;
                sum += x;
            }
            __syncthreads();  // the next round rebuilds tab/bst
        }
        const TaskRec nxt = decode_task(n0, n1);
        // in flight during the reduction and the next task's build
        load_part<JB, 0, SA>(sa, a.S + nxt.s0, umin32(nxt.ns, CAP), opaque_tid());
        // count and checksum reduced together: one barrier pair per task
        cnt = wave_sum(cnt);
        sum = wave_sum(sum);
        const int wid = threadIdx.x / 64;
        if (lane_id() == 0) {
            red[wid] = cnt;
            red[NW + wid] = sum;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t c = 0, x = 0;
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                c += red[w];
                x += red[NW + w];
            }
            // plain store of the task's partial result; k_seg_reduce adds them
            // up per segment (one shared accumulator would serialise every task)
            a.task_res[t] = make_ulonglong2(c, x);
        }
        __syncthreads();  // red[] is reused by the next task
        if (!has_next) break;
        t = tn;
        cur = nxt;
    }
}

size_t join_scan_tmp_bytes(uint64_t max_parts) { return scan_tmp_bytes(max_parts + 1, 4); }

uint32_t join_task_split(uint32_t split, int cap) {
    return split ? split : 8u * (uint32_t)cap;
}

uint64_t join_max_tasks(uint64_t max_parts, uint64_t max_S, uint32_t split) {
    return max_parts + (split ? max_S / split : max_S) + 1;
}

template <int JB, int CAP, bool PAIRS>
static int launch_join(const JoinArgs &a, int num_cu, hipStream_t st) {
    const size_t lds = join_lds_bytes(CAP);
    AJ_TRY(lds_limit_set((const void *)k_join<JB, CAP, PAIRS>, lds));
    // resident work-groups per CU, per device (LDS and registers decide it)
    static std::mutex mu;
    static int per_cu_dev[64] = {};
    int dev = 0;
    AJ_HIP(hipGetDevice(&dev));
    int per_cu;
    {
        std::lock_guard<std::mutex> g(mu);
        per_cu = dev >= 0 && dev < 64 ? per_cu_dev[dev] : 0;
    }
    if (!per_cu) {
        AJ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_join<JB, CAP, PAIRS>, JB, lds));
        if (per_cu < 1) per_cu = 1;
        std::lock_guard<std::mutex> g(mu);
        if (dev >= 0 && dev < 64) per_cu_dev[dev] = per_cu;
    }
    hipLaunchKernelGGL((k_join<JB, CAP, PAIRS>), dim3(num_cu * per_cu), dim3(JB), lds, st, a);
    return AJ_OK;
}

template <int JB, int CAP>
static int launch_join2(const JoinArgs &a, bool pairs, int num_cu, hipStream_t st) {
    return pairs ? launch_join<JB, CAP, true>(a, num_cu, st) : launch_join<JB, CAP, false>(a, num_cu, st);
}

int run_join(const JoinDesc &d, hipStream_t st, const Marks *mk, int m0) {
    if (d.split == 0 || d.max_parts >= (1ull << 31) || !d.task_info) {
        set_error(AJ_ERR_ARG, "bad join split/partition count (split %u, cap %d)", d.split,
                  d.cap);
        return AJ_ERR_ARG;
    }
    JoinArgs a;
    a.R = d.R;
    a.S = d.S;
    a.roff = d.roff;
    a.soff = d.soff;
    a.rlen = d.rlen;
    a.slen = d.slen;
    a.part_seg_base = d.part_seg_base;
    a.nseg = d.nseg;
    a.max_parts = d.max_parts;
    a.nreal = d.nreal;
    a.split = d.split;
    a.seg_count = d.seg_count;
    a.seg_sum = d.seg_sum;
    a.pairs = d.pairs;
    a.pairs_cap = d.pairs_cap;
    a.records = d.records;
    a.hash_shift = d.hash_shift;
    a.seg_bits = d.seg_bits;
    a.pair_cursor = d.pair_cursor;
    a.task_cnt = d.task_cnt;
    a.task_base = d.task_base;
    a.task_next = d.task_next;
    a.task_info = d.task_info;
    a.task_res = d.task_res;
    if (!d.task_res) {
        set_error(AJ_ERR_ARG, "join workspace without task results");
        return AJ_ERR_ARG;
    }
    if (!d.tasks_ready) {
        const unsigned g = (unsigned)std::min<uint64_t>((d.max_parts + 256) / 256, 4096);
        hipLaunchKernelGGL(k_task_count, dim3(g), dim3(256), 0, st, a);
        AJ_TRY(exclusive_scan_u32(d.task_cnt, d.task_base, d.max_parts + 1, d.scan_tmp,
                                  d.scan_tmp_bytes, st));
        const unsigned gt = (unsigned)std::min<uint64_t>((d.max_tasks + 255) / 256, 16384);
        hipLaunchKernelGGL(k_task_fill, dim3(gt), dim3(256), 0, st, a);
    }
    if (mk) mk->mark(m0);
    const bool pairs = d.pairs != nullptr;
    int rc;
    switch (d.cap) {
        case 2048: rc = launch_join2<256, 2048>(a, pairs, d.num_cu, st); break;
        case 4096: rc = launch_join2<256, 4096>(a, pairs, d.num_cu, st); break;
        case 8192: rc = launch_join2<512, 8192>(a, pairs, d.num_cu, st); break;
        case 8704: rc = launch_join2<512, 8704>(a, pairs, d.num_cu, st); break;
        case 12288: rc = launch_join2<1024, 12288>(a, pairs, d.num_cu, st); break;
        default:
            set_error(AJ_ERR_ARG, "cap must be 2048, 4096, 8192, 8704 or 12288 (got %d)", d.cap);
            return AJ_ERR_ARG;
    }
    if (rc) return rc;
    if (mk) mk->mark(m0 + 1);
    hipLaunchKernelGGL(k_seg_reduce, dim3((unsigned)std::min(d.nseg, 4096)), dim3(RBLOCK), 0, st,
                       a);
    AJ_HIP(hipGetLastError());
    return AJ_OK;
}

}  // namespace aj
