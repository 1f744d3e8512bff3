// npo.hip -- no-partitioning join on one HBM-resident hash table (NPO).
//
// Reference: NPO builds one shared chained table over all of R with a latch
// per bucket (build_hashtable_mt, common_functions.cpp:350-400) and probes it
// with every S tuple, counting each key-equal entry of the bucket chain
// (probe_hashtable_single, common_functions.cpp:237-280).
//
// CDNA4 form: open addressing over 16-byte slots, one slot per distinct key,
// linear probing, capacity = next power of two >= 2|R| (at most half the slots
// hold a key).  Word 0 of a slot is the key's first tuple (key << 32 |
// payload), claimed with one 64-bit atomicCAS (the latch and the slot write of
// the reference in one instruction).  Word 1 aggregates the key's further R
// tuples, (sum of their payloads mod 2^32) << 32 | their count, with one 64-bit
// atomicAdd each.  A counting probe that finds its key adds count = 1 + the
// further count and checksum = (sum of the key's R payloads) << 32 + count * S
// payload -- exactly the sum of (R.payload << 32 | S.payload) over its matches
// mod 2^64 -- and stops: neither an insert nor a probe walks past the copies of
// a key, so duplicates cost O(1) per tuple.  Joins that write pairs or records
// also chain the further tuples (R index + 1, head per slot in `head`, links in
// `next`, pushed with one atomicExch) and read their payloads from R.  The
// empty marker is the all-ones word; an R tuple with exactly that bit pattern
// (key -1, payload -1) is counted on the side and matched arithmetically, so
// every 64-bit tuple value stays joinable.
//
// The home slot is a murmur3-finalised key: the reference's `key & mask`
// (common_functions.h:57-59) would turn dense keys into one long run under
// linear probing.
//
// Building: random device-wide atomics run at about 13 G CAS/s over a table of
// this size (each is a round trip to memory), while the same table written
// out sequentially from LDS takes about a tenth of that time.  So for
// counting joins of large R the table is cut into regions of 4096 or 8192
// slots (NpoGeom: the hash picks the region and the slot in it, probing wraps
// inside the region), R is partitioned by hashed key into the regions with
// the staged two passes (keys replaced by fmix32(key) in pass 1: a bijection,
// undone here), and one work-group per region builds it in LDS -- inserts and
// aggregates are LDS atomics -- and writes all its slots out with coalesced
// 16-byte stores, so no separate fill of the table is needed.  The
// partitions are sized far below a region's slots, so a region never fills;
// a partition that overflows its capacity (heavy duplicates) sets a flag, and
// the flat build then runs on the device in its place (its kernels return at
// once when the flag is clear) and the probe reads the flat geometry.  The
// probe is the same either way: one random slot per S tuple.
//
// Bytes per tuple: R read 8 + slot write 16 (flat build), or R read 8 +
// partitions written 8 + read 8 + 32 of table slots (region build, every slot
// written once); S read 8 + slot read 16 (one slot per probe for unique keys
// at half load).
#include "aj_internal.h"
#include "device_common.h"

namespace aj {

constexpr int NBLOCK = 256;

__global__ __launch_bounds__(NBLOCK) void k_fill_u64(uint64_t *p, uint64_t n, uint64_t v) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ __launch_bounds__(NBLOCK) void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

static unsigned grid_for(uint64_t n, int num_cu) {
    uint64_t g = (n + NBLOCK - 1) / NBLOCK;
    uint64_t cap = (uint64_t)num_cu * 16;
    if (g > cap) g = cap;
    return (unsigned)(g ? g : 1);
}

// Plain streaming copy of 16-byte words (the traffic-calibration kernel).
__global__ __launch_bounds__(NBLOCK) void aj_copy_kernel(const uint4 *src, uint4 *dst, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

int copy_d2d(void *dst, const void *src, uint64_t bytes, int num_cu, hipStream_t st) {
    if ((bytes & 15) || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15)) {
        set_error(AJ_ERR_ARG, "aj_copy_d2d needs 16-byte aligned pointers and sizes");
        return AJ_ERR_ARG;
    }
    const uint64_t n = bytes / 16;
    if (n)
        hipLaunchKernelGGL(aj_copy_kernel, dim3(grid_for(n, num_cu)), dim3(NBLOCK), 0, st,
                           (const uint4 *)src, (uint4 *)dst, n);
    AJ_HIP(hipGetLastError());
    return AJ_OK;
}

// Random-access calibration: lane i loads the 16-byte slot fmix32(i + 1) mod
// the table size (KIND 0), claims slot fmix32((i + 1) ^ 0x5bd1e995) with a
// 64-bit CAS of an empty word (KIND 1), or both at once (KIND 2: the pattern
// of an SHJ step).  The loaded words are kept live through a store that the
// sum never triggers.
template <int KIND>
__global__ __launch_bounds__(NBLOCK) void k_access(uint64_t *tab, uint64_t mask, uint64_t n) {
    typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));
    uint64_t acc = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t k = (uint32_t)(i + 1);
        if (KIND != 1) {
            const u64x2 w = *reinterpret_cast<const u64x2 *>(tab + 2 * (fmix32(k) & mask));
            acc += w.x + w.y;
        }
        if (KIND != 0)
            acc += atomicCAS((unsigned long long *)&tab[2 * (fmix32(k ^ 0x5bd1e995u) & mask)],
                             (unsigned long long)EMPTY_SLOT,
                             (unsigned long long)(((uint64_t)k << 32) | (uint32_t)i));
    }
    if (acc == 0x5A5A5A5A5A5A5A5Aull) tab[1] = acc;
}

int access_kernel(int kind, uint64_t *tab, uint64_t slots, uint64_t n, int num_cu,
                  hipStream_t st) {
    if (kind < 0 || kind > 2 || slots < 2 || (slots & (slots - 1))) {
        set_error(AJ_ERR_ARG, "access kernel: kind 0..2 and a power-of-two slot count");
        return AJ_ERR_ARG;
    }
    const dim3 g(grid_for(n, num_cu)), b(NBLOCK);
    if (kind == 0) hipLaunchKernelGGL(k_access<0>, g, b, 0, st, tab, slots - 1, n);
    if (kind == 1) hipLaunchKernelGGL(k_access<1>, g, b, 0, st, tab, slots - 1, n);
    if (kind == 2) hipLaunchKernelGGL(k_access<2>, g, b, 0, st, tab, slots - 1, n);
    AJ_HIP(hipGetLastError());
    return AJ_OK;
}

int fill_u64(uint64_t *p, uint64_t n, uint64_t v, hipStream_t st) {
    if (!n) return AJ_OK;
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, 256)), dim3(NBLOCK), 0, st, p, n, v);
    AJ_HIP(hipGetLastError());
    return AJ_OK;
}

int fill_u32(uint32_t *p, uint64_t n, uint32_t v, hipStream_t st) {
    if (!n) return AJ_OK;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256)), dim3(NBLOCK), 0, st, p, n, v);
    AJ_HIP(hipGetLastError());
    return AJ_OK;
}

// Empty slots: word 0 all ones, aggregate 0.  With only_if, only when the
// region build overflowed (*only_if set), and then the sentinel count that
// build may have added is cleared too.
__global__ __launch_bounds__(NBLOCK) void k_fill_slots(uint4 *t, uint64_t n,
                                                     const uint32_t *only_if, uint64_t *acc) {
    if (only_if && !*only_if) return;
    if (only_if && blockIdx.x == 0 && threadIdx.x == 0) acc[2] = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        t[i] = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u);
}

// Insert tuple i (value v, key k): claim k's slot with its first tuple, or add
// v to the slot's aggregate of further tuples (and, when head is set, chain i
// behind the first).
__device__ __forceinline__ void npo_insert(uint64_t i, uint64_t v, uint32_t k, uint64_t *tab,
                                           uint64_t mask, uint32_t *head, uint32_t *next) {
    uint64_t h = fmix32(k) & mask;
    // terminates: fewer distinct keys than slots
    for (;;) {
        const uint64_t old = atomicCAS((unsigned long long *)&tab[2 * h],
                                       (unsigned long long)EMPTY_SLOT, (unsigned long long)v);
        if (old == EMPTY_SLOT) return;
        if (key_of(old) == k) {
            atomicAdd((unsigned long long *)&tab[2 * h + 1],
                      (unsigned long long)(((uint64_t)payload_of(v) << 32) | 1u));
            if (head) next[i] = atomicExch(&head[h], (uint32_t)(i + 1));
            return;
        }
        h = (h + 1) & mask;
    }
}

// Matches of key k in the table's slot for k (geometry g): calls agg(first
// tuple, count of further tuples, sum of their payloads) once if k has a
// slot, and returns the slot's index (~0 if none).
template <class F>
__device__ __forceinline__ uint64_t npo_find(uint32_t k, const uint64_t *tab, NpoGeom g,
                                             F &&agg) {
    typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));
    const uint32_t h = fmix32(k);
    const uint64_t region = ((uint64_t)(h & ((1u << g.b1) - 1u)) << g.b2) |
                            ((h >> g.b1) & ((1u << g.b2) - 1u));
    const uint64_t rmask = (1ull << g.rs_log2) - 1, base = region << g.rs_log2;
    for (uint64_t s = (h >> (g.b1 + g.b2)) & rmask;; s = (s + 1) & rmask) {
        const uint64_t i = base | s;
        const u64x2 w = *reinterpret_cast<const u64x2 *>(tab + 2 * i);  // one 16-byte slot
        if (w.x == EMPTY_SLOT) return ~0ull;
        if (key_of(w.x) == k) {
            agg(w.x, (uint32_t)w.y, (uint32_t)(w.y >> 32));
            return i;
        }
    }
}

// The flat build (with only_if: only when the region build overflowed).
__global__ __launch_bounds__(NBLOCK) void k_npo_build(const uint64_t *R, uint64_t n,
                                                    uint64_t *tab, uint64_t mask,
                                                    uint32_t *head, uint32_t *next,
                                                    uint64_t *acc, const uint32_t *only_if) {
    if (only_if && !*only_if) return;
    uint64_t sentinels = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t v = R[i];
        if (v == EMPTY_SLOT) {
            ++sentinels;
            continue;
        }
        npo_insert(i, v, key_of(v), tab, mask, head, next);
    }
    if (sentinels) atomicAdd((unsigned long long *)&acc[2], (unsigned long long)sentinels);
}

// Region build: work-group p builds region p from partition p of the hashed
// R tuples (parts + p * fcap, fcur[p] of them) -- the region's slots in LDS,
// each tuple's key recovered from its hash, inserts with LDS atomics as in
// npo_insert -- then writes every slot of the region out.  Returns at once when
// the staged partitioning overflowed (*flag); sets *flag itself if a region
// would be more than 15/16 full (the partition capacities keep regions under
// 5/8, so this guards the probe's termination, it does not happen).
template <int RS_LOG2, int RBLOCK>
__global__ __launch_bounds__(RBLOCK) void k_npo_region(const uint64_t *parts, const uint32_t *fcur,
                                                     uint32_t fcap, uint32_t rbits,
                                                     uint64_t *tab, uint64_t *acc,
                                                     uint32_t *flag) {
    constexpr uint32_t RS = 1u << RS_LOG2;
    constexpr int B = 4;  // tuples loaded per lane before their inserts: loads in flight
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    extern __shared__ uint64_t lds[];  // RS slots of 2 words
    __shared__ uint32_t s_sent, s_new, s_bad;
    // the flag, the partition's size and its first tuples are loaded together
    // (the tuple loads stay inside the partition whatever the flag says); the
    // work-group returns once the flag is in
    const uint32_t p = blockIdx.x;
    const uint32_t flagged = *flag;
    const uint32_t n = min(fcur[p], fcap);
    const uint64_t *in = parts + (uint64_t)p * fcap;
    uint64_t t[B];
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const uint32_t i = b * RBLOCK + threadIdx.x;
        t[b] = i < n ? in[i] : 0;
    }
    u32x4 *slots = reinterpret_cast<u32x4 *>(lds);
    for (uint32_t i = threadIdx.x; i < RS; i += RBLOCK)
        slots[i] = u32x4{0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u};
    if (threadIdx.x == 0) s_sent = s_new = s_bad = 0;
    if (flagged) return;  // uniform: every lane read the same word
    __syncthreads();
    uint32_t sent = 0, fresh = 0, bad = 0;
    for (uint32_t i0 = 0; i0 < n; i0 += RBLOCK * B) {
        if (i0) {
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const uint32_t i = i0 + b * RBLOCK + threadIdx.x;
                t[b] = i < n ? in[i] : 0;
            }
        }
#pragma unroll
        for (int b = 0; b < B; ++b) {
            if (i0 + b * RBLOCK + threadIdx.x >= n) continue;
            const uint32_t h = key_of(t[b]), k = fmix32_inv(h);
            const uint64_t v = ((uint64_t)k << 32) | (uint32_t)t[b];
            if (v == EMPTY_SLOT) {
                ++sent;
                continue;
            }
            uint32_t s = (h >> rbits) & (RS - 1);
            for (uint32_t j = 0;; ++j, s = (s + 1) & (RS - 1)) {
                if (j == RS) {  // region full
                    bad = 1;
                    break;
                }
                const uint64_t old = atomicCAS((unsigned long long *)&lds[2 * s],
                                               (unsigned long long)EMPTY_SLOT,
                                               (unsigned long long)v);
                if (old == EMPTY_SLOT) {
                    ++fresh;
                    break;
                }
                if (key_of(old) == k) {
                    atomicAdd((unsigned long long *)&lds[2 * s + 1],
                              (unsigned long long)(((uint64_t)(uint32_t)t[b] << 32) | 1u));
                    break;
                }
            }
        }
    }
    sent = wave_sum(sent);
    fresh = wave_sum(fresh);
    if (lane_id() == 0) {
        if (sent) atomicAdd(&s_sent, sent);
        if (fresh) atomicAdd(&s_new, fresh);
    }
    if (bad) s_bad = 1;
    __syncthreads();
    if (s_bad || s_new > RS - RS / 16) {
        if (threadIdx.x == 0) atomicOr(flag, 1u);
        return;
    }
    // the table is read next by random probes, not from cache: stream it out
    u32x4 *out = reinterpret_cast<u32x4 *>(tab) + ((uint64_t)p << RS_LOG2);
    for (uint32_t i = threadIdx.x; i < RS; i += RBLOCK) __builtin_nontemporal_store(slots[i], out + i);
    if (threadIdx.x == 0 && s_sent)
        atomicAdd((unsigned long long *)&acc[2], (unsigned long long)s_sent);
}

// Count and checksum of S tuple s: O(1) from the key's aggregate.
__device__ __forceinline__ void npo_count_one(uint64_t s, const uint64_t *tab, NpoGeom g,
                                              uint64_t nsent, uint64_t &cnt, uint64_t &sum) {
    const uint32_t k = key_of(s);
    npo_find(k, tab, g, [&](uint64_t first, uint32_t more, uint32_t more_sum) {
        const uint64_t c = 1ull + more;
        const uint32_t rsum = payload_of(first) + more_sum;  // R payloads, mod 2^32
        cnt += c;
        sum += ((uint64_t)rsum << 32) + c * payload_of(s);
    });
    if (k == 0xFFFFFFFFu && nsent) {
        cnt += nsent;
        sum += nsent * pair_value(EMPTY_SLOT, s);
    }
}

// Every match of S tuple s written from wpos on (the first tuple, then the
// chain of further ones), as pairs, S records or timed records.
__device__ __forceinline__ void npo_write_one(uint64_t s, const uint64_t *tab, NpoGeom g,
                                              const uint64_t *R, const uint32_t *head,
                                              const uint32_t *next, uint64_t nsent,
                                              uint64_t *pairs, uint64_t &wpos, uint64_t cap,
                                              int records, uint32_t tnow) {
    auto emit = [&](uint64_t r) {
        if (wpos < cap)
            pairs[wpos] = records == 2 ? ((uint64_t)tnow << 32) | (uint32_t)s
                          : records ? s : pair_value(r, s);
        ++wpos;
    };
    const uint32_t k = key_of(s);
    const uint64_t h = npo_find(k, tab, g, [&](uint64_t first, uint32_t, uint32_t) { emit(first); });
    if (h != ~0ull)
        for (uint32_t j = head[h]; j; j = next[j - 1]) emit(R[j - 1]);
    if (k == 0xFFFFFFFFu)
        for (uint64_t q = 0; q < nsent; ++q) emit(EMPTY_SLOT);
}

// The probe; with flag, the table has the flat geometry instead of g when the
// region build overflowed (*flag set).
template <bool PAIRS>
__global__ __launch_bounds__(NBLOCK) void k_npo_probe(const uint64_t *S, uint64_t n,
                                                    const uint64_t *tab, NpoGeom g,
                                                    const uint32_t *flag, const uint64_t *R,
                                                    const uint32_t *head, const uint32_t *next,
                                                    uint64_t *acc, uint64_t *pairs, uint64_t cap,
                                                    int records) {
    __shared__ uint64_t red[NBLOCK / 64];
    __shared__ uint64_t s_base;
    const uint64_t nsent = acc[2];
    if (flag && *flag) {
        g.rs_log2 += g.b1 + g.b2;
        g.b1 = g.b2 = 0;
    }
    uint64_t cnt = 0, sum = 0;
    // uniform trip count per block: every thread reaches the block scans
    for (uint64_t base = blockIdx.x * (uint64_t)NBLOCK; base < n;
         base += (uint64_t)gridDim.x * NBLOCK) {
        const uint64_t i = base + threadIdx.x;
        const bool valid = i < n;
        const uint64_t s = valid ? S[i] : 0;
        uint64_t c = 0, x = 0;
        if (valid) npo_count_one(s, tab, g, nsent, c, x);
        if (PAIRS) {
            uint64_t total;
            const uint64_t pre = block_exclusive_scan<NBLOCK>(c, red, &total);
            if (threadIdx.x == 0 && total)
                s_base = atomicAdd((unsigned long long *)&acc[3], (unsigned long long)total);
            __syncthreads();
            if (valid && c) {
                uint64_t wpos = s_base + pre;
                const uint32_t tnow =
                    records == 2 ? (uint32_t)__builtin_amdgcn_s_memrealtime() : 0u;
                npo_write_one(s, tab, g, R, head, next, nsent, pairs, wpos, cap, records, tnow);
            }
            __syncthreads();
        }
        cnt += c;
        sum += x;
    }
    cnt = block_sum<NBLOCK>(cnt, red);
    sum = block_sum<NBLOCK>(sum, red);
    // per-block partials (acc[8 + 2b], acc[9 + 2b]), added by k_npo_reduce: a
    // shared accumulator would serialise one atomic per block
    if (threadIdx.x == 0) {
        acc[8 + 2 * blockIdx.x] = cnt;
        acc[9 + 2 * blockIdx.x] = sum;
    }
}

__global__ __launch_bounds__(NBLOCK) void k_npo_reduce(uint64_t *acc, unsigned nblocks) {
    __shared__ uint64_t red[NBLOCK / 64];
    uint64_t c = 0, x = 0;
    for (unsigned b = threadIdx.x; b < nblocks; b += NBLOCK) {
        c += acc[8 + 2 * b];
        x += acc[9 + 2 * b];
    }
    c = block_sum<NBLOCK>(c, red);
    x = block_sum<NBLOCK>(x, red);
    if (threadIdx.x == 0) {
        acc[0] = c;
        acc[1] = x;
    }
}

size_t npo_acc_words(int num_cu) { return 8 + 2 * (size_t)num_cu * 16; }

template <int RS_LOG2, int RBLOCK>
static int launch_region(const NpoDesc &d, uint32_t *flag, hipStream_t st) {
    const StagedGeom &g = d.stage->g;
    const size_t lds = (size_t)16 << RS_LOG2;
    AJ_TRY(lds_limit_set((const void *)k_npo_region<RS_LOG2, RBLOCK>, lds));
    hipLaunchKernelGGL((k_npo_region<RS_LOG2, RBLOCK>), dim3(1u << (g.b1 + g.b2)), dim3(RBLOCK),
                       lds, st, (const uint64_t *)d.stage->out[0], (const uint32_t *)d.stage->fcur,
                       g.fcap[0], (uint32_t)(g.b1 + g.b2), d.table, d.acc, flag);
    return AJ_OK;
}

int run_npo(const NpoDesc &d, hipStream_t st, const Marks *mk, int m0) {
    if (d.capacity < 2 || (d.capacity & (d.capacity - 1)) || d.capacity <= d.nR) {
        set_error(AJ_ERR_ARG, "NPO table capacity must be a power of two above |R|");
        return AJ_ERR_ARG;
    }
    if (d.pairs && (!d.head || !d.next)) {
        set_error(AJ_ERR_ARG, "NPO with pairs needs the chain arrays");
        return AJ_ERR_ARG;
    }
    int cap_log2 = 0;
    while ((1ull << cap_log2) < d.capacity) ++cap_log2;
    NpoGeom flat;
    flat.rs_log2 = (uint32_t)cap_log2;
    NpoGeom geom = flat;
    uint32_t *flag = nullptr;
    if (d.stage) {
        const StagedGeom &g = d.stage->g;
        if (d.pairs || d.stage->n[0] != d.nR || d.stage->n[1] != 0 || (g.cache_mode & 16) == 0 ||
            d.geom.b1 != (uint32_t)g.b1 || d.geom.b2 != (uint32_t)g.b2 ||
            (d.geom.rs_log2 != 12 && d.geom.rs_log2 != 13) ||
            d.geom.b1 + d.geom.b2 + d.geom.rs_log2 != (uint32_t)cap_log2 ||
            d.stage->flag != reinterpret_cast<uint32_t *>(d.acc + 4)) {
            set_error(AJ_ERR_ARG, "NPO region build: geometry does not match the table");
            return AJ_ERR_ARG;
        }
        geom = d.geom;
        flag = d.stage->flag;
    }
    uint32_t *head = d.pairs ? d.head : nullptr;  // chains only when matches are written
    if (mk) mk->mark(m0);
    if (d.stage) {
        // hashed partitioning into the regions, the regions in LDS, and the flat
        // build in their place if a partition overflowed (returns at once if not)
        AJ_TRY(run_staged(*d.stage, st));
        if (d.geom.rs_log2 == 12)
            AJ_TRY((launch_region<12, 1024>(d, flag, st)));
        else
            AJ_TRY((launch_region<13, 1024>(d, flag, st)));
    }
    hipLaunchKernelGGL(k_fill_slots, dim3(grid_for(d.capacity, 256)), dim3(NBLOCK), 0, st,
                       reinterpret_cast<uint4 *>(d.table), d.capacity, (const uint32_t *)flag,
                       d.acc);
    if (head) AJ_TRY(fill_u32(head, d.capacity, 0, st));
    if (d.nR)
        hipLaunchKernelGGL(k_npo_build, dim3(grid_for(d.nR, d.num_cu)), dim3(NBLOCK), 0, st, d.R,
                           d.nR, d.table, d.capacity - 1, head, d.next, d.acc,
                           (const uint32_t *)flag);
    if (mk) mk->mark(m0 + 1);
    if (d.nS) {
        const unsigned g = grid_for(d.nS, d.num_cu);  // <= num_cu * 16 (npo_acc_words)
        if (d.pairs)
            hipLaunchKernelGGL(k_npo_probe<true>, dim3(g), dim3(NBLOCK), 0, st, d.S, d.nS,
                               (const uint64_t *)d.table, flat, (const uint32_t *)nullptr, d.R,
                               (const uint32_t *)head, (const uint32_t *)d.next, d.acc, d.pairs,
                               d.pairs_cap, d.records);
        else
            hipLaunchKernelGGL(k_npo_probe<false>, dim3(g), dim3(NBLOCK), 0, st, d.S, d.nS,
                               (const uint64_t *)d.table, geom, (const uint32_t *)flag, d.R,
                               (const uint32_t *)nullptr, (const uint32_t *)nullptr, d.acc,
                               nullptr, (uint64_t)0, 0);
        hipLaunchKernelGGL(k_npo_reduce, dim3(1), dim3(NBLOCK), 0, st, d.acc, g);
    }
    if (mk) mk->mark(m0 + 2);
    AJ_HIP(hipGetLastError());
    return AJ_OK;
}

}  // namespace aj
