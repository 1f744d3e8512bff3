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
//   scan          task bases (flat exclusive scan)
//   k_task_fill   task -> (R range, S slice, segment) records, so a worker
//                 finds its task with one load
//   k_join        persistent work-groups take tasks round-robin (t = wg,
//                 wg + grid, ...).  A task first loads part of its S slice
//                 into registers, then takes R_p
//                 in rounds of `cap` tuples; each round counting-sorts R by
//                 bucket into LDS (bucket starts + sorted tuples, about four
//                 tuples per bucket) and probes every S tuple against one
//                 bucket.  Count and checksum are reduced per work-group and
//                 added once to the segment's 64-bit accumulators.  With pairs
//                 enabled each round probes twice: count, reserve a range with
//                 one atomic, write.
//
// Bytes per tuple: R and S are each read once (S again per extra R round when
// the slice does not fit the registers).
#include <hipcub/hipcub.hpp>

#include "aj_internal.h"
#include "device_common.h"

namespace aj {

constexpr int SREG = 4;  // S tuples per thread loaded ahead of the build
constexpr int JUNROLL = 4;

// Buckets of a round of m tuples: nb = max(1, next_pow2(m) / 4).
__host__ __device__ constexpr uint32_t max_buckets(uint32_t cap) {
    uint32_t p = 1;
    while (p < cap) p *= 2;
    return p >= 4 ? p / 4 : 1;
}
__device__ __forceinline__ int bucket_lg(uint32_t m) {
    int lg = 0;
    while ((4u << lg) < m) ++lg;
    return lg;
}

struct JoinArgs {
    const uint64_t *R, *S, *roff, *soff;
    const uint32_t *part_seg_base;
    int nseg;
    uint64_t max_parts;
    uint32_t split;
    uint64_t *seg_count, *seg_sum;
    uint64_t *pairs;
    uint64_t pairs_cap;
    uint64_t *pair_cursor;
    uint32_t *task_cnt, *task_base, *task_next;
    uint4 *task_info;
    ulonglong2 *task_res;  // per task {count, checksum}
};

// Per segment k: count/checksum = sums over its tasks, which are contiguous:
// [task_base[part_seg_base[k]], task_base[part_seg_base[k+1]]).  One block per
// segment (grid-stride), plain stores: the accumulators need no zeroing.
constexpr int RBLOCK = 256;
__global__ __launch_bounds__(RBLOCK) void k_seg_reduce(JoinArgs a) {
    __shared__ uint64_t red[2 * (RBLOCK / 64)];
    for (int k = blockIdx.x; k < a.nseg; k += gridDim.x) {
        const uint32_t tb = a.task_base[a.part_seg_base[k]];
        const uint32_t te = a.task_base[a.part_seg_base[k + 1]];
        uint64_t c = 0, x = 0;
        for (uint32_t t = tb + threadIdx.x; t < te; t += RBLOCK) {
            const ulonglong2 r = a.task_res[t];
            c += r.x;
            x += r.y;
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
            const uint64_t nr = a.roff[p + 1] - a.roff[p];
            const uint64_t ns = a.soff[p + 1] - a.soff[p];
            if (nr && ns) n = (uint32_t)((ns + a.split - 1) / a.split);
        }
        a.task_cnt[p] = n;
    }
}

__global__ void k_task_fill(JoinArgs a) {
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < a.max_parts;
         p += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t n = a.task_cnt[p];
        if (!n) continue;
        const uint32_t b = a.task_base[p];
        const uint32_t seg =
            a.nseg == 1 ? 0u
                        : (uint32_t)last_le<uint32_t>(a.part_seg_base, (uint64_t)a.nseg + 1,
                                                      (uint32_t)p);
        const uint64_t r0 = a.roff[p], nr = a.roff[p + 1] - r0;
        const uint64_t sb = a.soff[p], se = a.soff[p + 1];
        for (uint32_t j = 0; j < n; ++j) {
            const uint64_t s0 = sb + (uint64_t)j * a.split;
            const uint64_t ns = min<uint64_t>(a.split, se - s0);
            a.task_info[2 * (uint64_t)(b + j)] =
                make_uint4((uint32_t)r0, (uint32_t)(r0 >> 32), (uint32_t)nr, seg);
            a.task_info[2 * (uint64_t)(b + j) + 1] =
                make_uint4((uint32_t)s0, (uint32_t)(s0 >> 32), (uint32_t)ns, 0u);
        }
    }
}

__device__ __forceinline__ uint32_t bucket_of(uint32_t key, int lg) {
    return lg ? (key * 0x9E3779B1u) >> (32 - lg) : 0u;
}

// Counting-sort m (<= CAP) tuples of R into `tab` by bucket; `bst` receives
// nb+1 bucket starts.  Only the ranks stay in registers across the scan; the
// tuples are read again (cache-hot) to place them.  Returns log2(nb).
template <int JB, int CAP>
__device__ __forceinline__ int build_round(const uint64_t *R, uint32_t m, uint64_t *tab,
                                           uint32_t *bst, uint32_t *scan_tmp) {
    constexpr int ITEMS = CAP / JB;
    const int lg = bucket_lg(m);
    const uint32_t nb = 1u << lg;
    const uint32_t *Rk = reinterpret_cast<const uint32_t *>(R) + 1;  // key word of each tuple
    uint32_t k[ITEMS];
#pragma unroll
    for (int q = 0; q < ITEMS; ++q) {
        const uint32_t i = threadIdx.x + q * JB;
        k[q] = i < m ? Rk[2 * i] : 0u;
    }
    for (uint32_t i = threadIdx.x; i <= nb; i += JB) bst[i] = 0;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < ITEMS; ++q) {
        const uint32_t i = threadIdx.x + q * JB;
        if (i < m) k[q] = atomicAdd(&bst[bucket_of(k[q], lg)], 1u);  // key -> rank
    }
    __syncthreads();
    lds_exclusive_scan<JB>(bst, nb + 1, scan_tmp);
#pragma unroll
    for (int q = 0; q < ITEMS; ++q) {
        const uint32_t i = threadIdx.x + q * JB;
        if (i < m) {
            const uint64_t t = R[i];
            tab[bst[bucket_of(key_of(t), lg)] + k[q]] = t;
        }
    }
    __syncthreads();
    return lg;
}

template <bool WRITE>
__device__ __forceinline__ void probe_one(uint64_t s, const uint64_t *tab, const uint32_t *bst,
                                          int lg, uint64_t &cnt, uint64_t &sum, uint64_t *pairs,
                                          uint64_t &wpos, uint64_t cap) {
    const uint32_t k = key_of(s);
    const uint32_t h = bucket_of(k, lg);
    const uint32_t e = bst[h + 1];
    for (uint32_t x = bst[h]; x < e; ++x) {
        const uint64_t r = tab[x];
        if (key_of(r) == k) {
            if (WRITE) {
                if (wpos < cap) pairs[wpos] = pair_value(r, s);
                ++wpos;
            } else {
                ++cnt;
                sum += pair_value(r, s);
            }
        }
    }
}

// Probe the task's S slice: the first SREG*JB tuples from registers, the rest
// streamed from memory.
template <int JB, bool WRITE>
__device__ __forceinline__ void probe_task(const uint64_t (&sreg)[SREG], const uint64_t *S,
                                           uint64_t ns, const uint64_t *tab, const uint32_t *bst,
                                           int lg, uint64_t &cnt, uint64_t &sum, uint64_t *pairs,
                                           uint64_t &wpos, uint64_t cap) {
#pragma unroll
    for (int q = 0; q < SREG; ++q)
        if (threadIdx.x + (uint64_t)q * JB < ns)
            probe_one<WRITE>(sreg[q], tab, bst, lg, cnt, sum, pairs, wpos, cap);
    uint64_t i = (uint64_t)SREG * JB + threadIdx.x;
    for (; i + (JUNROLL - 1) * JB < ns; i += JUNROLL * JB) {
        uint64_t s[JUNROLL];
#pragma unroll
        for (int u = 0; u < JUNROLL; ++u) s[u] = S[i + u * JB];
#pragma unroll
        for (int u = 0; u < JUNROLL; ++u)
            probe_one<WRITE>(s[u], tab, bst, lg, cnt, sum, pairs, wpos, cap);
    }
    for (; i < ns; i += JB) probe_one<WRITE>(S[i], tab, bst, lg, cnt, sum, pairs, wpos, cap);
}

// A task record is two uint4: {r0 lo, r0 hi, |R_p|, segment}, {s0 lo, s0 hi,
// |S slice|, 0}; everything a worker needs in one independent load.
struct TaskRec {
    uint64_t r0, s0;
    uint32_t nr, ns, seg;
};
__device__ __forceinline__ TaskRec decode_task(uint4 a, uint4 b) {
    TaskRec r;
    r.r0 = ((uint64_t)a.y << 32) | a.x;
    r.nr = a.z;
    r.seg = a.w;
    r.s0 = ((uint64_t)b.y << 32) | b.x;
    r.ns = b.z;
    return r;
}

template <int JB>
__device__ __forceinline__ void load_sreg(uint64_t (&sreg)[SREG], const uint64_t *S, uint64_t ns) {
#pragma unroll
    for (int q = 0; q < SREG; ++q) {
        const uint64_t i = threadIdx.x + (uint64_t)q * JB;
        sreg[q] = i < ns ? __builtin_nontemporal_load(S + i) : 0;
    }
}

template <int JB, int CAP, bool PAIRS>
__global__ __launch_bounds__(JB, 4) void k_join(JoinArgs a) {
    extern __shared__ uint64_t jsm[];
    uint64_t *tab = jsm;                                     // CAP tuples
    uint32_t *bst = reinterpret_cast<uint32_t *>(tab + CAP);  // max_buckets(CAP) + 1 starts
    __shared__ uint64_t red[2 * (JB / 64)];
    __shared__ uint32_t scan_tmp[JB / 64];
    __shared__ uint64_t s_base;
    constexpr int NW = JB / 64;

    // Static round-robin over tasks (a shared device-wide task counter costs one
    // serialised atomic per task).  The next task's record and its first S
    // tuples are fetched while the current task runs.
    const uint32_t total = a.task_base[a.max_parts];
    uint32_t t = blockIdx.x;
    if (t >= total) return;
    TaskRec cur = decode_task(a.task_info[2 * t], a.task_info[2 * t + 1]);
    uint64_t sreg[SREG];
    load_sreg<JB>(sreg, a.S + cur.s0, cur.ns);
    for (;;) {
        const uint32_t tn = t + gridDim.x;
        const bool has_next = tn < total;
        uint4 n0 = make_uint4(0, 0, 0, 0), n1 = n0;
        if (has_next) {
            n0 = a.task_info[2 * tn];
            n1 = a.task_info[2 * tn + 1];
        }
        const uint64_t *S = a.S + cur.s0;
        const uint64_t ns = cur.ns;
        const uint64_t r1 = cur.r0 + cur.nr;
        uint64_t cnt = 0, sum = 0;
        for (uint64_t rb = cur.r0; rb < r1; rb += CAP) {
            const uint32_t m = (uint32_t)min<uint64_t>(CAP, r1 - rb);
            const int lg = build_round<JB, CAP>(a.R + rb, m, tab, bst, scan_tmp);
            uint64_t c = 0, x = 0, wpos = 0;
            probe_task<JB, false>(sreg, S, ns, tab, bst, lg, c, x, nullptr, wpos, 0);
            if (PAIRS) {
                uint32_t total_c;
                const uint32_t pre = block_exclusive_scan<JB>((uint32_t)c, scan_tmp, &total_c);
                if (threadIdx.x == 0)
                    s_base = atomicAdd((unsigned long long *)a.pair_cursor, (unsigned long long)total_c);
                __syncthreads();
                wpos = s_base + pre;
                uint64_t dc = 0, dx = 0;
                probe_task<JB, true>(sreg, S, ns, tab, bst, lg, dc, dx, a.pairs, wpos, a.pairs_cap);
            }
            cnt += c;
            sum += x;
            __syncthreads();  // the next round rebuilds tab/bst
        }
        TaskRec nxt{};
        if (has_next) {
            nxt = decode_task(n0, n1);
            load_sreg<JB>(sreg, a.S + nxt.s0, nxt.ns);  // in flight during the reduction
        }
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

size_t join_scan_tmp_bytes(uint64_t max_parts) {
    size_t bytes = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (uint32_t *)nullptr, (uint32_t *)nullptr,
                                           (int)(max_parts + 1));
    return bytes;
}

uint64_t join_max_tasks(uint64_t max_parts, uint64_t max_S, uint32_t split) {
    return max_parts + (split ? max_S / split : max_S) + 1;
}

template <int JB, int CAP, bool PAIRS>
static int launch_join(const JoinArgs &a, int num_cu, hipStream_t st) {
    const size_t lds = sizeof(uint64_t) * CAP + sizeof(uint32_t) * (max_buckets(CAP) + 2);
    static bool attr = false;
    if (!attr) {
        AJ_HIP(hipFuncSetAttribute((const void *)k_join<JB, CAP, PAIRS>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        attr = true;
    }
    static int per_cu = 0;
    if (!per_cu) {
        AJ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_join<JB, CAP, PAIRS>, JB, lds));
        if (per_cu < 1) per_cu = 1;
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
        set_error(AJ_ERR_ARG, "bad join split/partition count");
        return AJ_ERR_ARG;
    }
    JoinArgs a;
    a.R = d.R;
    a.S = d.S;
    a.roff = d.roff;
    a.soff = d.soff;
    a.part_seg_base = d.part_seg_base;
    a.nseg = d.nseg;
    a.max_parts = d.max_parts;
    a.split = d.split;
    a.seg_count = d.seg_count;
    a.seg_sum = d.seg_sum;
    a.pairs = d.pairs;
    a.pairs_cap = d.pairs_cap;
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
    const unsigned g = (unsigned)std::min<uint64_t>((d.max_parts + 256) / 256, 4096);
    hipLaunchKernelGGL(k_task_count, dim3(g), dim3(256), 0, st, a);
    size_t tmp = d.scan_tmp_bytes;
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(d.scan_tmp, tmp, d.task_cnt, d.task_base,
                                                    (int)(d.max_parts + 1), st);
    if (e != hipSuccess) {
        set_error(AJ_ERR_HIP, "task scan failed: %s", hipGetErrorString(e));
        return AJ_ERR_HIP;
    }
    hipLaunchKernelGGL(k_task_fill, dim3(g), dim3(256), 0, st, a);
    if (mk) mk->mark(m0);
    const bool pairs = d.pairs != nullptr;
    int rc;
    switch (d.cap) {
        case 2048: rc = launch_join2<256, 2048>(a, pairs, d.num_cu, st); break;
        case 4096: rc = launch_join2<256, 4096>(a, pairs, d.num_cu, st); break;
        case 8192: rc = launch_join2<512, 8192>(a, pairs, d.num_cu, st); break;
        case 12288: rc = launch_join2<512, 12288>(a, pairs, d.num_cu, st); break;
        default:
            set_error(AJ_ERR_ARG, "cap must be 2048, 4096, 8192 or 12288 (got %d)", d.cap);
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
