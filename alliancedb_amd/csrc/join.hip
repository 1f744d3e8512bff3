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
//   k_join        persistent work-groups pull task ids from a device counter.
//                 Each task = (partition p, slice j of S_p).  R_p is taken in
//                 rounds of `cap` tuples; each round counting-sorts its tuples
//                 by bucket into LDS (bucket starts + sorted tuples) and then
//                 streams the S slice, scanning one bucket per probe.  Count
//                 and checksum are reduced per work-group and added once to
//                 the segment's 64-bit accumulators.  With pairs enabled, each
//                 round probes twice: count, reserve a range with one atomic,
//                 write.
//
// Bytes per tuple: R and S are each read once per round (8 B).
#include <hipcub/hipcub.hpp>

#include "aj_internal.h"
#include "device_common.h"

namespace aj {

constexpr int JBLOCK = 512;
constexpr int JUNROLL = 4;

// Bucket count of a full round: the smallest power of two nb with 2*nb >= CAP
// (build_round's rule at m = CAP); the bucket-start array holds nb+1 entries.
constexpr uint32_t max_buckets(uint32_t cap) {
    uint32_t nb = 1;
    while (2 * nb < cap) nb *= 2;
    return nb;
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
};

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

__device__ __forceinline__ uint32_t bucket_of(uint32_t key, int lg) {
    return lg ? (key * 0x9E3779B1u) >> (32 - lg) : 0u;
}

// Counting-sort m (<= CAP) tuples of R into `tab` by bucket; `bst` receives
// nb+1 bucket starts.  Returns log2(nb).
template <int CAP>
__device__ __forceinline__ int build_round(const uint64_t *R, uint32_t m, uint64_t *tab,
                                           uint32_t *bst, uint32_t *scan_tmp) {
    constexpr int ITEMS = CAP / JBLOCK;
    // about two tuples per bucket, as the reference's BUCKET_SIZE
    int lg = 0;
    while ((2u << lg) < m) ++lg;
    const uint32_t nb = 1u << lg;
    for (uint32_t i = threadIdx.x; i <= nb; i += JBLOCK) bst[i] = 0;
    __syncthreads();
    uint64_t t[ITEMS];
    uint32_t rk[ITEMS];
#pragma unroll
    for (int q = 0; q < ITEMS; ++q) {
        const uint32_t i = threadIdx.x + q * JBLOCK;
        t[q] = i < m ? R[i] : 0;
    }
#pragma unroll
    for (int q = 0; q < ITEMS; ++q) {
        const uint32_t i = threadIdx.x + q * JBLOCK;
        if (i < m) rk[q] = atomicAdd(&bst[bucket_of(key_of(t[q]), lg)], 1u);
    }
    __syncthreads();
    lds_exclusive_scan<JBLOCK>(bst, nb + 1, scan_tmp);
#pragma unroll
    for (int q = 0; q < ITEMS; ++q) {
        const uint32_t i = threadIdx.x + q * JBLOCK;
        if (i < m) tab[bst[bucket_of(key_of(t[q]), lg)] + rk[q]] = t[q];
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

template <bool WRITE>
__device__ __forceinline__ void probe_slice(const uint64_t *S, uint64_t ns, const uint64_t *tab,
                                            const uint32_t *bst, int lg, uint64_t &cnt,
                                            uint64_t &sum, uint64_t *pairs, uint64_t &wpos,
                                            uint64_t cap) {
    uint64_t i = threadIdx.x;
    for (; i + (JUNROLL - 1) * JBLOCK < ns; i += JUNROLL * JBLOCK) {
        uint64_t s[JUNROLL];
#pragma unroll
        for (int u = 0; u < JUNROLL; ++u) s[u] = S[i + u * JBLOCK];
#pragma unroll
        for (int u = 0; u < JUNROLL; ++u)
            probe_one<WRITE>(s[u], tab, bst, lg, cnt, sum, pairs, wpos, cap);
    }
    for (; i < ns; i += JBLOCK) probe_one<WRITE>(S[i], tab, bst, lg, cnt, sum, pairs, wpos, cap);
}

template <int CAP, bool PAIRS>
__global__ __launch_bounds__(JBLOCK) void k_join(JoinArgs a) {
    extern __shared__ uint64_t jsm[];
    uint64_t *tab = jsm;                                     // CAP tuples
    uint32_t *bst = reinterpret_cast<uint32_t *>(tab + CAP);  // max_buckets(CAP) + 1 starts
    __shared__ uint32_t s_task;
    __shared__ uint64_t red[JBLOCK / 64];
    __shared__ uint32_t scan_tmp[JBLOCK / 64];
    __shared__ uint64_t s_base;

    const uint32_t total = a.task_base[a.max_parts];
    for (;;) {
        if (threadIdx.x == 0) s_task = atomicAdd(a.task_next, 1u);
        __syncthreads();
        const uint32_t t = s_task;
        __syncthreads();
        if (t >= total) break;
        const uint64_t p = last_le<uint32_t>(a.task_base, a.max_parts + 1, t);
        const uint64_t j = t - a.task_base[p];
        const uint64_t r0 = a.roff[p], r1 = a.roff[p + 1];
        const uint64_t s0 = a.soff[p] + j * a.split;
        const uint64_t s1 = min(s0 + a.split, a.soff[p + 1]);
        const int seg = (int)last_le<uint32_t>(a.part_seg_base, (uint64_t)a.nseg + 1, (uint32_t)p);
        uint64_t cnt = 0, sum = 0;
        for (uint64_t rb = r0; rb < r1; rb += CAP) {
            const uint32_t m = (uint32_t)min<uint64_t>(CAP, r1 - rb);
            const int lg = build_round<CAP>(a.R + rb, m, tab, bst, scan_tmp);
            uint64_t c = 0, x = 0, wpos = 0;
            probe_slice<false>(a.S + s0, s1 - s0, tab, bst, lg, c, x, nullptr, wpos, 0);
            if (PAIRS) {
                uint32_t total_c;
                const uint32_t pre = block_exclusive_scan<JBLOCK>((uint32_t)c, scan_tmp, &total_c);
                if (threadIdx.x == 0) s_base = atomicAdd((unsigned long long *)a.pair_cursor,
                                                         (unsigned long long)total_c);
                __syncthreads();
                wpos = s_base + pre;
                uint64_t dc = 0, dx = 0;
                probe_slice<true>(a.S + s0, s1 - s0, tab, bst, lg, dc, dx, a.pairs, wpos,
                                  a.pairs_cap);
            }
            cnt += c;
            sum += x;
            __syncthreads();  // the next round rebuilds tab/bst
        }
        cnt = block_sum<JBLOCK>(cnt, red);
        sum = block_sum<JBLOCK>(sum, red);
        if (threadIdx.x == 0 && cnt) {
            atomicAdd((unsigned long long *)&a.seg_count[seg], (unsigned long long)cnt);
            atomicAdd((unsigned long long *)&a.seg_sum[seg], (unsigned long long)sum);
        }
    }
}

size_t join_scan_tmp_bytes(uint64_t max_parts) {
    size_t bytes = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (uint32_t *)nullptr, (uint32_t *)nullptr,
                                     (int)(max_parts + 1));
    return bytes;
}

template <int CAP, bool PAIRS> static int launch_join(const JoinArgs &a, int num_cu, hipStream_t st) {
    const size_t lds = sizeof(uint64_t) * CAP + sizeof(uint32_t) * (max_buckets(CAP) + 2);
    static bool attr = false;
    if (!attr) {
        AJ_HIP(hipFuncSetAttribute((const void *)k_join<CAP, PAIRS>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        attr = true;
    }
    int per_cu = 0;
    AJ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_join<CAP, PAIRS>, JBLOCK, lds));
    if (per_cu < 1) per_cu = 1;
    hipLaunchKernelGGL((k_join<CAP, PAIRS>), dim3(num_cu * per_cu), dim3(JBLOCK), lds, st, a);
    return AJ_OK;
}

int run_join(const JoinDesc &d, hipStream_t st, const Marks *mk, int m0) {
    if (d.split == 0 || d.max_parts >= (1ull << 31)) {
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
    const unsigned g = (unsigned)std::min<uint64_t>((d.max_parts + 256) / 256, 4096);
    hipLaunchKernelGGL(k_task_count, dim3(g), dim3(256), 0, st, a);
    size_t tmp = d.scan_tmp_bytes;
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(d.scan_tmp, tmp, d.task_cnt, d.task_base,
                                                    (int)(d.max_parts + 1), st);
    if (e != hipSuccess) {
        set_error(AJ_ERR_HIP, "task scan failed: %s", hipGetErrorString(e));
        return AJ_ERR_HIP;
    }
    if (mk) mk->mark(m0);
    const bool pairs = d.pairs != nullptr;
    int rc;
    switch (d.cap) {
        case 4096: rc = pairs ? launch_join<4096, true>(a, d.num_cu, st) : launch_join<4096, false>(a, d.num_cu, st); break;
        case 8192: rc = pairs ? launch_join<8192, true>(a, d.num_cu, st) : launch_join<8192, false>(a, d.num_cu, st); break;
        case 12288: rc = pairs ? launch_join<12288, true>(a, d.num_cu, st) : launch_join<12288, false>(a, d.num_cu, st); break;
        default:
            set_error(AJ_ERR_ARG, "cap must be 4096, 8192 or 12288 (got %d)", d.cap);
            return AJ_ERR_ARG;
    }
    if (rc) return rc;
    if (mk) mk->mark(m0 + 1);
    AJ_HIP(hipGetLastError());
    return AJ_OK;
}

}  // namespace aj
