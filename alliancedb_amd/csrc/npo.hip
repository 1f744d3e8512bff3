// npo.hip -- no-partitioning join on one HBM-resident hash table (NPO).
//
// Reference: NPO builds one shared chained table over all of R with a latch
// per bucket (build_hashtable_mt, common_functions.cpp:350-400) and probes it
// with every S tuple, counting each key-equal entry of the bucket chain
// (probe_hashtable_single, common_functions.cpp:237-280).
//
// CDNA4 form: open addressing over 64-bit slots that hold the whole tuple
// (key << 32 | payload), linear probing, capacity = next power of two >= 2|R|
// (load factor <= 1/2).  Insert is one 64-bit atomicCAS per probed slot (the
// latch and the slot write of the reference in one instruction).  Duplicate
// keys each take their own slot; a probe walks from the key's home slot to the
// first empty slot and counts every slot with an equal key.  The empty marker
// is the all-ones word; an R tuple with exactly that bit pattern (key -1,
// payload -1) is counted on the side and matched arithmetically, so every
// 64-bit tuple value stays joinable.
//
// The home slot is a murmur3-finalised key: the reference's `key & mask`
// (common_functions.h:57-59) would turn dense keys into one long run under
// linear probing.
//
// Bytes per tuple: R read 8 + slot write 8; S read 8 + slot read 8.
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

__global__ __launch_bounds__(NBLOCK) void k_npo_build(const uint64_t *R, uint64_t n, uint64_t *tab,
                                                    uint64_t mask, uint64_t *acc) {
    uint64_t sentinels = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t v = R[i];
        if (v == EMPTY_SLOT) {
            ++sentinels;
            continue;
        }
        uint64_t h = fmix32(key_of(v)) & mask;
        // terminates: fewer tuples than slots
        while (atomicCAS((unsigned long long *)&tab[h], (unsigned long long)EMPTY_SLOT,
                         (unsigned long long)v) != EMPTY_SLOT)
            h = (h + 1) & mask;
    }
    if (sentinels) atomicAdd((unsigned long long *)&acc[2], (unsigned long long)sentinels);
}

template <bool WRITE>
__device__ __forceinline__ void npo_probe_one(uint64_t s, const uint64_t *tab, uint64_t mask,
                                              uint64_t nsent, uint64_t &cnt, uint64_t &sum,
                                              uint64_t *pairs, uint64_t &wpos, uint64_t cap,
                                              int records = 0, uint32_t tnow = 0) {
    const uint32_t k = key_of(s);
    uint64_t h = fmix32(k) & mask;
    for (;;) {
        const uint64_t v = tab[h];
        if (v == EMPTY_SLOT) break;
        if (key_of(v) == k) {
            if (WRITE) {
                if (wpos < cap)
                    pairs[wpos] = records == 2 ? ((uint64_t)tnow << 32) | (uint32_t)s
                                  : records ? s : pair_value(v, s);
                ++wpos;
            } else {
                ++cnt;
                sum += pair_value(v, s);
            }
        }
        h = (h + 1) & mask;
    }
    if (k == 0xFFFFFFFFu && nsent) {
        const uint64_t pv = pair_value(EMPTY_SLOT, s);
        if (WRITE) {
            for (uint64_t q = 0; q < nsent; ++q) {
                if (wpos < cap)
                    pairs[wpos] = records == 2 ? ((uint64_t)tnow << 32) | (uint32_t)s
                                  : records ? s : pv;
                ++wpos;
            }
        } else {
            cnt += nsent;
            sum += nsent * pv;
        }
    }
}

template <bool PAIRS>
__global__ __launch_bounds__(NBLOCK) void k_npo_probe(const uint64_t *S, uint64_t n,
                                                    const uint64_t *tab, uint64_t mask,
                                                    uint64_t *acc, uint64_t *pairs, uint64_t cap,
                                                    int records) {
    __shared__ uint64_t red[NBLOCK / 64];
    __shared__ uint64_t s_base;
    const uint64_t nsent = acc[2];
    uint64_t cnt = 0, sum = 0;
    // uniform trip count per block: every thread reaches the block scans
    for (uint64_t base = blockIdx.x * (uint64_t)NBLOCK; base < n;
         base += (uint64_t)gridDim.x * NBLOCK) {
        const uint64_t i = base + threadIdx.x;
        const bool valid = i < n;
        const uint64_t s = valid ? S[i] : 0;
        uint64_t c = 0, x = 0, wpos = 0;
        if (valid) npo_probe_one<false>(s, tab, mask, nsent, c, x, nullptr, wpos, 0);
        if (PAIRS) {
            uint64_t total;
            const uint64_t pre = block_exclusive_scan<NBLOCK>(c, red, &total);
            if (threadIdx.x == 0 && total)
                s_base = atomicAdd((unsigned long long *)&acc[3], (unsigned long long)total);
            __syncthreads();
            if (valid && c) {
                wpos = s_base + pre;
                uint64_t dc = 0, dx = 0;
                const uint32_t tnow =
                    records == 2 ? (uint32_t)__builtin_amdgcn_s_memrealtime() : 0u;
                npo_probe_one<true>(s, tab, mask, nsent, dc, dx, pairs, wpos, cap, records, tnow);
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

int run_npo(const NpoDesc &d, hipStream_t st, const Marks *mk, int m0) {
    if (d.capacity < 2 || (d.capacity & (d.capacity - 1)) || d.capacity <= d.nR) {
        set_error(AJ_ERR_ARG, "NPO table capacity must be a power of two above |R|");
        return AJ_ERR_ARG;
    }
    AJ_TRY(fill_u64(d.table, d.capacity, EMPTY_SLOT, st));
    const uint64_t mask = d.capacity - 1;
    if (mk) mk->mark(m0);
    if (d.nR)
        hipLaunchKernelGGL(k_npo_build, dim3(grid_for(d.nR, d.num_cu)), dim3(NBLOCK), 0, st, d.R,
                           d.nR, d.table, mask, d.acc);
    if (mk) mk->mark(m0 + 1);
    if (d.nS) {
        const unsigned g = grid_for(d.nS, d.num_cu);  // <= num_cu * 16 (npo_acc_words)
        if (d.pairs)
            hipLaunchKernelGGL(k_npo_probe<true>, dim3(g), dim3(NBLOCK), 0, st, d.S, d.nS,
                               (const uint64_t *)d.table, mask, d.acc, d.pairs, d.pairs_cap,
                               d.records);
        else
            hipLaunchKernelGGL(k_npo_probe<false>, dim3(g), dim3(NBLOCK), 0, st, d.S, d.nS,
                               (const uint64_t *)d.table, mask, d.acc, nullptr, (uint64_t)0, 0);
        hipLaunchKernelGGL(k_npo_reduce, dim3(1), dim3(NBLOCK), 0, st, d.acc, g);
    }
    if (mk) mk->mark(m0 + 2);
    AJ_HIP(hipGetLastError());
    return AJ_OK;
}

}  // namespace aj
