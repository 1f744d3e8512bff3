// heavy.hip -- the heaviest keys of a skewed probe side, joined by aggregates.
//
// A counting join's result is a sum over keys: for key k with R tuples of
// payloads r_i and S tuples of payloads s_j, the matches are cR(k) * cS(k) and
// their checksum terms ((r_i << 32) | s_j) sum to (sumR(k) * cS(k)) << 32 +
// cR(k) * sumS(k) (mod 2^64).  So the keys that hold a large share of S (Zipf
// S at theta >= 1) need not be partitioned at all: these kernels take their
// four aggregates from one read of R and S, and copy the rest of S (the light
// keys) to a compacted relation that the partitioned join then takes.  The
// heavy keys come from a strided sample of S (host side, api.cpp).
#include <algorithm>

#include "aj_internal.h"
#include "device_common.h"

namespace aj {

constexpr int HEAVY_SLOTS = 4 * HEAVY_MAX;  // hash slots: at most a quarter of them taken
constexpr int HEAVY_TOP = 2;                // keys summed in registers (the hottest)
constexpr int HBLOCK = 512, HITEMS = 16;    // 8192-tuple tiles

__global__ void k_heavy_sample(const uint64_t *S, uint64_t stride, uint32_t m, uint32_t *keys) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < m; j += gridDim.x * blockDim.x)
        keys[j] = key_of(S[(uint64_t)j * stride]);
}

// 1 + index of key k among the heavy keys, 0 if it is not one.  A slot holds
// key << 32 | (1 + index), 0 when empty (open addressing, linear probing; an
// empty slot ends every probe since at most a quarter of the slots are taken).
__host__ __device__ __forceinline__ uint32_t heavy_hash(uint32_t k) {
    return (k * 0x9E3779B1u) >> (32 - 12);  // 12 bits: HEAVY_SLOTS
}
static_assert(HEAVY_SLOTS == 1 << 12, "heavy_hash gives 12 bits");

__device__ __forceinline__ uint32_t heavy_find(const uint64_t *slots, uint32_t k) {
    uint32_t h = heavy_hash(k);
    for (;;) {
        const uint64_t v = slots[h];
        if ((uint32_t)v == 0u || (uint32_t)(v >> 32) == k) return (uint32_t)v;
        h = (h + 1) & (HEAVY_SLOTS - 1);
    }
}

struct HeavyArgs {
    const uint64_t *slots;    // HEAVY_SLOTS
    uint32_t nheavy;
    const uint64_t *in;
    uint64_t n;
    unsigned long long *cnt;  // nheavy counts of this side
    unsigned long long *sum;  // nheavy payload sums of this side
    uint64_t *out;            // nullable: the other tuples, compacted
    unsigned long long *out_n;
};

// Grid-stride over tiles of HBLOCK * HITEMS tuples: heavy tuples are counted
// and summed (the HEAVY_TOP hottest keys in registers, the others by LDS
// atomics), the others, with `out` set, written to out in tile order (per item
// and wave a contiguous run, placed by one atomic per tile).  A work-group
// takes at most 2^32 / (HBLOCK * HITEMS) tiles (32-bit LDS counts).
__global__ __launch_bounds__(HBLOCK) void k_heavy_split(HeavyArgs a) {
    __shared__ uint64_t s_slots[HEAVY_SLOTS];
    __shared__ uint32_t s_cnt[HEAVY_MAX];
    __shared__ unsigned long long s_sum[HEAVY_MAX];
    constexpr int NW = HBLOCK / 64;
    static_assert(HITEMS * NW <= 128, "two scan entries per lane of wave 0");
    __shared__ uint32_t s_run[HITEMS * NW];  // kept tuples per (item, wave), then their starts
    __shared__ unsigned long long s_base;
    const uint32_t tid = threadIdx.x, lane = lane_id(), wid = tid / 64;
    for (uint32_t i = tid; i < HEAVY_SLOTS; i += HBLOCK) s_slots[i] = a.slots[i];
    for (uint32_t i = tid; i < HEAVY_MAX; i += HBLOCK) {
        s_cnt[i] = 0;
        s_sum[i] = 0;
    }
    __syncthreads();
    uint32_t tc[HEAVY_TOP] = {};
    uint64_t ts[HEAVY_TOP] = {};
    constexpr uint64_t TILE = (uint64_t)HBLOCK * HITEMS;
    const uint64_t lt_mask = (1ull << lane) - 1;
    for (uint64_t base = blockIdx.x * TILE; base < a.n; base += gridDim.x * TILE) {
        uint64_t t[HITEMS];
        const bool full = base + TILE <= a.n;
        if (full) {
#pragma unroll
            for (int j = 0; j < HITEMS; ++j)
                t[j] = __builtin_nontemporal_load(a.in + base + (uint64_t)j * HBLOCK + tid);
        } else {
#pragma unroll
            for (int j = 0; j < HITEMS; ++j) {
                const uint64_t i = base + (uint64_t)j * HBLOCK + tid;
                t[j] = i < a.n ? __builtin_nontemporal_load(a.in + i) : 0;
            }
        }
        uint32_t keepm = 0;  // bit j: item j is a tuple of a light key
#pragma unroll
        for (int j = 0; j < HITEMS; ++j) {
            const bool valid = full || base + (uint64_t)j * HBLOCK + tid < a.n;
            const uint32_t s = valid ? heavy_find(s_slots, key_of(t[j])) : ~0u;
            const uint32_t p = payload_of(t[j]);
#pragma unroll
            for (int q = 0; q < HEAVY_TOP; ++q) {
                const bool h = s == (uint32_t)q + 1;
                tc[q] += h ? 1u : 0u;
                ts[q] += h ? p : 0u;
            }
            if (s > HEAVY_TOP && s != ~0u) {
                atomicAdd(&s_cnt[s - 1], 1u);
                atomicAdd(&s_sum[s - 1], (unsigned long long)p);
            }
            if (s == 0) keepm |= 1u << j;
        }
        if (a.out) {
            // per item j, this wave's kept tuples form one run of the tile's
            // output, in (item, wave) order
#pragma unroll
            for (int j = 0; j < HITEMS; ++j) {
                const uint64_t b = __ballot((keepm >> j) & 1u);
                if (lane == 0) s_run[j * NW + wid] = (uint32_t)__popcll(b);
            }
            __syncthreads();
            if (tid < 64) {
                const uint32_t v0 = s_run[2 * tid], v1 = s_run[2 * tid + 1];
                const uint32_t inc = wave_inclusive_scan(v0 + v1);
                const uint32_t pre = inc - v0 - v1;
                s_run[2 * tid] = pre;
                s_run[2 * tid + 1] = pre + v0;
                if (tid == 63) s_base = atomicAdd(a.out_n, (unsigned long long)inc);
            }
            __syncthreads();
            const uint64_t ob = s_base;
#pragma unroll
            for (int j = 0; j < HITEMS; ++j) {
                const uint64_t b = __ballot((keepm >> j) & 1u);
                if ((keepm >> j) & 1u)
                    a.out[ob + s_run[j * NW + wid] + (uint64_t)__popcll(b & lt_mask)] = t[j];
            }
            __syncthreads();  // s_run and s_base are rewritten by the next tile
        }
    }
    // the hottest keys: lane sums, then wave sums, into LDS
#pragma unroll
    for (int q = 0; q < HEAVY_TOP; ++q) {
        const uint64_t c = wave_sum((uint64_t)tc[q]), x = wave_sum(ts[q]);
        if (lane == 0 && (uint32_t)q < a.nheavy && c) {
            atomicAdd(&s_cnt[q], (uint32_t)c);
            atomicAdd(&s_sum[q], (unsigned long long)x);
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < a.nheavy; i += HBLOCK)
        if (s_cnt[i]) {
            atomicAdd(&a.cnt[i], (unsigned long long)s_cnt[i]);
            atomicAdd(&a.sum[i], s_sum[i]);
        }
}

int heavy_sample(const uint64_t *dS, uint64_t nS, uint32_t m, uint32_t *d_keys, hipStream_t st) {
    const uint64_t stride = nS / m;
    hipLaunchKernelGGL(k_heavy_sample, dim3((m + 255) / 256), dim3(256), 0, st, dS, stride, m,
                       d_keys);
    AJ_HIP(hipGetLastError());
    return AJ_OK;
}

int heavy_split(const HeavyDesc &d, hipStream_t st) {
    if (d.nheavy == 0 || d.nheavy > HEAVY_MAX) {
        set_error(AJ_ERR_ARG, "heavy_split: 1..%d heavy keys (got %u)", HEAVY_MAX, d.nheavy);
        return AJ_ERR_ARG;
    }
    for (int side = 0; side < 2; ++side) {
        HeavyArgs a;
        a.slots = d.slots;
        a.nheavy = d.nheavy;
        a.in = side ? d.S : d.R;
        a.n = side ? d.nS : d.nR;
        a.cnt = d.agg + (side ? 2 : 0) * HEAVY_MAX;
        a.sum = d.agg + (side ? 3 : 1) * HEAVY_MAX;
        a.out = side ? d.light : nullptr;
        a.out_n = d.light_n;
        const uint64_t tiles = (a.n + (uint64_t)HBLOCK * HITEMS - 1) / ((uint64_t)HBLOCK * HITEMS);
        // enough work-groups that one takes < 2^32 / TILE tiles (32-bit counts)
        const unsigned grid = (unsigned)std::min<uint64_t>(std::max<uint64_t>(tiles, 1),
                                                           (uint64_t)d.num_cu * 4);
        if (a.n) hipLaunchKernelGGL(k_heavy_split, dim3(grid), dim3(HBLOCK), 0, st, a);
    }
    AJ_HIP(hipGetLastError());
    return AJ_OK;
}

void heavy_table(const uint32_t *keys, uint32_t n, uint64_t *slots) {
    std::fill(slots, slots + HEAVY_SLOTS, 0ull);
    for (uint32_t i = 0; i < n && i < (uint32_t)HEAVY_MAX; ++i) {
        uint32_t h = heavy_hash(keys[i]);
        while (slots[h]) h = (h + 1) & (HEAVY_SLOTS - 1);
        slots[h] = ((uint64_t)keys[i] << 32) | (i + 1);
    }
}

}  // namespace aj
