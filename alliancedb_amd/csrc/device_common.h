// device_common.h -- wave/block primitives shared by the join kernels (gfx950,
// 64-wide wavefronts).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <type_traits>
#include <utility>

namespace aj {

// f(std::integral_constant<int, q>) for q = 0..N-1 as straight-line code (a
// fold expression, not a loop), so per-lane register arrays indexed by q are
// always indexed by constants and stay in registers.
template <class F, int... Q>
__device__ __forceinline__ void unroll_seq(F &&f, std::integer_sequence<int, Q...>) {
    (f(std::integral_constant<int, Q>{}), ...);
}
template <int N, class F> __device__ __forceinline__ void unroll(F &&f) {
    unroll_seq(f, std::make_integer_sequence<int, N>{});
}

constexpr uint64_t EMPTY_SLOT = ~0ull;

__device__ __forceinline__ uint32_t key_of(uint64_t t) { return (uint32_t)(t >> 32); }
__device__ __forceinline__ uint32_t payload_of(uint64_t t) { return (uint32_t)t; }
// (R.payloadID << 32) | S.payloadID
__device__ __forceinline__ uint64_t pair_value(uint64_t r, uint64_t s) {
    return (r << 32) | (uint64_t)(uint32_t)s;
}

// Tuple i of a relation stored as 8-byte tuple_t or 12-byte {key, payload, ts}.
__device__ __forceinline__ uint64_t load_tuple(const void *base, int stride, uint64_t i) {
    if (stride == 8) return __builtin_nontemporal_load(static_cast<const uint64_t *>(base) + i);
    const uint32_t *p = static_cast<const uint32_t *>(base) + 3 * i;
    return ((uint64_t)p[0] << 32) | p[1];
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Lanes of the wave whose `d` equals this lane's (restricted to valid lanes),
// built from one ballot per digit bit.  Every lane of the wave must call it.
__device__ __forceinline__ uint64_t match_peers(uint32_t d, bool valid, int bits) {
    uint64_t peers = __ballot(valid);
    for (int i = 0; i < bits; ++i) {
        const bool b = (d >> i) & 1u;
        const uint64_t m = __ballot(b);
        peers &= b ? m : ~m;
    }
    return valid ? peers : 0ull;
}

// Wave-aggregated LDS counter increment: returns this lane's rank among the
// lanes that added to counter[d] in this call (valid lanes only).  One LDS
// atomic per distinct digit in the wave, so heavy duplication does not
// serialise on one address.
__device__ __forceinline__ uint32_t wave_rank_add(uint32_t *counter, uint32_t d, bool valid,
                                                  int bits) {
    const uint64_t peers = match_peers(d, valid, bits);
    const int lane = lane_id();
    const int leader = peers ? __ffsll((long long)peers) - 1 : 0;
    uint32_t base = 0;
    if (valid && lane == leader) base = atomicAdd(&counter[d], (uint32_t)__popcll(peers));
    base = __shfl(base, leader);
    const uint64_t below = (lane == 0) ? 0ull : (peers & ((1ull << lane) - 1ull));
    return base + (uint32_t)__popcll(below);
}

__device__ __forceinline__ void wave_count_add(uint32_t *counter, uint32_t d, bool valid,
                                               int bits) {
    const uint64_t peers = match_peers(d, valid, bits);
    const int lane = lane_id();
    if (valid && lane == __ffsll((long long)peers) - 1)
        atomicAdd(&counter[d], (uint32_t)__popcll(peers));
}

template <class T> __device__ __forceinline__ T wave_inclusive_scan(T x) {
    const int lane = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        T y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    return x;
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t y = __shfl_xor(x, o);
        x = x > y ? x : y;
    }
    return x;
}

// Maximum over the wave with DPP lane moves (no LDS traffic): row shifts by
// 1, 2, 4, 8 leave each row's maximum in its lane 15, two row broadcasts carry
// them to lane 63, which is read back as a wave-uniform value.
__device__ __forceinline__ uint32_t wave_max_u32_dpp(uint32_t x) {
    auto mx = [](uint32_t a, int b) { return a > (uint32_t)b ? a : (uint32_t)b; };
    x = mx(x, __builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false));  // row_shr:1
    x = mx(x, __builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false));  // row_shr:2
    x = mx(x, __builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false));  // row_shr:4
    x = mx(x, __builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false));  // row_shr:8
    x = mx(x, __builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false));  // row_bcast:15
    x = mx(x, __builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false));  // row_bcast:31
    return (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
}

template <class T> __device__ __forceinline__ T wave_sum(T x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

// Exclusive scan over the block; `tmp` is LDS of at least BLOCK/64 entries.
// Returns the exclusive prefix of v; *total receives the block sum.
template <int BLOCK, class T>
__device__ __forceinline__ T block_exclusive_scan(T v, T *tmp, T *total) {
    constexpr int NW = BLOCK / 64;
    const int lane = lane_id();
    const int wid = threadIdx.x / 64;
    T x = wave_inclusive_scan(v);
    if (lane == 63) tmp[wid] = x;
    __syncthreads();
    if (wid == 0) {
        T w = lane < NW ? tmp[lane] : T(0);
        w = wave_inclusive_scan(w);
        if (lane < NW) tmp[lane] = w;
    }
    __syncthreads();
    const T pre = wid > 0 ? tmp[wid - 1] : T(0);
    if (total) *total = tmp[NW - 1];
    __syncthreads();
    return pre + x - v;
}

template <int BLOCK, class T> __device__ __forceinline__ T block_sum(T v, T *tmp) {
    constexpr int NW = BLOCK / 64;
    const int lane = lane_id();
    const int wid = threadIdx.x / 64;
    v = wave_sum(v);
    if (lane == 0) tmp[wid] = v;
    __syncthreads();
    T s = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) s += tmp[i];
    __syncthreads();
    return s;
}

// In-place exclusive scan of an LDS array of n entries (n may exceed BLOCK).
// Returns the sum of all entries (to every thread).
template <int BLOCK>
__device__ __forceinline__ uint32_t lds_exclusive_scan(uint32_t *a, uint32_t n, uint32_t *tmp) {
    const uint32_t per = (n + BLOCK - 1) / BLOCK;
    const uint32_t b = threadIdx.x * per;
    const uint32_t e = min(b + per, n);
    uint32_t s = 0;
    for (uint32_t i = b; i < e; ++i) s += a[i];
    uint32_t total;
    uint32_t pre = block_exclusive_scan<BLOCK>(s, tmp, &total);
    for (uint32_t i = b; i < e; ++i) {
        uint32_t v = a[i];
        a[i] = pre;
        pre += v;
    }
    __syncthreads();
    return total;
}

// Exclusive scan of n LDS counters `a` (u32) into 16-bit starts `out` (the
// caller guarantees every prefix fits 16 bits).  Ends with a barrier.
template <int BLOCK>
__device__ __forceinline__ uint32_t lds_exclusive_scan_u16(const uint32_t *a, uint16_t *out,
                                                           uint32_t n, uint32_t *tmp) {
    const uint32_t per = (n + BLOCK - 1) / BLOCK;
    const uint32_t b = threadIdx.x * per;
    const uint32_t e = min(b + per, n);
    uint32_t s = 0;
    for (uint32_t i = b; i < e; ++i) s += a[i];
    uint32_t total;
    uint32_t pre = block_exclusive_scan<BLOCK>(s, tmp, &total);
    for (uint32_t i = b; i < e; ++i) {
        const uint32_t v = a[i];
        out[i] = (uint16_t)pre;
        pre += v;
    }
    __syncthreads();
    return total;
}

// Largest i in [0, n) with a[i] <= x (a non-decreasing, a[0] <= x).
template <class T>
__device__ __forceinline__ uint64_t last_le(const T *a, uint64_t n, T x) {
    uint64_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint64_t mid = (lo + hi) >> 1;
        if (a[mid] <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

__host__ __device__ __forceinline__ uint32_t fmix32(uint32_t h) {
    // murmur3 finalizer: a bijective avalanche mix of the key
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

__device__ __forceinline__ uint32_t fmix32_inv(uint32_t h) {
    // fmix32 undone step by step: x ^ x >> 16 is its own inverse, x ^ x >> 13
    // is undone by y ^ y >> 13 ^ y >> 26, and each multiplier by its inverse
    // modulo 2^32 (0x85ebca6b * 0xa5cb9243 = 0xc2b2ae35 * 0x7ed1b41d = 1 mod 2^32)
    h ^= h >> 16;
    h *= 0x7ed1b41du;
    h ^= (h >> 13) ^ (h >> 26);
    h *= 0xa5cb9243u;
    h ^= h >> 16;
    return h;
}

}  // namespace aj
