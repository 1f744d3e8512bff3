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
//   k_part_off  partition start offsets for the join
//
// Bytes per tuple: k_hist reads 8, k_scatter reads 8 and writes 8.
#include <hipcub/hipcub.hpp>

#include "aj_internal.h"
#include "device_common.h"

namespace aj {

constexpr int PITEMS = 16;                // tuples per lane per tile
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
};

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

// Zero the histogram tail beyond the entries this launch uses (so the flat
// scan over the host-side bound stays correct).
__global__ void k_zero_tail(uint32_t *hist, const uint64_t *header, uint64_t max_hist) {
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

// Tile of PTILE tuples starting at `tb` (tb even for 8-byte tuples): element
// j of this thread is tuple idx[j]; v[j] says whether it lies in [lo, hi).
// 8-byte tuples are read two per lane with 16-byte loads (consecutive lanes on
// consecutive 16 bytes); a load never reaches past `hi`, so nothing outside
// the segment's array is touched.  12-byte stream tuples are read one per lane.
template <int PBLOCK>
__device__ __forceinline__ void load_tile(const void *in, int stride, uint64_t tb, uint64_t lo,
                                          uint64_t hi, uint64_t (&t)[PITEMS],
                                          bool (&v)[PITEMS]) {
    if (stride == 8) {
        typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));
        const uint64_t *p = static_cast<const uint64_t *>(in);
        if (tb >= lo && tb + PBLOCK * PITEMS <= hi) {  // whole tile inside: unconditional 16-byte loads
#pragma unroll
            for (int j = 0; j < PITEMS / 2; ++j) {
                const u64x2 q = __builtin_nontemporal_load(
                    reinterpret_cast<const u64x2 *>(p + tb + 2 * ((uint64_t)j * PBLOCK + threadIdx.x)));
                t[2 * j] = q.x;
                t[2 * j + 1] = q.y;
                v[2 * j] = v[2 * j + 1] = true;
            }
            return;
        }
#pragma unroll
        for (int j = 0; j < PITEMS / 2; ++j) {
            const uint64_t i0 = tb + 2 * ((uint64_t)j * PBLOCK + threadIdx.x);
            uint64_t a = 0, b = 0;
            if (i0 + 1 < hi) {
                typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));
                const u64x2 q = __builtin_nontemporal_load(reinterpret_cast<const u64x2 *>(p + i0));
                a = q.x;
                b = q.y;
            } else if (i0 < hi) {
                a = p[i0];
            }
            t[2 * j] = a;
            t[2 * j + 1] = b;
            v[2 * j] = i0 >= lo && i0 < hi;
            v[2 * j + 1] = i0 + 1 >= lo && i0 + 1 < hi;
        }
    } else {
#pragma unroll
        for (int j = 0; j < PITEMS; ++j) {
            const uint64_t i = tb + (uint64_t)j * PBLOCK + threadIdx.x;
            v[j] = i >= lo && i < hi;
            t[j] = v[j] ? load_tuple(in, stride, i) : 0;
        }
    }
}

// first tile start of a chunk: 8-byte tuples start on an even index so that
// 16-byte loads are aligned (the element before `begin` is masked out)
__device__ __forceinline__ uint64_t tile_start(uint64_t begin, int stride) {
    return stride == 8 ? (begin & ~1ull) : begin;
}

// ---------------------------------------------------------------- k_hist
template <int STRIDE, int PBLOCK>
__global__ __launch_bounds__(PBLOCK) void k_hist(PassArgs a) {
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

// ---------------------------------------------------------------- k_part_off
__global__ void k_part_off(PassArgs a) {
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

// ---------------------------------------------------------------- host side
size_t pass_scan_tmp_bytes(uint64_t max_hist) {
    size_t bytes = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (uint32_t *)nullptr, (uint32_t *)nullptr,
                                     (int)max_hist);
    return bytes;
}

// Raises a kernel's dynamic-LDS limit once to the largest size requested so far
// (callers hold the context lock).
static int lds_limit_set(const void *fn, size_t bytes) {
    static const void *fns[16] = {};
    static size_t set[16] = {};
    int i = 0;
    while (i < 16 && fns[i] && fns[i] != fn) ++i;
    if (i == 16) i = 15;
    if (!fns[i]) fns[i] = fn;
    if (bytes > set[i]) {
        AJ_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
        set[i] = bytes;
    }
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
    if (mk) mk->mark(m0);
    if (d.max_chunks)
        hipLaunchKernelGGL((k_hist<STRIDE, PBLOCK>), dim3((unsigned)d.max_chunks), dim3(PBLOCK),
                           lds_hist, st, a);
    if (mk) mk->mark(m0 + 1);
    size_t tmp_bytes = w.scan_tmp_bytes;
    if (d.max_hist) {
        hipError_t e = hipcub::DeviceScan::ExclusiveSum(w.scan_tmp, tmp_bytes, a.hist,
                                                        a.hist_scan, (int)d.max_hist, st);
        if (e != hipSuccess) {
            set_error(AJ_ERR_HIP, "histogram scan failed: %s", hipGetErrorString(e));
            return AJ_ERR_HIP;
        }
    }
    if (mk) mk->mark(m0 + 2);
    if (d.max_chunks)
        hipLaunchKernelGGL((k_scatter<STRIDE, PBLOCK>), dim3((unsigned)d.max_chunks),
                           dim3(PBLOCK), lds_scatter, st, a);
    if (mk) mk->mark(m0 + 3);
    return AJ_OK;
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
    if (d.n_total[0] + d.n_total[1] >= (1ull << 32)) {
        set_error(AJ_ERR_ARG, "|R|+|S| per launch must be below 2^32");
        return AJ_ERR_ARG;
    }
    if (d.max_hist >= (1ull << 31)) {
        set_error(AJ_ERR_ARG, "histogram too large (%llu entries)", (unsigned long long)d.max_hist);
        return AJ_ERR_ARG;
    }
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

    hipLaunchKernelGGL(k_plan, dim3(1), dim3(PLAN_BLOCK), 0, st, a);
    hipLaunchKernelGGL(k_zero_tail, dim3(256), dim3(256), 0, st, a.hist, a.header, d.max_hist);
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

}  // namespace aj
