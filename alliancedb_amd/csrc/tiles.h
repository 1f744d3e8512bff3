// tiles.h -- tile loads shared by the partitioning kernels (8-byte tuple_t or
// 12-byte stream tuples, gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "device_common.h"

namespace aj {

constexpr int PITEMS = 16;  // tuples per lane per tile

// Spare counters and slots for lanes without a tuple, one per lane mod 32: a
// single spare would take the atomics and stores of up to 64 lanes of a wave
// on one LDS address (serialised), which in a partly filled tile costs more
// than the rest of the tile's LDS work.
constexpr int NSPARE = 32;

// A streaming (nontemporal) load, or a plain one that may be served by and
// allocate in the caches (NT = false).
template <bool NT, class T> __device__ __forceinline__ T load16(const T *p) {
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

// Tile of PTILE tuples starting at `tb` (tb even for 8-byte tuples): element
// j of this thread is tuple idx[j]; v[j] says whether it lies in [lo, hi).
// 8-byte tuples of a whole tile are read two per lane with 16-byte loads
// (consecutive lanes on consecutive 16 bytes).  In a partly filled tile every
// element is read with its own 8-byte load at its index clamped to hi - 1, so
// all loads are unconditional and in flight together (a load per element under
// a branch made each wait for the one before) and none reaches past `hi`.
// 12-byte stream tuples are read one per lane.

template <int PBLOCK, int NI = PITEMS, bool NT = true>
__device__ __forceinline__ void load_tile(const void *in, int stride, uint64_t tb, uint64_t lo,
                                          uint64_t hi, uint64_t (&t)[NI],
                                          bool (&v)[NI]) {
    if (stride == 8) {
        typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));
        const uint64_t *p = static_cast<const uint64_t *>(in);
        if (tb >= lo && tb + PBLOCK * NI <= hi) {  // whole tile inside: unconditional 16-byte loads
#pragma unroll
            for (int j = 0; j < NI / 2; ++j) {
                const u64x2 q = load16<NT>(
                    reinterpret_cast<const u64x2 *>(p + tb + 2 * ((uint64_t)j * PBLOCK + threadIdx.x)));
                t[2 * j] = q.x;
                t[2 * j + 1] = q.y;
                v[2 * j] = v[2 * j + 1] = true;
            }
            return;
        }
        // the lane index behind an empty asm statement: the per-item offsets
        // below are recomputed here instead of being hoisted out of the tile
        // loop and kept live (or spilled) across it.  Offsets are 32-bit and
        // relative to the tile (uniform base, lo_r <= 1, hi_r <= the tile).
        uint32_t lane = threadIdx.x;
        asm volatile("" : "+v"(lane));
        const uint64_t *pt = p + tb;
        const uint32_t lo_r = lo > tb ? (uint32_t)(lo - tb) : 0u;
        const uint32_t hi_r = (uint32_t)min<uint64_t>(hi - tb, (uint64_t)PBLOCK * NI);
#pragma unroll
        for (int j = 0; j < NI / 2; ++j) {
            const uint32_t r0 = 2 * ((uint32_t)j * PBLOCK + lane), r1 = r0 + 1;
            t[2 * j] = load16<NT>(pt + (r0 < hi_r ? r0 : hi_r - 1));
            t[2 * j + 1] = load16<NT>(pt + (r1 < hi_r ? r1 : hi_r - 1));
            v[2 * j] = r0 >= lo_r && r0 < hi_r;
            v[2 * j + 1] = r1 >= lo_r && r1 < hi_r;
        }
    } else {
        const uint32_t *p = static_cast<const uint32_t *>(in);
        if (tb >= lo && tb + PBLOCK * NI <= hi) {  // whole tile inside: unconditional loads
#pragma unroll
            for (int j = 0; j < NI; ++j) {
                const uint32_t *q = p + 3 * (tb + (uint64_t)j * PBLOCK + threadIdx.x);
                t[j] = ((uint64_t)load16<NT>(q) << 32) | load16<NT>(q + 1);
                v[j] = true;
            }
            return;
        }
        // a partly filled tile: conditional loads (the clamped form needs more
        // registers than the kernels that read stream tuples have left)
#pragma unroll
        for (int j = 0; j < NI; ++j) {
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

}  // namespace aj
