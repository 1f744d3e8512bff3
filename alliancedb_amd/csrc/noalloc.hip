// noalloc.hip -- aj_join_dev_noalloc: the device join without device allocation.
//
// Every workspace buffer of the join (partition copies, histograms, planning
// arrays, task records, the NPO table, counters) is carved out of `scratch`, a
// file-scope device array of AJ_NOALLOC_SCRATCH_BYTES, by a bump cursor that is
// reset at the start of each call; nothing persists between calls, and every
// counter the kernels need at zero is cleared by the join's own setup kernels.
// In host code `scratch` names a host-side stand-in, so its device address on
// each GPU is looked up once (hipGetSymbolAddress, with that device current) and
// only that address is handed to the kernels.  A call holds its device's lock
// until its result is on the host, so calls on one GPU never overlap on the
// array (without the lock they would race on it).
#include <mutex>

#include "aj_internal.h"

namespace aj {

static __device__ __align__(256) uint8_t scratch[AJ_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;

struct NoallocState {
    std::mutex mu;
    bool ready = false;
    uint8_t *ptr = nullptr;  // device address of this device's copy of scratch
    ScratchArena arena;
    Ctx ctx;
};

NoallocState g_noalloc[NOALLOC_MAX_DEVICES];

template <class F> void for_each_buf(Ctx &c, F f) {
    DevBuf *bufs[] = {&c.inR,   &c.inS,  &c.p1R,   &c.p1S,  &c.p2R,  &c.p2S,    &c.hist,
                      &c.scan_tmp, &c.plan, &c.seg, &c.tasks, &c.acc, &c.pairs, &c.table,
                      &c.misc,  &c.hist2d, &c.steps, &c.stage, &c.chain, &c.chain2, &c.heads,
                      &c.heads2, &c.carry, &c.carry2};
    for (DevBuf *b : bufs) f(*b);
}

// Once per device, with that device current: the array's device address, and a
// context whose buffers all come from it and which creates no events.
int noalloc_setup(NoallocState &s, int dev) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        return set_error(AJ_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                         hipGetErrorString(e));
    }
    int ncu = 0;
    AJ_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    s.ptr = static_cast<uint8_t *>(p);
    s.arena.base = s.ptr;
    s.arena.size = AJ_NOALLOC_SCRATCH_BYTES;
    s.arena.used = 0;
    s.ctx.device = dev;
    s.ctx.num_cu = ncu > 0 ? ncu : 256;
    s.ctx.stream = nullptr;
    s.ctx.no_events = true;
    s.ctx.arena_only = true;
    for_each_buf(s.ctx, [&](DevBuf &b) { b.arena = &s.arena; });
    s.ready = true;
    return AJ_OK;
}

}  // namespace
}  // namespace aj

using namespace aj;

extern "C" int aj_join_dev_noalloc(const tuple_t *dR, uint64_t nR, const tuple_t *dS,
                                   uint64_t nS, const aj_join_opts_t *opts, aj_join_out_t *out,
                                   void *stream) {
    if (!out || (nR && !dR) || (nS && !dS))
        return set_error(AJ_ERR_ARG, "aj_join_dev_noalloc: NULL argument");
    const int algo = opts ? opts->algo : AJ_ALGO_PRO;
    if (algo != AJ_ALGO_PRO && algo != AJ_ALGO_NPO)
        return set_error(AJ_ERR_ARG, "aj_join_dev_noalloc joins PRO or NPO (algo %d)", algo);
    if (nR > AJ_NOALLOC_MAX_TUPLES || nS > AJ_NOALLOC_MAX_TUPLES)
        return set_error(AJ_ERR_CAPACITY,
                         "aj_join_dev_noalloc: at most %u tuples per relation (|R| = %llu, "
                         "|S| = %llu)",
                         AJ_NOALLOC_MAX_TUPLES, (unsigned long long)nR, (unsigned long long)nS);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(AJ_ERR_NODEV, "aj_join_dev_noalloc: no current HIP device");
    }
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return set_error(AJ_ERR_ARG, "aj_join_dev_noalloc: device %d out of range", dev);
    NoallocState &s = g_noalloc[dev];
    std::lock_guard<std::mutex> lk(s.mu);
    if (!s.ready) AJ_TRY(noalloc_setup(s, dev));
    s.arena.used = 0;
    for_each_buf(s.ctx, [](DevBuf &b) {
        b.p = nullptr;
        b.cap = 0;
    });
    return join_dev(&s.ctx, reinterpret_cast<const uint64_t *>(dR), nR,
                    reinterpret_cast<const uint64_t *>(dS), nS, opts, out,
                    static_cast<hipStream_t>(stream));
}
