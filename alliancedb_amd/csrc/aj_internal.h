// aj_internal.h -- shared declarations of the HIP join library (not installed).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/alliancedb/aj.h"

namespace aj {

// ---------------------------------------------------------------- errors
int set_error(int code, const char *fmt, ...);  // returns code
int last_status();

#define AJ_HIP(call)                                                              \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess) {                                                   \
            ::aj::set_error(AJ_ERR_HIP, "%s failed at %s:%d: %s", #call, __FILE__, \
                            __LINE__, hipGetErrorString(e_));                     \
            return AJ_ERR_HIP;                                                    \
        }                                                                         \
    } while (0)

#define AJ_TRY(expr)                \
    do {                            \
        int rc_ = (expr);           \
        if (rc_ != AJ_OK) return rc_; \
    } while (0)

// ---------------------------------------------------------------- memory
// Device memory handed out by bumping a cursor through one fixed region (the
// no-allocation entry point's scratch array); reset before each call.
struct ScratchArena {
    uint8_t *base = nullptr;  // device address of the region
    size_t size = 0, used = 0;
};

// A device buffer that only grows; reused across calls so that repeated joins
// of the same size do not allocate.  With an arena, ensure() carves the buffer
// out of the arena instead (AJ_ERR_CAPACITY when it does not fit) and release()
// only forgets it.
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    ScratchArena *arena = nullptr;
    int ensure(size_t bytes);
    void release();
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// Allocation hook (tests inject failures / budgets through aj_set_alloc_hook).
int dev_malloc(void **p, size_t bytes);
void dev_free(void *p);

// ---------------------------------------------------------------- tuning
struct ProConfig {
    int bits1 = 0;        // radix bits of pass 1 (0 = derive from sizes)
    int bits2 = 0;        // radix bits of pass 2 (0 = single pass)
    int cap = 8704;       // R tuples per LDS build round
    int split = 0;        // S tuples per join task (0 = cap; at most cap)
    int chunk = 0;        // tuples per partition work-group (0 = derived from sizes)
    int block = 512;      // lanes per partition work-group (256 or 512)
    bool p2d = true;      // two passes: pass 2 from the 2-D pass-1 histogram when it fits
    bool staged = true;   // two passes of <= 8 bits each: staged partitioning first
    int stage_sets = 0;   // staged: cursor sets (0 = default)
    uint64_t stage_chunk = 0;  // staged: tuples per super-chunk (0 = default)
    int stage_block = 0;  // staged: lanes per work-group (0 = default)
    int stage_cache = -1; // staged: nontemporal access mode (-1 = default)
};
ProConfig choose_pro_config(uint64_t nR, uint64_t nS, const aj_join_opts_t *opts);
// the autotuner's entry for the size bucket of (nR, nS), if any (tune.cpp)
bool tune_lookup(uint64_t nR, uint64_t nS, aj_tune_entry_t *out);

// ---------------------------------------------------------------- context
struct BatchPlan;
struct Ctx {
    int device = 0;
    int num_cu = 256;
    hipStream_t stream = nullptr;
    std::mutex mu;
    // scratch reused across calls
    DevBuf inR, inS;            // uploaded inputs (drop-in path)
    DevBuf p1R, p1S, p2R, p2S;  // partitioned copies
    DevBuf hist, scan_tmp, plan, seg, tasks, acc, pairs, table, misc;
    DevBuf hist2d;              // per-chunk (d1, d2) counts of a two-pass PRO
    DevBuf steps;               // SHJ: per-step match counts
    DevBuf stage;               // staged PRO: staging regions
    DevBuf chain, chain2;       // NPO / SHJ with pairs: duplicate chain links (R, S)
    DevBuf heads, heads2;       // NPO / SHJ with pairs: chain heads per slot (R, S)
    DevBuf carry, carry2;       // SHJ: carries of the per-key payload sums (R, S)
    DevBuf heavy, light;        // PRO on skewed S: heavy-key tables and sums; light S tuples
    hipEvent_t ev0 = nullptr, ev1 = nullptr;  // none: no timing (device_ms 0)
    uint64_t *hres = nullptr;  // pinned host slots for a join's scalars (HRES_SLOTS); none: pageable
    bool no_events = false;     // never create events (no per-kernel timing either)
    bool arena_only = false;    // buffers come from a fixed scratch array (no allocation)
    BatchPlan *pro1 = nullptr;  // one-pass staged PRO: a one-segment window plan
    hipEvent_t evk[AJ_NPHASES + 1] = {};  // per-kernel timing (option kernel_timing)
    std::vector<hipEvent_t> evs;          // per-launch timing of the staged passes
};
bool kernel_timing_enabled();
bool staged_enabled();  // option "staged"
// option npo_region_min: smallest |R| whose NPO table is built by regions (0 = never)
uint64_t npo_region_min();
// A join's scalars (count, checksum, ...) are copied back into the context's
// pinned slots, which the copy engine writes directly: a copy to pageable
// memory goes through a staging buffer and keeps the device idle ~15-30 us
// longer between back-to-back joins.  A context without them (the
// allocation-free entry point's) reads into `local`.
constexpr int HRES_SLOTS = 8;
inline uint64_t *result_slots(Ctx *ctx, uint64_t *local) { return ctx->hres ? ctx->hres : local; }

// Events of a context that may have none: recording is skipped, times are 0.
inline hipError_t record_event(hipEvent_t e, hipStream_t st) {
    return e ? hipEventRecord(e, st) : hipSuccess;
}
inline float elapsed_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    if (!a || !b) return 0;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return ms;
}
inline bool phase_timing(const Ctx *c);
// The join behind aj_join_dev on a context the caller has locked and whose
// device is current.
int join_dev(Ctx *ctx, const uint64_t *dR, uint64_t nR, const uint64_t *dS, uint64_t nS,
             const aj_join_opts_t *opts, aj_join_out_t *out, hipStream_t st);
// Eager symmetric hash join (shj.hip): steps of `batch` tuples per side.
uint64_t shj_steps(uint64_t nR, uint64_t nS, uint64_t batch);
int shj_device(Ctx *ctx, const uint64_t *dR, uint64_t nR, const uint64_t *dS, uint64_t nS,
               uint64_t batch, uint64_t *progress, uint64_t *pairs, uint64_t pcap, int records,
               aj_join_out_t *out, hipStream_t st);
// Writes the device's constant-rate clock to *slot (stream-ordered).
int device_clock_stamp(uint64_t *slot, hipStream_t st);

// Optional event marks between kernels: mark(i) records event i on the stream.
struct Marks {
    hipEvent_t *ev = nullptr;  // nullptr = timing off
    hipStream_t st = nullptr;
    void mark(int i) const {
        if (ev) (void)hipEventRecord(ev[i], st);
    }
};
inline bool phase_timing(const Ctx *c) { return kernel_timing_enabled() && !c->no_events; }
Ctx *default_ctx(int *status);
Ctx *ctx_of(aj_ctx_t *c);

// ---------------------------------------------------------------- kernels
// Segment description of one partition pass, all arrays on the device.
//   in_begin/in_end[2][nseg]: input ranges of R (0) and S (1)
//   bits[nseg]: radix bits of each segment for this pass; shift: digit shift
struct PassDesc {
    const void *in[2];        // input tuples (8-byte or 12-byte stream tuples)
    int in_stride;            // 8 or 12 bytes
    uint64_t *out[2];         // output 8-byte tuples, segment-major / partition-major
    const uint64_t *in_begin[2];
    const uint64_t *in_end[2];
    const uint32_t *bits;     // per segment
    int nseg;
    int shift;
    int max_bits;
    uint32_t chunk;           // tuples per work-group (multiple of the largest tile)
    int block = 512;          // lanes per histogram/scatter work-group: 256 or 512
    bool exact_single = false;  // one segment of exactly n_total[0] / n_total[1] tuples
                                // (the plan is then computed from the host's sizes)
    uint64_t n_total[2];      // total tuples of R and S in this pass
    uint64_t max_chunks;      // upper bound of chunks over R+S (grid size)
    uint64_t max_hist;        // upper bound of histogram entries
    uint64_t max_parts;       // upper bound of partitions per relation
    // one-segment 8-byte pass 1 of a two-pass PRO: also count (d1, d2) pairs
    // per chunk into hist2d (chunks x 2^(max_bits + bits2)), see run_pass2_2d
    uint32_t *hist2d = nullptr;
    int bits2 = 0;
    // nullable: every kernel of the pass except the planning returns at once
    // while *guard is 0 (a fallback that runs only after an overflow)
    const uint32_t *guard = nullptr;
    bool skip_plan = false;  // the planning kernel already ran (run_partition_plan)
};

// Workspace carved out of ctx buffers by the pass (device pointers).
struct PassWork {
    uint32_t *hist;           // max_hist
    uint64_t *seg_chunk_pre[2];   // nseg+1 exclusive chunk prefix per relation
    uint64_t *seg_hist_base[2];   // nseg histogram base per relation
    uint64_t *seg_out_base[2];    // nseg+1 output base per relation
    uint32_t *seg_part_base;      // nseg+1 partition base (shared by R and S)
    uint64_t *header;             // [0]=chunksR [1]=chunksS [2]=histR [3]=histS [4]=parts
    uint64_t *part_off[2];        // max_parts+1 partition offsets per relation
    void *scan_tmp;
    size_t scan_tmp_bytes;
};

size_t pass_scan_tmp_bytes(uint64_t max_hist);
// marks (nullable): m0 after planning, m0+1 after the histogram, m0+2 after the
// scan, m0+3 after the scatter
int run_partition_pass(const PassDesc &d, const PassWork &w, hipStream_t st,
                       const Marks *marks = nullptr, int m0 = 0);
// Only the planning kernel of a multi-segment pass (chunk, output and partition
// bases per segment), for callers that partition by other means first.
int run_partition_plan(const PassDesc &d, const PassWork &w, hipStream_t st);

// Staged window pass (12-byte window tuples, one pass of bits[w] radix bits per
// window, no histogram): every partition of window w gets cap[w] slots at
// base[w] + d * cap[w]; a tile's run for a digit is placed by one atomic add on
// that partition's cursor.  An S run that does not fit its partition goes to a
// spill area (after the S layout) and is listed; run_window_parts then gathers
// each partition's spilled runs into one extension range (after the spill
// area).  When the layout does not fit the output buffers, an R run does not
// fit its partition, or the spill area or its list is full, *flag is set; the
// caller then runs the histogram pass with guard = flag, which rewrites the
// output exactly.
constexpr int WSTAGE_MAX_BITS = 11;  // the LDS of a 1024-lane tile holds 2^11 digits
struct WStageDesc {
    uint32_t *cap;            // 2 * W per-window partition capacities (R, then S)
    uint64_t *base;           // 2 * W per-window output bases
    uint32_t *cur[2];         // max_parts cursors per relation, zeroed by the caller
    uint32_t *flag, *lflag;   // zeroed by the caller
    uint64_t out_cap[2];      // tuples of each layout (the partitions + a trash tile)
    int W;                    // stride of cap / base
    uint64_t *pstart[2];      // out: 2 * max_parts partition starts per relation
    uint32_t *plen[2];        // out: 2 * max_parts partition lengths per relation
                              // (entry max_parts + p: partition p's S extension)
    uint32_t *strad;          // max_parts (no init needed)
    uint32_t *spill_cur, *spill_n, *ext_cur;  // zeroed by the caller
    uint4 *spill_list;        // spill_max entries
    uint64_t spill_cap;       // tuples of the spill area, and of the extension area
    uint32_t spill_max;
    uint32_t *ext_off;        // max_parts (no init needed)
};
// The S output buffer of a staged window pass holds out_cap[1] + 2 * spill_cap tuples.
// Tuples an output buffer needs for n tuples in at most `parts` partitions.
uint64_t window_stage_bound(uint64_t n, uint64_t parts);
// After run_partition_plan: capacities, bases and the scatter.
int run_window_stage(const PassDesc &d, const PassWork &w, const WStageDesc &ws,
                     hipStream_t st);
// After the guarded histogram pass: the partitions' starts and lengths, from
// the staged layout or (after an overflow) the exact one, and the spilled S runs
// gathered into their extensions.
int run_window_parts(const PassDesc &d, const PassWork &w, const WStageDesc &ws,
                     hipStream_t st);

// Pass 2 of a two-pass one-segment PRO driven by the 2-D histogram of pass 1
// (no pass-2 histogram kernel).  Reads the pass-1 output and workspace.
struct P2Desc {
    const uint64_t *in[2];       // pass-1 output
    uint64_t *out[2];            // pass-2 output
    int b1, b2;
    uint32_t *hist2d;            // as filled by pass 1
    uint32_t *unit_counts;       // workspace: max_units * 2^b2
    uint64_t *unit_span;         // workspace: 2 * max_units (input begin | rel << 63, end)
    uint64_t *poff2[2];          // out: 2^(b1+b2)+1 final partition offsets per relation
    uint32_t *ubase;             // workspace: 2 * 2^b1 + 1
    const PassWork *w1;          // pass-1 workspace (offsets, header)
    const uint32_t *hist_scan;   // pass-1 scanned histogram (PassWork::hist + max_hist)
    uint64_t unit;               // pass-2 tuples per work-group (target)
    uint64_t max_units;          // grid bound: (nR + nS) / unit + 2 * 2^b1
    int block;
};
constexpr int P2D_MAX_BITS = 15;  // b1 + b2 for the 2-D histogram path (128 KB of LDS)
int run_pass2_2d(const P2Desc &d, hipStream_t st, const Marks *marks = nullptr, int m0 = 0);

// Two-pass one-segment PRO partitioning streamed in super-chunks through a
// staging buffer, output space reserved by atomics (staged.hip).
struct StagedGeom {
    int b1 = 0, b2 = 0;       // radix bits of pass 1 and pass 2 (each 1..8)
    int nset = 1;             // staging cursor sets (work-group mod nset)
    int block = 1024;         // lanes per pass-2 work-group (512 or 1024)
    int block1 = 1024;        // lanes per pass-1 work-group (512 or 1024)
    int cache_mode = 9;       // nontemporal accesses: 1 input loads, 2 staging stores,
                              // 4 staging loads, 8 final stores; 16: keys hashed in pass 1
    uint64_t chunk = 0;       // tuples per super-chunk (a multiple of the tile)
    uint64_t tiles_per_wg = 0;
    uint32_t grid1 = 0;       // pass-1 work-groups of a full super-chunk
    uint32_t units = 1;       // pass-2 work items per staging region
    uint32_t part = 0;        // tuples per pass-2 work item (0: a region split evenly over units)
    uint32_t per_wg = 1;      // pass-2 work items per work-group
    uint32_t grid2 = 0;       // pass-2 work-groups
    uint32_t scap = 0;        // tuples per staging region (nset * 2^b1 regions)
    uint32_t fcap[2] = {0, 0};  // tuples per final partition of R and S
    uint64_t stage_tuples = 0;  // staging buffer size (regions + a tile of trash)
    uint64_t out_tuples[2] = {0, 0};
};
struct StagedDesc {
    StagedGeom g;
    const uint64_t *in[2];
    uint64_t n[2];
    uint64_t *stage;          // g.stage_tuples
    uint64_t *out[2];         // g.out_tuples[r]: partition p at p * g.fcap[r]
    uint32_t *scur;           // 2 * nset * 2^b1 staging cursors
    uint32_t *fcur;           // 2 * 2^(b1+b2) final cursors (R, then S)
    uint32_t *flag;           // set when a partition overflowed
};
// chunk 0 = 2^25 tuples
// block: lanes per pass-2 work-group; block1 those of pass 1 (0: block)
int staged_geometry(StagedGeom *g, uint64_t nR, uint64_t nS, int b1, int b2, uint64_t chunk,
                    int nset, int block, int num_cu, int block1 = 0);
// events recorded by run_staged when timing: before and after every pass-1
// launch (pass 2 runs between one super-chunk's "after" and the next "before")
int staged_events(const StagedDesc &d);
int run_staged(const StagedDesc &d, hipStream_t st, std::vector<hipEvent_t> *evs = nullptr);
uint64_t staged_max_tasks(const StagedGeom &g, uint32_t split);
// join task records and task_base[0] / task_base[max_parts] from the final cursors
int staged_tasks(const StagedDesc &d, uint32_t split, uint4 *task_info, uint32_t *task_base,
                 uint64_t max_parts, uint64_t max_tasks, hipStream_t st);

// Join of partition pairs (R part_off, S part_off) into per-segment results.
struct JoinDesc {
    const uint64_t *R, *S;          // partitioned tuples
    const uint64_t *roff, *soff;    // partition offsets (nparts+1), device
    const uint32_t *rlen = nullptr, *slen = nullptr;  // nullable: partition p is
                                    // [roff[p], + rlen[p]) instead of up to roff[p+1]
    const uint64_t *nparts_dev;     // number of partitions (device scalar)
    uint64_t max_parts;
    uint64_t nreal = 0;             // nonzero: entry p >= nreal is another S range of
                                    // partition p - nreal (its segment is that one's)
    const uint32_t *part_seg_base;  // nseg+1: first partition of each segment
    int nseg;
    int hash_shift;                 // radix bits of every partition (bucket hash skips them)
    const uint32_t *seg_bits = nullptr;  // nullable: radix bits per segment instead
    int cap;                        // R tuples per LDS round
    uint32_t split;                 // S tuples per task
    uint64_t *seg_count, *seg_sum;  // nseg accumulators (zeroed by the caller)
    uint64_t *pairs;                // nullable: materialise pairs
    uint64_t pairs_cap;
    int records = 0;                // 1: <S.key, S.payloadID> records; 2: timed records
    uint64_t *pair_cursor;          // device scalar (zeroed by the caller)
    uint32_t *task_cnt;             // workspace: max_parts+1
    uint32_t *task_base;            // workspace: max_parts+1
    uint32_t *task_next;            // device scalar (zeroed by the caller)
    uint4 *task_info;               // workspace: 2 * join_max_tasks() entries
    ulonglong2 *task_res;           // workspace: join_max_tasks() partial results
    uint64_t max_tasks;             // join_max_tasks() of this workspace
    void *scan_tmp;
    size_t scan_tmp_bytes;
    int num_cu;
    bool tasks_ready = false;       // task_info and task_base[0], task_base[max_parts]
                                    // filled by the caller (no task planning)
};
// Raises a kernel's dynamic-LDS limit on the current device (thread-safe).
int lds_limit_set(const void *fn, size_t bytes);
size_t join_scan_tmp_bytes(uint64_t max_parts);
// S tuples per join task (0 = 8 LDS rounds of `cap`).  A task probes its
// slice in register-resident sub-slices of `cap` tuples.
uint32_t join_task_split(uint32_t split, int cap);
uint64_t join_max_tasks(uint64_t max_parts, uint64_t max_S, uint32_t split);
// marks (nullable): m0 after task planning, m0+1 after the join kernel
int run_join(const JoinDesc &d, hipStream_t st, const Marks *marks = nullptr, int m0 = 0);

// NPO global table
// Slot geometry of the NPO table: 2^(b1+b2) regions of 2^rs_log2 slots each.
// A key with hash h = fmix32(key) lives in region ((h mod 2^b1) << b2) |
// ((h >> b1) mod 2^b2), from slot (h >> (b1+b2)) mod 2^rs_log2 of it on,
// probing linearly and wrapping inside the region.  b1 = b2 = 0: one region,
// the flat table.
struct NpoGeom {
    uint32_t b1 = 0, b2 = 0, rs_log2 = 0;
};

struct NpoDesc {
    const uint64_t *R, *S;
    uint64_t nR, nS;
    uint64_t *table;     // capacity 16-byte slots (2 words each)
    uint64_t capacity;   // power of two
    // region build (counting joins only): R partitioned by hashed key into the
    // table's regions (the staged passes with keys hashed in pass 1), each
    // region built in LDS; null = the flat build with device-wide atomics
    const StagedDesc *stage = nullptr;
    NpoGeom geom;        // the region geometry when stage is set
    uint32_t *head;      // with pairs: capacity chain heads (R index + 1, 0 = none)
    uint32_t *next;      // with pairs: nR chain links (R index + 1 of the next tuple)
    uint64_t *acc;       // [0]=count [1]=sum [2]=sentinel R tuples [3]=pair cursor
                         // [4]=region build overflow flag (u32; the staged flag)
    uint64_t *pairs;
    uint64_t pairs_cap;
    int records = 0;  // 1: <S.key, S.payloadID> records; 2: (clock << 32 | S.payloadID)
    int num_cu;
};
int run_npo(const NpoDesc &d, hipStream_t st, const Marks *marks = nullptr, int m0 = 0);
size_t npo_acc_words(int num_cu);  // u64 words of NpoDesc::acc

// ---------------------------------------------------------------- batches
// Workspace and launch geometry of batched window joins (windows.hip).
struct BatchLayout {
    size_t rb, re, sb, se, bits, cpre[2], hb[2], ob[2], pb, hdr, poff[2], tcnt, tbase, tinfo, tres, scal;
    size_t wcap, wbase, wcur[2], pstart[2], plen[2], strad, extoff, splist;  // the staged window pass
    size_t bytes;
};

struct BatchPlan {  // sized for at most max_R/max_S tuples in windows, max_win windows
    Ctx *ctx = nullptr;
    uint64_t max_R = 0, max_S = 0, max_window = 0;
    int max_win = 0;
    int max_bits = 0;
    uint32_t chunk = 0;
    int cap = 8192;
    uint32_t split = 1 << 16;
    uint32_t target = 4096;
    uint64_t max_chunks = 0, max_hist = 0, max_parts = 0;
    bool staged = false;           // histogram-free window pass first (option "staged")
    uint64_t out_cap[2] = {0, 0};  // tuples of the R / S layouts
    uint64_t spill_cap = 0;        // tuples of the S spill area (and of the extension area)
    uint32_t spill_max = 0;        // spilled runs listed
    BatchLayout L{};
    DevBuf work, hist, scan_tmp, outR, outS;
    size_t scan_bytes = 0;
    // graph
    bool use_graph = true;
    hipGraphExec_t exec = nullptr;
    hipStream_t cap_stream = nullptr;
    struct Key {
        const void *R, *S, *bounds;
        uint64_t nR, nS;
        int nwin;
        uint64_t *cnt, *sum;
    } key{};
    bool have_key = false;
};


// Sizes the plan (cap, split, target, chunk from opts, 0 = defaults) and
// allocates its workspace.
int batch_plan_setup(BatchPlan *p, Ctx *c, uint64_t max_R, uint64_t max_S, int max_windows,
                     uint64_t max_window_tuples, const aj_batch_opts_t *opts);
int batch_plan_init(BatchPlan *p);
void batch_plan_release(BatchPlan *p);
// Device block of the plan's window ranges: rb | re | sb | se, max_win u64
// entries each (tuple indices into the R and S arrays, end exclusive).
uint64_t *batch_ranges(BatchPlan *p);
// Device word that is nonzero after a run whose staged pass fell back to the
// histogram pass (plans with staged set).
const uint32_t *batch_fallback_flag(BatchPlan *p);
// Enqueues the joins of nwin windows whose ranges are in batch_ranges(p);
// nR/nS = tuples in those windows.  Per-window count/checksum into cnt/sum.
// The inputs are 12-byte aj_stream_tuple_t (stride 12) or 8-byte tuple_t
// (stride 8); marks (nullable): m0+1..m0+3 after the planning, m0+4..m0+8 after
// the partitioning, m0+9 and m0+10 from the join.
int enqueue_ranges(BatchPlan *p, const void *dR, uint64_t nR, const void *dS, uint64_t nS,
                   int nwin, uint64_t *cnt, uint64_t *sum, hipStream_t st, int stride = 12,
                   const Marks *mk = nullptr, int m0 = 0);

// ---------------------------------------------------------------- heavy keys (heavy.hip)
// The hottest keys of a skewed S joined by their aggregates (counting joins):
// cR, sumR, cS, sumS per key from one read of R and S, the other S tuples
// copied out for the partitioned join.
constexpr int HEAVY_MAX = 1024;
struct HeavyDesc {
    const uint64_t *R, *S;
    uint64_t nR, nS;
    const uint64_t *slots;         // heavy_table's slots of nheavy keys, hottest first
    uint32_t nheavy;
    unsigned long long *agg;       // 4 * HEAVY_MAX: cR | sumR | cS | sumS (zeroed by the caller)
    uint64_t *light;               // out: S tuples of the other keys (nS slots)
    unsigned long long *light_n;   // out: their number (zeroed by the caller)
    int num_cu;
};
// m keys of S at stride nS / m into d_keys.
int heavy_sample(const uint64_t *dS, uint64_t nS, uint32_t m, uint32_t *d_keys, hipStream_t st);
// Host: the 4 * HEAVY_MAX hash slots of n keys (key << 32 | 1 + key index, 0 empty).
void heavy_table(const uint32_t *keys, uint32_t n, uint64_t *slots);
int heavy_split(const HeavyDesc &d, hipStream_t st);

// fill kernel (graph-safe replacement of hipMemsetAsync)
int fill_u64(uint64_t *p, uint64_t n, uint64_t v, hipStream_t st);
int fill_u32(uint32_t *p, uint64_t n, uint32_t v, hipStream_t st);
int copy_d2d(void *dst, const void *src, uint64_t bytes, int num_cu, hipStream_t st);
// Prefix sums of n counters into out (not in place), modulo the type's width,
// on st; tmp of scan_tmp_bytes(n, sizeof element) bytes (partition.hip).
size_t scan_tmp_bytes(uint64_t n, size_t elem);
// guard (nullable): the scan's kernels return at once while *guard is 0
int exclusive_scan_u32(const uint32_t *in, uint32_t *out, uint64_t n, void *tmp,
                       size_t tmp_bytes, hipStream_t st, const uint32_t *guard = nullptr);
int inclusive_scan_u64(const uint64_t *in, uint64_t *out, uint64_t n, void *tmp,
                       size_t tmp_bytes, hipStream_t st);
// one random-access calibration launch (aj_access_kernel) over `slots` 16-byte slots
int access_kernel(int kind, uint64_t *tab, uint64_t slots, uint64_t n, int num_cu,
                  hipStream_t st);

}  // namespace aj
