/*
 * alliancedb/aj.h -- C ABI of the MI355X lazy intra-window equi-join library.
 *
 * Drop-in entry points (same names, same argument meaning as the reference's
 * joins; see each declaration for the interface it replaces), plus a device
 * API for callers that keep tuples resident in HBM, a batched-window API and a
 * serving loop.  Every non-drop-in call returns an aj_status (0 = AJ_OK);
 * aj_last_error() describes the last failure of the calling thread.
 *
 * No C++ or torch types cross this boundary.
 */
#ifndef ALLIANCEDB_AJ_H
#define ALLIANCEDB_AJ_H

#include <stddef.h>
#include <stdint.h>

#include "types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status codes */
enum aj_status {
    AJ_OK = 0,
    AJ_ERR_ARG = -1,      /* bad argument (NULL relation, size out of range, ...) */
    AJ_ERR_NOMEM = -2,    /* device or host allocation failed */
    AJ_ERR_HIP = -3,      /* HIP runtime error */
    AJ_ERR_CAPACITY = -4, /* output buffer too small / budget exceeded */
    AJ_ERR_STATE = -5,    /* call not valid in the object's current state */
    AJ_ERR_NODEV = -6     /* no usable GPU */
};

const char *aj_last_error(void);
int aj_last_status(void);

/* --------------------------------------------------------- drop-in joins */

/*
 * PRO: radix-partitioned hash join on the GPU.
 * Replaces `result_t *PRO(relation_t *relR, relation_t *relS, param_t)`
 * (hashing/joins/parallel_radix_join.h:37, parallel_radix_join.cpp:1704) in
 * its original ETH form `result_t *join(relation_t*, relation_t*, int)`
 * (hashing/test/test.cpp:70).  nthreads is accepted for signature
 * compatibility and recorded in result_t.nthreads; the GPU path ignores it.
 * Host relations are uploaded, joined on the current default device and
 * never modified (stricter than the reference, whose two-pass PRO overwrites
 * its inputs, parallel_radix_join.cpp:1176-1181; no RELATION_PADDING needed).
 * Returns a malloc'ed result, or NULL on error (see aj_last_error).
 * Extended results: aj_result_ext().  Release it with aj_free_result(), or
 * with free() when it carries no record lists (option "materialize" 0 or 1).
 */
result_t *PRO(relation_t *relR, relation_t *relS, int nthreads);

/*
 * NPO: no-partitioning hash join on the GPU (one global hash table).
 * Replaces `result_t *NPO(relation_t*, relation_t*, param_t)`
 * (hashing/joins/no_partitioning_join.h:36, no_partitioning_join.cpp:383).
 * Same contract as PRO.
 */
result_t *NPO(relation_t *relR, relation_t *relS, int nthreads);

/*
 * Extended result.  base is the first member, so the pointer PRO/NPO return
 * is the address of the whole block and free(result) releases everything,
 * pairs included (they live in the same allocation).
 */
typedef struct aj_result_t {
    result_t base;
    uint64_t checksum; /* sum of (R.payloadID << 32 | S.payloadID) over matches, mod 2^64 */
    uint64_t npairs;   /* number of materialised pairs (0 unless enabled) */
    uint64_t *pairs;   /* (R.payloadID << 32 | S.payloadID), unordered */
    int status;
    float device_ms;   /* join time on the device, inputs resident */
    float total_ms;    /* including host<->device copies */
    /* option "timing" = 1 (see aj_write_timing_files): */
    uint64_t ntimed;   /* timed records, one per match */
    uint64_t *timed;   /* (device clock ticks since the join started << 32) | S.payloadID */
    double clock_hz;   /* tick rate of timed (the device's constant-rate clock) */
    int algo;          /* AJ_ALGO_PRO, AJ_ALGO_NPO or AJ_ALGO_SHJ */
    float phase_ms[12]; /* per-phase device times, indexed by enum aj_phase */
    uint64_t ntimed_r; /* SHJ: the last ntimed_r timed records are R tuples' probes
                          (R.payloadID in the low word); 0 for PRO and NPO */
} aj_result_t;

/*
 * Eager symmetric hash join (SHJ) on the GPU: every tuple is inserted into its
 * side's hash table and probed against the other side's, in arrival (index)
 * order, in steps of aj_join_opts_t.batch tuples per side (R part first), so
 * each match is produced in the step of its later tuple.  Replaces the
 * reference's eager joins `result_t *SHJ_JM_NP(relation_t*, relation_t*,
 * param_t)` and `SHJ_JBCR_NP(...)` (hashing/joins/onlinejoins.h:33,62,
 * onlinejoins.cpp:166-181, 220-235) and the single-thread `SHJ_st`
 * (onlinejoins.h:94, onlinejoins.cpp:132-162): the join-matrix and
 * join-biclique/ContRand layouts only decide which CPU thread sees which tuple,
 * and all three return the same matches.  nthreads is recorded only.  Count
 * and checksum as PRO.  Option materialize 1 returns the pairs, 2 the
 * reference's records: per match the probing (later-arriving) tuple's
 * <key, payloadID>, as SHJJoiner::join's probe appends them
 * (common_functions.cpp:251-257), all in result list entry 0.  Option timing
 * 1 gives timed records of both probe sides (aj_result_t.ntimed_r), which
 * aj_write_timing_files turns into the reference's files as its merge does
 * (R probes against R's timestamps, S probes against S's).
 */
result_t *SHJ_JM_NP(relation_t *relR, relation_t *relS, int nthreads);
result_t *SHJ_JBCR_NP(relation_t *relR, relation_t *relS, int nthreads);
result_t *SHJ_st(relation_t *relR, relation_t *relS, int nthreads);

/*
 * The multithreaded CPU joins with the same signature and result contract:
 * the baseline a driver selects by name beside the GPU joins (the reference's
 * algos[] registry, hashing/main.cpp:306-338, selects PRO and NPO the same
 * way).  PRO_CPU restates the reference's PRO (parallel_radix_join.cpp:
 * 981-1473): pass 1 with per-thread histograms, pass 2 per pass-1 partition,
 * a bucket-chaining join per partition pair, radix bits and passes from the
 * options "radix_bits" / "passes" (0 = the reference's defaults, 8 bits in 2
 * passes, prj_params.h:16, 21).  NPO_CPU restates NPO
 * (no_partitioning_join.cpp:156-476): one chained table of next_pow2(|R| / 2)
 * buckets indexed by key & (buckets - 1), built and probed by all threads.
 * They run on nthreads host threads (<= 0: the host's hardware threads), never
 * modify their inputs, return count and checksum (aj_result_t.total_ms = wall
 * time including thread start-up and temporaries), and ignore the
 * materialize and timing options.  The GPU entry points never call them.
 */
result_t *PRO_CPU(relation_t *relR, relation_t *relS, int nthreads);
result_t *NPO_CPU(relation_t *relR, relation_t *relS, int nthreads);

/* Returns the extension of a result from PRO/NPO, or NULL if r is NULL. */
const aj_result_t *aj_result_ext(const result_t *r);

/*
 * Record lists in the reference's format (option "materialize" = 2): the
 * result's resultlist has max(nthreads, 1) separately malloc'ed entries, as the
 * reference's joins return (parallel_radix_join.cpp:1522-1524); entry i has
 * threadid i and its results point at an aj_chainedtuplebuffer_t (types.h).
 * Each match contributes one record, the matching S tuple <S.key, S.payloadID>
 * (parallel_radix_join.cpp:333-338, common_functions.cpp:252-257).  All
 * records are in entry 0; the other entries are empty chains.  Record order is
 * unspecified, as across the reference's threads.
 *
 * aj_free_result releases a result of PRO/NPO with everything it owns (record
 * chains, resultlist, pairs); NULL is a no-op.  Callers written for the
 * reference, which frees resultlist and the result (benchmark.cpp:255-268) and
 * not the chains, keep working but leak the chains.
 */
void aj_free_result(result_t *r);

/*
 * Writes a result's records as the reference's write_result_relation does
 * (tuple_buffer.h:155-236, unsorted variant; called for "Out.tbl" at
 * benchmark.cpp:262-265): per resultlist entry, newest record first, one
 * "%d %d\n" line of key and payloadID.  AJ_ERR_STATE if the result counts
 * matches but carries no record lists.
 */
int aj_write_result_file(const result_t *r, const char *path);

/*
 * Process-wide options of the drop-in entry points (the reference sets the
 * equivalents at compile time: JOIN_RESULT_MATERIALIZE, NUM_RADIX_BITS,
 * NUM_PASSES; common_functions.h:18, prj_params.h:16-21).
 *   "materialize"  0..2  1: also return the pair list (aj_result_t.pairs);
 *                        2: return the reference's per-thread record lists
 *                        (result_t.resultlist, see aj_free_result); default 0
 *   "radix_bits"   0..24 total radix bits (0 = tuned default)
 *   "passes"       0..2  partition passes (0 = tuned default)
 *   "device"       >=0   HIP device of the default context
 *   "kernel_timing" 0/1  record per-kernel device times (aj_join_out_t.phase_ms)
 *   "cap"          0, 2048, 4096, 8192, 8704, 12288  R tuples per LDS build
 *                        round (0 = default 8704)
 *   "split"        >=0   S tuples per join task, probed in slices of cap (0 = 8 x cap)
 *   "part_block"   0, 256, 512  lanes per partition work-group (0 = 512)
 *   "part_chunk"   multiple of 8192  tuples per partition work-group (0 = derived)
 *   "part_hist2d"  0/1   two-pass PRO: pass 2 from a (d1, d2) histogram taken in
 *                        pass 1 instead of its own histogram pass (default 1)
 *   "tuned"        0/1   use the autotuner's table for open geometries (default 1)
 *   "timing"       0/1   drop-in results also carry timed records and per-phase
 *                        device times for aj_write_timing_files (default 0)
 *   "device_budget" bytes of device memory the library may hold (0 = unlimited)
 *   "staged"       0/1   two-pass PRO with at most 8 bits per pass: partition both
 *                        passes super-chunk by super-chunk through a staging buffer
 *                        with atomically reserved, over-allocated partitions (no
 *                        histogram pass); an input whose partitions overflow is
 *                        detected on the device and joined again with the
 *                        histogram passes.  One PRO pass of at most 11 bits
 *                        (counting joins of >= 2^20 tuples) and window batches
 *                        (plans of at most 11 radix bits per window): one
 *                        histogram-free pass into
 *                        over-allocated partitions, S runs past a partition's
 *                        capacity spilled and gathered into one extension range
 *                        per partition; R overflow, or a full spill area, runs the
 *                        histogram pass on the device instead.  Read when a plan
 *                        or server is created (default 1)
 *   "heavy"        0/1   counting PRO joins of |S| >= 2^22 and >= 4 |R|: when keys
 *                        drawn 4+ times in a 2^11-tuple strided sample of S hold >=
 *                        40% of it, the keys drawn 8+ times in a 2^14-tuple sample
 *                        (the hottest 1024) are joined by their aggregates (count
 *                        and payload sum per key and side); the partitioned join
 *                        takes R and the other S tuples.  The result is the same
 *                        (default 1)
 *   "stage_sets"   0..64 staged: cursor sets of the staging regions (0 = 4)
 *   "stage_chunk"  >=0   staged: tuples per super-chunk (0 = 2^25)
 *   "stage_block"  0, 512, 1024  staged: lanes per work-group (0 = 1024)
 *   "stage_cache"  -1..15 staged: nontemporal accesses, a sum of 1 (input loads),
 *                        2 (staging stores), 4 (staging loads), 8 (partition
 *                        stores); -1 = the default (9); modes 0, 1, 5, 9, 13, 15
 *                        have kernels for 1024 lanes, 1, 5, 9 for 512, others are
 *                        AJ_ERR_ARG
 *   "npo_region_min" >=0 NPO: counting joins of at least this many R tuples build
 *                        the table region by region (R partitioned by hashed key,
 *                        each region built in LDS); 0 = always the flat build
 *                        (default 2^22)
 *   "shj_batch"    >=0   tuples per side per eager step of the SHJ drop-ins
 *                        (0 = max(65536, max(|R|, |S|) / 256)).  The reference's
 *                        shj alternates R[i], S[i] one tuple at a time
 *                        (localjoiner.cpp:43-52); 1 gives exactly its order, so
 *                        the records (materialize 2) and the timed records'
 *                        R/S split are the reference's.  With a larger step a
 *                        pair (r_i, s_j) of one step is recorded by its S tuple
 *                        (the R part of a step is inserted first), where the
 *                        reference records the later of the two; count and
 *                        checksum never change.  At most 2^20 steps per call.
 */
int aj_set_option(const char *name, int64_t value);
int64_t aj_get_option(const char *name);

/*
 * Progressive-result files of a drop-in result taken with option "timing" = 1,
 * in the layout the reference writes under its EXP_DIR when built with timing
 * (dump_timing, merge and sortRecords: timer/t_timer.cpp:31-99, 401-452;
 * breakdown_global: t_timer.cpp:113-140; call sites
 * parallel_radix_join.cpp:1628-1656 and no_partitioning_join.cpp:433-459):
 *   <exp_dir>/results/records/<name>_<exp_id>.txt     inputs, matches
 *   <exp_dir>/results/timestamps/<name>_<exp_id>.txt  per sampled match, the time
 *       it was produced (ms since the join started) + lastTS, ascending
 *   <exp_dir>/results/latency/<name>_<exp_id>.txt     per sampled match,
 *       (produced - S tuple's ts + lastTS) in ms, + lastTS again, ascending
 *   <exp_dir>/results/gaps/<name>_<exp_id>.txt        S.payloadID - sample index
 *       (+ lastTS), ascending
 *   <exp_dir>/results/breakdown/<name>_<exp_id>.txt   wait, partition, build,
 *       sort, merge, join, others in 2.1 GHz cycles per input tuple
 * name is the reference's "PRJ" (PRO) or "NPJ" (NPO); lastTS is the window in
 * ms (0 for a relation without timestamps); S timestamps are relS->payload->ts
 * in 2.1 GHz cycles as aj_load_relation gives them (0 when relS->payload is
 * NULL).  Every gap-th match in production order is sampled (the reference's
 * record_gap; 1 = all).  Matches are timed per join task (PRO) or per probe
 * block (NPO) with the device clock; the production order of records with the
 * same time is by S.payloadID.  Folders are created as needed.  Fills
 * summary (nullable) with what the reference prints: the time to 1, 5, 10, 25
 * and 50% of the sampled results and the 95th / 99th latency, in ms.
 */
typedef struct aj_timing_summary_t {
    double progress_ms[5]; /* 1, 5, 10, 25, 50 % */
    double latency95_ms, latency99_ms;
    uint64_t samples;
} aj_timing_summary_t;

int aj_write_timing_files(const result_t *r, const relation_t *relR, const relation_t *relS,
                          const char *exp_dir, const char *name, int exp_id, long lastTS,
                          int gap, aj_timing_summary_t *summary);

/* ------------------------------------------------------------ input files */

/*
 * Loads a relation from the reference's text format.  Replaces
 * `int load_relation(relation_t*, relation_payload_t*, int32_t keyby,
 * int32_t tsKey, char *filename, uint64_t num_tuples, uint32_t partitions)`
 * (hashing/utils/generator.h:124, generator.cpp:645-666, read_relation
 * :899-1027, smooth :1050-1067), the loader behind the driver's -R/-S files:
 *   - fields are separated by the first ',' or '|' found in the file; line i
 *     (no header) gives the int32 key in field keyby and, when tsKey != 0, the
 *     timestamp field tsKey * 2.1e6 (cycles at 2.1 GHz of a millisecond column);
 *   - runs of equal timestamps are spread linearly towards the next value;
 *   - line i lands in partition i % partitions of n / partitions tuples (the
 *     last partition takes the remainder and any overflow), payloadID = its
 *     index, pl->ts[index] = its timestamp.
 * rel->tuples is allocated here (64-byte aligned; release with free()).  pl
 * may be NULL (timestamps dropped); if pl->ts is NULL it is allocated here.
 * Returns AJ_OK or an aj_status; inputs on which the reference's behaviour is
 * undefined (no separator, too few lines, a missing or non-numeric field, a
 * key outside int32, a negative timestamp, smoothing overflow) are AJ_ERR_ARG.
 * With tsKey == 0 every timestamp is 0 (the reference leaves them unset).
 */
int aj_load_relation(relation_t *rel, relation_payload_t *pl, int32_t keyby, int32_t tsKey,
                     const char *filename, uint64_t num_tuples, uint32_t partitions);

/* ------------------------------------------------------------ device API */

/* Device memory helpers for FFI callers that have no HIP binding of their own
 * (no counterpart in the reference, whose relations live in host memory).
 * They act on the default context's device. */
int aj_device_alloc(void **ptr, uint64_t bytes);
int aj_device_free(void *ptr);
int aj_copy_h2d(void *dst_device, const void *src_host, uint64_t bytes);
int aj_copy_d2h(void *dst_host, const void *src_device, uint64_t bytes);
/* Device-to-device copy by a plain streaming kernel (aj_copy_kernel), also
 * the reference point for calibrating memory-traffic counters. */
int aj_copy_d2d(void *dst_device, const void *src_device, uint64_t bytes);
/* Random-access ceilings of the hash-table joins (no counterpart in the
 * reference; measurement only).  One timed launch over a caller's table of
 * `slots` 16-byte slots (a power of two): n random 16-byte slot loads
 * (AJ_ACCESS_LOAD, an NPO probe's pattern), n random 64-bit CAS claims of empty
 * words (AJ_ACCESS_CAS, an insert's), or both at independent slots
 * (AJ_ACCESS_LOAD_CAS, an SHJ step's).  The table is first filled with the
 * empty marker (not timed); *ms receives the launch's device time. */
enum aj_access_kind { AJ_ACCESS_LOAD = 0, AJ_ACCESS_CAS = 1, AJ_ACCESS_LOAD_CAS = 2 };
int aj_access_kernel(int kind, void *table_device, uint64_t slots, uint64_t n, float *ms);
int aj_device_sync(void);

typedef struct aj_ctx aj_ctx_t;

int aj_ctx_create(int device, aj_ctx_t **out);
int aj_ctx_destroy(aj_ctx_t *ctx);

enum aj_algo { AJ_ALGO_PRO = 0, AJ_ALGO_NPO = 1, AJ_ALGO_SHJ = 2 };

typedef struct aj_join_opts_t {
    int algo;          /* AJ_ALGO_PRO or AJ_ALGO_NPO */
    int radix_bits;    /* PRO: total radix bits, 0 = tuned */
    int passes;        /* PRO: 1 or 2, 0 = tuned */
    int cap;           /* PRO: R tuples per LDS build round, 0 = default */
    uint32_t split;    /* PRO: S tuples per join task (probed in slices of cap), 0 = 8 x cap */
    uint64_t *pairs;   /* device buffer for pairs, NULL = count/checksum only */
    uint64_t pairs_capacity;
    int records;       /* 1: write the reference's result records <S.key, S.payloadID>
                          (the matching S tuple, one per match; for SHJ the probing,
                          later-arriving tuple) into pairs instead of
                          (R.payloadID << 32 | S.payloadID); 2: timed records
                          (device clock << 32 | S.payloadID; SHJ: of the probing
                          tuple, R probes from the end, out.nrecords_r); count and
                          checksum unchanged */
    uint64_t batch;    /* SHJ: tuples per side per eager step, 0 = max(65536, n / 256) */
    uint64_t *progress; /* SHJ (nullable, host or device): matches after each step,
                          ceil(max(|R|, |S|) / batch) entries */
} aj_join_opts_t;

/* Per-kernel device times of the last join, filled when the option
 * "kernel_timing" is 1 (HIP events around each launch, on the join stream). */
enum aj_phase {
    AJ_PH_PLAN1 = 0,    /* PRO pass-1 planning (segment tables) */
    AJ_PH_HIST1 = 1,    /* PRO pass-1 histogram kernel */
    AJ_PH_SCAN1 = 2,    /* PRO pass-1 histogram scan */
    AJ_PH_SCATTER1 = 3, /* PRO pass-1 scatter kernel */
    AJ_PH_PLAN2 = 4,
    AJ_PH_HIST2 = 5,
    AJ_PH_SCAN2 = 6,
    AJ_PH_SCATTER2 = 7,
    AJ_PH_TASKS = 8,    /* PRO join task planning */
    AJ_PH_JOIN = 9,     /* PRO build/probe kernel */
    AJ_PH_NPO_BUILD = 10,
    AJ_PH_NPO_PROBE = 11,
    AJ_NPHASES = 12
};

typedef struct aj_join_out_t {
    uint64_t count;
    uint64_t checksum;
    uint64_t npairs;
    float device_ms;
    float phase_ms[AJ_NPHASES];
    uint64_t nrecords_r; /* SHJ timed records (records = 2): R tuples' probes, written
                            from the end of the pair buffer backwards */
    uint32_t flags;      /* PRO: AJ_OUT_STAGED when the staged partitioning produced
                            the result, AJ_OUT_FALLBACK when a staged attempt overflowed
                            and the histogram passes joined again (device_ms covers both).
                            NPO: AJ_OUT_STAGED when the table was built region by region,
                            AJ_OUT_FALLBACK when that overflowed and the flat build ran
                            in its place (on the device); 0 for the flat build alone.
                            PRO also AJ_OUT_HEAVY when the hottest keys of S were joined
                            by their aggregates (option "heavy"). */
    float heavy_ms;      /* PRO with AJ_OUT_HEAVY: device time of the pass that splits off
                            the hottest keys (one read of R and S, the rest of S written
                            out); part of device_ms.  0 otherwise. */
    uint64_t light;      /* PRO with AJ_OUT_HEAVY: S tuples left to the partitioned join
                            (those of the keys not joined by their aggregates); 0 otherwise. */
} aj_join_out_t;

enum aj_out_flags { AJ_OUT_STAGED = 1, AJ_OUT_FALLBACK = 2, AJ_OUT_HEAVY = 4 };

/*
 * Join device-resident relations (8-byte tuple_t arrays in HBM) on `stream`
 * (a hipStream_t, NULL = the context's stream).  Blocks until the result is
 * on the host.  If opts->pairs is set and too small, returns AJ_ERR_CAPACITY
 * with out->count holding the required size.
 */
int aj_join_dev(aj_ctx_t *ctx, const tuple_t *dR, uint64_t nR, const tuple_t *dS,
                uint64_t nS, const aj_join_opts_t *opts, aj_join_out_t *out,
                void *stream);

/*
 * aj_join_dev without any device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (a serving loop's hot
 * path, code that must not call hipMalloc or hipFree, harnesses that forbid
 * allocation).  Joins PRO or NPO (opts->algo; NULL opts = PRO) of device-resident
 * relations of at most AJ_NOALLOC_MAX_TUPLES tuples each, on `stream` (a
 * hipStream_t, NULL = the null stream) of the current device, and blocks until
 * the result is on the host, like aj_join_dev.  Every workspace buffer is carved
 * from one fixed array of AJ_NOALLOC_SCRATCH_BYTES in the library's device code:
 * each GPU the library runs on holds its own copy until the process ends, and the
 * first call on a device looks up that copy's address once.  No hipMalloc,
 * hipFree or other allocating call is made.  All calls on one GPU share its
 * array, which keeps the last call's contents; every buffer the kernels need
 * zeroed is cleared at the start of each call, and a call holds the device's lock
 * until its result is on the host, so calls on one GPU never overlap (calls that
 * overlapped on the array would race).  Larger relations, and geometries whose
 * workspace does not fit the array, are AJ_ERR_CAPACITY; algo SHJ is AJ_ERR_ARG.
 * No events are created: device_ms and phase_ms are 0.  The caller's opts->pairs
 * buffer (device memory it owns) may hold pairs or records as in aj_join_dev.
 */
#define AJ_NOALLOC_MAX_TUPLES (1u << 20)
#define AJ_NOALLOC_SCRATCH_BYTES (64u << 20)
int aj_join_dev_noalloc(const tuple_t *dR, uint64_t nR, const tuple_t *dS, uint64_t nS,
                        const aj_join_opts_t *opts, aj_join_out_t *out, void *stream);

/* ------------------------------------------------------- batched windows */
/*
 * Many tumbling windows of two timestamp-ordered streams of 12-byte
 * aj_stream_tuple_t, joined window by window in one graph launch.  Window w
 * holds the tuples with d_bounds[w] <= ts < d_bounds[w+1]; per-window count
 * and checksum land in d_count[w] / d_checksum[w] (device arrays of nwin).
 * Each window's result equals PRO on that window's tuples (the reference runs
 * one window per process, benchmark.cpp:164-257; there is no batched
 * counterpart in the reference).  The plan owns the scratch for its maximum
 * sizes; the kernel sequence is captured once into a hipGraph and replayed.
 * With the option "staged" (default) the windows are partitioned without a
 * histogram pass; whether a batch fell back to the histogram pass (skewed R
 * keys) is decided on the device and never changes a result.
 */
typedef struct aj_batch_opts_t {
    int cap;          /* R tuples per LDS round (4096/8192/8704/12288), 0 = 8704 */
    uint32_t split;   /* S tuples per join task, 0 = cap, at most cap */
    uint32_t target;  /* largest R tuples per partition aimed at, 0 = the largest
                         power of two <= cap */
    uint32_t chunk;   /* tuples per partition work-group, 0 = derived */
    int no_graph;     /* 1 = enqueue kernels directly instead of a graph */
} aj_batch_opts_t;

typedef struct aj_batch_plan aj_batch_plan_t;

int aj_batch_plan_create(aj_ctx_t *ctx, uint64_t max_R, uint64_t max_S, int max_windows,
                         uint64_t max_window_tuples, const aj_batch_opts_t *opts,
                         aj_batch_plan_t **out);
int aj_batch_plan_destroy(aj_batch_plan_t *plan);
/* Asynchronous on `stream` (NULL = the context's stream). */
int aj_batch_run(aj_batch_plan_t *plan, const aj_stream_tuple_t *dR, uint64_t nR,
                 const aj_stream_tuple_t *dS, uint64_t nS, const uint32_t *d_bounds, int nwin,
                 uint64_t *d_count, uint64_t *d_checksum, void *stream);

/*
 * Device stream generator for window workloads: Poisson arrivals at
 * rate_per_us tuples per microsecond (exponential gaps, prefix-summed), window
 * w lasting sizes[w] / rate microseconds, keys uniform on
 * [1, max(1, sizes[w] / dup)] of the tuple's window, payload = row id.
 * Writes up to `capacity` tuples to d_out, nwin+1 boundaries to d_bounds and
 * the tuple count to *n_out.  Blocks until done.
 */
int aj_gen_window_streams(aj_ctx_t *ctx, uint64_t seed, int nwin, const uint32_t *sizes,
                          double rate_per_us, uint32_t dup, aj_stream_tuple_t *d_out,
                          uint64_t capacity, uint32_t *d_bounds, uint64_t *n_out);

/*
 * Device generator of n window tuples (12-byte aj_stream_tuple_t, ts = 0):
 * keys uniform on [1, key_domain] (theta = 0) or Zipf(theta) ranks over
 * key_domain drawn by rejection-inversion and mapped through a seeded
 * permutation of 1..key_domain; payload = payload0 + row.  Value i depends
 * only on (seed, i), so any slice can be regenerated.  Runs on `stream`
 * (a hipStream_t); NULL = the context's stream, and then blocks until done.
 * No counterpart in the reference (its windows come from createRelation,
 * benchmark.cpp:38-126, on the host).
 */
int aj_gen_stream_tuples(aj_ctx_t *ctx, aj_stream_tuple_t *d_out, uint64_t n, uint64_t seed,
                         uint32_t key_domain, double theta, uint32_t payload0, void *stream);

/* ------------------------------------------------------------ serving loop */
/*
 * A request loop for window joins arriving from many streams (no counterpart
 * in the reference, which joins one window per process run).  Each window's
 * tuples live in device memory owned by the server ("window memory"); the
 * window memory plus the workspaces of the batches in flight are bounded by
 * an HBM budget.  Life of a request:
 *   aj_server_window_alloc   admission: reserves window memory for |R|, |S|
 *                            12-byte tuples; waits while the budget is taken
 *   fill / upload            the window's tuples arrive (device generator or
 *                            host copy, on the server's ingest stream)
 *   aj_server_submit         the window closed: queue it for joining
 *   (server)                 a dispatcher thread packs queued windows into a
 *                            batch whenever one of `slots` HIP streams is free
 *                            and launches the batched join on it, so batches
 *                            overlap on the device; a completer thread waits
 *                            for each batch, publishes per-window results and
 *                            returns their window memory
 *   aj_server_poll           results with close/start/done times
 * Results are exactly those of a plain join of each window alone.
 */
typedef struct aj_server aj_server_t;

typedef struct aj_server_opts_t {
    uint64_t hbm_budget;        /* device bytes for windows + batch workspaces, 0 = 200e9 */
    int slots;                  /* batches in flight (one HIP stream each), 0 = 3 */
    int max_batch_windows;      /* windows per batch, 0 = 256 */
    uint64_t max_batch_tuples;  /* tuples per side per batch, 0 = 2^27 */
    uint64_t max_window_tuples; /* tuples per side per window, 0 = 2^24 */
    int cap;                    /* LDS round as in aj_batch_opts_t, 0 = 8704 */
} aj_server_opts_t;

typedef struct aj_window_t {
    uint64_t id;
    aj_stream_tuple_t *dR; /* device: nR tuples inside the server's window memory */
    aj_stream_tuple_t *dS;
    uint64_t nR, nS;
} aj_window_t;

typedef struct aj_server_result_t {
    uint64_t id, tag;
    uint64_t count, checksum;
    int status;            /* AJ_OK, or the error of the batch that carried it */
    int batch_windows;     /* windows in that batch */
    double close_us;       /* window close (aj_server_submit), server clock */
    double start_us;       /* its batch was launched */
    double done_us;        /* its result was on the host */
} aj_server_result_t;

typedef struct aj_server_stats_t {
    uint64_t windows_done, batches, tuples_done;
    uint64_t window_bytes;      /* window memory (budget minus workspaces) */
    uint64_t window_bytes_used; /* currently reserved */
    uint64_t window_bytes_peak;
    uint64_t workspace_bytes;   /* all slots */
    uint64_t alloc_waits;       /* window_alloc calls that waited for memory */
    uint64_t pending;           /* submitted, not yet in a batch */
    uint64_t in_flight;         /* batches on the device */
} aj_server_stats_t;

int aj_server_create(aj_ctx_t *ctx, const aj_server_opts_t *opts, aj_server_t **out);
/* Waits for the batches in flight; windows still queued are dropped. */
int aj_server_destroy(aj_server_t *srv);
/* Reserves window memory; waits up to timeout_us (< 0: no limit, 0: no
 * wait) for other windows to finish.  AJ_ERR_CAPACITY on timeout or if the
 * window can never fit; AJ_ERR_ARG above max_window_tuples. */
int aj_server_window_alloc(aj_server_t *srv, uint64_t nR, uint64_t nS, int64_t timeout_us,
                           aj_window_t *win);
/* Cancels a window that was allocated and not submitted. */
int aj_server_window_release(aj_server_t *srv, const aj_window_t *win);
/* R keys uniform on [1, key_domain]; S keys uniform (theta_s = 0) or
 * Zipf(theta_s) over key_domain (aj_gen_stream_tuples with seeds
 * seed and seed ^ 0x5DEECE66D).  Asynchronous: fills are queued and launched
 * together on the ingest stream (up to 16 window sides per launch), at the
 * latest before the next batch, upload or release of window memory; batches
 * wait for them on the device.  key_domain 0 or theta_s outside [0, 10] is
 * AJ_ERR_ARG. */
int aj_server_window_fill(aj_server_t *srv, const aj_window_t *win, uint64_t seed,
                          uint32_t key_domain, double theta_s);
/* Copies host tuples into the window (blocking). */
int aj_server_window_upload(aj_server_t *srv, const aj_window_t *win,
                            const aj_stream_tuple_t *hostR, const aj_stream_tuple_t *hostS);
/* Queues the window for joining; close_us < 0 means now. */
int aj_server_submit(aj_server_t *srv, const aj_window_t *win, uint64_t tag, double close_us);
/* Up to max finished results into out[0..*n); waits up to timeout_us (< 0:
 * no limit) for the first one. */
int aj_server_poll(aj_server_t *srv, aj_server_result_t *out, int max, int64_t timeout_us,
                   int *n);
/* Waits until every submitted window has a result. */
int aj_server_drain(aj_server_t *srv);
double aj_server_now_us(aj_server_t *srv);
int aj_server_stats(aj_server_t *srv, aj_server_stats_t *out);

/* ------------------------------------------------------------ autotuner */
/*
 * PRO geometry per size bucket (floor log2 |R|, floor log2 |S|).  No
 * counterpart in the reference, which fixes NUM_RADIX_BITS / NUM_PASSES at
 * compile time (prj_params.h:16-21).  A PRO call whose options leave radix
 * bits, passes and cap all at 0 uses the bucket's entry when one exists and
 * the option "tuned" is 1 (the default).
 */
typedef struct aj_tune_entry_t {
    int log2_r, log2_s;   /* size bucket */
    int radix_bits;       /* total radix bits */
    int passes;           /* 1 or 2 (0 with 0 bits) */
    int cap;              /* LDS round */
    int part_block;       /* 256 or 512 */
    uint32_t part_chunk;  /* 0 = derived */
    float device_ms;      /* measured at tuning time */
    int one_pass_hist;    /* one pass of <= 11 bits: 1 = through the histogram pass,
                             0 = the histogram-free staged pass (option "staged") */
} aj_tune_entry_t;

/* Times candidate geometries (radix bits around the size heuristic, 1 or 2
 * passes, one pass with and without the staged pass, LDS rounds
 * 8192/8704/12288, then the partition work-group size) on
 * the given device relations, `reps` timed runs each after a warm-up; every
 * candidate must return the same count and checksum (else AJ_ERR_STATE).  The
 * fastest becomes the bucket's entry and is returned in *best. */
int aj_tune_pro(aj_ctx_t *ctx, const tuple_t *dR, uint64_t nR, const tuple_t *dS, uint64_t nS,
                int reps, aj_tune_entry_t *best);
int aj_tune_put(const aj_tune_entry_t *entry);
/* AJ_ERR_STATE when the bucket has no entry. */
int aj_tune_get(uint64_t nR, uint64_t nS, aj_tune_entry_t *out);
int aj_tune_clear(void);
/* JSON, one entry per line; load merges into the current table. */
int aj_tune_save(const char *path);
int aj_tune_load(const char *path);

/*
 * Autotuner of a window batch's geometry (no counterpart in the reference,
 * which joins one window per run): plans of LDS rounds {8192, 8704, 12288} x
 * partition targets {the round's power of two, half of it} x tuples per
 * partition work-group {derived, 65536, 262144} each join the caller's
 * resident batch (a warm-up, then reps timed launches with a sync each); every
 * candidate must return the same per-window counts and checksums
 * (AJ_ERR_STATE otherwise).  *best receives the fastest options (graph on),
 * *best_ms (nullable) its mean launch time.  Allocates a plan per candidate.
 */
int aj_tune_batch(aj_ctx_t *ctx, const aj_stream_tuple_t *dR, uint64_t nR,
                  const aj_stream_tuple_t *dS, uint64_t nS, const uint32_t *d_bounds, int nwin,
                  uint64_t max_window_tuples, int reps, aj_batch_opts_t *best, float *best_ms);

/* ------------------------------------------------------------ generators */
/*
 * Seeded host generators (counter-based splitmix64 stream: value i is output
 * number i of splitmix64 started at the seed, so the result does not depend on
 * the thread count), writing tuple_t arrays with payloadID = row id (0-based).
 *   unique:  keys are a seeded pseudo-random permutation of 1..n
 *   uniform: keys uniform on [lo, hi]
 *   zipf:    keys are ranks 1..nkeys drawn Zipf(theta) by rejection-inversion
 *            (Hoermann & Derflinger 1996), mapped through a seeded
 *            permutation of 1..nkeys so hot keys scatter
 */
int aj_gen_unique(tuple_t *out, uint64_t n, uint64_t seed);
int aj_gen_uniform(tuple_t *out, uint64_t n, uint32_t lo, uint32_t hi, uint64_t seed);
int aj_gen_zipf(tuple_t *out, uint64_t n, uint32_t nkeys, double theta, uint64_t seed);

#ifdef __cplusplus
}
#endif

#endif /* ALLIANCEDB_AJ_H */
