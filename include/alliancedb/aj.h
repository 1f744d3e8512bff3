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
 * Returns a malloc'ed result the caller frees with free(), or NULL on error
 * (see aj_last_error).  Extended results: aj_result_ext().
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
} aj_result_t;

/* Returns the extension of a result from PRO/NPO, or NULL if r is NULL. */
const aj_result_t *aj_result_ext(const result_t *r);

/*
 * Process-wide options of the drop-in entry points (the reference sets the
 * equivalents at compile time: JOIN_RESULT_MATERIALIZE, NUM_RADIX_BITS,
 * NUM_PASSES; common_functions.h:18, prj_params.h:16-21).
 *   "materialize"  0/1   also return the pair list (default 0)
 *   "radix_bits"   0..24 total radix bits (0 = tuned default)
 *   "passes"       0..2  partition passes (0 = tuned default)
 *   "device"       >=0   HIP device of the default context
 */
int aj_set_option(const char *name, int64_t value);
int64_t aj_get_option(const char *name);

/* ------------------------------------------------------------ device API */

/* Device memory helpers for FFI callers that have no HIP binding of their own
 * (no counterpart in the reference, whose relations live in host memory).
 * They act on the default context's device. */
int aj_device_alloc(void **ptr, uint64_t bytes);
int aj_device_free(void *ptr);
int aj_copy_h2d(void *dst_device, const void *src_host, uint64_t bytes);
int aj_copy_d2h(void *dst_host, const void *src_device, uint64_t bytes);
int aj_device_sync(void);

typedef struct aj_ctx aj_ctx_t;

int aj_ctx_create(int device, aj_ctx_t **out);
int aj_ctx_destroy(aj_ctx_t *ctx);

enum aj_algo { AJ_ALGO_PRO = 0, AJ_ALGO_NPO = 1 };

typedef struct aj_join_opts_t {
    int algo;          /* AJ_ALGO_PRO or AJ_ALGO_NPO */
    int radix_bits;    /* PRO: total radix bits, 0 = tuned */
    int passes;        /* PRO: 1 or 2, 0 = tuned */
    int cap;           /* PRO: R tuples per LDS build round, 0 = default */
    uint32_t split;    /* PRO: S tuples per join task, 0 = default */
    uint64_t *pairs;   /* device buffer for pairs, NULL = count/checksum only */
    uint64_t pairs_capacity;
} aj_join_opts_t;

typedef struct aj_join_out_t {
    uint64_t count;
    uint64_t checksum;
    uint64_t npairs;
    float device_ms;
} aj_join_out_t;

/*
 * Join device-resident relations (8-byte tuple_t arrays in HBM) on `stream`
 * (a hipStream_t, NULL = the context's stream).  Blocks until the result is
 * on the host.  If opts->pairs is set and too small, returns AJ_ERR_CAPACITY
 * with out->count holding the required size.
 */
int aj_join_dev(aj_ctx_t *ctx, const tuple_t *dR, uint64_t nR, const tuple_t *dS,
                uint64_t nS, const aj_join_opts_t *opts, aj_join_out_t *out,
                void *stream);

/* ------------------------------------------------------------ generators */
/*
 * Seeded host generators (splitmix64-seeded xoshiro256**), writing tuple_t
 * arrays with payloadID = row id (0-based).
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
