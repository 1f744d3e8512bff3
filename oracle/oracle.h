/*
 * oracle.h -- CPU oracle for the lazy equi-join path (TEST INFRASTRUCTURE).
 *
 * This is a clean-room C restatement of the reference's CPU joins, used only
 * as the checker by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg.  Nothing in the shipped library links or calls it.
 *
 *   oj_pro        PRO: two-pass parallel radix partitioning + bucket-chaining
 *                 join (hashing/joins/parallel_radix_join.cpp:291-354,
 *                 573-614, 688-806, 981-1473, 1488-1706).
 *   oj_npo        NPO: one shared chained hash table with per-bucket latches
 *                 (hashing/joins/no_partitioning_join.cpp:156-476,
 *                 hashing/joins/common_functions.cpp:74-102, 221-280, 334-400).
 *   oj_aggregate  closed-form count/checksum by grouping keys (no join at all).
 *   oj_pairs_*    sorted (R.payload<<32 | S.payload) pair lists: sort-merge and
 *                 brute force.
 *
 * Result contract (shared with the GPU path):
 *   count    = number of pairs (r, s) with r.key == s.key
 *   checksum = sum over those pairs of ((u64)r.payloadID << 32 | (u32)s.payloadID),
 *              wrapping mod 2^64.  The reference computes only the count; the
 *              checksum is accumulated at the same match site
 *              (parallel_radix_join.cpp:331, common_functions.cpp:250).
 */
#ifndef AJ_ORACLE_H
#define AJ_ORACLE_H

#include <stdint.h>
#include "../include/alliancedb/types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oj_out_t {
    uint64_t count;
    uint64_t checksum;
    uint64_t npairs;   /* pairs written (== count when a buffer was given) */
    double part_ms;    /* partition phase (PRO) / build phase (NPO) */
    double join_ms;    /* build+probe phase (PRO) / probe phase (NPO) */
    double total_ms;   /* whole call, temporaries included */
} oj_out_t;

/* Returns 0 on success, <0 on error (bad arguments, out of memory, pair
 * buffer too small).  pairs may be NULL. Inputs are never modified. */
int oj_pro(const tuple_t *R, uint64_t nR, const tuple_t *S, uint64_t nS,
           int nthreads, int radix_bits, int passes, oj_out_t *out,
           uint64_t *pairs, uint64_t pairs_cap);

int oj_npo(const tuple_t *R, uint64_t nR, const tuple_t *S, uint64_t nS,
           int nthreads, oj_out_t *out, uint64_t *pairs, uint64_t pairs_cap);

int oj_aggregate(const tuple_t *R, uint64_t nR, const tuple_t *S, uint64_t nS,
                 uint64_t *count, uint64_t *checksum);

/* Both return the number of pairs (sorted ascending) or -1 if cap is too small. */
int64_t oj_pairs_sortmerge(const tuple_t *R, uint64_t nR, const tuple_t *S,
                           uint64_t nS, uint64_t *out, uint64_t cap);
int64_t oj_pairs_bruteforce(const tuple_t *R, uint64_t nR, const tuple_t *S,
                            uint64_t nS, uint64_t *out, uint64_t cap);

/* LSD radix sort of u64 values, ascending (used to sort pair lists). */
int oj_sort_u64(uint64_t *a, uint64_t n);

#ifdef __cplusplus
}
#endif

#endif /* AJ_ORACLE_H */
