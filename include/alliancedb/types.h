/*
 * alliancedb/types.h -- ABI types of the lazy intra-window join path.
 *
 * These are byte-identical to the reference's relation/result types so that a
 * caller built against the reference links against this library unchanged:
 *
 *   tuple_t            hashing/utils/types.h:67-71   8 B: {int32 payloadID @0, int32 key @4}
 *   relation_t         hashing/utils/types.h:77-81
 *   relation_payload_t hashing/utils/types.h:84-88
 *   threadresult_t     hashing/utils/types.h:91-95
 *   result_t           hashing/utils/types.h:98-102
 *   tuplebuffer_t, chainedtuplebuffer_t   hashing/utils/tuple_buffer.h:31-47
 *
 * Read as a little-endian uint64, a tuple is (key << 32) | payloadID.
 * Keys and payloads are compared and combined as 32-bit bit patterns; the
 * signed declaration only matters to callers, never to join results.
 *
 * Next to the reference's own headers: the shared types (intkey_t through
 * result_t) are declared only when the reference's hashing/utils/types.h has
 * not been seen (its include guard is TYPES_H, types.h:11-12), so a
 * translation unit that includes the reference's headers first gets the
 * reference's declarations.  When this header comes first in C++ code built
 * with the reference's hashing/ directory on the include path (detected by
 * its utils/types.h and joins/parallel_radix_join.h), it includes the
 * reference's utils/types.h itself, so a later include of it is a no-op.
 * Define AJ_NO_REFERENCE_TYPES to turn that detection off.
 */
#ifndef ALLIANCEDB_TYPES_H
#define ALLIANCEDB_TYPES_H

#include <stdint.h>

#if defined(__cplusplus) && !defined(TYPES_H) && !defined(AJ_NO_REFERENCE_TYPES) && \
    defined(__has_include)
#if __has_include("utils/types.h") && __has_include("joins/parallel_radix_join.h")
#include "utils/types.h"
#endif
#endif

#ifdef __cplusplus
extern "C" {
#endif

#ifndef TYPES_H
typedef int32_t intkey_t;
typedef int32_t value_t;

typedef struct tuple_t {
    value_t payloadID; /* row id of the tuple's payload */
    intkey_t key;      /* join key */
} tuple_t;

typedef struct relation_payload_t {
    uint64_t *ts;        /* per-tuple timestamp, indexed by payloadID */
    uint64_t num_tuples;
} relation_payload_t;

typedef struct relation_t {
    tuple_t *tuples;
    uint64_t num_tuples;
    relation_payload_t *payload;
} relation_t;

typedef struct threadresult_t {
    int64_t nresults;
    void *results;
    uint32_t threadid;
} threadresult_t;

typedef struct result_t {
    int64_t totalresults;       /* number of matching (r, s) pairs */
    threadresult_t *resultlist; /* NULL unless option "materialize" is 2 (see aj.h) */
    int nthreads;
} result_t;
#endif /* TYPES_H */

/*
 * Chained result buffers, layout-identical to the reference's
 * (hashing/utils/tuple_buffer.h:31-47): threadresult_t.results points at an
 * aj_chainedtuplebuffer_t when a result carries record lists.  Buffers hold
 * AJ_CHAINEDBUFF_NUMTUPLESPERBUF tuples (CHAINEDBUFF_NUMTUPLESPERBUF,
 * tuple_buffer.h:23); buf is the newest buffer and holds writepos tuples, each
 * older buffer (next) is full.
 */
#define AJ_CHAINEDBUFF_NUMTUPLESPERBUF (1024 * 1024)

typedef struct aj_tuplebuffer_t {
    tuple_t *tuples;
    struct aj_tuplebuffer_t *next;
} aj_tuplebuffer_t;

typedef struct aj_chainedtuplebuffer_t {
    aj_tuplebuffer_t *buf;
    aj_tuplebuffer_t *readcursor;
    aj_tuplebuffer_t *writecursor;
    uint32_t writepos;
    uint32_t readpos;
    uint32_t readlen;
    uint32_t numbufs;
} aj_chainedtuplebuffer_t;

/*
 * 12-byte stream tuple used by the window paths (batched windows, serving).
 * The reference keeps the timestamp in a side array (relation_payload_t.ts);
 * the window paths keep it inline so one coalesced read gives the window id.
 */
typedef struct aj_stream_tuple_t {
    uint32_t key;
    uint32_t payload;
    uint32_t ts; /* arrival time in microseconds */
} aj_stream_tuple_t;

#ifdef __cplusplus
}
#endif

#endif /* ALLIANCEDB_TYPES_H */
