// ref_driver.cpp -- calls the reference's own single-threaded joins
// (TEST INFRASTRUCTURE, built into oracle/_ref/ by oracle/Makefile from the
// reference sources where they lie; nothing here is copied from them).
//
//   RJ_st   hashing/joins/parallel_radix_join.cpp:1719-1844: two-pass radix
//           clustering + bucket_chaining_join (the same JoinFunction PRO uses,
//           parallel_radix_join.cpp:291-354).
//   NPO_st  hashing/joins/no_partitioning_join.cpp:112-147: the same
//           allocate_hashtable / probe_hashtable code NPO uses
//           (common_functions.cpp:74-102, 221-280).
//
// The multithreaded PRO/NPO entry points cannot run outside the authors'
// machine (thread 0 programs Intel PCM MSRs and writes to /data1/xtra without
// a NULL check, parallel_radix_join.cpp:1036-1047), so these two single-thread
// variants are what pins the oracle's counts.  The objects are compiled with
// JOIN_RESULT_MATERIALIZE, so every match also yields the reference's own
// materialised <S.key, S.payloadID> record (parallel_radix_join.cpp:333-338,
// common_functions.cpp:252-258), which pins the S half of the pair list.
#include "benchmark.h"
#include "joins/no_partitioning_join.h"
#include "joins/parallel_radix_join.h"
#include "utils/tuple_buffer.h"
#include "timer/t_timer.h"
#include "helper/localjoiner.h"

#include <cstdint>
#include <cstdlib>
#include <cstring>

extern "C" {

// algo: 0 = RJ_st, 1 = NPO_st.  R/S are arrays of 8-byte tuples; they are
// copied first because RJ_st's second pass writes into its inputs.
// s_list (nullable) receives (S.key << 32 | S.payloadID) of every match.
// Returns 0, or -1 on bad arguments, -3 if s_list is too small.
int ref_join(int algo, const void *R, uint64_t nR, const void *S, uint64_t nS,
             int64_t *count, uint64_t *s_list, uint64_t cap, uint64_t *nlist) {
    if (algo < 0 || algo > 1 || !count) return -1;
    // NPO_st sizes its table as NEXT_POW_2(|R| / BUCKET_SIZE): for |R| < 2 that
    // is zero buckets and the reference dereferences outside the table
    // (no_partitioning_join.cpp:118, common_functions.cpp:80-101).
    if (algo == 1 && nR < 2) return -5;
    const size_t pad = 64 * 1024;
    relation_t r, s;
    r.tuples = (tuple_t *)malloc(nR * sizeof(tuple_t) + pad);
    s.tuples = (tuple_t *)malloc(nS * sizeof(tuple_t) + pad);
    if (nR) memcpy(r.tuples, R, nR * sizeof(tuple_t));
    if (nS) memcpy(s.tuples, S, nS * sizeof(tuple_t));
    r.num_tuples = nR;
    s.num_tuples = nS;
    r.payload = s.payload = nullptr;
    param_t p;
    memset(&p, 0, sizeof(p));
    result_t *res = algo == 0 ? RJ_st(&r, &s, p) : NPO_st(&r, &s, p);
    *count = res->totalresults;
    int rc = 0;
    uint64_t n = 0;
    chainedtuplebuffer_t *cb = (chainedtuplebuffer_t *)res->resultlist[0].results;
    // newest buffer first, holding writepos records; older buffers are full
    uint32_t len = cb->writepos;
    for (tuplebuffer_t *b = cb->buf; b; b = b->next, len = CHAINEDBUFF_NUMTUPLESPERBUF) {
        for (uint32_t i = 0; i < len; i++) {
            if (s_list) {
                if (n >= cap) { rc = -3; break; }
                s_list[n] = ((uint64_t)(uint32_t)b->tuples[i].key << 32) |
                            (uint32_t)b->tuples[i].payloadID;
            }
            n++;
        }
    }
    if (nlist) *nlist = n;
    chainedtuplebuffer_free(cb);
    free(res->resultlist);
    free(res);
    free(r.tuples);
    free(s.tuples);
    return rc;
}

// The reference's own load_relation (hashing/utils/generator.cpp:645-666) on a
// text file: the n tuples and timestamps it produces are copied out.  Only for
// tsKey != 0 (with tsKey == 0 the reference leaves the timestamps unset) and
// files that hold a ',' or '|' and at least n well-formed lines.
int ref_load_relation(const char *path, int keyby, int tsKey, uint64_t n, uint32_t partitions,
                      uint64_t *tuples_out, uint64_t *ts_out) {
    if (!path || tsKey == 0 || partitions == 0 || !tuples_out || !ts_out) return -1;
    relation_t rel;
    relation_payload_t pl;
    memset(&rel, 0, sizeof(rel));
    pl.ts = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
    pl.num_tuples = n;
    if (load_relation(&rel, &pl, keyby, tsKey, const_cast<char *>(path), n, partitions)) {
        free(pl.ts);
        return -2;
    }
    if (n) {
        memcpy(tuples_out, rel.tuples, n * sizeof(tuple_t));
        memcpy(ts_out, pl.ts, n * sizeof(uint64_t));
    }
    free(rel.tuples);
    free(pl.ts);
    return 0;
}

// The reference's own symmetric hash join of one thread, shj
// (helper/localjoiner.cpp:33-101): R and S tuples alternate in index order,
// each inserted into its side's table and probed against the other's.  The
// per-match records <S.key, S.payloadID> are not used here (the probing side
// varies), only the count.  Inputs are copied (the tables point into them).
int ref_shj(const void *R, uint64_t nR, const void *S, uint64_t nS, int64_t *count) {
    if (!count || nR < 2 || nS < 2) return -1;  // the tables need >= 1 bucket
    relation_t r, s;
    r.tuples = (tuple_t *)malloc(nR * sizeof(tuple_t));
    s.tuples = (tuple_t *)malloc(nS * sizeof(tuple_t));
    memcpy(r.tuples, R, nR * sizeof(tuple_t));
    memcpy(s.tuples, S, nS * sizeof(tuple_t));
    r.num_tuples = nR;
    s.num_tuples = nS;
    r.payload = s.payload = nullptr;
    chainedtuplebuffer_t *cb = chainedtuplebuffer_init();
    {
        SHJJoiner j = shj(0, &r, &s, cb);
        *count = j.matches;
    }
    chainedtuplebuffer_free(cb);
    free(r.tuples);
    free(s.tuples);
    return 0;
}

// The reference's own shj() with its records: every match appends the probing
// (later-arriving) tuple's <key, payloadID> to the chained buffer
// (common_functions.cpp:251-257, built with JOIN_RESULT_MATERIALIZE); list
// (nullable) receives them as (key << 32 | payloadID).  Returns 0, -1 on bad
// arguments, -3 if list is too small.
int ref_shj_records(const void *R, uint64_t nR, const void *S, uint64_t nS, int64_t *count,
                    uint64_t *list, uint64_t cap, uint64_t *nlist) {
    if (!count || nR < 2 || nS < 2) return -1;  // the tables need >= 1 bucket
    relation_t r, s;
    r.tuples = (tuple_t *)malloc(nR * sizeof(tuple_t));
    s.tuples = (tuple_t *)malloc(nS * sizeof(tuple_t));
    memcpy(r.tuples, R, nR * sizeof(tuple_t));
    memcpy(s.tuples, S, nS * sizeof(tuple_t));
    r.num_tuples = nR;
    s.num_tuples = nS;
    r.payload = s.payload = nullptr;
    chainedtuplebuffer_t *cb = chainedtuplebuffer_init();
    {
        SHJJoiner j = shj(0, &r, &s, cb);
        *count = j.matches;
    }
    int rc = 0;
    uint64_t n = 0;
    uint32_t len = cb->writepos;  // newest buffer first; older buffers are full
    for (tuplebuffer_t *b = cb->buf; b; b = b->next, len = CHAINEDBUFF_NUMTUPLESPERBUF) {
        for (uint32_t i = 0; i < len; i++) {
            if (list) {
                if (n >= cap) { rc = -3; break; }
                list[n] = ((uint64_t)(uint32_t)b->tuples[i].key << 32) |
                          (uint32_t)b->tuples[i].payloadID;
            }
            n++;
        }
    }
    if (nlist) *nlist = n;
    chainedtuplebuffer_free(cb);
    free(r.tuples);
    free(s.tuples);
    return rc;
}

// The reference's own breakdown_global of the lazy joins (timer/t_timer.cpp:
// 113-140) for the given phase cycles, written to path.  (merge / dump_timing,
// the progressive records, compile to nothing here: the reference's
// common_functions.h:21 defines NO_TIMING and its scripts flip that line in
// place, benchmark.sh:88.)
int ref_breakdown_file(const char *path, uint64_t nR, uint64_t nS, long lastTS,
                       uint64_t overall, uint64_t part, uint64_t build, uint64_t join) {
    if (!path) return -1;
    T_TIMER t;
    t.overall_timer = overall;
    t.partition_timer = part;
    t.buildtimer = build;
    t.join_timer = join;
    FILE *fp = fopen(path, "w");
    if (!fp) return -2;
    breakdown_global((int64_t)(nR + nS), 1, &t, lastTS, fp);
    fclose(fp);
    return 0;
}

// The reference's own write_result_relation (hashing/utils/tuple_buffer.h:155-236,
// unsorted variant) on a result_t whose resultlist entries point at chained
// tuple buffers -- the Out.tbl writer of benchmark.cpp:262-265.
int ref_write_result(result_t *res, const char *path) {
    if (!res || !path) return -1;
    write_result_relation(res, const_cast<char *>(path));
    return 0;
}

}  // extern "C"
