/*
 * oracle.c -- CPU oracle for the lazy equi-join path (TEST INFRASTRUCTURE).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this file's library, and only as the checker.  The product path never links
 * it.  See oracle.h for the result contract and the reference files each
 * function restates.
 *
 * Pinning (what checks each function against something it did not compute):
 *   oj_pro / oj_npo   counts: against the reference's own RJ_st / NPO_st built
 *                     from its sources (oracle/_ref), and all results against
 *                     oj_aggregate, brute force and the closed forms in tests/.
 *   oj_aggregate      against brute force on tiny inputs and the closed forms
 *                     (permutation x permutation, one shared key, disjoint keys).
 *   oj_pairs_*        against each other and against count/checksum.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

static inline uint32_t ukey(const tuple_t *t) { return (uint32_t)t->key; }
static inline uint64_t pair_of(const tuple_t *r, const tuple_t *s) {
    return ((uint64_t)(uint32_t)r->payloadID << 32) | (uint64_t)(uint32_t)s->payloadID;
}

/* ------------------------------------------------------------------ */
/* growable per-thread pair buffer                                     */
/* ------------------------------------------------------------------ */
typedef struct pairbuf_t {
    uint64_t *v;
    uint64_t n, cap;
    int oom;
} pairbuf_t;

static inline void pb_push(pairbuf_t *b, uint64_t x) {
    if (b->n == b->cap) {
        uint64_t nc = b->cap ? b->cap * 2 : 4096;
        uint64_t *nv = (uint64_t *)realloc(b->v, nc * sizeof(uint64_t));
        if (!nv) { b->oom = 1; return; }
        b->v = nv;
        b->cap = nc;
    }
    b->v[b->n++] = x;
}

static uint32_t next_pow2_u32(uint32_t v) {
    /* smallest power of two >= v (v >= 1); the NEXT_POW_2 macro of
     * parallel_radix_join.cpp:77-88 */
    v--;
    v |= v >> 1; v |= v >> 2; v |= v >> 4; v |= v >> 8; v |= v >> 16;
    return v + 1;
}

/* ------------------------------------------------------------------ */
/* PRO                                                                  */
/* ------------------------------------------------------------------ */

/* Bucket-chaining build/probe of one partition pair: the JoinFunction of PRO,
 * parallel_radix_join.cpp:291-354.  Chain heads and links are 1-based row
 * numbers; the hash is the key bits just above the radix bits. */
static void chain_join(const tuple_t *R, uint32_t nR, const tuple_t *S, uint32_t nS,
                       int radix_bits, uint64_t *cnt, uint64_t *sum, pairbuf_t *pb,
                       uint32_t *next, uint32_t *bucket) {
    uint32_t N = next_pow2_u32(nR);
    const uint32_t mask = (N - 1) << radix_bits;
    memset(bucket, 0, sizeof(uint32_t) * N);
    for (uint32_t i = 0; i < nR;) {
        uint32_t idx = (ukey(&R[i]) & mask) >> radix_bits;
        next[i] = bucket[idx];
        bucket[idx] = ++i;
    }
    uint64_t c = 0, x = 0;
    for (uint32_t i = 0; i < nS; i++) {
        uint32_t idx = (ukey(&S[i]) & mask) >> radix_bits;
        for (uint32_t hit = bucket[idx]; hit > 0; hit = next[hit - 1]) {
            if (S[i].key == R[hit - 1].key) {
                c++;
                x += pair_of(&R[hit - 1], &S[i]);
                if (pb) pb_push(pb, pair_of(&R[hit - 1], &S[i]));
            }
        }
    }
    *cnt += c;
    *sum += x;
}

/* Serial radix cluster of one relation piece by digit (key >> shift) & (F-1):
 * radix_cluster, parallel_radix_join.cpp:573-614 (without the L1 padding). */
static void cluster(tuple_t *out, const tuple_t *in, uint64_t n, int shift, int bits,
                    uint64_t *hist /* F+1 entries out: start offsets */) {
    const uint32_t F = 1u << bits;
    memset(hist, 0, sizeof(uint64_t) * (F + 1));
    for (uint64_t i = 0; i < n; i++) hist[((ukey(&in[i]) >> shift) & (F - 1)) + 1]++;
    for (uint32_t d = 0; d < F; d++) hist[d + 1] += hist[d];
    uint64_t *dst = (uint64_t *)malloc(sizeof(uint64_t) * F);
    memcpy(dst, hist, sizeof(uint64_t) * F);
    for (uint64_t i = 0; i < n; i++) out[dst[(ukey(&in[i]) >> shift) & (F - 1)]++] = in[i];
    free(dst);
}

typedef struct pro_task_t {
    uint64_t r0, nr, s0, ns; /* offsets into the pass-1 output */
} pro_task_t;

typedef struct pro_shared_t {
    const tuple_t *R, *S;
    uint64_t nR, nS;
    int T, bits, passes, bits1, bits2;
    tuple_t *tmpR, *tmpS;   /* pass-1 output */
    tuple_t *tmp2R, *tmp2S; /* pass-2 output */
    uint64_t *histR, *histS; /* T x F1 counts, then offsets */
    pthread_barrier_t bar;
    pro_task_t *ptasks;     /* pass-2 (or join, when passes==1) tasks */
    uint32_t nptasks;
    atomic_uint pnext;
    pro_task_t *jtasks;     /* join tasks */
    atomic_uint jcount;
    atomic_uint jnext;
    uint64_t max_sub;       /* scratch size for chain_join */
    double t_part_end;
    int want_pairs;
} pro_shared_t;

typedef struct pro_arg_t {
    pro_shared_t *sh;
    int tid;
    uint64_t cnt, sum;
    pairbuf_t pb;
} pro_arg_t;

/* Pass-1 parallel radix partition: parallel_radix_partition,
 * parallel_radix_join.cpp:739-806.  Output layout is partition-major,
 * thread-minor, exactly as the offsets at 783-791 give. */
static void pro_pass1(pro_shared_t *sh, int tid, const tuple_t *rel, uint64_t n,
                      tuple_t *tmp, uint64_t *hist) {
    const int T = sh->T;
    const uint32_t F = 1u << sh->bits1;
    uint64_t per = n / T;
    uint64_t b = per * tid, e = (tid == T - 1) ? n : b + per;
    uint64_t *my = hist + (uint64_t)tid * F;
    memset(my, 0, sizeof(uint64_t) * F);
    for (uint64_t i = b; i < e; i++) my[ukey(&rel[i]) & (F - 1)]++;
    pthread_barrier_wait(&sh->bar);
    /* dst[j] = sum_i sum_{k<j} hist[i][k] + sum_{i<tid} hist[i][j] */
    uint64_t *dst = (uint64_t *)calloc(F, sizeof(uint64_t));
    uint64_t before = 0;
    for (uint32_t j = 0; j < F; j++) {
        uint64_t o = before;
        for (int i = 0; i < T; i++) {
            uint64_t h = hist[(uint64_t)i * F + j];
            if (i < tid) o += h;
            before += h;
        }
        dst[j] = o;
    }
    for (uint64_t i = b; i < e; i++) tmp[dst[ukey(&rel[i]) & (F - 1)]++] = rel[i];
    free(dst);
}

static void *pro_thread(void *p) {
    pro_arg_t *a = (pro_arg_t *)p;
    pro_shared_t *sh = a->sh;
    const int tid = a->tid;
    const uint32_t F1 = 1u << sh->bits1;

    pro_pass1(sh, tid, sh->R, sh->nR, sh->tmpR, sh->histR);
    pro_pass1(sh, tid, sh->S, sh->nS, sh->tmpS, sh->histS);
    pthread_barrier_wait(&sh->bar);

    /* thread 0 turns non-empty pass-1 partition pairs into tasks
     * (parallel_radix_join.cpp:1116-1197) */
    if (tid == 0) {
        uint64_t ro = 0, so = 0;
        sh->nptasks = 0;
        for (uint32_t d = 0; d < F1; d++) {
            uint64_t nr = 0, ns = 0;
            for (int i = 0; i < sh->T; i++) {
                nr += sh->histR[(uint64_t)i * F1 + d];
                ns += sh->histS[(uint64_t)i * F1 + d];
            }
            if (nr > 0 && ns > 0) {
                pro_task_t t = {ro, nr, so, ns};
                sh->ptasks[sh->nptasks++] = t;
                if (nr > sh->max_sub) sh->max_sub = nr;
            }
            ro += nr;
            so += ns;
        }
    }
    pthread_barrier_wait(&sh->bar);

    if (sh->passes == 2) {
        /* pass 2: each task is radix-clustered on the next digit and every
         * non-empty sub-partition pair becomes a join task
         * (serial_radix_partition, parallel_radix_join.cpp:688-729) */
        const uint32_t F2 = 1u << sh->bits2;
        uint64_t *hr = (uint64_t *)malloc(sizeof(uint64_t) * (F2 + 1));
        uint64_t *hs = (uint64_t *)malloc(sizeof(uint64_t) * (F2 + 1));
        unsigned k;
        while ((k = atomic_fetch_add(&sh->pnext, 1)) < sh->nptasks) {
            pro_task_t t = sh->ptasks[k];
            cluster(sh->tmp2R + t.r0, sh->tmpR + t.r0, t.nr, sh->bits1, sh->bits2, hr);
            cluster(sh->tmp2S + t.s0, sh->tmpS + t.s0, t.ns, sh->bits1, sh->bits2, hs);
            for (uint32_t d = 0; d < F2; d++) {
                uint64_t nr = hr[d + 1] - hr[d], ns = hs[d + 1] - hs[d];
                if (nr > 0 && ns > 0) {
                    unsigned j = atomic_fetch_add(&sh->jcount, 1);
                    pro_task_t jt = {t.r0 + hr[d], nr, t.s0 + hs[d], ns};
                    sh->jtasks[j] = jt;
                }
            }
        }
        free(hr);
        free(hs);
    }
    pthread_barrier_wait(&sh->bar);
    if (tid == 0) sh->t_part_end = now_ms();

    /* join phase: drain the join queue (parallel_radix_join.cpp:1409-1416) */
    const tuple_t *baseR = sh->passes == 2 ? sh->tmp2R : sh->tmpR;
    const tuple_t *baseS = sh->passes == 2 ? sh->tmp2S : sh->tmpS;
    const pro_task_t *tasks = sh->passes == 2 ? sh->jtasks : sh->ptasks;
    const unsigned ntasks = sh->passes == 2 ? atomic_load(&sh->jcount) : sh->nptasks;
    uint32_t *next = NULL, *bucket = NULL;
    uint64_t scratch = 0;
    unsigned k;
    while ((k = atomic_fetch_add(&sh->jnext, 1)) < ntasks) {
        pro_task_t t = tasks[k];
        if (t.nr > scratch) {
            scratch = t.nr;
            free(next);
            free(bucket);
            next = (uint32_t *)malloc(sizeof(uint32_t) * t.nr);
            bucket = (uint32_t *)malloc(sizeof(uint32_t) * next_pow2_u32((uint32_t)t.nr));
        }
        chain_join(baseR + t.r0, (uint32_t)t.nr, baseS + t.s0, (uint32_t)t.ns, sh->bits,
                   &a->cnt, &a->sum, sh->want_pairs ? &a->pb : NULL, next, bucket);
    }
    free(next);
    free(bucket);
    return NULL;
}

int oj_pro(const tuple_t *R, uint64_t nR, const tuple_t *S, uint64_t nS, int nthreads,
           int radix_bits, int passes, oj_out_t *out, uint64_t *pairs, uint64_t pairs_cap) {
    if (!out || nthreads < 1 || radix_bits < 1 || radix_bits > 24 || passes < 1 || passes > 2)
        return -1;
    if ((nR && !R) || (nS && !S)) return -1;
    if (nR > (1ull << 31) || nS > (1ull << 31)) return -1;
    memset(out, 0, sizeof(*out));
    if (nR == 0 || nS == 0) return 0;
    double t0 = now_ms();
    pro_shared_t sh;
    memset(&sh, 0, sizeof(sh));
    sh.R = R; sh.S = S; sh.nR = nR; sh.nS = nS;
    sh.T = nthreads;
    sh.bits = radix_bits;
    sh.passes = passes;
    /* split of the radix bits over passes: prj_params.h:75-81 */
    sh.bits1 = passes == 2 ? radix_bits / 2 : radix_bits;
    sh.bits2 = radix_bits - sh.bits1;
    sh.want_pairs = pairs != NULL;
    const uint32_t F1 = 1u << sh.bits1;
    sh.tmpR = (tuple_t *)malloc(sizeof(tuple_t) * nR);
    sh.tmpS = (tuple_t *)malloc(sizeof(tuple_t) * nS);
    if (passes == 2) {
        sh.tmp2R = (tuple_t *)malloc(sizeof(tuple_t) * nR);
        sh.tmp2S = (tuple_t *)malloc(sizeof(tuple_t) * nS);
    }
    sh.histR = (uint64_t *)malloc(sizeof(uint64_t) * F1 * nthreads);
    sh.histS = (uint64_t *)malloc(sizeof(uint64_t) * F1 * nthreads);
    sh.ptasks = (pro_task_t *)malloc(sizeof(pro_task_t) * F1);
    sh.jtasks = (pro_task_t *)malloc(sizeof(pro_task_t) * ((uint64_t)1 << radix_bits));
    int rc = 0;
    if (!sh.tmpR || !sh.tmpS || (passes == 2 && (!sh.tmp2R || !sh.tmp2S)) || !sh.histR ||
        !sh.histS || !sh.ptasks || !sh.jtasks) {
        rc = -2;
        goto done;
    }
    pthread_barrier_init(&sh.bar, NULL, nthreads);
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
    pro_arg_t *args = (pro_arg_t *)calloc(nthreads, sizeof(pro_arg_t));
    for (int i = 0; i < nthreads; i++) {
        args[i].sh = &sh;
        args[i].tid = i;
        pthread_create(&th[i], NULL, pro_thread, &args[i]);
    }
    for (int i = 0; i < nthreads; i++) pthread_join(th[i], NULL);
    double t1 = now_ms();
    pthread_barrier_destroy(&sh.bar);
    uint64_t np = 0;
    for (int i = 0; i < nthreads; i++) {
        out->count += args[i].cnt;
        out->checksum += args[i].sum;
        if (args[i].pb.oom) rc = -2;
        if (pairs && rc == 0) {
            if (np + args[i].pb.n > pairs_cap) rc = -3;
            else if (args[i].pb.n) memcpy(pairs + np, args[i].pb.v, args[i].pb.n * sizeof(uint64_t));
            np += args[i].pb.n;
        }
        free(args[i].pb.v);
    }
    out->npairs = pairs ? np : 0;
    out->part_ms = sh.t_part_end - t0;
    out->join_ms = t1 - sh.t_part_end;
    free(th);
    free(args);
done:
    free(sh.tmpR); free(sh.tmpS); free(sh.tmp2R); free(sh.tmp2S);
    free(sh.histR); free(sh.histS); free(sh.ptasks); free(sh.jtasks);
    out->total_ms = now_ms() - t0;
    return rc;
}

/* ------------------------------------------------------------------ */
/* NPO                                                                  */
/* ------------------------------------------------------------------ */

#define NPO_BUCKET_SIZE 2      /* npj_params.h:19 */
#define NPO_OVERFLOW_BUF 1024  /* npj_params.h:29 */

typedef struct npo_bucket_t {  /* npj_types.h:32-38 */
    atomic_flag latch;
    uint32_t count;
    tuple_t tuples[NPO_BUCKET_SIZE];
    struct npo_bucket_t *next;
} npo_bucket_t;

typedef struct npo_pool_t {    /* bucket_buffer_t, npj_types.h:62-66 */
    struct npo_pool_t *next;
    uint32_t count;
    npo_bucket_t buf[NPO_OVERFLOW_BUF];
} npo_pool_t;

typedef struct npo_shared_t {
    const tuple_t *R, *S;
    uint64_t nR, nS;
    int T;
    npo_bucket_t *buckets;
    uint32_t mask;
    pthread_barrier_t bar;
    double t_build_end;
    int want_pairs;
} npo_shared_t;

typedef struct npo_arg_t {
    npo_shared_t *sh;
    int tid;
    uint64_t cnt, sum;
    pairbuf_t pb;
    int oom;
} npo_arg_t;

/* get_new_bucket, common_functions.cpp:334-348 */
static npo_bucket_t *npo_new_bucket(npo_pool_t **pool) {
    if ((*pool)->count < NPO_OVERFLOW_BUF) return &(*pool)->buf[(*pool)->count++];
    npo_pool_t *np = (npo_pool_t *)calloc(1, sizeof(npo_pool_t));
    if (!np) return NULL;
    np->count = 1;
    np->next = *pool;
    *pool = np;
    return &np->buf[0];
}

static void *npo_thread(void *p) {
    npo_arg_t *a = (npo_arg_t *)p;
    npo_shared_t *sh = a->sh;
    const int T = sh->T, tid = a->tid;
    npo_pool_t *pool = (npo_pool_t *)calloc(1, sizeof(npo_pool_t));
    if (!pool) a->oom = 1;
    pthread_barrier_wait(&sh->bar);

    /* build_hashtable_mt, common_functions.cpp:350-400: latch the head bucket,
     * fill it, else the first overflow bucket, else chain a new overflow
     * bucket right after the head. */
    uint64_t per = sh->nR / T;
    uint64_t b = per * tid, e = (tid == T - 1) ? sh->nR : b + per;
    for (uint64_t i = b; i < e && pool; i++) {
        npo_bucket_t *curr = sh->buckets + (ukey(&sh->R[i]) & sh->mask);
        while (atomic_flag_test_and_set_explicit(&curr->latch, memory_order_acquire)) {}
        npo_bucket_t *nxt = curr->next;
        tuple_t *dest;
        if (curr->count == NPO_BUCKET_SIZE) {
            if (!nxt || nxt->count == NPO_BUCKET_SIZE) {
                npo_bucket_t *nb = npo_new_bucket(&pool);
                if (!nb) { a->oom = 1; atomic_flag_clear_explicit(&curr->latch, memory_order_release); break; }
                curr->next = nb;
                nb->next = nxt;
                nb->count = 1;
                dest = nb->tuples;
            } else {
                dest = nxt->tuples + nxt->count;
                nxt->count++;
            }
        } else {
            dest = curr->tuples + curr->count;
            curr->count++;
        }
        *dest = sh->R[i];
        atomic_flag_clear_explicit(&curr->latch, memory_order_release);
    }
    pthread_barrier_wait(&sh->bar);
    if (tid == 0) sh->t_build_end = now_ms();

    /* probe_hashtable / probe_hashtable_single, common_functions.cpp:221-280 */
    per = sh->nS / T;
    b = per * tid;
    e = (tid == T - 1) ? sh->nS : b + per;
    uint64_t c = 0, x = 0;
    for (uint64_t i = b; i < e; i++) {
        const tuple_t *s = &sh->S[i];
        const npo_bucket_t *bk = sh->buckets + (ukey(s) & sh->mask);
        do {
            for (uint32_t j = 0; j < bk->count; j++) {
                if (s->key == bk->tuples[j].key) {
                    c++;
                    x += pair_of(&bk->tuples[j], s);
                    if (sh->want_pairs) pb_push(&a->pb, pair_of(&bk->tuples[j], s));
                }
            }
            bk = bk->next;
        } while (bk);
    }
    a->cnt = c;
    a->sum = x;
    /* The pools stay alive until every thread has probed: overflow buckets of
     * one thread are linked into chains that other threads walk. */
    pthread_barrier_wait(&sh->bar);
    while (pool) {
        npo_pool_t *n = pool->next;
        free(pool);
        pool = n;
    }
    return NULL;
}

int oj_npo(const tuple_t *R, uint64_t nR, const tuple_t *S, uint64_t nS, int nthreads,
           oj_out_t *out, uint64_t *pairs, uint64_t pairs_cap) {
    if (!out || nthreads < 1) return -1;
    if ((nR && !R) || (nS && !S)) return -1;
    if (nR > (1ull << 31) || nS > (1ull << 31)) return -1;
    memset(out, 0, sizeof(*out));
    if (nR == 0 || nS == 0) return 0;
    double t0 = now_ms();
    npo_shared_t sh;
    memset(&sh, 0, sizeof(sh));
    sh.R = R; sh.S = S; sh.nR = nR; sh.nS = nS;
    sh.T = nthreads;
    sh.want_pairs = pairs != NULL;
    /* nbuckets = next_pow2(|R| / BUCKET_SIZE), mask = nbuckets-1
     * (no_partitioning_join.cpp:406-408, common_functions.cpp:74-102);
     * at least one bucket so that |R| = 1 is well defined. */
    uint64_t want = nR / NPO_BUCKET_SIZE;
    uint32_t nb = want ? next_pow2_u32((uint32_t)want) : 1;
    sh.mask = nb - 1;
    sh.buckets = (npo_bucket_t *)calloc(nb, sizeof(npo_bucket_t));
    if (!sh.buckets) return -2;
    pthread_barrier_init(&sh.bar, NULL, nthreads);
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
    npo_arg_t *args = (npo_arg_t *)calloc(nthreads, sizeof(npo_arg_t));
    for (int i = 0; i < nthreads; i++) {
        args[i].sh = &sh;
        args[i].tid = i;
        pthread_create(&th[i], NULL, npo_thread, &args[i]);
    }
    for (int i = 0; i < nthreads; i++) pthread_join(th[i], NULL);
    double t1 = now_ms();
    pthread_barrier_destroy(&sh.bar);
    int rc = 0;
    uint64_t np = 0;
    for (int i = 0; i < nthreads; i++) {
        out->count += args[i].cnt;
        out->checksum += args[i].sum;
        if (args[i].oom || args[i].pb.oom) rc = -2;
        if (pairs && rc == 0) {
            if (np + args[i].pb.n > pairs_cap) rc = -3;
            else if (args[i].pb.n) memcpy(pairs + np, args[i].pb.v, args[i].pb.n * sizeof(uint64_t));
            np += args[i].pb.n;
        }
        free(args[i].pb.v);
    }
    out->npairs = pairs ? np : 0;
    out->part_ms = sh.t_build_end - t0;
    out->join_ms = t1 - sh.t_build_end;
    free(th);
    free(args);
    free(sh.buckets);
    out->total_ms = now_ms() - t0;
    return rc;
}

/* ------------------------------------------------------------------ */
/* sorting helpers and join-free oracles                               */
/* ------------------------------------------------------------------ */

int oj_sort_u64(uint64_t *a, uint64_t n) {
    if (n < 2) return 0;
    uint64_t *tmp = (uint64_t *)malloc(sizeof(uint64_t) * n);
    if (!tmp) return -2;
    uint64_t *src = a, *dst = tmp;
    static const int B = 16;
    uint64_t *cnt = (uint64_t *)malloc(sizeof(uint64_t) * (1u << B));
    for (int shift = 0; shift < 64; shift += B) {
        memset(cnt, 0, sizeof(uint64_t) * (1u << B));
        for (uint64_t i = 0; i < n; i++) cnt[(src[i] >> shift) & 0xFFFF]++;
        uint64_t s = 0;
        for (uint32_t d = 0; d < (1u << B); d++) {
            uint64_t c = cnt[d];
            cnt[d] = s;
            s += c;
        }
        for (uint64_t i = 0; i < n; i++) dst[cnt[(src[i] >> shift) & 0xFFFF]++] = src[i];
        uint64_t *t = src;
        src = dst;
        dst = t;
    }
    /* 4 passes: the sorted data is back in a */
    free(cnt);
    free(tmp);
    return 0;
}

/* tuples as u64 (key << 32 | payload), sorted: equal keys are contiguous */
static uint64_t *sorted_by_key(const tuple_t *T, uint64_t n) {
    uint64_t *v = (uint64_t *)malloc(sizeof(uint64_t) * (n ? n : 1));
    if (!v) return NULL;
    for (uint64_t i = 0; i < n; i++)
        v[i] = ((uint64_t)(uint32_t)T[i].key << 32) | (uint32_t)T[i].payloadID;
    oj_sort_u64(v, n);
    return v;
}

/* count = sum_k cR(k) cS(k);
 * checksum = sum_k cS(k) * sumR(k) * 2^32 + cR(k) * sumS(k)   (mod 2^64) */
int oj_aggregate(const tuple_t *R, uint64_t nR, const tuple_t *S, uint64_t nS,
                 uint64_t *count, uint64_t *checksum) {
    uint64_t *r = sorted_by_key(R, nR), *s = sorted_by_key(S, nS);
    if (!r || !s) { free(r); free(s); return -2; }
    uint64_t i = 0, j = 0, c = 0, x = 0;
    while (i < nR && j < nS) {
        uint32_t kr = (uint32_t)(r[i] >> 32), ks = (uint32_t)(s[j] >> 32);
        if (kr < ks) { i++; continue; }
        if (ks < kr) { j++; continue; }
        uint64_t cr = 0, sr = 0, cs = 0, ss = 0;
        while (i < nR && (uint32_t)(r[i] >> 32) == kr) { cr++; sr += (uint32_t)r[i]; i++; }
        while (j < nS && (uint32_t)(s[j] >> 32) == ks) { cs++; ss += (uint32_t)s[j]; j++; }
        c += cr * cs;
        x += ((cs * sr) << 32) + cr * ss;
    }
    free(r);
    free(s);
    *count = c;
    *checksum = x;
    return 0;
}

int64_t oj_pairs_sortmerge(const tuple_t *R, uint64_t nR, const tuple_t *S, uint64_t nS,
                           uint64_t *out, uint64_t cap) {
    uint64_t *r = sorted_by_key(R, nR), *s = sorted_by_key(S, nS);
    if (!r || !s) { free(r); free(s); return -2; }
    uint64_t i = 0, j = 0, n = 0;
    int64_t rc = 0;
    while (i < nR && j < nS) {
        uint32_t kr = (uint32_t)(r[i] >> 32), ks = (uint32_t)(s[j] >> 32);
        if (kr < ks) { i++; continue; }
        if (ks < kr) { j++; continue; }
        uint64_t i1 = i, j1 = j;
        while (i1 < nR && (uint32_t)(r[i1] >> 32) == kr) i1++;
        while (j1 < nS && (uint32_t)(s[j1] >> 32) == ks) j1++;
        for (uint64_t a = i; a < i1; a++)
            for (uint64_t b = j; b < j1; b++) {
                if (n >= cap) { rc = -1; goto done; }
                out[n++] = ((uint64_t)(uint32_t)r[a] << 32) | (uint32_t)s[b];
            }
        i = i1;
        j = j1;
    }
done:
    free(r);
    free(s);
    if (rc) return rc;
    oj_sort_u64(out, n);
    return (int64_t)n;
}

int64_t oj_pairs_bruteforce(const tuple_t *R, uint64_t nR, const tuple_t *S, uint64_t nS,
                            uint64_t *out, uint64_t cap) {
    uint64_t n = 0;
    for (uint64_t i = 0; i < nR; i++)
        for (uint64_t j = 0; j < nS; j++)
            if (R[i].key == S[j].key) {
                if (n >= cap) return -1;
                out[n++] = pair_of(&R[i], &S[j]);
            }
    oj_sort_u64(out, n);
    return (int64_t)n;
}
