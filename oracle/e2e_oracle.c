/*
 * End-to-end single-end harness of the CPU oracle -- TEST INFRASTRUCTURE.
 *
 * Only tests/ load this code, as the checker: it restates swiftbwa's single-
 * end `mem` path (file:line citations at each function) so that the GPU
 * library's extension and CIGAR results can be taken through the same
 * seeding, chaining, dedup, primary-marking and SAM steps as the CPU path.
 *
 * Parity pins (tests/test_oracle_e2e.py):
 *   index + SMEMs    brute force on the text: every SMEM's SA interval equals
 *                    the set of exact-match positions of its substring in
 *                    fwd + revcomp, it cannot be extended either way, and
 *                    every maximal match >= min_seed_len is reported; SA is
 *                    sorted; occ / count identities.
 *   whole pipeline   the reference's lambda fixtures
 *                    (TestData/e2e_expected_se.sam via IntegrationTests.swift:
 *                    319-358: POS, FLAG, CIGAR, MAPQ within 5, NM, MD).
 *   paired end       the reference's lambda PE fixture
 *                    (TestData/e2e_expected_pe.sam via IntegrationTests.swift:
 *                    430-479); the insert-size estimate against the simulated
 *                    fragment distribution; rescue and proper pairing on a
 *                    simulated set (tests/test_oracle_e2e.py).
 *   chainer, filter, dedup, markSecondary, MAPQ, pair resolution, MAPQ
 *                    adjustment, SAM fields: pinned only through the lambda
 *                    fixtures (parity partially unpinned: the fixtures
 *                    exercise unique loci, one chain per read, and four pairs,
 *                    too few for an insert-size estimate, so no rescue).
 */
#define _GNU_SOURCE
#include "e2e_oracle.h"

#include <pthread.h>
#include <unistd.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------------- */
/* options (ScoringParameters.swift:3-42 defaults)                         */
/* ---------------------------------------------------------------------- */
enum {
    MAX_OCC = 500, MIN_OUT_SCORE = 30, MAX_CHAIN_GAP = 10000, SPLIT_WIDTH = 10,
    MAX_XA_HITS = 5, MIN_CHAIN_WEIGHT = 0
};
static const float SEED_SPLIT_RATIO = 1.5f, MASK_LEVEL = 0.50f, MASK_LEVEL_REDUN = 0.95f,
                   CHAIN_DROP_RATIO = 0.50f;

/* ---------------------------------------------------------------------- */
/* stable merge sort (Swift's sort keeps equal elements in order)          */
/* ---------------------------------------------------------------------- */
typedef int (*before_fn)(const void *a, const void *b, void *ctx);

static void msort(char *a, char *tmp, size_t n, size_t sz, before_fn before, void *ctx)
{
    if (n < 2) return;
    size_t h = n / 2;
    msort(a, tmp, h, sz, before, ctx);
    msort(a + h * sz, tmp, n - h, sz, before, ctx);
    size_t i = 0, j = h, o = 0;
    while (i < h && j < n) {
        if (before(a + j * sz, a + i * sz, ctx)) memcpy(tmp + (o++) * sz, a + (j++) * sz, sz);
        else memcpy(tmp + (o++) * sz, a + (i++) * sz, sz);
    }
    while (i < h) memcpy(tmp + (o++) * sz, a + (i++) * sz, sz);
    while (j < n) memcpy(tmp + (o++) * sz, a + (j++) * sz, sz);
    memcpy(a, tmp, n * sz);
}

static void stable_sort(void *base, size_t n, size_t sz, before_fn before, void *ctx)
{
    if (n < 2) return;
    char *tmp = (char *)malloc(n * sz);
    msort((char *)base, tmp, n, sz, before, ctx);
    free(tmp);
}

/* ---------------------------------------------------------------------- */
/* types                                                                   */
/* ---------------------------------------------------------------------- */
typedef struct { int64_t k, l; int32_t qb, qe; } smem_t;          /* SMEM.swift:5-25 */
typedef struct { int64_t k, l, s; int32_t m, n; } bidi_t;          /* SMEMFinder.swift:62-68 */
typedef struct { int64_t rbeg; int32_t qbeg, len, score; } seed_t; /* Seed.swift */

typedef struct {                                                   /* Chain.swift:4-22 */
    seed_t *seeds;
    int n, cap;
    int32_t weight, rid, kept, first;
    int64_t pos;
    float frac_rep;
} chain_t;

typedef struct {                                                   /* AlignmentRegion.swift */
    int64_t rb, re;
    int32_t qb, qe, rid, score, true_score, sub, csub, sub_n, w, seed_cov, secondary;
    uint64_t hash;
    float frac_rep;
    int64_t cig;                                                   /* CIGAR job, -1: dummy */
} region_t;

typedef struct { smem_t *v; int n, cap; } smem_vec;
typedef struct { chain_t *v; int n, cap; } chain_vec;
typedef struct { region_t *v; int n, cap; } region_vec;

typedef struct {                                                   /* one CIGAR job */
    int read;
    int64_t q_off, r_off;
    int32_t q_len, r_len, qb, qe, read_len, true_score, w;
    uint8_t rev;
    int64_t ref_pos;
    /* result */
    uint32_t *cigar;
    int n_cigar;
    int32_t nm;
    char *md;
    int64_t pos;
    int done;
} cjob_t;

typedef struct {                                   /* OrientationStats (InsertSizeEstimator) */
    int64_t low, high, proper_low, proper_high;
    double mean, stddev;
    int count, failed;
} ori_stats_t;

typedef struct {                                   /* MateRescue.RescueCandidate */
    int pair, mate_read, is_rev, rid, mate_len;
    int64_t clamped;
    int64_t raw_rb, raw_re;                        /* the window before fetchSequence */
    int64_t q_off, t_off;
    int32_t q_len, t_len;
    int32_t out6[6];
    int ok;
} rcand_t;

typedef struct {                                                   /* one extension call */
    int64_t q_off, t_off;
    int32_t q_len, t_len, h0, w, eb;
    or_swres_t r;
} call_t;

struct or_e2e {
    or_scoring_t sc;
    /* index */
    int64_t gl, n;                  /* n = 2 gl (text without '$') */
    uint8_t *fwd, *text, *bwt;
    int64_t *sa;                    /* n + 1 rows, sa[0] = n */
    int64_t cnt[5];
    int64_t sentinel;
    int32_t *occ;                   /* (n + 2) x 4 prefix counts */
    char name[256];
    /* reference sequences (ReferenceMetadata.annotations): offsets into the
       forward text, lengths, names; one sequence unless or_e2e_set_contigs */
    int n_ctg;
    int64_t *ctg_off, *ctg_len;
    char (*ctg_name)[256];
    /* batch */
    int n_reads;
    uint8_t *reads;
    int64_t *read_off;
    int32_t *read_len;
    chain_vec chains;
    int32_t *chain_read;            /* per chain */
    region_vec *regs;               /* per read */
    cjob_t *jobs;
    int64_t n_jobs, jobs_cap;
    uint8_t *cq, *cr;               /* CIGAR segment arenas */
    int64_t cq_n, cq_cap, cr_n, cr_cap;
    call_t *calls;
    int64_t n_calls, calls_cap;
    uint8_t *lq, *lt;
    int64_t lq_n, lq_cap, lt_n, lt_cap;
    /* paired end: reads [0, n_pairs) are read 1, [n_pairs, 2 n_pairs) read 2 */
    int n_pairs, do_rescue, primary_ori;
    ori_stats_t dist[4];
    rcand_t *resc;
    int64_t n_resc, resc_cap;
    uint8_t *rq, *rt;
    int64_t rq_n, rq_cap, rt_n, rt_cap;
    int raw_resc;                   /* rescue windows left to the caller's fetchSequence */
    int shared;                     /* index borrowed from another handle (or_e2e_fork) */
};

#define VEC_PUSH(vec, x)                                                              \
    do {                                                                              \
        if ((vec).n == (vec).cap) {                                                   \
            (vec).cap = (vec).cap ? 2 * (vec).cap : 16;                               \
            (vec).v = realloc((vec).v, (size_t)(vec).cap * sizeof(*(vec).v));         \
        }                                                                             \
        (vec).v[(vec).n++] = (x);                                                     \
    } while (0)

static void *grow(void *p, int64_t *cap, int64_t need, size_t sz)
{
    if (need <= *cap) return p;
    int64_t c = *cap ? *cap : 1024;
    while (c < need) c *= 2;
    *cap = c;
    return realloc(p, (size_t)c * sz);
}

/* ---------------------------------------------------------------------- */
/* index: IndexBuilder.swift:20-110, FMIndexWriter.swift:18-60             */
/* ---------------------------------------------------------------------- */
static int suffix_cmp(const void *a, const void *b, void *ctx)
{
    const or_e2e_t *e = (const or_e2e_t *)ctx;
    int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    const uint8_t *t = e->text;
    while (x < e->n && y < e->n && t[x] == t[y]) { ++x; ++y; }
    if (x == e->n) return y == e->n ? 0 : -1;                  /* '$' sorts first */
    if (y == e->n) return 1;
    return (int)t[x] - (int)t[y];
}

/* Suffix order of text[0, n) (then '$'): rows 1..n of sa.  Suffixes go to
 * 65,536 buckets by their first 8 bases (one counting pass and one scatter,
 * each over text slices in parallel); each bucket is sorted by its suffixes'
 * first 32 bases packed in 64 bits, and runs of equal keys by the full
 * comparison.  A suffix shorter than 32 bases is padded with A (code 0): its
 * key is then <= that of every suffix it prefixes, and equal keys are
 * resolved by suffix_cmp, which sorts '$' first -- the order is qsort's. */
typedef struct { uint64_t key; int64_t pos; } skey_t;

static int skey_cmp(const void *a, const void *b)
{
    const skey_t *x = (const skey_t *)a, *y = (const skey_t *)b;
    return x->key < y->key ? -1 : x->key > y->key;
}

typedef struct {
    const or_e2e_t *e;
    int64_t *sa;                     /* rows 1..n */
    int64_t lo, hi;                  /* text slice (count, scatter) */
    int64_t *cnt;                    /* 65,536 per slice */
    int64_t *bstart;                 /* 65,537 bucket starts (rows) */
    int64_t *next_bucket;            /* shared work counter (sort) */
    pthread_mutex_t *mu;
    int phase;
} sa_work_t;

static uint64_t key_at(const uint8_t *t, int64_t n, int64_t i)
{
    uint64_t k = 0;
    for (int j = 0; j < 32; ++j) k = (k << 2) | (i + j < n ? t[i + j] : 0);
    return k;
}

static void *sa_worker(void *arg)
{
    sa_work_t *w = (sa_work_t *)arg;
    const or_e2e_t *e = w->e;
    const uint8_t *t = e->text;
    const int64_t n = e->n;
    if (w->phase == 0 || w->phase == 1) {
        if (w->lo >= w->hi) return NULL;
        uint64_t k = key_at(t, n, w->lo);
        for (int64_t i = w->lo; i < w->hi; ++i) {
            if (i > w->lo) k = (k << 2) | (i + 31 < n ? t[i + 31] : 0);
            const uint32_t b = (uint32_t)(k >> 48);
            if (w->phase == 0) w->cnt[b]++;
            else w->sa[w->cnt[b]++] = i;
        }
        return NULL;
    }
    skey_t *buf = NULL;
    int64_t cap = 0;
    for (;;) {
        pthread_mutex_lock(w->mu);
        const int64_t b = (*w->next_bucket)++;
        pthread_mutex_unlock(w->mu);
        if (b >= 65536) break;
        const int64_t r0 = w->bstart[b], r1 = w->bstart[b + 1], m = r1 - r0;
        if (m < 2) continue;
        if (m > cap) { cap = m; buf = (skey_t *)realloc(buf, sizeof(skey_t) * (size_t)cap); }
        for (int64_t k = 0; k < m; ++k) buf[k] = (skey_t){key_at(t, n, w->sa[r0 + k]), w->sa[r0 + k]};
        qsort(buf, (size_t)m, sizeof(skey_t), skey_cmp);
        for (int64_t k = 0; k < m; ++k) w->sa[r0 + k] = buf[k].pos;
        for (int64_t a = 0; a < m;) {                          /* runs of equal keys */
            int64_t z = a + 1;
            while (z < m && buf[z].key == buf[a].key) ++z;
            if (z - a > 1) qsort_r(w->sa + r0 + a, (size_t)(z - a), sizeof(int64_t), suffix_cmp, (void *)e);
            a = z;
        }
    }
    free(buf);
    return NULL;
}

static void suffix_sort(or_e2e_t *e)
{
    const int64_t n = e->n;
    int nt = (int)sysconf(_SC_NPROCESSORS_ONLN);
    nt = nt < 1 ? 1 : nt > 16 ? 16 : nt;
    if (n < (1 << 16)) nt = 1;
    int64_t *sa = e->sa + 1;
    sa_work_t w[16];
    pthread_t th[16];
    int64_t *cnt = (int64_t *)calloc((size_t)nt * 65536, sizeof(int64_t));
    int64_t *bstart = (int64_t *)calloc(65537, sizeof(int64_t));
    int64_t next = 0;
    pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;
    for (int ph = 0; ph < 3; ++ph) {
        for (int k = 0; k < nt; ++k) {
            w[k] = (sa_work_t){e, sa, n * k / nt, n * (k + 1) / nt, cnt + (size_t)k * 65536, bstart,
                               &next, &mu, ph};
            pthread_create(&th[k], NULL, sa_worker, &w[k]);
        }
        for (int k = 0; k < nt; ++k) pthread_join(th[k], NULL);
        if (ph == 0) {                 /* slice counts -> each slice's write cursors */
            int64_t at = 0;
            for (int b = 0; b < 65536; ++b) {
                bstart[b] = at;
                for (int k = 0; k < nt; ++k) {
                    const int64_t c = cnt[(size_t)k * 65536 + b];
                    cnt[(size_t)k * 65536 + b] = at;
                    at += c;
                }
            }
            bstart[65536] = at;
        }
    }
    free(cnt);
    free(bstart);
}

or_e2e_t *or_e2e_create(const uint8_t *fwd, int64_t gl, const char *name, const or_scoring_t *sc)
{
    if (gl <= 0) return NULL;
    or_e2e_t *e = (or_e2e_t *)calloc(1, sizeof *e);
    e->sc = *sc;
    e->gl = gl;
    e->n = 2 * gl;
    snprintf(e->name, sizeof e->name, "%s", name ? name : "ref");
    e->fwd = (uint8_t *)malloc((size_t)gl);
    memcpy(e->fwd, fwd, (size_t)gl);
    e->text = (uint8_t *)malloc((size_t)e->n);
    for (int64_t i = 0; i < gl; ++i) e->text[i] = fwd[i] & 3;
    for (int64_t i = 0; i < gl; ++i) e->text[gl + i] = (uint8_t)(3 - (fwd[gl - 1 - i] & 3));
    e->sa = (int64_t *)malloc(sizeof(int64_t) * (size_t)(e->n + 1));
    e->sa[0] = e->n;                                           /* the '$' suffix */
    suffix_sort(e);
    e->bwt = (uint8_t *)malloc((size_t)(e->n + 1));
    e->sentinel = -1;
    for (int64_t i = 0; i <= e->n; ++i) {
        if (e->sa[i] > 0) e->bwt[i] = e->text[e->sa[i] - 1];
        else { e->bwt[i] = 4; e->sentinel = i; }
    }
    e->occ = (int32_t *)malloc(sizeof(int32_t) * (size_t)(e->n + 2) * 4);
    {
        int32_t o[4] = {0, 0, 0, 0};
        for (int64_t i = 0; i <= e->n; ++i) {
            memcpy(e->occ + i * 4, o, sizeof o);
            if (e->bwt[i] < 4) o[e->bwt[i]]++;
        }
        memcpy(e->occ + (e->n + 1) * 4, o, sizeof o);
    }
    int64_t f[4] = {0, 0, 0, 0};
    for (int64_t i = 0; i < e->n; ++i) f[e->text[i]]++;
    e->cnt[0] = 1;                                             /* count[c] + 1 (BWT.swift:10-11) */
    for (int c = 0; c < 4; ++c) e->cnt[c + 1] = e->cnt[c] + f[c];
    e->n_ctg = 1;
    e->ctg_off = (int64_t *)calloc(1, sizeof(int64_t));
    e->ctg_len = (int64_t *)malloc(sizeof(int64_t));
    e->ctg_len[0] = gl;
    e->ctg_name = (char (*)[256])calloc(1, 256);
    snprintf(e->ctg_name[0], 256, "%s", e->name);
    return e;
}

/* Several reference sequences concatenated in the forward text, as the
 * index's annotations describe them (offset, length, name per sequence;
 * names NUL-separated).  Returns 0, or -1 when they do not tile [0, gl). */
int or_e2e_set_contigs(or_e2e_t *e, int n, const int64_t *off, const int64_t *len, const char *names)
{
    if (n < 1) return -1;
    int64_t at = 0;
    for (int k = 0; k < n; ++k) {
        if (off[k] != at || len[k] <= 0) return -1;
        at += len[k];
    }
    if (at != e->gl) return -1;
    free(e->ctg_off); free(e->ctg_len); free(e->ctg_name);
    e->n_ctg = n;
    e->ctg_off = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    e->ctg_len = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    e->ctg_name = (char (*)[256])calloc((size_t)n, 256);
    for (int k = 0; k < n; ++k) {
        e->ctg_off[k] = off[k];
        e->ctg_len[k] = len[k];
        snprintf(e->ctg_name[k], 256, "%s", names);
        names += strlen(names) + 1;
    }
    return 0;
}

/* ReferenceMetadata.sequenceID (ReferenceMetadata.swift:77-89): the last
 * sequence whose offset is <= pos */
static int seq_id(const or_e2e_t *e, int64_t pos)
{
    int lo = 0, hi = e->n_ctg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (e->ctg_off[mid] <= pos) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

int or_e2e_seq_id(const or_e2e_t *e, int64_t pos) { return seq_id(e, pos); }

static void batch_clear(or_e2e_t *e)
{
    for (int i = 0; i < e->chains.n; ++i) free(e->chains.v[i].seeds);
    free(e->chains.v);
    memset(&e->chains, 0, sizeof e->chains);
    free(e->chain_read); e->chain_read = NULL;
    if (e->regs) {
        for (int i = 0; i < e->n_reads; ++i) free(e->regs[i].v);
        free(e->regs); e->regs = NULL;
    }
    for (int64_t j = 0; j < e->n_jobs; ++j) { free(e->jobs[j].cigar); free(e->jobs[j].md); }
    e->n_jobs = 0;
    e->cq_n = e->cr_n = 0;
    e->n_calls = 0; e->lq_n = e->lt_n = 0;
    e->n_resc = 0; e->rq_n = e->rt_n = 0; e->n_pairs = 0; e->do_rescue = 0;
    free(e->reads); e->reads = NULL;
    free(e->read_off); e->read_off = NULL;
    free(e->read_len); e->read_len = NULL;
    e->n_reads = 0;
}

void or_e2e_free(or_e2e_t *e)
{
    if (!e) return;
    batch_clear(e);
    free(e->jobs); free(e->cq); free(e->cr); free(e->calls); free(e->lq); free(e->lt);
    free(e->resc); free(e->rq); free(e->rt);
    if (!e->shared) {
        free(e->ctg_off); free(e->ctg_len); free(e->ctg_name);
        free(e->fwd); free(e->text); free(e->bwt); free(e->sa); free(e->occ);
    }
    free(e);
}

/* A handle with its own batch state over the index of `base`, which must
 * outlive it and keep its sequences unchanged; batches on different handles
 * may run in different threads. */
or_e2e_t *or_e2e_fork(const or_e2e_t *base)
{
    or_e2e_t *e = (or_e2e_t *)calloc(1, sizeof *e);
    e->sc = base->sc;
    e->gl = base->gl; e->n = base->n;
    e->fwd = base->fwd; e->text = base->text; e->bwt = base->bwt; e->sa = base->sa;
    memcpy(e->cnt, base->cnt, sizeof e->cnt);
    e->sentinel = base->sentinel;
    e->occ = base->occ;
    memcpy(e->name, base->name, sizeof e->name);
    e->n_ctg = base->n_ctg;
    e->ctg_off = base->ctg_off; e->ctg_len = base->ctg_len; e->ctg_name = base->ctg_name;
    e->shared = 1;
    return e;
}

int64_t or_e2e_sa(const or_e2e_t *e, int64_t i) { return e->sa[i]; }
int or_e2e_bwt(const or_e2e_t *e, int64_t i) { return e->bwt[i]; }
void or_e2e_counts(const or_e2e_t *e, int64_t *cnt5) { memcpy(cnt5, e->cnt, sizeof e->cnt); }

/* BWT.occ (BWT.swift:58-77): occurrences of c in bwt[0, pos) */
int64_t or_e2e_occ(const or_e2e_t *e, int64_t pos, int c) { return e->occ[pos * 4 + c]; }

/* BackwardSearch.backwardExt (BackwardSearch.swift:22-71) */
static bidi_t backward_ext(const or_e2e_t *e, int64_t k, int64_t l, int64_t s, int a)
{
    bidi_t r = {0, 0, 0, 0, 0};
    if (a >= 4) return r;
    const int64_t sp = k, ep = k + s;
    int64_t kk[4], ss[4], ll[4];
    for (int c = 0; c < 4; ++c) {
        int64_t o0 = or_e2e_occ(e, sp, c), o1 = or_e2e_occ(e, ep, c);
        kk[c] = e->cnt[c] + o0;
        ss[c] = o1 - o0;
    }
    int64_t sent = (sp <= e->sentinel && ep > e->sentinel) ? 1 : 0;
    ll[3] = l + sent;
    ll[2] = ll[3] + ss[3];
    ll[1] = ll[2] + ss[2];
    ll[0] = ll[1] + ss[1];
    r.k = kk[a]; r.l = ll[a]; r.s = ss[a];
    return r;
}

/* ---------------------------------------------------------------------- */
/* SMEMFinder.findSMEMsAtPosition (SMEMFinder.swift:90-270)                */
/* ---------------------------------------------------------------------- */
static int smems_at(const or_e2e_t *e, const uint8_t *q, int qlen, int start, int32_t min_seed,
                    int64_t min_intv, smem_vec *out)
{
    int next = start + 1;
    const int a = q[start];
    if (a >= 4) return next;                                   /* :104-108 */
    bidi_t smem = {e->cnt[a], e->cnt[3 - a], e->cnt[a + 1] - e->cnt[a], start, start};
    bidi_t *prev = (bidi_t *)malloc(sizeof(bidi_t) * (size_t)(qlen + 1));
    int n_prev = 0;
    for (int j = start + 1; j < qlen; ++j) {                   /* forward phase :124-157 */
        const int b = q[j];
        next = j + 1;
        if (b >= 4) break;
        bidi_t ext = backward_ext(e, smem.l, smem.k, smem.s, 3 - b);
        bidi_t ns = {ext.l, ext.k, ext.s, smem.m, j};
        if (ns.s != smem.s) prev[n_prev++] = smem;
        if (ns.s < min_intv) { next = j; break; }
        smem = ns;
    }
    if (smem.s >= min_intv) prev[n_prev++] = smem;             /* :160-163 */
    for (int lo = 0, hi = n_prev - 1; lo < hi; ++lo, --hi) {   /* :166 reverse */
        bidi_t t = prev[lo]; prev[lo] = prev[hi]; prev[hi] = t;
    }
    for (int bj = start - 1; bj >= 0; --bj) {                  /* backward phase :169-232 */
        const int b = q[bj];
        if (b >= 4) break;
        int n_curr = 0;
        int64_t curr_s = -1;
        int p = 0;
        for (; p < n_prev; ++p) {
            bidi_t iv = prev[p];
            bidi_t ext = backward_ext(e, iv.k, iv.l, iv.s, b);
            bidi_t ni = {ext.k, ext.l, ext.s, bj, iv.n};
            if (ni.s < min_intv && (iv.n - iv.m + 1) >= min_seed) {
                smem_t sm = {iv.k, iv.k + iv.s - 1, iv.m, iv.n + 1};
                VEC_PUSH(*out, sm);
                break;
            }
            if (ni.s >= min_intv && ni.s != curr_s) {
                curr_s = ni.s;
                prev[n_curr++] = ni;
                break;
            }
        }
        for (++p; p < n_prev; ++p) {
            bidi_t iv = prev[p];
            bidi_t ext = backward_ext(e, iv.k, iv.l, iv.s, b);
            bidi_t ni = {ext.k, ext.l, ext.s, bj, iv.n};
            if (ni.s >= min_intv && ni.s != curr_s) {
                curr_s = ni.s;
                prev[n_curr++] = ni;
            }
        }
        n_prev = n_curr;
        if (n_curr == 0) break;
    }
    if (n_prev != 0) {                                         /* :234-244 */
        bidi_t iv = prev[0];
        if ((iv.n - iv.m + 1) >= min_seed) {
            smem_t sm = {iv.k, iv.k + iv.s - 1, iv.m, iv.n + 1};
            VEC_PUSH(*out, sm);
        }
    }
    free(prev);
    return next;
}

/* order of SMEMFinder.swift:53-58: query begin ascending, then length descending */
static int smem_before(const void *a, const void *b, void *ctx)
{
    const smem_t *x = (const smem_t *)a, *y = (const smem_t *)b;
    if (x->qb != y->qb) return x->qb < y->qb;
    return (x->qe - x->qb) > (y->qe - y->qb);
}

/* SMEMFinder.findAllSMEMs (SMEMFinder.swift:28-60, arena form :407-444) */
static void find_all_smems(const or_e2e_t *e, const uint8_t *q, int qlen, int32_t min_seed,
                           int64_t min_intv, smem_vec *out)
{
    int pos = 0;
    while (pos < qlen) pos = smems_at(e, q, qlen, pos, min_seed, min_intv, out);
    stable_sort(out->v, (size_t)out->n, sizeof(smem_t), smem_before, NULL);
}

int or_e2e_smems(const or_e2e_t *e, const uint8_t *q, int qlen, int32_t min_seed_len,
                 int64_t min_intv, int64_t *rows, int cap)
{
    smem_vec v = {0};
    if (qlen > 0) find_all_smems(e, q, qlen, min_seed_len, min_intv, &v);
    for (int i = 0; i < v.n && i < cap; ++i) {
        rows[i * 4 + 0] = v.v[i].k; rows[i * 4 + 1] = v.v[i].l;
        rows[i * 4 + 2] = v.v[i].qb; rows[i * 4 + 3] = v.v[i].qe;
    }
    int n = v.n;
    free(v.v);
    return n;
}

/* ---------------------------------------------------------------------- */
/* phases 1-3: BWAMemAligner.alignReadPhase1to3 (:130-205)                 */
/* ---------------------------------------------------------------------- */

/* internalReseed, arena form (BWAMemAligner.swift:260-312) */
static void internal_reseed(const or_e2e_t *e, const uint8_t *q, int qlen, smem_vec *sm)
{
    const int split_len = (int)((float)e->sc.min_seed_len * SEED_SPLIT_RATIO + 0.499f);
    const int orig = sm->n;
    smem_vec extra = {0};
    for (int i = 0; i < orig; ++i) {
        smem_t s = sm->v[i];
        int len = s.qe - s.qb;
        int64_t count = s.l - s.k + 1;
        if (len < split_len || count > SPLIT_WIDTH) continue;
        int mid = (s.qb + s.qe) / 2;
        smems_at(e, q, qlen, mid, e->sc.min_seed_len, count + 1, &extra);
    }
    if (extra.n == 0) { free(extra.v); return; }
    for (int i = 0; i < extra.n; ++i) VEC_PUSH(*sm, extra.v[i]);
    free(extra.v);
    stable_sort(sm->v, (size_t)sm->n, sizeof(smem_t), smem_before, NULL);
}

/* computeFracRep (BWAMemAligner.swift:208-229) */
static float frac_rep(const smem_vec *sm, int read_len)
{
    int32_t l_rep = 0, rep_b = 0, rep_e = 0;
    for (int i = 0; i < sm->n; ++i) {
        const smem_t *s = &sm->v[i];
        if (s->l - s->k + 1 <= MAX_OCC) continue;
        if (s->qb > rep_e) { l_rep += rep_e - rep_b; rep_b = s->qb; rep_e = s->qe; }
        else if (s->qe > rep_e) rep_e = s->qe;
    }
    l_rep += rep_e - rep_b;
    return read_len > 0 ? (float)l_rep / (float)read_len : 0.0f;
}

static void chain_push(chain_t *c, seed_t s)
{
    if (c->n == c->cap) {
        c->cap = c->cap ? 2 * c->cap : 8;
        c->seeds = (seed_t *)realloc(c->seeds, sizeof(seed_t) * (size_t)c->cap);
    }
    c->seeds[c->n++] = s;
}

static int seed_before(const void *a, const void *b, void *ctx)
{
    return ((const seed_t *)a)->rbeg < ((const seed_t *)b)->rbeg;
}

/* SeedChainer.chain (SeedChainer.swift:16-107); no contig is ALT */
static void seed_chain(const or_e2e_t *e, const smem_vec *sm, chain_vec *out)
{
    for (int i = 0; i < sm->n; ++i) {
        const smem_t *s = &sm->v[i];
        int32_t seed_len = s->qe - s->qb;
        if (seed_len < e->sc.min_seed_len) continue;
        int64_t occs = s->l - s->k + 1;
        int64_t step = occs > MAX_OCC ? occs / MAX_OCC : 1;
        int64_t count = 0;
        for (int64_t sa_idx = 0; sa_idx < occs && count < MAX_OCC; sa_idx += step, ++count) {
            int64_t ref_pos = e->sa[s->k + sa_idx];
            /* :44-52: reverse-strand positions map to forward coordinates,
               and the sequence is the one holding min(rbeg, gl - 1) */
            int64_t rbeg = ref_pos;
            if (rbeg >= e->gl) rbeg = e->gl * 2 - 1 - rbeg - (int64_t)seed_len + 1;
            const int32_t rid = seq_id(e, rbeg < e->gl - 1 ? rbeg : e->gl - 1);
            seed_t seed = {ref_pos, s->qb, seed_len, seed_len * e->sc.a};
            int merged = 0;
            for (int ci = out->n - 1; ci >= 0; --ci) {
                chain_t *c = &out->v[ci];
                if (c->rid != rid || c->n == 0) continue;
                const seed_t *last = &c->seeds[c->n - 1];
                int64_t rgap = llabs(ref_pos - last->rbeg - (int64_t)last->len);
                int64_t qgap = llabs((int64_t)seed.qbeg - last->qbeg - (int64_t)last->len);
                int64_t gap = rgap > qgap ? rgap : qgap;
                if (gap < MAX_CHAIN_GAP) {
                    chain_push(c, seed);
                    c->weight += seed_len;
                    merged = 1;
                    break;
                }
            }
            if (!merged) {
                chain_t c = {0};
                chain_push(&c, seed);
                c.weight = seed_len;
                c.rid = rid;
                c.pos = ref_pos;
                c.kept = 3;
                c.first = -1;
                VEC_PUSH(*out, c);
            }
        }
    }
    for (int i = 0; i < out->n; ++i) {                         /* :97-103 */
        chain_t *c = &out->v[i];
        stable_sort(c->seeds, (size_t)c->n, sizeof(seed_t), seed_before, NULL);
        if (c->n) c->pos = c->seeds[0].rbeg;
    }
}

/* ChainFilter.chainWeight */
static int32_t chain_weight(const chain_t *c)
{
    int32_t w = 0;
    int64_t end = 0;
    for (int i = 0; i < c->n; ++i) {
        int64_t sb = c->seeds[i].qbeg, se = sb + c->seeds[i].len;
        if (sb >= end) w += c->seeds[i].len;
        else if (se > end) w += (int32_t)(se - end);
        if (se > end) end = se;
    }
    int32_t tmp = w;
    w = 0;
    end = 0;
    for (int i = 0; i < c->n; ++i) {
        int64_t sb = c->seeds[i].rbeg, se = sb + c->seeds[i].len;
        if (sb >= end) w += c->seeds[i].len;
        else if (se > end) w += (int32_t)(se - end);
        if (se > end) end = se;
    }
    w = w < tmp ? w : tmp;
    return w < (1 << 30) ? w : (1 << 30) - 1;
}

static int chain_heavier(const void *a, const void *b, void *ctx)
{
    return ((const chain_t *)a)->weight > ((const chain_t *)b)->weight;
}

/* ChainFilter.filter */
static void chain_filter(const or_e2e_t *e, chain_vec *cv)
{
    if (cv->n == 0) return;
    const int32_t min_w = MIN_CHAIN_WEIGHT > 0 ? MIN_CHAIN_WEIGHT : e->sc.min_seed_len;
    for (int i = 0; i < cv->n; ++i) cv->v[i].weight = chain_weight(&cv->v[i]);
    int m = 0;
    for (int i = 0; i < cv->n; ++i) {
        if (cv->v[i].weight < min_w) free(cv->v[i].seeds);
        else cv->v[m++] = cv->v[i];
    }
    cv->n = m;
    if (m == 0) return;
    stable_sort(cv->v, (size_t)m, sizeof(chain_t), chain_heavier, NULL);
    for (int i = 0; i < m; ++i) { cv->v[i].kept = 0; cv->v[i].first = -1; }
    cv->v[0].kept = 3;
    int *kept = (int *)malloc(sizeof(int) * (size_t)m);
    int nk = 0;
    kept[nk++] = 0;
    for (int i = 1; i < m; ++i) {
        chain_t *ci = &cv->v[i];
        int32_t ib = ci->n ? ci->seeds[0].qbeg : 0;
        int32_t ie = ci->n ? ci->seeds[ci->n - 1].qbeg + ci->seeds[ci->n - 1].len : 0;
        int32_t il = ie - ib;
        int large = 0, suppressed = 0;
        for (int x = 0; x < nk; ++x) {
            chain_t *ck = &cv->v[kept[x]];
            int32_t kb = ck->n ? ck->seeds[0].qbeg : 0;
            int32_t ke = ck->n ? ck->seeds[ck->n - 1].qbeg + ck->seeds[ck->n - 1].len : 0;
            int32_t kl = ke - kb;
            int32_t ob = ib > kb ? ib : kb, oe = ie < ke ? ie : ke;
            if (ob < oe) {                                     /* no ALT chains here */
                int32_t ov = oe - ob, ml = il < kl ? il : kl;
                if (ml > 0 && (float)ov >= (float)ml * MASK_LEVEL && ml < MAX_CHAIN_GAP) {
                    large = 1;
                    if (ck->first < 0) ck->first = i;
                    if ((float)ci->weight < (float)ck->weight * CHAIN_DROP_RATIO &&
                        ck->weight - ci->weight >= (e->sc.min_seed_len << 1)) {
                        suppressed = 1;
                        break;
                    }
                }
            }
        }
        if (!suppressed) {
            kept[nk++] = i;
            ci->kept = large ? 2 : 3;
        }
    }
    for (int x = 0; x < nk; ++x) {
        int f = cv->v[kept[x]].first;
        if (f >= 0 && f < m) cv->v[f].kept = 1;
    }
    free(kept);
    int n2 = 0;
    for (int i = 0; i < m; ++i) {
        if (cv->v[i].kept == 0) free(cv->v[i].seeds);
        else cv->v[n2++] = cv->v[i];
    }
    cv->n = n2;
}

/* BWAMemAligner.alignRead phases 1-3 (:130-200) for reads [lo, hi) */
typedef struct {
    const or_e2e_t *e;
    chain_vec *per;
    int lo, hi;
} seed_work_t;

static void *seed_worker(void *arg)
{
    seed_work_t *w = (seed_work_t *)arg;
    const or_e2e_t *e = w->e;
    for (int r = w->lo; r < w->hi; ++r) {
        const uint8_t *q = e->reads + e->read_off[r];
        const int qlen = e->read_len[r];
        smem_vec sm = {0};
        if (qlen > 0) find_all_smems(e, q, qlen, e->sc.min_seed_len, 1, &sm);
        if (sm.n == 0) { free(sm.v); continue; }               /* :140 */
        internal_reseed(e, q, qlen, &sm);                      /* :143 */
        /* -y reseeding (:149-171) is off at the default reseedLength 0 */
        stable_sort(sm.v, (size_t)sm.n, sizeof(smem_t), smem_before, NULL);   /* :174-177 */
        float fr = frac_rep(&sm, qlen);                        /* :181 */
        chain_vec cv = {0};
        seed_chain(e, &sm, &cv);                               /* :184-192 */
        for (int i = 0; i < cv.n; ++i) cv.v[i].frac_rep = fr;  /* :195-197 */
        chain_filter(e, &cv);                                  /* :200 */
        w->per[r] = cv;
        free(sm.v);
    }
    return NULL;
}

int64_t or_e2e_seed(or_e2e_t *e, int n_reads, const uint8_t *bases, const int64_t *off,
                    const int32_t *len)
{
    batch_clear(e);
    e->n_reads = n_reads;
    int64_t total = 0;
    for (int i = 0; i < n_reads; ++i) total += len[i];
    e->reads = (uint8_t *)malloc((size_t)(total ? total : 1));
    e->read_off = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n_reads + 1));
    e->read_len = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n_reads + 1));
    int64_t o = 0;
    for (int i = 0; i < n_reads; ++i) {
        memcpy(e->reads + o, bases + off[i], (size_t)len[i]);
        e->read_off[i] = o;
        e->read_len[i] = len[i];
        o += len[i];
    }
    e->regs = (region_vec *)calloc((size_t)(n_reads > 0 ? n_reads : 1), sizeof(region_vec));
    /* phases 1-3 of each read (independent), on threads over read slices */
    chain_vec *per = (chain_vec *)calloc((size_t)(n_reads > 0 ? n_reads : 1), sizeof(chain_vec));
    int nt = (int)sysconf(_SC_NPROCESSORS_ONLN);
    nt = nt < 1 ? 1 : nt > 16 ? 16 : nt;
    if (n_reads < 256) nt = 1;
    seed_work_t w[16];
    pthread_t th[16];
    for (int k = 0; k < nt; ++k) {
        w[k] = (seed_work_t){e, per, (int)((int64_t)n_reads * k / nt), (int)((int64_t)n_reads * (k + 1) / nt)};
        if (nt > 1) pthread_create(&th[k], NULL, seed_worker, &w[k]);
        else seed_worker(&w[k]);
    }
    if (nt > 1)
        for (int k = 0; k < nt; ++k) pthread_join(th[k], NULL);
    int32_t *cr = NULL;
    int64_t cr_cap = 0;
    for (int r = 0; r < n_reads; ++r) {
        for (int i = 0; i < per[r].n; ++i) {
            VEC_PUSH(e->chains, per[r].v[i]);
            cr = (int32_t *)grow(cr, &cr_cap, e->chains.n, sizeof(int32_t));
            cr[e->chains.n - 1] = r;
        }
        free(per[r].v);
    }
    free(per);
    e->chain_read = cr;
    return e->chains.n;
}

int64_t or_e2e_n_seeds(const or_e2e_t *e)
{
    int64_t n = 0;
    for (int i = 0; i < e->chains.n; ++i) n += e->chains.v[i].n;
    return n;
}

void or_e2e_get_chains(const or_e2e_t *e, int32_t *chain_read, int64_t *chain_off,
                       int64_t *rbeg, int32_t *qbeg, int32_t *slen, int32_t *sscore)
{
    int64_t o = 0;
    for (int i = 0; i < e->chains.n; ++i) {
        const chain_t *c = &e->chains.v[i];
        chain_read[i] = e->chain_read[i];
        chain_off[i] = o;
        for (int k = 0; k < c->n; ++k, ++o) {
            rbeg[o] = c->seeds[k].rbeg; qbeg[o] = c->seeds[k].qbeg;
            slen[o] = c->seeds[k].len; sscore[o] = c->seeds[k].score;
        }
    }
    chain_off[e->chains.n] = o;
}

/* ---------------------------------------------------------------------- */
/* phase 4: extension (ExtensionAligner.extend via or_extend_chain_ex)     */
/* ---------------------------------------------------------------------- */
static void log_call(void *ctx, const uint8_t *q, int qlen, const uint8_t *t, int tlen,
                     int32_t h0, int32_t w, int32_t eb, const or_swres_t *r)
{
    or_e2e_t *e = (or_e2e_t *)ctx;
    e->calls = (call_t *)grow(e->calls, &e->calls_cap, e->n_calls + 1, sizeof(call_t));
    e->lq = (uint8_t *)grow(e->lq, &e->lq_cap, e->lq_n + qlen + 1, 1);
    e->lt = (uint8_t *)grow(e->lt, &e->lt_cap, e->lt_n + tlen + 1, 1);
    call_t *c = &e->calls[e->n_calls++];
    c->q_off = e->lq_n; c->t_off = e->lt_n;
    c->q_len = qlen; c->t_len = tlen; c->h0 = h0; c->w = w; c->eb = eb; c->r = *r;
    memcpy(e->lq + e->lq_n, q, (size_t)qlen); e->lq_n += qlen;
    memcpy(e->lt + e->lt_n, t, (size_t)tlen); e->lt_n += tlen;
}

static void add_regions(or_e2e_t *e, int ci, const int64_t *rows, int n)
{
    const chain_t *c = &e->chains.v[ci];
    region_vec *rv = &e->regs[e->chain_read[ci]];
    for (int k = 0; k < n; ++k) {
        const int64_t *o = rows + (size_t)k * 10;
        region_t g;
        memset(&g, 0, sizeof g);
        g.qb = (int32_t)o[0]; g.qe = (int32_t)o[1]; g.rb = o[2]; g.re = o[3];
        g.score = (int32_t)o[4]; g.true_score = (int32_t)o[5]; g.sub = (int32_t)o[6];
        g.sub_n = (int32_t)o[7]; g.w = (int32_t)o[8]; g.seed_cov = (int32_t)o[9];
        g.rid = c->rid; g.frac_rep = c->frac_rep;               /* ExtensionAligner.swift:191-195 */
        g.secondary = -1;
        g.cig = -1;
        VEC_PUSH(*rv, g);
    }
}

int64_t or_e2e_extend_cpu(or_e2e_t *e)
{
    e->n_calls = 0; e->lq_n = e->lt_n = 0;
    int64_t total = 0;
    for (int ci = 0; ci < e->chains.n; ++ci) {
        const chain_t *c = &e->chains.v[ci];
        int r = e->chain_read[ci];
        int32_t *sq = (int32_t *)malloc(sizeof(int32_t) * (size_t)c->n * 3);
        int32_t *sl = sq + c->n, *ss = sq + 2 * c->n;
        int64_t *sr = (int64_t *)malloc(sizeof(int64_t) * (size_t)c->n);
        for (int k = 0; k < c->n; ++k) {
            sq[k] = c->seeds[k].qbeg; sl[k] = c->seeds[k].len;
            ss[k] = c->seeds[k].score; sr[k] = c->seeds[k].rbeg;
        }
        int64_t *rows = (int64_t *)malloc(sizeof(int64_t) * 10 * (size_t)c->n);
        int nreg = or_extend_chain_ex(c->n, sq, sl, sr, ss, e->reads + e->read_off[r],
                                      e->read_len[r], e->fwd, e->gl, e->gl, &e->sc, rows, c->n,
                                      NULL, NULL, log_call, e);
        add_regions(e, ci, rows, nreg);
        total += nreg;
        free(sq); free(sr); free(rows);
    }
    return total;
}

void or_e2e_set_regions(or_e2e_t *e, const int32_t *n_per_chain, const int64_t *rows)
{
    for (int i = 0; i < e->n_reads; ++i) e->regs[i].n = 0;
    int64_t o = 0;
    for (int ci = 0; ci < e->chains.n; ++ci) {
        add_regions(e, ci, rows + o * 10, n_per_chain[ci]);
        o += n_per_chain[ci];
    }
}

int64_t or_e2e_n_calls(const or_e2e_t *e, int64_t *q_bytes, int64_t *t_bytes)
{
    if (q_bytes) *q_bytes = e->lq_n;
    if (t_bytes) *t_bytes = e->lt_n;
    return e->n_calls;
}

void or_e2e_get_calls(const or_e2e_t *e, uint8_t *q, int64_t *q_off, int32_t *q_len,
                      uint8_t *t, int64_t *t_off, int32_t *t_len, int32_t *h0, int32_t *w,
                      int32_t *end_bonus, int32_t *out6)
{
    if (e->lq_n) memcpy(q, e->lq, (size_t)e->lq_n);
    if (e->lt_n) memcpy(t, e->lt, (size_t)e->lt_n);
    for (int64_t i = 0; i < e->n_calls; ++i) {
        const call_t *c = &e->calls[i];
        q_off[i] = c->q_off; q_len[i] = c->q_len; t_off[i] = c->t_off; t_len[i] = c->t_len;
        h0[i] = c->h0; w[i] = c->w; end_bonus[i] = c->eb;
        memcpy(out6 + i * 6, &c->r, sizeof(or_swres_t));
    }
}

/* ---------------------------------------------------------------------- */
/* phases 4.5-5: RegionDedup.sortDedupPatch, ChainFilter.markSecondary     */
/* ---------------------------------------------------------------------- */

/* FMIndex.getReference, array form (FMIndex.swift:52-59) */
static int get_reference(const or_e2e_t *e, int64_t pos, int len, uint8_t *buf)
{
    const int64_t gl = e->gl;
    if (pos >= gl) {
        int64_t fe = 2 * gl - pos, fs = fe - len;
        int64_t s0 = fs > 0 ? fs : 0, e0 = fe < gl ? fe : gl;
        int n = e0 > s0 ? (int)(e0 - s0) : 0;
        for (int k = 0; k < n; ++k) buf[k] = (uint8_t)(3 - e->fwd[e0 - 1 - k]);
        return n;
    }
    int64_t n = (int64_t)len < gl - pos ? (int64_t)len : gl - pos;
    if (n <= 0 || pos < 0) return 0;
    memcpy(buf, e->fwd + pos, (size_t)n);
    return (int)n;
}

/* RegionDedup.patchRegion */
static int32_t patch_region(const or_e2e_t *e, const region_t *a, const region_t *b,
                            const uint8_t *query, int32_t w)
{
    const int64_t gl = e->gl;
    if ((a->rb >= gl) != (b->rb >= gl)) return 0;
    if (!(a->qb < b->qb && a->qe < b->qe && a->re < b->re)) return 0;
    int64_t ref_gap = b->rb - a->re;
    int64_t q_gap = (int64_t)b->qb - (int64_t)a->qe;
    int32_t bw = abs((int32_t)(ref_gap - q_gap));
    float rel;
    if (a->re > b->rb) rel = (float)bw / (float)(a->re - b->rb);
    else rel = (float)bw / (float)(b->rb - a->re);
    if (a->re <= b->rb) {
        if (bw > w << 1 || rel >= 0.05f) return 0;
    } else {
        if (bw > w << 2 || rel >= 0.05f * 2) return 0;
    }
    int32_t mqb = a->qb, mqe = b->qe;
    int64_t mrb = a->rb, mre = b->re;
    if (!(mqe > mqb && mre > mrb)) return 0;
    int qn = mqe - mqb;
    int rl = (int)(mre - mrb);
    uint8_t *rbuf = (uint8_t *)malloc((size_t)rl + 1);
    int rn = get_reference(e, mrb, rl, rbuf);
    if (qn <= 0 || rn <= 0) { free(rbuf); return 0; }
    int32_t mw = w > bw ? w : bw;
    int32_t score = 0;
    int64_t cells = 0;
    or_ksw_global(qn, query + mqb, rn, rbuf, &e->sc, mw, &score, NULL, 0, &cells);
    free(rbuf);
    int32_t ql = mqe - mqb, rlen = (int32_t)(mre - mrb);
    int32_t d = abs(ql - rlen);
    int32_t ex_q = e->sc.a * ql - e->sc.o_ins - e->sc.e_ins * d;
    int32_t ex_r = e->sc.a * rlen - e->sc.o_ins - e->sc.e_ins * d;
    int32_t ex = ex_q > ex_r ? ex_q : ex_r;
    if (ex <= 0) return 0;
    if ((float)score / (float)ex >= 0.90f) return score;
    return 0;
}

static int reg_by_rid_re(const void *a, const void *b, void *ctx)
{
    const region_t *x = (const region_t *)a, *y = (const region_t *)b;
    if (x->rid != y->rid) return x->rid < y->rid;
    return x->re < y->re;
}

static int reg_by_score(const void *a, const void *b, void *ctx)
{
    return ((const region_t *)a)->score > ((const region_t *)b)->score;
}

static void remove_deleted(region_vec *rv)
{
    int m = 0;
    for (int i = 0; i < rv->n; ++i)
        if (rv->v[i].qe != rv->v[i].qb) rv->v[m++] = rv->v[i];
    rv->n = m;
}

/* RegionDedup.sortDedupPatch */
static void sort_dedup_patch(const or_e2e_t *e, region_vec *rv, const uint8_t *query)
{
    if (rv->n <= 1) return;
    region_t *R = rv->v;
    stable_sort(R, (size_t)rv->n, sizeof(region_t), reg_by_rid_re, NULL);
    const int n = rv->n;
    for (int i = 0; i < n; ++i) {
        if (!(R[i].qe > R[i].qb)) continue;
        for (int j = i - 1; j >= 0; --j) {
            if (R[j].rid != R[i].rid) break;
            if (!(R[j].rb < R[i].re + (int64_t)MAX_CHAIN_GAP)) continue;
            if (!(R[j].qe > R[j].qb)) continue;
            int64_t rob = R[i].rb > R[j].rb ? R[i].rb : R[j].rb;
            int64_t roe = R[i].re < R[j].re ? R[i].re : R[j].re;
            int32_t qob = R[i].qb > R[j].qb ? R[i].qb : R[j].qb;
            int32_t qoe = R[i].qe < R[j].qe ? R[i].qe : R[j].qe;
            if (rob < roe && qob < qoe) {
                float ro = (float)(roe - rob), qo = (float)(qoe - qob);
                int64_t rli = R[i].re - R[i].rb, rlj = R[j].re - R[j].rb;
                int32_t qli = R[i].qe - R[i].qb, qlj = R[j].qe - R[j].qb;
                float mr = (float)(rli < rlj ? rli : rlj), mq = (float)(qli < qlj ? qli : qlj);
                if (mr > 0 && mq > 0 && ro > MASK_LEVEL_REDUN * mr && qo > MASK_LEVEL_REDUN * mq) {
                    if (R[i].score < R[j].score) { R[i].qe = R[i].qb; break; }
                    R[j].qe = R[j].qb;
                    continue;
                }
            }
            int p, qx;
            if (R[j].rb < R[i].rb) { p = i; qx = j; } else { p = j; qx = i; }
            if (R[qx].qb < R[p].qb) {
                int32_t ww = R[qx].w > R[p].w ? R[qx].w : R[p].w;
                int32_t ms = patch_region(e, &R[qx], &R[p], query, ww);
                if (ms > 0) {
                    R[p].rb = R[p].rb < R[qx].rb ? R[p].rb : R[qx].rb;
                    R[p].re = R[p].re > R[qx].re ? R[p].re : R[qx].re;
                    R[p].qb = R[p].qb < R[qx].qb ? R[p].qb : R[qx].qb;
                    R[p].qe = R[p].qe > R[qx].qe ? R[p].qe : R[qx].qe;
                    R[p].score = ms;
                    R[p].true_score = ms;
                    R[p].w = R[p].w > R[qx].w ? R[p].w : R[qx].w;
                    R[qx].qe = R[qx].qb;
                    if (qx == i) break;
                }
            }
        }
    }
    remove_deleted(rv);
    stable_sort(rv->v, (size_t)rv->n, sizeof(region_t), reg_by_score, NULL);
    if (rv->n > 1) {
        R = rv->v;
        for (int i = 1; i < rv->n; ++i)
            if (R[i].score == R[i - 1].score && R[i].rb == R[i - 1].rb && R[i].qb == R[i - 1].qb)
                R[i].qe = R[i].qb;
        remove_deleted(rv);
    }
}

/* hash64 (AlignmentRegion.swift:3-14), wrapping 64-bit arithmetic */
static uint64_t hash64(uint64_t k)
{
    k += ~(k << 32);
    k ^= (k >> 22);
    k += ~(k << 13);
    k ^= (k >> 8);
    k += (k << 3);
    k ^= (k >> 15);
    k += ~(k << 27);
    k ^= (k >> 31);
    return k;
}

static int reg_by_score_hash(const void *a, const void *b, void *ctx)
{
    const region_t *x = (const region_t *)a, *y = (const region_t *)b;
    if (x->score != y->score) return x->score > y->score;
    return x->hash < y->hash;
}

/* ChainFilter.markSecondary + markSecondaryCore (no ALT regions) */
static void mark_secondary(const or_e2e_t *e, region_vec *rv, uint64_t read_id)
{
    if (rv->n <= 1) return;
    region_t *R = rv->v;
    for (int i = 0; i < rv->n; ++i) {
        R[i].sub = 0; R[i].sub_n = 0; R[i].secondary = -1;
        R[i].hash = hash64(read_id + (uint64_t)i);
    }
    stable_sort(R, (size_t)rv->n, sizeof(region_t), reg_by_score_hash, NULL);
    const or_scoring_t *sc = &e->sc;
    int32_t tmp = sc->a + sc->b;
    int32_t g1 = sc->o_ins + sc->e_ins, g2 = sc->o_del + sc->e_del;
    if (g1 > tmp) tmp = g1;
    if (g2 > tmp) tmp = g2;
    int *pri = (int *)malloc(sizeof(int) * (size_t)rv->n);
    int np = 0;
    pri[np++] = 0;
    for (int i = 1; i < rv->n; ++i) {
        int found = 0;
        for (int x = 0; x < np; ++x) {
            int pi = pri[x];
            int32_t ob = R[pi].qb > R[i].qb ? R[pi].qb : R[i].qb;
            int32_t oe = R[pi].qe < R[i].qe ? R[pi].qe : R[i].qe;
            if (ob < oe) {
                float ov = (float)(oe - ob);
                int32_t lp = R[pi].qe - R[pi].qb, li = R[i].qe - R[i].qb;
                float ml = (float)(lp < li ? lp : li);
                if (ml > 0 && ov / ml > MASK_LEVEL) {
                    if (R[pi].sub == 0) R[pi].sub = R[i].score;
                    if (R[pi].score - R[i].score <= tmp) R[pi].sub_n += 1;   /* no ALT */
                    R[i].secondary = pi;
                    found = 1;
                    break;
                }
            }
        }
        if (!found) pri[np++] = i;
    }
    free(pri);
}

/* generateCIGAR's segment preparation (BWAMemAligner.swift:841-889) */
static int64_t add_cigar_job(or_e2e_t *e, int r, const region_t *g)
{
    const int64_t gl = e->gl;
    const uint8_t *read = e->reads + e->read_off[r];
    const int rl = e->read_len[r];
    const int rev = g->rb >= gl;
    cjob_t j;
    memset(&j, 0, sizeof j);
    j.read = r;
    j.q_len = g->qe - g->qb;
    e->cq = (uint8_t *)grow(e->cq, &e->cq_cap, e->cq_n + j.q_len + 1, 1);
    j.q_off = e->cq_n;
    if (rev) {
        int rq0 = rl - g->qe;                                 /* rc(read)[rl-qe ..< rl-qb] */
        for (int k = 0; k < j.q_len; ++k) {
            uint8_t b = read[rl - 1 - (rq0 + k)];
            e->cq[e->cq_n + k] = b < 4 ? (uint8_t)(3 - b) : 4;
        }
    } else {
        memcpy(e->cq + e->cq_n, read + g->qb, (size_t)j.q_len);
    }
    e->cq_n += j.q_len;
    int64_t rb = g->rb, re = g->re;
    if (rev) { int64_t frb = 2 * gl - re, fre = 2 * gl - rb; rb = frb; re = fre; }
    {                                                          /* :866-872 chromosome end */
        const int rid = seq_id(e, rb);
        const int64_t chr_end = e->ctg_off[rid] + e->ctg_len[rid];
        if (re > chr_end) re = chr_end;
    }
    int64_t ref_len = re - rb;
    int64_t safe = ref_len < gl - rb ? ref_len : gl - rb;
    j.r_off = e->cr_n;
    if (safe > 0 && rb >= 0) {
        e->cr = (uint8_t *)grow(e->cr, &e->cr_cap, e->cr_n + safe + 1, 1);
        memcpy(e->cr + e->cr_n, e->fwd + rb, (size_t)safe);
        j.r_len = (int32_t)safe;
        e->cr_n += safe;
    }
    j.qb = g->qb; j.qe = g->qe; j.read_len = rl; j.rev = (uint8_t)rev;
    j.true_score = g->true_score; j.w = g->w; j.ref_pos = rb;
    e->jobs = (cjob_t *)grow(e->jobs, &e->jobs_cap, e->n_jobs + 1, sizeof(cjob_t));
    e->jobs[e->n_jobs] = j;
    return e->n_jobs++;
}

/* phases 4.5-5 of one read (BWAMemAligner.swift:402-421) */
static void dedup_read(or_e2e_t *e, int r, uint64_t read_id)
{
    region_vec *rv = &e->regs[r];
    if (rv->n == 0) return;
    sort_dedup_patch(e, rv, e->reads + e->read_off[r]);
    mark_secondary(e, rv, read_id);
}

/* generateFilteredCIGARs (BWAMemAligner.swift:434-476), default flags (no -a):
 * the regions that get a CIGAR become jobs, the rest keep the dummy */
static void cigar_plan_read(or_e2e_t *e, int r)
{
    region_vec *rv = &e->regs[r];
    region_t *R = rv->v;
    int sec = 0;
    for (int i = 0; i < rv->n; ++i) {
        if (i > 0 && R[i].score < MIN_OUT_SCORE) continue;
        if (R[i].secondary >= 0) {
            int pi = R[i].secondary;
            if (pi < rv->n && R[i].score < R[pi].score / 2) continue;
            sec++;
        }
    }
    const int skip_all = sec > MAX_XA_HITS;
    for (int i = 0; i < rv->n; ++i) {
        R[i].cig = -1;
        if (i > 0 && R[i].score < MIN_OUT_SCORE) continue;
        if (R[i].secondary >= 0) {
            int pi = R[i].secondary;
            if (pi < rv->n && R[i].score < R[pi].score / 2) continue;
            if (i > 0 && skip_all) continue;
        }
        R[i].cig = add_cigar_job(e, r, &R[i]);
    }
}

int64_t or_e2e_finish(or_e2e_t *e)
{
    for (int64_t j = 0; j < e->n_jobs; ++j) { free(e->jobs[j].cigar); free(e->jobs[j].md); }
    e->n_jobs = 0; e->cq_n = e->cr_n = 0;
    for (int r = 0; r < e->n_reads; ++r) {
        dedup_read(e, r, (uint64_t)r);                         /* SE: readId = batch index */
        cigar_plan_read(e, r);
    }
    return e->n_jobs;
}

void or_e2e_cigar_sizes(const or_e2e_t *e, int64_t *q_bytes, int64_t *r_bytes)
{
    *q_bytes = e->cq_n;
    *r_bytes = e->cr_n;
}

void or_e2e_get_cigar_jobs(const or_e2e_t *e, uint8_t *qseg, int64_t *q_off, int32_t *q_len,
                           uint8_t *rseg, int64_t *r_off, int32_t *r_len, int32_t *qb,
                           int32_t *qe, int32_t *read_len, uint8_t *rev, int32_t *true_score,
                           int32_t *w, int64_t *ref_pos)
{
    if (e->cq_n) memcpy(qseg, e->cq, (size_t)e->cq_n);
    if (e->cr_n) memcpy(rseg, e->cr, (size_t)e->cr_n);
    for (int64_t i = 0; i < e->n_jobs; ++i) {
        const cjob_t *j = &e->jobs[i];
        q_off[i] = j->q_off; q_len[i] = j->q_len; r_off[i] = j->r_off; r_len[i] = j->r_len;
        qb[i] = j->qb; qe[i] = j->qe; read_len[i] = j->read_len; rev[i] = j->rev;
        true_score[i] = j->true_score; w[i] = j->w; ref_pos[i] = j->ref_pos;
    }
}

void or_e2e_set_cigar(or_e2e_t *e, int64_t jx, const uint32_t *cigar, int n_cigar, int32_t nm,
                      const char *md, int64_t pos)
{
    cjob_t *j = &e->jobs[jx];
    free(j->cigar); free(j->md);
    j->cigar = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n_cigar ? n_cigar : 1));
    memcpy(j->cigar, cigar, sizeof(uint32_t) * (size_t)n_cigar);
    j->n_cigar = n_cigar;
    j->nm = nm;
    j->md = strdup(md ? md : "");
    j->pos = pos;
    j->done = 1;
}

void or_e2e_cigar_cpu(or_e2e_t *e)
{
    enum { CAP = 4096, MDCAP = 8192 };
    uint32_t *cig = (uint32_t *)malloc(sizeof(uint32_t) * CAP);
    char *md = (char *)malloc(MDCAP);
    for (int64_t i = 0; i < e->n_jobs; ++i) {
        const cjob_t *j = &e->jobs[i];
        or_cigar_res_t res;
        int n = or_cigar_generate(e->cq + j->q_off, j->q_len, e->cr + j->r_off, j->r_len, j->qb,
                                  j->qe, j->read_len, j->rev, j->true_score, j->w, &e->sc,
                                  j->ref_pos, cig, CAP, md, MDCAP, &res);
        or_e2e_set_cigar(e, i, cig, n < CAP ? n : CAP, res.nm, md, res.pos);
    }
    free(cig); free(md);
}

/* ---------------------------------------------------------------------- */
/* classification, MAPQ and SAM text                                       */
/* ---------------------------------------------------------------------- */

/* MappingQuality.compute */
static int mapq_of(const or_e2e_t *e, const region_t *g)
{
    if (g->score <= 0) return 0;
    const or_scoring_t *sc = &e->sc;
    int32_t sub = g->sub > 0 ? g->sub : sc->min_seed_len * sc->a;
    if (g->csub > sub) sub = g->csub;
    if (sub >= g->score) return 0;
    int32_t ql = g->qe - g->qb, rl = (int32_t)(g->re - g->rb);
    int32_t l = ql > rl ? ql : rl;
    double a = sc->a, b = sc->b;
    double identity = 1.0 - ((double)l * a - (double)g->score) / (a + b) / (double)l;
    int32_t mapq;
    {
        double dl = l;
        double tmp = dl < 50.0 ? 1.0 : 3.0 / log(dl);
        tmp *= identity * identity;
        mapq = (int32_t)(6.02 * (double)(g->score - sub) / a * tmp * tmp + 0.499);
    }
    if (g->sub_n > 0) mapq -= (int32_t)(4.343 * log((double)(g->sub_n + 1)) + 0.499);
    if (mapq > 60) mapq = 60;
    if (mapq < 0) mapq = 0;
    mapq = (int32_t)((double)mapq * (1.0 - (double)g->frac_rep) + 0.499);
    if (mapq < 0) mapq = 0;
    if (mapq > 255) mapq = 255;
    return mapq;
}

typedef struct { char *s; int64_t n, cap; } sbuf_t;

static void sb_put(sbuf_t *b, const char *s, int64_t n)
{
    if (b->n + n + 1 > b->cap) {
        int64_t c = b->cap ? b->cap : 4096;
        while (c < b->n + n + 1) c *= 2;
        b->s = (char *)realloc(b->s, (size_t)c);
        b->cap = c;
    }
    memcpy(b->s + b->n, s, (size_t)n);
    b->n += n;
    b->s[b->n] = 0;
}

static void sb_printf(sbuf_t *b, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
static void sb_printf(sbuf_t *b, const char *fmt, ...)
{
    char tmp[1024];
    va_list ap;
    va_start(ap, fmt);
    int n = vsnprintf(tmp, sizeof tmp, fmt, ap);
    va_end(ap);
    if (n < (int)sizeof tmp) { sb_put(b, tmp, n); return; }
    char *big = (char *)malloc((size_t)n + 1);
    va_start(ap, fmt);
    vsnprintf(big, (size_t)n + 1, fmt, ap);
    va_end(ap);
    sb_put(b, big, n);
    free(big);
}

static void cigar_str(sbuf_t *b, const uint32_t *c, int n, int hard)
{
    static const char OPS[] = "MIDNSHP=X";
    if (n == 0) { sb_put(b, "*", 1); return; }
    for (int k = 0; k < n; ++k) {
        uint32_t op = c[k] & 0xF, len = c[k] >> 4;
        if (hard && op == 4 && (k == 0 || (k == n - 1 && n > 1))) op = 5;   /* convertToHardClip */
        sb_printf(b, "%u%c", len, op < 9 ? OPS[op] : '?');
    }
}

static const char BASES[] = "ACGTN";

static void seq_qual(sbuf_t *b, const uint8_t *read, const uint8_t *qual, int rl, int rev,
                     int t0, int t1)
{
    for (int k = t0; k < rl - t1; ++k) {
        uint8_t x = rev ? read[rl - 1 - k] : read[k];
        char c = rev ? (x < 4 ? BASES[3 - x] : 'N') : BASES[x < 4 ? x : 4];
        sb_put(b, &c, 1);
    }
    sb_put(b, "\t", 1);
    for (int k = t0; k < rl - t1; ++k) {
        char c = (char)(rev ? qual[rl - 1 - k] : qual[k]);
        sb_put(b, &c, 1);
    }
}


/* CIGARInfo (BWAMemAligner.swift:817-839, the dummy :425-429) */
typedef struct {
    const uint32_t *cig;
    int n;
    int32_t nm;
    const char *md;
    int64_t pos;
    int rev;
    int64_t ref_len;                   /* reference consumed */
    int dummy;                         /* cigarString "*" */
    int rid;                           /* decodePosition(pos) */
    int64_t lpos;
} cinfo_t;

static cinfo_t cinfo_of(const or_e2e_t *e, const region_t *g)
{
    cinfo_t c;
    memset(&c, 0, sizeof c);
    c.md = "";
    if (g->cig < 0) { c.dummy = 1; return c; }                 /* pos 0: rid 0, local 0 */
    const cjob_t *j = &e->jobs[g->cig];
    c.cig = j->cigar; c.n = j->n_cigar; c.nm = j->nm; c.md = j->md ? j->md : "";
    c.pos = j->pos; c.rev = j->rev;
    for (int k = 0; k < c.n; ++k) {
        uint32_t op = c.cig[k] & 0xF;
        if (op == 0 || op == 2 || op == 3 || op == 7 || op == 8) c.ref_len += c.cig[k] >> 4;
    }
    c.rid = seq_id(e, c.pos);                                  /* ReferenceMetadata.swift:92-96 */
    c.lpos = c.pos - e->ctg_off[c.rid];
    return c;
}

static void cinfo_str(sbuf_t *b, const cinfo_t *c, int hard)
{
    if (c->dummy) sb_put(b, "*", 1);
    else cigar_str(b, c->cig, c->n, hard);
}

/* PairedEndInfo (SAMOutputBuilder.swift:5-31) */
typedef struct {
    int on, is_r1, proper, mate_tid, mate_rev, mate_unmapped;
    int64_t mate_pos, tlen;
    int has_mc;
    cinfo_t mate_cig;
} pe_t;

/* RecordSpec */
typedef struct {
    int r1, reg;                       /* reg -1: unmapped record */
    cinfo_t ci;
    int mapq, primary, supp, secondary;
    pe_t pe;
    char *sa, *xa;                     /* owned; NULL: no tag */
} spec_t;

typedef struct { spec_t *v; int n, cap; } spec_vec;

static void specs_free(spec_vec *sv)
{
    for (int i = 0; i < sv->n; ++i) { free(sv->v[i].sa); free(sv->v[i].xa); }
    free(sv->v);
    memset(sv, 0, sizeof *sv);
}

static char *sb_take(sbuf_t *b)
{
    if (!b->s) { sb_put(b, "", 0); }
    return b->s;
}

/* classifyAlignments (BWAMemAligner.swift:923-1099) for default flags: the
 * regions that pass minOutputScore and are not secondaries become the primary
 * and supplementaries; secondaries not below half their parent go to XA. */
static void classify(const or_e2e_t *e, const region_t *R, int nreg, int is_r1, const pe_t *pe,
                     spec_vec *out)
{
    int base = out->n, nonsec = 0;
    const region_t *xa_src[64];
    int nxa = 0, xa_over = 0;
    for (int i = 0; i < nreg; ++i) {
        const region_t *g = &R[i];
        if (g->score < MIN_OUT_SCORE) continue;
        if (g->secondary >= 0) {
            int pi = g->secondary;
            if (pi < nreg && g->score < R[pi].score / 2) continue;
            if (nxa < 64) xa_src[nxa++] = g; else xa_over = 1;
            continue;
        }
        spec_t s;
        memset(&s, 0, sizeof s);
        s.r1 = is_r1; s.reg = i; s.ci = cinfo_of(e, g); s.mapq = mapq_of(e, g);
        s.primary = nonsec == 0; s.supp = nonsec > 0;
        if (pe) s.pe = *pe;
        VEC_PUSH(*out, s);
        nonsec++;
    }
    if (nonsec == 0) {                                         /* unmapped record */
        spec_t s;
        memset(&s, 0, sizeof s);
        s.r1 = is_r1; s.reg = -1;
        if (pe) s.pe = *pe;
        VEC_PUSH(*out, s);
        return;
    }
    spec_t *S = out->v + base;
    for (int k = 1; k < nonsec; ++k)                           /* supplementary MAPQ cap */
        if (S[k].mapq > S[0].mapq) S[k].mapq = S[0].mapq;
    if (!xa_over && nxa <= MAX_XA_HITS) {                      /* buildXATag */
        sbuf_t xa = {0};
        for (int k = 0; k < nxa; ++k) {
            const region_t *g = xa_src[k];
            cinfo_t c = cinfo_of(e, g);
            sb_printf(&xa, "%s,%c%lld,", e->ctg_name[c.rid], c.rev ? '-' : '+',
                      (long long)(c.lpos + 1));
            cinfo_str(&xa, &c, 0);
            sb_printf(&xa, ",%d;", c.nm);
        }
        S[0].xa = sb_take(&xa);
    }
    if (nonsec > 1) {                                          /* buildSATag */
        for (int k = 0; k < nonsec; ++k) {
            sbuf_t sa = {0};
            for (int x = 0; x < nonsec; ++x) {
                if (x == k) continue;
                sb_printf(&sa, "%s,%lld,%c,", e->ctg_name[S[x].ci.rid], (long long)(S[x].ci.lpos + 1),
                          S[x].ci.rev ? '-' : '+');
                cinfo_str(&sa, &S[x].ci, 0);
                sb_printf(&sa, ",%d,%d;", S[x].mapq, S[x].ci.nm);
            }
            S[k].sa = sb_take(&sa);
        }
    }
}

/* buildRecord / buildUnmappedRecord (SAMOutputBuilder.swift:103-296, 370-438)
 * rendered as SAM text: RNAME/POS from decodePosition, RNEXT "=" when the
 * mate is on the same sequence (htslib's SAM text form). */
static void write_record(sbuf_t *b, const or_e2e_t *e, const char *name, const uint8_t *read,
                         const uint8_t *qual, int rl, const spec_t *s, const region_t *R)
{
    const pe_t *pe = s->pe.on ? &s->pe : NULL;
    int flag = 0;
    if (pe) {
        flag |= 0x1 | (pe->is_r1 ? 0x40 : 0x80);
        if (pe->mate_rev) flag |= 0x20;
        if (pe->mate_unmapped) flag |= 0x8;
    }
    const int mtid = pe && !pe->mate_unmapped ? pe->mate_tid : -1;
    const int64_t mpos = pe && !pe->mate_unmapped ? pe->mate_pos : -1;
    if (s->reg < 0) {
        flag |= 0x4;
        sb_printf(b, "%s\t%d\t*\t0\t0\t*\t%s\t%lld\t0\t", name, flag,
                  mtid < 0 ? "*" : e->ctg_name[mtid],
                  (long long)(mpos + 1));
        seq_qual(b, read, qual, rl, 0, 0, 0);
        sb_put(b, "\n", 1);
        return;
    }
    const region_t *g = &R[s->reg];
    const cinfo_t *c = &s->ci;
    const int rev = g->rb >= e->gl;
    if (pe && pe->proper) flag |= 0x2;
    if (rev) flag |= 0x10;
    if (s->supp) flag |= 0x800;
    else if (!s->primary) flag |= 0x100;
    int t0 = 0, t1 = 0;
    if (s->supp && !c->dummy && c->n) {                        /* convertToHardClip */
        if ((c->cig[0] & 0xF) == 4) t0 = (int)(c->cig[0] >> 4);
        if (c->n > 1 && (c->cig[c->n - 1] & 0xF) == 4) t1 = (int)(c->cig[c->n - 1] >> 4);
        if (!(t0 < rl - t1)) { t0 = 0; t1 = 0; }
    }
    sb_printf(b, "%s\t%d\t%s\t%lld\t%d\t", name, flag, e->ctg_name[c->rid], (long long)(c->lpos + 1),
              s->mapq);
    cinfo_str(b, c, s->supp);
    sb_printf(b, "\t%s\t%lld\t%lld\t",
              mtid < 0 ? "*" : mtid == c->rid ? "=" : e->ctg_name[mtid], (long long)(mpos + 1),
              (long long)(pe ? pe->tlen : 0));
    seq_qual(b, read, qual, rl, rev, t0, t1);
    int32_t xs = g->sub > g->csub ? g->sub : g->csub;
    sb_printf(b, "\tAS:i:%d\tXS:i:%d\tNM:i:%d\tMD:Z:%s", g->score, xs, c->nm, c->md);
    if (pe && pe->has_mc) {
        sb_put(b, "\tMC:Z:", 6);
        cinfo_str(b, &pe->mate_cig, 0);
    }
    if (s->sa) sb_printf(b, "\tSA:Z:%s", s->sa);
    /* buildXATag returns "" without secondaries, and buildRecord writes any
     * non-nil tag (SAMOutputBuilder.swift:262-264) */
    if (s->xa) sb_printf(b, "\tXA:Z:%s", s->xa);
    sb_put(b, "\n", 1);
}

static int64_t emit(sbuf_t *b, char *out, int64_t cap)
{
    int64_t n = b->n;
    if (out && cap > 0) {
        int64_t m = n < cap - 1 ? n : cap - 1;
        if (b->s) memcpy(out, b->s, (size_t)m);
        out[m] = 0;
    }
    free(b->s);
    return n;
}

int64_t or_e2e_sam(const or_e2e_t *e, const char *names, const uint8_t *quals, char *out,
                   int64_t cap)
{
    sbuf_t b = {0};
    const char *nm = names;
    for (int r = 0; r < e->n_reads; ++r) {
        const char *name = nm;
        nm += strlen(nm) + 1;
        const region_vec *rv = &e->regs[r];
        spec_vec sv = {0};
        classify(e, rv->v, rv->n, 1, NULL, &sv);
        for (int k = 0; k < sv.n; ++k)
            write_record(&b, e, name, e->reads + e->read_off[r], quals + e->read_off[r],
                         e->read_len[r], &sv.v[k], rv->v);
        specs_free(&sv);
    }
    return emit(&b, out, cap);
}

/* ---------------------------------------------------------------------- */
/* paired end: BWAMemAligner.alignPairedBatch / processFullBatch (CPU      */
/* path, :1348-1543, :1857-2128), InsertSizeEstimator, MateRescue,         */
/* PairedEndResolver, preparePairOutput (:1584-1856)                       */
/* ---------------------------------------------------------------------- */
enum { UNPAIRED_PENALTY = 17, MAX_MATESW = 50, PE_DIST_PAIRS = 4096 };

/* InsertSizeEstimator.inferOrientation; returns 0 when the rids differ */
static int infer_orientation(const or_e2e_t *e, const region_t *r1, const region_t *r2, int *ori,
                             int64_t *dist)
{
    if (r1->rid != r2->rid) return 0;
    const int64_t gl = e->gl;
    int same = (r1->rb >= gl) == (r2->rb >= gl);
    int64_t p2 = same ? r2->rb : (2 * gl - 1 - r2->rb);
    *dist = llabs(p2 - r1->rb);
    *ori = (same ? 0 : 1) ^ (p2 > r1->rb ? 0 : 3);
    return 1;
}

static int cmp_i64(const void *a, const void *b)
{
    int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return x < y ? -1 : x > y;
}

static const region_t *first_nonsec(const region_vec *rv)
{
    for (int i = 0; i < rv->n; ++i)
        if (rv->v[i].secondary < 0) return &rv->v[i];
    return NULL;
}

/* InsertSizeEstimator.estimate over the first min(4096, n) pairs: the first
 * sub-batch pair that alignPairedBatch receives (:1465-1472) */
static void estimate_dist(or_e2e_t *e)
{
    const int np = e->n_pairs < PE_DIST_PAIRS ? e->n_pairs : PE_DIST_PAIRS;
    int64_t *smp[4];
    int ns[4] = {0, 0, 0, 0};
    for (int d = 0; d < 4; ++d) smp[d] = (int64_t *)malloc(sizeof(int64_t) * (size_t)(np + 1));
    for (int i = 0; i < np; ++i) {
        const region_t *r1 = first_nonsec(&e->regs[i]), *r2 = first_nonsec(&e->regs[e->n_pairs + i]);
        if (!r1 || !r2) continue;
        if (mapq_of(e, r1) < 20 || mapq_of(e, r2) < 20) continue;   /* default scoring */
        int ori;
        int64_t d;
        if (!infer_orientation(e, r1, r2, &ori, &d)) continue;
        smp[ori][ns[ori]++] = d;
    }
    memset(e->dist, 0, sizeof e->dist);
    int max_count = 0, primary = 0;
    for (int d = 0; d < 4; ++d) {
        ori_stats_t *st = &e->dist[d];
        st->failed = 1;
        int n = ns[d];
        if (n < 10) continue;
        int64_t *sz = smp[d];
        qsort(sz, (size_t)n, sizeof(int64_t), cmp_i64);
        int64_t q25 = sz[n / 4], q75 = sz[(n * 3) / 4], iqr = q75 - q25;
        int64_t low = q25 - 2 * iqr, high = q75 + 2 * iqr;
        int nf = 0;
        double sum = 0.0;
        for (int k = 0; k < n; ++k)
            if (sz[k] >= low && sz[k] <= high) { sum += (double)sz[k]; nf++; }
        if (nf == 0) continue;
        double mean = sum / (double)nf, var = 0.0;
        for (int k = 0; k < n; ++k)
            if (sz[k] >= low && sz[k] <= high) var += ((double)sz[k] - mean) * ((double)sz[k] - mean);
        var /= (double)(nf - 1 > 1 ? nf - 1 : 1);
        double sd = sqrt(var);
        st->low = low; st->high = high; st->mean = mean; st->stddev = sd; st->count = nf;
        st->failed = 0;
        int64_t plow = q25 - 3 * iqr, phigh = q75 + 3 * iqr;
        int64_t a = (int64_t)(mean - 4.0 * sd + 0.5), b = (int64_t)(mean + 4.0 * sd + 0.5);
        st->proper_low = plow < a ? plow : a;
        if (st->proper_low < 1) st->proper_low = 1;
        st->proper_high = phigh > b ? phigh : b;
        if (nf > max_count) { max_count = nf; primary = d; }
    }
    e->primary_ori = primary;
    if (max_count > 0)
        for (int d = 0; d < 4; ++d)
            if (!e->dist[d].failed && (double)e->dist[d].count < (double)max_count * 0.05)
                e->dist[d].failed = 1;
    for (int d = 0; d < 4; ++d) free(smp[d]);
}

void or_e2e_pe_dist(const or_e2e_t *e, double *out8x4, int32_t *primary)
{
    for (int d = 0; d < 4; ++d) {
        const ori_stats_t *st = &e->dist[d];
        double *o = out8x4 + d * 8;
        o[0] = st->failed; o[1] = (double)st->low; o[2] = (double)st->high; o[3] = st->mean;
        o[4] = st->stddev; o[5] = (double)st->proper_low; o[6] = (double)st->proper_high;
        o[7] = st->count;
    }
    *primary = e->primary_ori;
}

int64_t or_e2e_pe_finish(or_e2e_t *e, int n_pairs)
{
    for (int64_t j = 0; j < e->n_jobs; ++j) { free(e->jobs[j].cigar); free(e->jobs[j].md); }
    e->n_jobs = 0; e->cq_n = e->cr_n = 0;
    if (2 * n_pairs != e->n_reads) return -1;
    e->n_pairs = n_pairs;
    for (int r = 0; r < e->n_reads; ++r) {                     /* phase 1: readId = gi << 1 | mate */
        uint64_t gi = (uint64_t)(r % n_pairs), mate = (uint64_t)(r / n_pairs);
        dedup_read(e, r, (gi << 1) | mate);
    }
    estimate_dist(e);
    e->do_rescue = !e->dist[e->primary_ori].failed;
    e->n_resc = 0;
    return e->do_rescue;
}

/* MateRescue.hasMateHit */
static int has_mate_hit(const or_e2e_t *e, const region_vec *mate, const region_t *a, int dir,
                        const ori_stats_t *st)
{
    const int64_t gl = e->gl;
    for (int i = 0; i < mate->n; ++i) {
        const region_t *m = &mate->v[i];
        if (m->rid != a->rid) continue;
        int same = (a->rb >= gl) == (m->rb >= gl);
        int64_t p2 = same ? m->rb : (2 * gl - 1 - m->rb);
        int d = (same ? 0 : 1) ^ (p2 > a->rb ? 0 : 3);
        if (d != dir) continue;
        int64_t dist = llabs(p2 - a->rb);
        if (dist >= st->proper_low && dist <= st->proper_high) return 1;
    }
    return 0;
}

/* BWAMemAligner.selectRescueCandidates (:2143-2159) */
static int rescue_templates(const region_vec *rv, const region_t **out)
{
    const region_t *best = first_nonsec(rv);
    if (!best) return 0;
    int32_t thr = best->score - UNPAIRED_PENALTY;
    int n = 0;
    for (int i = 0; i < rv->n; ++i) {
        if (!(rv->v[i].secondary < 0 && rv->v[i].score >= thr)) continue;
        out[n++] = &rv->v[i];
        if (n >= MAX_MATESW) break;
    }
    return n;
}

/* MateRescue.collectCandidates for pair i, direction dir (1: read 2 from
 * read 1's regions, 2: read 1 from read 2's), with PackedReference.
 * fetchSequence windows (PackedReference.swift:72-121). */
static void collect_rescue(or_e2e_t *e, int i, int dir)
{
    const int tr = dir == 1 ? i : e->n_pairs + i;              /* template read */
    const int mr = dir == 1 ? e->n_pairs + i : i;              /* mate read */
    const region_vec *tv = &e->regs[tr], *mv = &e->regs[mr];
    const region_t *tp[MAX_MATESW];
    const int nt = rescue_templates(tv, tp);
    const int64_t gl = e->gl;
    const int mlen = e->read_len[mr];
    const uint8_t *mread = e->reads + e->read_off[mr];
    for (int t = 0; t < nt; ++t) {
        const region_t *a = tp[t];
        int all_sat = 1;
        for (int r = 0; r < 4; ++r) {
            const ori_stats_t *st = &e->dist[r];
            if (st->failed) continue;
            if (has_mate_hit(e, mv, a, r, st)) continue;
            all_sat = 0;
            const int is_rev = (r >> 1) != (r & 1), larger = (r >> 1) == 0;
            int64_t rb, re;
            if (!is_rev) {
                if (larger) { rb = a->rb + st->proper_low; re = a->rb + st->proper_high + mlen; }
                else { rb = a->rb - st->proper_high - mlen; re = a->rb - st->proper_low; }
            } else {
                if (larger) { rb = a->rb + st->proper_low - mlen; re = a->rb + st->proper_high; }
                else { rb = a->rb - st->proper_high; re = a->rb - st->proper_low + mlen; }
            }
            if (rb < 0) rb = 0;
            if (re > 2 * gl) re = 2 * gl;
            if (!(rb < re)) continue;
            if (e->raw_resc) {                                 /* the caller fetches */
                e->resc = (rcand_t *)grow(e->resc, &e->resc_cap, e->n_resc + 1, sizeof(rcand_t));
                rcand_t *c = &e->resc[e->n_resc++];
                memset(c, 0, sizeof *c);
                c->pair = i; c->mate_read = mr; c->is_rev = is_rev; c->rid = a->rid;
                c->mate_len = mlen; c->raw_rb = rb; c->raw_re = re; c->clamped = -1;
                e->rq = (uint8_t *)grow(e->rq, &e->rq_cap, e->rq_n + mlen + 1, 1);
                c->q_off = e->rq_n; c->q_len = mlen;
                for (int k = 0; k < mlen; ++k) {
                    uint8_t x = is_rev ? mread[mlen - 1 - k] : mread[k];
                    e->rq[e->rq_n + k] = is_rev ? (x < 4 ? (uint8_t)(3 - x) : 4) : x;
                }
                e->rq_n += mlen;
                c->t_off = e->rt_n; c->t_len = 0;
                continue;
            }
            /* PackedReference.fetchSequence (PackedReference.swift:72-121):
               clamp to the sequence holding the window's start, on its strand */
            const int fwd = rb < gl;
            const int64_t fb = fwd ? rb : 2 * gl - re, fe = fwd ? re : 2 * gl - rb;
            const int rid = seq_id(e, fb);
            const int64_t lo_c = e->ctg_off[rid], hi_c = lo_c + e->ctg_len[rid];
            const int64_t cb = fb > lo_c ? fb : lo_c, ce = fe < hi_c ? fe : hi_c;
            if (!(cb < ce)) continue;
            if (rid != a->rid) continue;                       /* MateRescue.swift:111, :256 */
            const int64_t clamped = fwd ? cb : 2 * gl - ce;
            e->resc = (rcand_t *)grow(e->resc, &e->resc_cap, e->n_resc + 1, sizeof(rcand_t));
            rcand_t *c = &e->resc[e->n_resc++];
            memset(c, 0, sizeof *c);
            c->pair = i; c->mate_read = mr; c->is_rev = is_rev; c->rid = a->rid;
            c->mate_len = mlen; c->clamped = clamped;
            e->rq = (uint8_t *)grow(e->rq, &e->rq_cap, e->rq_n + mlen + 1, 1);
            e->rt = (uint8_t *)grow(e->rt, &e->rt_cap, e->rt_n + (ce - cb) + 1, 1);
            c->q_off = e->rq_n; c->q_len = mlen;
            for (int k = 0; k < mlen; ++k) {
                uint8_t x = is_rev ? mread[mlen - 1 - k] : mread[k];
                e->rq[e->rq_n + k] = is_rev ? (x < 4 ? (uint8_t)(3 - x) : 4) : x;
            }
            e->rq_n += mlen;
            c->t_off = e->rt_n; c->t_len = (int32_t)(ce - cb);
            for (int64_t k = 0; k < ce - cb; ++k)
                e->rt[e->rt_n + k] = fwd ? e->fwd[cb + k] : (uint8_t)(3 - e->fwd[ce - 1 - k]);
            e->rt_n += ce - cb;
        }
        if (all_sat) break;
    }
}

int64_t or_e2e_pe_rescue_jobs(or_e2e_t *e, int dir)
{
    e->n_resc = 0; e->rq_n = e->rt_n = 0;
    e->raw_resc = 0;
    if (!e->do_rescue) return 0;
    for (int i = 0; i < e->n_pairs; ++i) collect_rescue(e, i, dir);
    return e->n_resc;
}

/* The same candidates with their windows before fetchSequence: BWT-space
 * [rb, re) and the template's sequence id, for a caller that clamps and
 * fetches them itself; the target arena stays empty. */
int64_t or_e2e_pe_rescue_jobs_raw(or_e2e_t *e, int dir)
{
    e->n_resc = 0; e->rq_n = e->rt_n = 0;
    e->raw_resc = 1;
    if (e->do_rescue)
        for (int i = 0; i < e->n_pairs; ++i) collect_rescue(e, i, dir);
    e->raw_resc = 0;
    return e->n_resc;
}

void or_e2e_pe_get_rescue_windows(const or_e2e_t *e, int64_t *rb, int64_t *re, int32_t *rid)
{
    for (int64_t i = 0; i < e->n_resc; ++i) {
        rb[i] = e->resc[i].raw_rb; re[i] = e->resc[i].raw_re; rid[i] = e->resc[i].rid;
    }
}

/* fetchSequence's clampedBegin of each raw candidate (ignored where ok = 0). */
void or_e2e_pe_set_rescue_windows(or_e2e_t *e, const int64_t *clamped)
{
    for (int64_t i = 0; i < e->n_resc; ++i) e->resc[i].clamped = clamped[i];
}

void or_e2e_pe_rescue_sizes(const or_e2e_t *e, int64_t *q_bytes, int64_t *t_bytes)
{
    *q_bytes = e->rq_n;
    *t_bytes = e->rt_n;
}

void or_e2e_pe_get_rescue_jobs(const or_e2e_t *e, uint8_t *q, int64_t *q_off, int32_t *q_len,
                               uint8_t *t, int64_t *t_off, int32_t *t_len)
{
    if (e->rq_n) memcpy(q, e->rq, (size_t)e->rq_n);
    if (e->rt_n) memcpy(t, e->rt, (size_t)e->rt_n);
    for (int64_t i = 0; i < e->n_resc; ++i) {
        q_off[i] = e->resc[i].q_off; q_len[i] = e->resc[i].q_len;
        t_off[i] = e->resc[i].t_off; t_len[i] = e->resc[i].t_len;
    }
}

void or_e2e_pe_set_rescue(or_e2e_t *e, const int32_t *out6, const int32_t *ok)
{
    for (int64_t i = 0; i < e->n_resc; ++i) {
        memcpy(e->resc[i].out6, out6 + i * 6, sizeof(int32_t) * 6);
        e->resc[i].ok = ok[i];
    }
}

/* LocalSWAligner.align of every pending candidate (minSubScore =
 * minSeedLength * matchScore clamped to u8, MateRescue.swift:125) */
void or_e2e_pe_rescue_cpu(or_e2e_t *e)
{
    int ms = e->sc.min_seed_len * e->sc.a;
    ms = ms < 0 ? 0 : ms > 255 ? 255 : ms;
    for (int64_t i = 0; i < e->n_resc; ++i) {
        rcand_t *c = &e->resc[i];
        int64_t cells = 0;
        c->ok = or_local_sw(e->rq + c->q_off, c->q_len, e->rt + c->t_off, c->t_len, &e->sc, ms,
                            c->out6, &cells);
    }
}

/* MateRescue.applyResult, appended in candidate order (the CPU path's
 * rescue() appends per template and direction in the same order) */
int64_t or_e2e_pe_apply(or_e2e_t *e)
{
    const int32_t min_score = e->sc.min_seed_len * e->sc.a;
    int64_t added = 0;
    for (int64_t i = 0; i < e->n_resc; ++i) {
        const rcand_t *c = &e->resc[i];
        if (!c->ok) continue;
        const int32_t score = c->out6[0], qb = c->out6[1], qe = c->out6[2], tb = c->out6[3],
                      te = c->out6[4], sc2 = c->out6[5];
        if (score < min_score) continue;
        region_t g;
        memset(&g, 0, sizeof g);
        g.score = score; g.true_score = score; g.csub = sc2; g.rid = c->rid; g.secondary = -1;
        g.cig = -1;
        if (!c->is_rev) {
            g.rb = c->clamped + tb; g.re = c->clamped + te + 1; g.qb = qb; g.qe = qe + 1;
        } else {
            g.rb = 2 * e->gl - (c->clamped + te + 1); g.re = 2 * e->gl - (c->clamped + tb);
            g.qb = c->mate_len - (qe + 1); g.qe = c->mate_len - qb;
        }
        int64_t rs = (int64_t)te - tb + 1, qs = (int64_t)(qe - qb + 1);
        g.seed_cov = (int32_t)((rs < qs ? rs : qs) / 2);
        VEC_PUSH(e->regs[c->mate_read], g);
        added++;
    }
    return added;
}

/* re-marking after rescue (processFullBatch's CPU path, rescuePair :2104-2123) */
void or_e2e_pe_remark(or_e2e_t *e)
{
    if (!e->do_rescue) return;
    for (int r = 0; r < e->n_reads; ++r) {
        uint64_t gi = (uint64_t)(r % e->n_pairs), mate = (uint64_t)(r / e->n_pairs);
        if (e->regs[r].n > 1) mark_secondary(e, &e->regs[r], (gi << 1) | mate);
    }
}

int64_t or_e2e_pe_cigar_plan(or_e2e_t *e)
{
    for (int64_t j = 0; j < e->n_jobs; ++j) { free(e->jobs[j].cigar); free(e->jobs[j].md); }
    e->n_jobs = 0; e->cq_n = e->cr_n = 0;
    for (int r = 0; r < e->n_reads; ++r) cigar_plan_read(e, r);
    return e->n_jobs;
}

/* PairedEndResolver.computeZPenalty */
static int z_penalty(double z, int32_t a)
{
    double ec = erfc(z / sqrt(2.0));
    if (!(ec > 0)) return (int)a * 10;
    double pen = 0.721 * log(2.0 * ec) * (double)a + 0.499;
    int v = (int)floor(-pen);
    return v > 0 ? v : 0;
}

typedef struct { int score, idx1, idx2, proper; int64_t ins; int ori; } phit_t;

static int phit_before(const void *a, const void *b, void *ctx)
{
    return ((const phit_t *)a)->score > ((const phit_t *)b)->score;
}

typedef struct { int idx1, idx2, score, subo, n_sub, proper; } pdec_t;

/* PairedEndResolver.resolve; returns 0 for nil */
static int resolve_pair(const or_e2e_t *e, const region_vec *v1, const region_vec *v2, pdec_t *d)
{
    if (v1->n == 0 || v2->n == 0) return 0;
    int c1[4096], c2[4096], n1 = 0, n2 = 0;
    const region_vec *vs[2] = {v1, v2};
    int *cs[2] = {c1, c2}, *ns[2] = {&n1, &n2};
    for (int s = 0; s < 2; ++s) {                              /* topCandidates */
        const region_t *best = first_nonsec(vs[s]);
        for (int i = 0; i < vs[s]->n && *ns[s] < 4096; ++i)
            if (!best || vs[s]->v[i].score >= best->score - UNPAIRED_PENALTY) cs[s][(*ns[s])++] = i;
    }
    phit_t *hits = (phit_t *)malloc(sizeof(phit_t) * (size_t)(n1 * n2 + 1));
    int nh = 0;
    for (int a = 0; a < n1; ++a)
        for (int b = 0; b < n2; ++b) {
            const region_t *r1 = &v1->v[c1[a]], *r2 = &v2->v[c2[b]];
            int ori;
            int64_t ins;
            if (!infer_orientation(e, r1, r2, &ori, &ins)) continue;
            const ori_stats_t *st = &e->dist[ori];
            if (st->failed || !(st->stddev > 0) || ins < st->proper_low || ins > st->proper_high)
                continue;
            double z = fabs((double)ins - st->mean) / st->stddev;
            phit_t h = {r1->score + r2->score - z_penalty(z, e->sc.a), c1[a], c2[b],
                        ori == e->primary_ori, ins, ori};
            hits[nh++] = h;
        }
    if (nh == 0) { free(hits); return 0; }
    stable_sort(hits, (size_t)nh, sizeof(phit_t), phit_before, NULL);
    d->idx1 = hits[0].idx1; d->idx2 = hits[0].idx2; d->score = hits[0].score;
    d->proper = hits[0].proper;
    d->subo = nh > 1 ? hits[1].score : 0;
    const int tmp = e->sc.a + e->sc.b;
    d->n_sub = 0;
    for (int i = 1; i < nh; ++i)
        if (d->subo - hits[i].score <= tmp) d->n_sub++;
    free(hits);
    return 1;
}

/* PairedEndResolver.computeTLEN */
static void compute_tlen(int64_t p1, int rev1, int64_t l1, int64_t p2, int rev2, int64_t l2,
                         int64_t *t1, int64_t *t2)
{
    (void)rev2;
    int64_t e1 = p1 + l1, e2 = p2 + l2;
    int64_t lo = p1 < p2 ? p1 : p2, hi = e1 > e2 ? e1 : e2, span = hi - lo;
    if (p1 < p2 || (p1 == p2 && !rev1)) { *t1 = span; *t2 = -span; }
    else { *t1 = -span; *t2 = span; }
}

/* PairedEndResolver.adjustMAPQ */
static void adjust_mapq(const or_e2e_t *e, int *q1p, int *q2p, const region_t *g1,
                        const region_t *g2, const pdec_t *d)
{
    const int o = d->score;
    int subo = d->subo;
    const int score_un = g1->score + g2->score - UNPAIRED_PENALTY;
    if (subo < score_un) subo = score_un;
    if (*q1p == 0 && *q2p == 0) {
        if (g1->sub > 0) { int s = o - (g1->score - g1->sub); if (s > subo) subo = s; }
        if (g2->sub > 0) { int s = o - (g2->score - g2->sub); if (s > subo) subo = s; }
    }
    const double a = e->sc.a;
    int qpe = (int)(6.02 * (double)(o - subo) / a + 0.499);
    if (d->n_sub > 0) qpe -= (int)(4.343 * log((double)(d->n_sub + 1)) + 0.499);
    if (qpe < 0) qpe = 0;
    if (qpe > 60) qpe = 60;
    qpe = (int)((double)qpe * (1.0 - 0.5 * ((double)g1->frac_rep + (double)g2->frac_rep)) + 0.499);
    int q1 = *q1p, q2 = *q2p;
    if (o > score_un) {
        q1 = q1 > qpe ? q1 : (qpe < q1 + 40 ? qpe : q1 + 40);
        q2 = q2 > qpe ? q2 : (qpe < q2 + 40 ? qpe : q2 + 40);
        if (g1->csub > 0) {
            int cap = (int)(6.02 * (double)(g1->score - g1->csub) / a + 0.499);
            if (q1 > cap) q1 = cap;
        }
        if (g2->csub > 0) {
            int cap = (int)(6.02 * (double)(g2->score - g2->csub) / a + 0.499);
            if (q2 > cap) q2 = cap;
        }
    }
    *q1p = q1 < 0 ? 0 : q1 > 60 ? 60 : q1;
    *q2p = q2 < 0 ? 0 : q2 > 60 ? 60 : q2;
}

/* BWAMemAligner.promotePairedRegion (:1549-1583) */
static void promote(region_t *R, int n, int idx)
{
    if (idx < 0 || idx >= n) return;
    if (R[idx].secondary >= 0 && R[idx].secondary < n) R[idx].sub = R[R[idx].secondary].score;
    R[idx].secondary = -1;
    if (idx > 0) {
        region_t t = R[0]; R[0] = R[idx]; R[idx] = t;
        for (int i = 0; i < n; ++i) {
            if (R[i].secondary == idx) R[i].secondary = 0;
            else if (R[i].secondary == 0) R[i].secondary = idx;
        }
    }
    R[0].secondary = -1;
}

static int is_multi(const region_vec *rv)
{
    int c = 0;
    for (int i = 0; i < rv->n; ++i)
        if (rv->v[i].secondary < 0 && rv->v[i].score >= MIN_OUT_SCORE && ++c > 1) return 1;
    return 0;
}

static void pe_pair(pe_t *pe, int is_r1, int proper, const cinfo_t *mate, int64_t tlen)
{
    memset(pe, 0, sizeof *pe);
    pe->on = 1; pe->is_r1 = is_r1; pe->proper = proper;
    pe->mate_tid = mate->rid; pe->mate_pos = mate->lpos; pe->mate_rev = mate->rev;
    pe->mate_unmapped = 0;
    pe->tlen = tlen; pe->has_mc = 1; pe->mate_cig = *mate;
}

/* preparePairOutput (:1584-1799) and prepareMappedUnmappedPlan (:1800-1856)
 * of pair i, then the records in plan order */
static void pair_records(or_e2e_t *e, int i, const char *name, const uint8_t *quals, sbuf_t *b)
{
    const int r1 = i, r2 = e->n_pairs + i;
    region_vec *v1 = &e->regs[r1], *v2 = &e->regs[r2];
    spec_vec sv = {0};
    region_t *R1 = v1->v, *R2 = v2->v;
    region_t *P1 = NULL, *P2 = NULL;                           /* paired copies */
    const int e1 = v1->n == 0, e2 = v2->n == 0;
    if (e1 && e2) {
        for (int s = 0; s < 2; ++s) {
            spec_t sp;
            memset(&sp, 0, sizeof sp);
            sp.r1 = s == 0; sp.reg = -1;
            sp.pe.on = 1; sp.pe.is_r1 = s == 0; sp.pe.mate_tid = -1; sp.pe.mate_pos = -1;
            sp.pe.mate_unmapped = 1;
            VEC_PUSH(sv, sp);
        }
    } else {
        pdec_t d;
        const int has_dec = resolve_pair(e, v1, v2, &d);
        if (has_dec && !is_multi(v1) && !is_multi(v2)) {
            P1 = (region_t *)malloc(sizeof(region_t) * (size_t)v1->n);
            P2 = (region_t *)malloc(sizeof(region_t) * (size_t)v2->n);
            memcpy(P1, v1->v, sizeof(region_t) * (size_t)v1->n);
            memcpy(P2, v2->v, sizeof(region_t) * (size_t)v2->n);
            promote(P1, v1->n, d.idx1);
            promote(P2, v2->n, d.idx2);
            for (int k = 1; k < v1->n; ++k) if (P1[k].secondary < 0) P1[k].secondary = 0;
            for (int k = 1; k < v2->n; ++k) if (P2[k].secondary < 0) P2[k].secondary = 0;
            /* a promoted region whose CIGAR was skipped gets one now (CPU) */
            region_t *pp[2] = {&P1[0], &P2[0]};
            const int rr[2] = {r1, r2};
            for (int s = 0; s < 2; ++s) {
                if (pp[s]->cig >= 0) continue;
                int64_t jx = add_cigar_job(e, rr[s], pp[s]);
                cjob_t *j = &e->jobs[jx];
                uint32_t cig[4096];
                char md[8192];
                or_cigar_res_t res;
                int n = or_cigar_generate(e->cq + j->q_off, j->q_len, e->cr + j->r_off, j->r_len,
                                          j->qb, j->qe, j->read_len, j->rev, j->true_score, j->w,
                                          &e->sc, j->ref_pos, cig, 4096, md, 8192, &res);
                or_e2e_set_cigar(e, jx, cig, n < 4096 ? n : 4096, res.nm, md, res.pos);
                pp[s]->cig = jx;
            }
            cinfo_t c1 = cinfo_of(e, &P1[0]), c2 = cinfo_of(e, &P2[0]);
            int64_t t1, t2;
            compute_tlen(c1.lpos, c1.rev, c1.ref_len, c2.lpos, c2.rev, c2.ref_len, &t1, &t2);
            pe_t pe1, pe2;
            pe_pair(&pe1, 1, d.proper, &c2, t1);
            pe_pair(&pe2, 0, d.proper, &c1, t2);
            classify(e, P1, v1->n, 1, &pe1, &sv);
            int n_r1 = sv.n;
            classify(e, P2, v2->n, 0, &pe2, &sv);
            int i1 = -1, i2 = -1;
            for (int k = 0; k < sv.n; ++k) {
                if (sv.v[k].primary && sv.v[k].r1 && i1 < 0) i1 = k;
                if (sv.v[k].primary && !sv.v[k].r1 && i2 < 0) i2 = k;
            }
            (void)n_r1;
            if (i1 >= 0 && i2 >= 0) {
                int q1 = sv.v[i1].mapq, q2 = sv.v[i2].mapq;
                adjust_mapq(e, &q1, &q2, &P1[0], &P2[0], &d);
                sv.v[i1].mapq = q1; sv.v[i2].mapq = q2;
                for (int k = 0; k < sv.n; ++k)
                    if (sv.v[k].supp) {
                        int cap = sv.v[k].r1 ? q1 : q2;
                        if (sv.v[k].mapq > cap) sv.v[k].mapq = cap;
                    }
            }
            R1 = P1; R2 = P2;
        } else if (!e1 && e2) {
            goto mapped_unmapped;
        } else if (e1 && !e2) {
            goto mapped_unmapped;
        } else {
            /* both mapped, no pairing: default primaries */
            cinfo_t c1 = cinfo_of(e, &v1->v[0]), c2 = cinfo_of(e, &v2->v[0]);
            int proper = 0;
            int64_t t1 = 0, t2 = 0;
            int ori;
            int64_t ins;
            if (c1.rid == c2.rid) {                            /* :1752-1771 */
                if (infer_orientation(e, &v1->v[0], &v2->v[0], &ori, &ins)) {
                    const ori_stats_t *st = &e->dist[ori];
                    if (!st->failed && ins >= st->proper_low && ins <= st->proper_high) proper = 1;
                }
                compute_tlen(c1.lpos, c1.rev, c1.ref_len, c2.lpos, c2.rev, c2.ref_len, &t1, &t2);
            }
            pe_t pe1, pe2;
            pe_pair(&pe1, 1, proper, &c2, t1);
            pe_pair(&pe2, 0, proper, &c1, t2);
            classify(e, v1->v, v1->n, 1, &pe1, &sv);
            classify(e, v2->v, v2->n, 0, &pe2, &sv);
        }
    }
    goto write;
mapped_unmapped: {
        const int m_is_r1 = !e1;
        region_vec *mv = m_is_r1 ? v1 : v2;
        const int mr = m_is_r1 ? r1 : r2;
        region_t *g = &mv->v[0];
        if (g->cig < 0) {                                      /* cigarCache.first ?? generateCIGAR */
            int64_t jx = add_cigar_job(e, mr, g);
            cjob_t *j = &e->jobs[jx];
            uint32_t cig[4096];
            char md[8192];
            or_cigar_res_t res;
            int n = or_cigar_generate(e->cq + j->q_off, j->q_len, e->cr + j->r_off, j->r_len, j->qb,
                                      j->qe, j->read_len, j->rev, j->true_score, j->w, &e->sc,
                                      j->ref_pos, cig, 4096, md, 8192, &res);
            or_e2e_set_cigar(e, jx, cig, n < 4096 ? n : 4096, res.nm, md, res.pos);
            g->cig = jx;
        }
        cinfo_t c = cinfo_of(e, g);
        spec_t ms, us;
        memset(&ms, 0, sizeof ms);
        memset(&us, 0, sizeof us);
        ms.r1 = m_is_r1; ms.reg = 0; ms.ci = c; ms.mapq = mapq_of(e, g); ms.primary = 1;
        ms.pe.on = 1; ms.pe.is_r1 = m_is_r1; ms.pe.mate_tid = c.rid; ms.pe.mate_pos = c.lpos;
        ms.pe.mate_unmapped = 1;
        us.r1 = !m_is_r1; us.reg = -1;
        us.pe.on = 1; us.pe.is_r1 = !m_is_r1; us.pe.mate_tid = c.rid; us.pe.mate_pos = c.lpos;
        us.pe.mate_rev = c.rev; us.pe.mate_unmapped = 0;
        if (m_is_r1) { VEC_PUSH(sv, ms); VEC_PUSH(sv, us); }
        else { VEC_PUSH(sv, us); VEC_PUSH(sv, ms); }
        /* the plan keeps the mapped read's regions only */
        if (m_is_r1) R1 = v1->v; else R2 = v2->v;
    }
write:
    for (int k = 0; k < sv.n; ++k) {
        const spec_t *s = &sv.v[k];
        const int rd = s->r1 ? r1 : r2;
        write_record(b, e, name, e->reads + e->read_off[rd], quals + e->read_off[rd],
                     e->read_len[rd], s, s->r1 ? R1 : R2);
    }
    specs_free(&sv);
    free(P1);
    free(P2);
}

int64_t or_e2e_pe_sam(or_e2e_t *e, const char *names, const uint8_t *quals, char *out, int64_t cap)
{
    sbuf_t b = {0};
    const char *nm = names;
    for (int i = 0; i < e->n_pairs; ++i) {
        const char *name = nm;
        nm += strlen(nm) + 1;
        pair_records(e, i, name, quals, &b);
    }
    return emit(&b, out, cap);
}
