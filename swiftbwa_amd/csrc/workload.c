/*
 * Deterministic synthetic workloads for the Smith-Waterman hot path.
 *
 * Random numbers: splitmix64 seeds xoshiro256** (Blackman & Vigna).  Every
 * job set is a pure function of (kind, seed, n), so the GPU host regenerates
 * the same inputs the tests were written against.
 *
 * Mutation model (shared by every kind): per source base, a substitution with
 * probability p_sub picks one of the other three bases uniformly; an indel
 * event with probability p_indel is an insertion or a deletion with
 * probability 1/2 each, of geometric length (p = 1/2, mean 2).
 *
 * Kinds:
 *   WL_EXT_UNIFORM  banded extension, uniform: query 150 bp; target = the
 *                   query's source window plus a 0..100 bp flank, mutated at
 *                   1% substitutions / 0.1% indels; h0 ~ U[19,60]; w = 100;
 *                   end bonus 5.
 *   WL_NW           banded global NW: qlen ~ U[50,250]; target = query
 *                   mutated at 1% / 0.2%, plus one planted indel block of
 *                   0..20 bp; w = min(100, 2|dlen| + 3); true_score = score
 *                   of the planted alignment.
 *   WL_EXT_MIXED    uneven extension: half left jobs (q ~ U[0,150]), half
 *                   right jobs (q ~ U[0,250]); w ~ U{20,50,100}; tlen = q +
 *                   min(calMaxGap(q), 2w); 5% of jobs at 10% substitutions;
 *                   2% with h0 ~ U[61,250], else U[19,60]; end bonus 5.
 *   WL_LOCAL        mate-rescue local SW: 150-bp mate (either strand) against
 *                   a 600-bp window holding its true locus (0.5% / 0.05%).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { WL_EXT_UNIFORM = 2, WL_NW = 3, WL_EXT_MIXED = 4, WL_LOCAL = 5 };

typedef struct { uint64_t s[4]; } rng_t;

static uint64_t splitmix64(uint64_t *x)
{
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static void rng_seed(rng_t *r, uint64_t seed)
{
    uint64_t x = seed;
    for (int k = 0; k < 4; ++k) r->s[k] = splitmix64(&x);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static uint64_t next_u64(rng_t *r)
{
    uint64_t *s = r->s;
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

/* uniform in [lo, hi] (inclusive) */
static int64_t uni(rng_t *r, int64_t lo, int64_t hi)
{
    uint64_t span = (uint64_t)(hi - lo + 1);
    return lo + (int64_t)(next_u64(r) % span);
}

static double unif01(rng_t *r) { return (double)(next_u64(r) >> 11) * (1.0 / 9007199254740992.0); }

static int geom_len(rng_t *r)
{
    int len = 1;
    while (next_u64(r) & 1) len++;
    return len;
}

static void random_seq(rng_t *r, uint8_t *dst, int n)
{
    for (int k = 0; k < n; ++k) dst[k] = (uint8_t)(next_u64(r) & 3);
}

/* ins_*: bases present only in the mutated copy (D runs of the alignment);
 * del_*: source bases dropped from the copy (I runs). */
typedef struct { int subs, ins_events, ins_bases, del_events, del_bases; } mut_stats_t;

/* Mutates src[0..n) into dst (capacity cap).  Returns the output length. */
static int mutate(rng_t *r, const uint8_t *src, int n, double p_sub, double p_indel,
                  uint8_t *dst, int cap, mut_stats_t *st)
{
    int o = 0;
    for (int k = 0; k < n && o < cap; ++k) {
        double u = unif01(r);
        if (u < p_indel) {
            int len = geom_len(r);
            if (next_u64(r) & 1) {               /* insertion before src[k] */
                for (int x = 0; x < len && o < cap; ++x) dst[o++] = (uint8_t)(next_u64(r) & 3);
                if (st) { st->ins_events++; st->ins_bases += len; }
                if (o < cap) dst[o++] = src[k];
            } else {                             /* deletion of src[k..k+len) */
                if (st) { st->del_events++; st->del_bases += (k + len <= n ? len : n - k); }
                k += len - 1;
            }
        } else if (u < p_indel + p_sub) {
            uint8_t b = (uint8_t)((src[k] + 1 + uni(r, 0, 2)) & 3);
            dst[o++] = b;
            if (st) st->subs++;
        } else {
            dst[o++] = src[k];
        }
    }
    return o;
}

typedef struct {
    int kind;
    int64_t n;
    uint8_t *q, *t;                 /* byte arenas (one base per byte) */
    int64_t q_bytes, t_bytes;
    uint32_t *q_off, *t_off;
    uint16_t *q_len, *t_len;
    int32_t *h0, *w, *eb, *true_score, *dir;
} wl_set_t;

static int cal_max_gap(int qlen, int band)
{
    /* ExtensionAligner.calMaxGap with A=1, O=6, E=1 (ins and del) */
    int l = (qlen * 1 - 6) / 1 + 1;
    if (l < 1) l = 1;
    return l < 2 * band ? l : 2 * band;
}

wl_set_t *wl_generate(int kind, uint64_t seed, int64_t n)
{
    if (n < 0 || n > 100000000) return NULL;
    int qcap, tcap;
    switch (kind) {
    case WL_EXT_UNIFORM: qcap = 150; tcap = 400; break;
    case WL_NW: qcap = 250; tcap = 320; break;
    case WL_EXT_MIXED: qcap = 250; tcap = 460; break;
    case WL_LOCAL: qcap = 150; tcap = 600; break;
    default: return NULL;
    }
    wl_set_t *s = (wl_set_t *)calloc(1, sizeof *s);
    s->kind = kind;
    s->n = n;
    s->q = (uint8_t *)malloc((size_t)n * qcap + 1);
    s->t = (uint8_t *)malloc((size_t)n * tcap + 1);
    s->q_off = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n + 1));
    s->t_off = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n + 1));
    s->q_len = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)(n + 1));
    s->t_len = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)(n + 1));
    s->h0 = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
    s->w = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
    s->eb = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
    s->true_score = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
    s->dir = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
    rng_t r;
    rng_seed(&r, seed * 0x100000001B3ull + (uint64_t)kind);
    uint8_t src[1024], tmp[1024];
    int64_t qo = 0, to = 0;
    for (int64_t i = 0; i < n; ++i) {
        int ql = 0, tl = 0;
        uint8_t *qd = s->q + qo, *td = s->t + to;
        if (kind == WL_EXT_UNIFORM) {
            int flank = (int)uni(&r, 0, 100);
            random_seq(&r, src, 150 + flank);
            memcpy(qd, src, 150);
            ql = 150;
            tl = mutate(&r, src, 150 + flank, 0.01, 0.001, td, tcap, NULL);
            s->h0[i] = (int32_t)uni(&r, 19, 60);
            s->w[i] = 100;
            s->eb[i] = 5;
        } else if (kind == WL_NW) {
            ql = (int)uni(&r, 50, 250);
            random_seq(&r, qd, ql);
            mut_stats_t st = {0};
            int m = mutate(&r, qd, ql, 0.01, 0.002, tmp, 300, &st);
            int blk = (int)uni(&r, 0, 20);
            int pos = (int)uni(&r, 0, m);
            int gap_open = 6, gap_ext = 1;
            if (next_u64(&r) & 1) {              /* planted insertion into the target */
                memcpy(td, tmp, (size_t)pos);
                random_seq(&r, td + pos, blk);
                memcpy(td + pos + blk, tmp + pos, (size_t)(m - pos));
                tl = m + blk;
                if (blk) { st.ins_events++; st.ins_bases += blk; }
            } else {                             /* planted deletion from the target */
                if (blk > m - 1) blk = m - 1 > 0 ? m - 1 : 0;
                if (pos + blk > m) pos = m - blk;
                memcpy(td, tmp, (size_t)pos);
                memcpy(td + pos, tmp + pos + blk, (size_t)(m - pos - blk));
                tl = m - blk;
                if (blk) { st.del_events++; st.del_bases += blk; }
            }
            if (tl < 1) { td[0] = 0; tl = 1; }
            int dl = ql - tl;
            if (dl < 0) dl = -dl;
            int w = 2 * dl + 3;
            s->w[i] = w < 100 ? w : 100;
            /* score of the planted alignment: matched bases - 4 per substitution
             * - (6 + len) per gap event (A=1, B=4, O=6, E=1). */
            int aligned = ql - st.del_bases;   /* query bases not in an I run */
            s->true_score[i] = (aligned - st.subs) * 1 - st.subs * 4
                               - (st.ins_events + st.del_events) * gap_open
                               - (st.ins_bases + st.del_bases) * gap_ext;
        } else if (kind == WL_EXT_MIXED) {
            int left = (int)(i & 1) == 0;
            int q = left ? (int)uni(&r, 0, 150) : (int)uni(&r, 0, 250);
            static const int wset[3] = {20, 50, 100};
            int w = wset[uni(&r, 0, 2)];
            int gap = cal_max_gap(q, 100);
            if (gap > 2 * w) gap = 2 * w;
            int tlen = q + gap;
            random_seq(&r, src, tlen);
            memcpy(td, src, (size_t)tlen);
            tl = tlen;
            double u = unif01(&r);
            double p_sub = u < 0.05 ? 0.10 : 0.005;
            int m = mutate(&r, src, q, p_sub, 0.0005, tmp, 300, NULL);
            ql = m < q ? m : q;
            memcpy(qd, tmp, (size_t)ql);
            double v = unif01(&r);
            s->h0[i] = v < 0.02 ? (int32_t)uni(&r, 61, 250) : (int32_t)uni(&r, 19, 60);
            s->w[i] = w;
            s->eb[i] = 5;
            s->dir[i] = left ? 0 : 1;
        } else {                                 /* WL_LOCAL */
            random_seq(&r, src, 600);
            memcpy(td, src, 600);
            tl = 600;
            int pos = (int)uni(&r, 100, 350);
            int m = mutate(&r, src + pos, 150, 0.005, 0.0005, tmp, 300, NULL);
            ql = m < 150 ? m : 150;
            if (next_u64(&r) & 1) {
                for (int k = 0; k < ql; ++k) qd[k] = (uint8_t)(3 - tmp[ql - 1 - k]);
                s->dir[i] = 1;
            } else {
                memcpy(qd, tmp, (size_t)ql);
            }
        }
        s->q_off[i] = (uint32_t)qo;
        s->t_off[i] = (uint32_t)to;
        s->q_len[i] = (uint16_t)ql;
        s->t_len[i] = (uint16_t)tl;
        qo += ql;
        to += tl;
    }
    s->q_bytes = qo;
    s->t_bytes = to;
    return s;
}

/* Field accessors for ctypes callers. */
void *wl_field(wl_set_t *s, int which)
{
    switch (which) {
    case 0: return s->q;
    case 1: return s->q_off;
    case 2: return s->q_len;
    case 3: return s->t;
    case 4: return s->t_off;
    case 5: return s->t_len;
    case 6: return s->h0;
    case 7: return s->w;
    case 8: return s->eb;
    case 9: return s->true_score;
    case 10: return s->dir;
    default: return NULL;
    }
}

int64_t wl_count(wl_set_t *s, int which)
{
    if (which == 0) return s->q_bytes;
    if (which == 3) return s->t_bytes;
    return s->n;
}

void wl_free(wl_set_t *s)
{
    if (!s) return;
    free(s->q); free(s->t); free(s->q_off); free(s->t_off); free(s->q_len); free(s->t_len);
    free(s->h0); free(s->w); free(s->eb); free(s->true_score); free(s->dir);
    free(s);
}

/* ------------------------------------------------------------------------
 * Serving workload: a random reference with planted repeat families, and
 * chunks of read pairs turned into the hot-path jobs of each pair.
 * ---------------------------------------------------------------------- */

/* Random ACGT reference of n bases with two planted families: a 300-bp and a
 * 6-kb element, `copies` copies each at `div` substitution divergence. */
uint8_t *wl_reference(uint64_t seed, int64_t n, int copies, double div)
{
    if (n <= 0) return NULL;
    rng_t r;
    rng_seed(&r, seed * 0x9E3779B97F4A7C15ull + 11);
    uint8_t *ref = (uint8_t *)malloc((size_t)n);
    for (int64_t k = 0; k < n; ++k) ref[k] = (uint8_t)(next_u64(&r) & 3);
    const int fam_len[2] = {300, 6000};
    for (int f = 0; f < 2; ++f) {
        int L = fam_len[f];
        if (L * 4 > n) continue;
        uint8_t *elem = (uint8_t *)malloc((size_t)L);
        random_seq(&r, elem, L);
        for (int c = 0; c < copies; ++c) {
            int64_t pos = uni(&r, 0, n - L);
            for (int k = 0; k < L; ++k)
                ref[pos + k] = unif01(&r) < div ? (uint8_t)((elem[k] + 1 + uni(&r, 0, 2)) & 3) : elem[k];
        }
        free(elem);
    }
    return ref;
}

/* .pac layout: 4 bases per byte, first base in bits 7-6 (PackedReference.swift:4-5). */
void wl_pack_pac(const uint8_t *ref, int64_t n, uint8_t *pac)
{
    int64_t nb = (n + 3) / 4;
    memset(pac, 0, (size_t)nb);
    for (int64_t k = 0; k < n; ++k) pac[k >> 2] |= (uint8_t)((ref[k] & 3) << ((3 - (k & 3)) << 1));
}

void wl_free_bytes(uint8_t *p) { free(p); }

/* Read pairs of `len` bases in FR orientation: a fragment of length
 * round(N(ins_mean, ins_sd)) (at least len) starts at a uniform position of
 * ref (one sequence of gl bases); its source window is mutated with the shared
 * model; read 1 is the fragment's first len bases and read 2 the reverse
 * complement of its last len bases, or the other way round (probability 1/2:
 * a fragment from the reverse strand).  A fraction `hard` of pairs get a
 * substitution in read 2 every 15 bases (no exact 19-mer: seeding finds
 * nothing, local alignment still scores ~100).  pos: fragment start; rev: 1
 * when read 1 is the reverse-strand read. */
int wl_pe_reads(const uint8_t *ref, int64_t gl, uint64_t seed, int64_t n, int len,
                double ins_mean, double ins_sd, double p_sub, double p_indel, double hard,
                uint8_t *bases1, uint8_t *bases2, int64_t *pos, uint8_t *rev)
{
    const int slack = 64;
    const int max_frag = (int)(ins_mean + 8 * ins_sd) + len;
    if (gl < max_frag + slack + 1 || len <= 0) return -1;
    rng_t r;
    rng_seed(&r, seed * 0x9E3779B97F4A7C15ull + 23);
    uint8_t *buf = (uint8_t *)malloc((size_t)(max_frag + 2 * slack));
    for (int64_t i = 0; i < n; ++i) {
        int L, m;
        int64_t p;
        do {
            double u1 = unif01(&r), u2 = unif01(&r);
            double z = sqrt(-2.0 * log(u1 > 1e-300 ? u1 : 1e-300)) * cos(6.283185307179586 * u2);
            L = (int)floor(ins_mean + ins_sd * z + 0.5);
            if (L < len) L = len;
            if (L > max_frag) L = max_frag;
            p = uni(&r, 0, gl - L - slack);
            m = mutate(&r, ref + p, L + slack, p_sub, p_indel, buf, max_frag + 2 * slack, NULL);
        } while (m < L);
        const int rv = (int)(next_u64(&r) & 1);
        uint8_t *f = bases1 + i * len, *b = bases2 + i * len;
        uint8_t *left = rv ? b : f, *right = rv ? f : b;
        memcpy(left, buf, (size_t)len);
        for (int k = 0; k < len; ++k) right[k] = (uint8_t)(3 - buf[L - 1 - k]);
        if (unif01(&r) < hard)
            for (int k = 7; k < len; k += 15) b[k] = (uint8_t)((b[k] + 1 + uni(&r, 0, 2)) & 3);
        pos[i] = p;
        rev[i] = (uint8_t)rv;
    }
    free(buf);
    return 0;
}

/* Single-end reads of `len` bases sampled uniformly from both strands of ref
 * (one sequence of gl bases): the source window starts at a uniform position,
 * is mutated at p_sub / p_indel (the shared model), cut to `len` bases, and
 * reverse-complemented for reverse-strand reads.  bases: n x len codes;
 * pos: 0-based source start; rev: 1 for reverse strand. */
int wl_se_reads(const uint8_t *ref, int64_t gl, uint64_t seed, int64_t n, int len,
                double p_sub, double p_indel, uint8_t *bases, int64_t *pos, uint8_t *rev)
{
    const int slack = 64;
    if (gl < len + slack || len <= 0) return -1;
    rng_t r;
    rng_seed(&r, seed * 0x9E3779B97F4A7C15ull + 17);
    uint8_t *buf = (uint8_t *)malloc((size_t)(len + 2 * slack));
    for (int64_t i = 0; i < n; ++i) {
        int64_t p;
        int m;
        do {
            p = uni(&r, 0, gl - len - slack);
            m = mutate(&r, ref + p, len + slack, p_sub, p_indel, buf, len + 2 * slack, NULL);
        } while (m < len);
        const int rv = (int)(next_u64(&r) & 1);
        uint8_t *o = bases + i * len;
        if (rv) for (int k = 0; k < len; ++k) o[k] = (uint8_t)(3 - buf[len - 1 - k]);
        else memcpy(o, buf, (size_t)len);
        pos[i] = p;
        rev[i] = (uint8_t)rv;
    }
    free(buf);
    return 0;
}

typedef struct {
    int64_t n;                       /* jobs */
    uint8_t *q; int64_t q_bytes;
    uint32_t *q_off; uint16_t *q_len;
    int64_t *t_pos; uint16_t *t_len; uint8_t *t_rev;
    int32_t *a, *b, *c;              /* ext: h0, w, eb; nw: w, true_score */
    int32_t *r;                      /* read of the job: 2 * pair + mate */
} wl_jobs_t;

typedef struct {
    int64_t pairs;
    wl_jobs_t ext, loc, nw;
} wl_chunk_t;

static void jobs_alloc(wl_jobs_t *j, int64_t n, int qcap)
{
    j->n = 0;
    j->q = (uint8_t *)malloc((size_t)n * qcap + 1);
    j->q_bytes = 0;
    j->q_off = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n + 1));
    j->q_len = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)(n + 1));
    j->t_pos = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
    j->t_len = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)(n + 1));
    j->t_rev = (uint8_t *)calloc((size_t)n + 1, 1);
    j->a = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
    j->b = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
    j->c = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
    j->r = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
}

static int64_t jobs_add(wl_jobs_t *j, const uint8_t *q, int ql, int64_t tpos, int tlen, int rev)
{
    int64_t k = j->n++;
    j->q_off[k] = (uint32_t)j->q_bytes;
    j->q_len[k] = (uint16_t)ql;
    memcpy(j->q + j->q_bytes, q, (size_t)ql);
    j->q_bytes += ql;
    j->t_pos[k] = tpos;
    j->t_len[k] = (uint16_t)tlen;
    j->t_rev[k] = (uint8_t)rev;
    return k;
}

/* Chunk of `pairs` read pairs (2 x 150 bp, insert ~ N(400, 50), 0.5% subs /
 * 0.05% indels; mate 1 forward, mate 2 reverse strand) on a reference of gl
 * bases.  Per pair: for each mate, a 30-bp seed at offset U[20,100] of the
 * read and its left and right extension jobs (h0 = seed score, w = 100, end
 * bonus 5; windows sized as ExtensionAligner sizes them for a one-seed chain);
 * a CIGAR NW job per mate (w = 5); and a mate-rescue local-SW job of mate 2
 * (forward-oriented) against a 600-bp window.  Reverse-strand windows use
 * BWT-space positions in [gl, 2gl). */
wl_chunk_t *wl_chunk(const uint8_t *ref, int64_t gl, uint64_t seed, int64_t pairs)
{
    if (pairs < 0 || gl < 10000) return NULL;
    wl_chunk_t *c = (wl_chunk_t *)calloc(1, sizeof *c);
    c->pairs = pairs;
    jobs_alloc(&c->ext, pairs * 4, 150);
    jobs_alloc(&c->nw, pairs * 2, 150);
    jobs_alloc(&c->loc, pairs, 150);
    rng_t r;
    rng_seed(&r, seed * 0xD1342543DE82EF95ull + 29);
    uint8_t m[2][320], tmp[320];
    for (int64_t p = 0; p < pairs; ++p) {
        double u1 = unif01(&r) + 1e-12, u2 = unif01(&r);
        double z = __builtin_sqrt(-2.0 * __builtin_log(u1)) * __builtin_cos(6.283185307179586 * u2);
        int ins = (int)(400.0 + 50.0 * z);
        if (ins < 200) ins = 200;
        if (ins > 600) ins = 600;
        int64_t p1 = uni(&r, 1000, gl - 1000 - ins);
        /* mate 1: forward segment; mate 2: reverse complement of [p1+ins-150, p1+ins) */
        int l1 = mutate(&r, ref + p1, 150, 0.005, 0.0005, m[0], 300, NULL);
        int l2m = mutate(&r, ref + p1 + ins - 150, 150, 0.005, 0.0005, tmp, 300, NULL);
        if (l1 > 150) l1 = 150;
        if (l2m > 150) l2m = 150;
        for (int k = 0; k < l2m; ++k) m[1][k] = (uint8_t)(3 - tmp[l2m - 1 - k]);
        int lens[2] = {l1, l2m};
        /* locus of each mate in BWT space, read orientation */
        int64_t locus[2] = {p1, 2 * gl - (p1 + ins)};
        for (int mt = 0; mt < 2; ++mt) {
            const uint8_t *rd = m[mt];
            int rl = lens[mt];
            int o = (int)uni(&r, 20, 100);
            if (o + 30 > rl) o = rl - 30 > 0 ? rl - 30 : 0;
            int64_t rbeg = locus[mt] + o;
            uint8_t qrev[160];
            if (o > 0) {
                for (int k = 0; k < o; ++k) qrev[k] = rd[o - 1 - k];
                int L = o + cal_max_gap(o, 100);
                int64_t k = jobs_add(&c->ext, qrev, o, rbeg - L, L, 1);
                c->ext.a[k] = 30; c->ext.b[k] = 100; c->ext.c[k] = 5;
                c->ext.r[k] = (int32_t)(2 * p + mt);
            }
            int rem = rl - o - 30;
            if (rem > 0) {
                int R = rem + cal_max_gap(rem, 100);
                int64_t k = jobs_add(&c->ext, rd + o + 30, rem, rbeg + 30, R, 0);
                c->ext.a[k] = 30 + o; c->ext.b[k] = 100; c->ext.c[k] = 5;
                c->ext.r[k] = (int32_t)(2 * p + mt);
            }
            int64_t k = jobs_add(&c->nw, rd, rl, locus[mt], 150, 0);
            c->nw.a[k] = 5;
            c->nw.r[k] = (int32_t)(2 * p + mt);
            c->nw.b[k] = rl - 30;
        }
        /* rescue mate 2 (forward orientation) in a 600-bp window after mate 1 */
        for (int k = 0; k < l2m; ++k) tmp[k] = (uint8_t)(3 - m[1][l2m - 1 - k]);
        int64_t wstart = p1 + 400 - 150 - 225;
        if (wstart < 0) wstart = 0;
        if (wstart + 600 > gl) wstart = gl - 600;
        c->loc.r[jobs_add(&c->loc, tmp, l2m, wstart, 600, 0)] = (int32_t)(2 * p + 1);
    }
    return c;
}

void *wl_chunk_field(wl_chunk_t *c, int kind, int which)
{
    wl_jobs_t *j = kind == 0 ? &c->ext : kind == 1 ? &c->loc : &c->nw;
    switch (which) {
    case 0: return j->q;
    case 1: return j->q_off;
    case 2: return j->q_len;
    case 3: return j->t_pos;
    case 4: return j->t_len;
    case 5: return j->t_rev;
    case 6: return j->a;
    case 7: return j->b;
    case 8: return j->c;
    case 9: return j->r;
    default: return NULL;
    }
}

int64_t wl_chunk_count(wl_chunk_t *c, int kind, int which)
{
    wl_jobs_t *j = kind == 0 ? &c->ext : kind == 1 ? &c->loc : &c->nw;
    return which == 0 ? j->q_bytes : j->n;
}

static void jobs_free(wl_jobs_t *j)
{
    free(j->q); free(j->q_off); free(j->q_len); free(j->t_pos); free(j->t_len); free(j->t_rev);
    free(j->a); free(j->b); free(j->c); free(j->r);
}

void wl_chunk_free(wl_chunk_t *c)
{
    if (!c) return;
    jobs_free(&c->ext); jobs_free(&c->loc); jobs_free(&c->nw);
    free(c);
}
