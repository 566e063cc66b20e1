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
