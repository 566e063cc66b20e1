/*
 * CPU oracle for the swiftbwa Smith-Waterman hot path -- TEST INFRASTRUCTURE.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library, and only as the checker (or as the timed CPU baseline).  The
 * product library libhipsw never links it.
 *
 * Each function restates one Swift function of swiftbwa's Alignment module;
 * the citation at each function names the file:line ranges it follows.
 *
 * Parity pins (what checks each function against something other than itself;
 * the checks live in tests/test_oracle_*.py):
 *   or_ksw_extend      reference known answers (AlignmentTests.swift:48-63,
 *                      :82-96, :115-136, :140-162, :208-237, :239-283 closed
 *                      form), closed forms (perfect match, all-mismatch,
 *                      h0 decay), invariants, and an independent Python
 *                      restatement of published bwa ksw_extend2 (eh_t form)
 *                      on exhaustive/random tiny inputs.
 *   or_ksw_global      reference known answers (AlignmentTests.swift:313-446),
 *                      brute-force full-matrix Gotoh NW score when the band
 *                      cannot bind, CIGAR consistency (consumes qlen/tlen,
 *                      CIGAR re-scores to the reported score).
 *   or_cigar_generate  reference known answers (AlignmentTests.swift:450-688).
 *   or_local_sw        reference known answers (AlignmentTests.swift:715-780),
 *                      brute-force full local SW score, start/end re-scoring.
 *   or_extend_chain    reference known answers (AlignmentTests.swift:783-826,
 *                      :1140-1380).
 */
#include "sw_oracle.h"

#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------------- */
/* ScoringParameters.swift:2-45 defaults; scoringMatrix() :65-73.          */
/* ---------------------------------------------------------------------- */
void or_scoring_fill_matrix(or_scoring_t *sc)
{
    for (int i = 0; i < 25; ++i) sc->mat[i] = -1;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            sc->mat[i * 5 + j] = (int8_t)(i == j ? sc->a : -sc->b);
}

void or_scoring_default(or_scoring_t *sc)
{
    sc->a = 1; sc->b = 4;
    sc->o_ins = 6; sc->e_ins = 1;
    sc->o_del = 6; sc->e_del = 1;
    sc->w = 100; sc->zdrop = 100;
    sc->pen_clip5 = 5; sc->pen_clip3 = 5;
    sc->min_seed_len = 19;
    or_scoring_fill_matrix(sc);
}

static inline int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }
static inline int32_t imin(int32_t a, int32_t b) { return a < b ? a : b; }

/* ---------------------------------------------------------------------- */
/* BandedSWScalar.align -- BandedSWScalar.swift:21-191                     */
/* ---------------------------------------------------------------------- */
int64_t or_ksw_extend(int qlen, const uint8_t *query, int tlen, const uint8_t *target,
                      const or_scoring_t *sc, int32_t w, int32_t h0, int32_t end_bonus,
                      or_swres_t *out)
{
    memset(out, 0, sizeof *out);
    if (qlen <= 0 || tlen <= 0) return 0;                      /* :33-35 */
    const int m = 5;
    const int8_t *mat = sc->mat;
    const int32_t o_del = sc->o_del, e_del = sc->e_del;
    const int32_t o_ins = sc->o_ins, e_ins = sc->e_ins;
    const int32_t zdrop = sc->zdrop;

    /* query profile qp[k*qlen+j] = mat[k*5+query[j]]  (:44-50) */
    int32_t *qp = (int32_t *)malloc(sizeof(int32_t) * (size_t)m * (size_t)qlen);
    for (int k = 0; k < m; ++k)
        for (int j = 0; j < qlen; ++j) qp[k * qlen + j] = mat[k * m + query[j]];

    /* w cap (:52-62): Int32(Double(...)/Double(e) + 1.0) truncates to zero */
    int32_t max_mat = 0;
    for (int i = 0; i < m * m; ++i)
        if (mat[i] > max_mat) max_mat = mat[i];
    int32_t max_ins = (int32_t)((double)(qlen * max_mat + end_bonus - o_ins) / (double)e_ins + 1.0);
    w = imin(w, imax(1, max_ins));
    int32_t max_del = (int32_t)((double)(qlen * max_mat + end_bonus - o_del) / (double)e_del + 1.0);
    w = imin(w, imax(1, max_del));

    /* h[], e[] of qlen+1 entries, zero-filled (:64-66) */
    int32_t *h = (int32_t *)calloc((size_t)qlen + 1, sizeof(int32_t));
    int32_t *e = (int32_t *)calloc((size_t)qlen + 1, sizeof(int32_t));

    /* first row (:68-77) */
    h[0] = h0;
    if (h0 > o_ins + e_ins) {
        h[1] = h0 - o_ins - e_ins;
        for (int j = 2; j <= qlen && h[j - 1] > e_ins; ++j) h[j] = h[j - 1] - e_ins;
    }

    int32_t max_score = h0, max_i = -1, max_j = -1;           /* :79-84 */
    int32_t g_score = -1, g_tle = -1, max_off = 0;
    int beg = 0, end = qlen;                                   /* :86-87 */
    int64_t cells = 0;

    for (int i = 0; i < tlen; ++i) {                           /* :90 */
        const int32_t *qrow = qp + (int)target[i] * qlen;
        if (beg < i - w) beg = i - w;                          /* :95-97 */
        if (end > i + w + 1) end = i + w + 1;
        if (end > qlen) end = qlen;

        int32_t f = 0, h1;                                     /* :99-107 */
        if (beg == 0) {
            h1 = h0 - (o_del + e_del * (i + 1));
            if (h1 < 0) h1 = 0;
        } else {
            h1 = 0;
        }
        int32_t row_max = 0, row_max_j = -1;

        /* An empty row (beg >= end) runs no iterations: bwa's C semantics.
         * (The reference's Swift range would trap for beg > end, which the
         * narrowing rule below never produces.) */
        for (int j = beg; j < end; ++j) {                      /* :112-143 */
            int32_t M = h[j], ev = e[j];
            h[j] = h1;
            int32_t diag = M != 0 ? M + qrow[j] : 0;
            int32_t hv = diag > ev ? diag : ev;
            hv = hv > f ? hv : f;
            h1 = hv;
            row_max_j = row_max > hv ? row_max_j : j;
            row_max = row_max > hv ? row_max : hv;
            int32_t t = diag - (o_del + e_del);
            if (t < 0) t = 0;
            int32_t ne = ev - e_del;
            if (ne < t) ne = t;
            e[j] = ne;
            t = diag - (o_ins + e_ins);
            if (t < 0) t = 0;
            int32_t nf = f - e_ins;
            if (nf < t) nf = t;
            f = nf;
        }
        if (end > beg) cells += end - beg;
        h[end] = h1;                                           /* :145-146 */
        e[end] = 0;

        if (end == qlen) {                                     /* :149-152 */
            g_tle = g_score > h1 ? g_tle : i;
            g_score = g_score > h1 ? g_score : h1;
        }
        if (row_max == 0) break;                               /* :155 */
        if (row_max > max_score) {                             /* :158-163 */
            max_score = row_max;
            max_i = i;
            max_j = row_max_j;
            int32_t off = row_max_j - i;
            if (off < 0) off = -off;
            if (off > max_off) max_off = off;
        } else if (zdrop > 0) {                                /* :164-172 */
            int32_t di = i - max_i, dj = row_max_j - max_j;
            if (di > dj) {
                if (max_score - row_max - (di - dj) * e_del > zdrop) break;
            } else {
                if (max_score - row_max - (dj - di) * e_ins > zdrop) break;
            }
        }
        /* dynamic band narrowing (:175-180) */
        int nb = beg;
        while (nb < end && h[nb] == 0 && e[nb] == 0) ++nb;
        beg = nb;
        int ne = end;
        while (ne >= beg && h[ne] == 0 && e[ne] == 0) --ne;
        end = ne + 2 < qlen ? ne + 2 : qlen;
    }
    free(qp); free(h); free(e);
    out->score = max_score;                                    /* :183-190 */
    out->qle = max_j + 1;
    out->tle = max_i + 1;
    out->gtle = g_tle + 1;
    out->gscore = g_score;
    out->max_off = max_off;
    return cells;
}

/* ---------------------------------------------------------------------- */
/* CIGARBuilder.swift:4-38 : append merges equal consecutive ops.          */
/* ---------------------------------------------------------------------- */
typedef struct { uint32_t *ops; int n, cap; } cig_t;

static void cig_append(cig_t *c, uint32_t op, uint32_t len)
{
    if (c->n > 0 && (c->ops[c->n - 1] & 0xF) == op) {
        c->ops[c->n - 1] += len << 4;
        return;
    }
    if (c->n == c->cap) {
        c->cap = c->cap ? c->cap * 2 : 16;
        c->ops = (uint32_t *)realloc(c->ops, sizeof(uint32_t) * (size_t)c->cap);
    }
    c->ops[c->n++] = (len << 4) | op;
}

/* ---------------------------------------------------------------------- */
/* GlobalAligner.align -- GlobalAligner.swift:19-151; traceback :160-232   */
/* ---------------------------------------------------------------------- */
static void global_fill(int qlen, const uint8_t *query, int tlen, const uint8_t *target,
                        const or_scoring_t *sc, int32_t w, int nCol, uint8_t *z,
                        int32_t *score, int64_t *cells)
{
    const int m = 5;
    const int32_t o_del = sc->o_del, e_del = sc->e_del;
    const int32_t o_ins = sc->o_ins, e_ins = sc->e_ins;
    const int32_t oe_del = o_del + e_del, oe_ins = o_ins + e_ins;
    const int32_t neg_inf = INT32_MIN / 2;                     /* :52 */
    int32_t *qp = (int32_t *)malloc(sizeof(int32_t) * (size_t)m * (size_t)qlen);
    int32_t *h = (int32_t *)malloc(sizeof(int32_t) * ((size_t)qlen + 1));
    int32_t *e = (int32_t *)malloc(sizeof(int32_t) * ((size_t)qlen + 1));
    for (int k = 0; k < m; ++k)                                /* :64-72 */
        for (int j = 0; j < qlen; ++j) qp[k * qlen + j] = sc->mat[k * m + query[j]];
    for (int j = 0; j <= qlen; ++j) { h[j] = neg_inf; e[j] = neg_inf; }   /* :75-77 */
    memset(z, 0, (size_t)tlen * (size_t)nCol);
    h[0] = 0;                                                  /* :80-84 */
    int init_end = w < qlen ? w : qlen;
    for (int j = 1; j <= init_end; ++j) h[j] = 0 - (o_ins + e_ins * j);
    int64_t nc = 0;
    for (int i = 0; i < tlen; ++i) {                           /* :89-142 */
        int beg = i - w > 0 ? i - w : 0;
        int end = i + w + 1 < qlen ? i + w + 1 : qlen;
        const int32_t *qrow = qp + (int)target[i] * qlen;
        int32_t f = neg_inf;
        int32_t h_diag = h[beg];
        h[beg] = beg == 0 ? 0 - (oe_del + e_del * i) : neg_inf;
        uint8_t *zi = z + (size_t)i * nCol - beg;
        for (int j = beg; j < end; ++j) {
            int j1 = j + 1;
            int32_t h_cur = h_diag + qrow[j];
            h_diag = h[j1];
            int32_t e_open = h_diag - oe_del;
            int32_t e_ext = e[j1] - e_del;
            uint8_t e_flag = e_open > e_ext ? 0 : 0x01;
            int32_t ev = e_open > e_ext ? e_open : e_ext;
            e[j1] = ev;
            int32_t f_open = h[j] - oe_ins;
            int32_t f_ext = f - e_ins;
            uint8_t f_flag = f_open > f_ext ? 0 : 0x02;
            int32_t fv = f_open > f_ext ? f_open : f_ext;
            f = fv;
            uint8_t from_e = ev > h_cur ? 0x04 : 0;
            h_cur = ev > h_cur ? ev : h_cur;
            uint8_t from_f = fv > h_cur ? 0x08 : 0;
            h_cur = fv > h_cur ? fv : h_cur;
            h[j1] = h_cur;
            zi[j] = e_flag | f_flag | from_e | from_f;
        }
        if (end > beg) nc += end - beg;
    }
    *score = h[qlen];                                          /* :144 */
    if (cells) *cells += nc;
    free(qp); free(h); free(e);
}

static void global_traceback(const uint8_t *z, int nCol, int qlen, int tlen, int32_t w, cig_t *c)
{
    int i = tlen - 1, j = qlen - 1, state = 0;                 /* :168-170 */
    while (i >= 0 || j >= 0) {                                 /* :172 */
        if (i < 0) { cig_append(c, 1, 1); j -= 1; continue; }  /* :173-178 */
        if (j < 0) { cig_append(c, 2, 1); i -= 1; continue; }  /* :179-184 */
        int beg = i - w > 0 ? i - w : 0;                       /* :186-191 */
        int col = j - beg;
        if (!(col >= 0 && col < nCol)) break;
        uint8_t d = z[(size_t)i * nCol + col];
        if (state == 0) {                                      /* :196-209 */
            int hdir = (d >> 2) & 0x03;
            if (hdir == 0) { cig_append(c, 0, 1); i -= 1; j -= 1; }
            else if (hdir == 1) state = 1;
            else state = 2;
        } else if (state == 1) {                               /* :210-216 */
            cig_append(c, 2, 1);
            if ((d & 0x01) == 0) state = 0;
            i -= 1;
        } else {                                               /* :217-223 */
            cig_append(c, 1, 1);
            if ((d & 0x02) == 0) state = 0;
            j -= 1;
        }
    }
    for (int a = 0, b = c->n - 1; a < b; ++a, --b) {          /* :230 reverse */
        uint32_t t = c->ops[a]; c->ops[a] = c->ops[b]; c->ops[b] = t;
    }
}

/* Full GlobalAligner.align: returns CIGAR in a cig_t owned by the caller. */
static void global_align(int qlen, const uint8_t *query, int tlen, const uint8_t *target,
                         const or_scoring_t *sc, int32_t w_in, int32_t *score, cig_t *c,
                         int64_t *cells)
{
    c->n = 0;
    if (qlen <= 0 || tlen <= 0) { *score = 0; return; }        /* :29-31 */
    int32_t d = qlen - tlen;
    if (d < 0) d = -d;
    int32_t w = w_in > d ? w_in : d;                           /* :47 */
    int nCol = qlen < 2 * w + 1 ? qlen : 2 * w + 1;            /* :48 */
    uint8_t *z = (uint8_t *)malloc((size_t)tlen * (size_t)nCol);
    global_fill(qlen, query, tlen, target, sc, w, nCol, z, score, cells);
    global_traceback(z, nCol, qlen, tlen, w, c);
    free(z);
}

int or_ksw_global(int qlen, const uint8_t *query, int tlen, const uint8_t *target,
                  const or_scoring_t *sc, int32_t w, int32_t *score,
                  uint32_t *cigar, int cap, int64_t *cells)
{
    cig_t c = {0};
    int64_t nc = 0;
    global_align(qlen, query, tlen, target, sc, w, score, &c, &nc);
    if (cells) *cells = nc;
    int n = c.n;
    for (int k = 0; k < n && k < cap; ++k) cigar[k] = c.ops[k];
    free(c.ops);
    return n;
}

/* CIGARGenerator.inferBandWidth -- CIGARGenerator.swift:147-170 */
static int32_t infer_band_width(int32_t qlen, int32_t rlen, int32_t score, const or_scoring_t *sc)
{
    int32_t max_len = imax(qlen, rlen), min_len = imin(qlen, rlen);
    int32_t diff = max_len - min_len;
    int32_t expected = min_len * sc->a;
    if (score >= expected) return diff + 3;
    int32_t deficit = expected - score;
    int32_t gap_cost = sc->o_ins + sc->e_ins;
    int32_t mm_delta = sc->a + sc->b;
    int32_t est = deficit / imin(gap_cost, mm_delta);
    return imax(diff + 3, est);
}

int or_global_with_retry(const uint8_t *q, int qlen, const uint8_t *t, int tlen,
                         int32_t w0, int32_t true_score, int max_retries,
                         const or_scoring_t *sc, int32_t *score, uint32_t *cigar,
                         int cap, int32_t *final_w, int64_t *cells)
{
    /* CIGARGenerator.swift:82-96: NW at w, then double w while the score is
     * below trueScore, at most max_retries times. */
    cig_t c = {0};
    int64_t nc = 0;
    int32_t w = w0, sc_out = 0;
    global_align(qlen, q, tlen, t, sc, w, &sc_out, &c, &nc);
    int retries = 0;
    while (sc_out < true_score && retries < max_retries) {
        w *= 2;
        retries += 1;
        global_align(qlen, q, tlen, t, sc, w, &sc_out, &c, &nc);
    }
    *score = sc_out;
    if (final_w) *final_w = w;
    if (cells) *cells = nc;
    int n = c.n;
    for (int k = 0; k < n && k < cap; ++k) cigar[k] = c.ops[k];
    free(c.ops);
    return n;
}

/* ---------------------------------------------------------------------- */
/* CIGARGenerator.generate -- CIGARGenerator.swift:30-111, helpers :172-331 */
/* ---------------------------------------------------------------------- */
typedef struct { char *s; int n, cap; } str_t;

static void str_putc(str_t *s, char ch)
{
    if (s->n + 1 >= s->cap) {
        s->cap = s->cap ? s->cap * 2 : 64;
        s->s = (char *)realloc(s->s, (size_t)s->cap);
    }
    s->s[s->n++] = ch;
    s->s[s->n] = 0;
}

static void str_putint(str_t *s, int v)
{
    char buf[24];
    int k = 0;
    if (v == 0) buf[k++] = '0';
    while (v > 0) { buf[k++] = (char)('0' + v % 10); v /= 10; }
    while (k > 0) str_putc(s, buf[--k]);
}

static char base_char(uint8_t b)                               /* :323-331 */
{
    switch (b) { case 0: return 'A'; case 1: return 'C'; case 2: return 'G'; case 3: return 'T'; default: return 'N'; }
}

static int32_t nm_from_cigar(const uint32_t *cig, int n, const uint8_t *q, int qn,
                             const uint8_t *t, int tn, int toff)   /* :173-210 */
{
    int32_t nm = 0;
    int qi = 0, ti = toff;
    for (int k = 0; k < n; ++k) {
        uint32_t op = cig[k] & 0xF;
        int len = (int)(cig[k] >> 4);
        if (op == 0) {
            for (int x = 0; x < len; ++x)
                if (qi < qn && ti < tn) { if (q[qi] != t[ti]) nm += 1; qi++; ti++; }
        } else if (op == 1) { nm += len; qi += len; }
        else if (op == 2) { nm += len; ti += len; }
    }
    return nm;
}

static void md_from_cigar(const uint32_t *cig, int n, const uint8_t *q, int qn,
                          const uint8_t *t, int tn, int toff, str_t *md)  /* :219-268 */
{
    int mc = 0, qi = 0, ti = toff;
    for (int k = 0; k < n; ++k) {
        uint32_t op = cig[k] & 0xF;
        int len = (int)(cig[k] >> 4);
        if (op == 0) {
            for (int x = 0; x < len; ++x) {
                if (qi < qn && ti < tn) {
                    if (q[qi] == t[ti]) mc += 1;
                    else { str_putint(md, mc); str_putc(md, base_char(t[ti])); mc = 0; }
                    qi++; ti++;
                }
            }
        } else if (op == 1) {
            qi += len;
        } else if (op == 2) {
            str_putint(md, mc);
            str_putc(md, '^');
            for (int x = 0; x < len; ++x)
                if (ti < tn) { str_putc(md, base_char(t[ti])); ti++; }
            mc = 0;
        }
    }
    str_putint(md, mc);
}

static void squeeze_deletions(cig_t *c, int64_t *pos_adj, int *ref_off)  /* :272-295 */
{
    *pos_adj = 0; *ref_off = 0;
    int lead = 0;
    while (lead < c->n && (c->ops[lead] & 0xF) == 2) {
        int len = (int)(c->ops[lead] >> 4);
        *pos_adj += len; *ref_off += len;
        lead++;
    }
    if (lead) { memmove(c->ops, c->ops + lead, sizeof(uint32_t) * (size_t)(c->n - lead)); c->n -= lead; }
    while (c->n > 0 && (c->ops[c->n - 1] & 0xF) == 2) c->n -= 1;
}

static void add_soft_clips(cig_t *c, int32_t qb, int32_t qe, int32_t read_len, int is_rev)  /* :298-320 */
{
    int32_t lead, trail;
    if (is_rev) { lead = read_len - qe; trail = qb; }
    else { lead = qb; trail = read_len - qe; }
    if (lead > 0) {
        if (c->n + 1 > c->cap) { c->cap = c->n + 8; c->ops = (uint32_t *)realloc(c->ops, sizeof(uint32_t) * (size_t)c->cap); }
        memmove(c->ops + 1, c->ops, sizeof(uint32_t) * (size_t)c->n);
        c->ops[0] = ((uint32_t)lead << 4) | 4;
        c->n += 1;
    }
    if (trail > 0) {
        if (c->n + 1 > c->cap) { c->cap = c->n + 8; c->ops = (uint32_t *)realloc(c->ops, sizeof(uint32_t) * (size_t)c->cap); }
        c->ops[c->n++] = ((uint32_t)trail << 4) | 4;
    }
}

int or_cigar_generate(const uint8_t *qseg, int qlen, const uint8_t *rseg, int rlen,
                      int32_t qb, int32_t qe, int32_t read_len, int is_reverse,
                      int32_t true_score, int32_t initial_w, const or_scoring_t *sc,
                      int64_t ref_pos, uint32_t *cigar, int cap, char *md, int md_cap,
                      or_cigar_res_t *res)
{
    memset(res, 0, sizeof *res);
    int32_t w = infer_band_width(qe - qb, rlen, true_score, sc);   /* :44-50 */
    w = imax(w, initial_w);
    cig_t c = {0};
    str_t mds = {0};
    str_putc(&mds, 'x'); mds.n = 0; mds.s[0] = 0;
    int32_t min_indel = sc->o_ins + sc->e_ins + sc->o_del + sc->e_del;   /* :65-67 */
    int32_t deficit = qlen * sc->a - true_score;
    int32_t score;
    int64_t pos = ref_pos;
    if (qlen == rlen && qlen > 0 && deficit < min_indel) {     /* :69-80 fast path */
        cig_append(&c, 0, (uint32_t)qlen);
        res->nm = nm_from_cigar(c.ops, c.n, qseg, qlen, rseg, rlen, 0);
        md_from_cigar(c.ops, c.n, qseg, qlen, rseg, rlen, 0, &mds);
        add_soft_clips(&c, qb, qe, read_len, is_reverse);
        score = true_score;
    } else {
        int64_t cells = 0;
        global_align(qlen, qseg, rlen, rseg, sc, w, &score, &c, &cells);   /* :82-86 */
        res->nw_calls = 1;
        int retries = 0;
        while (score < true_score && retries < 3) {            /* :88-96 */
            w *= 2;
            retries += 1;
            global_align(qlen, qseg, rlen, rseg, sc, w, &score, &c, &cells);
            res->nw_calls += 1;
        }
        res->cells = cells;
        int64_t pos_adj; int ref_off;
        squeeze_deletions(&c, &pos_adj, &ref_off);             /* :98-101 */
        pos += pos_adj;
        res->nm = nm_from_cigar(c.ops, c.n, qseg, qlen, rseg, rlen, ref_off);
        md_from_cigar(c.ops, c.n, qseg, qlen, rseg, rlen, ref_off, &mds);
        add_soft_clips(&c, qb, qe, read_len, is_reverse);
    }
    res->score = score;
    res->pos = pos;
    res->n_cigar = c.n;
    for (int k = 0; k < c.n && k < cap; ++k) cigar[k] = c.ops[k];
    if (md && md_cap > 0) {
        int L = mds.n < md_cap - 1 ? mds.n : md_cap - 1;
        memcpy(md, mds.s, (size_t)L);
        md[L] = 0;
    }
    free(c.ops);
    free(mds.s);
    return res->n_cigar;
}

/* ---------------------------------------------------------------------- */
/* LocalSWAligner -- LocalSWAligner.swift:42-492                            */
/* 16-lane uint8 Farrar striped pass emulated lane by lane.                 */
/* ---------------------------------------------------------------------- */
typedef struct { uint8_t v[16]; } u8x16;

static inline uint8_t u8max(uint8_t a, uint8_t b) { return a > b ? a : b; }
static inline u8x16 shift_right(u8x16 a)                       /* :316-323 */
{
    u8x16 r;
    r.v[0] = 0;
    for (int l = 1; l < 16; ++l) r.v[l] = a.v[l - 1];
    return r;
}

typedef struct { uint64_t *p; int n, cap; } peaks_t;

static void peaks_record(peaks_t *pk, uint64_t score, int32_t axis)
{
    /* LocalSWAligner.swift:266-286 / :451-471 */
    uint64_t entry = (score << 32) | (uint64_t)(uint32_t)axis;
    if (pk->n == 0 || (int32_t)(pk->p[pk->n - 1] & 0xFFFFFFFFu) + 1 != axis) {
        if (pk->n == pk->cap) {
            pk->cap = pk->cap == 0 ? 8 : pk->cap << 1;
            pk->p = (uint64_t *)realloc(pk->p, sizeof(uint64_t) * (size_t)pk->cap);
        }
        pk->p[pk->n++] = entry;
    } else if ((pk->p[pk->n - 1] >> 32) < score) {
        pk->p[pk->n - 1] = entry;
    }
}

static int32_t peaks_score2(const peaks_t *pk, int32_t best, int32_t max_base, int32_t axis)
{
    /* LocalSWAligner.swift:289-303 / :474-489 */
    int32_t score2 = 0;
    if (pk->p && best > 0 && max_base > 0) {
        int32_t r = (best + max_base - 1) / max_base;
        int32_t lo = axis - r, hi = axis + r;
        for (int k = 0; k < pk->n; ++k) {
            int32_t col = (int32_t)(uint32_t)(pk->p[k] & 0xFFFFFFFFu);
            int32_t s = (int32_t)(pk->p[k] >> 32);
            if ((col < lo || col > hi) && s > score2) score2 = s;
        }
    }
    return score2;
}

/* simdSWPass -- :133-306.  Returns 0 on 8-bit overflow (the reference's nil). */
static int simd_sw_pass(const uint8_t *query, int qlen, const uint8_t *target, int tlen,
                        const int8_t *mat, int32_t gap_oe, int32_t gap_e, uint8_t bias,
                        uint8_t min_sub, int32_t *o_score, int32_t *o_qend, int32_t *o_tend,
                        int32_t *o_score2, int64_t *cells)
{
    const int lanes = 16;
    const int sn = (qlen + lanes - 1) / lanes;
    u8x16 *profile = (u8x16 *)calloc((size_t)5 * (size_t)sn, sizeof(u8x16));
    u8x16 *H = (u8x16 *)calloc((size_t)sn, sizeof(u8x16));
    u8x16 *E = (u8x16 *)calloc((size_t)sn, sizeof(u8x16));
    for (int k = 0; k < 5; ++k)                                /* :152-167 */
        for (int s = 0; s < sn; ++s)
            for (int l = 0; l < lanes; ++l) {
                int j = s + l * sn;
                if (j < qlen) {
                    int v = (int)mat[k * 5 + query[j]] + (int)bias;
                    profile[k * sn + s].v[l] = (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v);
                }
            }
    const uint8_t goe = (uint8_t)gap_oe, ge = (uint8_t)gap_e;
    uint8_t max_score = 0;
    int32_t max_i = -1, max_j = -1;
    peaks_t pk = {0};
    const int track = min_sub > 0;
    uint8_t max_base = 0;
    if (track) {                                               /* :191-199 */
        for (int k = 0; k < 25; ++k) {
            int v = (int)mat[k] + (int)bias;
            uint8_t u = (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v);
            if (u > max_base) max_base = u;
        }
        max_base = max_base > bias ? (uint8_t)(max_base - bias) : 0;
    }
    int ok = 1;
    for (int i = 0; i < tlen && ok; ++i) {                     /* :201 */
        int tb = target[i] < 4 ? target[i] : 4;
        const u8x16 *prof = profile + tb * sn;
        u8x16 vH = shift_right(H[sn - 1]);
        u8x16 f = {{0}};
        for (int s = 0; s < sn; ++s) {                         /* :209-231 */
            u8x16 hn, hp = H[s];
            for (int l = 0; l < lanes; ++l) {
                uint8_t x = (uint8_t)(vH.v[l] + prof[s].v[l]);          /* wrapping &+ */
                x = (uint8_t)(u8max(x, bias) - bias);
                uint8_t eo = (uint8_t)(u8max(hp.v[l], goe) - goe);
                uint8_t ev = u8max(eo, E[s].v[l]);
                ev = (uint8_t)(u8max(ev, ge) - ge);
                E[s].v[l] = ev;
                x = u8max(x, ev);
                x = u8max(x, f.v[l]);
                hn.v[l] = x;
                uint8_t fo = (uint8_t)(u8max(x, goe) - goe);
                uint8_t fv = u8max(fo, f.v[l]);
                f.v[l] = (uint8_t)(u8max(fv, ge) - ge);
            }
            vH = hp;
            H[s] = hn;
        }
        u8x16 cf = shift_right(f);                             /* :234-242 lazy-F */
        for (int s = 0; s < sn; ++s) {
            u8x16 nh;
            int same = 1;
            for (int l = 0; l < lanes; ++l) {
                nh.v[l] = u8max(H[s].v[l], cf.v[l]);
                if (nh.v[l] != H[s].v[l]) same = 0;
            }
            if (same) break;
            H[s] = nh;
            for (int l = 0; l < lanes; ++l) {
                uint8_t fo = (uint8_t)(u8max(H[s].v[l], goe) - goe);
                uint8_t c = u8max(fo, cf.v[l]);
                cf.v[l] = (uint8_t)(u8max(c, ge) - ge);
            }
        }
        if (cells) *cells += qlen;
        uint8_t col_max = 0;                                   /* :245-263 */
        for (int s = 0; s < sn; ++s) {
            uint8_t lm = 0;
            for (int l = 0; l < lanes; ++l) lm = u8max(lm, H[s].v[l]);
            if (lm > 250) { ok = 0; break; }
            if (lm > col_max) col_max = lm;
            if (lm > max_score) {
                max_score = lm;
                max_i = i;
                for (int l = 0; l < lanes; ++l) {
                    if (H[s].v[l] == lm) {
                        int j = s + l * sn;
                        if (j < qlen) { max_j = j; break; }
                    }
                }
            }
        }
        if (!ok) break;
        if (track && col_max >= min_sub) peaks_record(&pk, col_max, i);   /* :266-286 */
    }
    if (ok) {
        int32_t s2 = track ? peaks_score2(&pk, max_score, max_base, max_i) : 0;
        *o_score = max_score; *o_qend = max_j; *o_tend = max_i; *o_score2 = s2;
    }
    free(pk.p); free(profile); free(H); free(E);
    return ok;
}

/* scalarSWPass -- :383-492 (query-outer, strict '>' ties) */
static void scalar_sw_pass(const uint8_t *query, int qlen, const uint8_t *target, int tlen,
                           const int8_t *mat, int32_t gap_oe, int32_t gap_e, int32_t min_sub,
                           int32_t *o_score, int32_t *o_qend, int32_t *o_tend, int32_t *o_score2,
                           int64_t *cells)
{
    int32_t *h = (int32_t *)calloc((size_t)tlen + 1, sizeof(int32_t));
    int32_t *e = (int32_t *)calloc((size_t)tlen + 1, sizeof(int32_t));
    int32_t best = 0, bq = -1, bt = -1;
    const int track = min_sub > 0;
    peaks_t pk = {0};
    int32_t max_base = 0;
    if (track)
        for (int k = 0; k < 25; ++k) if (mat[k] > max_base) max_base = mat[k];
    for (int i = 0; i < qlen; ++i) {
        int qb = query[i] < 4 ? query[i] : 4;
        const int8_t *mrow = mat + qb * 5;
        int32_t f = 0, hprev = 0, row_max = 0;
        for (int j = 0; j < tlen; ++j) {
            int j1 = j + 1;
            int tb = target[j] < 4 ? target[j] : 4;
            int32_t hc = hprev + mrow[tb];
            int32_t fo = h[j1] - gap_oe, fx = f - gap_e;
            f = fo > fx ? fo : fx;
            int32_t eo = h[j] - gap_oe, ex = e[j1] - gap_e;
            int32_t ev = eo > ex ? eo : ex;
            e[j1] = ev;
            hc = hc > f ? hc : f;
            hc = hc > ev ? hc : ev;
            hc = hc > 0 ? hc : 0;
            hprev = h[j1];
            h[j1] = hc;
            if (hc > row_max) row_max = hc;
            if (hc > best) { best = hc; bq = i; bt = j; }
        }
        if (cells) *cells += tlen;
        if (track && row_max >= min_sub) peaks_record(&pk, (uint64_t)(uint32_t)row_max, i);
    }
    *o_score2 = track ? peaks_score2(&pk, best, max_base, bq) : 0;
    *o_score = best; *o_qend = bq; *o_tend = bt;
    free(pk.p); free(h); free(e);
}

static void reversed_prefix(const uint8_t *s, int n, uint8_t *out)
{
    for (int k = 0; k < n; ++k) out[k] = s[n - 1 - k];
}

int or_local_sw(const uint8_t *query, int qlen, const uint8_t *target, int tlen,
                const or_scoring_t *sc, int min_sub_score, int32_t *out6, int64_t *cells)
{
    if (qlen <= 0 || tlen <= 0) return 0;                      /* :48 */
    const int8_t *mat = sc->mat;
    const int32_t gap_oe = sc->o_ins + sc->e_ins, gap_e = sc->e_ins;   /* :51-53 */
    const uint8_t bias = (uint8_t)sc->b;
    const uint8_t min_sub = (uint8_t)min_sub_score;
    int64_t nc = 0;
    if (qlen * sc->a <= 250) {                                  /* :56-71 */
        int32_t fs, fq, ft, f2;
        if (simd_sw_pass(query, qlen, target, tlen, mat, gap_oe, gap_e, bias, min_sub,
                         &fs, &fq, &ft, &f2, &nc) && fs > 0) {
            int rq = fq + 1, rt = ft + 1;                       /* :100-115 */
            uint8_t *rq_s = (uint8_t *)malloc((size_t)rq + 1), *rt_s = (uint8_t *)malloc((size_t)rt + 1);
            reversed_prefix(query, rq, rq_s);
            reversed_prefix(target, rt, rt_s);
            int32_t rs, rqe, rte, r2;
            int okr = simd_sw_pass(rq_s, rq, rt_s, rt, mat, gap_oe, gap_e, bias, 0,
                                   &rs, &rqe, &rte, &r2, &nc);
            free(rq_s); free(rt_s);
            if (okr) {
                out6[0] = fs; out6[1] = fq - rqe; out6[2] = fq;  /* :117-124 */
                out6[3] = ft - rte; out6[4] = ft; out6[5] = f2;
                if (cells) *cells += nc;
                return 1;
            }
        }
    }
    /* scalarAlign -- :327-381 */
    int32_t fs, fq, ft, f2;
    scalar_sw_pass(query, qlen, target, tlen, mat, gap_oe, gap_e, min_sub, &fs, &fq, &ft, &f2, &nc);
    if (fs <= 0) { if (cells) *cells += nc; return 0; }
    int rq = fq + 1, rt = ft + 1;
    uint8_t *rq_s = (uint8_t *)malloc((size_t)rq), *rt_s = (uint8_t *)malloc((size_t)rt);
    reversed_prefix(query, rq, rq_s);
    reversed_prefix(target, rt, rt_s);
    int32_t rs, rqe, rte, r2;
    scalar_sw_pass(rq_s, rq, rt_s, rt, mat, gap_oe, gap_e, 0, &rs, &rqe, &rte, &r2, &nc);
    free(rq_s); free(rt_s);
    out6[0] = fs; out6[1] = fq - rqe; out6[2] = fq;
    out6[3] = ft - rte; out6[4] = ft; out6[5] = f2;
    if (cells) *cells += nc;
    return 1;
}

/* ---------------------------------------------------------------------- */
/* ExtensionAligner.extend -- ExtensionAligner.swift:114-346, :352-363      */
/* ---------------------------------------------------------------------- */
static int cal_max_gap(int qlen, const or_scoring_t *sc)       /* :352-363 */
{
    int a = sc->a;
    int l_del = sc->e_del > 0 ? (qlen * a - sc->o_del) / sc->e_del + 1 : 1;
    int l_ins = sc->e_ins > 0 ? (qlen * a - sc->o_ins) / sc->e_ins + 1 : 1;
    int l = l_del > l_ins ? l_del : l_ins;
    if (l < 1) l = 1;
    int cap = sc->w << 1;
    return l < cap ? l : cap;
}

/* Two getReference semantics.  gl == 0: the reference tests' closure
 * (AlignmentTests.swift:806-816), start = max(0,pos), end = min(count,
 * start+len) over a flat array.  gl > 0: FMIndex.getReference
 * (FMIndex.swift:25-50) over the forward genome `ref` of length gl, where
 * positions >= gl address the reverse complement: a window is clipped at the
 * strand boundary and at both genome ends. */
static int get_ref(const uint8_t *ref, int64_t ref_len, int64_t gl, int64_t pos, int len, uint8_t *buf)
{
    if (gl > 0) {
        if (pos >= gl) {
            int64_t fwd_end = 2 * gl - pos, fwd_start = fwd_end - len;
            int64_t s0 = fwd_start > 0 ? fwd_start : 0, e0 = fwd_end < gl ? fwd_end : gl;
            int n = e0 > s0 ? (int)(e0 - s0) : 0;
            for (int k = 0; k < n; ++k) buf[k] = (uint8_t)(3 - ref[e0 - 1 - k]);
            return n;
        }
        int64_t n64 = (int64_t)len < gl - pos ? (int64_t)len : gl - pos;
        if (n64 <= 0 || pos < 0) return 0;
        memcpy(buf, ref + pos, (size_t)n64);
        return (int)n64;
    }
    int64_t start = pos > 0 ? pos : 0;
    int64_t end = start + len < ref_len ? start + len : ref_len;
    int n = end > start ? (int)(end - start) : 0;
    if (n) memcpy(buf, ref + start, (size_t)n);
    return n;
}

typedef struct { int64_t qb, qe, rb, re, score, true_score, sub, sub_n, w, seed_cov; } reg_t;

int or_extend_chain(int n_seeds, const int32_t *sq, const int32_t *sl, const int64_t *sr,
                    const int32_t *ss, const uint8_t *read, int read_len,
                    const uint8_t *ref, int64_t ref_len, const or_scoring_t *sc,
                    int64_t *regions_out, int max_regions, int64_t *ext_calls, int64_t *cells)
{
    return or_extend_chain_ex(n_seeds, sq, sl, sr, ss, read, read_len, ref, ref_len, 0, sc,
                              regions_out, max_regions, ext_calls, cells, NULL, NULL);
}

int or_extend_chain_ex(int n_seeds, const int32_t *sq, const int32_t *sl, const int64_t *sr,
                       const int32_t *ss, const uint8_t *read, int read_len,
                       const uint8_t *ref, int64_t ref_len, int64_t gl, const or_scoring_t *sc,
                       int64_t *regions_out, int max_regions, int64_t *ext_calls, int64_t *cells,
                       or_ext_log_fn log, void *log_ctx)
{
    if (n_seeds <= 0) return 0;
    const int32_t band = sc->w;
    int64_t rmax0 = INT64_MAX, rmax1 = 0;                      /* :134-144 */
    for (int i = 0; i < n_seeds; ++i) {
        int64_t b = sr[i] - (int64_t)(sq[i] + cal_max_gap(sq[i], sc));
        int rem = read_len - sq[i] - sl[i];
        int64_t e = sr[i] + sl[i] + (int64_t)(rem + cal_max_gap(rem, sc));
        if (b < rmax0) rmax0 = b;
        if (e > rmax1) rmax1 = e;
    }
    if (rmax0 < 0) rmax0 = 0;
    int32_t tmp = imax(sc->a + sc->b, imax(sc->o_ins + sc->e_ins, sc->o_del + sc->e_del));

    /* seeds by score, descending; stable (:155-158) */
    int *order = (int *)malloc(sizeof(int) * (size_t)n_seeds);
    for (int i = 0; i < n_seeds; ++i) order[i] = i;
    for (int i = 1; i < n_seeds; ++i) {                        /* insertion sort = stable */
        int x = order[i], k = i - 1;
        while (k >= 0 && ss[order[k]] < ss[x]) { order[k + 1] = order[k]; --k; }
        order[k + 1] = x;
    }
    reg_t *regs = (reg_t *)calloc((size_t)n_seeds, sizeof(reg_t));
    int nreg = 0;
    uint8_t *qbuf = (uint8_t *)malloc((size_t)read_len + 1);
    uint8_t *tbuf = (uint8_t *)malloc(10200);
    int64_t calls = 0, nc = 0;

    for (int oi = 0; oi < n_seeds; ++oi) {
        int si = order[oi];
        int32_t qbeg = sq[si], slen = sl[si], sscore = ss[si];
        int64_t rbeg = sr[si];
        int32_t qend = qbeg + slen;
        int64_t rend = rbeg + slen;
        int cover = -1;                                        /* :172-189 */
        for (int r = 0; r < nreg; ++r)
            if (qbeg >= regs[r].qb && qend <= regs[r].qe) { cover = r; break; }
        if (cover >= 0) {
            int32_t s_aln = slen * sc->a;
            if (s_aln > regs[cover].sub) regs[cover].sub = s_aln;
            if (s_aln >= regs[cover].score - tmp) regs[cover].sub_n += 1;
            continue;
        }
        reg_t reg = {0};
        reg.w = band;
        int32_t acc_h0 = sscore;
        int32_t l_score = 0, l_qlen = 0, l_tlen = 0;
        if (qbeg > 0) {                                        /* :206-262 */
            int lq = qbeg;
            for (int i = 0; i < lq; ++i) qbuf[i] = read[lq - 1 - i];
            int64_t tl64 = rbeg - rmax0 < 10000 ? rbeg - rmax0 : 10000;
            int tl = (int)tl64;
            if (tl > 0) {
                int at = get_ref(ref, ref_len, gl, rbeg - tl, tl, tbuf);
                for (int lo = 0, hi = at - 1; lo < hi; ++lo, --hi) { uint8_t t = tbuf[lo]; tbuf[lo] = tbuf[hi]; tbuf[hi] = t; }
                int32_t w = band, prev = sscore;
                or_swres_t r = {0};
                for (int tr = 0; tr < 2; ++tr) {
                    nc += or_ksw_extend(lq, qbuf, at, tbuf, sc, w, sscore, sc->pen_clip5, &r);
                    calls++;
                    if (log) log(log_ctx, qbuf, lq, tbuf, at, sscore, w, sc->pen_clip5, &r);
                    if (r.score == prev || r.max_off < (w >> 1) + (w >> 2) || tr == 1) break;
                    prev = r.score;
                    w = w << 1;
                }
                acc_h0 = r.score;
                reg.w = reg.w > w ? reg.w : w;
                if (r.gscore <= 0 || r.gscore <= r.score - sc->pen_clip5) {
                    l_qlen = r.qle; l_tlen = r.tle; l_score = r.score;
                } else {
                    l_qlen = qbeg; l_tlen = r.gtle; l_score = r.gscore;
                }
            }
        }
        int32_t r_score = 0, r_qlen = 0, r_tlen = 0;
        if (qend < read_len) {                                 /* :269-316 */
            int rq = read_len - qend;
            for (int i = 0; i < rq; ++i) qbuf[i] = read[qend + i];
            int64_t tr64 = rmax1 - rend < 10000 ? rmax1 - rend : 10000;
            int tr_len = (int)tr64;
            if (tr_len > 0) {
                int at = get_ref(ref, ref_len, gl, rend, tr_len, tbuf);
                int32_t w = band, prev = acc_h0;
                or_swres_t r = {0};
                for (int tr = 0; tr < 2; ++tr) {
                    nc += or_ksw_extend(rq, qbuf, at, tbuf, sc, w, acc_h0, sc->pen_clip3, &r);
                    calls++;
                    if (log) log(log_ctx, qbuf, rq, tbuf, at, acc_h0, w, sc->pen_clip3, &r);
                    if (r.score == prev || r.max_off < (w >> 1) + (w >> 2) || tr == 1) break;
                    prev = r.score;
                    w = w << 1;
                }
                reg.w = reg.w > w ? reg.w : w;
                if (r.gscore <= 0 || r.gscore <= r.score - sc->pen_clip3) {
                    r_qlen = r.qle; r_tlen = r.tle; r_score = r.score;
                } else {
                    r_qlen = read_len - qend; r_tlen = r.gtle; r_score = r.gscore;
                }
            }
        }
        reg.qb = qbeg - l_qlen;                                /* :318-331 */
        reg.qe = qend + r_qlen;
        reg.rb = rbeg - l_tlen;
        reg.re = rend + r_tlen;
        int32_t ts = slen * sc->a;
        if (l_score > 0) ts = l_score;
        if (r_score > 0) ts += r_score - acc_h0;
        reg.score = ts;
        reg.true_score = ts;
        int32_t cov = 0;                                       /* :333, computeSeedCov */
        for (int k = 0; k < n_seeds; ++k)
            if (sq[k] >= reg.qb && sq[k] + sl[k] <= reg.qe && sr[k] >= reg.rb && sr[k] + sl[k] <= reg.re)
                cov += sl[k];
        reg.seed_cov = cov;
        regs[nreg++] = reg;
    }
    int32_t sub_thr = sc->min_seed_len * sc->a;                /* :338-343 */
    for (int r = 0; r < nreg; ++r)
        if (regs[r].sub > 0 && regs[r].sub < sub_thr) regs[r].sub = 0;
    int nout = nreg < max_regions ? nreg : max_regions;
    for (int r = 0; r < nout; ++r) {
        int64_t *o = regions_out + (size_t)r * 10;
        o[0] = regs[r].qb; o[1] = regs[r].qe; o[2] = regs[r].rb; o[3] = regs[r].re;
        o[4] = regs[r].score; o[5] = regs[r].true_score; o[6] = regs[r].sub;
        o[7] = regs[r].sub_n; o[8] = regs[r].w; o[9] = regs[r].seed_cov;
    }
    if (ext_calls) *ext_calls = calls;
    if (cells) *cells = nc;
    free(order); free(regs); free(qbuf); free(tbuf);
    return nreg;
}

/* ---------------------------------------------------------------------- */
/* Threaded batch drivers (blocks of 1024 jobs from an atomic counter).     */
/* ---------------------------------------------------------------------- */
typedef struct {
    int kind, n;
    const uint8_t *seqs;
    const uint64_t *q_off, *t_off;
    const int32_t *q_len, *t_len, *h0, *w, *eb, *ts;
    int max_retries, min_sub, cap;
    const or_scoring_t *sc;
    int32_t *out6, *score, *n_cigar, *ok;
    uint32_t *cigar;
    atomic_int next;
    atomic_llong cells;
} batch_t;

static void *batch_worker(void *arg)
{
    batch_t *b = (batch_t *)arg;
    int64_t cells = 0;
    for (;;) {
        int s = atomic_fetch_add(&b->next, 1024);
        if (s >= b->n) break;
        int e = s + 1024 < b->n ? s + 1024 : b->n;
        for (int i = s; i < e; ++i) {
            const uint8_t *q = b->seqs + b->q_off[i], *t = b->seqs + b->t_off[i];
            if (b->kind == 0) {
                or_swres_t r;
                cells += or_ksw_extend(b->q_len[i], q, b->t_len[i], t, b->sc, b->w[i], b->h0[i], b->eb[i], &r);
                memcpy(b->out6 + (size_t)i * 6, &r, sizeof r);
            } else if (b->kind == 1) {
                int64_t c = 0;
                int n = or_global_with_retry(q, b->q_len[i], t, b->t_len[i], b->w[i], b->ts[i],
                                             b->max_retries, b->sc, b->score + i,
                                             b->cigar + (size_t)i * b->cap, b->cap, NULL, &c);
                b->n_cigar[i] = n;
                cells += c;
            } else {
                int64_t c = 0;
                b->ok[i] = or_local_sw(q, b->q_len[i], t, b->t_len[i], b->sc, b->min_sub,
                                       b->out6 + (size_t)i * 6, &c);
                cells += c;
            }
        }
    }
    atomic_fetch_add(&b->cells, cells);
    return NULL;
}

static int64_t run_batch(batch_t *b, int nthreads)
{
    atomic_init(&b->next, 0);
    atomic_init(&b->cells, 0);
    if (nthreads < 1) nthreads = 1;
    if (nthreads == 1) { batch_worker(b); return atomic_load(&b->cells); }
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int k = 0; k < nthreads; ++k) pthread_create(&th[k], NULL, batch_worker, b);
    for (int k = 0; k < nthreads; ++k) pthread_join(th[k], NULL);
    free(th);
    return atomic_load(&b->cells);
}

int64_t or_ksw_extend_batch(int n, const uint8_t *seqs, const uint64_t *q_off,
                            const int32_t *q_len, const uint64_t *t_off,
                            const int32_t *t_len, const int32_t *h0, const int32_t *w,
                            const int32_t *end_bonus, const or_scoring_t *sc,
                            int32_t *out6, int nthreads)
{
    batch_t b;
    memset(&b, 0, sizeof b);
    b.kind = 0; b.n = n; b.seqs = seqs; b.q_off = q_off; b.t_off = t_off;
    b.q_len = q_len; b.t_len = t_len; b.h0 = h0; b.w = w; b.eb = end_bonus;
    b.sc = sc; b.out6 = out6;
    return run_batch(&b, nthreads);
}

int64_t or_global_batch(int n, const uint8_t *seqs, const uint64_t *q_off,
                        const int32_t *q_len, const uint64_t *t_off, const int32_t *t_len,
                        const int32_t *w, const int32_t *true_score, int max_retries,
                        const or_scoring_t *sc, int32_t *score, uint32_t *cigar, int cap,
                        int32_t *n_cigar, int nthreads)
{
    batch_t b;
    memset(&b, 0, sizeof b);
    b.kind = 1; b.n = n; b.seqs = seqs; b.q_off = q_off; b.t_off = t_off;
    b.q_len = q_len; b.t_len = t_len; b.w = w; b.ts = true_score;
    b.max_retries = max_retries; b.sc = sc; b.score = score; b.cigar = cigar;
    b.cap = cap; b.n_cigar = n_cigar;
    return run_batch(&b, nthreads);
}

int64_t or_local_batch(int n, const uint8_t *seqs, const uint64_t *q_off,
                       const int32_t *q_len, const uint64_t *t_off, const int32_t *t_len,
                       int min_sub_score, const or_scoring_t *sc, int32_t *out6,
                       int32_t *ok, int nthreads)
{
    batch_t b;
    memset(&b, 0, sizeof b);
    b.kind = 2; b.n = n; b.seqs = seqs; b.q_off = q_off; b.t_off = t_off;
    b.q_len = q_len; b.t_len = t_len; b.min_sub = min_sub_score; b.sc = sc;
    b.out6 = out6; b.ok = ok;
    return run_batch(&b, nthreads);
}
