/*
 * CPU oracle for the swiftbwa Smith-Waterman hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, line for line, the CPU
 * algorithms of swiftbwa's `Alignment` module (Swift) in plain C so that the
 * HIP kernels can be checked for bit-exact parity.  Only the test suite, the
 * smoke check and the CPU-baseline leg of the benchmark load it; the product
 * library (libhipsw) never links or calls it.
 *
 * Reference files restated (paths relative to the swiftbwa source tree):
 *   Sources/Alignment/BandedSWScalar.swift     -> or_ksw_extend
 *   Sources/Alignment/GlobalAligner.swift      -> or_ksw_global
 *   Sources/Alignment/CIGARGenerator.swift     -> or_cigar_generate
 *   Sources/Alignment/CIGARBuilder.swift       -> CIGAR packing (len<<4 | op)
 *   Sources/Alignment/LocalSWAligner.swift     -> or_local_sw
 *   Sources/Alignment/ExtensionAligner.swift   -> or_extend_chain
 *   Sources/BWACore/ScoringParameters.swift    -> or_scoring_default / matrix
 *
 * Parity status of each function (what pins it) is listed in the header
 * comment of sw_oracle.c.
 */
#ifndef SW_ORACLE_H
#define SW_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ScoringParameters.swift:2-45 (only the fields the hot path reads). */
typedef struct {
    int32_t a;          /* matchScore            (default 1)   */
    int32_t b;          /* mismatchPenalty       (default 4)   */
    int32_t o_ins;      /* gapOpenPenalty        (default 6)   */
    int32_t e_ins;      /* gapExtendPenalty      (default 1)   */
    int32_t o_del;      /* gapOpenPenaltyDeletion   (6)        */
    int32_t e_del;      /* gapExtendPenaltyDeletion (1)        */
    int32_t w;          /* bandWidth             (100)         */
    int32_t zdrop;      /* zDrop                 (100)         */
    int32_t pen_clip5;  /* penClip5              (5)           */
    int32_t pen_clip3;  /* penClip3              (5)           */
    int32_t min_seed_len; /* minSeedLength       (19)          */
    int8_t  mat[25];    /* scoringMatrix(), ScoringParameters.swift:65-73 */
} or_scoring_t;

/* SWResult.swift:2-25, field order = MetalSW result row order. */
typedef struct {
    int32_t score, qle, tle, gtle, gscore, max_off;
} or_swres_t;

void or_scoring_default(or_scoring_t *sc);
void or_scoring_fill_matrix(or_scoring_t *sc);

/* BandedSWScalar.align.  Returns the number of inner-loop cell iterations
 * (BandedSWScalar.swift:112) executed. */
int64_t or_ksw_extend(int qlen, const uint8_t *query, int tlen, const uint8_t *target,
                      const or_scoring_t *sc, int32_t w, int32_t h0, int32_t end_bonus,
                      or_swres_t *out);

/* GlobalAligner.align + traceback.  Writes up to `cap` packed CIGAR ops.
 * Returns the number of ops the traceback produced (may exceed cap; only
 * min(n, cap) are written).  *score receives h[qlen].  *cells receives the
 * fill-loop iterations (GlobalAligner.swift:108). */
int or_ksw_global(int qlen, const uint8_t *query, int tlen, const uint8_t *target,
                  const or_scoring_t *sc, int32_t w, int32_t *score,
                  uint32_t *cigar, int cap, int64_t *cells);

/* CIGARGenerator.generate.  Writes cigar (<= cap ops), md (NUL-terminated,
 * <= md_cap bytes incl. NUL), and returns the number of ops. */
typedef struct {
    int32_t score;
    int32_t nm;
    int64_t pos;
    int32_t n_cigar;
    int32_t nw_calls;   /* GlobalAligner.align invocations (0 on the fast path) */
    int64_t cells;      /* NW fill cells over all invocations */
} or_cigar_res_t;

int or_cigar_generate(const uint8_t *qseg, int qlen, const uint8_t *rseg, int rlen,
                      int32_t qb, int32_t qe, int32_t read_len, int is_reverse,
                      int32_t true_score, int32_t initial_w, const or_scoring_t *sc,
                      int64_t ref_pos, uint32_t *cigar, int cap, char *md, int md_cap,
                      or_cigar_res_t *res);

/* GPU-side NW dispatch semantics: generate()'s band choice, NW, retry loop,
 * but no fast path and no post-processing.  Mirrors what hipsw_global_nw
 * returns (score, raw CIGAR, final w). */
int or_global_with_retry(const uint8_t *q, int qlen, const uint8_t *t, int tlen,
                         int32_t w0, int32_t true_score, int max_retries,
                         const or_scoring_t *sc, int32_t *score, uint32_t *cigar,
                         int cap, int32_t *final_w, int64_t *cells);

/* LocalSWAligner.align.  Returns 1 and fills out[6] = {score, qb, qe, tb, te,
 * score2} on success; returns 0 where the reference returns nil. */
int or_local_sw(const uint8_t *query, int qlen, const uint8_t *target, int tlen,
                const or_scoring_t *sc, int min_sub_score, int32_t *out6,
                int64_t *cells);

/* ExtensionAligner.extend (one chain).  Seeds are given as SoA arrays.
 * `ref` is a flat forward reference of length ref_len; getReference(pos,len)
 * returns ref[max(0,pos) .. min(ref_len, pos+len)).  Regions are written as
 * rows of 10 int64: {qb, qe, rb, re, score, true_score, sub, sub_n, w, seed_cov}.
 * Returns the number of regions. */
int or_extend_chain(int n_seeds, const int32_t *seed_qbeg, const int32_t *seed_len,
                    const int64_t *seed_rbeg, const int32_t *seed_score,
                    const uint8_t *read, int read_len,
                    const uint8_t *ref, int64_t ref_len,
                    const or_scoring_t *sc, int64_t *regions_out, int max_regions,
                    int64_t *ext_calls, int64_t *cells);

/* or_extend_chain with a choice of getReference semantics (gl > 0: the
 * FMIndex form over the forward genome `ref` of length gl, positions >= gl on
 * the reverse strand; gl == 0: the flat form above) and an optional sink that
 * sees every BandedSWScalar.align call: query, target window, h0, w, end
 * bonus and the result. */
typedef void (*or_ext_log_fn)(void *ctx, const uint8_t *q, int qlen, const uint8_t *t, int tlen,
                              int32_t h0, int32_t w, int32_t end_bonus, const or_swres_t *r);
int or_extend_chain_ex(int n_seeds, const int32_t *seed_qbeg, const int32_t *seed_len,
                       const int64_t *seed_rbeg, const int32_t *seed_score,
                       const uint8_t *read, int read_len,
                       const uint8_t *ref, int64_t ref_len, int64_t gl, const or_scoring_t *sc,
                       int64_t *regions_out, int max_regions, int64_t *ext_calls, int64_t *cells,
                       or_ext_log_fn log, void *log_ctx);

/* Batch entry points used by the tests and the CPU baseline.  Sequences are
 * one byte per base (0..4) in one arena; offsets are byte offsets. */
int64_t or_ksw_extend_batch(int n, const uint8_t *seqs, const uint64_t *q_off,
                            const int32_t *q_len, const uint64_t *t_off,
                            const int32_t *t_len, const int32_t *h0, const int32_t *w,
                            const int32_t *end_bonus, const or_scoring_t *sc,
                            int32_t *out6, int nthreads);

int64_t or_global_batch(int n, const uint8_t *seqs, const uint64_t *q_off,
                        const int32_t *q_len, const uint64_t *t_off, const int32_t *t_len,
                        const int32_t *w, const int32_t *true_score, int max_retries,
                        const or_scoring_t *sc, int32_t *score, uint32_t *cigar, int cap,
                        int32_t *n_cigar, int nthreads);

int64_t or_local_batch(int n, const uint8_t *seqs, const uint64_t *q_off,
                       const int32_t *q_len, const uint64_t *t_off, const int32_t *t_len,
                       int min_sub_score, const or_scoring_t *sc, int32_t *out6,
                       int32_t *ok, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
