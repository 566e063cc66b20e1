/*
 * End-to-end single-end harness of the CPU oracle -- TEST INFRASTRUCTURE.
 *
 * Restates, in plain C, swiftbwa's single-end `mem` path around the
 * Smith-Waterman hot path, so that a read set can be taken through seeding,
 * chaining, extension, region dedup, primary marking, CIGAR generation and SAM
 * fields on the CPU, with the extension and CIGAR phases replaceable by the
 * GPU library's results (imported per phase).  Only the tests load it.
 *
 * Reference files restated (swiftbwa source tree):
 *   Sources/FMIndexBuilder/IndexBuilder.swift:20-110  text = fwd + revcomp, SA, '$' smallest
 *   Sources/FMIndexBuilder/FMIndexWriter.swift:18-60  BWT, sentinel row, count[] + 1
 *   Sources/FMIndex/BWT.swift:58-77                  occ(pos, c) = #c in bwt[0, pos)
 *   Sources/FMIndex/BackwardSearch.swift:22-83        backwardExt
 *   Sources/FMIndex/SMEMFinder.swift:28-446           findAllSMEMs / findSMEMsAtPosition
 *   Sources/FMIndex/FMIndex.swift:25-50               getReference
 *   Sources/BWAMem/BWAMemAligner.swift:116-432        alignRead phases 1-5, internalReseed,
 *                                                     computeFracRep
 *   Sources/Alignment/SeedChainer.swift:16-107        chain
 *   Sources/Alignment/ChainFilter.swift               filter, markSecondary(+Core), chainWeight
 *   Sources/Alignment/RegionDedup.swift               sortDedupPatch, patchRegion
 *   Sources/BWAMem/BWAMemAligner.swift:434-476, 841-921, 923-1099
 *                                                     generateFilteredCIGARs, generateCIGAR,
 *                                                     classifyAlignments
 *   Sources/BWAMem/MappingQuality.swift               compute
 *   Sources/BWAMem/SAMOutputBuilder.swift:103-438     buildRecord, hard clips, SA/XA tags,
 *                                                     buildUnmappedRecord
 *   Sources/BWACore/AlignmentRegion.swift:3-14        hash64
 *   paired end:
 *   Sources/BWAMem/BWAMemAligner.swift:1348-1543, 1857-2159
 *                                                     alignPairedBatch (pair ids, distribution
 *                                                     from the first 4096 pairs), processFullBatch
 *                                                     and rescuePair (CPU rescue order),
 *                                                     selectRescueCandidates, hasMultipleNonSecondary
 *   Sources/BWAMem/BWAMemAligner.swift:1549-1856      promotePairedRegion, preparePairOutput,
 *                                                     prepareMappedUnmappedPlan
 *   Sources/BWAMem/InsertSizeEstimator.swift          estimate, inferOrientation
 *   Sources/BWAMem/MateRescue.swift                   collectCandidates, applyResult, hasMateHit
 *   Sources/FMIndex/PackedReference.swift:72-121      fetchSequence
 *   Sources/BWAMem/PairedEndResolver.swift            resolve, computeTLEN, adjustMAPQ, z penalty
 *   Sources/BWAMem/SAMOutputBuilder.swift:5-57        PairedEndInfo, makePair
 *
 * Scope: one reference sequence without ambiguous bases (the config-1 random
 * reference, the lambda fixture genome, the simulated paired-end genome),
 * default options (flag 0: no -a, -M, -Y, -5, -q, -P, -S; no -y reseeding; no
 * ALT contigs; no manual insert size).
 */
#ifndef E2E_ORACLE_H
#define E2E_ORACLE_H

#include <stdint.h>
#include "sw_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_e2e or_e2e_t;

/* Builds the index of one sequence (bases 0..3, gl of them) named `name`. */
or_e2e_t *or_e2e_create(const uint8_t *fwd, int64_t gl, const char *name, const or_scoring_t *sc);
void or_e2e_free(or_e2e_t *e);
/* Splits the indexed text into n sequences (the index's annotations: offset,
 * length, NUL-separated names) that tile [0, gl); 0, or -1 if they do not. */
int or_e2e_set_contigs(or_e2e_t *e, int n, const int64_t *off, const int64_t *len, const char *names);
/* A handle with its own batch state sharing base's index (base outlives it). */
or_e2e_t *or_e2e_fork(const or_e2e_t *base);
/* ReferenceMetadata.sequenceID of a forward-text position. */
int or_e2e_seq_id(const or_e2e_t *e, int64_t pos);

/* Index inspection for the tests: SA row i (0..2gl), BWT char of row i (4 =
 * '$'), count tuple (5 values), occ(pos, c). */
int64_t or_e2e_sa(const or_e2e_t *e, int64_t i);
int or_e2e_bwt(const or_e2e_t *e, int64_t i);
void or_e2e_counts(const or_e2e_t *e, int64_t *cnt5);
int64_t or_e2e_occ(const or_e2e_t *e, int64_t pos, int c);

/* SMEMFinder.findAllSMEMs of one query: rows {k, l, qb, qe}.  Returns the
 * number of SMEMs (rows written up to cap). */
int or_e2e_smems(const or_e2e_t *e, const uint8_t *q, int qlen, int32_t min_seed_len,
                 int64_t min_intv, int64_t *rows, int cap);

/* Phases 1-3 (SMEMs, internal reseeding, sort, frac_rep, chaining, chain
 * filter) of a batch of reads (bases 0..4 in one arena).  Keeps the reads
 * and chains in the context; returns the number of chains. */
int64_t or_e2e_seed(or_e2e_t *e, int n_reads, const uint8_t *bases, const int64_t *off,
                    const int32_t *len);
int64_t or_e2e_n_seeds(const or_e2e_t *e);
/* chain_read[c], chain_off[c..c+1] = its seeds (rbeg, qbeg, len, score). */
void or_e2e_get_chains(const or_e2e_t *e, int32_t *chain_read, int64_t *chain_off,
                       int64_t *rbeg, int32_t *qbeg, int32_t *slen, int32_t *sscore);

/* Phase 4 on the CPU: ExtensionAligner.extend of every chain.  Records every
 * BandedSWScalar.align call (the call set).  Returns the number of regions. */
int64_t or_e2e_extend_cpu(or_e2e_t *e);
/* Phase 4 imported: n_per_chain[c] regions per chain, rows of 10 int64
 * {qb, qe, rb, re, score, true_score, sub, sub_n, w, seed_cov}. */
void or_e2e_set_regions(or_e2e_t *e, const int32_t *n_per_chain, const int64_t *rows);
/* Call set of the last or_e2e_extend_cpu: n calls; q/t bytes in arenas. */
int64_t or_e2e_n_calls(const or_e2e_t *e, int64_t *q_bytes, int64_t *t_bytes);
void or_e2e_get_calls(const or_e2e_t *e, uint8_t *q, int64_t *q_off, int32_t *q_len,
                      uint8_t *t, int64_t *t_off, int32_t *t_len, int32_t *h0, int32_t *w,
                      int32_t *end_bonus, int32_t *out6);

/* Phases 4.5-5 (sortDedupPatch, markSecondary) per read, then the CIGAR job
 * list of generateFilteredCIGARs.  Returns the number of CIGAR jobs. */
int64_t or_e2e_finish(or_e2e_t *e);
void or_e2e_cigar_sizes(const or_e2e_t *e, int64_t *q_bytes, int64_t *r_bytes);
/* job j: query segment, reference segment, qb, qe, read length, reverse flag,
 * true score, initial w, reference position. */
void or_e2e_get_cigar_jobs(const or_e2e_t *e, uint8_t *qseg, int64_t *q_off, int32_t *q_len,
                           uint8_t *rseg, int64_t *r_off, int32_t *r_len, int32_t *qb,
                           int32_t *qe, int32_t *read_len, uint8_t *rev, int32_t *true_score,
                           int32_t *w, int64_t *ref_pos);
/* CIGAR generation of every job on the CPU (or_cigar_generate). */
void or_e2e_cigar_cpu(or_e2e_t *e);
/* Imported CIGAR of job j (packed ops, NUL-terminated MD). */
void or_e2e_set_cigar(or_e2e_t *e, int64_t j, const uint32_t *cigar, int n_cigar, int32_t nm,
                      const char *md, int64_t pos);

/* SAM records of the batch, one line each, in read order: names are
 * NUL-separated, qualities are Phred+33 text in the bases' arena layout.
 * Returns the text length (written up to cap bytes, NUL-terminated when it
 * fits). */
int64_t or_e2e_sam(const or_e2e_t *e, const char *names, const uint8_t *quals, char *out,
                   int64_t cap);

/* ---- paired end (reads [0, n) are read 1 of pairs 0..n-1, [n, 2n) read 2) ---- */

/* Phases 4.5-5 per read with readId = pair << 1 | mate, then the insert-size
 * distribution of the first min(4096, n) pairs.  Returns 1 when rescue runs
 * (the primary orientation's estimate did not fail), 0 when not, -1 when the
 * batch does not hold 2n reads. */
int64_t or_e2e_pe_finish(or_e2e_t *e, int n_pairs);
/* per orientation FF FR RF RR: failed, low, high, mean, stddev, properLow,
 * properHigh, count; and the primary orientation */
void or_e2e_pe_dist(const or_e2e_t *e, double *out8x4, int32_t *primary);
/* Rescue candidates of direction 1 (read 2 from read 1's regions) or 2 (read 1
 * from read 2's regions, rescued ones included): local-SW jobs (mate sequence,
 * reference window).  Returns the count. */
int64_t or_e2e_pe_rescue_jobs(or_e2e_t *e, int dir);
void or_e2e_pe_rescue_sizes(const or_e2e_t *e, int64_t *q_bytes, int64_t *t_bytes);
void or_e2e_pe_get_rescue_jobs(const or_e2e_t *e, uint8_t *q, int64_t *q_off, int32_t *q_len,
                               uint8_t *t, int64_t *t_off, int32_t *t_len);
/* LocalSWAligner.align of the pending candidates on the CPU, or imported
 * results (n x {score, qb, qe, tb, te, score2} and the success flags). */
void or_e2e_pe_rescue_cpu(or_e2e_t *e);
void or_e2e_pe_set_rescue(or_e2e_t *e, const int32_t *out6, const int32_t *ok);
/* Raw candidates (windows before PackedReference.fetchSequence, BWT-space
 * [rb, re), with the template's sequence id); the caller clamps and fetches,
 * then passes each candidate's clampedBegin and its local-SW result (ok 0 for
 * a window the clamp dropped). */
int64_t or_e2e_pe_rescue_jobs_raw(or_e2e_t *e, int dir);
void or_e2e_pe_get_rescue_windows(const or_e2e_t *e, int64_t *rb, int64_t *re, int32_t *rid);
void or_e2e_pe_set_rescue_windows(or_e2e_t *e, const int64_t *clamped);
/* MateRescue.applyResult of the pending candidates; returns regions added. */
int64_t or_e2e_pe_apply(or_e2e_t *e);
/* markSecondary after rescue (only when rescue runs). */
void or_e2e_pe_remark(or_e2e_t *e);
/* CIGAR job list over all 2n reads (generateFilteredCIGARs). */
int64_t or_e2e_pe_cigar_plan(or_e2e_t *e);
/* SAM records of the pairs in pair order (one name per pair). */
int64_t or_e2e_pe_sam(or_e2e_t *e, const char *names, const uint8_t *quals, char *out,
                      int64_t cap);

#ifdef __cplusplus
}
#endif
#endif
