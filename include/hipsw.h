/*
 * hipsw.h -- C ABI of the MI355X (gfx950) Smith-Waterman engine that replaces
 * swiftbwa's MetalSW module behind its Swift call surface.
 *
 * Exposed to Swift as a system-library target (module map beside this header),
 * in the style of swiftbwa's CHelpers target (Package.swift:54-58,
 * Sources/CHelpers/include/CHelpers.h).
 *
 * Conventions
 *  - The caller owns all host memory; nothing returned aliases engine memory.
 *  - Status 0 = success.  A negative status means "use the CPU path": the Swift
 *    shim maps it to nil / [] exactly as MetalSW's dispatchers signal failure
 *    (BandedSWDispatcher.swift:70,83,129,142; GlobalSWDispatcher.swift:81).
 *  - Sequences are one byte per base, 0..4 (A C G T N), the layout swiftbwa
 *    holds (Base.swift:2-9).  Queries and targets are packed the way MetalSW's
 *    QueryTargetPacker packs them (MetalDispatchHelper.swift:30-62): one byte
 *    arena each, u32 offsets, u16 lengths.  The engine packs to 2 bits + an N
 *    mask before the upload.
 *  - Every call is thread-safe.  The one-shot entry points (hipsw_banded_extend,
 *    hipsw_global_nw, hipsw_local_sw, hipsw_cigar_generate, hipsw_extend_chains)
 *    each lease one of the engine's stream contexts (stream, events, counters,
 *    grow-only device buffers) for the call, so calls from different host threads
 *    pack, upload, run and copy back concurrently; a call waits only when every
 *    context is leased.  Calls on device-resident sets (hipsw_ext_* / hipsw_nw_*)
 *    and the autotuner run on the engine's own stream under its mutex.
 *  - hipsw_banded_extend, hipsw_banded_extend_noalloc and hipsw_ext_upload
 *    pack and plan into host buffers the engine keeps for reuse (at most four,
 *    grow-only, about 190 bytes per job of the largest call; freed by
 *    hipsw_engine_destroy), so repeated calls do not fault in fresh host
 *    memory.  The arrays copied to the device (about 140 bytes per 150-base
 *    job) are page-locked (hipHostMalloc) when the runtime grants it, so their
 *    upload is a direct DMA; blocks above 1 GiB stay pageable.
 *  - hipsw_last_error reports the calling thread's last failure on the engine
 *    (errno-like), so concurrent callers do not see each other's messages.
 */
#ifndef HIPSW_H
#define HIPSW_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define HIPSW_OK              0
#define HIPSW_E_INVALID      -1   /* bad argument (null pointer, bad scoring) */
#define HIPSW_E_DEVICE       -2   /* HIP runtime error (alloc, launch, copy)  */
#define HIPSW_E_UNSUPPORTED  -3   /* input outside what the kernels handle    */
#define HIPSW_E_BUDGET       -4   /* request exceeds the engine's HBM budget  */
#define HIPSW_E_PENDING      -5   /* ticket not finished (poll)               */

/* Dispatch modes of the extension and NW entry points. */
#define HIPSW_MODE_PLAIN      0   /* one launch per call, input order            */
#define HIPSW_MODE_BATCHED    1   /* area-binned persistent launch, tuned per bin */
#define HIPSW_MODE_INPUT_ORDER 2  /* hipsw_ext_run only: the batched form's kernel over the
                                     jobs in input order, one launch, one work item per
                                     block -- no binning, sorting or work stealing (a
                                     measurement baseline) */

typedef struct hipsw_engine hipsw_engine;

/* Scoring: the 5x5 matrix of ScoringParameters.scoringMatrix()
 * (ScoringParameters.swift:65-73) plus the gap penalties and z-drop.
 * `mat` is indexed [target_base * 5 + query_base] (BandedSWScalar.swift:45-50). */
typedef struct {
    int8_t  mat[25];
    int32_t o_ins, e_ins;     /* gapOpenPenalty, gapExtendPenalty            */
    int32_t o_del, e_del;     /* gapOpenPenaltyDeletion, gapExtendPenaltyDeletion */
    int32_t zdrop;            /* zDrop                                        */
} hipsw_scoring;

/* Per-call timing/counters, filled when the caller passes a non-NULL pointer. */
typedef struct {
    double  kernel_ms;        /* device time of the DP kernels (HIP events)   */
    double  total_ms;         /* wall time of the call incl. pack/copies      */
    int64_t n_jobs;
    int64_t launches;         /* kernel launches issued                        */
    int64_t promoted;         /* jobs run in a wider score tier than int16     */
    int64_t tier_jobs[3];     /* jobs per score tier: u8, u16, i32             */
    int64_t cells;            /* DP cells evaluated, counted as the CPU loop    */
                              /* counts them (BandedSWScalar.swift:112)         */
} hipsw_stats;

/* ---- engine (replaces MetalSWEngine.shared, MetalSWEngine.swift:8-17) ---- */

/* Create an engine on HIP device `device`.  NULL when no usable device exists
 * or the kernels cannot be loaded -- the same "engine is nil" signal MetalSW
 * gives (MetalSWEngine.swift:8-17); callers then stay on the CPU. */
hipsw_engine *hipsw_engine_create(int device);
void          hipsw_engine_destroy(hipsw_engine *eng);
/* Text of the calling thread's last error on this engine (valid until that
 * thread's next call on it). */
const char   *hipsw_last_error(const hipsw_engine *eng);
/* MetalSWEngine.minBatchSize (MetalSWEngine.swift:39). */
int32_t       hipsw_min_batch_size(void);
/* Fault injection, for testing callers' fallbacks (no reference counterpart).
 * Arms one fault: the countdown-th device allocation (site 1), kernel launch
 * (site 2) or host<->device copy (site 3) that a call on this engine makes from
 * now on fails as if the HIP runtime had refused it; the call then returns
 * HIPSW_E_DEVICE with "injected" in hipsw_last_error, after draining its
 * stream, and the engine stays usable.  The fault disarms once it fires;
 * countdown 0 (or site 0) disarms it. */
#define HIPSW_FAULT_NONE   0
#define HIPSW_FAULT_ALLOC  1
#define HIPSW_FAULT_LAUNCH 2
#define HIPSW_FAULT_COPY   3
int           hipsw_fault_inject(hipsw_engine *eng, int32_t site, int32_t countdown);
/* Device name and compute-unit count, for reports. */
int           hipsw_device_info(const hipsw_engine *eng, char *name, int name_cap,
                                int32_t *n_cu);

/* ---- banded extension (replaces BandedSWDispatcher.dispatchBatch8 and
 *      dispatchBatch16, BandedSWDispatcher.swift:48-158) ----
 * Semantics are BandedSWScalar.align (BandedSWScalar.swift:21-191) per task,
 * with per-task h0, w and end bonus (the Metal path took w from tasks[0] only,
 * BandedSWDispatcher.swift:61, and had no end bonus).  out is n rows of six
 * int32 in SWResult order {score, queryEnd, targetEnd, globalTargetEnd,
 * globalScore, maxOff} (SWResult.swift:2-25, BandedSWDispatcher.swift:92-99).
 * Scores are exact in every case: the engine picks an 8-, 16- or 32-bit score
 * tier per task from the bound h0 + qlen * max(mat).  Every w must be >= 0
 * (HIPSW_E_INVALID otherwise): a negative band makes the scalar loop index its
 * row below column 0 (BandedSWScalar.swift:97-99, :145). */
int hipsw_banded_extend(hipsw_engine *eng, int32_t n,
                        const uint8_t *queries, const uint32_t *q_off, const uint16_t *q_len,
                        const uint8_t *targets, const uint32_t *t_off, const uint16_t *t_len,
                        const int32_t *h0, const int32_t *w, const int32_t *end_bonus,
                        const hipsw_scoring *sc, int32_t mode, int32_t *out,
                        hipsw_stats *stats);

/* Allocation-free variant of hipsw_banded_extend: same arguments, same
 * results.  No counterpart in the reference; it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * hosts that forbid hipMalloc).  Every device buffer beyond the caller's
 * inputs and outputs -- the packed job arrays, the batched order, work items
 * and descriptors, the counters and the HBM column state -- lives in one
 * static device array of 64 MiB per device, held until the process ends; the
 * call makes no hipMalloc/hipFree.  A set whose layout needs more than that
 * is refused with HIPSW_E_BUDGET (hipsw_last_error gives the bytes needed)
 * before anything runs.  Calls on one device share the array: the library
 * serialises them with a per-device mutex, so overlapping calls from
 * several threads wait rather than race. */
int hipsw_banded_extend_noalloc(hipsw_engine *eng, int32_t n,
                                const uint8_t *queries, const uint32_t *q_off,
                                const uint16_t *q_len, const uint8_t *targets,
                                const uint32_t *t_off, const uint16_t *t_len,
                                const int32_t *h0, const int32_t *w, const int32_t *end_bonus,
                                const hipsw_scoring *sc, int32_t mode, int32_t *out,
                                hipsw_stats *stats);

/* Device-resident job sets: upload once, run many times (benchmarks, serving).
 * hipsw_ext_run leaves results on the device; hipsw_ext_download copies them
 * out.  No reference counterpart: MetalSW re-uploads every call. */
typedef struct hipsw_ext_set hipsw_ext_set;
hipsw_ext_set *hipsw_ext_upload(hipsw_engine *eng, int32_t n,
                                const uint8_t *queries, const uint32_t *q_off,
                                const uint16_t *q_len, const uint8_t *targets,
                                const uint32_t *t_off, const uint16_t *t_len,
                                const int32_t *h0, const int32_t *w, const int32_t *end_bonus,
                                const hipsw_scoring *sc);
int  hipsw_ext_run(hipsw_engine *eng, hipsw_ext_set *set, int32_t mode, hipsw_stats *stats);

/* Autotuner of the batched form (no reference counterpart).  For each bin of
 * the set in the 8-bit score tier (one query-length class each) it times the
 * candidate kernels on the bin's first max_jobs jobs -- the packed layout (two
 * matrices per lane, packed int16 arithmetic) with each built split of the
 * column state between LDS and HBM (lds_blocks: 16-column blocks in LDS; fewer
 * blocks, more waves per CU), then one matrix per lane (int32 arithmetic) --
 * and for each bin of the 16-bit tier the one-matrix-per-lane kernel at each
 * built split; keeps the fastest per (tier, class) in the engine (later
 * uploads, the server and the chain driver plan with it) and re-plans the set.
 * Results do not change, only speed.  report receives up to report_cap
 * entries, one per (bin, candidate): tier (0 = 8-bit, 1 = 16-bit),
 * query-length class (0: 0-63, 1: 64-159, 2: 160-255),
 * jobs timed, kind (1 packed, 2 one per lane), lds_blocks, chosen (1 for the
 * winner of its bin) and the best of two timed runs in ms. */
typedef struct {
    int32_t tier, length_class, jobs, kind, lds_blocks, chosen;
    double ms;
} hipsw_tune_entry;
int  hipsw_ext_autotune(hipsw_engine *eng, hipsw_ext_set *set, int32_t max_jobs,
                        hipsw_tune_entry *report, int32_t report_cap, int32_t *n_report);
/* Sets the engine's layout for one query-length class of the 8-bit tier, as
 * the autotuner would (no reference counterpart): lets a host restore tuned
 * choices without timing again.  kind 0 restores the default (packed, column
 * state in LDS up to 160 columns); kind 1 (packed) or 2 (one job per lane)
 * with lds_blocks among the splits built for the class (class 0: 4; class 1:
 * 10, 6, 5; class 2: 10, 8, 6; one job per lane the same).  Applies to
 * plans made afterwards.  HIPSW_E_INVALID for a layout not built. */
int  hipsw_ext_set_layout(hipsw_engine *eng, int32_t length_class, int32_t kind,
                          int32_t lds_blocks);
/* The same for the 16-bit tier with the one-job-per-lane kernel (no
 * reference counterpart): lds_blocks among the splits built for the class
 * (class 0: 4; class 1: 10, 6, 5; class 2: 10, 8, 6), 0 for the default (all
 * column state in LDS up to 160 columns).  Applies to plans made afterwards.
 * HIPSW_E_INVALID for a layout not built. */
int  hipsw_ext_set_layout16(hipsw_engine *eng, int32_t length_class, int32_t lds_blocks);
/* The 16-bit tier's layout with its kernel kind (no reference counterpart):
 * kind 0 restores the default (one job per lane), 2 = one job per lane with
 * lds_blocks as above, 3 = two jobs per lane in 16-bit halves (h and e
 * planes) with lds_blocks among 4, 2 (class 0) or 5, 3, 2 (classes 1, 2), 0
 * for its default.  Kind 3 runs a bin only when every job's score bound
 * h0 + qlen * max(mat) is below 2048 and e_del, e_ins <= 31; other bins keep
 * the one-job-per-lane kernel.  Applies to plans made afterwards. */
int  hipsw_ext_set_layout16_kind(hipsw_engine *eng, int32_t length_class, int32_t kind,
                                 int32_t lds_blocks);
/* The batched plan of a set, one entry per bin (no reference counterpart):
 * tier (0 8-bit, 1 16-bit, 2 32-bit), query-length class, has_n, jobs, kind
 * (0 one job per lane, generic; 1 packed; 2 one job per lane, register query;
 * 3 packed, 16-bit tier),
 * lds_blocks (16-column blocks of column state in LDS, 0 for the generic
 * kind), work items and grid.  Lets a host check which layout a pinned or
 * tuned choice gave each bin. */
typedef struct {
    int32_t tier, length_class, has_n, jobs, kind, lds_blocks, items, grid;
} hipsw_plan_entry;
int  hipsw_ext_plan(hipsw_engine *eng, const hipsw_ext_set *set, hipsw_plan_entry *entries,
                    int32_t cap, int32_t *n_entries);
int  hipsw_ext_download(hipsw_engine *eng, hipsw_ext_set *set, int32_t *out);
void hipsw_ext_free(hipsw_engine *eng, hipsw_ext_set *set);

/* ---- banded global NW + traceback (replaces GlobalSWDispatcher.dispatchBatch,
 *      GlobalSWDispatcher.swift:6-118) ----
 * Per task, the NW loop of CIGARGenerator.generate (CIGARGenerator.swift:82-96):
 * GlobalAligner.align (GlobalAligner.swift:19-232) at band w[k]; while the
 * score is below true_score[k] and fewer than max_retries retries were made,
 * w doubles and the alignment is redone.  Unlike the Metal kernel, trueScore
 * and w are not clamped to int16 (GlobalSWDispatcher.swift:55-56) and the band
 * widening follows the CPU (w' = max(w, |qlen-tlen|) inside every attempt).
 * Results: out_score[k]; the packed CIGAR (len<<4 | op, M0 I1 D2) in
 * out_cigar[k*cap ...], out_n_cigar[k] ops (the traceback's full count, which
 * may exceed cap); out_flags[k] bit 0 = more ops than cap (the caller redoes
 * that task on the CPU); out_final_w[k] (nullable) = band of the last attempt. */
int hipsw_global_nw(hipsw_engine *eng, int32_t n,
                    const uint8_t *queries, const uint32_t *q_off, const uint16_t *q_len,
                    const uint8_t *targets, const uint32_t *t_off, const uint16_t *t_len,
                    const int32_t *w, const int32_t *true_score, int32_t max_retries,
                    int32_t cap, const hipsw_scoring *sc, int32_t mode,
                    int32_t *out_score, uint32_t *out_cigar, int32_t *out_n_cigar,
                    int32_t *out_flags, int32_t *out_final_w, hipsw_stats *stats);

/* ---- CIGAR generation (replaces CIGARGenerator.generate, CIGARGenerator.swift:
 *      30-111, and its postProcess, :117-142, for a batch of regions) ----
 * Region k: query segment qsegs[q_off[k] ..+ q_len[k]) (already reverse-
 * complemented for reverse-strand regions) and reference segment
 * rsegs[r_off[k] ..+ r_len[k]), one byte per base 0-4; qb/qe/read_len/
 * is_reverse/true_score/initial_w/ref_pos as generate() takes them.
 * match_score / mismatch_penalty are ScoringParameters.matchScore /
 * mismatchPenalty (inferBandWidth and the fast path use them).  The band is
 * max(inferBandWidth, initialW); equal-length regions whose score deficit is
 * below the indel cost take the fast path (lenM, score = trueScore); the rest
 * run NW with up to 3 doublings on the device.  Outputs per region: score, NM,
 * pos (refPos plus squeezed leading deletions), the packed CIGAR with soft
 * clips in out_cigar[k*cap ...] (out_n_cigar[k] = full op count), the MD tag
 * NUL-terminated in out_md[k*md_cap ...].  out_flags[k]: bit 0 = more than cap
 * ops (redo the region on the CPU), bit 1 = MD truncated to md_cap - 1 chars. */
int hipsw_cigar_generate(hipsw_engine *eng, int32_t n,
                         const uint8_t *qsegs, const uint32_t *q_off, const uint16_t *q_len,
                         const uint8_t *rsegs, const uint32_t *r_off, const uint16_t *r_len,
                         const int32_t *qb, const int32_t *qe, const int32_t *read_len,
                         const uint8_t *is_reverse, const int32_t *true_score,
                         const int32_t *initial_w, const int64_t *ref_pos,
                         const hipsw_scoring *sc, int32_t match_score, int32_t mismatch_penalty,
                         int32_t cap, int32_t md_cap, int32_t *out_score, int32_t *out_nm,
                         int64_t *out_pos, uint32_t *out_cigar, int32_t *out_n_cigar,
                         char *out_md, int32_t *out_flags, hipsw_stats *stats);

/* The same generator in two host steps around an NW the caller runs (the
 * serving loop's NW jobs, targets fetched from the resident .pac).
 * hipsw_cigar_route writes each region's band (out_w) and route: 0 the fast
 * path, 1 NW (query segment vs reference segment, band out_w, up to 3
 * doublings while the score is below true_score, as hipsw_global_nw with
 * max_retries 3 runs it), 2 an empty segment.  hipsw_cigar_finish takes the
 * NW results of the route-1 regions in region order -- score, the used ops
 * back to back at nw_cigar[nw_cig_off[s] ..), the full op count, the flags,
 * and the op cap the NW ran with -- and writes what hipsw_cigar_generate
 * writes. */
int hipsw_cigar_route(int32_t n, const uint16_t *q_len, const uint16_t *r_len,
                      const int32_t *qb, const int32_t *qe, const int32_t *true_score,
                      const int32_t *initial_w, const hipsw_scoring *sc, int32_t match_score,
                      int32_t mismatch_penalty, int32_t *out_w, uint8_t *out_route);
int hipsw_cigar_finish(int32_t n,
                       const uint8_t *qsegs, const uint32_t *q_off, const uint16_t *q_len,
                       const uint8_t *rsegs, const uint32_t *r_off, const uint16_t *r_len,
                       const int32_t *qb, const int32_t *qe, const int32_t *read_len,
                       const uint8_t *is_reverse, const int32_t *true_score,
                       const int64_t *ref_pos, const uint8_t *route,
                       const int32_t *nw_score, const uint32_t *nw_cigar,
                       const uint32_t *nw_cig_off, const int32_t *nw_n_cigar,
                       const int32_t *nw_flags, int32_t nw_cap, int32_t cap, int32_t md_cap,
                       int32_t *out_score, int32_t *out_nm, int64_t *out_pos,
                       uint32_t *out_cigar, int32_t *out_n_cigar, char *out_md,
                       int32_t *out_flags);

/* ---- chain extension (replaces ExtensionAligner.extend,
 *      ExtensionAligner.swift:114-346, for a batch of chains) ----
 * Chain c owns seeds [chain_seed_off[c], chain_seed_off[c+1]) (MemSeed fields
 * rbeg / qbeg / len / score, Seed.swift:4-19) and read chain_read[c] (bases
 * 0-4 at reads[read_off[r] ..+ read_len[r]]).  Reference windows come from
 * get_ref(ref_ctx, pos, len, buf), the getReference closure of the Swift call
 * (returns the bases written, <= len); hipsw_get_reference_pac is
 * FMIndex.getReference over a .pac.  Results: the regions of chain c in the
 * order the CPU loop appends them, at out[chain_seed_off[c] ...], out_n[c] of
 * them.  Exact with the CPU loop (windows, stable score order, covered-seed
 * skip, left then right with h0 chaining, two band tries, clip or extend,
 * trueScore, seedCov, the sub threshold); rid / isAlt / fracRep are the
 * chain's and stay with the caller. */
typedef struct {
    int64_t rb, re;
    int32_t qb, qe, score, true_score, sub, sub_n, w, seed_cov, seed_len0, reserved;
} hipsw_region;

typedef struct {   /* ScoringParameters fields the driver reads */
    int32_t match_score, mismatch_penalty, band_width, pen_clip5, pen_clip3, min_seed_len;
} hipsw_ext_opts;

typedef int32_t (*hipsw_get_reference_fn)(void *ctx, int64_t pos, int32_t len, uint8_t *buf);

typedef struct {   /* context of hipsw_get_reference_pac */
    const uint8_t *pac;     /* 2-bit forward strand, first base in bits 7-6 */
    int64_t genome_len;     /* forward length; positions >= it read the reverse strand */
} hipsw_pac_ref;

int32_t hipsw_get_reference_pac(void *ctx, int64_t pos, int32_t len, uint8_t *buf);

int hipsw_extend_chains(hipsw_engine *eng, int32_t n_chains, const uint32_t *chain_seed_off,
                        const int64_t *seed_rbeg, const int32_t *seed_qbeg,
                        const int32_t *seed_len, const int32_t *seed_score,
                        const uint32_t *chain_read, const uint8_t *reads,
                        const uint32_t *read_off, const int32_t *read_len,
                        hipsw_get_reference_fn get_ref, void *ref_ctx, const hipsw_scoring *sc,
                        const hipsw_ext_opts *opt, hipsw_region *out, int32_t *out_n,
                        hipsw_stats *stats);

/* The same driver as a four-pass state machine, for callers that run the
 * extensions themselves (the serving loop: hipsw_chunk windows fetched on the
 * device).  hipsw_chain_begin copies the chains; then four times
 * hipsw_chain_next (the pass's jobs: queries, and targets as getReference
 * windows (t_pos, t_len), reversed after fetching when t_rev) and
 * hipsw_chain_feed (n x 6 results in SWResult order, as hipsw_banded_extend
 * returns them); then hipsw_chain_finish writes the regions as
 * hipsw_extend_chains does.  The pass's arrays stay valid until the next
 * call on the driver.  Passes: left try 0, left try 1 (w doubled), right
 * try 0 (h0 = the left score), right try 1; a pass may have no jobs. */
typedef struct hipsw_chain_driver hipsw_chain_driver;

typedef struct {
    int32_t index;                  /* 0..3 */
    int32_t n;
    const uint8_t *q; const uint32_t *q_off; const uint16_t *q_len;
    const int64_t *t_pos; const int32_t *t_len; const uint8_t *t_rev;
    const int32_t *h0, *w, *eb;
} hipsw_chain_pass;

hipsw_chain_driver *hipsw_chain_begin(int32_t n_chains, const uint32_t *chain_seed_off,
                                      const int64_t *seed_rbeg, const int32_t *seed_qbeg,
                                      const int32_t *seed_len, const int32_t *seed_score,
                                      const uint32_t *chain_read, int32_t n_reads,
                                      const uint8_t *reads, const uint32_t *read_off,
                                      const int32_t *read_len, const hipsw_scoring *sc,
                                      const hipsw_ext_opts *opt);
int  hipsw_chain_next(hipsw_chain_driver *drv, hipsw_chain_pass *pass);
int  hipsw_chain_feed(hipsw_chain_driver *drv, const int32_t *out6);
int  hipsw_chain_finish(hipsw_chain_driver *drv, hipsw_region *out, int32_t *out_n);
void hipsw_chain_free(hipsw_chain_driver *drv);

/* Device-resident NW job sets (benchmarks, serving); see hipsw_ext_upload. */
typedef struct hipsw_nw_set hipsw_nw_set;
hipsw_nw_set *hipsw_nw_upload(hipsw_engine *eng, int32_t n,
                              const uint8_t *queries, const uint32_t *q_off,
                              const uint16_t *q_len, const uint8_t *targets,
                              const uint32_t *t_off, const uint16_t *t_len,
                              const int32_t *w, const int32_t *true_score,
                              int32_t max_retries, int32_t cap, const hipsw_scoring *sc);
int  hipsw_nw_run(hipsw_engine *eng, hipsw_nw_set *set, int32_t mode, hipsw_stats *stats);
int  hipsw_nw_download(hipsw_engine *eng, hipsw_nw_set *set, int32_t *out_score,
                       uint32_t *out_cigar, int32_t *out_n_cigar, int32_t *out_flags,
                       int32_t *out_final_w);
void hipsw_nw_free(hipsw_engine *eng, hipsw_nw_set *set);
/* Row-state form of the NW fill's 16-bit tier on this engine (no reference
 * counterpart; results are identical in every form): form 0, the default,
 * keeps a lane's row state in registers when its launch's bands are at most
 * 128 columns wide, and fills two jobs per lane in int16 pairs when the bands
 * of at most 64 columns, but not those of at most 32, hold all but 1/32 of
 * those cells; form 1 keeps the row state in an LDS ring (HBM for long
 * queries) throughout; form 2 keeps it
 * in registers, one job per lane; form 3 as form 2 but two jobs per lane for
 * every band of at most 64 columns.  Applies to NW runs and serving chunks
 * started afterwards.  HIPSW_E_INVALID for another form. */
int  hipsw_nw_set_state(hipsw_engine *eng, int32_t form);

/* ---- mate-rescue local SW (replaces LocalSWDispatcher.dispatchBatch,
 *      LocalSWDispatcher.swift:31-95) ----
 * Per task LocalSWAligner.align (LocalSWAligner.swift:42-79): the 16-lane
 * uint8 striped pass with its reverse pass, or the int32 scalar pass when
 * qLen*matchScore > 250 or the striped pass overflows.  min_sub_score enables
 * score2 (KSW_XSUBO; MateRescue passes minSeedLength*matchScore,
 * MateRescue.swift:125).  out6 rows: {score, qb, qe (inclusive), tb, te
 * (inclusive), score2}; ok[k] = 0 where the reference returns nil (the row is
 * then all zeros).  Unlike MetalSW's LocalSWResult, score2 is returned. */
int hipsw_local_sw(hipsw_engine *eng, int32_t n,
                   const uint8_t *queries, const uint32_t *q_off, const uint16_t *q_len,
                   const uint8_t *targets, const uint32_t *t_off, const uint16_t *t_len,
                   int32_t min_sub_score, int32_t match_score, int32_t mismatch_penalty,
                   const hipsw_scoring *sc, int32_t *out6, int32_t *ok, hipsw_stats *stats);

/* ---- serving loop (continuous batching over a resident reference) ----
 * No MetalSW counterpart (MetalSW dispatches one synchronous batch per call,
 * MetalDispatchHelper.swift:67-89).  A chunk carries the hot-path jobs of a
 * batch of reads; every target is a window of the resident reference given as
 * (BWT-space position, length), fetched on the device exactly as
 * FMIndex.getReference does (FMIndex.swift:25-65), and reversed afterwards when
 * ext_t_rev[k] is set (left extensions, ExtensionAligner.swift:219-225).
 * Chunks are admitted FIFO onto n_slots streams while their device memory fits
 * hbm_budget next to the running chunks, so chunks overlap on the device.
 * Results are those of hipsw_banded_extend / hipsw_local_sw / hipsw_global_nw
 * (max_retries 3) on the same jobs. */
typedef struct hipsw_server hipsw_server;

typedef struct {
    int32_t n_ext;                  /* banded extensions */
    const uint8_t *ext_q; const uint32_t *ext_q_off; const uint16_t *ext_q_len;
    const int64_t *ext_t_pos; const uint16_t *ext_t_len; const uint8_t *ext_t_rev;
    const int32_t *ext_h0, *ext_w, *ext_eb;
    int32_t n_loc;                  /* mate-rescue local SW */
    const uint8_t *loc_q; const uint32_t *loc_q_off; const uint16_t *loc_q_len;
    const int64_t *loc_t_pos; const uint16_t *loc_t_len;
    int32_t loc_min_sub;
    int32_t n_nw;                   /* CIGAR NW */
    const uint8_t *nw_q; const uint32_t *nw_q_off; const uint16_t *nw_q_len;
    const int64_t *nw_t_pos; const uint16_t *nw_t_len;
    const int32_t *nw_w, *nw_true_score;
    /* Non-NULL: the local-SW targets are mate-rescue windows in BWT space,
     * [loc_t_pos, loc_t_pos + loc_t_len), which the server clamps as
     * PackedReference.fetchSequence does (PackedReference.swift:72-121) to
     * the reference sequence holding the window's start; a job whose clamp is
     * empty or lands on a sequence other than loc_rid[k] (MateRescue.swift:
     * 105-111) is skipped (loc_ok 0, loc_win -1). */
    const int32_t *loc_rid;
} hipsw_chunk;

typedef struct {
    int32_t *ext_out;               /* n_ext x 6 (nullable)                     */
    int32_t *loc_out, *loc_ok;      /* n_loc x 6, n_loc                         */
    int32_t *nw_score;              /* n_nw                                     */
    uint32_t *nw_cigar;             /* n_nw x nw_cap                            */
    int32_t *nw_n_cigar, *nw_flags; /* n_nw                                     */
    double latency_ms;              /* submit -> results ready                  */
    double queue_ms, run_ms;
    int64_t cells;                  /* extension + NW fill cells                */
    double host_ms;                 /* part of run_ms spent packing and planning
                                       the chunk on the host                     */
    uint32_t *nw_cigar_off;         /* NULL: nw_cigar receives rows of nw_cap ops
                                       (the used ones, then zeros).  Otherwise
                                       n_nw + 1 offsets: nw_cigar (same size)
                                       receives the used ops back to back, job
                                       k's at [off[k], off[k + 1]), and only
                                       that prefix is written                    */
    int64_t est_bytes;              /* device bytes the chunk was admitted with  */
    int64_t dev_peak_bytes;         /* largest device footprint of its run; never
                                       above est_bytes                           */
    int64_t *loc_win;               /* n_loc clamped window starts in BWT space
                                       (fetchSequence's clampedBegin), -1 for a
                                       skipped job; written when loc_rid was set */
} hipsw_chunk_result;

hipsw_server *hipsw_server_create(hipsw_engine *eng, int32_t n_slots, int64_t hbm_budget,
                                  const hipsw_scoring *sc, int32_t match_score,
                                  int32_t mismatch_penalty, int32_t nw_cap);
/* Uploads the 2-bit .pac (4 bases per byte, first base in the high bits). */
int     hipsw_server_set_reference(hipsw_server *srv, const uint8_t *pac, int64_t n_bases);
/* The reference's sequences (ReferenceMetadata.annotations: offset, length),
 * which must tile [0, n_bases); set_reference resets them to one sequence. */
int     hipsw_server_set_contigs(hipsw_server *srv, int32_t n, const int64_t *off,
                                 const int64_t *len);
/* Copies the chunk and queues it; returns a ticket (> 0) or a negative status
 * (HIPSW_E_BUDGET when the chunk alone exceeds the budget). */
int64_t hipsw_server_submit(hipsw_server *srv, const hipsw_chunk *chunk);
/* HIPSW_OK when the ticket's results are ready, HIPSW_E_PENDING otherwise. */
int     hipsw_server_poll(hipsw_server *srv, int64_t ticket);
/* Blocks until ready, copies results into `res` and retires the ticket. */
int     hipsw_server_wait(hipsw_server *srv, int64_t ticket, hipsw_chunk_result *res);
int     hipsw_server_stats(hipsw_server *srv, int64_t *in_use, int64_t *peak_in_use,
                           int32_t *running, int32_t *queued);
/* Stops the slots; call after every ticket has been waited for. */
void    hipsw_server_destroy(hipsw_server *srv);

#ifdef __cplusplus
}
#endif
#endif
