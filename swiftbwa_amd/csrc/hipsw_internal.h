// Internal definitions shared by the HIP kernels and the host engine.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hipsw {

// Score tiers.  Every H/E/F value of a banded extension lies in
// [0, h0 + qlen * max(mat)] (BandedSWScalar.swift:112-143 clamps E and F at 0
// and H = max(diag, E, F)), so the tier is an exact bound test per job.
enum Tier : int { TIER_U8 = 0, TIER_U16 = 1, TIER_I32 = 2 };

constexpr uint32_t kNoN = 0xFFFFFFFFu;   // "sequence has no N" marker

// Device view of an uploaded extension job set (structure of arrays).
// Sequences are 2-bit codes, 16 per 32-bit word, base k of a word at bits
// 2k..2k+1; every query and target starts on a word boundary.  N bases carry
// code 0 in the 2-bit stream and a set bit in the N-mask stream (32 per word).
// One job of the packed form, stored in batched order so that a work item's
// descriptors are one contiguous read.
struct PkMeta {
    uint32_t job;             // index into the caller's job arrays / output rows
    uint32_t q_word, t_word;  // first 2-bit word of query / target
    uint16_t q_len, t_len;
    int32_t  h0, w, eb;
    uint32_t has_n;           // non-zero: query or target holds an N
};

struct ExtDev {
    const uint32_t *seq;      // 2-bit words
    const uint32_t *nmask;    // N-bit words (never dereferenced when no N)
    const uint32_t *q_word;   // first word of each query
    const uint32_t *t_word;   // first word of each target
    const uint32_t *qn_word;  // first N word of each query, kNoN if none
    const uint32_t *tn_word;  // first N word of each target, kNoN if none
    const int32_t  *q_len;
    const int32_t  *t_len;
    const int32_t  *h0;
    const int32_t  *w;
    const int32_t  *eb;       // end bonus
    int32_t        *out;      // 6 per job, SWResult order
    unsigned long long *cells; // DP cells evaluated (one atomic add per wave)
    const PkMeta   *pk;       // packed-form descriptors in batched order (batched form only)
};

// Scoring constants broadcast to every lane.
struct ExtParams {
    int8_t  mat[32];          // 25 used, [target*5 + query]
    int32_t o_del, e_del, o_ins, e_ins, zdrop;
    int32_t oe_del, oe_ins, max_mat;
    // Fast path: the 4x4 A/C/G/T block of mat is match_sc on the diagonal and
    // mis_sc off it (the shape ScoringParameters.scoringMatrix() builds).
    int32_t std_mat, match_sc, mis_sc;
};

// A unit of persistent work: `count` (<= 64) consecutive entries of the
// sorted job order starting at `start`.
struct WorkItem {
    uint32_t start;
    uint32_t count;
};

}  // namespace hipsw

// Kernel launchers (ext_kernel.hip).
namespace hipsw {
hipError_t launch_ext_plain(int tier, bool global_state, const ExtDev &d, const ExtParams &p,
                            const uint32_t *order, int n, int cols, int grid, void *gscratch,
                            hipStream_t s);
hipError_t launch_ext_persistent(int tier, bool global_state, const ExtDev &d,
                                 const ExtParams &p, const uint32_t *order,
                                 const WorkItem *items, int n_items, uint32_t *counter,
                                 int cols, int grid, void *gscratch, hipStream_t s);
hipError_t launch_zero_u32(uint32_t *p, int n, hipStream_t s);
size_t ext_state_bytes(int tier);   // bytes of h/e state per column per lane
size_t ext_lds_bytes(int tier, bool global_state, int cols);   // dynamic LDS per block
// Kernel families of a batched extension bin.
enum ExtKind {
    EXT_UNPACKED = 0,     // one job per lane, any matrix / N / tier (ext_persistent_kernel)
    EXT_PACKED = 1,       // two jobs per lane, 8-bit tier, items of 128 jobs
    EXT_LANE = 2,         // one job per lane, 8/16-bit tier, packed-form structure
    EXT_PACKED16 = 3,     // two jobs per lane, 16-bit tier (scores < 2048), items of 128 jobs
};
// Packed and one-per-lane forms: a bin whose longest query needs `cols`
// columns runs the instance for ext_pk_nq(cols) query words with `lb`
// 16-column blocks of its state in LDS and the rest in HBM.
int ext_pk_nq(int cols);
int ext_pk_default_lb(int cols);
int ext_pk_default_lb_k(int kind, int cols);
bool ext_pk_lb_supported(int kind, int cols, int lb);
int ext_pk_lb_candidates(int kind, int cols, int *out);   // out: up to 5 entries
size_t ext_pk_lds_bytes(int cols, int lb);                // LDS bytes per wave
size_t ext_pk_hbm_bytes_per_wave(int cols, int lb);       // HBM state bytes per resident wave
size_t ext_pk_lds_bytes_k(int kind, int cols, int lb);    // the same for a kind (EXT_PACKED16: two planes)
size_t ext_pk_hbm_bytes_per_wave_k(int kind, int cols, int lb);
// Descriptors of the packed form, in `order`.
void build_pk_meta(int n, const uint32_t *order, const uint32_t *q_word, const uint32_t *t_word,
                   const uint32_t *qn_word, const uint32_t *tn_word, const int32_t *q_len,
                   const int32_t *t_len, const int32_t *h0, const int32_t *w, const int32_t *eb,
                   PkMeta *out);
hipError_t launch_ext_persistent_pk(const ExtDev &d, const ExtParams &p, const WorkItem *items,
                                    int n_items, uint32_t *counter, int cols, int lb, int grid,
                                    void *gscratch, hipStream_t s);
// Two jobs per lane in the 16-bit tier (h and e planes of 16-bit pairs):
// scores below 2048 and e_del, e_ins <= 31 (checked by the planner).
hipError_t launch_ext_persistent_pk16(const ExtDev &d, const ExtParams &p, const WorkItem *items,
                                      int n_items, uint32_t *counter, int cols, int lb, int grid,
                                      void *gscratch, hipStream_t s);
// One-job-per-lane form (8- and 16-bit tiers, no N, standard matrix, items of
// <= 64 jobs, descriptors in d.pk); same LDS size as the packed form.
hipError_t launch_ext_persistent_ln(const ExtDev &d, const ExtParams &p, const WorkItem *items,
                                    int n_items, uint32_t *counter, int cols, int lb, int grid,
                                    void *gscratch, hipStream_t s);
}  // namespace hipsw
