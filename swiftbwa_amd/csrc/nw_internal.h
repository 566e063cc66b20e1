// Internal definitions of the banded global NW path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "hipsw_internal.h"

namespace hipsw {

constexpr int32_t NW_FLAG_TRUNCATED = 1;   // more CIGAR ops than the output row holds

// NW state tiers.  GlobalAligner keeps int32 H/E with the Int32.min/2 sentinel
// (GlobalAligner.swift:52).  The 16-bit tier stores both as int16 with the
// sentinel remapped to -16384; every comparison of the fill keeps its outcome
// while |real value| + |sentinel drift| < 16384, which the host checks per job
// with the bound (qlen + tlen) * max(max|mat|, o_del+e_del, o_ins+e_ins) <= 8000.
enum NwTier : int { NW_I16 = 0, NW_I32 = 1 };

struct NwDev {
    const uint32_t *seq;
    const uint32_t *nmask;
    const uint32_t *q_word, *t_word, *qn_word, *tn_word;
    const int32_t *q_len, *t_len;
    const int32_t *w;          // band of this pass (doubled on retries)
    int32_t *score;
    uint32_t *cigar;           // cap ops per job
    int32_t *n_cigar;
    int32_t *flags;
    unsigned long long *cells; // fill cells (GlobalAligner.swift:108 iterations)
    int cap;
};

struct NwParams {
    int8_t mat[32];
    int32_t o_del, e_del, o_ins, e_ins, oe_del, oe_ins;
};

// LDS bytes of one block: scoring table, the staged query (qw8 words of 8
// 4-bit codes per lane) and, unless held in HBM, the S-slot row state.
size_t nw_lds_bytes(int tier, bool global_state, int S, int qw8);
// Register-state form of the 16-bit tier: the instance (K = 32, 64, 96 or 128
// columns) for bands of at most ncol_max columns, 0 when none holds them; its
// LDS bytes; resident blocks per CU (registers and LDS), 0 if unknown.
int nw_reg_cols(int ncol_max);
size_t nw_reg_lds_bytes(int K, int qw8);
int nw_reg_blocks_per_cu(int K, int qw8);
// Two jobs per lane (K = 32, 64, 96): LDS bytes and resident blocks per CU.
size_t nw_pk_lds_bytes(int K, int qw8);
int nw_pk_blocks_per_cu(int K, int qw8);
// reg > 0 runs the register-state form with K = reg, reg < 0 its two-jobs-per-lane
// form with K = -reg (16-bit tier, |reg| >= 8 * ncw; items of up to 128 jobs).
hipError_t launch_nw(int tier, bool global_state, int reg, const NwDev &d, const NwParams &p,
                     const uint32_t *order, const WorkItem *items, int n_items,
                     uint32_t *counter, int S, uint32_t *zbase, int zrows, int ncw, int grid,
                     bool persistent, void *gstate, int qw8, hipStream_t s);

// Retry selection on the device: jobs of list[0..m) with score < ts go to
// next[0..*count) (count zeroed by the caller) with w_next[j] = 2 w_cur[j].
hipError_t launch_nw_select(const uint32_t *list, int m, const int32_t *score, const int32_t *ts,
                            const int32_t *w_cur, int32_t *w_next, uint32_t *next,
                            uint32_t *count, hipStream_t s);

// Wave-per-job form (nw_kernel.hip): jobs[0..n), query lengths <= 512 (the
// register chunks of the instance chosen by nw_wide_kc(qmax); 0 = none fits),
// o_ins >= 0.  Flag matrices of at most zmax bytes each, in LDS (zlds) or at
// zg + block * zmax.
int nw_wide_kc(int qmax);
hipError_t launch_nw_wide(const NwDev &d, const NwParams &p, const uint32_t *jobs, int n,
                          uint32_t *counter, int qmax, uint8_t *zg, size_t zmax, bool zlds,
                          int grid, hipStream_t s);

// One launch of an NW pass (host plan).
struct NwLaunch {
    int tier, S, ncol_max, zrows, item_begin, item_count, grid;
    bool global_state;
    int reg;     // register-state form: K columns (0: the LDS/HBM ring)
    int qw8;     // staged query words (8 bases each) per lane
    int wide;    // 1: run the bin's jobs with the wave-per-job form
    int qmax;    // longest query of the bin
    size_t zmax; // wave-per-job form: largest tlen * ncol of the bin
};

// Backpointer bytes of a lane-per-job launch: one word per row, 8-column
// block and job slot (64 per wave, 128 with two jobs per lane) per resident wave.
inline size_t nw_launch_zbytes(const NwLaunch &L)
{
    const size_t ncw = (size_t)(L.ncol_max + 7) / 8;
    return (size_t)L.grid * L.zrows * ncw * (L.reg < 0 ? 128 : 64) * 4;
}

// Launches of a pass queued on the side stream, ahead of the main stream's:
// wave-per-job bins with their flags in LDS, when the pass has other
// launches, and two-jobs-per-lane bins (backpointers in a region of their
// own) when a one-job-per-lane launch holds the main stream.
inline bool nw_side_launch(const NwLaunch &L, const std::vector<NwLaunch> &all)
{
    if (all.size() < 2) return false;
    if (L.wide) return L.zmax <= 49152;
    if (L.reg >= 0) return false;
    for (const NwLaunch &o : all)
        if (!o.wide && o.reg >= 0) return true;
    return false;
}

}  // namespace hipsw

struct hipsw_engine;
namespace hipsw {
// Plans one NW pass over the job ids in `active` with bands w[] (indexed by job).
// wide: 0 = lane-per-job launches only; 1 = bins too small to give every CU a
// wave go to the wave-per-job form; 2 = the whole pass does (if every query
// fits it).
void nw_plan_pass(const hipsw_engine *eng, const int32_t *ql, const int32_t *tl, int max_pen,
                  const std::vector<uint32_t> &active, const std::vector<int32_t> &w, int mode,
                  std::vector<uint32_t> &order, std::vector<WorkItem> &items,
                  std::vector<NwLaunch> &launches, int wide = 0);
}  // namespace hipsw
