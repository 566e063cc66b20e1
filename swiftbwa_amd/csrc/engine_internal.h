// Host-side engine state shared by the host translation units of libhipsw.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "hipsw.h"
#include "hipsw_internal.h"

struct hipsw_engine {
    int device = 0;
    int n_cu = 0;
    int lds_per_cu = 0;
    int lds_per_block = 0;
    std::string name, arch, last_error;
    hipStream_t stream = nullptr;
    hipStream_t side[3] = {nullptr, nullptr, nullptr};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t side_ev[4] = {nullptr, nullptr, nullptr, nullptr};
    uint32_t *counters = nullptr;     // 64 work-queue counters
    void *scratch = nullptr;          // global DP state for very long queries
    size_t scratch_bytes = 0;
    std::recursive_mutex mu;

    // NW arenas (grow-only), see nw_host.hip.
    void *nw_dev = nullptr;
    size_t nw_dev_bytes = 0;

    // Autotuner choice per query-length class of the 8-bit tier: 0 = default
    // (packed form), otherwise an ExtKind, and the 16-column blocks of state
    // held in LDS (0 = default) (hipsw_ext_autotune).
    int tuned_kind[4] = {0, 0, 0, 0};
    int tuned_lb[4] = {0, 0, 0, 0};
    int tuned_lb16[4] = {0, 0, 0, 0};   // 16-bit tier: LDS blocks of the lane kernel

    bool fail(const char *what, hipError_t e);
    bool check(hipError_t e, const char *what);
    void *scratch_get(size_t bytes);
    void free_nw_buffers();
    ~hipsw_engine();
};

namespace hipsw {

double now_ms();
void parallel_for(int64_t n, int threads, const std::function<void(int64_t, int64_t)> &fn);
int host_threads();
bool scoring_valid(const hipsw_scoring *sc);
ExtParams make_params(const hipsw_scoring *sc);
bool pack2(const uint8_t *src, int len, uint32_t *dst, uint32_t *ndst);
bool has_n(const uint8_t *src, int len);

// 2-bit packed query/target arenas of a job set (word-aligned per sequence).
struct PackedJobs {
    std::vector<uint32_t> seq, nmask, qw, tw, qnw, tnw;
    std::vector<int32_t> ql, tl;
};
// Packs n jobs; returns false (with a message) on a base code > 4 or an arena
// beyond 2^32 words.
bool pack_jobs(int n, const uint8_t *queries, const uint32_t *q_off, const uint16_t *q_len,
               const uint8_t *targets, const uint32_t *t_off, const uint16_t *t_len,
               PackedJobs &pj, std::string &err);

// Stable LSD radix sort of idx by 32-bit keys (two 16-bit passes).
void radix_sort_by_key(std::vector<uint32_t> &keys, std::vector<uint32_t> &idx);

// One persistent launch of the batched extension form.
struct ExtLaunch {
    int tier;
    bool global_state;
    int kind;             // ExtKind
    int cls;              // query-length class
    int job_begin, job_end;   // the bin's range in the batched order
    int cols;
    int lb;               // packed / one-per-lane: 16-column blocks of state in LDS
    int item_begin, item_count;
    int grid;
    size_t scratch_off;   // bytes into the engine scratch
    size_t scratch_bytes; // HBM DP state of the launch (0: all in LDS)
};

size_t plan_ext_batched(const hipsw_engine *eng, int n, const int32_t *ql, const int32_t *tl,
                        const int32_t *h0, const int32_t *w, const uint8_t *has_n, int max_mat,
                        bool std_mat, std::vector<uint32_t> &order, std::vector<WorkItem> &items,
                        std::vector<ExtLaunch> &launches);
// Work items and grid of a bin for its kind.
void ext_bin_items(const hipsw_engine *eng, ExtLaunch &L, std::vector<WorkItem> &items);
// Work items, grids and scratch regions of every bin (kinds fixed); returns
// the scratch bytes the launches need.
size_t ext_layout_launches(const hipsw_engine *eng, std::vector<ExtLaunch> &launches,
                           std::vector<WorkItem> &items);
hipError_t launch_ext_bin(const ExtLaunch &L, const ExtDev &d, const ExtParams &p,
                          const uint32_t *order, const WorkItem *items, uint32_t *counter,
                          void *gscratch, hipStream_t s);

}  // namespace hipsw

struct hipsw_ext_set {
    int n = 0;
    hipsw::ExtParams params;
    void *dmem = nullptr;            // one allocation holding every array below
    size_t dmem_bytes = 0;
    hipsw::ExtDev dev;
    uint32_t *d_order = nullptr;     // batched order
    hipsw::WorkItem *d_items = nullptr;
    // plain form
    int plain_tier = 0, plain_cols = 1;
    bool plain_global = false;
    int plain_grid = 1;
    size_t plain_scratch = 0;
    // batched form
    std::vector<hipsw::ExtLaunch> launches;
    size_t batched_scratch = 0;
    size_t items_cap = 0;            // capacity of d_items (work items)
    int64_t tier_jobs[3] = {0, 0, 0};
    // batched form as one graph (counter resets, then the bins); rebuilt when
    // the plan or the scratch region changes
    hipGraphExec_t graph = nullptr;
    void *graph_scratch = nullptr;
};
