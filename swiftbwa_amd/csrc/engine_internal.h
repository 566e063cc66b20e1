// Host-side engine state shared by the host translation units of libhipsw.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <condition_variable>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "hipsw.h"
#include "hipsw_internal.h"

namespace hipsw {

// std::vector whose growth leaves new elements of trivial types
// uninitialised, for host arrays whose every element is written before it is
// read: a resize then costs no fill pass.
template <class T> struct NoInitAlloc : std::allocator<T> {
    template <class U> struct rebind {
        using other = NoInitAlloc<U>;
    };
    NoInitAlloc() = default;
    template <class U> NoInitAlloc(const NoInitAlloc<U> &) noexcept {}
    template <class U> void construct(U *p) noexcept { ::new ((void *)p) U; }
    template <class U, class... A> void construct(U *p, A &&...a)
    {
        ::new ((void *)p) U(std::forward<A>(a)...);
    }
};
template <class T> using NoFillVec = std::vector<T, NoInitAlloc<T>>;

// Page-locked host memory (hipHostMalloc) for arrays the device copies from:
// their uploads run as direct DMA rather than through the runtime's pageable
// staging.  Elements are default-initialised like NoInitAlloc's.  Blocks
// above 1 GiB, and every block when pinning fails (no device), come from the
// ordinary heap; a 64-byte header records which.
void *pinned_or_heap_alloc(size_t bytes);
void pinned_or_heap_free(void *p) noexcept;
template <class T> struct PinnedAlloc {
    using value_type = T;
    PinnedAlloc() = default;
    template <class U> PinnedAlloc(const PinnedAlloc<U> &) noexcept {}
    T *allocate(size_t n) { return static_cast<T *>(pinned_or_heap_alloc(n * sizeof(T))); }
    void deallocate(T *p, size_t) noexcept { pinned_or_heap_free(p); }
    template <class U> void construct(U *p) noexcept { ::new ((void *)p) U; }
    template <class U, class... A> void construct(U *p, A &&...a)
    {
        ::new ((void *)p) U(std::forward<A>(a)...);
    }
    template <class U> bool operator==(const PinnedAlloc<U> &) const noexcept { return true; }
    template <class U> bool operator!=(const PinnedAlloc<U> &) const noexcept { return false; }
};
template <class T> using PinnedVec = std::vector<T, PinnedAlloc<T>>;

struct ExtPrep;

// A grow-only device buffer (hipMalloc only when a request outgrows it).
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    void *get(size_t bytes);   // nullptr when the allocation fails
    void release();
};

// The device resources one call runs on: a stream, two timing events, the
// persistent kernels' work-queue counters and grow-only buffers for the call's
// job arrays, HBM DP state and NW backpointers.  The engine owns one context
// for its device-resident sets and the autotuner (used under the engine
// mutex) and a pool of them for the one-shot entry points, so that calls from
// several host threads run concurrently, each on its own stream.
struct ExecCtx {
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // a second stream for launches that run beside the main one (small NW
    // bins next to the large one), forked from and joined back to `stream`
    // with two events
    hipStream_t side = nullptr;
    hipEvent_t evf = nullptr, evj = nullptr;
    uint32_t *counters = nullptr;   // 64 work-queue counters
    DevBuf arena;                   // job arrays of a one-shot call
    DevBuf scratch;                 // HBM DP state (extension hybrid layouts, local SW scalar)
    DevBuf nw;                      // NW backpointers and HBM row state
    bool init();
    void destroy();
};

// The autotuner's layout choice per query-length class: kind[c] 0 = default
// (packed form) or an ExtKind for the 8-bit tier, kind16[c] 0 = default (one
// job per lane) or an ExtKind for the 16-bit tier, and the 16-column blocks of
// column state held in LDS for the 8-bit (lb) and 16-bit (lb16) tiers, 0 =
// default.  Planners read a copy taken under the table's own mutex.
struct TuneTable {
    int kind[4] = {0, 0, 0, 0};
    int lb[4] = {0, 0, 0, 0};
    int kind16[4] = {0, 0, 0, 0};
    int lb16[4] = {0, 0, 0, 0};
};

}  // namespace hipsw

struct hipsw_engine {
    int device = 0;
    int n_cu = 0;
    int lds_per_cu = 0;
    int lds_per_block = 0;
    std::string name, arch;
    // device-resident sets (hipsw_ext_* / hipsw_nw_*) and the autotuner run on
    // `main` while holding `mu`
    hipsw::ExecCtx main;
    std::recursive_mutex mu;
    // one-shot entry points lease a context from this pool; the lock covers
    // only taking and returning a context
    std::vector<std::unique_ptr<hipsw::ExecCtx>> slots;
    std::vector<int> free_slots;
    std::mutex slot_mu;
    std::condition_variable slot_cv;
    // NW fill of the 16-bit tier (hipsw_nw_set_state): 0 = row state in
    // registers where the band fits, two jobs per lane when the bands of at
    // most 64 columns, but not those of at most 32, hold all but 1/32 of those
    // cells; 1 = the LDS ring throughout; 2 = registers, one job per lane;
    // 3 = registers, two jobs per lane for every band of at most 64 columns
    std::atomic<int> nw_form{0};
    // host halves of one-shot extension calls, kept for reuse: a call takes
    // one, packs and plans into it and gives it back, so repeated calls
    // neither fault in nor fill fresh host memory (at most kEngineSlots kept)
    std::mutex prep_mu;
    std::vector<std::unique_ptr<hipsw::ExtPrep>> prep_pool;
    // autotuner table, read by every planner as a copy
    mutable std::mutex tune_mu;
    hipsw::TuneTable tune;
    // device address of the allocation-free entry point's static scratch
    // array on this engine's device (hipGetSymbolAddress at creation)
    void *noalloc_scratch = nullptr;
    // fault injection (hipsw_fault_inject): the armed site and its countdown
    std::mutex fault_mu;
    int fault_site = 0, fault_countdown = 0;

    hipsw::TuneTable tune_snapshot() const;
    // True when an armed fault of `site` fires at this check (it then disarms);
    // the caller fails the step as if the HIP runtime had.
    bool inject(int site);
    // Error text of the calling thread's last failed call on this engine.
    void set_error(const std::string &msg) const;
    bool fail(const char *what, hipError_t e) const;
    bool check(hipError_t e, const char *what) const;
    ~hipsw_engine();
};

namespace hipsw {

// A context leased from the engine's pool for the duration of one call.
class SlotLease {
public:
    explicit SlotLease(hipsw_engine *eng);
    ~SlotLease();
    ExecCtx &ctx() { return *eng_->slots[idx_]; }
    SlotLease(const SlotLease &) = delete;
    SlotLease &operator=(const SlotLease &) = delete;

private:
    hipsw_engine *eng_;
    int idx_;
};

constexpr int kEngineSlots = 4;


double now_ms();
void parallel_for(int64_t n, int threads, const std::function<void(int64_t, int64_t)> &fn);
// fn(c) for c in [0, chunks), one thread each (in the caller's thread when
// chunks == 1)
void parallel_chunks(int chunks, const std::function<void(int)> &fn);
int host_threads();
bool scoring_valid(const hipsw_scoring *sc);
ExtParams make_params(const hipsw_scoring *sc);
bool pack2(const uint8_t *src, int len, uint32_t *dst, uint32_t *ndst);
bool has_n(const uint8_t *src, int len);
// pack2 without the N mask, in the same pass telling whether the sequence
// holds an N: -1 for a code above 4, 1 with an N, 0 without.
int pack2_scan(const uint8_t *src, int len, uint32_t *dst);
// The N mask words of a sequence (bit k of word k / 32 set where base k is N).
void nmask_fill(const uint8_t *src, int len, uint32_t *ndst);

// 2-bit packed query/target arenas of a job set (word-aligned per sequence).
struct PackedJobs {
    NoFillVec<uint32_t> seq, qw, tw, qnw, tnw;
    std::vector<uint32_t> nmask;
    NoFillVec<int32_t> ql, tl;
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
    int has_n;            // the bin's jobs hold an N base
    int job_begin, job_end;   // the bin's range in the batched order
    int cols;
    int lb;               // packed / one-per-lane: 16-column blocks of state in LDS
    bool w16_ok;          // 16-bit tier: the two-jobs-per-lane form may run the bin
                          // (every score below 2048, e_del and e_ins <= 31)
    int item_begin, item_count;
    int grid;
    size_t scratch_bytes; // HBM DP state of the launch (0: all in LDS); the
                          // bins run one after another on one stream, so all
                          // of them use the same region from its start
};

size_t plan_ext_batched(const hipsw_engine *eng, const TuneTable &tune, int n, const int32_t *ql,
                        const int32_t *tl, const int32_t *h0, const int32_t *w,
                        const uint8_t *has_n, int max_mat, bool std_mat, int e_max,
                        std::vector<uint32_t> &order, std::vector<WorkItem> &items,
                        std::vector<ExtLaunch> &launches);
// Work items and grid of a bin for its kind.
void ext_bin_items(const hipsw_engine *eng, ExtLaunch &L, std::vector<WorkItem> &items);
// Work items, grids and scratch sizes of every bin (kinds fixed); returns the
// scratch bytes the launches need (the largest bin's: bins share the region).
size_t ext_layout_launches(const hipsw_engine *eng, std::vector<ExtLaunch> &launches,
                           std::vector<WorkItem> &items);
// Upper bound of the scratch bytes plan_ext_batched can return for jobs with
// these query lengths and h0, whatever the tuning table says (every built
// LDS/HBM split is considered); used for admission before planning.
size_t ext_scratch_bound(const hipsw_engine *eng, int n, const int32_t *ql, const int32_t *h0,
                         const uint8_t *has_n, int max_mat, bool std_mat);
hipError_t launch_ext_bin(const ExtLaunch &L, const ExtDev &d, const ExtParams &p,
                          const uint32_t *order, const WorkItem *items, uint32_t *counter,
                          void *gscratch, hipStream_t s);

}  // namespace hipsw

struct hipsw_ext_set {
    int n = 0;
    hipsw::ExtParams params;
    void *dmem = nullptr;            // the arrays below; owned when own_dmem
    bool own_dmem = false;
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
    // batched form as one graph (counter resets, then the bins), for sets
    // that are run many times; rebuilt when the plan or the scratch region
    // changes
    hipGraphExec_t graph = nullptr;
    void *graph_scratch = nullptr;
    // input-order form (HIPSW_MODE_INPUT_ORDER): the batched form's register-
    // query kernel over the jobs in input order, one work item per block;
    // io_ok is false when some job needs the generic kernel.  Its descriptors
    // and items are built on the device at the first such run.
    bool io_ok = false;
    hipsw::ExtLaunch io_launch{};
    size_t io_scratch = 0;
    void *io_dmem = nullptr;
};

namespace hipsw {

// Host half of an extension set: packed sequences, per-job arrays, tiers,
// both plans and the packed-form descriptors, with the device arena's size.
// Built without touching the device.
struct ExtPrep {
    int n = 0;
    ExtParams params;
    // the arrays uploaded as they are: page-locked (kept for reuse by PrepLease)
    PinnedVec<uint32_t> seq, qw, tw, qnw, tnw;
    std::vector<uint32_t> nmask, order;
    PinnedVec<int32_t> ql, tl, h0, w, eb;
    std::vector<WorkItem> items;
    PinnedVec<PkMeta> pk;
    hipsw_ext_set shape;             // tiers, plans and scratch sizes (no device pointers)
    size_t arena_bytes = 0;          // device bytes of the set's arrays
};
// An ExtPrep taken from the engine's pool (a new one when the pool is empty)
// for the duration of one call; given back on destruction.
class PrepLease {
public:
    explicit PrepLease(hipsw_engine *eng);
    ~PrepLease();
    ExtPrep &get() { return *p_; }
    PrepLease(const PrepLease &) = delete;
    PrepLease &operator=(const PrepLease &) = delete;

private:
    hipsw_engine *eng_;
    std::unique_ptr<ExtPrep> p_;
};
// Validates, packs and plans n jobs.  Returns HIPSW_OK or a negative status
// with the reason in err.
int ext_prepare(const hipsw_engine *eng, const TuneTable &tune, int n, const uint8_t *queries,
                const uint32_t *q_off, const uint16_t *q_len, const uint8_t *targets,
                const uint32_t *t_off, const uint16_t *t_len, const int32_t *h0,
                const int32_t *w, const int32_t *end_bonus, const hipsw_scoring *sc,
                ExtPrep &P, std::string &err);
// Points `set` at an arena of P.arena_bytes at dmem and queues the uploads of
// P's arrays on stream s.
hipError_t ext_bind_upload(const ExtPrep &P, void *dmem, hipsw_ext_set *set, hipStream_t s);
// Queues one run of a bound set on stream s: the cell counter reset, event e0,
// the launches (plain: one; batched: one per bin) with work-queue counters
// `counters` (64) and HBM state at gs, event e1.
hipError_t ext_enqueue(const hipsw_ext_set *set, int mode, void *gs, uint32_t *counters,
                       hipStream_t s, hipEvent_t e0, hipEvent_t e1);

// Error text of the calling thread's last failed call on eng.
const char *thread_error(const hipsw_engine *eng);

// Device address of the allocation-free entry point's static scratch array
// on `device` (noalloc.hip), looked up once per device with the device
// current; nullptr on failure.
void *noalloc_scratch_address(int device);

// One-shot run of a prepared set on context c (hipsw_banded_extend): the
// job arrays go to `arena` (the context's grow-only arena when null), HBM
// state to gs_fixed (the context's scratch when null), the work-queue
// counters to `counters` (the context's when null); results into out.
int ext_run_once(hipsw_engine *eng, ExecCtx &c, const ExtPrep &P, int mode, void *arena,
                 void *gs_fixed, uint32_t *counters, int32_t *out, hipsw_stats *stats, double t0);

}  // namespace hipsw
