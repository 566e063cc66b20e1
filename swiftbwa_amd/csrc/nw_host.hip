// Host side of the banded global NW entry points (include/hipsw.h).
//
// Replaces GlobalSWDispatcher.dispatchBatch (Sources/MetalSW/GlobalSWDispatcher.swift:
// 6-118).  The retry loop is the CPU one (CIGARGenerator.swift:88-96): after each
// pass the tasks whose score is still below trueScore get w doubled and run
// again, as a compacted pass over just those tasks.
#include "engine_internal.h"
#include "nw_internal.h"

#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

using namespace hipsw;

struct hipsw_nw_set {
    int n = 0;
    int max_retries = 3;
    int cap = 512;
    NwParams params;
    std::vector<int32_t> ql, tl, w0, ts;
    int max_pen = 7;
    void *dmem = nullptr;            // owned when own_dmem (device-resident sets)
    bool own_dmem = false;
    NwDev dev;
    int32_t *d_w = nullptr;
    uint32_t *d_order = nullptr;
    WorkItem *d_items = nullptr;
    // retry passes: true scores, the doubled bands, the pass's order and
    // items, and the jobs the device selects for the next pass (count last)
    int32_t *d_ts = nullptr, *d_w_r = nullptr;
    uint32_t *d_order_r = nullptr, *d_next = nullptr;
    WorkItem *d_items_r = nullptr;
    // which mode's first-pass plan d_order / d_items hold (-1: none), and
    // whether d_w holds w0
    int p0_on_device = -1;
    bool w0_on_device = false;
    std::vector<int32_t> final_w;
    // first-pass plans, per mode (the retry passes are planned on the fly)
    bool plan_ready[2] = {false, false};
    int plan_form[2] = {0, 0};            // the engine's NW state form the plan was made for
    std::vector<uint32_t> plan_order[2];
    std::vector<hipsw::WorkItem> plan_items[2];
    std::vector<hipsw::NwLaunch> plan_launches[2];
};

namespace {

template <typename T>
T *carve(char *&cur, size_t count)
{
    T *p = reinterpret_cast<T *>(cur);
    cur += (count * sizeof(T) + 255) & ~(size_t)255;
    return p;
}
template <typename T>
size_t span(size_t count) { return (count * sizeof(T) + 255) & ~(size_t)255; }

int s_class(int S)
{
    if (S <= 32) return 0;
    if (S <= 64) return 1;
    if (S <= 128) return 2;
    if (S <= 256) return 3;
    return 4;
}

}  // namespace

namespace {

// Host half of an NW set: validation, scoring, packed sequences.
struct NwPrep {
    int n = 0, max_retries = 3, cap = 512, max_pen = 7;
    NwParams params;
    PackedJobs pj;
    std::vector<int32_t> w0, ts;
    size_t items_cap = 0, arena_bytes = 0;
};

int nw_prepare(int n, const uint8_t *queries, const uint32_t *q_off, const uint16_t *q_len,
               const uint8_t *targets, const uint32_t *t_off, const uint16_t *t_len,
               const int32_t *w, const int32_t *true_score, int32_t max_retries, int32_t cap,
               const hipsw_scoring *sc, NwPrep &P, std::string &err)
{
    if (n < 0 || cap < 1 || max_retries < 0 ||
        (n > 0 && (!queries || !q_off || !q_len || !targets || !t_off || !t_len || !w ||
                   !true_score))) {
        err = "invalid argument";
        return HIPSW_E_INVALID;
    }
    if (!scoring_valid(sc)) {
        err = "gap extension penalties must be > 0";
        return HIPSW_E_INVALID;
    }
    P.n = n;
    P.max_retries = max_retries;
    P.cap = cap;
    memset(&P.params, 0, sizeof P.params);
    int mat_abs = 0;
    for (int k = 0; k < 25; ++k) {
        P.params.mat[k] = sc->mat[k];
        mat_abs = std::max(mat_abs, std::abs((int)sc->mat[k]));
    }
    P.params.o_del = sc->o_del; P.params.e_del = sc->e_del;
    P.params.o_ins = sc->o_ins; P.params.e_ins = sc->e_ins;
    P.params.oe_del = sc->o_del + sc->e_del;
    P.params.oe_ins = sc->o_ins + sc->e_ins;
    P.max_pen = std::max(mat_abs, std::max(P.params.oe_del, P.params.oe_ins));
    for (int k = 0; k < n; ++k) {
        if (w[k] < 1) {
            // GlobalAligner's first-row loop `for j in 1...initEnd` needs w >= 1
            err = "w must be >= 1";
            return HIPSW_E_INVALID;
        }
    }
    if (!pack_jobs(n, queries, q_off, q_len, targets, t_off, t_len, P.pj, err))
        return HIPSW_E_INVALID;
    P.w0.assign(w, w + n);
    P.ts.assign(true_score, true_score + n);
    const size_t words = P.pj.seq.size(), nwords = P.pj.nmask.size();
    P.items_cap = (size_t)n / 64 + 64;
    P.arena_bytes = span<uint32_t>(words) + span<uint32_t>(nwords) + 4 * span<uint32_t>(n) +
                    3 * span<int32_t>(n) + span<int32_t>(n) + span<uint32_t>((size_t)n * cap) +
                    2 * span<int32_t>(n) + span<unsigned long long>(1) + span<uint32_t>(n) +
                    span<WorkItem>(P.items_cap) + 2 * span<int32_t>(n) + span<uint32_t>(n) +
                    span<WorkItem>(P.items_cap) + span<uint32_t>((size_t)n + 1);
    return HIPSW_OK;
}

hipError_t nw_bind_upload(NwPrep &P, void *dmem, hipsw_nw_set *set, hipStream_t st)
{
    const int n = P.n, cap = P.cap;
    set->n = n;
    set->max_retries = P.max_retries;
    set->cap = cap;
    set->params = P.params;
    set->max_pen = P.max_pen;
    set->ql.assign(P.pj.ql.begin(), P.pj.ql.end());
    set->tl.assign(P.pj.tl.begin(), P.pj.tl.end());
    set->w0 = P.w0;
    set->ts = P.ts;
    set->final_w = P.w0;
    set->dmem = dmem;
    const size_t words = P.pj.seq.size(), nwords = P.pj.nmask.size();
    char *cur = reinterpret_cast<char *>(dmem);
    uint32_t *d_seq = carve<uint32_t>(cur, words);
    uint32_t *d_nmask = carve<uint32_t>(cur, nwords);
    uint32_t *d_qw = carve<uint32_t>(cur, n);
    uint32_t *d_tw = carve<uint32_t>(cur, n);
    uint32_t *d_qn = carve<uint32_t>(cur, n);
    uint32_t *d_tn = carve<uint32_t>(cur, n);
    int32_t *d_ql = carve<int32_t>(cur, n);
    int32_t *d_tl = carve<int32_t>(cur, n);
    set->d_w = carve<int32_t>(cur, n);
    int32_t *d_score = carve<int32_t>(cur, n);
    uint32_t *d_cig = carve<uint32_t>(cur, (size_t)n * cap);
    int32_t *d_ncig = carve<int32_t>(cur, n);
    int32_t *d_flags = carve<int32_t>(cur, n);
    unsigned long long *d_cells = carve<unsigned long long>(cur, 1);
    set->d_order = carve<uint32_t>(cur, n);
    set->d_items = carve<WorkItem>(cur, P.items_cap);
    set->d_ts = carve<int32_t>(cur, n);
    set->d_w_r = carve<int32_t>(cur, n);
    set->d_order_r = carve<uint32_t>(cur, n);
    set->d_items_r = carve<WorkItem>(cur, P.items_cap);
    set->d_next = carve<uint32_t>(cur, (size_t)n + 1);
    set->p0_on_device = -1;
    set->w0_on_device = false;
    set->dev = NwDev{d_seq, d_nmask, d_qw, d_tw, d_qn, d_tn, d_ql, d_tl, set->d_w, d_score,
                     d_cig, d_ncig, d_flags, d_cells, cap};
    hipError_t e = hipSuccess;
    auto up = [&](void *dst, const void *src, size_t b) {
        if (e == hipSuccess && b) e = hipMemcpyAsync(dst, src, b, hipMemcpyHostToDevice, st);
    };
    up(d_seq, P.pj.seq.data(), words * 4);
    up(d_nmask, P.pj.nmask.data(), nwords * 4);
    up(d_qw, P.pj.qw.data(), (size_t)n * 4);
    up(d_tw, P.pj.tw.data(), (size_t)n * 4);
    up(d_qn, P.pj.qnw.data(), (size_t)n * 4);
    up(d_tn, P.pj.tnw.data(), (size_t)n * 4);
    up(d_ql, P.pj.ql.data(), (size_t)n * 4);
    up(d_tl, P.pj.tl.data(), (size_t)n * 4);
    up(set->d_ts, P.ts.data(), (size_t)n * 4);
    return e;
}

}  // namespace

extern "C" hipsw_nw_set *hipsw_nw_upload(hipsw_engine *eng, int32_t n, const uint8_t *queries,
                                         const uint32_t *q_off, const uint16_t *q_len,
                                         const uint8_t *targets, const uint32_t *t_off,
                                         const uint16_t *t_len, const int32_t *w,
                                         const int32_t *true_score, int32_t max_retries,
                                         int32_t cap, const hipsw_scoring *sc)
{
    if (!eng) return nullptr;
    NwPrep P;
    std::string err;
    if (nw_prepare(n, queries, q_off, q_len, targets, t_off, t_len, w, true_score, max_retries,
                   cap, sc, P, err) != HIPSW_OK) {
        eng->set_error("hipsw_nw_upload: " + err);
        return nullptr;
    }
    std::lock_guard<std::recursive_mutex> lock(eng->mu);
    hipSetDevice(eng->device);
    void *dmem = nullptr;
    if (!eng->check(hipMalloc(&dmem, P.arena_bytes), "hipMalloc(nw set)")) return nullptr;
    auto *set = new hipsw_nw_set();
    set->own_dmem = true;
    if (!eng->check(nw_bind_upload(P, dmem, set, eng->main.stream), "upload") ||
        !eng->check(hipStreamSynchronize(eng->main.stream), "upload sync")) {
        hipFree(dmem);
        delete set;
        return nullptr;
    }
    return set;
}

// Builds the launches of one pass over `active` (job ids) with bands w[].
void hipsw::nw_plan_pass(const hipsw_engine *eng, const int32_t *qlv, const int32_t *tlv,
                         int max_pen, const std::vector<uint32_t> &active,
                         const std::vector<int32_t> &w, int mode, std::vector<uint32_t> &order,
                         std::vector<WorkItem> &items, std::vector<NwLaunch> &launches, int wide)
{
    const int m = (int)active.size();
    std::vector<int> ncol(m), S(m), tier(m);
    for (int a = 0; a < m; ++a) {
        const int k = (int)active[a];
        const int q = qlv[k], t = tlv[k];
        const int wp = std::max(std::min(w[k], 1 << 20), std::abs(q - t));
        ncol[a] = std::min(q, 2 * wp + 1);
        S[a] = std::max(ncol[a] + 2, 10);   // the fill reads 8 distinct slots per block
        const int64_t bound = (int64_t)(q + t) * max_pen + 64;
        tier[a] = bound <= 8000 ? NW_I16 : NW_I32;
    }
    order.clear();
    items.clear();
    launches.clear();
    const int form = eng->nw_form.load();
    auto add_launch = [&](int tr, int s_max, int c_max, int z_rows, int ib, int ic, int q_max,
                          size_t z_max, bool pk) {
        NwLaunch L;
        L.qmax = q_max;
        L.zmax = z_max;
        L.wide = 0;
        if (wide == 2 || (wide == 1 && ic < eng->n_cu)) L.wide = nw_wide_kc(q_max) > 0 ? 1 : 0;
        L.qw8 = std::max(1, (q_max + 7) / 8);
        L.tier = tr;
        L.S = s_max;
        L.ncol_max = c_max;
        L.zrows = z_rows;
        L.item_begin = ib;
        L.item_count = ic;
        const size_t lds = nw_lds_bytes(tr, false, s_max, L.qw8);
        L.global_state = lds > (size_t)eng->lds_per_block;
        L.reg = 0;
        const int kreg = tr == NW_I16 && form != 1 ? nw_reg_cols(c_max) : 0;
        int reg_bpc = 0;
        if (pk && kreg && kreg <= 64 && nw_pk_lds_bytes(kreg, L.qw8) <= (size_t)eng->lds_per_block)
            reg_bpc = nw_pk_blocks_per_cu(kreg, L.qw8);
        else if (!pk && kreg && nw_reg_lds_bytes(kreg, L.qw8) <= (size_t)eng->lds_per_block)
            reg_bpc = nw_reg_blocks_per_cu(kreg, L.qw8);
        if (reg_bpc > 0) {
            L.reg = pk ? -kreg : kreg;
            L.global_state = false;
            L.grid = std::max(1, std::min(ic, eng->n_cu * std::min(reg_bpc, 32)));
        } else {
            const size_t lds_used = nw_lds_bytes(tr, L.global_state, s_max, L.qw8);
            const int wpc = std::max(1, std::min(16, (int)(eng->lds_per_cu / (lds_used + 512))));
            L.grid = std::max(1, std::min(ic, eng->n_cu * wpc));
        }
        if (L.wide) {
            // one wave per job; flag matrices in LDS when they fit
            int jobs = 0;
            for (int t = ib; t < ib + ic; ++t) jobs += (int)items[t].count;
            const bool zl = z_max <= 49152;
            const int per_cu = zl ? std::max(1, std::min(8, (int)(eng->lds_per_cu / (z_max + 1024)))) : 8;
            L.grid = std::max(1, std::min(jobs, eng->n_cu * per_cu));
        }
        launches.push_back(L);
    };
    if (mode == HIPSW_MODE_PLAIN) {
        int tr = NW_I16, smax = 3, cmax = 1, zr = 1, qm = 1;
        size_t zm = 1;
        for (int a = 0; a < m; ++a) {
            tr = std::max(tr, tier[a]);
            smax = std::max(smax, S[a]);
            cmax = std::max(cmax, ncol[a]);
            zr = std::max(zr, tlv[active[a]]);
            qm = std::max(qm, qlv[active[a]]);
            zm = std::max(zm, (size_t)tlv[active[a]] * ncol[a]);
            order.push_back(active[a]);
        }
        for (int g = 0; g < m; g += 64) items.push_back(WorkItem{(uint32_t)g, (uint32_t)std::min(64, m - g)});
        add_launch(tr, smax, cmax, zr, 0, (int)items.size(), qm, zm, false);
        return;
    }
    // Batched: bins by (tier, S class); inside a bin the lanes of one wave
    // run the same number of 8-cell blocks per row and have similar target
    // lengths (blocks desc, tlen desc).  With the register-state form the
    // 16-bit jobs go by register instance (bands of <= 32, 64, 96, 128
    // columns, 5 to 2 waves per SIMD): the classes up to the smallest one
    // that leaves at most 1/32 of their cells above it share one bin, so one
    // persistent launch balances them (its items longest first) instead of
    // each class ending on its own tail, and the few wider jobs do not pull
    // the shared launch down to a larger instance's occupancy.
    // Two jobs per lane (bins 4, 5): form 0 runs the merged bin that way when
    // it is the 64-column instance (with bands of at most 32 columns only, one
    // job per lane at 5 waves per SIMD is faster; the wider classes, at most
    // 1/32 of the cells, go to bins of their own beside it); form 3 always runs the
    // classes of at most 64 columns that way, merged among themselves, and
    // merges the wider ones as before (bins 2, 3).  Both persistent launches
    // then share the GPU, the packed one on the side stream.
    const bool reg_form = form != 1;
    bool pk_form = form == 3;
    std::vector<int> rcls(m, -1);
    int kstar = -1, kpk = -1;
    if (reg_form) {
        int64_t cells[4] = {0, 0, 0, 0};
        for (int a = 0; a < m; ++a) {
            const int K = tier[a] == NW_I16 ? nw_reg_cols(ncol[a]) : 0;
            if (!K) continue;
            rcls[a] = K / 32 - 1;
            cells[rcls[a]] += (int64_t)tlv[active[a]] * ncol[a];
        }
        // the smallest class of [lo, hi] with at most 1/32 of the group's cells above it
        auto merge_class = [&](int lo, int hi) {
            int64_t total = 0;
            for (int k = lo; k <= hi; ++k) total += cells[k];
            int64_t above = total;
            int k = lo;
            for (; k < hi; ++k) {
                above -= cells[k];
                if (above * 32 <= total) break;
            }
            return k;
        };
        if (pk_form) {
            kpk = merge_class(0, 1);
            kstar = merge_class(2, 3);
        } else {
            kstar = merge_class(0, 3);
            if (form == 0 && kstar == 1) {
                pk_form = true;
                kpk = kstar;
                kstar = 2;                                     // classes 2, 3 in bins of their own
            }
        }
    }
    std::vector<uint32_t> keys(m), idx(m);
    for (int a = 0; a < m; ++a) {
        int sc = s_class(S[a]);
        if (rcls[a] >= 0 && pk_form && rcls[a] <= 1) sc = rcls[a] <= kpk ? 4 : 5;   // 4, 5
        else if (rcls[a] >= 0 && pk_form) sc = rcls[a] <= kstar ? 2 : 3;            // 2, 3
        else if (rcls[a] >= 0) sc = rcls[a] <= kstar ? 0 : rcls[a];                 // 0..3
        else if (reg_form && tier[a] == NW_I16) sc = 3 + s_class(S[a]);  // 6, 7: S > 130
        const uint32_t bin = (uint32_t)(tier[a] * 8 + sc);
        const uint32_t nb = (uint32_t)std::min((ncol[a] + 7) / 8, 4095);
        const uint32_t tl = (uint32_t)std::min(tlv[active[a]], 65535);
        keys[a] = (bin << 28) | ((4095u - nb) << 16) | (65535u - tl);
        idx[a] = (uint32_t)a;
    }
    radix_sort_by_key(keys, idx);
    for (int a = 0; a < m; ++a) order.push_back(active[idx[a]]);
    for (int s = 0; s < m;) {
        const uint32_t bin = keys[s] >> 28;
        int e = s, smax = 3, cmax = 1, zr = 1, qm = 1;
        size_t zm = 1;
        while (e < m && (keys[e] >> 28) == bin) {
            smax = std::max(smax, S[idx[e]]);
            cmax = std::max(cmax, ncol[idx[e]]);
            zr = std::max(zr, tlv[active[idx[e]]]);
            qm = std::max(qm, qlv[active[idx[e]]]);
            zm = std::max(zm, (size_t)tlv[active[idx[e]]] * ncol[idx[e]]);
            ++e;
        }
        const int ib = (int)items.size();
        const bool pk = pk_form && bin / 8 == NW_I16 && (bin % 8 == 4 || bin % 8 == 5);
        const int G = pk ? 128 : 64;
        for (int g = s; g < e; g += G) items.push_back(WorkItem{(uint32_t)g, (uint32_t)std::min(G, e - g)});
        add_launch((int)bin / 8, smax, cmax, zr, ib, (int)items.size() - ib, qm, zm, pk);
        // Items longest first (a wave's time is its longest target times its
        // widest band in blocks), so the persistent waves end together.  The
        // wave-per-job form reads the bin's jobs as one contiguous range of
        // the order and keeps its items as they are.
        if (!launches.back().wide) {
            std::vector<std::pair<int64_t, WorkItem>> cost;
            for (int t = ib; t < (int)items.size(); ++t) {
                int tm = 0, nb = 0;
                for (uint32_t u = items[t].start; u < items[t].start + items[t].count; ++u) {
                    tm = std::max(tm, tlv[active[idx[u]]]);
                    nb = std::max(nb, (ncol[idx[u]] + 7) / 8);
                }
                cost.push_back({(int64_t)tm * nb, items[t]});
            }
            std::stable_sort(cost.begin(), cost.end(),
                             [](const std::pair<int64_t, WorkItem> &x, const std::pair<int64_t, WorkItem> &y) {
                                 return x.first > y.first;
                             });
            for (size_t t = 0; t < cost.size(); ++t) items[ib + t] = cost[t].second;
        }
        s = e;
    }
}

// The passes of the CPU retry loop over a bound set, on context c.
static int nw_run_ctx(hipsw_engine *eng, ExecCtx &c, hipsw_nw_set *set, int mode,
                      hipsw_stats *stats)
{
    const double t0 = now_ms();
    const int n = set->n;
    double kernel_ms = 0;
    int64_t launches_total = 0;
    std::vector<int32_t> w = set->w0;
    std::vector<uint32_t> active(n);
    std::iota(active.begin(), active.end(), 0u);
    std::vector<uint32_t> order;
    std::vector<WorkItem> items;
    std::vector<NwLaunch> launches;
    if (n > 0 && !eng->check(launch_zero_u32(reinterpret_cast<uint32_t *>(set->dev.cells), 2,
                                             c.stream), "zero"))
        return HIPSW_E_DEVICE;
    for (int pass = 0; !active.empty(); ++pass) {
        if (pass == 0 && set->plan_ready[mode] && set->plan_form[mode] == eng->nw_form.load()) {
            order = set->plan_order[mode];
            items = set->plan_items[mode];
            launches = set->plan_launches[mode];
        } else {
            // Retry passes hold few tasks: one launch each.  The batched form
            // runs its small bins and its retry passes one wave per job (the
            // F recurrence needs o_ins >= 0); the plain form keeps one lane
            // per job throughout.
            const int wide = mode == HIPSW_MODE_BATCHED && set->params.o_ins >= 0 ? (pass == 0 ? 1 : 2) : 0;
            nw_plan_pass(eng, set->ql.data(), set->tl.data(), set->max_pen, active, w,
                         pass == 0 ? mode : HIPSW_MODE_PLAIN, order, items, launches, wide);
            if (pass == 0) {
                set->plan_order[mode] = order;
                set->plan_items[mode] = items;
                set->plan_launches[mode] = launches;
                set->plan_ready[mode] = true;
                set->plan_form[mode] = eng->nw_form.load();
                set->p0_on_device = -1;                        // the device holds an older plan
            }
        }
        if (launches.size() > 64) {
            eng->set_error("hipsw_nw_run: too many bins");
            return HIPSW_E_UNSUPPORTED;
        }
        // Scratch: backpointers [row][c/8][lane] per resident wave, plus DP
        // state for launches whose state does not fit in LDS.  Launches run in
        // order on their stream: those of one stream share one region; the
        // side stream's two-jobs-per-lane launches have a region of their own.
        size_t zbytes = 0, zside = 0, gbytes = 0;
        for (const NwLaunch &L : launches) {
            if (L.wide) {
                if (L.zmax > 49152) zbytes = std::max(zbytes, (size_t)L.grid * L.zmax);
                continue;
            }
            if (nw_side_launch(L, launches)) zside = std::max(zside, nw_launch_zbytes(L));
            else zbytes = std::max(zbytes, nw_launch_zbytes(L));
            if (L.global_state)
                gbytes = std::max(gbytes, (size_t)L.grid * L.S * 64 * (L.tier == NW_I16 ? 4 : 8));
        }
        zbytes = (zbytes + 255) & ~(size_t)255;
        const size_t need = zbytes + zside + ((gbytes + 255) & ~(size_t)255) + 256;
        void *nwbuf = eng->inject(HIPSW_FAULT_ALLOC) ? nullptr : c.nw.get(need);
        if (!nwbuf) {
            (void)hipStreamSynchronize(c.stream);
            eng->set_error("hipsw_nw_run: scratch allocation failed (injected or out of memory)");
            return HIPSW_E_DEVICE;
        }
        uint32_t *zbase = reinterpret_cast<uint32_t *>(nwbuf);
        uint32_t *zbase_side = reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(nwbuf) + zbytes);
        void *gstate = reinterpret_cast<char *>(nwbuf) + zbytes + ((zside + 255) & ~(size_t)255);
        // The first pass reads d_w (w0), d_order and d_items, uploaded once
        // per plan; a retry pass reads d_w_r (its bands, doubled on the
        // device) and its own small order and items.
        bool ok = true;
        uint32_t *d_ord = pass == 0 ? set->d_order : set->d_order_r;
        WorkItem *d_it = pass == 0 ? set->d_items : set->d_items_r;
        NwDev dev = set->dev;
        if (pass == 0) {
            if (!set->w0_on_device) {
                ok = eng->check(hipMemcpyAsync(set->d_w, set->w0.data(), (size_t)n * 4,
                                               hipMemcpyHostToDevice, c.stream), "upload w");
                set->w0_on_device = ok;
            }
            if (ok && set->p0_on_device != mode) {
                ok = eng->check(hipMemcpyAsync(d_ord, order.data(), order.size() * 4,
                                               hipMemcpyHostToDevice, c.stream), "upload order") &&
                     eng->check(hipMemcpyAsync(d_it, items.data(), items.size() * sizeof(WorkItem),
                                               hipMemcpyHostToDevice, c.stream), "upload items");
                set->p0_on_device = ok ? mode : -1;
            }
        } else {
            dev.w = set->d_w_r;
            ok = eng->check(hipMemcpyAsync(d_ord, order.data(), order.size() * 4,
                                           hipMemcpyHostToDevice, c.stream), "upload order") &&
                 eng->check(hipMemcpyAsync(d_it, items.data(), items.size() * sizeof(WorkItem),
                                           hipMemcpyHostToDevice, c.stream), "upload items");
        }
        if (ok && eng->inject(HIPSW_FAULT_LAUNCH)) {
            eng->set_error("hipsw_nw_run: kernel launch failed (injected)");
            ok = false;
        }
        if (ok) ok = eng->check(launch_zero_u32(c.counters, 64, c.stream), "zero");
        if (ok) ok = eng->check(hipEventRecord(c.ev0, c.stream), "event");
        // Small bins sent to the wave-per-job form with their flags in LDS
        // touch no shared scratch: they run on the side stream, queued first
        // (their waves start before the lane-per-job launch fills the GPU and
        // end inside it; queued after it, they wait for its first waves to
        // retire and run past its end).  The two-jobs-per-lane bins go there
        // too, with their own backpointer region: the two persistent launches
        // then fill each other's tails.
        bool forked = false;
        auto side_ok = [&](const NwLaunch &L) { return nw_side_launch(L, launches); };
        auto run_wide = [&](const NwLaunch &L, int k, hipStream_t ls) {
            // the bin's jobs are order[items[ib].start ..) contiguously
            const WorkItem &first = items[L.item_begin];
            const WorkItem &last = items[L.item_begin + L.item_count - 1];
            const int nj = (int)(last.start + last.count - first.start);
            return eng->check(launch_nw_wide(dev, set->params, d_ord + first.start, nj,
                                             c.counters + k, L.qmax,
                                             reinterpret_cast<uint8_t *>(zbase), L.zmax,
                                             L.zmax <= 49152, L.grid, ls),
                              "launch nw wide");
        };
        for (size_t li = 0; ok && li < launches.size(); ++li) {
            if (!side_ok(launches[li])) continue;
            if (!forked) {
                ok = eng->check(hipEventRecord(c.evf, c.stream), "event") &&
                     eng->check(hipStreamWaitEvent(c.side, c.evf, 0), "fork");
                forked = true;
            }
            if (ok && launches[li].wide) ok = run_wide(launches[li], (int)li, c.side);
            else if (ok) ok = eng->check(launch_nw(launches[li].tier, false, launches[li].reg, dev,
                                                   set->params, d_ord, d_it + launches[li].item_begin,
                                                   launches[li].item_count, c.counters + li, 0,
                                                   zbase_side, launches[li].zrows,
                                                   (launches[li].ncol_max + 7) / 8, launches[li].grid,
                                                   pass == 0 && mode == HIPSW_MODE_BATCHED, nullptr,
                                                   launches[li].qw8, c.side),
                                         "launch nw (side)");
            ++launches_total;
        }
        int k = -1;
        for (const NwLaunch &L : launches) {
            ++k;
            if (!ok) break;
            if (side_ok(L)) continue;
            if (L.wide) {
                ok = run_wide(L, k, c.stream);
                ++launches_total;
                continue;
            }
            const int ncw = (L.ncol_max + 7) / 8;
            ok = eng->check(launch_nw(L.tier, L.global_state, L.reg, dev, set->params, d_ord,
                                      d_it + L.item_begin, L.item_count,
                                      c.counters + k, L.S, zbase, L.zrows, ncw, L.grid,
                                      pass == 0 && mode == HIPSW_MODE_BATCHED, gstate, L.qw8,
                                      c.stream),
                            "launch nw");
            ++launches_total;
        }
        if (forked) {
            // join even after a failed launch, so nothing is left queued on the side
            const bool j = eng->check(hipEventRecord(c.evj, c.side), "event") &&
                           eng->check(hipStreamWaitEvent(c.stream, c.evj, 0), "join");
            ok = ok && j;
        }
        if (ok) ok = eng->check(hipEventRecord(c.ev1, c.stream), "event");
        if (ok) ok = eng->check(hipEventSynchronize(c.ev1), "nw kernel");
        if (!ok) {
            (void)hipStreamSynchronize(c.stream);
            (void)hipStreamSynchronize(c.side);
            return HIPSW_E_DEVICE;
        }
        float ms = 0.f;
        hipEventElapsedTime(&ms, c.ev0, c.ev1);
        kernel_ms += ms;
        if (eng->inject(HIPSW_FAULT_COPY)) {
            eng->set_error("hipsw_nw_run: download failed (injected)");
            return HIPSW_E_DEVICE;
        }
        // CIGARGenerator.swift:88-96: retry while score < trueScore, at most
        // max_retries times (every job of pass p has been tried p times).
        for (uint32_t j : active) set->final_w[j] = w[j];
        if (pass >= set->max_retries) break;
        uint32_t *d_count = set->d_next + n;
        uint32_t cnt = 0;
        if (!eng->check(launch_zero_u32(d_count, 1, c.stream), "zero") ||
            !eng->check(launch_nw_select(d_ord, (int)active.size(), set->dev.score, set->d_ts,
                                         pass == 0 ? set->d_w : set->d_w_r, set->d_w_r,
                                         set->d_next, d_count, c.stream), "retry select") ||
            !eng->check(hipMemcpyAsync(&cnt, d_count, 4, hipMemcpyDeviceToHost, c.stream),
                        "download retry count") ||
            !eng->check(hipStreamSynchronize(c.stream), "retry count"))
            return HIPSW_E_DEVICE;
        std::vector<uint32_t> next(cnt);
        if (cnt && (!eng->check(hipMemcpyAsync(next.data(), set->d_next, (size_t)cnt * 4,
                                               hipMemcpyDeviceToHost, c.stream), "download retries") ||
                    !eng->check(hipStreamSynchronize(c.stream), "retries")))
            return HIPSW_E_DEVICE;
        std::sort(next.begin(), next.end());                   // the device appends in no fixed order
        for (uint32_t j : next) w[j] = w[j] < (1 << 20) ? w[j] * 2 : w[j];
        active.swap(next);
    }
    if (stats) {
        memset(stats, 0, sizeof *stats);
        stats->kernel_ms = kernel_ms;
        stats->total_ms = now_ms() - t0;
        stats->n_jobs = n;
        stats->launches = launches_total;
        unsigned long long cells = 0;
        if (n > 0 && (!eng->check(hipMemcpyAsync(&cells, set->dev.cells, sizeof cells,
                                                 hipMemcpyDeviceToHost, c.stream), "cells") ||
                      !eng->check(hipStreamSynchronize(c.stream), "cells")))
            return HIPSW_E_DEVICE;
        stats->cells = (int64_t)cells;
    }
    return HIPSW_OK;
}

extern "C" int hipsw_nw_run(hipsw_engine *eng, hipsw_nw_set *set, int32_t mode,
                            hipsw_stats *stats)
{
    if (!eng || !set) return HIPSW_E_INVALID;
    if (mode != HIPSW_MODE_PLAIN && mode != HIPSW_MODE_BATCHED) return HIPSW_E_INVALID;
    std::lock_guard<std::recursive_mutex> lock(eng->mu);
    hipSetDevice(eng->device);
    return nw_run_ctx(eng, eng->main, set, mode, stats);
}

// Results of a bound set into the caller's arrays, copied on stream st.
static int nw_download_ctx(hipsw_engine *eng, hipStream_t st, hipsw_nw_set *set,
                           int32_t *out_score, uint32_t *out_cigar, int32_t *out_n_cigar,
                           int32_t *out_flags, int32_t *out_final_w)
{
    const int n = set->n;
    if (n == 0) return HIPSW_OK;
    if (eng->inject(HIPSW_FAULT_COPY)) {
        eng->set_error("hipsw_nw_download: download failed (injected)");
        return HIPSW_E_DEVICE;
    }
    bool ok = true;
    auto down = [&](void *dst, const void *src, size_t b) {
        if (ok && dst) ok = eng->check(hipMemcpyAsync(dst, src, b, hipMemcpyDeviceToHost, st), "download");
    };
    std::vector<int32_t> ncig(n);
    down(out_score, set->dev.score, (size_t)n * 4);
    down(out_cigar, set->dev.cigar, (size_t)n * set->cap * 4);
    down(ncig.data(), set->dev.n_cigar, (size_t)n * 4);
    down(out_flags, set->dev.flags, (size_t)n * 4);
    if (ok) ok = eng->check(hipStreamSynchronize(st), "download");
    if (!ok) {
        (void)hipStreamSynchronize(st);
        return HIPSW_E_DEVICE;
    }
    if (out_n_cigar) memcpy(out_n_cigar, ncig.data(), (size_t)n * 4);
    if (out_cigar)                                       // entries past the CIGAR are zero
        for (int k = 0; k < n; ++k) {
            const int used = std::min(std::max(ncig[k], 0), set->cap);
            std::fill(out_cigar + (size_t)k * set->cap + used, out_cigar + (size_t)(k + 1) * set->cap, 0u);
        }
    if (out_final_w) memcpy(out_final_w, set->final_w.data(), (size_t)n * 4);
    return HIPSW_OK;
}

extern "C" int hipsw_nw_download(hipsw_engine *eng, hipsw_nw_set *set, int32_t *out_score,
                                 uint32_t *out_cigar, int32_t *out_n_cigar, int32_t *out_flags,
                                 int32_t *out_final_w)
{
    if (!eng || !set) return HIPSW_E_INVALID;
    std::lock_guard<std::recursive_mutex> lock(eng->mu);
    hipSetDevice(eng->device);
    return nw_download_ctx(eng, eng->main.stream, set, out_score, out_cigar, out_n_cigar,
                           out_flags, out_final_w);
}

extern "C" void hipsw_nw_free(hipsw_engine *eng, hipsw_nw_set *set)
{
    if (!set) return;
    if (eng) {
        std::lock_guard<std::recursive_mutex> lock(eng->mu);
        hipSetDevice(eng->device);
        hipStreamSynchronize(eng->main.stream);
        if (set->dmem && set->own_dmem) hipFree(set->dmem);
    }
    delete set;
}

extern "C" int hipsw_global_nw(hipsw_engine *eng, int32_t n, const uint8_t *queries,
                               const uint32_t *q_off, const uint16_t *q_len,
                               const uint8_t *targets, const uint32_t *t_off,
                               const uint16_t *t_len, const int32_t *w,
                               const int32_t *true_score, int32_t max_retries, int32_t cap,
                               const hipsw_scoring *sc, int32_t mode, int32_t *out_score,
                               uint32_t *out_cigar, int32_t *out_n_cigar, int32_t *out_flags,
                               int32_t *out_final_w, hipsw_stats *stats)
{
    if (!eng) return HIPSW_E_INVALID;
    if (n == 0) return HIPSW_OK;
    if (n < 0 || !out_score || !out_cigar || !out_n_cigar || !out_flags ||
        (mode != HIPSW_MODE_PLAIN && mode != HIPSW_MODE_BATCHED))
        return HIPSW_E_INVALID;
    const double t0 = now_ms();
    NwPrep P;
    std::string err;
    const int rc0 = nw_prepare(n, queries, q_off, q_len, targets, t_off, t_len, w, true_score,
                               max_retries, cap, sc, P, err);
    if (rc0 != HIPSW_OK) {
        eng->set_error("hipsw_global_nw: " + err);
        return rc0;
    }
    SlotLease lease(eng);
    ExecCtx &c = lease.ctx();
    hipSetDevice(eng->device);
    void *dmem = eng->inject(HIPSW_FAULT_ALLOC) ? nullptr : c.arena.get(P.arena_bytes);
    if (!dmem) {
        eng->set_error("hipsw_global_nw: arena allocation failed (injected or out of memory)");
        return HIPSW_E_DEVICE;
    }
    hipsw_nw_set set;
    if (eng->inject(HIPSW_FAULT_COPY)) {
        eng->set_error("hipsw_global_nw: upload failed (injected)");
        return HIPSW_E_DEVICE;
    }
    if (!eng->check(nw_bind_upload(P, dmem, &set, c.stream), "hipsw_global_nw: upload")) {
        (void)hipStreamSynchronize(c.stream);
        return HIPSW_E_DEVICE;
    }
    int rc = nw_run_ctx(eng, c, &set, mode, stats);
    if (rc == HIPSW_OK)
        rc = nw_download_ctx(eng, c.stream, &set, out_score, out_cigar, out_n_cigar, out_flags,
                             out_final_w);
    if (stats) stats->total_ms = now_ms() - t0;
    return rc;
}

extern "C" int hipsw_nw_set_state(hipsw_engine *eng, int32_t form)
{
    if (!eng || form < 0 || form > 3) return HIPSW_E_INVALID;
    eng->nw_form.store(form);
    return HIPSW_OK;
}
