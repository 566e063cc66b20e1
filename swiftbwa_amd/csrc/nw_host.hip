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
    void *dmem = nullptr;
    NwDev dev;
    int32_t *d_w = nullptr;
    uint32_t *d_order = nullptr;
    WorkItem *d_items = nullptr;
    std::vector<int32_t> final_w;
    // first-pass plans, per mode (the retry passes are planned on the fly)
    bool plan_ready[2] = {false, false};
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

extern "C" hipsw_nw_set *hipsw_nw_upload(hipsw_engine *eng, int32_t n, const uint8_t *queries,
                                         const uint32_t *q_off, const uint16_t *q_len,
                                         const uint8_t *targets, const uint32_t *t_off,
                                         const uint16_t *t_len, const int32_t *w,
                                         const int32_t *true_score, int32_t max_retries,
                                         int32_t cap, const hipsw_scoring *sc)
{
    if (!eng) return nullptr;
    std::lock_guard<std::recursive_mutex> lock(eng->mu);
    if (n < 0 || cap < 1 || max_retries < 0 ||
        (n > 0 && (!queries || !q_off || !q_len || !targets || !t_off || !t_len || !w ||
                   !true_score))) {
        eng->last_error = "hipsw_nw_upload: invalid argument";
        return nullptr;
    }
    if (!scoring_valid(sc)) {
        eng->last_error = "hipsw_nw_upload: gap extension penalties must be > 0";
        return nullptr;
    }
    hipSetDevice(eng->device);
    auto *set = new hipsw_nw_set();
    set->n = n;
    set->max_retries = max_retries;
    set->cap = cap;
    memset(&set->params, 0, sizeof set->params);
    int mat_abs = 0;
    for (int k = 0; k < 25; ++k) {
        set->params.mat[k] = sc->mat[k];
        mat_abs = std::max(mat_abs, std::abs((int)sc->mat[k]));
    }
    set->params.o_del = sc->o_del; set->params.e_del = sc->e_del;
    set->params.o_ins = sc->o_ins; set->params.e_ins = sc->e_ins;
    set->params.oe_del = sc->o_del + sc->e_del;
    set->params.oe_ins = sc->o_ins + sc->e_ins;
    set->max_pen = std::max(mat_abs, std::max(set->params.oe_del, set->params.oe_ins));

    PackedJobs pj;
    std::string err;
    if (!pack_jobs(n, queries, q_off, q_len, targets, t_off, t_len, pj, err)) {
        eng->last_error = "hipsw_nw_upload: " + err;
        delete set;
        return nullptr;
    }
    set->ql = pj.ql;
    set->tl = pj.tl;
    set->w0.assign(w, w + n);
    set->ts.assign(true_score, true_score + n);
    for (int k = 0; k < n; ++k) {
        if (set->w0[k] < 1) {
            // GlobalAligner's first-row loop `for j in 1...initEnd` needs w >= 1
            eng->last_error = "hipsw_nw_upload: w must be >= 1";
            delete set;
            return nullptr;
        }
    }
    const size_t words = pj.seq.size(), nwords = pj.nmask.size();
    const size_t items_cap = (size_t)n / 64 + 64;
    const size_t bytes = span<uint32_t>(words) + span<uint32_t>(nwords) + 4 * span<uint32_t>(n) +
                         3 * span<int32_t>(n) + span<int32_t>(n) + span<uint32_t>((size_t)n * cap) +
                         2 * span<int32_t>(n) + span<unsigned long long>(1) +
                         span<uint32_t>(n) + span<WorkItem>(items_cap);
    if (!eng->check(hipMalloc(&set->dmem, bytes), "hipMalloc(nw set)")) {
        delete set;
        return nullptr;
    }
    char *cur = reinterpret_cast<char *>(set->dmem);
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
    set->d_items = carve<WorkItem>(cur, items_cap);
    bool ok = true;
    auto up = [&](void *dst, const void *src, size_t b) {
        if (ok && b) ok = eng->check(hipMemcpyAsync(dst, src, b, hipMemcpyHostToDevice, eng->stream), "upload");
    };
    up(d_seq, pj.seq.data(), words * 4);
    up(d_nmask, pj.nmask.data(), nwords * 4);
    up(d_qw, pj.qw.data(), (size_t)n * 4);
    up(d_tw, pj.tw.data(), (size_t)n * 4);
    up(d_qn, pj.qnw.data(), (size_t)n * 4);
    up(d_tn, pj.tnw.data(), (size_t)n * 4);
    up(d_ql, pj.ql.data(), (size_t)n * 4);
    up(d_tl, pj.tl.data(), (size_t)n * 4);
    if (ok) ok = eng->check(hipStreamSynchronize(eng->stream), "upload sync");
    if (!ok) {
        hipFree(set->dmem);
        delete set;
        return nullptr;
    }
    set->dev = NwDev{d_seq, d_nmask, d_qw, d_tw, d_qn, d_tn, d_ql, d_tl, set->d_w, d_score,
                     d_cig, d_ncig, d_flags, d_cells, cap};
    set->final_w = set->w0;
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
    auto add_launch = [&](int tr, int s_max, int c_max, int z_rows, int ib, int ic, int q_max,
                          size_t z_max) {
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
        const size_t lds_used = nw_lds_bytes(tr, L.global_state, s_max, L.qw8);
        const int wpc = std::max(1, std::min(16, (int)(eng->lds_per_cu / (lds_used + 512))));
        L.grid = std::max(1, std::min(ic, eng->n_cu * wpc));
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
        add_launch(tr, smax, cmax, zr, 0, (int)items.size(), qm, zm);
        return;
    }
    // Batched: bins by (tier, S class); inside a bin lanes of one wave share
    // the band width and have similar target lengths (ncol desc, tlen desc).
    std::vector<uint32_t> keys(m), idx(m);
    for (int a = 0; a < m; ++a) {
        const uint32_t bin = (uint32_t)(tier[a] * 8 + s_class(S[a]));
        const uint32_t nc = (uint32_t)std::min(ncol[a], 4095);
        const uint32_t tl = (uint32_t)std::min(tlv[active[a]], 65535);
        keys[a] = (bin << 28) | ((4095u - nc) << 16) | (65535u - tl);
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
        for (int g = s; g < e; g += 64) items.push_back(WorkItem{(uint32_t)g, (uint32_t)std::min(64, e - g)});
        add_launch((int)bin / 8, smax, cmax, zr, ib, (int)items.size() - ib, qm, zm);
        s = e;
    }
}

extern "C" int hipsw_nw_run(hipsw_engine *eng, hipsw_nw_set *set, int32_t mode,
                            hipsw_stats *stats)
{
    if (!eng || !set) return HIPSW_E_INVALID;
    if (mode != HIPSW_MODE_PLAIN && mode != HIPSW_MODE_BATCHED) return HIPSW_E_INVALID;
    std::lock_guard<std::recursive_mutex> lock(eng->mu);
    hipSetDevice(eng->device);
    const double t0 = now_ms();
    const int n = set->n;
    double kernel_ms = 0;
    int64_t launches_total = 0;
    std::vector<int32_t> w = set->w0;
    std::vector<int> tries(n, 0);
    std::vector<uint32_t> active(n);
    std::iota(active.begin(), active.end(), 0u);
    std::vector<int32_t> score(n);
    std::vector<uint32_t> order;
    std::vector<WorkItem> items;
    std::vector<NwLaunch> launches;
    if (n > 0 && !eng->check(launch_zero_u32(reinterpret_cast<uint32_t *>(set->dev.cells), 2,
                                             eng->stream), "zero"))
        return HIPSW_E_DEVICE;
    for (int pass = 0; !active.empty(); ++pass) {
        if (pass == 0 && set->plan_ready[mode]) {
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
            }
        }
        if (launches.size() > 64) {
            eng->last_error = "hipsw_nw_run: too many bins";
            return HIPSW_E_UNSUPPORTED;
        }
        // Scratch: backpointers [row][c/8][lane] per resident wave, plus DP
        // state for launches whose state does not fit in LDS.  Launches run in
        // order on one stream, so they share one region.
        size_t zbytes = 0, gbytes = 0;
        for (const NwLaunch &L : launches) {
            if (L.wide) {
                if (L.zmax > 49152) zbytes = std::max(zbytes, (size_t)L.grid * L.zmax);
                continue;
            }
            const size_t ncw = (size_t)(L.ncol_max + 7) / 8;
            zbytes = std::max(zbytes, (size_t)L.grid * L.zrows * ncw * 64 * 4);
            if (L.global_state)
                gbytes = std::max(gbytes, (size_t)L.grid * L.S * 64 * (L.tier == NW_I16 ? 4 : 8));
        }
        const size_t need = zbytes + ((gbytes + 255) & ~(size_t)255) + 256;
        if (need > eng->nw_dev_bytes) {
            eng->free_nw_buffers();
            if (!eng->check(hipMalloc(&eng->nw_dev, need), "hipMalloc(nw scratch)")) return HIPSW_E_DEVICE;
            eng->nw_dev_bytes = need;
        }
        uint32_t *zbase = reinterpret_cast<uint32_t *>(eng->nw_dev);
        void *gstate = reinterpret_cast<char *>(eng->nw_dev) + ((zbytes + 255) & ~(size_t)255);
        bool ok = eng->check(hipMemcpyAsync(set->d_w, w.data(), (size_t)n * 4,
                                            hipMemcpyHostToDevice, eng->stream), "upload w");
        if (ok) ok = eng->check(hipMemcpyAsync(set->d_order, order.data(), order.size() * 4,
                                               hipMemcpyHostToDevice, eng->stream), "upload order");
        if (ok) ok = eng->check(hipMemcpyAsync(set->d_items, items.data(), items.size() * sizeof(WorkItem),
                                               hipMemcpyHostToDevice, eng->stream), "upload items");
        if (ok) ok = eng->check(launch_zero_u32(eng->counters, 64, eng->stream), "zero");
        if (ok) ok = eng->check(hipEventRecord(eng->ev0, eng->stream), "event");
        int k = 0;
        for (const NwLaunch &L : launches) {
            if (!ok) break;
            if (L.wide) {
                // the bin's jobs are order[items[ib].start ..) contiguously
                const WorkItem &first = items[L.item_begin];
                const WorkItem &last = items[L.item_begin + L.item_count - 1];
                const int nj = (int)(last.start + last.count - first.start);
                ok = eng->check(launch_nw_wide(set->dev, set->params, set->d_order + first.start, nj,
                                               eng->counters + k, L.qmax,
                                               reinterpret_cast<uint8_t *>(zbase), L.zmax,
                                               L.zmax <= 49152, L.grid, eng->stream),
                                "launch nw wide");
                ++k;
                ++launches_total;
                continue;
            }
            const int ncw = (L.ncol_max + 7) / 8;
            ok = eng->check(launch_nw(L.tier, L.global_state, set->dev, set->params, set->d_order,
                                      set->d_items + L.item_begin, L.item_count,
                                      eng->counters + k, L.S, zbase, L.zrows, ncw, L.grid,
                                      pass == 0 && mode == HIPSW_MODE_BATCHED, gstate, L.qw8,
                                      eng->stream),
                            "launch nw");
            ++k;
            ++launches_total;
        }
        if (ok) ok = eng->check(hipEventRecord(eng->ev1, eng->stream), "event");
        if (ok) ok = eng->check(hipEventSynchronize(eng->ev1), "nw kernel");
        if (!ok) return HIPSW_E_DEVICE;
        float ms = 0.f;
        hipEventElapsedTime(&ms, eng->ev0, eng->ev1);
        kernel_ms += ms;
        if (!eng->check(hipMemcpy(score.data(), set->dev.score, (size_t)n * 4, hipMemcpyDeviceToHost),
                        "download scores"))
            return HIPSW_E_DEVICE;
        // CIGARGenerator.swift:88-96: retry while score < trueScore.
        std::vector<uint32_t> next;
        for (uint32_t j : active) {
            set->final_w[j] = w[j];
            if (score[j] < set->ts[j] && tries[j] < set->max_retries) {
                w[j] = w[j] < (1 << 20) ? w[j] * 2 : w[j];
                tries[j] += 1;
                next.push_back(j);
            }
        }
        active.swap(next);
    }
    if (stats) {
        memset(stats, 0, sizeof *stats);
        stats->kernel_ms = kernel_ms;
        stats->total_ms = now_ms() - t0;
        stats->n_jobs = n;
        stats->launches = launches_total;
        unsigned long long c = 0;
        if (n > 0) hipMemcpy(&c, set->dev.cells, sizeof c, hipMemcpyDeviceToHost);
        stats->cells = (int64_t)c;
    }
    return HIPSW_OK;
}

extern "C" int hipsw_nw_download(hipsw_engine *eng, hipsw_nw_set *set, int32_t *out_score,
                                 uint32_t *out_cigar, int32_t *out_n_cigar, int32_t *out_flags,
                                 int32_t *out_final_w)
{
    if (!eng || !set) return HIPSW_E_INVALID;
    std::lock_guard<std::recursive_mutex> lock(eng->mu);
    const int n = set->n;
    if (n == 0) return HIPSW_OK;
    hipSetDevice(eng->device);
    bool ok = true;
    auto down = [&](void *dst, const void *src, size_t b) {
        if (ok && dst) ok = eng->check(hipMemcpy(dst, src, b, hipMemcpyDeviceToHost), "download");
    };
    std::vector<int32_t> ncig(n);
    down(out_score, set->dev.score, (size_t)n * 4);
    down(out_cigar, set->dev.cigar, (size_t)n * set->cap * 4);
    down(ncig.data(), set->dev.n_cigar, (size_t)n * 4);
    if (ok && out_n_cigar) memcpy(out_n_cigar, ncig.data(), (size_t)n * 4);
    if (ok && out_cigar)                                       // entries past the CIGAR are zero
        for (int k = 0; k < n; ++k) {
            const int used = std::min(std::max(ncig[k], 0), set->cap);
            std::fill(out_cigar + (size_t)k * set->cap + used, out_cigar + (size_t)(k + 1) * set->cap, 0u);
        }
    down(out_flags, set->dev.flags, (size_t)n * 4);
    if (out_final_w) memcpy(out_final_w, set->final_w.data(), (size_t)n * 4);
    return ok ? HIPSW_OK : HIPSW_E_DEVICE;
}

extern "C" void hipsw_nw_free(hipsw_engine *eng, hipsw_nw_set *set)
{
    if (!set) return;
    if (eng) {
        std::lock_guard<std::recursive_mutex> lock(eng->mu);
        hipSetDevice(eng->device);
        hipStreamSynchronize(eng->stream);
        if (set->dmem) hipFree(set->dmem);
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
    if (n < 0 || !out_score || !out_cigar || !out_n_cigar || !out_flags) return HIPSW_E_INVALID;
    std::lock_guard<std::recursive_mutex> lock(eng->mu);
    const double t0 = now_ms();
    hipsw_nw_set *set = hipsw_nw_upload(eng, n, queries, q_off, q_len, targets, t_off, t_len, w,
                                        true_score, max_retries, cap, sc);
    if (!set) return HIPSW_E_INVALID;
    int rc = hipsw_nw_run(eng, set, mode, stats);
    if (rc == HIPSW_OK)
        rc = hipsw_nw_download(eng, set, out_score, out_cigar, out_n_cigar, out_flags, out_final_w);
    hipsw_nw_free(eng, set);
    if (stats) stats->total_ms = now_ms() - t0;
    return rc;
}
