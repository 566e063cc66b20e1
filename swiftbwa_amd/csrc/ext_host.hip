// Host side of the banded-extension entry points (include/hipsw.h).
//
// Replaces BandedSWDispatcher.dispatchBatch8/16 (Sources/MetalSW/
// BandedSWDispatcher.swift:48-158).  A call packs the tasks to 2-bit words,
// picks an exact score tier per task, uploads one arena, runs the DP kernels
// and copies SWResult rows back.
#include "engine_internal.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>

using namespace hipsw;

namespace {

// Query-length classes of the batched form: each class is one persistent
// launch whose LDS footprint is sized to the class' longest query.  The bounds
// end one short of a 16-column multiple, so a class fills its packed-form LDS
// exactly (columns 0..qlen): 150-bp queries take 160 columns, four waves a CU.
constexpr int kClassBounds[] = {63, 159, 255, 1 << 30};
constexpr int kClasses = 4;

int class_of(int qlen)
{
    for (int c = 0; c < kClasses; ++c)
        if (qlen <= kClassBounds[c]) return c;
    return kClasses - 1;
}

int tier_of(int64_t h0, int64_t qlen, int max_mat)
{
    const int64_t bound = h0 + qlen * (int64_t)max_mat;
    if (h0 >= 0 && bound <= 255) return TIER_U8;
    if (h0 >= 0 && bound <= 65535) return TIER_U16;
    return TIER_I32;
}

// Waves (64-thread blocks) resident per CU for a given dynamic LDS size.
int waves_per_cu(const hipsw_engine *eng, size_t lds)
{
    const size_t per_cu = eng->lds_per_cu > 0 ? (size_t)eng->lds_per_cu : 65536;
    int w = (int)(per_cu / std::max<size_t>(lds + 512, 1));
    return std::max(1, std::min(w, 8));
}

template <typename T>
T *carve(char *&cur, size_t count)
{
    T *p = reinterpret_cast<T *>(cur);
    cur += (count * sizeof(T) + 255) & ~(size_t)255;
    return p;
}

template <typename T>
size_t span(size_t count)
{
    return (count * sizeof(T) + 255) & ~(size_t)255;
}

}  // namespace

namespace hipsw {

// Batched form: bin by (tier, query-length class, N present); inside a bin
// order by band w, widest first (in steps of 4 columns, 255 and wider as one),
// then by estimated DP area, largest first (steps of 256 cells, 2 M and more
// as one), and within an area step by h0, largest first.  A wave sweeps the union of its jobs' bands row
// by row, and a row's band starts at i - w (:90-91): jobs of one w share its
// left edge, and h0 sets how fast that edge moves once z-drop and zero cells
// narrow it.  Then cut into work items (the bin stays close to longest
// first); one persistent launch per bin.  Bins without N, with the standard matrix and queries below
// 256 bases run the register-query kernels: the packed form for the 8-bit
// tier (unless the engine's tuning table picks the one-job-per-lane form for
// that class), the one-job-per-lane form for the 16-bit tier.  Returns the
// global-state scratch bytes the launches need.
size_t plan_ext_batched(const hipsw_engine *eng, const TuneTable &tune, int n, const int32_t *ql,
                        const int32_t *tl, const int32_t *h0, const int32_t *w,
                        const uint8_t *has_n, int max_mat, bool std_mat, int e_max,
                        std::vector<uint32_t> &order, std::vector<WorkItem> &items,
                        std::vector<ExtLaunch> &launches)
{
    std::vector<uint32_t> keys(n), idx(n);
    parallel_for(n, host_threads(), [&](int64_t lo, int64_t hi) {
        for (int64_t k = lo; k < hi; ++k) {
            const uint32_t bin = (uint32_t)((tier_of(h0[k], ql[k], max_mat) * kClasses + class_of(ql[k])) * 2 +
                                            (has_n && has_n[k] ? 1 : 0));
            const int64_t wk = std::max(w[k], 0);
            const uint64_t band = (uint64_t)std::min<int64_t>(ql[k], 2 * wk + 1);
            const uint32_t wq = (uint32_t)std::min<int64_t>(wk, 255) >> 2;               // 6 bits
            const uint32_t aq = (uint32_t)std::min<uint64_t>((uint64_t)tl[k] * band >> 8, 8191);   // 13 bits
            const uint32_t h0c = (uint32_t)std::min(std::max(h0[k], 0), 255);
            keys[k] = (bin << 27) | ((63u - wq) << 21) | ((8191u - aq) << 8) | (255u - h0c);
            idx[k] = (uint32_t)k;
        }
    });
    radix_sort_by_key(keys, idx);
    order = idx;
    launches.clear();
    for (int s = 0; s < n;) {
        const uint32_t bin = keys[s] >> 27;
        int e = s, cmax = 0;
        int64_t bmax = 0;                                      // the bin's largest score bound
        while (e < n && (keys[e] >> 27) == bin) {
            cmax = std::max(cmax, ql[idx[e]]);
            bmax = std::max(bmax, (int64_t)h0[idx[e]] + (int64_t)ql[idx[e]] * max_mat);
            ++e;
        }
        ExtLaunch L;
        const bool with_n = (bin & 1) != 0;
        L.tier = (int)(bin >> 1) / kClasses;
        L.cls = (int)(bin >> 1) % kClasses;
        L.has_n = with_n ? 1 : 0;
        L.cols = cmax + 1;
        const bool fast = !with_n && std_mat && cmax <= 255 && L.tier != TIER_I32 &&
                          ext_pk_lds_bytes(L.cols, ext_pk_default_lb(L.cols)) <=
                              (size_t)eng->lds_per_block;
        L.w16_ok = fast && L.tier == TIER_U16 && bmax <= 2047 && e_max <= 31;
        L.kind = !fast ? EXT_UNPACKED
                 : L.tier == TIER_U16 ? (tune.kind16[L.cls] == EXT_PACKED16 && L.w16_ok ? EXT_PACKED16 : EXT_LANE)
                                      : tune.kind[L.cls] > 0 ? tune.kind[L.cls] : EXT_PACKED;
        L.lb = 0;
        if (L.kind != EXT_UNPACKED) {
            const int t = L.tier == TIER_U8    ? tune.lb[L.cls]
                          : L.tier == TIER_U16 ? tune.lb16[L.cls]
                                               : 0;
            L.lb = t > 0 && ext_pk_lb_supported(L.kind, L.cols, t) ? t : ext_pk_default_lb_k(L.kind, L.cols);
        }
        L.global_state = L.kind == EXT_UNPACKED &&
                         ext_lds_bytes(L.tier, false, L.cols) > (size_t)eng->lds_per_block;
        L.job_begin = s;
        L.job_end = e;
        launches.push_back(L);
        s = e;
    }
    return ext_layout_launches(eng, launches, items);
}

// Waves of a packed / one-per-lane bin that one CU holds (whole multiples of
// the wave's LDS share, at most 8).
static int pk_waves_per_cu(const hipsw_engine *eng, const ExtLaunch &L)
{
    const size_t per_cu = eng->lds_per_cu > 0 ? (size_t)eng->lds_per_cu : 65536;
    return (int)std::max<size_t>(1, std::min<size_t>(8, per_cu / ext_pk_lds_bytes_k(L.kind, L.cols, L.lb)));
}

size_t ext_layout_launches(const hipsw_engine *eng, std::vector<ExtLaunch> &launches,
                           std::vector<WorkItem> &items)
{
    items.clear();
    size_t need = 0;
    for (ExtLaunch &L : launches) {
        ext_bin_items(eng, L, items);
        // one region per resident block of the launch (the kernels index it
        // by blockIdx.x); the bins run in turn on one stream and share it
        if (L.kind == EXT_UNPACKED)
            L.scratch_bytes = L.global_state ? (size_t)L.grid * L.cols * 64 * ext_state_bytes(L.tier) : 0;
        else
            L.scratch_bytes = (size_t)L.grid * ext_pk_hbm_bytes_per_wave_k(L.kind, L.cols, L.lb);
        need = std::max(need, (L.scratch_bytes + 255) & ~(size_t)255);
    }
    return need;
}

size_t ext_scratch_bound(const hipsw_engine *eng, int n, const int32_t *ql, const int32_t *h0,
                         const uint8_t *has_n, int max_mat, bool std_mat)
{
    // per (tier, class, N) bin: its job count and longest query (per-thread
    // tallies merged under a lock)
    constexpr int kBins = 3 * kClasses * 2;
    int64_t cnt[kBins] = {0};
    int cmax[kBins] = {0};
    std::mutex mu;
    parallel_for(n, host_threads(), [&](int64_t lo, int64_t hi) {
        int64_t c[kBins] = {0};
        int m[kBins] = {0};
        for (int64_t k = lo; k < hi; ++k) {
            const int b = (tier_of(h0[k], ql[k], max_mat) * kClasses + class_of(ql[k])) * 2 +
                          (has_n && has_n[k] ? 1 : 0);
            c[b]++;
            m[b] = std::max(m[b], ql[k]);
        }
        std::lock_guard<std::mutex> lk(mu);
        for (int b = 0; b < kBins; ++b) {
            cnt[b] += c[b];
            cmax[b] = std::max(cmax[b], m[b]);
        }
    });
    size_t bound = 0;
    for (int b = 0; b < kBins; ++b) {
        if (!cnt[b]) continue;
        ExtLaunch L{};
        L.tier = (b >> 1) / kClasses;
        L.cls = (b >> 1) % kClasses;
        L.has_n = b & 1;
        L.cols = cmax[b] + 1;
        L.job_begin = 0;
        L.job_end = (int)cnt[b];
        const bool fast = !(b & 1) && std_mat && cmax[b] <= 255 && L.tier != TIER_I32 &&
                          ext_pk_lds_bytes(L.cols, ext_pk_default_lb(L.cols)) <=
                              (size_t)eng->lds_per_block;
        std::vector<ExtLaunch> cand;
        if (!fast) {
            L.kind = EXT_UNPACKED;
            L.lb = 0;
            L.global_state = ext_lds_bytes(L.tier, false, L.cols) > (size_t)eng->lds_per_block;
            cand.push_back(L);
        } else {
            for (int kind : {EXT_PACKED, EXT_LANE, EXT_PACKED16}) {
                if (kind == EXT_PACKED16 && L.tier != TIER_U16) continue;
                int lbs[5];
                const int nlb = ext_pk_lb_candidates(kind, L.cols, lbs);
                for (int c = 0; c < nlb; ++c) {
                    ExtLaunch K = L;
                    K.kind = kind;
                    K.lb = lbs[c];
                    K.global_state = false;
                    cand.push_back(K);
                }
            }
        }
        for (ExtLaunch &K : cand) {
            std::vector<WorkItem> items;
            std::vector<ExtLaunch> one{K};
            bound = std::max(bound, ext_layout_launches(eng, one, items));
        }
    }
    return bound;
}

void ext_bin_items(const hipsw_engine *eng, ExtLaunch &L, std::vector<WorkItem> &items)
{
    const int per_item = L.kind == EXT_PACKED || L.kind == EXT_PACKED16 ? 128 : 64;
    L.item_begin = (int)items.size();
    for (int g = L.job_begin; g < L.job_end; g += per_item)
        items.push_back(WorkItem{(uint32_t)g, (uint32_t)std::min(per_item, L.job_end - g)});
    L.item_count = (int)items.size() - L.item_begin;
    int wpc;
    if (L.kind != EXT_UNPACKED)   // no static shared memory: the CU holds whole multiples
        wpc = pk_waves_per_cu(eng, L);
    else
        wpc = L.global_state ? 8 : waves_per_cu(eng, ext_lds_bytes(L.tier, false, L.cols));
    L.grid = std::max(1, std::min(L.item_count, eng->n_cu * wpc));
}

// One bin's launch.
hipError_t launch_ext_bin(const ExtLaunch &L, const ExtDev &d, const ExtParams &p,
                          const uint32_t *order, const WorkItem *items, uint32_t *counter,
                          void *gscratch, hipStream_t s)
{
    void *gs = L.scratch_bytes ? gscratch : nullptr;   // every bin starts at the region's base
    switch (L.kind) {
    case EXT_PACKED:
        return launch_ext_persistent_pk(d, p, items + L.item_begin, L.item_count, counter, L.cols,
                                        L.lb, L.grid, gs, s);
    case EXT_LANE:
        return launch_ext_persistent_ln(d, p, items + L.item_begin, L.item_count, counter, L.cols,
                                        L.lb, L.grid, gs, s);
    case EXT_PACKED16:
        return launch_ext_persistent_pk16(d, p, items + L.item_begin, L.item_count, counter, L.cols,
                                          L.lb, L.grid, gs, s);
    default:
        return launch_ext_persistent(L.tier, L.global_state, d, p, order, items + L.item_begin,
                                     L.item_count, counter, L.cols, L.grid, gs, s);
    }
}

}  // namespace hipsw

namespace hipsw {

int ext_prepare(const hipsw_engine *eng, const TuneTable &tune, int n, const uint8_t *queries,
                const uint32_t *q_off, const uint16_t *q_len, const uint8_t *targets,
                const uint32_t *t_off, const uint16_t *t_len, const int32_t *h0,
                const int32_t *w, const int32_t *end_bonus, const hipsw_scoring *sc,
                ExtPrep &P, std::string &err)
{
    if (n < 0 || (n > 0 && (!queries || !q_off || !q_len || !targets || !t_off || !t_len ||
                            !h0 || !w || !end_bonus))) {
        err = "null argument";
        return HIPSW_E_INVALID;
    }
    if (!scoring_valid(sc)) {
        err = "gap extension penalties must be > 0";
        return HIPSW_E_INVALID;
    }
    for (int k = 0; k < n; ++k) {
        // a negative band puts the row's end below column 0 (BandedSWScalar.swift:97-99)
        if (w[k] < 0) {
            err = "band width w must be >= 0";
            return HIPSW_E_INVALID;
        }
    }
    // P may come back from an earlier call: every array below is rewritten
    // in full (the no-fill ones element by element), the shape reset
    P.shape = hipsw_ext_set();
    hipsw_ext_set &S = P.shape;
    P.n = S.n = n;
    P.params = S.params = make_params(sc);
    const int max_mat = P.params.max_mat;
    const int threads = host_threads();

    // Word layout: every query and target starts on a 16-base word boundary,
    // sequences back to back.  Chunks of jobs sum their words, the chunk
    // totals are scanned, then each chunk writes its jobs' offsets.  The
    // packing pass reads each base once and also finds the sequences holding
    // an N; their mask words (placed by a second scan of the flags) are
    // filled afterwards from those sequences alone.
    P.qw.resize(n); P.tw.resize(n); P.qnw.resize(n); P.tnw.resize(n);
    std::vector<uint8_t> qn_flag(n), tn_flag(n);
    const int nch = (threads > 1 && n >= 4096) ? threads : 1;
    const int64_t chunk = ((int64_t)n + nch - 1) / nch;
    std::vector<uint64_t> cw(nch + 1, 0), cn(nch + 1, 0);
    parallel_chunks(nch, [&](int c) {
        uint64_t a = 0;
        for (int64_t k = c * chunk, e = std::min<int64_t>(n, (c + 1) * chunk); k < e; ++k)
            a += (q_len[k] + 15) / 16 + (t_len[k] + 15) / 16;
        cw[c + 1] = a;
    });
    for (int c = 0; c < nch; ++c) cw[c + 1] += cw[c];
    uint64_t words = cw[nch];
    if (words + 64 > 0xFFFFFFF0ull) {
        err = "job set exceeds 2^32 sequence words";
        return HIPSW_E_UNSUPPORTED;
    }
    parallel_chunks(nch, [&](int c) {
        uint64_t a = cw[c];
        for (int64_t k = c * chunk, e = std::min<int64_t>(n, (c + 1) * chunk); k < e; ++k) {
            P.qw[k] = (uint32_t)a;
            a += (q_len[k] + 15) / 16;
            P.tw[k] = (uint32_t)a;
            a += (t_len[k] + 15) / 16;
        }
    });
    words += 64;    // tail padding: whole-word reads past the last sequence stay in bounds
    // every sequence word is written by the packing below; only the padding
    // is cleared here
    P.seq.resize(words);
    std::fill(P.seq.end() - 64, P.seq.end(), 0u);
    P.ql.resize(n); P.tl.resize(n); P.h0.resize(n); P.w.resize(n); P.eb.resize(n);
    std::atomic<bool> bad{false};
    parallel_for(n, threads, [&](int64_t lo, int64_t hi) {
        for (int64_t k = lo; k < hi; ++k) {
            P.ql[k] = q_len[k];
            P.tl[k] = t_len[k];
            P.h0[k] = h0[k];
            P.w[k] = w[k];
            P.eb[k] = end_bonus[k];
            const int rq = pack2_scan(queries + q_off[k], q_len[k], P.seq.data() + P.qw[k]);
            const int rt = pack2_scan(targets + t_off[k], t_len[k], P.seq.data() + P.tw[k]);
            if (rq < 0 || rt < 0) bad = true;
            qn_flag[k] = rq > 0;
            tn_flag[k] = rt > 0;
        }
    });
    if (bad) {
        err = "base code > 4";
        return HIPSW_E_INVALID;
    }
    parallel_chunks(nch, [&](int c) {
        uint64_t b = 0;
        for (int64_t k = c * chunk, e = std::min<int64_t>(n, (c + 1) * chunk); k < e; ++k)
            b += (qn_flag[k] ? (q_len[k] + 31) / 32 : 0) + (tn_flag[k] ? (t_len[k] + 31) / 32 : 0);
        cn[c + 1] = b;
    });
    for (int c = 0; c < nch; ++c) cn[c + 1] += cn[c];
    const uint64_t nwords = cn[nch] + 64;
    P.nmask.assign(nwords, 0);
    parallel_chunks(nch, [&](int c) {
        uint64_t b = cn[c];
        for (int64_t k = c * chunk, e = std::min<int64_t>(n, (c + 1) * chunk); k < e; ++k) {
            P.qnw[k] = qn_flag[k] ? (uint32_t)b : kNoN;
            if (qn_flag[k]) {
                nmask_fill(queries + q_off[k], q_len[k], P.nmask.data() + b);
                b += (q_len[k] + 31) / 32;
            }
            P.tnw[k] = tn_flag[k] ? (uint32_t)b : kNoN;
            if (tn_flag[k]) {
                nmask_fill(targets + t_off[k], t_len[k], P.nmask.data() + b);
                b += (t_len[k] + 31) / 32;
            }
        }
    });

    // Exact score tier per job; jobs holding an N.
    std::vector<uint8_t> any_n(n);
    std::vector<int64_t> ctier((size_t)nch * 3, 0);
    std::vector<int> cq(nch, 0);
    std::vector<uint8_t> csome(nch, 0);
    parallel_chunks(nch, [&](int c) {
        int64_t t3[3] = {0, 0, 0};
        int q = 0;
        uint8_t some = 0;
        for (int64_t k = c * chunk, e = std::min<int64_t>(n, (c + 1) * chunk); k < e; ++k) {
            t3[tier_of(P.h0[k], P.ql[k], max_mat)]++;
            q = std::max(q, (int)P.ql[k]);
            any_n[k] = qn_flag[k] | tn_flag[k];
            some |= any_n[k];
        }
        for (int t = 0; t < 3; ++t) ctier[(size_t)c * 3 + t] = t3[t];
        cq[c] = q;
        csome[c] = some;
    });
    int plain_tier = TIER_U8, qmax = 0;
    bool some_n = false;
    for (int c = 0; c < nch; ++c) {
        for (int t = 0; t < 3; ++t) {
            S.tier_jobs[t] += ctier[(size_t)c * 3 + t];
            if (ctier[(size_t)c * 3 + t]) plain_tier = std::max(plain_tier, t);
        }
        qmax = std::max(qmax, cq[c]);
        some_n |= csome[c] != 0;
    }

    // Plain form: input order, one tier and one LDS size for the whole launch.
    S.plain_tier = plain_tier;
    S.plain_cols = qmax + 1;
    const size_t plain_lds = ext_lds_bytes(plain_tier, false, S.plain_cols);
    S.plain_global = plain_lds > (size_t)eng->lds_per_block;
    const int groups = (n + 63) / 64;
    if (S.plain_global) {
        S.plain_grid = std::max(1, std::min(groups, eng->n_cu * 8));
        S.plain_scratch = (size_t)S.plain_grid * S.plain_cols * 64 * ext_state_bytes(plain_tier);
    } else {
        S.plain_grid = std::max(1, groups);
    }

    // Batched form (see plan_ext_batched).
    // Input-order form: one launch of the register-query kernel at the set's
    // widest tier and longest query, items in input order.
    if (n > 0 && !some_n && P.params.std_mat && qmax <= 255 && plain_tier != TIER_I32) {
        ExtLaunch &L = S.io_launch;
        L.tier = plain_tier;
        L.kind = plain_tier == TIER_U8 ? EXT_PACKED : EXT_LANE;
        L.cls = class_of(qmax);
        L.has_n = 0;
        L.cols = qmax + 1;
        L.lb = ext_pk_default_lb(L.cols);
        L.global_state = false;
        L.job_begin = 0;
        L.job_end = n;
        if (ext_pk_lds_bytes(L.cols, L.lb) <= (size_t)eng->lds_per_block) {
            const int per_item = L.kind == EXT_PACKED ? 128 : 64;
            L.item_begin = 0;
            L.item_count = (n + per_item - 1) / per_item;
            L.grid = L.item_count;                  // one item per block
            L.scratch_bytes = (size_t)L.grid * ext_pk_hbm_bytes_per_wave(L.cols, L.lb);
            S.io_scratch = (L.scratch_bytes + 255) & ~(size_t)255;
            S.io_ok = true;
        }
    }
    S.batched_scratch = plan_ext_batched(eng, tune, n, P.ql.data(), P.tl.data(), P.h0.data(),
                                         P.w.data(), any_n.data(), max_mat,
                                         P.params.std_mat != 0,
                                         std::max(P.params.e_del, P.params.e_ins), P.order,
                                         P.items, S.launches);
    // Work items of any later re-plan (hipsw_ext_autotune) fit: 64 jobs or more per item.
    S.items_cap = (size_t)n / 64 + S.launches.size() + 1;
    P.pk.resize(n);
    parallel_for(n, threads, [&](int64_t lo, int64_t hi) {
        build_pk_meta((int)(hi - lo), P.order.data() + lo, P.qw.data(), P.tw.data(), P.qnw.data(),
                      P.tnw.data(), P.ql.data(), P.tl.data(), P.h0.data(), P.w.data(),
                      P.eb.data(), P.pk.data() + lo);
    });
    P.arena_bytes = span<uint32_t>(words) + span<uint32_t>(nwords) + 4 * span<uint32_t>(n) +
                    6 * span<int32_t>(n) + span<int32_t>((size_t)n * 6) + span<uint32_t>(n) +
                    span<WorkItem>(S.items_cap) + span<PkMeta>(n);
    return HIPSW_OK;
}

PrepLease::PrepLease(hipsw_engine *eng) : eng_(eng)
{
    {
        std::lock_guard<std::mutex> lk(eng->prep_mu);
        if (!eng->prep_pool.empty()) {
            p_ = std::move(eng->prep_pool.back());
            eng->prep_pool.pop_back();
        }
    }
    if (!p_) p_.reset(new ExtPrep());
}

PrepLease::~PrepLease()
{
    std::lock_guard<std::mutex> lk(eng_->prep_mu);
    if (eng_->prep_pool.size() < (size_t)kEngineSlots) eng_->prep_pool.push_back(std::move(p_));
}

hipError_t ext_bind_upload(const ExtPrep &P, void *dmem, hipsw_ext_set *set, hipStream_t st)
{
    const hipsw_ext_set &S = P.shape;
    const int n = P.n;
    set->n = n;
    set->params = P.params;
    set->plain_tier = S.plain_tier;
    set->plain_cols = S.plain_cols;
    set->plain_global = S.plain_global;
    set->plain_grid = S.plain_grid;
    set->plain_scratch = S.plain_scratch;
    set->launches = S.launches;
    set->batched_scratch = S.batched_scratch;
    set->items_cap = S.items_cap;
    set->io_ok = S.io_ok;
    set->io_launch = S.io_launch;
    set->io_scratch = S.io_scratch;
    for (int k = 0; k < 3; ++k) set->tier_jobs[k] = S.tier_jobs[k];
    set->dmem = dmem;
    set->dmem_bytes = P.arena_bytes;
    char *cur = reinterpret_cast<char *>(dmem);
    uint32_t *d_seq = carve<uint32_t>(cur, P.seq.size());
    uint32_t *d_nmask = carve<uint32_t>(cur, P.nmask.size());
    uint32_t *d_qw = carve<uint32_t>(cur, n);
    uint32_t *d_tw = carve<uint32_t>(cur, n);
    uint32_t *d_qn = carve<uint32_t>(cur, n);
    uint32_t *d_tn = carve<uint32_t>(cur, n);
    int32_t *d_ql = carve<int32_t>(cur, n);
    int32_t *d_tl = carve<int32_t>(cur, n);
    int32_t *d_h0 = carve<int32_t>(cur, n);
    int32_t *d_w = carve<int32_t>(cur, n);
    int32_t *d_eb = carve<int32_t>(cur, n);
    int32_t *d_cells = carve<int32_t>(cur, std::max(n, 2));   // 8-byte cell counter
    int32_t *d_out = carve<int32_t>(cur, (size_t)n * 6);
    set->d_order = carve<uint32_t>(cur, n);
    set->d_items = carve<WorkItem>(cur, S.items_cap);
    PkMeta *d_pk = carve<PkMeta>(cur, n);
    set->dev = ExtDev{d_seq, d_nmask, d_qw, d_tw, d_qn, d_tn, d_ql, d_tl, d_h0, d_w, d_eb, d_out,
                      reinterpret_cast<unsigned long long *>(d_cells), d_pk};
    hipError_t e = hipSuccess;
    auto up = [&](void *dst, const void *src, size_t b) {
        if (e == hipSuccess && b) e = hipMemcpyAsync(dst, src, b, hipMemcpyHostToDevice, st);
    };
    up(d_seq, P.seq.data(), P.seq.size() * 4);
    up(d_nmask, P.nmask.data(), P.nmask.size() * 4);
    up(d_qw, P.qw.data(), (size_t)n * 4);
    up(d_tw, P.tw.data(), (size_t)n * 4);
    up(d_qn, P.qnw.data(), (size_t)n * 4);
    up(d_tn, P.tnw.data(), (size_t)n * 4);
    up(d_ql, P.ql.data(), (size_t)n * 4);
    up(d_tl, P.tl.data(), (size_t)n * 4);
    up(d_h0, P.h0.data(), (size_t)n * 4);
    up(d_w, P.w.data(), (size_t)n * 4);
    up(d_eb, P.eb.data(), (size_t)n * 4);
    up(set->d_order, P.order.data(), (size_t)n * 4);
    up(set->d_items, P.items.data(), P.items.size() * sizeof(WorkItem));
    up(d_pk, P.pk.data(), (size_t)n * sizeof(PkMeta));
    return e;
}

// Descriptors and work items of the input-order form: job k's descriptor at
// position k, item i = jobs [i * per_item, ...).
__global__ void build_io_kernel(ExtDev d, int n, int per_item, PkMeta *pk, WorkItem *items)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        pk[k] = PkMeta{(uint32_t)k, d.q_word[k], d.t_word[k], (uint16_t)d.q_len[k],
                       (uint16_t)d.t_len[k], d.h0[k], d.w[k], d.eb[k], 0u};
        if (k % per_item == 0)
            items[k / per_item] = WorkItem{(uint32_t)k, (uint32_t)min(per_item, n - k)};
    }
}

static size_t io_pk_bytes(int n) { return ((size_t)n * sizeof(PkMeta) + 255) & ~(size_t)255; }

hipError_t ext_enqueue(const hipsw_ext_set *set, int mode, void *gs, uint32_t *counters,
                       hipStream_t st, hipEvent_t e0, hipEvent_t e1)
{
    if (mode == HIPSW_MODE_INPUT_ORDER) {
        if (!set->io_ok || !set->io_dmem) return hipErrorInvalidValue;
        ExtDev d = set->dev;
        d.pk = static_cast<const PkMeta *>(set->io_dmem);
        const WorkItem *items = reinterpret_cast<const WorkItem *>(
            static_cast<const char *>(set->io_dmem) + io_pk_bytes(set->n));
        hipError_t e = launch_zero_u32(reinterpret_cast<uint32_t *>(set->dev.cells), 2, st);
        if (e == hipSuccess && e0) e = hipEventRecord(e0, st);
        if (e == hipSuccess)
            e = launch_ext_bin(set->io_launch, d, set->params, nullptr, items, nullptr, gs, st);
        if (e == hipSuccess && e1) e = hipEventRecord(e1, st);
        return e;
    }
    hipError_t e = launch_zero_u32(reinterpret_cast<uint32_t *>(set->dev.cells), 2, st);
    if (e == hipSuccess && mode == HIPSW_MODE_BATCHED) e = launch_zero_u32(counters, 64, st);
    if (e == hipSuccess && e0) e = hipEventRecord(e0, st);
    if (e != hipSuccess) return e;
    if (mode == HIPSW_MODE_PLAIN) {
        e = launch_ext_plain(set->plain_tier, set->plain_global, set->dev, set->params, nullptr,
                             set->n, set->plain_cols, set->plain_grid, gs, st);
    } else {
        for (size_t b = 0; e == hipSuccess && b < set->launches.size(); ++b)
            e = launch_ext_bin(set->launches[b], set->dev, set->params, set->d_order, set->d_items,
                               counters + b, gs, st);
    }
    if (e == hipSuccess && e1) e = hipEventRecord(e1, st);
    return e;
}

}  // namespace hipsw

extern "C" hipsw_ext_set *hipsw_ext_upload(hipsw_engine *eng, int32_t n,
                                           const uint8_t *queries, const uint32_t *q_off,
                                           const uint16_t *q_len, const uint8_t *targets,
                                           const uint32_t *t_off, const uint16_t *t_len,
                                           const int32_t *h0, const int32_t *w,
                                           const int32_t *end_bonus, const hipsw_scoring *sc)
{
    if (!eng) return nullptr;
    PrepLease prep(eng);                                       // page-locked host halves, reused
    ExtPrep &P = prep.get();
    std::string err;
    if (ext_prepare(eng, eng->tune_snapshot(), n, queries, q_off, q_len, targets, t_off, t_len,
                    h0, w, end_bonus, sc, P, err) != HIPSW_OK) {
        eng->set_error("hipsw_ext_upload: " + err);
        return nullptr;
    }
    std::lock_guard<std::recursive_mutex> lock(eng->mu);
    hipSetDevice(eng->device);
    auto *set = new hipsw_ext_set();
    void *dmem = nullptr;
    if (!eng->check(hipMalloc(&dmem, P.arena_bytes), "hipMalloc(ext set)")) {
        delete set;
        return nullptr;
    }
    set->own_dmem = true;
    if (!eng->check(ext_bind_upload(P, dmem, set, eng->main.stream), "upload") ||
        !eng->check(hipStreamSynchronize(eng->main.stream), "upload sync")) {
        hipFree(dmem);
        delete set;
        return nullptr;
    }
    return set;
}

// The batched form of a set as one graph: the counter resets, then the bins
// back to back on the engine's main stream.  Captured once per plan (and again
// when the scratch region has moved); replays cost one launch per run.  The
// work-queue counters are reset by a kernel inside the graph (no memset node).
// Running the bins on concurrent streams instead, so that one bin's tail
// overlaps the next one's start, measured slower on the uneven 4M set (26.8
// against 25.5 ms): every persistent launch already fills the device.
static bool build_batched_graph(hipsw_engine *eng, hipsw_ext_set *set, void *gs)
{
    hipGraph_t g = nullptr;
    hipStream_t st = eng->main.stream;
    bool ok = eng->check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "begin capture");
    if (ok) ok = eng->check(ext_enqueue(set, HIPSW_MODE_BATCHED, gs, eng->main.counters, st, nullptr,
                                        nullptr), "capture ext_persistent");
    const hipError_t ee = hipStreamEndCapture(st, &g);
    ok = ok && eng->check(ee, "end capture");
    if (ok) ok = eng->check(hipGraphInstantiate(&set->graph, g, nullptr, nullptr, 0), "instantiate");
    if (g) hipGraphDestroy(g);
    if (!ok) {
        set->graph = nullptr;
        return false;
    }
    set->graph_scratch = gs;
    return true;
}

static void fill_ext_stats(hipsw_stats *stats, const hipsw_ext_set *set, float ms, double t0,
                           int64_t launches, unsigned long long cells)
{
    if (!stats) return;
    memset(stats, 0, sizeof *stats);
    stats->kernel_ms = ms;
    stats->total_ms = now_ms() - t0;
    stats->n_jobs = set->n;
    stats->launches = launches;
    stats->promoted = set->tier_jobs[TIER_I32];
    for (int k = 0; k < 3; ++k) stats->tier_jobs[k] = set->tier_jobs[k];
    stats->cells = (int64_t)cells;
}

extern "C" int hipsw_ext_run(hipsw_engine *eng, hipsw_ext_set *set, int32_t mode,
                             hipsw_stats *stats)
{
    if (!eng || !set) return HIPSW_E_INVALID;
    if (mode != HIPSW_MODE_PLAIN && mode != HIPSW_MODE_BATCHED && mode != HIPSW_MODE_INPUT_ORDER)
        return HIPSW_E_INVALID;
    std::lock_guard<std::recursive_mutex> lock(eng->mu);
    hipSetDevice(eng->device);
    ExecCtx &c = eng->main;
    const double t0 = now_ms();
    int64_t launches = 0;
    float ms = 0.f;
    unsigned long long cells = 0;
    if (set->n > 0) {
        if (mode == HIPSW_MODE_BATCHED && set->launches.size() > 64) {
            eng->set_error("hipsw_ext_run: too many bins");
            return HIPSW_E_UNSUPPORTED;
        }
        if (mode == HIPSW_MODE_INPUT_ORDER) {
            if (!set->io_ok) {
                eng->set_error("hipsw_ext_run: the input-order form needs every job in the 8- or "
                               "16-bit tier, the standard matrix, no N and queries below 256 bases");
                return HIPSW_E_UNSUPPORTED;
            }
            if (!set->io_dmem) {
                const int per_item = set->io_launch.kind == EXT_PACKED ? 128 : 64;
                const size_t b = io_pk_bytes(set->n) + (size_t)set->io_launch.item_count * sizeof(WorkItem);
                if (!eng->check(hipMalloc(&set->io_dmem, b), "hipMalloc(input-order form)"))
                    return HIPSW_E_DEVICE;
                hipLaunchKernelGGL(build_io_kernel, dim3(std::min((set->n + 255) / 256, 4096)), dim3(256),
                                   0, c.stream, set->dev, set->n, per_item,
                                   static_cast<PkMeta *>(set->io_dmem),
                                   reinterpret_cast<WorkItem *>(static_cast<char *>(set->io_dmem) +
                                                                io_pk_bytes(set->n)));
                if (!eng->check(hipGetLastError(), "build input-order form")) return HIPSW_E_DEVICE;
            }
        }
        const size_t need = mode == HIPSW_MODE_PLAIN ? (set->plain_global ? set->plain_scratch : 0)
                            : mode == HIPSW_MODE_BATCHED ? set->batched_scratch
                                                         : set->io_scratch;
        void *gs = nullptr;
        if (need && !(gs = c.scratch.get(need))) {
            eng->set_error("hipsw_ext_run: scratch allocation failed");
            return HIPSW_E_DEVICE;
        }
        if (mode != HIPSW_MODE_BATCHED) {
            if (!eng->check(ext_enqueue(set, mode, gs, c.counters, c.stream, c.ev0, c.ev1),
                            "launch extension"))
                return HIPSW_E_DEVICE;
            launches = 1;
        } else {
            if (set->graph && set->graph_scratch != gs) {
                hipGraphExecDestroy(set->graph);
                set->graph = nullptr;
            }
            if (!set->graph && !build_batched_graph(eng, set, gs)) return HIPSW_E_DEVICE;
            if (!eng->check(hipEventRecord(c.ev0, c.stream), "event") ||
                !eng->check(hipGraphLaunch(set->graph, c.stream), "graph launch") ||
                !eng->check(hipEventRecord(c.ev1, c.stream), "event"))
                return HIPSW_E_DEVICE;
            launches = (int64_t)set->launches.size();
        }
        if (!eng->check(hipEventSynchronize(c.ev1), "kernel")) return HIPSW_E_DEVICE;
        hipEventElapsedTime(&ms, c.ev0, c.ev1);
        if (stats && !eng->check(hipMemcpy(&cells, set->dev.cells, sizeof cells, hipMemcpyDeviceToHost),
                                 "cells"))
            return HIPSW_E_DEVICE;
    }
    fill_ext_stats(stats, set, ms, t0, launches, cells);
    return HIPSW_OK;
}

extern "C" int hipsw_ext_download(hipsw_engine *eng, hipsw_ext_set *set, int32_t *out)
{
    if (!eng || !set || (set->n > 0 && !out)) return HIPSW_E_INVALID;
    std::lock_guard<std::recursive_mutex> lock(eng->mu);
    if (set->n == 0) return HIPSW_OK;
    hipSetDevice(eng->device);
    if (!eng->check(hipMemcpy(out, set->dev.out, (size_t)set->n * 6 * sizeof(int32_t),
                              hipMemcpyDeviceToHost),
                    "download"))
        return HIPSW_E_DEVICE;
    return HIPSW_OK;
}

extern "C" void hipsw_ext_free(hipsw_engine *eng, hipsw_ext_set *set)
{
    if (!set) return;
    if (eng) {
        std::lock_guard<std::recursive_mutex> lock(eng->mu);
        hipSetDevice(eng->device);
        hipStreamSynchronize(eng->main.stream);
        if (set->graph) hipGraphExecDestroy(set->graph);
        if (set->dmem && set->own_dmem) hipFree(set->dmem);
        if (set->io_dmem) hipFree(set->io_dmem);
    }
    delete set;
}

namespace hipsw {

// One-shot run of a prepared set on a leased context: the job arrays go to
// the context's grow-only arena, the bins are launched directly (a graph is
// not worth capturing for a set that runs once), and the results come back
// on the context's stream.  `arena`, `gs_fixed` and `counters` (nullable)
// replace the context's buffers.
int ext_run_once(hipsw_engine *eng, ExecCtx &c, const ExtPrep &P, int mode, void *arena,
                 void *gs_fixed, uint32_t *counters, int32_t *out, hipsw_stats *stats, double t0)
{
    // every failure drains the stream first: copies from the caller's and the
    // prepared host arrays may still be queued when a later step fails
    // (what: the failed step; a runtime failure keeps the runtime's text)
    auto fail = [&](const char *what, bool runtime = false) {
        const std::string rt = runtime ? std::string(" (") + thread_error(eng) + ")" : "";
        (void)hipStreamSynchronize(c.stream);
        eng->set_error(std::string("hipsw_banded_extend: ") + what + rt);
        return HIPSW_E_DEVICE;
    };
    hipsw_ext_set set;
    if (!arena && eng->inject(HIPSW_FAULT_ALLOC)) return fail("arena allocation failed (injected)");
    if (!arena && !(arena = c.arena.get(P.arena_bytes))) return fail("arena allocation failed");
    if (eng->inject(HIPSW_FAULT_COPY)) return fail("upload failed (injected)");
    if (!eng->check(ext_bind_upload(P, arena, &set, c.stream), "upload"))
        return fail("upload failed", true);
    if (mode == HIPSW_MODE_BATCHED && set.launches.size() > 64) {
        (void)hipStreamSynchronize(c.stream);
        eng->set_error("hipsw_banded_extend: too many bins");
        return HIPSW_E_UNSUPPORTED;
    }
    const size_t need = mode == HIPSW_MODE_PLAIN ? (set.plain_global ? set.plain_scratch : 0)
                                                 : set.batched_scratch;
    void *gs = gs_fixed;
    if (need && !gs && eng->inject(HIPSW_FAULT_ALLOC))
        return fail("scratch allocation failed (injected)");
    if (need && !gs && !(gs = c.scratch.get(need))) return fail("scratch allocation failed");
    unsigned long long cells = 0;
    if (eng->inject(HIPSW_FAULT_LAUNCH)) return fail("kernel launch failed (injected)");
    if (!eng->check(ext_enqueue(&set, mode, gs, counters ? counters : c.counters, c.stream, c.ev0,
                                c.ev1), "launch extension"))
        return fail("kernel launch failed", true);
    if (eng->inject(HIPSW_FAULT_COPY)) return fail("download failed (injected)");
    if (!eng->check(hipMemcpyAsync(out, set.dev.out, (size_t)P.n * 6 * sizeof(int32_t),
                                   hipMemcpyDeviceToHost, c.stream), "download") ||
        !eng->check(hipMemcpyAsync(&cells, set.dev.cells, sizeof cells, hipMemcpyDeviceToHost,
                                   c.stream), "download cells") ||
        !eng->check(hipStreamSynchronize(c.stream), "extension"))
        return fail("download failed", true);
    float ms = 0.f;
    hipEventElapsedTime(&ms, c.ev0, c.ev1);
    fill_ext_stats(stats, &set, ms, t0,
                   mode == HIPSW_MODE_PLAIN ? 1 : (int64_t)set.launches.size(), cells);
    return HIPSW_OK;
}

}  // namespace hipsw

extern "C" int hipsw_banded_extend(hipsw_engine *eng, int32_t n, const uint8_t *queries,
                                   const uint32_t *q_off, const uint16_t *q_len,
                                   const uint8_t *targets, const uint32_t *t_off,
                                   const uint16_t *t_len, const int32_t *h0, const int32_t *w,
                                   const int32_t *end_bonus, const hipsw_scoring *sc,
                                   int32_t mode, int32_t *out, hipsw_stats *stats)
{
    if (!eng) return HIPSW_E_INVALID;
    if (n == 0) return HIPSW_OK;
    if (n < 0 || !out || (mode != HIPSW_MODE_PLAIN && mode != HIPSW_MODE_BATCHED))
        return HIPSW_E_INVALID;
    const double t0 = now_ms();
    PrepLease prep(eng);
    ExtPrep &P = prep.get();
    std::string err;
    const int rc = ext_prepare(eng, eng->tune_snapshot(), n, queries, q_off, q_len, targets, t_off,
                               t_len, h0, w, end_bonus, sc, P, err);
    if (rc != HIPSW_OK) {
        eng->set_error("hipsw_banded_extend: " + err);
        return rc;
    }
    SlotLease lease(eng);
    hipSetDevice(eng->device);
    return ext_run_once(eng, lease.ctx(), P, mode, nullptr, nullptr, nullptr, out, stats, t0);
}

// Autotuner: for every 8-bit bin of the set that the register-query kernels
// can run, time the packed form (two jobs per lane, packed 16-bit arithmetic)
// against the one-job-per-lane form (32-bit arithmetic) on the bin's first
// `max_jobs` jobs in batched order, keep the faster in the engine's table for
// that query-length class (later uploads plan with it) and re-plan the set.
extern "C" int hipsw_ext_set_layout(hipsw_engine *eng, int32_t length_class, int32_t kind,
                                   int32_t lds_blocks)
{
    if (!eng || length_class < 0 || length_class > 2) return HIPSW_E_INVALID;
    if (kind == 0) {
        std::lock_guard<std::mutex> lk(eng->tune_mu);
        eng->tune.kind[length_class] = 0;
        eng->tune.lb[length_class] = 0;
        return HIPSW_OK;
    }
    const int cols = length_class == 0 ? 64 : length_class == 1 ? 160 : 256;   // class's widest
    if ((kind != EXT_PACKED && kind != EXT_LANE) || !ext_pk_lb_supported(kind, cols, lds_blocks)) {
        eng->set_error("hipsw_ext_set_layout: layout not built for this length class");
        return HIPSW_E_INVALID;
    }
    std::lock_guard<std::mutex> lk(eng->tune_mu);
    eng->tune.kind[length_class] = kind;
    eng->tune.lb[length_class] = lds_blocks;
    return HIPSW_OK;
}

extern "C" int hipsw_ext_set_layout16(hipsw_engine *eng, int32_t length_class, int32_t lds_blocks)
{
    if (!eng || length_class < 0 || length_class > 2) return HIPSW_E_INVALID;
    const int cols = length_class == 0 ? 64 : length_class == 1 ? 160 : 256;   // class's widest
    if (lds_blocks != 0 && !ext_pk_lb_supported(EXT_LANE, cols, lds_blocks)) {
        eng->set_error("hipsw_ext_set_layout16: layout not built for this length class");
        return HIPSW_E_INVALID;
    }
    std::lock_guard<std::mutex> lk(eng->tune_mu);
    eng->tune.kind16[length_class] = 0;
    eng->tune.lb16[length_class] = lds_blocks;
    return HIPSW_OK;
}

extern "C" int hipsw_ext_set_layout16_kind(hipsw_engine *eng, int32_t length_class, int32_t kind,
                                           int32_t lds_blocks)
{
    if (!eng || length_class < 0 || length_class > 2) return HIPSW_E_INVALID;
    if (kind != 0 && kind != EXT_LANE && kind != EXT_PACKED16) {
        eng->set_error("hipsw_ext_set_layout16_kind: kind must be 0, 2 (one job per lane) or 3 (two jobs per lane)");
        return HIPSW_E_INVALID;
    }
    const int cols = length_class == 0 ? 64 : length_class == 1 ? 160 : 256;   // class's widest
    const int k = kind == 0 ? EXT_LANE : kind;
    if (lds_blocks != 0 && !ext_pk_lb_supported(k, cols, lds_blocks)) {
        eng->set_error("hipsw_ext_set_layout16_kind: layout not built for this length class");
        return HIPSW_E_INVALID;
    }
    std::lock_guard<std::mutex> lk(eng->tune_mu);
    eng->tune.kind16[length_class] = kind == 0 ? 0 : k;
    eng->tune.lb16[length_class] = lds_blocks;
    return HIPSW_OK;
}

extern "C" int hipsw_ext_autotune(hipsw_engine *eng, hipsw_ext_set *set, int32_t max_jobs,
                                  hipsw_tune_entry *report, int32_t report_cap,
                                  int32_t *n_report)
{
    if (!eng || !set || max_jobs < 1) return HIPSW_E_INVALID;
    std::lock_guard<std::recursive_mutex> lock(eng->mu);
    hipSetDevice(eng->device);
    if (n_report) *n_report = 0;
    int nrep = 0;
    WorkItem *d_tmp = nullptr;
    const size_t tmp_cap = (size_t)max_jobs / 64 + 2;
    if (!eng->check(hipMalloc(&d_tmp, tmp_cap * sizeof(WorkItem)), "hipMalloc(tune items)"))
        return HIPSW_E_DEVICE;
    int rc = HIPSW_OK;
    for (ExtLaunch &L : set->launches) {
        if ((L.tier != TIER_U8 && L.tier != TIER_U16) || L.kind == EXT_UNPACKED) continue;
        // candidates, 8-bit tier: the packed form at each LDS share built for
        // this query length, then one job per lane (last: its results stay on
        // the device); 16-bit tier: one job per lane at each LDS share
        struct Cand { int kind, lb; };
        std::vector<Cand> cands;
        int lbs[5];
        if (L.tier == TIER_U8) {
            const int nlb = ext_pk_lb_candidates(EXT_PACKED, L.cols, lbs);
            for (int c = 0; c < nlb; ++c) cands.push_back(Cand{EXT_PACKED, lbs[c]});
            cands.push_back(Cand{EXT_LANE, ext_pk_default_lb(L.cols)});
        } else {
            const int nlb = ext_pk_lb_candidates(EXT_LANE, L.cols, lbs);
            for (int c = 0; c < nlb; ++c) cands.push_back(Cand{EXT_LANE, lbs[c]});
            if (L.w16_ok) {
                const int n16 = ext_pk_lb_candidates(EXT_PACKED16, L.cols, lbs);
                for (int c = 0; c < n16; ++c) cands.push_back(Cand{EXT_PACKED16, lbs[c]});
            }
        }
        ExtLaunch S = L;
        S.job_end = std::min(L.job_end, L.job_begin + max_jobs);
        double best = 1e30;
        Cand win{L.kind, L.lb};
        const int first_rep = nrep;
        for (const Cand &c : cands) {
            ExtLaunch K = S;
            K.kind = c.kind;
            K.lb = c.lb;
            std::vector<WorkItem> items;
            std::vector<ExtLaunch> one{K};
            const size_t need = ext_layout_launches(eng, one, items);
            K = one[0];
            void *gs = nullptr;
            if (need && !(gs = eng->main.scratch.get(need))) {
                eng->set_error("hipsw_ext_autotune: scratch allocation failed");
                rc = HIPSW_E_DEVICE;
                break;
            }
            if (items.size() > tmp_cap ||
                !eng->check(hipMemcpy(d_tmp, items.data(), items.size() * sizeof(WorkItem),
                                      hipMemcpyHostToDevice), "tune items")) {
                rc = HIPSW_E_DEVICE;
                break;
            }
            float t = 1e30f;
            for (int rep = 0; rep < 2; ++rep) {     // the first run warms caches and code
                ExecCtx &c = eng->main;
                if (!eng->check(launch_zero_u32(c.counters, 1, c.stream), "zero") ||
                    !eng->check(hipEventRecord(c.ev0, c.stream), "event") ||
                    !eng->check(launch_ext_bin(K, set->dev, set->params, set->d_order, d_tmp,
                                               c.counters, gs, c.stream), "tune launch") ||
                    !eng->check(hipEventRecord(c.ev1, c.stream), "event") ||
                    !eng->check(hipEventSynchronize(c.ev1), "tune sync")) {
                    rc = HIPSW_E_DEVICE;
                    break;
                }
                float e = 0;
                hipEventElapsedTime(&e, eng->main.ev0, eng->main.ev1);
                t = std::min(t, e);
            }
            if (rc != HIPSW_OK) break;
            if (report && nrep < report_cap)
                report[nrep] = hipsw_tune_entry{L.tier, L.cls, S.job_end - S.job_begin, c.kind, c.lb, 0,
                                                (double)t};
            ++nrep;
            if (t < best) { best = t; win = c; }
        }
        if (rc != HIPSW_OK) break;
        for (int r = first_rep; r < nrep && r < report_cap; ++r)
            if (report && report[r].kind == win.kind && report[r].lds_blocks == win.lb) report[r].chosen = 1;
        L.kind = win.kind;
        L.lb = win.lb;
        std::lock_guard<std::mutex> lk(eng->tune_mu);
        if (L.tier == TIER_U8) {
            eng->tune.kind[L.cls] = win.kind;
            eng->tune.lb[L.cls] = win.lb;
        } else {
            eng->tune.kind16[L.cls] = win.kind;
            eng->tune.lb16[L.cls] = win.lb;
        }
    }
    hipFree(d_tmp);
    if (rc != HIPSW_OK) return rc;
    // re-plan the set's work items, grids and scratch regions with the choices
    if (set->graph) {
        hipGraphExecDestroy(set->graph);
        set->graph = nullptr;
    }
    std::vector<WorkItem> items;
    set->batched_scratch = ext_layout_launches(eng, set->launches, items);
    if (items.size() > set->items_cap) return HIPSW_E_UNSUPPORTED;
    if (!eng->check(hipMemcpy(set->d_items, items.data(), items.size() * sizeof(WorkItem),
                              hipMemcpyHostToDevice), "tune re-plan"))
        return HIPSW_E_DEVICE;
    if (n_report) *n_report = nrep;
    return HIPSW_OK;
}

extern "C" int hipsw_ext_plan(hipsw_engine *eng, const hipsw_ext_set *set,
                              hipsw_plan_entry *entries, int32_t cap, int32_t *n_entries)
{
    if (!eng || !set || cap < 0 || (cap > 0 && !entries)) return HIPSW_E_INVALID;
    std::lock_guard<std::recursive_mutex> lock(eng->mu);
    int k = 0;
    for (const ExtLaunch &L : set->launches) {
        if (k < cap) {
            entries[k] = hipsw_plan_entry{L.tier, L.cls, L.has_n, L.job_end - L.job_begin, L.kind,
                                          L.kind == EXT_UNPACKED ? 0 : L.lb, L.item_count, L.grid};
        }
        ++k;
    }
    if (n_entries) *n_entries = k;
    return HIPSW_OK;
}
