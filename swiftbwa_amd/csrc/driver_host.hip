// Chain extension driver (include/hipsw.h, hipsw_chain_* and
// hipsw_extend_chains).
//
// Replaces ExtensionAligner.extend (Sources/Alignment/ExtensionAligner.swift:
// 114-346) for a batch of chains, with the CPU path's exact semantics and the
// banded extensions on the device.  The CPU loop is sequential per chain only
// through the covered-seed skip (:172-189), which reads the regions of
// earlier seeds; a seed's own extension depends on nothing but its chain's
// window (rmax0 / rmax1, :134-144) and its own left result.  So every seed is
// extended speculatively, in four passes over all chains -- left try 0,
// left try 1 for the seeds the band-retry rule sends back (:229-246), right
// try 0 with h0 = the left score, right try 1 -- and the host then walks each
// chain in seed-score order, keeps the seeds the skip rule keeps and drops the
// rest.  Results equal the sequential loop's; the device spends extra cells
// on skipped seeds.
//
// The passes are a small state machine (hipsw_chain_driver): each pass hands
// out its jobs with their reference windows as (position, length, reverse)
// specs, and takes the six result fields back.  hipsw_extend_chains runs the
// four passes through hipsw_banded_extend with windows from a getReference
// callback; the serving loop runs them as chunks whose windows the device
// fetches from the resident .pac.
#include "engine_internal.h"

#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

using namespace hipsw;

namespace {

constexpr int kMaxWindow = 10000;   // :214, :277

// calMaxGap (:352-363)
int cal_max_gap(int qlen, const hipsw_scoring *sc, const hipsw_ext_opts *o)
{
    const int a = o->match_score;
    const int ld = sc->e_del > 0 ? (qlen * a - sc->o_del) / sc->e_del + 1 : 1;
    const int li = sc->e_ins > 0 ? (qlen * a - sc->o_ins) / sc->e_ins + 1 : 1;
    return std::min(std::max(std::max(ld, li), 1), o->band_width << 1);
}

// One pass's banded-extension jobs: queries one byte per base, targets as
// getReference windows.
struct Side {
    std::vector<uint8_t> q;
    std::vector<uint32_t> q_off;
    std::vector<uint16_t> q_len;
    std::vector<int64_t> t_pos;
    std::vector<int32_t> t_len;          // requested; the window may return fewer
    std::vector<uint8_t> t_rev;
    std::vector<int32_t> h0, w, eb;
    std::vector<uint32_t> seed;          // global seed index of each job
    void push(const Side &s, size_t k, int32_t new_w)
    {
        seed.push_back(s.seed[k]);
        q_off.push_back(s.q_off[k]); q_len.push_back(s.q_len[k]);
        t_pos.push_back(s.t_pos[k]); t_len.push_back(s.t_len[k]); t_rev.push_back(s.t_rev[k]);
        h0.push_back(s.h0[k]); w.push_back(new_w); eb.push_back(s.eb[k]);
    }
};

// The retry subset of a side: jobs whose try 0 neither kept the previous
// score nor stayed inside three quarters of the band (:238-242); try 1 runs
// them again at twice the band.  Queries stay in the try-0 arena.
Side retry_side(const Side &s, const std::vector<int32_t> &r0, std::vector<int32_t> &slot_of_job)
{
    Side t;
    slot_of_job.assign(s.seed.size(), -1);
    for (size_t k = 0; k < s.seed.size(); ++k) {
        const int32_t score = r0[k * 6], max_off = r0[k * 6 + 5], w = s.w[k];
        if (score == s.h0[k] || max_off < (w >> 1) + (w >> 2)) continue;
        slot_of_job[k] = (int32_t)t.seed.size();
        t.push(s, k, w << 1);
    }
    return t;
}

}  // namespace

struct hipsw_chain_driver {
    // inputs (owned copies)
    int32_t n_chains = 0;
    std::vector<uint32_t> seed_off, chain_read, read_off;
    std::vector<int64_t> rbeg;
    std::vector<int32_t> qbeg, len, score, read_len;
    std::vector<uint8_t> reads;
    hipsw_scoring sc;
    hipsw_ext_opts opt;
    // state
    int next_pass = 0;                   // the pass next() hands out
    bool fed = true;                     // the handed-out pass has its results
    std::vector<int64_t> rmax0, rmax1;
    Side L, L2, R, R2;
    std::vector<int32_t> l0, l1, lslot, r0, r1, rslot, left_job, right_job, acc_h0;

    const Side &side(int p) const { return p == 0 ? L : p == 1 ? L2 : p == 2 ? R : R2; }
    std::vector<int32_t> &result(int p) { return p == 0 ? l0 : p == 1 ? l1 : p == 2 ? r0 : r1; }

    // final left result and band of seed s
    const int32_t *left_res(uint32_t s, int32_t &w) const
    {
        const int j = left_job[s];
        if (lslot[j] >= 0) { w = opt.band_width << 1; return &l1[(size_t)lslot[j] * 6]; }
        w = opt.band_width;
        return &l0[(size_t)j * 6];
    }

    // Chain windows (:134-144) and the left jobs of every seed (:206-227).
    void build_left()
    {
        const uint32_t n_seeds = seed_off[n_chains];
        rmax0.assign(n_chains, 0);
        rmax1.assign(n_chains, 0);
        left_job.assign(n_seeds, -1);
        right_job.assign(n_seeds, -1);
        for (int c = 0; c < n_chains; ++c) {
            const uint32_t s0 = seed_off[c], s1 = seed_off[c + 1];
            const int rl = read_len[chain_read[c]];
            const uint8_t *rd = reads.data() + read_off[chain_read[c]];
            int64_t lo = INT64_MAX, hi = 0;
            for (uint32_t s = s0; s < s1; ++s) {
                const int64_t b = rbeg[s] - (int64_t)(qbeg[s] + cal_max_gap(qbeg[s], &sc, &opt));
                const int rem = rl - qbeg[s] - len[s];
                const int64_t e = rbeg[s] + len[s] + (int64_t)(rem + cal_max_gap(rem, &sc, &opt));
                lo = std::min(lo, b);
                hi = std::max(hi, e);
            }
            rmax0[c] = std::max<int64_t>(lo, 0);
            rmax1[c] = hi;
            for (uint32_t s = s0; s < s1; ++s) {
                const int32_t qb = qbeg[s];
                if (qb <= 0) continue;
                const int64_t tl = std::min<int64_t>(rbeg[s] - rmax0[c], kMaxWindow);
                if (tl <= 0) continue;
                left_job[s] = (int32_t)L.seed.size();
                L.seed.push_back(s);
                L.q_off.push_back((uint32_t)L.q.size());
                L.q_len.push_back((uint16_t)qb);
                for (int i = qb - 1; i >= 0; --i) L.q.push_back(rd[i]);       // reversed query
                L.t_pos.push_back(rbeg[s] - tl);
                L.t_len.push_back((int32_t)tl);
                L.t_rev.push_back(1);                                         // reversed window
                L.h0.push_back(score[s]);
                L.w.push_back(opt.band_width);
                L.eb.push_back(opt.pen_clip5);
            }
        }
    }

    // Right jobs, h0 = the left score (or the seed score without a left side).
    void build_right()
    {
        acc_h0.assign(seed_off[n_chains], 0);
        for (int c = 0; c < n_chains; ++c) {
            const int rl = read_len[chain_read[c]];
            const uint8_t *rd = reads.data() + read_off[chain_read[c]];
            for (uint32_t s = seed_off[c]; s < seed_off[c + 1]; ++s) {
                int32_t wl;
                acc_h0[s] = left_job[s] >= 0 ? left_res(s, wl)[0] : score[s];
                const int32_t qend = qbeg[s] + len[s];
                if (qend >= rl) continue;
                const int64_t rend = rbeg[s] + len[s];
                const int64_t tl = std::min<int64_t>(rmax1[c] - rend, kMaxWindow);
                if (tl <= 0) continue;
                right_job[s] = (int32_t)R.seed.size();
                R.seed.push_back(s);
                R.q_off.push_back((uint32_t)R.q.size());
                R.q_len.push_back((uint16_t)(rl - qend));
                R.q.insert(R.q.end(), rd + qend, rd + rl);
                R.t_pos.push_back(rend);
                R.t_len.push_back((int32_t)tl);
                R.t_rev.push_back(0);
                R.h0.push_back(acc_h0[s]);
                R.w.push_back(opt.band_width);
                R.eb.push_back(opt.pen_clip3);
            }
        }
    }

    // Sequential pass per chain (:155-343): seed order, covered skip, regions.
    void walk(hipsw_region *out, int32_t *out_n) const
    {
        const int32_t a = opt.match_score, band = opt.band_width;
        const int32_t tmp = std::max(a + opt.mismatch_penalty,
                                     std::max(sc.o_ins + sc.e_ins, sc.o_del + sc.e_del));
        parallel_for(n_chains, host_threads(), [&](int64_t lo, int64_t hi) {
            std::vector<uint32_t> order;
            for (int64_t c = lo; c < hi; ++c) {
                const uint32_t s0 = seed_off[c], s1 = seed_off[c + 1];
                const int32_t rl = read_len[chain_read[c]];
                order.resize(s1 - s0);
                std::iota(order.begin(), order.end(), s0);
                std::stable_sort(order.begin(), order.end(),
                                 [&](uint32_t x, uint32_t y) { return score[x] > score[y]; });
                hipsw_region *regs = out + s0;
                int nreg = 0;
                for (uint32_t s : order) {
                    const int32_t qb = qbeg[s], sl = len[s];
                    const int32_t qend = qb + sl;
                    const int64_t rb = rbeg[s], rend = rb + sl;
                    int cover = -1;
                    for (int r = 0; r < nreg; ++r)
                        if (qb >= regs[r].qb && qend <= regs[r].qe) { cover = r; break; }
                    if (cover >= 0) {
                        const int32_t sa = sl * a;
                        regs[cover].sub = std::max(regs[cover].sub, sa);
                        if (sa >= regs[cover].score - tmp) regs[cover].sub_n += 1;
                        continue;
                    }
                    hipsw_region g;
                    std::memset(&g, 0, sizeof g);
                    g.w = band;
                    g.seed_len0 = sl;
                    int32_t ls = 0, lq = 0, lt = 0;
                    if (left_job[s] >= 0) {                                      // :248-260
                        int32_t w;
                        const int32_t *r = left_res(s, w);
                        g.w = std::max(g.w, w);
                        if (r[4] <= 0 || r[4] <= r[0] - opt.pen_clip5) { lq = r[1]; lt = r[2]; ls = r[0]; }
                        else { lq = qb; lt = r[3]; ls = r[4]; }
                    }
                    int32_t rs = 0, rq = 0, rt = 0;
                    if (right_job[s] >= 0) {                                     // :294-316
                        const int j = right_job[s];
                        const int32_t *r = rslot[j] >= 0 ? &r1[(size_t)rslot[j] * 6] : &r0[(size_t)j * 6];
                        g.w = std::max(g.w, rslot[j] >= 0 ? band << 1 : band);
                        if (r[4] <= 0 || r[4] <= r[0] - opt.pen_clip3) { rq = r[1]; rt = r[2]; rs = r[0]; }
                        else { rq = rl - qend; rt = r[3]; rs = r[4]; }
                    }
                    g.qb = qb - lq;                                              // :318-331
                    g.qe = qend + rq;
                    g.rb = rb - lt;
                    g.re = rend + rt;
                    int32_t ts = sl * a;
                    if (ls > 0) ts = ls;
                    if (rs > 0) ts += rs - acc_h0[s];
                    g.score = ts;
                    g.true_score = ts;
                    int32_t cov = 0;                                             // computeSeedCov, :36-70
                    for (uint32_t t = s0; t < s1; ++t)
                        if (qbeg[t] >= g.qb && qbeg[t] + len[t] <= g.qe && rbeg[t] >= g.rb &&
                            rbeg[t] + len[t] <= g.re)
                            cov += len[t];
                    g.seed_cov = cov;
                    regs[nreg++] = g;
                }
                const int32_t sub_min = opt.min_seed_len * a;                    // :338-343
                for (int r = 0; r < nreg; ++r)
                    if (regs[r].sub > 0 && regs[r].sub < sub_min) regs[r].sub = 0;
                out_n[c] = nreg;
            }
        });
    }
};

extern "C" hipsw_chain_driver *hipsw_chain_begin(int32_t n_chains, const uint32_t *chain_seed_off,
                                                 const int64_t *seed_rbeg, const int32_t *seed_qbeg,
                                                 const int32_t *seed_len, const int32_t *seed_score,
                                                 const uint32_t *chain_read, int32_t n_reads,
                                                 const uint8_t *reads, const uint32_t *read_off,
                                                 const int32_t *read_len, const hipsw_scoring *sc,
                                                 const hipsw_ext_opts *opt)
{
    if (n_chains < 0 || n_reads < 0 || !sc || !opt || !scoring_valid(sc) || opt->band_width < 0)
        return nullptr;
    if (n_chains > 0 && (!chain_seed_off || !chain_read || !reads || !read_off || !read_len))
        return nullptr;
    const uint32_t n_seeds = n_chains ? chain_seed_off[n_chains] : 0;
    if (n_seeds && (!seed_rbeg || !seed_qbeg || !seed_len || !seed_score)) return nullptr;
    auto d = std::make_unique<hipsw_chain_driver>();
    d->n_chains = n_chains;
    d->sc = *sc;
    d->opt = *opt;
    if (n_chains) {
        d->seed_off.assign(chain_seed_off, chain_seed_off + n_chains + 1);
        for (int c = 0; c < n_chains; ++c) {
            if (d->seed_off[c + 1] < d->seed_off[c]) return nullptr;
            if (chain_read[c] >= (uint32_t)n_reads) return nullptr;
        }
        d->chain_read.assign(chain_read, chain_read + n_chains);
        d->read_off.assign(read_off, read_off + n_reads);
        d->read_len.assign(read_len, read_len + n_reads);
        size_t rbytes = 0;
        for (int r = 0; r < n_reads; ++r) {
            if (read_len[r] < 0 || read_len[r] > 65535) return nullptr;
            rbytes = std::max(rbytes, (size_t)read_off[r] + (size_t)read_len[r]);
        }
        d->reads.assign(reads, reads + rbytes);
        d->rbeg.assign(seed_rbeg, seed_rbeg + n_seeds);
        d->qbeg.assign(seed_qbeg, seed_qbeg + n_seeds);
        d->len.assign(seed_len, seed_len + n_seeds);
        d->score.assign(seed_score, seed_score + n_seeds);
    } else {
        d->seed_off.assign(1, 0);
    }
    d->build_left();
    if (d->L.q.size() > 0xFFFFFFF0ull) return nullptr;
    return d.release();
}

extern "C" int hipsw_chain_next(hipsw_chain_driver *d, hipsw_chain_pass *pass)
{
    if (!d || !pass || !d->fed || d->next_pass >= 4) return HIPSW_E_INVALID;
    const int p = d->next_pass;
    if (p == 2) {
        d->build_right();
        if (d->R.q.size() > 0xFFFFFFF0ull) return HIPSW_E_UNSUPPORTED;
    }
    const Side &s = d->side(p);
    static const uint8_t one = 0;
    std::memset(pass, 0, sizeof *pass);
    pass->index = p;
    pass->n = (int32_t)s.seed.size();
    pass->q = s.q.empty() ? &one : s.q.data();
    pass->q_off = s.q_off.data();
    pass->q_len = s.q_len.data();
    pass->t_pos = s.t_pos.data();
    pass->t_len = s.t_len.data();
    pass->t_rev = s.t_rev.data();
    pass->h0 = s.h0.data();
    pass->w = s.w.data();
    pass->eb = s.eb.data();
    d->fed = false;
    return HIPSW_OK;
}

extern "C" int hipsw_chain_feed(hipsw_chain_driver *d, const int32_t *out6)
{
    if (!d || d->fed || d->next_pass >= 4) return HIPSW_E_INVALID;
    const int p = d->next_pass;
    const size_t n = d->side(p).seed.size();
    if (n && !out6) return HIPSW_E_INVALID;
    d->result(p).assign(out6, out6 + n * 6);
    if (p == 0) d->L2 = retry_side(d->L, d->l0, d->lslot);
    if (p == 2) d->R2 = retry_side(d->R, d->r0, d->rslot);
    d->fed = true;
    d->next_pass++;
    return HIPSW_OK;
}

extern "C" int hipsw_chain_finish(hipsw_chain_driver *d, hipsw_region *out, int32_t *out_n)
{
    if (!d || d->next_pass != 4 || !d->fed) return HIPSW_E_INVALID;
    if (d->n_chains == 0) return HIPSW_OK;
    if (!out || !out_n) return HIPSW_E_INVALID;
    d->walk(out, out_n);
    return HIPSW_OK;
}

extern "C" void hipsw_chain_free(hipsw_chain_driver *d) { delete d; }

extern "C" int32_t hipsw_get_reference_pac(void *ctx, int64_t pos, int32_t len, uint8_t *buf)
{
    // FMIndex.getReference (Sources/FMIndex/FMIndex.swift:25-49) on a .pac
    // (2-bit, first base of a byte in bits 7-6, PackedReference.swift:34-38).
    const hipsw_pac_ref *r = static_cast<const hipsw_pac_ref *>(ctx);
    const int64_t gl = r->genome_len;
    auto base = [&](int64_t i) { return (uint8_t)((r->pac[i >> 2] >> ((3 - (i & 3)) * 2)) & 3); };
    if (pos >= gl) {
        const int64_t fend = 2 * gl - pos, fstart = fend - len;
        const int64_t s = std::max<int64_t>(0, fstart), e = std::min(fend, gl);
        if (e <= s) return 0;
        const int32_t n = (int32_t)(e - s);
        for (int32_t k = 0; k < n; ++k) buf[k] = (uint8_t)(3 - base(e - 1 - k));
        return n;
    }
    const int64_t n = std::min<int64_t>(len, gl - pos);
    if (n <= 0 || pos < 0) return 0;
    for (int64_t k = 0; k < n; ++k) buf[k] = base(pos + k);
    return (int32_t)n;
}

extern "C" int hipsw_extend_chains(hipsw_engine *eng, int32_t n_chains,
                                   const uint32_t *chain_seed_off, const int64_t *seed_rbeg,
                                   const int32_t *seed_qbeg, const int32_t *seed_len,
                                   const int32_t *seed_score, const uint32_t *chain_read,
                                   const uint8_t *reads, const uint32_t *read_off,
                                   const int32_t *read_len, hipsw_get_reference_fn get_ref,
                                   void *ref_ctx, const hipsw_scoring *sc,
                                   const hipsw_ext_opts *opt, hipsw_region *out, int32_t *out_n,
                                   hipsw_stats *stats)
{
    if (!eng) return HIPSW_E_INVALID;
    if (n_chains == 0) return HIPSW_OK;
    if (n_chains < 0 || !chain_seed_off || !chain_read || !reads || !read_off || !read_len ||
        !get_ref || !out || !out_n)
        return HIPSW_E_INVALID;
    const double t0 = now_ms();
    int32_t n_reads = 0;
    for (int c = 0; c < n_chains; ++c) n_reads = std::max(n_reads, (int32_t)chain_read[c] + 1);
    std::unique_ptr<hipsw_chain_driver> d(
        hipsw_chain_begin(n_chains, chain_seed_off, seed_rbeg, seed_qbeg, seed_len, seed_score,
                          chain_read, n_reads, reads, read_off, read_len, sc, opt));
    if (!d) return HIPSW_E_INVALID;
    hipsw_stats acc;
    std::memset(&acc, 0, sizeof acc);
    std::vector<uint8_t> tbuf(kMaxWindow), t;
    std::vector<uint32_t> t_off;
    std::vector<uint16_t> t_len;
    std::vector<int32_t> res;
    for (int p = 0; p < 4; ++p) {
        hipsw_chain_pass ps;
        int rc = hipsw_chain_next(d.get(), &ps);
        if (rc != HIPSW_OK) return rc;
        res.assign((size_t)ps.n * 6, 0);
        if (ps.n) {
            // windows through the caller's getReference, reversed for left jobs
            t.clear();
            t_off.assign(ps.n, 0);
            t_len.assign(ps.n, 0);
            for (int k = 0; k < ps.n; ++k) {
                const int32_t got = get_ref(ref_ctx, ps.t_pos[k], ps.t_len[k], tbuf.data());
                if (got < 0 || got > ps.t_len[k]) return HIPSW_E_INVALID;
                t_off[k] = (uint32_t)t.size();
                t_len[k] = (uint16_t)got;
                if (ps.t_rev[k]) t.insert(t.end(), tbuf.rend() - got, tbuf.rend());
                else t.insert(t.end(), tbuf.begin(), tbuf.begin() + got);
            }
            if (t.size() > 0xFFFFFFF0ull) return HIPSW_E_UNSUPPORTED;
            const uint8_t one = 0;
            hipsw_stats st;
            std::memset(&st, 0, sizeof st);
            rc = hipsw_banded_extend(eng, ps.n, ps.q, ps.q_off, ps.q_len, t.empty() ? &one : t.data(),
                                     t_off.data(), t_len.data(), ps.h0, ps.w, ps.eb, sc,
                                     HIPSW_MODE_BATCHED, res.data(), &st);
            if (rc != HIPSW_OK) return rc;
            acc.kernel_ms += st.kernel_ms;
            acc.launches += st.launches;
            acc.cells += st.cells;
            acc.n_jobs += st.n_jobs;
            for (int k = 0; k < 3; ++k) acc.tier_jobs[k] += st.tier_jobs[k];
        }
        rc = hipsw_chain_feed(d.get(), res.data());
        if (rc != HIPSW_OK) return rc;
    }
    const int rc = hipsw_chain_finish(d.get(), out, out_n);
    if (rc != HIPSW_OK) return rc;
    if (stats) {
        *stats = acc;
        stats->total_ms = now_ms() - t0;
    }
    return HIPSW_OK;
}
