// Chain extension driver (include/hipsw.h, hipsw_extend_chains).
//
// Replaces ExtensionAligner.extend (Sources/Alignment/ExtensionAligner.swift:
// 114-346) for a batch of chains, with the CPU path's exact semantics and the
// banded extensions on the device.  The CPU loop is sequential per chain only
// through the covered-seed skip (:172-189), which reads the regions of
// earlier seeds; a seed's own extension depends on nothing but its chain's
// window (rmax0 / rmax1, :134-144) and its own left result.  So every seed is
// extended speculatively, in four device passes over all chains -- left try 0,
// left try 1 for the seeds the band-retry rule sends back (:229-246), right
// try 0 with h0 = the left score, right try 1 -- and the host then walks each
// chain in seed-score order, keeps the seeds the skip rule keeps and drops the
// rest.  Results equal the sequential loop's; the device spends extra cells
// on skipped seeds.
#include "engine_internal.h"

#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

using namespace hipsw;

namespace {

// calMaxGap (:352-363)
int cal_max_gap(int qlen, const hipsw_scoring *sc, const hipsw_ext_opts *o)
{
    const int a = o->match_score;
    const int ld = sc->e_del > 0 ? (qlen * a - sc->o_del) / sc->e_del + 1 : 1;
    const int li = sc->e_ins > 0 ? (qlen * a - sc->o_ins) / sc->e_ins + 1 : 1;
    return std::min(std::max(std::max(ld, li), 1), o->band_width << 1);
}

// One side's banded-extension jobs, one byte per base.
struct Side {
    std::vector<uint8_t> q, t;
    std::vector<uint32_t> q_off, t_off;
    std::vector<uint16_t> q_len, t_len;
    std::vector<int32_t> h0, w, eb;
    std::vector<uint32_t> seed;          // global seed index of each job
    std::vector<int32_t> out;            // 6 per job, SWResult order
};

int run_side(hipsw_engine *eng, const Side &s, const hipsw_scoring *sc, std::vector<int32_t> &out,
             hipsw_stats *acc)
{
    const int n = (int)s.seed.size();
    out.assign((size_t)n * 6, 0);
    if (n == 0) return HIPSW_OK;
    const uint8_t one = 0;
    hipsw_stats st;
    std::memset(&st, 0, sizeof st);
    const int rc = hipsw_banded_extend(eng, n, s.q.empty() ? &one : s.q.data(), s.q_off.data(),
                                       s.q_len.data(), s.t.empty() ? &one : s.t.data(),
                                       s.t_off.data(), s.t_len.data(), s.h0.data(), s.w.data(),
                                       s.eb.data(), sc, HIPSW_MODE_BATCHED, out.data(), &st);
    if (rc == HIPSW_OK && acc) {
        acc->kernel_ms += st.kernel_ms;
        acc->launches += st.launches;
        acc->cells += st.cells;
        acc->n_jobs += st.n_jobs;
        for (int k = 0; k < 3; ++k) acc->tier_jobs[k] += st.tier_jobs[k];
    }
    return rc;
}

// The retry subset of a side: jobs whose try 0 neither kept the previous
// score nor stayed inside three quarters of the band (:238-242); try 1 runs
// them again at twice the band.
Side retry_side(const Side &s, const std::vector<int32_t> &r0, const std::vector<int32_t> &prev,
                std::vector<int32_t> &slot_of_job)
{
    Side t;
    slot_of_job.assign(s.seed.size(), -1);
    for (size_t k = 0; k < s.seed.size(); ++k) {
        const int32_t score = r0[k * 6], max_off = r0[k * 6 + 5], w = s.w[k];
        if (score == prev[k] || max_off < (w >> 1) + (w >> 2)) continue;
        slot_of_job[k] = (int32_t)t.seed.size();
        t.seed.push_back(s.seed[k]);
        t.q_off.push_back(s.q_off[k]); t.q_len.push_back(s.q_len[k]);
        t.t_off.push_back(s.t_off[k]); t.t_len.push_back(s.t_len[k]);
        t.h0.push_back(s.h0[k]); t.w.push_back(w << 1); t.eb.push_back(s.eb[k]);
    }
    return t;
}

}  // namespace

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
        !get_ref || !sc || !opt || !out || !out_n || !scoring_valid(sc) || opt->band_width < 0)
        return HIPSW_E_INVALID;
    const uint32_t n_seeds = chain_seed_off[n_chains];
    if (n_seeds && (!seed_rbeg || !seed_qbeg || !seed_len || !seed_score)) return HIPSW_E_INVALID;
    const double t0 = now_ms();
    hipsw_stats acc;
    std::memset(&acc, 0, sizeof acc);
    const int32_t band = opt->band_width;
    constexpr int kMaxWindow = 10000;   // :214, :277

    // Chain windows (:134-144) and the left jobs of every seed (:206-227).
    std::vector<int64_t> rmax0(n_chains), rmax1(n_chains);
    std::vector<uint8_t> tbuf(kMaxWindow);
    Side L, R;
    std::vector<int32_t> left_job(n_seeds, -1), right_job(n_seeds, -1);
    for (int c = 0; c < n_chains; ++c) {
        const uint32_t s0 = chain_seed_off[c], s1 = chain_seed_off[c + 1];
        if (s1 < s0 || s1 > n_seeds) return HIPSW_E_INVALID;
        const int rl = read_len[chain_read[c]];
        const uint8_t *rd = reads + read_off[chain_read[c]];
        int64_t lo = INT64_MAX, hi = 0;
        for (uint32_t s = s0; s < s1; ++s) {
            const int64_t b = seed_rbeg[s] - (int64_t)(seed_qbeg[s] + cal_max_gap(seed_qbeg[s], sc, opt));
            const int rem = rl - seed_qbeg[s] - seed_len[s];
            const int64_t e = seed_rbeg[s] + seed_len[s] + (int64_t)(rem + cal_max_gap(rem, sc, opt));
            lo = std::min(lo, b);
            hi = std::max(hi, e);
        }
        rmax0[c] = std::max<int64_t>(lo, 0);
        rmax1[c] = hi;
        for (uint32_t s = s0; s < s1; ++s) {
            const int32_t qbeg = seed_qbeg[s];
            if (qbeg <= 0) continue;
            const int64_t tl = std::min<int64_t>(seed_rbeg[s] - rmax0[c], kMaxWindow);
            if (tl <= 0) continue;
            const int32_t got = get_ref(ref_ctx, seed_rbeg[s] - tl, (int32_t)tl, tbuf.data());
            if (got < 0 || got > tl) return HIPSW_E_INVALID;
            left_job[s] = (int32_t)L.seed.size();
            L.seed.push_back(s);
            L.q_off.push_back((uint32_t)L.q.size());
            L.q_len.push_back((uint16_t)qbeg);
            for (int i = qbeg - 1; i >= 0; --i) L.q.push_back(rd[i]);           // reversed query
            L.t_off.push_back((uint32_t)L.t.size());
            L.t_len.push_back((uint16_t)got);
            for (int i = got - 1; i >= 0; --i) L.t.push_back(tbuf[i]);          // reversed window
            L.h0.push_back(seed_score[s]);
            L.w.push_back(band);
            L.eb.push_back(opt->pen_clip5);
        }
    }

    if (L.t.size() > 0xFFFFFFF0ull || L.q.size() > 0xFFFFFFF0ull) return HIPSW_E_UNSUPPORTED;

    // Left tries 0 and 1.
    std::vector<int32_t> l0, l1, lslot;
    int rc = run_side(eng, L, sc, l0, &acc);
    if (rc != HIPSW_OK) return rc;
    std::vector<int32_t> lprev(L.seed.size());
    for (size_t k = 0; k < L.seed.size(); ++k) lprev[k] = L.h0[k];
    const Side L2 = retry_side(L, l0, lprev, lslot);
    rc = run_side(eng, L2, sc, l1, &acc);
    if (rc != HIPSW_OK) return rc;
    // final left result and band of seed s
    auto left_res = [&](uint32_t s, int32_t &w) -> const int32_t * {
        const int j = left_job[s];
        if (lslot[j] >= 0) { w = band << 1; return &l1[(size_t)lslot[j] * 6]; }
        w = band;
        return &l0[(size_t)j * 6];
    };

    // Right jobs, h0 = the left score (or the seed score without a left side).
    std::vector<int32_t> acc_h0(n_seeds);
    for (int c = 0; c < n_chains; ++c) {
        const int rl = read_len[chain_read[c]];
        const uint8_t *rd = reads + read_off[chain_read[c]];
        for (uint32_t s = chain_seed_off[c]; s < chain_seed_off[c + 1]; ++s) {
            int32_t wl;
            acc_h0[s] = left_job[s] >= 0 ? left_res(s, wl)[0] : seed_score[s];
            const int32_t qend = seed_qbeg[s] + seed_len[s];
            if (qend >= rl) continue;
            const int64_t rend = seed_rbeg[s] + seed_len[s];
            const int64_t tl = std::min<int64_t>(rmax1[c] - rend, kMaxWindow);
            if (tl <= 0) continue;
            const int32_t got = get_ref(ref_ctx, rend, (int32_t)tl, tbuf.data());
            if (got < 0 || got > tl) return HIPSW_E_INVALID;
            right_job[s] = (int32_t)R.seed.size();
            R.seed.push_back(s);
            R.q_off.push_back((uint32_t)R.q.size());
            R.q_len.push_back((uint16_t)(rl - qend));
            R.q.insert(R.q.end(), rd + qend, rd + rl);
            R.t_off.push_back((uint32_t)R.t.size());
            R.t_len.push_back((uint16_t)got);
            R.t.insert(R.t.end(), tbuf.data(), tbuf.data() + got);
            R.h0.push_back(acc_h0[s]);
            R.w.push_back(band);
            R.eb.push_back(opt->pen_clip3);
        }
    }
    if (R.t.size() > 0xFFFFFFF0ull || R.q.size() > 0xFFFFFFF0ull) return HIPSW_E_UNSUPPORTED;
    std::vector<int32_t> r0, r1, rslot;
    rc = run_side(eng, R, sc, r0, &acc);
    if (rc != HIPSW_OK) return rc;
    const Side R2 = retry_side(R, r0, R.h0, rslot);
    rc = run_side(eng, R2, sc, r1, &acc);
    if (rc != HIPSW_OK) return rc;

    // Sequential pass per chain (:155-343): seed order, covered skip, regions.
    const int32_t a = opt->match_score;
    const int32_t tmp = std::max(a + opt->mismatch_penalty,
                                 std::max(sc->o_ins + sc->e_ins, sc->o_del + sc->e_del));
    parallel_for(n_chains, host_threads(), [&](int64_t lo, int64_t hi) {
        std::vector<uint32_t> order;
        for (int64_t c = lo; c < hi; ++c) {
            const uint32_t s0 = chain_seed_off[c], s1 = chain_seed_off[c + 1];
            const int32_t rl = read_len[chain_read[c]];
            order.resize(s1 - s0);
            std::iota(order.begin(), order.end(), s0);
            std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
                return seed_score[x] > seed_score[y];
            });
            hipsw_region *regs = out + s0;
            int nreg = 0;
            for (uint32_t s : order) {
                const int32_t qbeg = seed_qbeg[s], len = seed_len[s];
                const int32_t qend = qbeg + len;
                const int64_t rbeg = seed_rbeg[s], rend = rbeg + len;
                int cover = -1;
                for (int r = 0; r < nreg; ++r)
                    if (qbeg >= regs[r].qb && qend <= regs[r].qe) { cover = r; break; }
                if (cover >= 0) {
                    const int32_t sa = len * a;
                    regs[cover].sub = std::max(regs[cover].sub, sa);
                    if (sa >= regs[cover].score - tmp) regs[cover].sub_n += 1;
                    continue;
                }
                hipsw_region g;
                std::memset(&g, 0, sizeof g);
                g.w = band;
                g.seed_len0 = len;
                int32_t ls = 0, lq = 0, lt = 0;
                if (left_job[s] >= 0) {                                      // :248-260
                    int32_t w;
                    const int32_t *r = left_res(s, w);
                    g.w = std::max(g.w, w);
                    if (r[4] <= 0 || r[4] <= r[0] - opt->pen_clip5) { lq = r[1]; lt = r[2]; ls = r[0]; }
                    else { lq = qbeg; lt = r[3]; ls = r[4]; }
                }
                int32_t rs = 0, rq = 0, rt = 0;
                if (right_job[s] >= 0) {                                     // :294-316
                    const int j = right_job[s];
                    const int32_t *r = rslot[j] >= 0 ? &r1[(size_t)rslot[j] * 6] : &r0[(size_t)j * 6];
                    g.w = std::max(g.w, rslot[j] >= 0 ? band << 1 : band);
                    if (r[4] <= 0 || r[4] <= r[0] - opt->pen_clip3) { rq = r[1]; rt = r[2]; rs = r[0]; }
                    else { rq = rl - qend; rt = r[3]; rs = r[4]; }
                }
                g.qb = qbeg - lq;                                            // :318-331
                g.qe = qend + rq;
                g.rb = rbeg - lt;
                g.re = rend + rt;
                int32_t ts = len * a;
                if (ls > 0) ts = ls;
                if (rs > 0) ts += rs - acc_h0[s];
                g.score = ts;
                g.true_score = ts;
                int32_t cov = 0;                                             // computeSeedCov, :36-70
                for (uint32_t t = s0; t < s1; ++t)
                    if (seed_qbeg[t] >= g.qb && seed_qbeg[t] + seed_len[t] <= g.qe &&
                        seed_rbeg[t] >= g.rb && seed_rbeg[t] + seed_len[t] <= g.re)
                        cov += seed_len[t];
                g.seed_cov = cov;
                regs[nreg++] = g;
            }
            const int32_t sub_min = opt->min_seed_len * a;                   // :338-343
            for (int r = 0; r < nreg; ++r)
                if (regs[r].sub > 0 && regs[r].sub < sub_min) regs[r].sub = 0;
            out_n[c] = nreg;
        }
    });
    if (stats) {
        *stats = acc;
        stats->total_ms = now_ms() - t0;
    }
    return HIPSW_OK;
}
