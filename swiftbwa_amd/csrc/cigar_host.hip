// CIGAR generation for alignment regions (include/hipsw.h).
//
// Replaces CIGARGenerator.generate (Sources/Alignment/CIGARGenerator.swift:
// 30-111) for a batch of regions.  Per region, on the host: the band
// (inferBandWidth, :147-170, then max with initialW, :44-50) and the
// equal-length fast path (:63-80).  Every other region runs banded NW plus
// traceback with the doubling retry loop (:82-96) on the device through the
// NW pipeline of hipsw_global_nw.  The host then applies postProcess (:117-142):
// squeeze leading / trailing deletions (:272-295), NM (:173-210), MD
// (:219-268) and soft clips (:298-320).
#include "engine_internal.h"

#include <algorithm>
#include <functional>
#include <cstring>
#include <string>
#include <vector>

using namespace hipsw;

namespace {

constexpr uint32_t kOpM = 0, kOpI = 1, kOpD = 2, kOpS = 4;   // CIGARBuilder.swift:4-10

// CIGARGenerator.inferBandWidth (:147-170): only the insertion gap cost enters.
int32_t infer_band(int32_t qlen, int32_t rlen, int32_t score, int32_t a, int32_t b,
                   const hipsw_scoring *sc)
{
    const int32_t diff = std::max(qlen, rlen) - std::min(qlen, rlen);
    const int32_t expected = std::min(qlen, rlen) * a;
    if (score >= expected) return diff + 3;
    const int32_t per_error = std::min(sc->o_ins + sc->e_ins, a + b);
    return std::max(diff + 3, (expected - score) / per_error);
}

char base_letter(uint8_t c) { return c < 4 ? "ACGT"[c] : 'N'; }   // :323-331

struct Region {
    const uint8_t *q, *r;
    int ql, rl;
};

// Walks M/I/D ops over query and reference (reference from `roff`), as NM and
// MD both do: M steps only while both sequences have bases left.
int32_t edit_distance(const std::vector<uint32_t> &cig, const Region &g, int roff)
{
    int32_t nm = 0;
    int qi = 0, ri = roff;
    for (uint32_t c : cig) {
        const uint32_t op = c & 0xF, len = c >> 4;
        if (op == kOpM) {
            for (uint32_t x = 0; x < len && qi < g.ql && ri < g.rl; ++x, ++qi, ++ri)
                nm += g.q[qi] != g.r[ri];
        } else if (op == kOpI) {
            nm += (int32_t)len;
            qi += (int)len;
        } else if (op == kOpD) {
            nm += (int32_t)len;
            ri += (int)len;
        }
    }
    return nm;
}

std::string md_tag(const std::vector<uint32_t> &cig, const Region &g, int roff)
{
    std::string md;
    int run = 0, qi = 0, ri = roff;
    for (uint32_t c : cig) {
        const uint32_t op = c & 0xF, len = c >> 4;
        if (op == kOpM) {
            for (uint32_t x = 0; x < len && qi < g.ql && ri < g.rl; ++x, ++qi, ++ri) {
                if (g.q[qi] == g.r[ri]) {
                    ++run;
                } else {
                    md += std::to_string(run);
                    md += base_letter(g.r[ri]);
                    run = 0;
                }
            }
        } else if (op == kOpI) {
            qi += (int)len;
        } else if (op == kOpD) {
            md += std::to_string(run);
            md += '^';
            for (uint32_t x = 0; x < len && ri < g.rl; ++x, ++ri) md += base_letter(g.r[ri]);
            run = 0;
        }
    }
    md += std::to_string(run);
    return md;
}

// Band and route of each region (CIGARGenerator.generate, :44-80).
void cigar_route(int32_t n, const uint16_t *q_len, const uint16_t *r_len, const int32_t *qb,
                 const int32_t *qe, const int32_t *true_score, const int32_t *initial_w,
                 const hipsw_scoring *sc, int32_t match_score, int32_t mismatch_penalty,
                 int32_t *w, uint8_t *route)
{
    const int32_t min_indel = sc->o_ins + sc->e_ins + sc->o_del + sc->e_del;   // :65-67
    for (int k = 0; k < n; ++k) {
        w[k] = std::max(infer_band(qe[k] - qb[k], r_len[k], true_score[k], match_score,
                                   mismatch_penalty, sc),
                        initial_w[k]);
        const int32_t deficit = (int32_t)q_len[k] * match_score - true_score[k];
        if (q_len[k] == r_len[k] && q_len[k] > 0 && deficit < min_indel) route[k] = 0;
        else if (q_len[k] == 0 || r_len[k] == 0) route[k] = 2;   // GlobalAligner.swift:29-31
        else route[k] = 1;
    }
}

// The NW result of a routed region: score, ops (the first min(n_ops, cap) of
// them), the full op count and the kernel's flags.
struct NwRow {
    int32_t score;
    const uint32_t *ops;
    int32_t n_ops, cap, flags;
};

// postProcess on the host (:117-142), regions in parallel.
void cigar_post(int32_t n, const uint8_t *qsegs, const uint32_t *q_off, const uint16_t *q_len,
                const uint8_t *rsegs, const uint32_t *r_off, const uint16_t *r_len,
                const int32_t *qb, const int32_t *qe, const int32_t *read_len,
                const uint8_t *is_reverse, const int32_t *true_score, const int64_t *ref_pos,
                const uint8_t *route, const std::function<NwRow(int64_t)> &nw_row, int32_t cap,
                int32_t md_cap, int32_t *out_score, int32_t *out_nm, int64_t *out_pos,
                uint32_t *out_cigar, int32_t *out_n_cigar, char *out_md, int32_t *out_flags)
{
    parallel_for(n, host_threads(), [&](int64_t lo, int64_t hi) {
        std::vector<uint32_t> cig;
        for (int64_t k = lo; k < hi; ++k) {
            const Region g{qsegs + q_off[k], rsegs + r_off[k], (int)q_len[k], (int)r_len[k]};
            int32_t flags = 0, score;
            cig.clear();
            int64_t pos = ref_pos[k];
            int roff = 0;
            if (route[k] == 0) {
                cig.push_back(((uint32_t)g.ql << 4) | kOpM);              // :69-80
                score = true_score[k];
            } else {
                if (route[k] == 1) {
                    const NwRow r = nw_row(k);
                    score = r.score;
                    if (r.n_ops > r.cap || (r.flags & 1)) flags |= 1;
                    cig.assign(r.ops, r.ops + std::max(0, std::min(r.n_ops, r.cap)));
                } else {
                    score = 0;
                }
                // squeezeDeletions (:272-295)
                size_t lead = 0;
                while (lead < cig.size() && (cig[lead] & 0xF) == kOpD) {
                    pos += cig[lead] >> 4;
                    roff += (int)(cig[lead] >> 4);
                    ++lead;
                }
                cig.erase(cig.begin(), cig.begin() + (long)lead);
                while (!cig.empty() && (cig.back() & 0xF) == kOpD) cig.pop_back();
            }
            const int32_t nm = edit_distance(cig, g, roff);
            const std::string md = md_tag(cig, g, roff);
            // addSoftClips (:298-320)
            const int32_t clip_front = is_reverse[k] ? read_len[k] - qe[k] : qb[k];
            const int32_t clip_back = is_reverse[k] ? qb[k] : read_len[k] - qe[k];
            if (clip_front > 0) cig.insert(cig.begin(), ((uint32_t)clip_front << 4) | kOpS);
            if (clip_back > 0) cig.push_back(((uint32_t)clip_back << 4) | kOpS);
            if ((int)cig.size() > cap) flags |= 1;
            if ((int)md.size() > md_cap - 1) flags |= 2;
            out_score[k] = score;
            out_nm[k] = nm;
            out_pos[k] = pos;
            out_n_cigar[k] = (int32_t)cig.size();
            uint32_t *oc = out_cigar + (size_t)k * cap;
            const int nc = std::min((int)cig.size(), (int)cap);
            std::copy(cig.begin(), cig.begin() + nc, oc);
            std::fill(oc + nc, oc + cap, 0u);
            char *om = out_md + (size_t)k * md_cap;
            const size_t nmd = std::min(md.size(), (size_t)md_cap - 1);
            std::memcpy(om, md.data(), nmd);
            std::memset(om + nmd, 0, (size_t)md_cap - nmd);
            out_flags[k] = flags;
        }
    });
}

}  // namespace

extern "C" int hipsw_cigar_generate(hipsw_engine *eng, int32_t n, const uint8_t *qsegs,
                                    const uint32_t *q_off, const uint16_t *q_len,
                                    const uint8_t *rsegs, const uint32_t *r_off,
                                    const uint16_t *r_len, const int32_t *qb, const int32_t *qe,
                                    const int32_t *read_len, const uint8_t *is_reverse,
                                    const int32_t *true_score, const int32_t *initial_w,
                                    const int64_t *ref_pos, const hipsw_scoring *sc,
                                    int32_t match_score, int32_t mismatch_penalty, int32_t cap,
                                    int32_t md_cap, int32_t *out_score, int32_t *out_nm,
                                    int64_t *out_pos, uint32_t *out_cigar, int32_t *out_n_cigar,
                                    char *out_md, int32_t *out_flags, hipsw_stats *stats)
{
    if (!eng) return HIPSW_E_INVALID;
    if (n == 0) return HIPSW_OK;
    if (n < 0 || cap < 1 || md_cap < 1 || !qsegs || !q_off || !q_len || !rsegs || !r_off ||
        !r_len || !qb || !qe || !read_len || !is_reverse || !true_score || !initial_w ||
        !ref_pos || !sc || !out_score || !out_nm || !out_pos || !out_cigar || !out_n_cigar ||
        !out_md || !out_flags)
        return HIPSW_E_INVALID;
    if (!scoring_valid(sc)) return HIPSW_E_INVALID;
    const double t0 = now_ms();

    std::vector<int32_t> w(n);
    std::vector<uint8_t> route(n);   // 0 fast path, 1 device NW, 2 empty segment
    cigar_route(n, q_len, r_len, qb, qe, true_score, initial_w, sc, match_score, mismatch_penalty,
                w.data(), route.data());
    std::vector<uint32_t> dev;       // regions that run NW
    for (int k = 0; k < n; ++k)
        if (route[k] == 1) dev.push_back((uint32_t)k);

    // Device NW with the retry loop for the routed subset.  Squeezing drops at
    // most two ops (merged leading and trailing deletions), so a raw CIGAR
    // longer than cap + 2 ops cannot fit `cap` and is flagged.
    const int nd = (int)dev.size();
    int raw_cap = 1;
    std::vector<uint32_t> sq_off(nd), sr_off(nd);
    std::vector<uint16_t> sq_len(nd), sr_len(nd);
    std::vector<int32_t> sw(nd), sts(nd);
    for (int s = 0; s < nd; ++s) {
        const uint32_t k = dev[s];
        sq_off[s] = q_off[k]; sq_len[s] = q_len[k];
        sr_off[s] = r_off[k]; sr_len[s] = r_len[k];
        sw[s] = w[k]; sts[s] = true_score[k];
        raw_cap = std::max(raw_cap, (int)q_len[k] + (int)r_len[k]);
    }
    raw_cap = std::min(raw_cap, cap + 2);
    std::vector<int32_t> d_score(nd), d_ncig(nd), d_flags(nd);
    std::vector<uint32_t> d_cig((size_t)nd * raw_cap);
    hipsw_stats nw_stats;
    std::memset(&nw_stats, 0, sizeof nw_stats);
    if (nd) {
        const int rc = hipsw_global_nw(eng, nd, qsegs, sq_off.data(), sq_len.data(), rsegs,
                                       sr_off.data(), sr_len.data(), sw.data(), sts.data(), 3,
                                       raw_cap, sc, HIPSW_MODE_BATCHED, d_score.data(),
                                       d_cig.data(), d_ncig.data(), d_flags.data(), nullptr,
                                       &nw_stats);
        if (rc != HIPSW_OK) return rc;
    }
    std::vector<int32_t> slot(n, -1);
    for (int s = 0; s < nd; ++s) slot[dev[s]] = s;
    cigar_post(n, qsegs, q_off, q_len, rsegs, r_off, r_len, qb, qe, read_len, is_reverse,
               true_score, ref_pos, route.data(),
               [&](int64_t k) {
                   const int s = slot[k];
                   return NwRow{d_score[s], d_cig.data() + (size_t)s * raw_cap, d_ncig[s], raw_cap,
                                d_flags[s]};
               },
               cap, md_cap, out_score, out_nm, out_pos, out_cigar, out_n_cigar, out_md, out_flags);
    if (stats) {
        *stats = nw_stats;
        stats->n_jobs = n;
        stats->total_ms = now_ms() - t0;
    }
    return HIPSW_OK;
}

extern "C" int hipsw_cigar_route(int32_t n, const uint16_t *q_len, const uint16_t *r_len,
                                 const int32_t *qb, const int32_t *qe, const int32_t *true_score,
                                 const int32_t *initial_w, const hipsw_scoring *sc,
                                 int32_t match_score, int32_t mismatch_penalty, int32_t *out_w,
                                 uint8_t *out_route)
{
    if (n == 0) return HIPSW_OK;
    if (n < 0 || !q_len || !r_len || !qb || !qe || !true_score || !initial_w || !sc || !out_w ||
        !out_route || !scoring_valid(sc))
        return HIPSW_E_INVALID;
    cigar_route(n, q_len, r_len, qb, qe, true_score, initial_w, sc, match_score, mismatch_penalty,
                out_w, out_route);
    return HIPSW_OK;
}

extern "C" int hipsw_cigar_finish(int32_t n, const uint8_t *qsegs, const uint32_t *q_off,
                                  const uint16_t *q_len, const uint8_t *rsegs,
                                  const uint32_t *r_off, const uint16_t *r_len, const int32_t *qb,
                                  const int32_t *qe, const int32_t *read_len,
                                  const uint8_t *is_reverse, const int32_t *true_score,
                                  const int64_t *ref_pos, const uint8_t *route,
                                  const int32_t *nw_score, const uint32_t *nw_cigar,
                                  const uint32_t *nw_cig_off, const int32_t *nw_n_cigar,
                                  const int32_t *nw_flags, int32_t nw_cap, int32_t cap,
                                  int32_t md_cap, int32_t *out_score, int32_t *out_nm,
                                  int64_t *out_pos, uint32_t *out_cigar, int32_t *out_n_cigar,
                                  char *out_md, int32_t *out_flags)
{
    if (n == 0) return HIPSW_OK;
    if (n < 0 || cap < 1 || md_cap < 1 || nw_cap < 1 || !qsegs || !q_off || !q_len || !rsegs ||
        !r_off || !r_len || !qb || !qe || !read_len || !is_reverse || !true_score || !ref_pos ||
        !route || !out_score || !out_nm || !out_pos || !out_cigar || !out_n_cigar || !out_md ||
        !out_flags)
        return HIPSW_E_INVALID;
    // NW rows come in region order over the route-1 regions
    std::vector<int32_t> slot(n, -1);
    int n_nw = 0;
    for (int k = 0; k < n; ++k) {
        if (route[k] > 2) return HIPSW_E_INVALID;
        if (route[k] == 1) slot[k] = n_nw++;
    }
    if (n_nw && (!nw_score || !nw_cigar || !nw_cig_off || !nw_n_cigar || !nw_flags))
        return HIPSW_E_INVALID;
    for (int s = 0; s < n_nw; ++s)
        if (nw_cig_off[s + 1] - nw_cig_off[s] != (uint32_t)std::min(std::max(nw_n_cigar[s], 0), nw_cap))
            return HIPSW_E_INVALID;
    cigar_post(n, qsegs, q_off, q_len, rsegs, r_off, r_len, qb, qe, read_len, is_reverse,
               true_score, ref_pos, route,
               [&](int64_t k) {
                   const int s = slot[k];
                   return NwRow{nw_score[s], nw_cigar + nw_cig_off[s], nw_n_cigar[s], nw_cap,
                                nw_flags[s]};
               },
               cap, md_cap, out_score, out_nm, out_pos, out_cigar, out_n_cigar, out_md, out_flags);
    return HIPSW_OK;
}
