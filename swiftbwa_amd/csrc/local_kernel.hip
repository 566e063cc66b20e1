// Mate-rescue local Smith-Waterman for gfx950, exact with LocalSWAligner.
//
// Semantics: LocalSWAligner.align, Sources/Alignment/LocalSWAligner.swift:42-79:
// a 16-lane uint8 Farrar striped pass (:133-306) with a single-pass lazy-F
// (:234-242), stripe-then-lane tie order (:250-262) and score2 peaks (:266-303);
// a reverse pass on the reversed prefixes recovers the start (:98-124); any
// overflow, an over-long query (qLen*A > 250) or a zero score falls back to the
// int32 scalar pass (:327-492).
//
// Layout of the striped pass: a 16-lane DPP row of the wave is one task and
// lane l of the row is SIMD byte-lane l of the reference's SIMD16<UInt8>, so
// every vector op of the reference is one VALU op, `shiftRight` is a DPP
// row_shr:1 with zero fill, and the per-stripe horizontal max is a 4-step
// in-row reduction.  Four tasks share a wave.  The u8 state lives in 32-bit
// registers; `pmax(x, c) &- c` is one clamped unsigned subtract.  Launches are
// per stripe count (the host sorts the tasks by sn = ceil(qlen/16)), so the
// stripe loops are unrolled to the exact count; a wave that mixes counts runs
// the padded form.  The best cell of a row is found from the row maximum
// directly, with the reference's stripe walk as the fallback.
//
// Layout of the scalar fallback: one task per lane, h/e rows in a per-wave HBM
// region interleaved as [column][lane].
#include "hipsw_internal.h"
#include "local_internal.h"

namespace hipsw {

__device__ __forceinline__ int row_shr1(int v)
{
    // DPP row_shr:1 (0x111), bound_ctrl: lane 0 of each 16-lane row reads 0.
    return __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, true);
}

// Maximum over each 16-lane row, in every lane of the row (values >= 0): four
// DPP steps, each fused into one v_max_i32_dpp by passing the identity as the
// "old" value.
__device__ __forceinline__ int row_max16(int v)
{
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false));    // quad_perm 1,0,3,2
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false));    // quad_perm 2,3,0,1
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false));   // row_half_mirror
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, false));   // row_mirror
    return v;
}

// u8 `pmax(x, c) &- c` for x, c in [0, 255]: a clamped unsigned subtract.
__device__ __forceinline__ uint32_t usat(uint32_t x, uint32_t c)
{
    return __builtin_elementwise_sub_sat(x, c);
}

// Base k of a 2-bit + N-mask sequence (k must be in range).
__device__ __forceinline__ int base_at(const LocalDev &d, uint32_t w0, uint32_t n0, int k)
{
    int b = (int)((d.seq[w0 + (k >> 4)] >> ((k & 15) * 2)) & 3u);
    if (n0 != kNoN && ((d.nmask[n0 + (k >> 5)] >> (k & 31)) & 1u)) b = 4;
    return b;
}

// One striped pass (simdSWPass, :133-306) for the task of this lane's 16-lane
// row.  `rev` reads the reversed prefixes query[0..qlen) / target[0..tlen).
// Returns false on 8-bit overflow.  colmax[i] receives the column maxima when
// peaks are tracked (min_sub > 0).  Every stripe 0..SNMAX-1 is computed in
// every row; with EXACT the caller guarantees that every active task of the
// wave has sn == SNMAX, otherwise stripes s >= sn of a task are padding whose
// cells are never read (the next row's diagonal, the lazy-F carry-in and the
// column maxima take stripes < sn only).
template <int SNMAX, bool EXACT>
__device__ bool striped_pass(const LocalDev &d, const LocalParams &p, bool act, uint32_t qw0,
                             uint32_t qn0, uint32_t tw0, uint32_t tn0, int qlen, int tlen,
                             bool rev, int min_sub, uint8_t *colmax, uint32_t *prof_lds,
                             int &o_score, int &o_qend,
                             int &o_tend, int &o_score2)
{
    const int lane16 = threadIdx.x & 15;
    const int grp = (threadIdx.x & 63) >> 4;
    const int sn = act ? (qlen + 15) / 16 : 0;
    const int bias = p.bias, goe = p.gap_oe, ge = p.gap_e;
    // striped profile, :152-167: lane l of stripe s holds sat8(mat[k*5+q[s+l*sn]] + bias),
    // 0 for padding.  Packed 4 stripes per register, one set per target code k.
    constexpr int NR = (SNMAX + 3) / 4;
    int qc[SNMAX];                                             // query code per stripe, -1 = padding
#pragma unroll
    for (int s = 0; s < SNMAX; ++s) {
        const int j = s + lane16 * sn;
        qc[s] = (act && s < sn && j < qlen) ? base_at(d, qw0, qn0, rev ? qlen - 1 - j : j) : -1;
    }
    // prof[k][r] of this lane lives in LDS at prof_lds[(k*NR + r)*64 + lane]:
    // the row's target code selects a set with a per-lane address.
    uint32_t *prof = prof_lds + (threadIdx.x & 63);
#pragma unroll
    for (int k = 0; k < 5; ++k) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            uint32_t x = 0;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int s = r * 4 + b;
                if (s < SNMAX) {
                    const int q = qc[s];
                    const int v = q < 0 ? 0 : min(max((int)p.mat[k * 5 + max(q, 0)] + bias, 0), 255);
                    x |= (uint32_t)v << (b * 8);
                }
            }
            prof[(k * NR + r) * 64] = x;
        }
    }
    // H and E are non-negative and at most 255 (u8 lanes held in 32-bit
    // registers); `usat` is the u8 `pmax(x, c) &- c` of the reference, one
    // clamped v_sub_u32.
    uint32_t H[SNMAX], E[SNMAX];
#pragma unroll
    for (int s = 0; s < SNMAX; ++s) { H[s] = 0; E[s] = 0; }
    int max_score = 0, max_i = -1, max_j = -1;
    bool ok = true;
    auto in_s = [&](int s) { return EXACT || s < sn; };            // stripe s is the task's
    auto last_s = [&](int s) { return EXACT ? s == SNMAX - 1 : s == sn - 1; };
    // wave-wide target length, as a scalar: the row loop is uniform
    int tmax_w = act ? tlen : 0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) tmax_w = max(tmax_w, __shfl_xor(tmax_w, o, 64));
    tmax_w = __builtin_amdgcn_readfirstlane(tmax_w);
    const uint32_t ubias = (uint32_t)bias, ugoe = (uint32_t)goe, uge = (uint32_t)ge;
    const uint64_t rowmask = 0xFFFFull << (grp * 16);
    // the target base is fetched one row ahead, so the global load overlaps
    // a row of work; ti = tlen - 1 - i reads the reversed prefix
    auto tbase = [&](int i) {
        return (act && i < tlen) ? min(base_at(d, tw0, tn0, rev ? tlen - 1 - i : i), 4) : 0;
    };
    int tbn = tbase(0);
    for (int i = 0; i < tmax_w; ++i) {
        const bool live = act && ok && i < tlen;
        const int tb = live ? tbn : 0;
        tbn = tbase(i + 1);
        uint32_t pr[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) pr[r] = prof[(tb * NR + r) * 64];
        // vH = shiftRight(H[sn-1]), :206
        uint32_t hlast = 0;
#pragma unroll
        for (int s = 0; s < SNMAX; ++s) hlast = last_s(s) ? H[s] : hlast;
        uint32_t vH = (uint32_t)row_shr1((int)hlast);
        uint32_t f = 0, fend = 0;
#pragma unroll
        for (int s = 0; s < SNMAX; ++s) {                      // :209-231
            // wrapping u8 &+ of the diagonal and the profile byte
            const uint32_t hd = (vH + (pr[s >> 2] >> ((s & 3) * 8))) & 0xFFu;
            const uint32_t hp = H[s];
            vH = hp;
            const uint32_t ev = usat(max(usat(hp, ugoe), E[s]), uge);
            E[s] = ev;
            const uint32_t hn = max(max(usat(hd, ubias), ev), f);
            H[s] = hn;
            f = usat(max(usat(hn, ugoe), f), uge);
            fend = last_s(s) ? f : fend;
        }
        // lazy-F, one pass with an early break (:234-242): a task goes on to
        // the next stripe while some lane of its 16-lane row changed
        uint32_t cf = (uint32_t)row_shr1((int)fend);
        bool go = live;
#pragma unroll
        for (int s = 0; s < SNMAX; ++s) {
            if (__builtin_amdgcn_ballot_w64(go) == 0) break;
            const bool in = go && in_s(s);
            const uint32_t nh = max(H[s], cf);
            const uint64_t changed = __builtin_amdgcn_ballot_w64(in && nh != H[s]);
            go = (changed & rowmask) != 0;
            H[s] = go ? nh : H[s];
            cf = usat(max(usat(nh, ugoe), cf), uge);
        }
        // column max and best cell (:245-263).  The column maximum (for the
        // peaks) and the overflow test need only the row maximum cmax over
        // all stripes.  The best cell changes only in a row where cmax beats
        // the task's best; the reference's stripe walk then ends at the first
        // stripe whose maximum is cmax, at its first lane holding cmax.  Such
        // a cell is a real query position when gap_oe >= 1 (a padding cell is
        // below both the previous best and the row's real maximum: it takes a
        // diagonal minus the bias, or a gap from an earlier row or an earlier
        // real cell), so that lane is found directly; otherwise (free gaps)
        // the search can fail and the wave replays the walk itself.
        uint32_t lmax = 0;
#pragma unroll
        for (int s = 0; s < SNMAX; ++s) lmax = in_s(s) ? max(lmax, H[s]) : lmax;
        const int cmax = row_max16((int)lmax);
        int col_max = cmax;
        if (__builtin_amdgcn_ballot_w64(live && ok && (cmax > max_score || cmax > 250)) != 0) {
            // first stripe holding cmax, per lane, then over the row
            int fs = 255;
#pragma unroll
            for (int s = SNMAX - 1; s >= 0; --s) fs = (in_s(s) && (int)H[s] == cmax) ? s : fs;
            const int sst = 255 - row_max16(255 - fs);
            const bool up = live && ok && cmax > max_score && cmax <= 250;
            const uint64_t hit = __ballot(up && fs == sst && sst + lane16 * sn < qlen);
            const uint32_t m16 = (uint32_t)((hit >> (grp * 16)) & 0xFFFFull);
            if (__builtin_amdgcn_ballot_w64(up && m16 == 0) == 0) {
                if (live && cmax > 250) ok = false;
                if (up) {
                    max_score = cmax;
                    max_i = i;
                    max_j = sst + (__ffs((int)m16) - 1) * sn;
                }
            } else {
                // the stripe walk itself (:245-263)
                col_max = 0;
#pragma unroll
                for (int s = 0; s < SNMAX; ++s) {
                    const bool in = live && in_s(s);
                    const int lm = row_max16(in ? (int)H[s] : 0);
                    if (in) {
                        if (lm > 250) ok = false;
                        if (ok) col_max = max(col_max, lm);
                    }
                    const bool better = in && ok && lm > max_score;
                    const int j = s + lane16 * sn;
                    const uint64_t hw = __ballot(better && (int)H[s] == lm && j < qlen);
                    if (better) {
                        max_score = lm;
                        max_i = i;
                        const uint32_t w16 = (uint32_t)((hw >> (grp * 16)) & 0xFFFFull);
                        if (w16) max_j = s + (__ffs((int)w16) - 1) * sn;
                    }
                }
            }
        }
        if (live && ok && min_sub > 0 && lane16 == 0) colmax[i] = (uint8_t)col_max;
    }
    o_score = max_score;
    o_qend = max_j;
    o_tend = max_i;
    o_score2 = 0;
    // score2 from the peak list (:266-303), replayed from the column maxima
    if (act && ok && min_sub > 0 && lane16 == 0 && max_score > 0 && p.max_base_u8 > 0) {
        const int r = (max_score + p.max_base_u8 - 1) / p.max_base_u8;
        const int lo = max_i - r, hi = max_i + r;
        int s2 = 0;
        int pk_col = -2, pk_sc = -1;
        auto close = [&]() {
            if (pk_sc >= 0 && (pk_col < lo || pk_col > hi) && pk_sc > s2) s2 = pk_sc;
        };
        for (int i = 0; i < tlen; ++i) {
            const int cm = colmax[i];
            if (cm < min_sub) continue;
            if (pk_sc < 0 || pk_col + 1 != i) {
                close();
                pk_col = i;
                pk_sc = cm;
            } else if (pk_sc < cm) {
                pk_col = i;
                pk_sc = cm;
            }
        }
        close();
        o_score2 = s2;
    }
    o_score2 = __shfl(o_score2, (threadIdx.x & 63) & ~15, 64);
    return ok;
}

// Tasks of one launch have at most SN stripes (sn = ceil(qlen/16); the host
// sorts them by sn, so most waves are uniform).  Each pass runs the exact form
// when every task of the wave has sn == SN and the padded form otherwise; the
// reverse pass, whose query prefix may need fewer stripes, decides again.
template <int SN>
__global__ __launch_bounds__(64) void local_striped_kernel(LocalDev d, LocalParams p,
                                                           const uint32_t *order, int n,
                                                           uint8_t *colmax_base, int tcap)
{
    const int grp = threadIdx.x >> 4;
    const int slot = blockIdx.x * 4 + grp;
    const bool valid = slot < n;
    const int job = valid ? (int)order[slot] : 0;
    const int qlen = valid ? d.q_len[job] : 0;
    const int tlen = valid ? d.t_len[job] : 0;
    const uint32_t qw0 = valid ? d.q_word[job] : 0, tw0 = valid ? d.t_word[job] : 0;
    const uint32_t qn0 = valid ? d.qn_word[job] : kNoN, tn0 = valid ? d.tn_word[job] : kNoN;
    uint8_t *colmax = colmax_base + (size_t)slot * tcap;
    // striped profile (prof_lds[(k * NR + r) * 64 + lane]), one buffer for
    // both forms of the pass
    __shared__ uint32_t prof_lds[5 * 4 * 64];
    // a task with more stripes than the launch's is left to the scalar pass
    const bool act = valid && qlen > 0 && tlen > 0 && (qlen + 15) / 16 <= SN;
    // pass 0: forward; pass 1: reverse on the reversed prefixes (:100-115)
    int fs = 0, fq = 0, ft = 0, f2 = 0, rq = 0, rt = 0;
    bool ok = false, ok2 = false;
    bool pact = act;
    int pql = qlen, ptl = tlen;
#pragma nounroll
    for (int pass = 0; pass < 2; ++pass) {
        const bool rev = pass == 1;
        const int msub = rev ? 0 : p.min_sub;
        int sc, qe, te, s2;
        bool r;
        if (__builtin_amdgcn_ballot_w64(pact && (pql + 15) / 16 != SN) == 0)
            r = striped_pass<SN, true>(d, p, pact, qw0, qn0, tw0, tn0, pql, ptl, rev, msub, colmax, prof_lds,
                                       sc, qe, te, s2);
        else
            r = striped_pass<SN, false>(d, p, pact, qw0, qn0, tw0, tn0, pql, ptl, rev, msub,
                                        colmax, prof_lds, sc, qe, te, s2);
        if (!rev) {
            fs = sc; fq = qe; ft = te; f2 = s2; ok = r;
            pact = act && ok && fs > 0 && fq >= 0;
            pql = pact ? fq + 1 : 0;
            ptl = pact ? ft + 1 : 0;
        } else {
            rq = qe; rt = te; ok2 = r;
        }
    }
    const bool act2 = pact;
    if (!valid || (threadIdx.x & 15) != 0) return;
    int32_t *o = d.out + (size_t)job * 6;
    if (act2 && ok2) {
        o[0] = fs; o[1] = fq - rq; o[2] = fq; o[3] = ft - rt; o[4] = ft; o[5] = f2;   // :117-124
        d.status[job] = LOCAL_OK;
    } else {
        d.status[job] = LOCAL_NEEDS_SCALAR;                    // :70-78
    }
}

// scalarSWPass (:383-492): query-outer int32 DP, strict '>' ties; h/e rows of
// tlen+1 in HBM as [column][lane]; peaks over query rows from rowmax[].
__device__ void scalar_pass(const LocalDev &d, const LocalParams &p, bool act, uint32_t qw0,
                            uint32_t qn0, uint32_t tw0, uint32_t tn0, int qcut, int tcut,
                            bool rev, int min_sub, int *hcol, int *ecol, int *rowmax,
                            int &o_score, int &o_qend, int &o_tend, int &o_score2)
{
    // qcut/tcut: lengths of the query/target prefixes aligned (reversed if rev)
    int best = 0, bq = -1, bt = -1;
    if (act) {
        for (int j = 0; j <= tcut; ++j) { hcol[j * 64] = 0; ecol[j * 64] = 0; }
        for (int i = 0; i < qcut; ++i) {
            const int qi = rev ? qcut - 1 - i : i;
            const int qb = min(base_at(d, qw0, qn0, qi), 4);
            int f = 0, hprev = 0, row_max = 0;
            for (int j = 0; j < tcut; ++j) {
                const int tj = rev ? tcut - 1 - j : j;
                const int tb = min(base_at(d, tw0, tn0, tj), 4);
                int hc = hprev + (int)p.mat[qb * 5 + tb];
                const int hj1 = hcol[(j + 1) * 64];
                const int fo = hj1 - p.gap_oe, fx = f - p.gap_e;
                f = fo > fx ? fo : fx;
                const int eo = hcol[j * 64] - p.gap_oe, ex = ecol[(j + 1) * 64] - p.gap_e;
                const int ev = eo > ex ? eo : ex;
                ecol[(j + 1) * 64] = ev;
                hc = hc > f ? hc : f;
                hc = hc > ev ? hc : ev;
                hc = hc > 0 ? hc : 0;
                hprev = hj1;
                hcol[(j + 1) * 64] = hc;
                if (hc > row_max) row_max = hc;
                if (hc > best) { best = hc; bq = i; bt = j; }
            }
            rowmax[i * 64] = row_max;
        }
    }
    o_score = best; o_qend = bq; o_tend = bt; o_score2 = 0;
    if (act && min_sub > 0 && best > 0 && p.max_base_i > 0) {
        const int r = (best + p.max_base_i - 1) / p.max_base_i;
        const int lo = bq - r, hi = bq + r;
        int s2 = 0, pk_col = -2, pk_sc = -1;
        auto close = [&]() {
            if (pk_sc >= 0 && (pk_col < lo || pk_col > hi) && pk_sc > s2) s2 = pk_sc;
        };
        for (int i = 0; i < qcut; ++i) {
            const int rm = rowmax[i * 64];
            if (rm < min_sub) continue;
            if (pk_sc < 0 || pk_col + 1 != i) { close(); pk_col = i; pk_sc = rm; }
            else if (pk_sc < rm) { pk_col = i; pk_sc = rm; }
        }
        close();
        o_score2 = s2;
    }
}

__global__ __launch_bounds__(64) void local_scalar_kernel(LocalDev d, LocalParams p,
                                                          const uint32_t *order, int n,
                                                          int *scratch, int tcap, int qcap)
{
    const int lane = threadIdx.x;
    const int slot = blockIdx.x * 64 + lane;
    const bool valid = slot < n;
    const int job = valid ? (int)order[slot] : 0;
    const int qlen = valid ? d.q_len[job] : 0;
    const int tlen = valid ? d.t_len[job] : 0;
    const uint32_t qw0 = valid ? d.q_word[job] : 0, tw0 = valid ? d.t_word[job] : 0;
    const uint32_t qn0 = valid ? d.qn_word[job] : kNoN, tn0 = valid ? d.tn_word[job] : kNoN;
    const bool act = valid && qlen > 0 && tlen > 0;
    int *base = scratch + (size_t)blockIdx.x * 64 * (2 * (tcap + 1) + qcap);
    int *hcol = base + lane, *ecol = base + (tcap + 1) * 64 + lane;
    int *rowmax = base + 2 * (tcap + 1) * 64 + lane;
    int fs, fq, ft, f2;
    scalar_pass(d, p, act, qw0, qn0, tw0, tn0, qlen, tlen, false, p.min_sub, hcol, ecol, rowmax,
                fs, fq, ft, f2);
    const bool act2 = act && fs > 0;
    int rs, rq, rt, r2;
    scalar_pass(d, p, act2, qw0, qn0, tw0, tn0, act2 ? fq + 1 : 0, act2 ? ft + 1 : 0, true, 0,
                hcol, ecol, rowmax, rs, rq, rt, r2);
    if (!valid) return;
    int32_t *o = d.out + (size_t)job * 6;
    if (act2) {
        o[0] = fs; o[1] = fq - rq; o[2] = fq; o[3] = ft - rt; o[4] = ft; o[5] = f2;
        d.status[job] = LOCAL_OK;
    } else {
        o[0] = o[1] = o[2] = o[3] = o[4] = o[5] = 0;
        d.status[job] = LOCAL_NIL;                             // :348
    }
}

void group_by_stripes(const uint32_t *ids, size_t n, const int32_t *q_len, uint32_t *out,
                      int32_t off[17])
{
    int32_t cnt[17] = {0};
    for (size_t k = 0; k < n; ++k) ++cnt[(q_len[ids[k]] + 15) / 16];
    off[0] = 0;
    for (int sn = 1; sn <= 16; ++sn) off[sn] = off[sn - 1] + cnt[sn];
    int32_t pos[17];
    for (int sn = 1; sn <= 16; ++sn) pos[sn] = off[sn - 1];
    for (size_t k = 0; k < n; ++k) out[pos[(q_len[ids[k]] + 15) / 16]++] = ids[k];
}

// Groups smaller than this many tasks are merged into the next larger group
// (and run padded there): a launch of a few waves costs a whole wave's
// latency, which a few hundred waves hide.
static constexpr int kMinLocalLaunch = 1024;

hipError_t launch_local_striped(const LocalDev &d, const LocalParams &p, const uint32_t *order,
                                const int32_t off[17], uint8_t *colmax, int tcap, hipStream_t s)
{
    int top = 0;                                               // largest non-empty group
    for (int sn = 1; sn <= 16; ++sn)
        if (off[sn] > off[sn - 1]) top = sn;
    int lo = 0;
    for (int sn = 1; sn <= top; ++sn) {
        const int n = off[sn] - lo;
        if (off[sn] == off[sn - 1] || (n < kMinLocalLaunch && sn < top)) continue;
        const uint32_t *o = order + lo;
        uint8_t *cm = colmax + (size_t)lo * tcap;
        const dim3 grid((n + 3) / 4), block(64);
#define HIPSW_LOCAL_SN(K) \
        case K: hipLaunchKernelGGL((local_striped_kernel<K>), grid, block, 0, s, d, p, o, n, cm, tcap); break;
        switch (sn) {
            HIPSW_LOCAL_SN(1) HIPSW_LOCAL_SN(2) HIPSW_LOCAL_SN(3) HIPSW_LOCAL_SN(4)
            HIPSW_LOCAL_SN(5) HIPSW_LOCAL_SN(6) HIPSW_LOCAL_SN(7) HIPSW_LOCAL_SN(8)
            HIPSW_LOCAL_SN(9) HIPSW_LOCAL_SN(10) HIPSW_LOCAL_SN(11) HIPSW_LOCAL_SN(12)
            HIPSW_LOCAL_SN(13) HIPSW_LOCAL_SN(14) HIPSW_LOCAL_SN(15) HIPSW_LOCAL_SN(16)
        }
#undef HIPSW_LOCAL_SN
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        lo = off[sn];
    }
    return hipSuccess;
}

hipError_t launch_local_scalar(const LocalDev &d, const LocalParams &p, const uint32_t *order,
                               int n, int *scratch, int tcap, int qcap, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(local_scalar_kernel, dim3((n + 63) / 64), dim3(64), 0, s, d, p, order, n,
                       scratch, tcap, qcap);
    return hipGetLastError();
}

}  // namespace hipsw
