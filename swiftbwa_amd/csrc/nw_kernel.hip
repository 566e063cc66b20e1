// Banded global Needleman-Wunsch with traceback, for gfx950.
//
// Semantics: GlobalAligner.align + traceback, Sources/Alignment/GlobalAligner.swift
// (band w' = max(w, |qlen - tlen|) :47, nCol = min(qlen, 2w'+1) :48, int32 state
// with the Int32.min/2 sentinel :52, first row :80-84, fill :89-142, score
// h[qlen] :144, traceback :160-232) restated per lane.
//
// State: int16 (sentinel remapped, exact under a per-job bound checked on the
// host) or int32, as the tier of the launch says.
//
// Layout: one wave64 fills 64 independent DP matrices, one per lane, row by
// row.  Lanes walk their own band in band-relative columns c = j - beg, so a
// wave whose lanes share w' runs every lane on every step.  The h/e row state
// of a lane is a circular buffer of S >= nCol + 2 slots in LDS (only the band
// is ever live: column j lives in slot j mod S), so the LDS footprint follows
// the band, not the query.  The 4-bit backpointers of each cell are packed 8
// per dword and stored band-relative as [row][c/8][lane] in a per-wave HBM
// region: every store of the wave writes 256 consecutive bytes.  Each lane then
// walks its own traceback from that region and writes its packed CIGAR
// (len << 4 | op, CIGARBuilder.swift:4-38) to its output row.
#include "hipsw_internal.h"
#include "nw_internal.h"

#include <type_traits>

namespace hipsw {

template <int T> struct NwSt;
template <> struct NwSt<NW_I32> {
    using W = int2;
    static constexpr int kNeg = INT32_MIN / 2;             // GlobalAligner.swift:52
    static __device__ __forceinline__ int h(W v) { return v.x; }
    static __device__ __forceinline__ int e(W v) { return v.y; }
    static __device__ __forceinline__ W pack(int hv, int ev) { return make_int2(hv, ev); }
};
template <> struct NwSt<NW_I16> {
    using W = uint32_t;
    static constexpr int kNeg = -16384;
    static __device__ __forceinline__ int h(W v) { return (int)(int16_t)(v & 0xFFFFu); }
    static __device__ __forceinline__ int e(W v) { return (int)(int16_t)(v >> 16); }
    static __device__ __forceinline__ W pack(int hv, int ev)
    {
        return ((uint32_t)hv & 0xFFFFu) | ((uint32_t)ev << 16);
    }
};

__device__ __forceinline__ int wave_max_nw(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// Wave-uniform minima of two per-lane u32 (all 64 lanes active), the two
// chains interleaved; fused v_min_u32_dpp steps (quad_perm and mirror steps within rows, row_bcast:15/31 into row 3; lane 63 holds the result).
__device__ __forceinline__ void wave_min2_u32(uint32_t &a, uint32_t &b)
{
#define HIPSW_NW_MIN2(CTRL, RM)                                                               \
    a = min(a, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)a, CTRL, RM, 0xF, false));      \
    b = min(b, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)b, CTRL, RM, 0xF, false));
    HIPSW_NW_MIN2(0xB1, 0xF)
    HIPSW_NW_MIN2(0x4E, 0xF)
    HIPSW_NW_MIN2(0x141, 0xF)
    HIPSW_NW_MIN2(0x140, 0xF)
    HIPSW_NW_MIN2(0x142, 0xA)
    HIPSW_NW_MIN2(0x143, 0xC)
#undef HIPSW_NW_MIN2
    a = __builtin_amdgcn_readlane(a, 63);
    b = __builtin_amdgcn_readlane(b, 63);
}


// Traceback (GlobalAligner.swift:160-232) of one job by one lane, with
// CIGARBuilder merging (CIGARBuilder.swift:4-38); ops come out last-first.
// They go to a ring of `cap` slots, so when the CIGAR has more than cap ops
// the row still ends up holding the first cap ops of the final CIGAR.
// zf(i, c) returns the 4 flag bits of row i, band column c = j - max(0, i - w):
// bit 0 E opened from H (e_open > e_ext), bit 1 F opened from H
// (f_open > f_ext), bit 2 H taken from E, bit 3 H taken from F -- the
// reference's z byte with bits 0 and 1 negated (:114-136).
//
// wr: this lane writes the output.  A wave may walk one job's traceback with
// every lane at once (wave-uniform state, flags read as uniform values) and
// lane 0 alone writing.
template <class ZF>
__device__ __forceinline__ void nw_traceback(const NwDev &d, int job, int qlen, int tlen, int w,
                                             int ncol, ZF zf, bool wr = true)
{
    uint32_t *cig = d.cigar + (size_t)job * d.cap;
    int i = tlen - 1, jj = qlen - 1, state = 0;
    int nops = 0;
    uint32_t cur_op = 0xFu, cur_len = 0;
    const int cap = d.cap;
    int slot = 0;
    auto push = [&]() {
        if (wr) cig[slot] = (cur_len << 4) | cur_op;
        slot = slot + 1 == cap ? 0 : slot + 1;
        ++nops;
    };
    auto emit = [&](uint32_t op) {
        if (op == cur_op) {
            ++cur_len;
        } else {
            if (cur_len) push();
            cur_op = op;
            cur_len = 1;
        }
    };
    while (i >= 0 || jj >= 0) {
        if (i < 0) { emit(1u); --jj; continue; }
        if (jj < 0) { emit(2u); --i; continue; }
        const int beg = max(0, i - w);
        const int c = jj - beg;
        if (!(c >= 0 && c < ncol)) break;
        const uint32_t dd = zf(i, c);
        if (state == 0) {
            const uint32_t hdir = (dd >> 2) & 3u;
            if (hdir == 0) { emit(0u); --i; --jj; }
            else if (hdir == 1) state = 1;
            else state = 2;
        } else if (state == 1) {
            emit(2u);
            if (dd & 1u) state = 0;                            // E was opened from H
            --i;
        } else {
            emit(1u);
            if (dd & 2u) state = 0;                            // F was opened from H
            --jj;
        }
    }
    if (cur_len) push();
    // Final order = reverse of emission order (:230).  With the ring, slots
    // [0, p) and [p, kept) are each reversed in place, p = nops mod cap.
    if (!wr) return;
    const int kept = min(nops, cap);
    const int p_split = nops > cap ? nops % cap : kept;
    for (int a = 0, b = p_split - 1; a < b; ++a, --b) {
        const uint32_t tmp = cig[a];
        cig[a] = cig[b];
        cig[b] = tmp;
    }
    for (int a = p_split, b = kept - 1; a < b; ++a, --b) {
        const uint32_t tmp = cig[a];
        cig[a] = cig[b];
        cig[b] = tmp;
    }
    d.n_cigar[job] = nops;
    d.flags[job] = nops > d.cap ? NW_FLAG_TRUNCATED : 0;
}

// Fill + traceback for the (<= 64) jobs idx[0..count).
template <int T>
__device__ void nw_wave(const NwDev &d, const NwParams &p, const uint32_t *idx, int count,
                        typename NwSt<T>::W *slots, int S, uint32_t *zreg, int zrows, int ncw,
                        const int8_t *smat, uint32_t *qlds, int qw8)
{
    using St = NwSt<T>;
    using Wd = typename St::W;
    constexpr int kNegInf = St::kNeg;
    const int lane = threadIdx.x & 63;
    const bool valid = lane < count;
    const int job = valid ? (int)idx[lane] : 0;
    const int qlen = valid ? d.q_len[job] : 0;
    const int tlen = valid ? d.t_len[job] : 0;
    const bool nonempty = valid && qlen > 0 && tlen > 0;     // :29-31
    int w = valid ? min(d.w[job], 1 << 20) : 0;   // wider than any band: same cells
    const int dl = qlen > tlen ? qlen - tlen : tlen - qlen;
    w = max(w, dl);                                            // :47
    const int ncol = min(qlen, 2 * w + 1);                     // :48
    const uint32_t qw0 = valid ? d.q_word[job] : 0;
    const uint32_t tw0 = valid ? d.t_word[job] : 0;
    const uint32_t qn0 = valid ? d.qn_word[job] : kNoN;
    const uint32_t tn0 = valid ? d.tn_word[job] : kNoN;
    Wd *col = slots + lane;                                    // slot s at col[s * 64]

    // Stage the query as 4-bit codes, 8 bases per LDS word: qlds[b*64 + lane]
    // holds bases 8b..8b+7 (code 4 = N).
    uint32_t *ql = qlds + lane;
    for (int b = 0; b < qw8; ++b) {
        uint32_t x = 0;
        if (nonempty && b * 8 < qlen) {
            x = (d.seq[qw0 + (b >> 1)] >> ((b & 1) * 16)) & 0xFFFFu;
            x = (x | (x << 8)) & 0x00FF00FFu;                   // spread 2-bit codes
            x = (x | (x << 4)) & 0x0F0F0F0Fu;                   // to 4-bit fields
            x = (x | (x << 2)) & 0x33333333u;
            if (qn0 != kNoN) {
                const uint32_t nb = (d.nmask[qn0 + (b >> 2)] >> ((b & 3) * 8)) & 0xFFu;
                for (int k = 0; k < 8; ++k)
                    if ((nb >> k) & 1u) x = (x & ~(0xFu << (4 * k))) | (4u << (4 * k));
            }
        }
        ql[b * 64] = x;
    }

    // First row (:75-84): h[0] = 0, h[j] = -(o_ins + e_ins*j) for j <= min(w,qlen),
    // everything else (and all of e) the sentinel.
    if (nonempty) {
        const int lim = min(qlen, S - 1);
        const int init_end = min(w, qlen);
        for (int s = 0; s <= lim; ++s) {
            const int hv = s == 0 ? 0 : (s <= init_end ? -(p.o_ins + p.e_ins * s) : kNegInf);
            col[s * 64] = St::pack(hv, kNegInf);
        }
    }
    const int cmax = wave_max_nw(nonempty ? ncol : 0);        // wave-uniform row width
    const int nwords = (cmax + 7) >> 3;
    uint32_t *zl = zreg + lane;                                // word (i, cw) at zl[(i*ncw+cw)*64]
    const int tmax = wave_max_nw(nonempty ? tlen : 0);

    uint32_t tw = 0, tn = 0;
    uint64_t ncells = 0;
    int sb = 0;                                                // slot of column beg
    int beg_prev = 0;
    for (int i = 0; i < tmax; ++i) {                           // :89
        const bool act = nonempty && i < tlen;
        if ((i & 15) == 0 && act) tw = d.seq[tw0 + (i >> 4)];
        int t = (tw >> ((i & 15) * 2)) & 3;
        if (tn0 != kNoN) {
            if ((i & 31) == 0 && act) tn = d.nmask[tn0 + (i >> 5)];
            if ((tn >> (i & 31)) & 1) t = 4;
        }
        const int t5 = t * 5;
        const int beg = max(0, i - w);                         // :90-91
        const int end = min(qlen, i + w + 1);
        const int width = end - beg;
        if (act && width > 0) ncells += (uint64_t)width;
        if (act) {
            for (int b = beg_prev; b < beg; ++b) sb = sb + 1 == S ? 0 : sb + 1;
        }
        beg_prev = beg;
        int f = kNegInf, h_diag = 0, h_left = 0;
        if (act) {                                             // :96-104
            const Wd v = col[sb * 64];
            h_diag = St::h(v);
            h_left = beg == 0 ? 0 - (p.oe_del + p.e_del * i) : kNegInf;
            col[sb * 64] = St::pack(h_left, St::e(v));
        }
        // Cells go in blocks of 8 (one backpointer dword per lane): the 8 row
        // state words and the 8 query codes of a block are read before any of
        // the block's writes, so only the F/H chain is serial.  S >= 10 keeps
        // the 8 slots of a block distinct.
        int s1 = sb + 1 == S ? 0 : sb + 1;                     // slot of column beg+1
        const int qlim = qw8 - 1;
        for (int cb = 0; cb * 8 < cmax; ++cb) {                // :108-141
            int sl[8];
            Wd v[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                sl[k] = s1;
                s1 = s1 + 1 == S ? 0 : s1 + 1;
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) v[k] = col[sl[k] * 64];
            const int j0 = beg + cb * 8;
            const int wq = max(0, min(j0 >> 3, qlim));
            const uint64_t win = (uint64_t)ql[wq * 64] | ((uint64_t)ql[min(wq + 1, qlim) * 64] << 32);
            const uint32_t codes = (uint32_t)(win >> ((j0 & 7) * 4));
            // the block's eight scores are read before any of its writes too:
            // a read after a write to the same memory would wait for it, once
            // per cell of the serial chain
            int sc[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) sc[k] = smat[t5 + (int)((codes >> (4 * k)) & 0xFu)];
            // Flag bits (the serial loop's, :114-136, with bits 0 and 1
            // stored as "opened": E opened = e_open > e_ext, F opened =
            // f_open > f_ext; bit 2 H from E, bit 3 H from F) are the sign
            // bits of differences, shifted into zword with one alignbit each:
            // cell k's nibble ends up at bits 28 - 4k .. 31 - 4k.
            uint32_t zword = 0;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                // Lanes past their band end (or without a job) keep their LDS
                // state and record no flags.
                const bool inb = act && cb * 8 + k < width;
                const int h_cur = h_diag + sc[k];
                const int hd = St::h(v[k]);
                const int e_open = hd - p.oe_del;
                const int e_ext = St::e(v[k]) - p.e_del;
                const int ev = max(e_open, e_ext);
                const int f_open = h_left - p.oe_ins;
                const int f_ext = f - p.e_ins;
                f = max(f_open, f_ext);
                const int h1 = max(ev, h_cur);
                const int hn = max(f, h1);
                zword = __builtin_amdgcn_alignbit(zword, (uint32_t)(h1 - f), 31);        // f > h1
                zword = __builtin_amdgcn_alignbit(zword, (uint32_t)(h_cur - ev), 31);    // ev > h_cur
                zword = __builtin_amdgcn_alignbit(zword, (uint32_t)(f_ext - f_open), 31);
                zword = __builtin_amdgcn_alignbit(zword, (uint32_t)(e_ext - e_open), 31);
                col[sl[k] * 64] = inb ? St::pack(hn, ev) : v[k];
                h_diag = hd;
                h_left = hn;
            }
            {
                // cells past the band end record no flags
                const int nv = act ? min(max(width - cb * 8, 0), 8) : 0;
                zword &= nv >= 8 ? 0xFFFFFFFFu : ~(0xFFFFFFFFu >> (4 * nv)) & (nv ? 0xFFFFFFFFu : 0u);
            }
            if (i < zrows) zl[((size_t)i * ncw + cb) * 64] = zword;
        }
        if (act) {
            // Column end+1 is read by the next row; in the full-width array it
            // still holds its first-row value, the sentinel (end+1 > w').
            const int s_next = (sb + width + 1) % S;
            if (end + 1 <= qlen) col[s_next * 64] = St::pack(kNegInf, kNegInf);
        }
    }
    (void)nwords;
    if (d.cells) {
        unsigned long long c = ncells;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
        if (lane == 0) atomicAdd(d.cells, c);
    }

    if (!valid) return;
    int32_t *o_score = d.score + job;
    if (!nonempty) {
        *o_score = 0;
        d.n_cigar[job] = 0;
        d.flags[job] = 0;
        return;
    }
    // score = h[qlen] (:144); column qlen sits in slot (sb + (qlen - beg_last)) mod S.
    {
        const int beg_last = max(0, tlen - 1 - w);
        const int s = (sb + (qlen - beg_last)) % S;
        *o_score = St::h(col[s * 64]);
    }
    nw_traceback(d, job, qlen, tlen, w, ncol, [&](int i, int c) {
        return (zl[((size_t)i * ncw + (c >> 3)) * 64] >> ((7 - (c & 7)) * 4)) & 0xFu;
    });
}

// The 16-bit tier with the row state in registers, for bands of at most K
// columns (K a multiple of 8).  R[k] holds the packed (h, e) of column beg + k
// of the lane's current row: the row's h_diag is R[0] (R[1] when the band
// start advances, which it does by one column per row once i > w'), cell
// beg + k reads the previous row's column beg + k + 1 at R[k + 1] (R[k + 2]
// when advancing) and writes R[k + 1], so the frame moves with the band by a
// per-lane select instead of a ring.  A row runs the 8-cell blocks that hold
// every active lane's band and the cell after it (at most ceil(cmax / 8),
// cmax the wave's widest band); cells past a lane's band end copy what they
// read, so the columns right of the band keep the values they had in the
// first row (the sentinel: they lie beyond min(w', qlen)), and entries past
// the blocks a row runs hold that sentinel before and after the row.
//
// Two forms of a block: on rows where every active lane advances, a block
// that lies inside every active lane's band runs without the shift select and
// without masks (a lane past its last row may then overwrite its state, so
// each lane's score h[qlen] is taken at its last row); other blocks select
// and mask per lane.  The score of a cell comes from the target base's row of
// the matrix held in two registers (codes 0-3 in the low word, 4 in the
// high), looked up four cells at a time with a byte permute of the staged
// query codes.  Only the query and the scoring tables stay in LDS.  Flags,
// traceback and output are those of nw_wave.
template <int K>
__device__ void nw_wave_reg(const NwDev &d, const NwParams &p, const uint32_t *idx, int count,
                            uint32_t *zreg, int zrows, int ncw, const uint32_t *rowtab,
                            uint32_t *qlds, int qw8)
{
    using St = NwSt<NW_I16>;
    constexpr int kNegInf = St::kNeg;
    const int lane = threadIdx.x & 63;
    const bool valid = lane < count;
    const int job = valid ? (int)idx[lane] : 0;
    const int qlen = valid ? d.q_len[job] : 0;
    const int tlen = valid ? d.t_len[job] : 0;
    const bool nonempty = valid && qlen > 0 && tlen > 0;     // :29-31
    int w = valid ? min(d.w[job], 1 << 20) : 0;
    const int dl = qlen > tlen ? qlen - tlen : tlen - qlen;
    w = max(w, dl);                                            // :47
    const int ncol = min(qlen, 2 * w + 1);                     // :48
    const uint32_t qw0 = valid ? d.q_word[job] : 0;
    const uint32_t tw0 = valid ? d.t_word[job] : 0;
    const uint32_t qn0 = valid ? d.qn_word[job] : kNoN;
    const uint32_t tn0 = valid ? d.tn_word[job] : kNoN;

    // Query staged as in nw_wave: 4-bit codes, 8 per word, qlds[b*64 + lane],
    // then K/8 zero words: a row reads words beg/8 .. beg/8 + ceil(cmax/8)
    // with beg <= qlen - 1, so at most word qw8 - 1 + K/8.
    uint32_t *ql = qlds + lane;
    for (int b = qw8; b < qw8 + K / 8; ++b) ql[b * 64] = 0;
    for (int b = 0; b < qw8; ++b) {
        uint32_t x = 0;
        if (nonempty && b * 8 < qlen) {
            x = (d.seq[qw0 + (b >> 1)] >> ((b & 1) * 16)) & 0xFFFFu;
            x = (x | (x << 8)) & 0x00FF00FFu;
            x = (x | (x << 4)) & 0x0F0F0F0Fu;
            x = (x | (x << 2)) & 0x33333333u;
            if (qn0 != kNoN) {
                const uint32_t nb = (d.nmask[qn0 + (b >> 2)] >> ((b & 3) * 8)) & 0xFFu;
                for (int k = 0; k < 8; ++k)
                    if ((nb >> k) & 1u) x = (x & ~(0xFu << (4 * k))) | (4u << (4 * k));
            }
        }
        ql[b * 64] = x;
    }

    // First row (:75-84) in the frame of beg = 0: R[s] is column s.
    uint32_t R[K + 2];
    {
        const int init_end = min(w, qlen);
#pragma unroll
        for (int s = 0; s < K + 2; ++s) {
            const int hv = s == 0 ? 0 : (s <= init_end ? -(p.o_ins + p.e_ins * s) : kNegInf);
            R[s] = St::pack(hv, kNegInf);
        }
    }
    const int cmax = wave_max_nw(nonempty ? ncol : 0);
    const int wmax = wave_max_nw(nonempty ? w : 0);
    uint32_t *zl = zreg + lane;                                // word (i, cw) at zl[(i*ncw+cw)*64]
    const int tmax = wave_max_nw(nonempty ? tlen : 0);
    // score = h[qlen] (:144): column qlen of the last row is R[qlen - beg_last]
    const int sidx = qlen - max(0, tlen - 1 - w);
    int score = 0;

    uint32_t tw = 0, tn = 0;
    uint64_t ncells = 0;
    for (int i = 0; i < tmax; ++i) {                           // :89
        const bool act = nonempty && i < tlen;
        if ((i & 15) == 0 && act) tw = d.seq[tw0 + (i >> 4)];
        int t = (tw >> ((i & 15) * 2)) & 3;
        if (tn0 != kNoN) {
            if ((i & 31) == 0 && act) tn = d.nmask[tn0 + (i >> 5)];
            if ((tn >> (i & 31)) & 1) t = 4;
        }
        const uint32_t tab_lo = rowtab[2 * t], tab_hi = rowtab[2 * t + 1];
        const int beg = max(0, i - w);                         // :90-91
        const int end = min(qlen, i + w + 1);
        const int width = act ? end - beg : 0;                 // cells this row
        ncells += (uint64_t)max(width, 0);
        const bool adv = act && i > w;                         // beg = i - w moved by one
        // Blocks [0, nfull) lie inside every active lane's band; on rows
        // past every lane's ramp they run unmasked.
        // A row runs the blocks holding every active lane's band and the
        // cell after it (the shift carries the sentinel, and the traceback
        // may read the flags one cell past the band, which must be zero),
        // at most ceil(cmax / 8); the entries beyond hold the sentinel in
        // both frames.
        uint32_t wlo = act ? (uint32_t)width : 0xFFFFFFFFu;
        uint32_t whi = act ? ~(uint32_t)width : 0xFFFFFFFFu;
        wave_min2_u32(wlo, whi);
        const int wmin = (int)wlo, wmaxr = (int)~whi;
        const int nblk = min((wmaxr + 8) >> 3, (cmax + 7) >> 3);
        const int nfull = i > wmax ? min(wmin, cmax) >> 3 : 0;
        // this row's query words: from beg / 8 (0 for lanes without a row)
        const int qb = act ? beg : 0;
        const uint32_t *qrow = ql + (qb >> 3) * 64;
        const uint32_t qsh = (uint32_t)(qb & 7) * 4;
        uint32_t qlo = qrow[0];
        int f = kNegInf, h_diag, h_left;
        {                                                      // :96-104
            const uint32_t v0 = adv ? R[1] : R[0];
            h_diag = St::h(v0);
            h_left = beg == 0 ? 0 - (p.oe_del + p.e_del * i) : kNegInf;
            if (act) R[0] = St::pack(h_left, kNegInf);         // column beg's e is never read again
        }
#pragma unroll
        for (int cb = 0; cb < K / 8; ++cb) {                   // :108-141
            // a guard, not a break: with an exit the loop has a run-time trip
            // count, is not unrolled, and R goes to scratch memory
            if (cb < nblk) {
                const uint32_t qhi = qrow[(cb + 1) * 64];
                const uint32_t codes = __builtin_amdgcn_alignbit(qhi, qlo, qsh);
                qlo = qhi;
                // codes of cells 0..7 as byte selectors, then the row's scores
                const uint32_t xe = codes & 0x0F0F0F0Fu;           // cells 0, 2, 4, 6
                const uint32_t xo = (codes >> 4) & 0x0F0F0F0Fu;    // cells 1, 3, 5, 7
                const uint32_t s03 = __builtin_amdgcn_perm(tab_hi, tab_lo,
                                                           __builtin_amdgcn_perm(xo, xe, 0x05010400u));
                const uint32_t s47 = __builtin_amdgcn_perm(tab_hi, tab_lo,
                                                           __builtin_amdgcn_perm(xo, xe, 0x07030602u));
                uint32_t zword = 0;
                if (cb < nfull) {
    #pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        const int kk = cb * 8 + k;
                        const uint32_t v = R[kk + 2];
                        const int sc = (int)(int8_t)((k < 4 ? s03 : s47) >> (8 * (k & 3)));
                        const int h_cur = h_diag + sc;
                        const int hd = St::h(v);
                        const int e_open = hd - p.oe_del;
                        const int e_ext = St::e(v) - p.e_del;
                        const int ev = max(e_open, e_ext);
                        const int f_open = h_left - p.oe_ins;
                        const int f_ext = f - p.e_ins;
                        f = max(f_open, f_ext);
                        const int h1 = max(ev, h_cur);
                        const int hn = max(f, h1);
                        zword = __builtin_amdgcn_alignbit(zword, (uint32_t)(h1 - f), 31);
                        zword = __builtin_amdgcn_alignbit(zword, (uint32_t)(h_cur - ev), 31);
                        zword = __builtin_amdgcn_alignbit(zword, (uint32_t)(f_ext - f_open), 31);
                        zword = __builtin_amdgcn_alignbit(zword, (uint32_t)(e_ext - e_open), 31);
                        R[kk + 1] = __builtin_amdgcn_perm((uint32_t)ev, (uint32_t)hn, 0x05040100u);
                        // the new state is computed here, not sunk to the row's
                        // end: sinking keeps every cell's h and e live across the
                        // row and the old and new arrays in separate registers
                        asm volatile("" : "+v"(R[kk + 1]));
                        h_diag = hd;
                        h_left = hn;
                    }
                } else {
    #pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        const int kk = cb * 8 + k;
                        const uint32_t v = adv ? R[kk + 2] : R[kk + 1];
                        const bool inb = kk < width;
                        const int sc = (int)(int8_t)((k < 4 ? s03 : s47) >> (8 * (k & 3)));
                        const int h_cur = h_diag + sc;
                        const int hd = St::h(v);
                        const int e_open = hd - p.oe_del;
                        const int e_ext = St::e(v) - p.e_del;
                        const int ev = max(e_open, e_ext);
                        const int f_open = h_left - p.oe_ins;
                        const int f_ext = f - p.e_ins;
                        f = max(f_open, f_ext);
                        const int h1 = max(ev, h_cur);
                        const int hn = max(f, h1);
                        zword = __builtin_amdgcn_alignbit(zword, (uint32_t)(h1 - f), 31);        // f > h1
                        zword = __builtin_amdgcn_alignbit(zword, (uint32_t)(h_cur - ev), 31);    // ev > h_cur
                        zword = __builtin_amdgcn_alignbit(zword, (uint32_t)(f_ext - f_open), 31);
                        zword = __builtin_amdgcn_alignbit(zword, (uint32_t)(e_ext - e_open), 31);
                        R[kk + 1] = inb ? __builtin_amdgcn_perm((uint32_t)ev, (uint32_t)hn, 0x05040100u) : v;
                        asm volatile("" : "+v"(R[kk + 1]));
                        h_diag = hd;
                        h_left = hn;
                    }
                    // cells past the band end record no flags
                    const int nv = min(max(width - cb * 8, 0), 8);
                    zword &= nv >= 8 ? 0xFFFFFFFFu : ~(0xFFFFFFFFu >> (4 * nv)) & (nv ? 0xFFFFFFFFu : 0u);
                }
                zl[((size_t)i * ncw + cb) * 64] = zword;           // i < tmax <= zrows
            }
        }
        // lanes on their last row take h[qlen] before a later unmasked block
        // may overwrite their state
        if (__builtin_amdgcn_ballot_w64(act && i == tlen - 1)) {
            const int si = act && i == tlen - 1 ? sidx : -1;
#pragma unroll
            for (int s = 0; s < K + 2; ++s)
                if (s == si) score = St::h(R[s]);
        }
    }
    if (d.cells) {
        unsigned long long c = ncells;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
        if (lane == 0) atomicAdd(d.cells, c);
    }

    if (!valid) return;
    if (!nonempty) {
        d.score[job] = 0;
        d.n_cigar[job] = 0;
        d.flags[job] = 0;
        return;
    }
    d.score[job] = score;
    nw_traceback(d, job, qlen, tlen, w, ncol, [&](int i, int c) {
        return (zl[((size_t)i * ncw + (c >> 3)) * 64] >> ((7 - (c & 7)) * 4)) & 0xFu;
    });
}

__device__ __forceinline__ const int8_t *nw_load_smat(const NwParams &p)
{
    extern __shared__ __align__(16) unsigned char nw_lds[];
    int8_t *smat = reinterpret_cast<int8_t *>(nw_lds);
    if (threadIdx.x < 32) smat[threadIdx.x] = p.mat[threadIdx.x];
    __syncthreads();
    return smat;
}

// Persistent work-stealing kernel (items ordered by decreasing area, taken
// from an atomic counter).  The plain form runs the same body over items in
// input order with a static round-robin assignment of items to blocks.
template <int T, bool GLOBAL>
__global__ __launch_bounds__(64) void nw_kernel(NwDev d, NwParams p, const uint32_t *order,
                                                const WorkItem *items, int n_items,
                                                uint32_t *counter, int S, uint32_t *zbase,
                                                int zrows, int ncw, int persistent, void *gstate,
                                                int qw8)
{
    using Wd = typename NwSt<T>::W;
    extern __shared__ __align__(16) unsigned char nw_lds[];
    const int8_t *smat = nw_load_smat(p);
    uint32_t *qlds = reinterpret_cast<uint32_t *>(nw_lds + 32);
    Wd *slots = GLOBAL ? reinterpret_cast<Wd *>(gstate) + (size_t)blockIdx.x * S * 64
                       : reinterpret_cast<Wd *>(nw_lds + 32 + (size_t)qw8 * 256);
    uint32_t *zreg = zbase + (size_t)blockIdx.x * zrows * ncw * 64;
    __shared__ uint32_t idx[64];
    __shared__ int item_sh;
    int it = blockIdx.x;
    for (;;) {
        if (persistent) {
            if (threadIdx.x == 0) item_sh = (int)atomicAdd(counter, 1u);
            __syncthreads();
            it = item_sh;
        }
        if (it >= n_items) break;
        const WorkItem wi = items[it];
        if (threadIdx.x < wi.count) idx[threadIdx.x] = order[wi.start + threadIdx.x];
        __syncthreads();
        nw_wave<T>(d, p, idx, (int)wi.count, slots, S, zreg, zrows, ncw, smat, qlds, qw8);
        __syncthreads();
        if (!persistent) it += gridDim.x;      // plain form: static round-robin
    }
}

// The register-state form of nw_kernel (16-bit tier, bands of at most K
// columns); LDS holds the scoring table and the staged query only.  Each
// instance is held to the register budget of 5, 4, 3 or 2 waves per SIMD
// (K = 32, 64, 96, 128): the row state takes K + 2 registers, and the few
// values the limit pushes out are per-job ones, reloaded outside the cells.
template <int K>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(K <= 32 ? 5 : K <= 64 ? 4 : K <= 96 ? 3 : 2)))
void nw_reg_kernel(NwDev d, NwParams p, const uint32_t *order,
                                                    const WorkItem *items, int n_items,
                                                    uint32_t *counter, uint32_t *zbase, int zrows,
                                                    int ncw, int persistent, int qw8)
{
    extern __shared__ __align__(16) unsigned char nw_lds[];
    // rows of the scoring matrix by target code t: word 2t holds the scores
    // of query codes 0-3 (bytes 0-3), word 2t + 1 that of code 4 (byte 0)
    uint32_t *rowtab = reinterpret_cast<uint32_t *>(nw_lds);
    if (threadIdx.x < 10) {
        const int t = threadIdx.x >> 1;
        uint32_t v = 0;
        if (threadIdx.x & 1) {
            v = (uint8_t)p.mat[t * 5 + 4];
        } else {
            for (int c = 3; c >= 0; --c) v = (v << 8) | (uint8_t)p.mat[t * 5 + c];
        }
        rowtab[threadIdx.x] = v;
    }
    __syncthreads();
    uint32_t *qlds = reinterpret_cast<uint32_t *>(nw_lds + 64);
    uint32_t *zreg = zbase + (size_t)blockIdx.x * zrows * ncw * 64;
    __shared__ uint32_t idx[64];
    __shared__ int item_sh;
    int it = blockIdx.x;
    for (;;) {
        if (persistent) {
            if (threadIdx.x == 0) item_sh = (int)atomicAdd(counter, 1u);
            __syncthreads();
            it = item_sh;
        }
        if (it >= n_items) break;
        const WorkItem wi = items[it];
        if (threadIdx.x < wi.count) idx[threadIdx.x] = order[wi.start + threadIdx.x];
        __syncthreads();
        nw_wave_reg<K>(d, p, idx, (int)wi.count, zreg, zrows, ncw, rowtab, qlds, qw8);
        __syncthreads();
        if (!persistent) it += gridDim.x;
    }
}

// ---------------------------------------------------------------------------
// Two jobs per lane: the register-state form with int16 pair arithmetic.
//
// Lane l fills jobs idx[l] (half 0 of every pair) and idx[l + 64] (half 1).
// H[k] and E[k] hold column beg + k of both jobs' current rows as int16 pairs
// (the frame of nw_wave_reg, moved per half), so every step of the cell --
// the diagonal add, the E and F updates and the maxima -- is one
// v_pk_*_i16 for both jobs.  The 16-bit tier's bound keeps every value in
// [-16384 - 8000, 8000], so the pair differences below do not wrap.
//
// Scores: per job and 8-cell block, four byte-permute lookups of the target
// base's matrix row put the scores of cells (0, 4), (1, 5), (2, 6), (3, 7) in
// bytes 1 and 3; one more permute per cell joins both jobs' scores as a
// sign-extended int16 pair (selectors 8-11 replicate the sign of bytes 1, 3,
// 5, 7).
//
// Flags: the four flag bits of the cell (those of nw_traceback) are the sign
// bits of four pair differences; a permute turns the signs of two of them into
// byte masks (flag a of job 0, flag b of job 0, flag a of job 1, flag b of
// job 1), and bit c of each byte is kept for cell c.
// Per row and block each job stores one word: byte f holds flag f of its 8
// cells.  The words of a wave are [row][block][128 job slots].
typedef short nws2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ nws2 nw_s2(uint32_t v) { return __builtin_bit_cast(nws2, v); }
__device__ __forceinline__ uint32_t nw_u(nws2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ uint32_t nw_pair(int a, int b)
{
    return ((uint32_t)a & 0xFFFFu) | ((uint32_t)b << 16);
}
__device__ __forceinline__ uint32_t nw_sel(uint32_t m, uint32_t x, uint32_t y) { return y ^ ((x ^ y) & m); }

// 4 flag bits of cell c (0..7) of a packed block word, in nw_traceback's order
__device__ __forceinline__ uint32_t nw_pk_flags(uint32_t z, int c)
{
    return ((((z >> c) & 0x01010101u) * 0x01020408u) >> 24) & 0xFu;
}

template <int K>
__device__ void nw_wave_pk(const NwDev &d, const NwParams &p, const uint32_t *idx, int count,
                           uint32_t *zreg, int zrows, int ncw, const uint32_t *rowtab,
                           uint32_t *qlds, int qw8)
{
    constexpr int NEG = NwSt<NW_I16>::kNeg;
    const int lane = threadIdx.x & 63;
    uint32_t *ql = qlds + lane;
    int qlen[2], tlen[2], wj[2];
    bool ne[2];
    uint32_t tw0[2], tn0[2];
    int cm = 0, wm = 0, tm = 0;
#pragma unroll
    for (int J = 0; J < 2; ++J) {
        const int slot = lane + 64 * J;
        const bool valid = slot < count;
        const int job = valid ? (int)idx[slot] : 0;
        qlen[J] = valid ? d.q_len[job] : 0;
        tlen[J] = valid ? d.t_len[job] : 0;
        ne[J] = valid && qlen[J] > 0 && tlen[J] > 0;           // :29-31
        int w = valid ? min(d.w[job], 1 << 20) : 0;
        w = max(w, abs(qlen[J] - tlen[J]));                    // :47
        wj[J] = w;
        const int ncol = min(qlen[J], 2 * w + 1);              // :48
        tw0[J] = valid ? d.t_word[job] : 0;
        tn0[J] = valid ? d.tn_word[job] : kNoN;
        const uint32_t qw0 = valid ? d.q_word[job] : 0;
        const uint32_t qn0 = valid ? d.qn_word[job] : kNoN;
        if (ne[J]) {
            cm = max(cm, ncol);
            wm = max(wm, w);
            tm = max(tm, tlen[J]);
        }
        // query J at words J * qw8 .. J * qw8 + qw8 - 1, 4-bit codes
        for (int b = 0; b < qw8; ++b) {
            uint32_t x = 0;
            if (ne[J] && b * 8 < qlen[J]) {
                x = (d.seq[qw0 + (b >> 1)] >> ((b & 1) * 16)) & 0xFFFFu;
                x = (x | (x << 8)) & 0x00FF00FFu;
                x = (x | (x << 4)) & 0x0F0F0F0Fu;
                x = (x | (x << 2)) & 0x33333333u;
                if (qn0 != kNoN) {
                    const uint32_t nb = (d.nmask[qn0 + (b >> 2)] >> ((b & 3) * 8)) & 0xFFu;
                    for (int k = 0; k < 8; ++k)
                        if ((nb >> k) & 1u) x = (x & ~(0xFu << (4 * k))) | (4u << (4 * k));
                }
            }
            ql[(J * qw8 + b) * 64] = x;
        }
    }
    // A row reads words beg/8 .. beg/8 + ceil(cmax/8) of its job's area, at
    // most J * qw8 + qw8 - 1 + K/8; cells those words feed beyond the band are
    // not kept, so job 0's reads may run into job 1's words.
    for (int b = 2 * qw8; b < 2 * qw8 + K / 8; ++b) ql[b * 64] = 0;
    const int cmax = wave_max_nw(cm);
    const int wmax = wave_max_nw(wm);
    const int tmax = wave_max_nw(tm);

    // First row (:75-84) in the frame of beg = 0: H[s] is column s.
    uint32_t H[K + 2], E[K + 2];
    {
        const int ie0 = min(wj[0], qlen[0]), ie1 = min(wj[1], qlen[1]);
#pragma unroll
        for (int s = 0; s < K + 2; ++s) {
            const int h0 = s == 0 ? 0 : (s <= ie0 ? -(p.o_ins + p.e_ins * s) : NEG);
            const int h1 = s == 0 ? 0 : (s <= ie1 ? -(p.o_ins + p.e_ins * s) : NEG);
            H[s] = nw_pair(h0, h1);
            E[s] = nw_pair(NEG, NEG);
        }
    }
    // score = h[qlen] (:144): column qlen of the last row is H[qlen - beg_last]
    const int sidx0 = qlen[0] - max(0, tlen[0] - 1 - wj[0]);
    const int sidx1 = qlen[1] - max(0, tlen[1] - 1 - wj[1]);
    uint32_t score = 0;
    const nws2 OED = {(short)p.oe_del, (short)p.oe_del}, ED = {(short)p.e_del, (short)p.e_del};
    const nws2 OEI = {(short)p.oe_ins, (short)p.oe_ins}, EI = {(short)p.e_ins, (short)p.e_ins};
    uint32_t *zl = zreg + lane;                  // word (i, cw, J) at zl[(i*ncw+cw)*128 + 64 J]
    uint32_t tw[2] = {0, 0}, tn[2] = {0, 0};
    uint32_t ncells = 0;
    for (int i = 0; i < tmax; ++i) {                           // :89
        uint32_t tabl[2], tabh[2], qlo[2], qsh[2];
        const uint32_t *qrow[2];
        int width[2], hl[2];
        bool act[2], adv[2];
#pragma unroll
        for (int J = 0; J < 2; ++J) {
            act[J] = ne[J] && i < tlen[J];
            if ((i & 15) == 0 && act[J]) tw[J] = d.seq[tw0[J] + (i >> 4)];
            int t = (tw[J] >> ((i & 15) * 2)) & 3;
            if (tn0[J] != kNoN) {
                if ((i & 31) == 0 && act[J]) tn[J] = d.nmask[tn0[J] + (i >> 5)];
                if ((tn[J] >> (i & 31)) & 1) t = 4;
            }
            tabl[J] = rowtab[2 * t];
            tabh[J] = rowtab[2 * t + 1];
            const int beg = max(0, i - wj[J]);                 // :90-91
            const int end = min(qlen[J], i + wj[J] + 1);
            width[J] = act[J] ? end - beg : 0;
            ncells += (uint32_t)width[J];
            adv[J] = act[J] && i > wj[J];
            hl[J] = beg == 0 ? 0 - (p.oe_del + p.e_del * i) : NEG;   // :96-104
            const int qb = act[J] ? beg : 0;
            qrow[J] = ql + (J * qw8 + (qb >> 3)) * 64;
            qsh[J] = (uint32_t)(qb & 7) * 4;
            qlo[J] = qrow[J][0];
        }
        // blocks: those holding every active band and the cell after it
        uint32_t wlo = min(act[0] ? (uint32_t)width[0] : 0xFFFFFFFFu,
                           act[1] ? (uint32_t)width[1] : 0xFFFFFFFFu);
        uint32_t whi = min(act[0] ? ~(uint32_t)width[0] : 0xFFFFFFFFu,
                           act[1] ? ~(uint32_t)width[1] : 0xFFFFFFFFu);
        wave_min2_u32(wlo, whi);
        const int wmin = (int)wlo, wmaxr = (int)~whi;
        const int nblk = min((wmaxr + 8) >> 3, (cmax + 7) >> 3);
        const int nfull = i > wmax ? min(wmin, cmax) >> 3 : 0;
        const uint32_t advM = (adv[0] ? 0xFFFFu : 0u) | (adv[1] ? 0xFFFF0000u : 0u);
        const uint32_t actM = (act[0] ? 0xFFFFu : 0u) | (act[1] ? 0xFFFF0000u : 0u);
        const uint32_t wpair = nw_pair(width[0], width[1]);
        uint32_t h_diag = nw_sel(advM, H[1], H[0]);
        uint32_t h_left = nw_pair(hl[0], hl[1]);
        H[0] = nw_sel(actM, h_left, H[0]);                     // column beg's e is never read again
        uint32_t f = nw_pair(NEG, NEG);
#pragma unroll
        for (int cb = 0; cb < K / 8; ++cb) {                   // :108-141
            if (cb < nblk) {                                   // a guard: see nw_wave_reg
                uint32_t lk[2][4];
#pragma unroll
                for (int J = 0; J < 2; ++J) {
                    const uint32_t qhi = qrow[J][(cb + 1) * 64];
                    const uint32_t codes = __builtin_amdgcn_alignbit(qhi, qlo[J], qsh[J]);
                    qlo[J] = qhi;
                    const uint32_t xe = codes & 0x0F0F0F0Fu;           // cells 0, 2, 4, 6
                    const uint32_t xo = (codes >> 4) & 0x0F0F0F0Fu;    // cells 1, 3, 5, 7
                    lk[J][0] = __builtin_amdgcn_perm(tabh[J], tabl[J], xe << 8);   // bytes 1, 3: cells 0, 4
                    lk[J][1] = __builtin_amdgcn_perm(tabh[J], tabl[J], xo << 8);   // 1, 5
                    lk[J][2] = __builtin_amdgcn_perm(tabh[J], tabl[J], xe);        // 2, 6
                    lk[J][3] = __builtin_amdgcn_perm(tabh[J], tabl[J], xo);        // 3, 7
                }
                uint32_t accP = 0, accQ = 0;
                auto cell = [&](auto full_t, int k) {
                    constexpr bool full = decltype(full_t)::value;
                    const int kk = cb * 8 + k;
                    const nws2 sc = nw_s2(__builtin_amdgcn_perm(lk[1][k & 3], lk[0][k & 3],
                                                                k < 4 ? 0x0A050801u : 0x0B070903u));
                    const uint32_t vh = full ? H[kk + 2] : nw_sel(advM, H[kk + 2], H[kk + 1]);
                    const uint32_t ve = full ? E[kk + 2] : nw_sel(advM, E[kk + 2], E[kk + 1]);
                    const nws2 hd = nw_s2(vh);
                    const nws2 h_cur = nw_s2(h_diag) + sc;                     // :110
                    const nws2 e_open = hd - OED, e_ext = nw_s2(ve) - ED;      // :116-120
                    const nws2 ev = __builtin_elementwise_max(e_open, e_ext);
                    const nws2 f_open = nw_s2(h_left) - OEI, f_ext = nw_s2(f) - EI;   // :121-127
                    const nws2 fv = __builtin_elementwise_max(f_open, f_ext);
                    const nws2 h1 = __builtin_elementwise_max(ev, h_cur);      // :129-134
                    const nws2 hn = __builtin_elementwise_max(fv, h1);
                    // flags (E opened, F opened, H from E, H from F): byte masks
                    // from the signs of the differences, [f0 or f2, f1 or f3] of
                    // job 0 then of job 1, one bit per cell
                    const uint32_t P = __builtin_amdgcn_perm(nw_u(f_ext - f_open), nw_u(e_ext - e_open), 0x0B090A08u);
                    const uint32_t Q = __builtin_amdgcn_perm(nw_u(h1 - fv), nw_u(h_cur - ev), 0x0B090A08u);
                    const uint32_t bit = 0x01010101u << k;
                    if constexpr (full) {
                        accP |= P & bit;
                        accQ |= Q & bit;
                        H[kk + 1] = nw_u(hn);
                        E[kk + 1] = nw_u(ev);
                    } else {
                        // halves whose band ends before this cell keep what they read
                        const uint32_t inM = nw_u((nws2{(short)kk, (short)kk} - nw_s2(wpair)) >> nws2{15, 15});
                        const uint32_t m = inM & bit;
                        accP |= P & m;
                        accQ |= Q & m;
                        H[kk + 1] = nw_sel(inM, nw_u(hn), vh);
                        E[kk + 1] = nw_sel(inM, nw_u(ev), ve);
                    }
                    // pinned where computed (see nw_wave_reg)
                    asm volatile("" : "+v"(H[kk + 1]));
                    asm volatile("" : "+v"(E[kk + 1]));
                    h_diag = vh;
                    // the next cell's left neighbour; in a half whose band has
                    // ended every later cell of the row is discarded as well
                    h_left = H[kk + 1];
                    f = nw_u(fv);
                };
                if (cb < nfull) {
#pragma unroll
                    for (int k = 0; k < 8; ++k) cell(std::true_type{}, k);
                } else {
#pragma unroll
                    for (int k = 0; k < 8; ++k) cell(std::false_type{}, k);
                }
                uint32_t *zw = zl + ((size_t)i * ncw + cb) * 128;      // i < tmax <= zrows
                zw[0] = __builtin_amdgcn_perm(accQ, accP, 0x05040100u);
                zw[64] = __builtin_amdgcn_perm(accQ, accP, 0x07060302u);
            }
        }
        // jobs on their last row take h[qlen] before a later unmasked block
        // may overwrite their state: a select tree over the bits of the index
        const bool end0 = act[0] && i == tlen[0] - 1, end1 = act[1] && i == tlen[1] - 1;
        if (__builtin_amdgcn_ballot_w64(end0 || end1)) {
            const int s0 = end0 ? sidx0 : 0, s1 = end1 ? sidx1 : 0;
            uint32_t T[K + 2];
#pragma unroll
            for (int s = 0; s < K + 2; ++s) T[s] = H[s];
#pragma unroll
            for (int b = 0; (1 << b) < K + 2; ++b) {
                const uint32_t mb = (((s0 >> b) & 1) ? 0xFFFFu : 0u) | (((s1 >> b) & 1) ? 0xFFFF0000u : 0u);
#pragma unroll
                for (int s = 0; s + (1 << b) < K + 2; s += 2 << b) T[s] = nw_sel(mb, T[s + (1 << b)], T[s]);
            }
            const uint32_t em = (end0 ? 0xFFFFu : 0u) | (end1 ? 0xFFFF0000u : 0u);
            score = nw_sel(em, T[0], score);
        }
    }
    if (d.cells) {
        unsigned long long c = ncells;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
        if (lane == 0) atomicAdd(d.cells, c);
    }
#pragma unroll
    for (int J = 0; J < 2; ++J) {
        const int slot = lane + 64 * J;
        if (slot >= count) continue;
        const int job = (int)idx[slot];
        if (!ne[J]) {
            d.score[job] = 0;
            d.n_cigar[job] = 0;
            d.flags[job] = 0;
            continue;
        }
        d.score[job] = (int)(int16_t)(score >> (16 * J));
        const int ncol = min(qlen[J], 2 * wj[J] + 1);
        const uint32_t *zj = zl + 64 * J;
        nw_traceback(d, job, qlen[J], tlen[J], wj[J], ncol, [&](int i, int c) {
            return nw_pk_flags(zj[((size_t)i * ncw + (c >> 3)) * 128], c & 7);
        });
    }
}

// The two-jobs-per-lane form: work items of up to 128 jobs; LDS holds the
// scoring table and both staged queries.
size_t nw_pk_lds_bytes(int K, int qw8) { return 64 + (size_t)(2 * qw8 + K / 8) * 256; }

template <int K>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(K <= 32 ? 3 : 2)))
void nw_pk_kernel(NwDev d, NwParams p, const uint32_t *order, const WorkItem *items, int n_items,
                  uint32_t *counter, uint32_t *zbase, int zrows, int ncw, int persistent, int qw8)
{
    extern __shared__ __align__(16) unsigned char nw_lds[];
    uint32_t *rowtab = reinterpret_cast<uint32_t *>(nw_lds);
    if (threadIdx.x < 10) {
        const int t = threadIdx.x >> 1;
        uint32_t v = 0;
        if (threadIdx.x & 1) {
            v = (uint8_t)p.mat[t * 5 + 4];
        } else {
            for (int c = 3; c >= 0; --c) v = (v << 8) | (uint8_t)p.mat[t * 5 + c];
        }
        rowtab[threadIdx.x] = v;
    }
    __syncthreads();
    uint32_t *qlds = reinterpret_cast<uint32_t *>(nw_lds + 64);
    uint32_t *zreg = zbase + (size_t)blockIdx.x * zrows * ncw * 128;
    __shared__ uint32_t idx[128];
    __shared__ int item_sh;
    int it = blockIdx.x;
    for (;;) {
        if (persistent) {
            if (threadIdx.x == 0) item_sh = (int)atomicAdd(counter, 1u);
            __syncthreads();
            it = item_sh;
        }
        if (it >= n_items) break;
        const WorkItem wi = items[it];
        if (threadIdx.x < wi.count) idx[threadIdx.x] = order[wi.start + threadIdx.x];
        if (threadIdx.x + 64 < wi.count) idx[threadIdx.x + 64] = order[wi.start + threadIdx.x + 64];
        __syncthreads();
        nw_wave_pk<K>(d, p, idx, (int)wi.count, zreg, zrows, ncw, rowtab, qlds, qw8);
        __syncthreads();
        if (!persistent) it += gridDim.x;
    }
}

// ---------------------------------------------------------------------------
// One wave per job ("wide" form), for jobs whose band is too wide to leave to
// one lane: the bins holding few of them and the band-doubling retry passes,
// whose time a lane-per-job kernel spends on its single slowest job.
//
// Lane l holds the columns j = 64k + l (chunk k < KC, query length <= 64 KC)
// of the row state H(i, j), E(i, j) in registers, int32 with the reference's
// Int32.min/2 sentinel (GlobalAligner.swift:52).  A row runs the chunks that
// meet its band [beg, end) (:90-91).  The diagonal input of column j is
// column j - 1 of the previous row (a one-lane DPP shift, the chunk's lane 0
// taking the previous chunk's lane 63); E is lane-local.  F, the row's only
// left-to-right dependence, is the exclusive prefix maximum
//   F(j) = max(C0 - e_ins (j - 1), max_{beg <= k < j} (G(k) - oe_ins - e_ins (j - 1 - k)))
// with G = max(H_diag + s, E): since o_ins >= 0, H(j-1) - oe_ins <= F(j-1) - e_ins
// whenever F(j-1) > G(j-1), so the F-free G carries the recurrence (:121-127);
// C0 folds the row-start h[beg] and the initial f = Int32.min/2.  It is one
// max-scan of G(k) - oe_ins + e_ins k (fused v_max_i32_dpp steps).  The four
// flag bits of every cell are the serial loop's (:114-136, in nw_traceback's
// encoding); they go to LDS
// (bytes, [row][j - beg] as the reference's z) when tlen * ncol fits, else to
// the wave's HBM region, and lane 0 walks the traceback from there.
typedef __attribute__((address_space(3))) uint8_t nw_lds_u8;
typedef __attribute__((address_space(1))) uint8_t nw_glb_u8;

template <int KC, typename ZB>
__device__ void nw_wide_job(const NwDev &d, const NwParams &p, int job, ZB *z)
{
    constexpr int NEG = INT32_MIN / 2;                         // :52
    constexpr int LOW = NEG - (1 << 28);                       // below every scanned term
    const int lane = threadIdx.x & 63;
    const int qlen = d.q_len[job], tlen = d.t_len[job];
    if (qlen <= 0 || tlen <= 0) {                              // :29-31
        if (lane == 0) { d.score[job] = 0; d.n_cigar[job] = 0; d.flags[job] = 0; }
        return;
    }
    const int w = max(min(d.w[job], 1 << 20), abs(qlen - tlen));   // :47
    const int ncol = min(qlen, 2 * w + 1);                     // :48
    const uint32_t qw0 = d.q_word[job], tw0 = d.t_word[job];
    const uint32_t qn0 = d.qn_word[job], tn0 = d.tn_word[job];
    // Per chunk: this lane's query code (N = 4) and the first row (:75-84):
    // h[j+1] = -(o_ins + e_ins (j+1)) for j + 1 <= min(w, qlen), e = sentinel.
    // Scores of this lane's query base against target bases 0-3 (signed
    // bytes of spk) and N (s4): mat[t * 5 + q] (GlobalAligner.swift:63-70).
    uint32_t colw[5];
#pragma unroll
    for (int q = 0; q < 5; ++q)
        colw[q] = (uint32_t)(uint8_t)p.mat[q] | ((uint32_t)(uint8_t)p.mat[5 + q] << 8) |
                  ((uint32_t)(uint8_t)p.mat[10 + q] << 16) | ((uint32_t)(uint8_t)p.mat[15 + q] << 24);
    int spk[KC], s4[KC], Hc[KC], Ec[KC];
    const int init_end = min(w, qlen);
#pragma unroll
    for (int k = 0; k < KC; ++k) {
        const int j = 64 * k + lane;
        int c = 0;
        if (j < qlen) {
            c = (int)((d.seq[qw0 + (j >> 4)] >> ((j & 15) * 2)) & 3u);
            if (qn0 != kNoN && ((d.nmask[qn0 + (j >> 5)] >> (j & 31)) & 1u)) c = 4;
        }
        spk[k] = (int)(c == 0 ? colw[0] : c == 1 ? colw[1] : c == 2 ? colw[2] : c == 3 ? colw[3] : colw[4]);
        s4[k] = c == 0 ? p.mat[20] : c == 1 ? p.mat[21] : c == 2 ? p.mat[22] : c == 3 ? p.mat[23] : p.mat[24];
        Hc[k] = j + 1 <= init_end ? -(p.o_ins + p.e_ins * (j + 1)) : NEG;
        Ec[k] = NEG;
    }
    int hvirt = 0;                                             // h[0] (:81)
    unsigned long long ncells = 0;
    uint32_t tw = 0, tn = 0;
    for (int i = 0; i < tlen; ++i) {                           // :89
        if ((i & 15) == 0) tw = d.seq[tw0 + (i >> 4)];
        int t = (int)((tw >> ((i & 15) * 2)) & 3u);
        if (tn0 != kNoN) {
            if ((i & 31) == 0) tn = d.nmask[tn0 + (i >> 5)];
            if ((tn >> (i & 31)) & 1u) t = 4;
        }
        const int beg = max(0, i - w), end = min(qlen, i + w + 1);   // :90-91
        ncells += (unsigned long long)(end - beg);
        const int hb = beg == 0 ? 0 - (p.oe_del + p.e_del * i) : NEG;   // :97-103
        int Cin = max(hb - p.oe_ins + p.e_ins * (beg - 1), NEG + p.e_ins * (beg - 2));
        int hcarry = hb, fcarry = NEG;                        // H, F of the column left of a chunk
        const int k0 = beg >> 6, k1 = (end - 1) >> 6;
        int prev63 = hvirt;                                    // old H of column 64 k0 - 1
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            if (k < k0 || k > k1) continue;                    // wave-uniform
            if (k > 0 && k == k0) prev63 = __builtin_amdgcn_readlane(Hc[k - 1], 63);
            const int j = 64 * k + lane;
            const bool act = j >= beg && j < end;
            const int hold = Hc[k], eold = Ec[k];
            // H(i-1, j-1): lane l - 1, lane 0 from the chunk to the left
            int dsrc = __builtin_amdgcn_update_dpp(0, hold, 0x138, 0xF, 0xF, false);   // wave_shr:1
            if (lane == 0) dsrc = prev63;
            prev63 = __builtin_amdgcn_readlane(hold, 63);
            const int sc = t < 4 ? __builtin_amdgcn_sbfe(spk[k], 8 * t, 8) : s4[k];
            const int hcur = dsrc + sc;                           // :110
            const int e_open = hold - p.oe_del, e_ext = eold - p.e_del;   // :116-120
            const uint32_t e_flag = e_open > e_ext ? 1u : 0u;          // E opened
            const int ev = e_open > e_ext ? e_open : e_ext;
            const int g = max(hcur, ev);
            // exclusive prefix max of G(k) - oe_ins + e_ins k over the band
            int x = act ? g - p.oe_ins + p.e_ins * j : LOW;
            x = max(x, __builtin_amdgcn_update_dpp(LOW, x, 0x111, 0xF, 0xF, false));   // row_shr:1
            x = max(x, __builtin_amdgcn_update_dpp(LOW, x, 0x112, 0xF, 0xF, false));   // row_shr:2
            x = max(x, __builtin_amdgcn_update_dpp(LOW, x, 0x114, 0xF, 0xF, false));   // row_shr:4
            x = max(x, __builtin_amdgcn_update_dpp(LOW, x, 0x118, 0xF, 0xF, false));   // row_shr:8
            x = max(x, __builtin_amdgcn_update_dpp(LOW, x, 0x142, 0xA, 0xF, false));   // row_bcast:15
            x = max(x, __builtin_amdgcn_update_dpp(LOW, x, 0x143, 0xC, 0xF, false));   // row_bcast:31
            int px = __builtin_amdgcn_update_dpp(LOW, x, 0x138, 0xF, 0xF, false);      // wave_shr:1
            px = max(px, Cin);
            Cin = max(Cin, __builtin_amdgcn_readlane(x, 63));
            const int fv = px - p.e_ins * (j - 1);                 // F(i, j)
            const uint32_t from_e = ev > hcur ? 4u : 0u;           // :129-134
            const int h1 = ev > hcur ? ev : hcur;
            const uint32_t from_f = fv > h1 ? 8u : 0u;
            const int hv = fv > h1 ? fv : h1;
            // f flag (:122-126): H(i, j-1) - oe_ins against F(i, j-1) - e_ins
            int hl = __builtin_amdgcn_update_dpp(0, hv, 0x138, 0xF, 0xF, false);
            int fl = __builtin_amdgcn_update_dpp(0, fv, 0x138, 0xF, 0xF, false);
            if (lane == 0) { hl = hcarry; fl = fcarry; }
            if (j == beg) { hl = hb; fl = NEG; }
            const uint32_t f_flag = (hl - p.oe_ins) > (fl - p.e_ins) ? 2u : 0u;   // F opened
            hcarry = __builtin_amdgcn_readlane(hv, 63);
            fcarry = __builtin_amdgcn_readlane(fv, 63);
            if (act) {
                Hc[k] = hv;                                    // :138
                Ec[k] = ev;
                z[(size_t)i * ncol + (j - beg)] = (uint8_t)(e_flag | f_flag | from_e | from_f);
            }
        }
        if (beg == 0) hvirt = hb;
    }
    if (d.cells && lane == 0) atomicAdd(d.cells, ncells);
    // score = h[qlen] (:144): column qlen - 1 of the last row
    const int kq = (qlen - 1) >> 6;
    int score = 0;
#pragma unroll
    for (int k = 0; k < KC; ++k)
        if (k == kq) score = __builtin_amdgcn_readlane(Hc[k], (qlen - 1) & 63);
    __syncthreads();                                           // flags visible to every lane
    if (lane == 0) d.score[job] = score;
    // every lane walks the traceback (its state is wave-uniform), lane 0 writes
    nw_traceback(d, job, qlen, tlen, w, ncol, [&](int i, int c) {
        return (uint32_t)__builtin_amdgcn_readfirstlane((int)z[(size_t)i * ncol + c]);
    }, lane == 0);
}

// Jobs jobs[0..n) of one pass, one wave each, taken from an atomic counter.
// zmax: bytes of one job's flag matrix at most (tlen * ncol); in LDS when
// zlds, else at zg + blockIdx.x * zmax.
template <int KC>
__global__ __launch_bounds__(64) void nw_wide_kernel(NwDev d, NwParams p, const uint32_t *jobs,
                                                     int n, uint32_t *counter, uint8_t *zg,
                                                     size_t zmax, int zlds)
{
    extern __shared__ __align__(16) unsigned char nw_lds[];
    // the flag matrix in LDS or in this block's HBM region, addressed in its
    // own address space (ds_ / global_ accesses rather than flat ones)
    nw_lds_u8 *zl = (nw_lds_u8 *)nw_lds;
    nw_glb_u8 *zh = (nw_glb_u8 *)(zg + (size_t)blockIdx.x * zmax);
    __shared__ int item_sh;
    for (;;) {
        if (threadIdx.x == 0) item_sh = (int)atomicAdd(counter, 1u);
        __syncthreads();
        const int it = item_sh;
        __syncthreads();
        if (it >= n) break;
        if (zlds) nw_wide_job<KC>(d, p, (int)jobs[it], zl);
        else nw_wide_job<KC>(d, p, (int)jobs[it], zh);
        __syncthreads();
    }
}

// The CPU retry loop's test (CIGARGenerator.swift:88-96) on the device: of
// the m jobs list[0..m), those whose score is below their true score are
// appended to next[0..*count) (in no fixed order), with their band doubled
// into w_next (w < 2^20: the cap of the band the kernels accept).
__global__ void nw_select_kernel(const uint32_t *list, int m, const int32_t *score,
                                 const int32_t *ts, const int32_t *w_cur, int32_t *w_next,
                                 uint32_t *next, uint32_t *count)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const uint32_t j = list[k];
    if (score[j] < ts[j]) {
        const uint32_t slot = atomicAdd(count, 1u);
        next[slot] = j;
        const int32_t wv = w_cur[j];
        w_next[j] = wv < (1 << 20) ? wv * 2 : wv;
    }
}

hipError_t launch_nw_select(const uint32_t *list, int m, const int32_t *score, const int32_t *ts,
                            const int32_t *w_cur, int32_t *w_next, uint32_t *next,
                            uint32_t *count, hipStream_t s)
{
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(nw_select_kernel, dim3((m + 255) / 256), dim3(256), 0, s, list, m, score,
                       ts, w_cur, w_next, next, count);
    return hipGetLastError();
}

int nw_wide_kc(int qmax) { return qmax <= 128 ? 2 : qmax <= 256 ? 4 : qmax <= 512 ? 8 : 0; }

hipError_t launch_nw_wide(const NwDev &d, const NwParams &p, const uint32_t *jobs, int n,
                          uint32_t *counter, int qmax, uint8_t *zg, size_t zmax, bool zlds,
                          int grid, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int kc = nw_wide_kc(qmax);
    const size_t lds = zlds ? zmax : 0;
    if (kc == 2) hipLaunchKernelGGL((nw_wide_kernel<2>), dim3(grid), dim3(64), lds, s, d, p, jobs, n, counter, zg, zmax, zlds ? 1 : 0);
    else if (kc == 4) hipLaunchKernelGGL((nw_wide_kernel<4>), dim3(grid), dim3(64), lds, s, d, p, jobs, n, counter, zg, zmax, zlds ? 1 : 0);
    else if (kc == 8) hipLaunchKernelGGL((nw_wide_kernel<8>), dim3(grid), dim3(64), lds, s, d, p, jobs, n, counter, zg, zmax, zlds ? 1 : 0);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

size_t nw_lds_bytes(int tier, bool global_state, int S, int qw8)
{
    const size_t w = tier == NW_I16 ? 4 : 8;
    return 32 + (size_t)qw8 * 256 + (global_state ? 0 : (size_t)S * 64 * w);
}

template <int T, bool G>
static hipError_t nw_t(const NwDev &d, const NwParams &p, const uint32_t *order,
                       const WorkItem *items, int n_items, uint32_t *counter, int S,
                       uint32_t *zbase, int zrows, int ncw, int grid, bool persistent,
                       void *gstate, int qw8, hipStream_t s)
{
    hipLaunchKernelGGL((nw_kernel<T, G>), dim3(grid), dim3(64), nw_lds_bytes(T, G, S, qw8), s, d,
                       p, order, items, n_items, counter, S, zbase, zrows, ncw,
                       persistent ? 1 : 0, gstate, qw8);
    return hipGetLastError();
}

int nw_reg_cols(int ncol_max)
{
    return ncol_max <= 32 ? 32 : ncol_max <= 64 ? 64 : ncol_max <= 96 ? 96 : ncol_max <= 128 ? 128 : 0;
}

size_t nw_reg_lds_bytes(int K, int qw8) { return 64 + (size_t)(qw8 + K / 8) * 256; }

template <int K>
static int nw_reg_occ(int qw8)
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, nw_reg_kernel<K>, 64,
                                                     nw_reg_lds_bytes(K, qw8)) != hipSuccess)
        return 0;
    return nb;
}

int nw_reg_blocks_per_cu(int K, int qw8)
{
    switch (K) {
    case 32: return nw_reg_occ<32>(qw8);
    case 64: return nw_reg_occ<64>(qw8);
    case 96: return nw_reg_occ<96>(qw8);
    case 128: return nw_reg_occ<128>(qw8);
    default: return 0;
    }
}

template <int K>
static hipError_t nw_reg_t(const NwDev &d, const NwParams &p, const uint32_t *order,
                           const WorkItem *items, int n_items, uint32_t *counter, uint32_t *zbase,
                           int zrows, int ncw, int grid, bool persistent, int qw8, hipStream_t s)
{
    hipLaunchKernelGGL((nw_reg_kernel<K>), dim3(grid), dim3(64), nw_reg_lds_bytes(K, qw8), s, d, p,
                       order, items, n_items, counter, zbase, zrows, ncw, persistent ? 1 : 0, qw8);
    return hipGetLastError();
}

template <int K>
static int nw_pk_occ(int qw8)
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, nw_pk_kernel<K>, 64,
                                                     nw_pk_lds_bytes(K, qw8)) != hipSuccess)
        return 0;
    return nb;
}

int nw_pk_blocks_per_cu(int K, int qw8)
{
    switch (K) {
    case 32: return nw_pk_occ<32>(qw8);
    case 64: return nw_pk_occ<64>(qw8);
    case 96: return nw_pk_occ<96>(qw8);
    default: return 0;
    }
}

template <int K>
static hipError_t nw_pk_t(const NwDev &d, const NwParams &p, const uint32_t *order,
                          const WorkItem *items, int n_items, uint32_t *counter, uint32_t *zbase,
                          int zrows, int ncw, int grid, bool persistent, int qw8, hipStream_t s)
{
    hipLaunchKernelGGL((nw_pk_kernel<K>), dim3(grid), dim3(64), nw_pk_lds_bytes(K, qw8), s, d, p,
                       order, items, n_items, counter, zbase, zrows, ncw, persistent ? 1 : 0, qw8);
    return hipGetLastError();
}

hipError_t launch_nw(int tier, bool global_state, int reg, const NwDev &d, const NwParams &p,
                     const uint32_t *order, const WorkItem *items, int n_items,
                     uint32_t *counter, int S, uint32_t *zbase, int zrows, int ncw, int grid,
                     bool persistent, void *gstate, int qw8, hipStream_t s)
{
    if (n_items <= 0) return hipSuccess;
    if (reg < 0) {
        // two jobs per lane, K = -reg; blocks cover 8 * ncw >= every band
        if (tier != NW_I16 || -reg < 8 * ncw) return hipErrorInvalidValue;
        switch (-reg) {
        case 32: return nw_pk_t<32>(d, p, order, items, n_items, counter, zbase, zrows, ncw, grid, persistent, qw8, s);
        case 64: return nw_pk_t<64>(d, p, order, items, n_items, counter, zbase, zrows, ncw, grid, persistent, qw8, s);
        case 96: return nw_pk_t<96>(d, p, order, items, n_items, counter, zbase, zrows, ncw, grid, persistent, qw8, s);
        default: return hipErrorInvalidValue;
        }
    }
    if (reg) {
        // the kernel's blocks cover 8 * ncw >= every band of the launch
        if (tier != NW_I16 || reg < 8 * ncw) return hipErrorInvalidValue;
        switch (reg) {
        case 32: return nw_reg_t<32>(d, p, order, items, n_items, counter, zbase, zrows, ncw, grid, persistent, qw8, s);
        case 64: return nw_reg_t<64>(d, p, order, items, n_items, counter, zbase, zrows, ncw, grid, persistent, qw8, s);
        case 96: return nw_reg_t<96>(d, p, order, items, n_items, counter, zbase, zrows, ncw, grid, persistent, qw8, s);
        case 128: return nw_reg_t<128>(d, p, order, items, n_items, counter, zbase, zrows, ncw, grid, persistent, qw8, s);
        default: return hipErrorInvalidValue;
        }
    }
    switch (tier * 2 + (global_state ? 1 : 0)) {
    case 0: return nw_t<NW_I16, false>(d, p, order, items, n_items, counter, S, zbase, zrows, ncw, grid, persistent, gstate, qw8, s);
    case 1: return nw_t<NW_I16, true>(d, p, order, items, n_items, counter, S, zbase, zrows, ncw, grid, persistent, gstate, qw8, s);
    case 2: return nw_t<NW_I32, false>(d, p, order, items, n_items, counter, S, zbase, zrows, ncw, grid, persistent, gstate, qw8, s);
    default: return nw_t<NW_I32, true>(d, p, order, items, n_items, counter, S, zbase, zrows, ncw, grid, persistent, gstate, qw8, s);
    }
}

}  // namespace hipsw
