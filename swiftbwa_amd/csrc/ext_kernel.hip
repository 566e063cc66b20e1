// Banded seed-extension kernels for gfx950 (ksw_extend2 semantics).
//
// Semantics: BandedSWScalar.align, Sources/Alignment/BandedSWScalar.swift:21-191
// (w cap :52-62, first row :68-77, row loop :90-181, cells :112-143, band
// narrowing :175-180), restated per lane.
//
// Layout: one wave64 holds 64 independent DP matrices, one per lane, and sweeps
// them row by row in lock step (inter-task SIMD).  Each lane owns the whole
// row recurrence of its matrix, so the row-sequential parts of ksw_extend2 --
// the band [beg,end) recomputed from the previous row, stale entries beyond
// `end`, last-j row-max ties, z-drop and the rowMax==0 break -- are exact by
// construction.  The per-column state h[j]/e[j] of all 64 matrices lives in LDS
// as one word per (column, lane), column-major, so a column access by the wave
// touches 64 consecutive words (conflict-free).  Scores are held in the
// narrowest exact tier: two u8 (16-bit word), two u16 (32-bit word) or two i32.
// Queries and targets are read as 2-bit words from HBM (16 bases per load).
#include <type_traits>

#include "hipsw_internal.h"

namespace hipsw {

template <int TIER> struct St;
template <> struct St<TIER_U8> {
    using W = uint16_t;
    static __device__ __forceinline__ int h(W v) { return v & 0xFF; }
    static __device__ __forceinline__ int e(W v) { return v >> 8; }
    static __device__ __forceinline__ W pack(int hv, int ev) { return (W)(hv | (ev << 8)); }
    static __device__ __forceinline__ bool zero(W v) { return v == 0; }
};
template <> struct St<TIER_U16> {
    using W = uint32_t;
    static __device__ __forceinline__ int h(W v) { return (int)(v & 0xFFFFu); }
    static __device__ __forceinline__ int e(W v) { return (int)(v >> 16); }
    static __device__ __forceinline__ W pack(int hv, int ev) { return (W)hv | ((W)ev << 16); }
    static __device__ __forceinline__ bool zero(W v) { return v == 0; }
};
template <> struct St<TIER_I32> {
    using W = int2;
    static __device__ __forceinline__ int h(W v) { return v.x; }
    static __device__ __forceinline__ int e(W v) { return v.y; }
    static __device__ __forceinline__ W pack(int hv, int ev) { return make_int2(hv, ev); }
    static __device__ __forceinline__ bool zero(W v) { return v.x == 0 && v.y == 0; }
};

size_t ext_state_bytes(int tier)
{
    return tier == TIER_U8 ? 2 : tier == TIER_U16 ? 4 : 8;
}

__device__ __forceinline__ int wave_min(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_max(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// Generic form: any 5x5 matrix, N bases, every tier.  Extends the (<= 64) jobs
// idx[0..count) with one wave.  `col` points at this lane's column-0 state
// word; column j is at col[j * 64].
template <int TIER>
__device__ __forceinline__ void ext_wave_generic(const ExtDev &d, const ExtParams &p,
                                         const uint32_t *idx, int count,
                                         typename St<TIER>::W *col, const int8_t *smat)
{
    using S = St<TIER>;
    using W = typename S::W;
    const int lane = threadIdx.x & 63;
    const bool valid = lane < count;
    const int job = valid ? (int)idx[lane] : 0;
    const int qlen = valid ? d.q_len[job] : 0;
    const int tlen = valid ? d.t_len[job] : 0;
    const int h0 = valid ? d.h0[job] : 0;
    int w = valid ? d.w[job] : 0;
    const int eb = valid ? d.eb[job] : 0;
    const uint32_t qw0 = valid ? d.q_word[job] : 0;
    const uint32_t tw0 = valid ? d.t_word[job] : 0;
    const uint32_t qn0 = valid ? d.qn_word[job] : kNoN;
    const uint32_t tn0 = valid ? d.tn_word[job] : kNoN;
    const bool nonempty = valid && qlen > 0 && tlen > 0;     // :33-35
    bool alive = nonempty;

    if (alive) {                                               // w cap, :52-62
        int mi = (int)((double)(qlen * p.max_mat + eb - p.o_ins) / (double)p.e_ins + 1.0);
        w = min(w, max(1, mi));
        int md = (int)((double)(qlen * p.max_mat + eb - p.o_del) / (double)p.e_del + 1.0);
        w = min(w, max(1, md));
    }

    // First row (:68-77): h[0] = h0, h[j] = max(0, h0 - o_ins - e_ins*j), e = 0.
    const int qmax = wave_max(alive ? qlen : 0);
    for (int j = 0; j <= qmax; ++j) {
        if (alive && j <= qlen) {
            int hv = j == 0 ? h0 : max(0, h0 - p.o_ins - p.e_ins * j);
            col[j * 64] = S::pack(hv, 0);
        }
    }

    int beg = 0, end = qlen;                                   // :79-87
    int maxv = h0, max_i = -1, max_j = -1;
    int gsc = -1, gtle = -1, max_off = 0;
    uint32_t tw = 0, tn = 0;
    uint64_t ncells = 0;

    for (int i = 0;; ++i) {
        if (alive && i >= tlen) alive = false;
        if (__ballot(alive) == 0) break;
        if ((i & 15) == 0 && alive) tw = d.seq[tw0 + (i >> 4)];
        int t = (tw >> ((i & 15) * 2)) & 3;
        if (tn0 != kNoN) {
            if ((i & 31) == 0 && alive) tn = d.nmask[tn0 + (i >> 5)];
            if ((tn >> (i & 31)) & 1) t = 4;
        }
        const int t5 = t * 5;

        int f = 0, h1 = 0;
        uint32_t key = 0;                 // (rowMax << 16) | rowMaxJ, narrow tiers
        int row_max = 0, row_max_j = -1;  // i32 tier
        if (alive) {                                           // :95-107
            beg = max(beg, i - w);
            end = min(end, i + w + 1);
            end = min(end, qlen);
            h1 = beg == 0 ? max(0, h0 - (p.o_del + p.e_del * (i + 1))) : 0;
        }
        if (alive && end > beg) ncells += (uint64_t)(end - beg);
        const int jlo = wave_min(alive ? beg : 0x7fffffff);
        const int jhi = wave_max(alive ? end : 0);
        uint32_t qw = 0, qn = 0;
        for (int j = jlo; j < jhi; ++j) {                      // :112-143
            if (((j & 15) == 0 || j == jlo) && alive && j < qlen) qw = d.seq[qw0 + (j >> 4)];
            if (qn0 != kNoN && ((j & 31) == 0 || j == jlo) && alive && j < qlen)
                qn = d.nmask[qn0 + (j >> 5)];
            if (alive && j >= beg && j < end) {
                int q = (qw >> ((j & 15) * 2)) & 3;
                if (qn0 != kNoN && ((qn >> (j & 31)) & 1)) q = 4;
                const W wd = col[j * 64];
                const int M = S::h(wd), E = S::e(wd);
                const int diag = M != 0 ? M + (int)smat[t5 + q] : 0;
                const int H = max(max(diag, E), f);
                if (TIER == TIER_I32) {
                    row_max_j = row_max > H ? row_max_j : j;
                    row_max = max(row_max, H);
                } else {
                    key = max(key, ((uint32_t)H << 16) | (uint32_t)j);
                }
                const int ne = max(max(E - p.e_del, diag - p.oe_del), 0);
                f = max(max(f - p.e_ins, diag - p.oe_ins), 0);
                col[j * 64] = S::pack(h1, ne);
                h1 = H;
            }
        }
        if (alive) {
            if (TIER != TIER_I32) {
                row_max = (int)(key >> 16);
                row_max_j = (int)(key & 0xFFFFu);
            }
            col[end * 64] = S::pack(h1, 0);                    // :145-146
            if (end == qlen) {                                 // :149-152
                gtle = gsc > h1 ? gtle : i;
                gsc = max(gsc, h1);
            }
            if (row_max == 0) {                                // :155
                alive = false;
            } else if (row_max > maxv) {                       // :158-163
                maxv = row_max;
                max_i = i;
                max_j = row_max_j;
                max_off = max(max_off, abs(row_max_j - i));
            } else if (p.zdrop > 0) {                          // :164-172
                const int di = i - max_i, dj = row_max_j - max_j;
                const int drop = di > dj ? maxv - row_max - (di - dj) * p.e_del
                                         : maxv - row_max - (dj - di) * p.e_ins;
                if (drop > p.zdrop) alive = false;
            }
            if (alive) {                                       // :175-180
                int nb = beg;
                while (nb < end && S::zero(col[nb * 64])) ++nb;
                beg = nb;
                int ne = end;
                while (ne >= beg && S::zero(col[ne * 64])) --ne;
                end = min(ne + 2, qlen);
            }
        }
    }
    if (d.cells) {
        unsigned long long c = ncells;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
        if (lane == 0) atomicAdd(d.cells, (unsigned long long)c);
    }
    if (valid) {                                               // :183-190
        int32_t *o = d.out + (size_t)job * 6;
        if (nonempty) {
            o[0] = maxv; o[1] = max_j + 1; o[2] = max_i + 1;
            o[3] = gtle + 1; o[4] = gsc; o[5] = max_off;
        } else {
            o[0] = 0; o[1] = 0; o[2] = 0; o[3] = 0; o[4] = 0; o[5] = 0;
        }
    }
}


// Fast form for the common case -- no N base in any of the wave's jobs, the
// standard A/-B matrix, 8- or 16-bit tier.  Same recurrence as the generic
// form; what changes is how a row is swept:
//  - each lane's query is staged once per job in LDS as 2-bit words
//    (qlds[b*64 + lane] = bases 16b..16b+15);
//  - per 16-column block the match bits of the whole block are one XOR with
//    the row's target base replicated 16 times;
//  - the row range [jlo, jhi) is split into masked edges and an unmasked
//    middle of whole 16-column blocks where every live lane is inside its band
//    ([max beg, min end)); the middle is fully unrolled with no exec-mask
//    changes.  Lanes already finished compute there too, harmlessly: they only
//    touch their own LDS column state, which is re-initialised per job.
template <int TIER>
__device__ __forceinline__ void ext_wave_fast(const ExtDev &d, const ExtParams &p,
                                              const uint32_t *idx, int count,
                                              typename St<TIER>::W *col, uint32_t *qlds)
{
    using S = St<TIER>;
    using W = typename S::W;
    const int lane = threadIdx.x & 63;
    const bool valid = lane < count;
    const int job = valid ? (int)idx[lane] : 0;
    const int qlen = valid ? d.q_len[job] : 0;
    const int tlen = valid ? d.t_len[job] : 0;
    const int h0 = valid ? d.h0[job] : 0;
    int w = valid ? d.w[job] : 0;
    const int eb = valid ? d.eb[job] : 0;
    const uint32_t qw0 = valid ? d.q_word[job] : 0;
    const uint32_t tw0 = valid ? d.t_word[job] : 0;
    const bool nonempty = valid && qlen > 0 && tlen > 0;     // :33-35
    bool alive = nonempty;
    const int A = p.match_sc, MIS = p.mis_sc;
    const int e_del = p.e_del, e_ins = p.e_ins, oe_del = p.oe_del, oe_ins = p.oe_ins;

    if (alive) {                                               // w cap, :52-62
        int mi = (int)((double)(qlen * p.max_mat + eb - p.o_ins) / (double)e_ins + 1.0);
        w = min(w, max(1, mi));
        int md = (int)((double)(qlen * p.max_mat + eb - p.o_del) / (double)e_del + 1.0);
        w = min(w, max(1, md));
    }
    const int qmax = wave_max(alive ? qlen : 0);
    // stage the query's 2-bit words
    const int qblocks = (qmax + 15) >> 4;
    uint32_t *ql = qlds + lane;
    for (int b = 0; b < qblocks; ++b)
        ql[b * 64] = (alive && b * 16 < qlen) ? d.seq[qw0 + b] : 0u;
    // first row (:68-77)
    for (int j = 0; j <= qmax; ++j) {
        if (alive && j <= qlen) {
            const int hv = j == 0 ? h0 : max(0, h0 - p.o_ins - e_ins * j);
            col[j * 64] = S::pack(hv, 0);
        }
    }

    int beg = 0, end = qlen;                                   // :79-87
    int maxv = h0, max_i = -1, max_j = -1;
    int gsc = -1, gtle = -1, max_off = 0;
    uint32_t tw = 0;
    uint64_t ncells = 0;

    for (int i = 0;; ++i) {
        if (alive && i >= tlen) alive = false;
        if (__ballot(alive) == 0) break;
        if ((i & 15) == 0 && alive) tw = d.seq[tw0 + (i >> 4)];
        const uint32_t t = (tw >> ((i & 15) * 2)) & 3u;
        const uint32_t tpat = t * 0x55555555u;

        int f = 0, h1 = 0;
        uint32_t key = 0;                 // (rowMax << 16) | rowMaxJ, last j wins ties
        if (alive) {                                           // :95-107
            beg = max(beg, i - w);
            end = min(end, i + w + 1);
            end = min(end, qlen);
            h1 = beg == 0 ? max(0, h0 - (p.o_del + e_del * (i + 1))) : 0;
            if (end > beg) ncells += (uint64_t)(end - beg);
        }
        const int jlo = wave_min(alive ? beg : 0x7fffffff);
        const int jhi = wave_max(alive ? end : 0);
        const int mlo = (wave_max(alive ? beg : 0) + 15) & ~15;   // middle: whole blocks
        const int mhi = wave_min(alive ? end : 0x7fffffff) & ~15;
        const bool has_mid = mlo < mhi && mlo >= jlo && mhi <= jhi;
        const int e1 = has_mid ? mlo : jhi;                   // masked [jlo, e1)
        const int s2 = has_mid ? mhi : jhi;                   // masked [s2, jhi)

        // one cell with an explicit in-band test (edges of the row)
        auto masked_cell = [&](int j) {
            const bool inb = alive && (unsigned)(j - beg) < (unsigned)(end - beg);
            const uint32_t qw = ql[(j >> 4) * 64];
            const uint32_t q = (qw >> ((j & 15) * 2)) & 3u;
            const W wd = col[j * 64];
            const int M = S::h(wd), E = S::e(wd);
            const int sc = q == t ? A : MIS;
            const int diag = M != 0 ? M + sc : 0;
            const int H = max(max(diag, E), f);
            const uint32_t kk = ((uint32_t)H << 16) | (uint32_t)j;
            const int ne = max(max(E - e_del, diag - oe_del), 0);
            const int nf = max(max(f - e_ins, diag - oe_ins), 0);
            col[j * 64] = inb ? S::pack(h1, ne) : wd;
            key = inb ? max(key, kk) : key;
            f = inb ? nf : f;
            h1 = inb ? H : h1;
        };
        for (int j = jlo; j < e1; ++j) masked_cell(j);
        if (has_mid) {
            for (int b = mlo >> 4; b < (mhi >> 4); ++b) {
                const uint32_t x = ql[b * 64] ^ tpat;
                const uint32_t mb = ~(x | (x >> 1)) & 0x55555555u;   // bit 2k: base k matches
                W *cb = col + b * 16 * 64;
                const uint32_t jb = (uint32_t)(b * 16);
#pragma unroll
                for (int k = 0; k < 16; ++k) {
                    const W wd = cb[k * 64];
                    const int M = S::h(wd), E = S::e(wd);
                    const int sc = ((mb >> (2 * k)) & 1u) ? A : MIS;
                    const int diag = M != 0 ? M + sc : 0;
                    const int H = max(max(diag, E), f);
                    key = max(key, ((uint32_t)H << 16) | (jb + k));
                    const int ne = max(max(E - e_del, diag - oe_del), 0);
                    f = max(max(f - e_ins, diag - oe_ins), 0);
                    cb[k * 64] = S::pack(h1, ne);
                    h1 = H;
                }
            }
        }
        for (int j = s2; j < jhi; ++j) masked_cell(j);

        if (alive) {
            const int row_max = (int)(key >> 16);
            const int row_max_j = (int)(key & 0xFFFFu);
            col[end * 64] = S::pack(h1, 0);                    // :145-146
            if (end == qlen) {                                 // :149-152
                gtle = gsc > h1 ? gtle : i;
                gsc = max(gsc, h1);
            }
            if (row_max == 0) {                                // :155
                alive = false;
            } else if (row_max > maxv) {                       // :158-163
                maxv = row_max;
                max_i = i;
                max_j = row_max_j;
                max_off = max(max_off, abs(row_max_j - i));
            } else if (p.zdrop > 0) {                          // :164-172
                const int di = i - max_i, dj = row_max_j - max_j;
                const int drop = di > dj ? maxv - row_max - (di - dj) * e_del
                                         : maxv - row_max - (dj - di) * e_ins;
                if (drop > p.zdrop) alive = false;
            }
            if (alive) {                                       // :175-180
                int nb = beg;
                while (nb < end && S::zero(col[nb * 64])) ++nb;
                beg = nb;
                int ne = end;
                while (ne >= beg && S::zero(col[ne * 64])) --ne;
                end = min(ne + 2, qlen);
            }
        }
    }
    if (d.cells) {
        unsigned long long c = ncells;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
        if (lane == 0) atomicAdd(d.cells, c);
    }
    if (valid) {                                               // :183-190
        int32_t *o = d.out + (size_t)job * 6;
        if (nonempty) {
            o[0] = maxv; o[1] = max_j + 1; o[2] = max_i + 1;
            o[3] = gtle + 1; o[4] = gsc; o[5] = max_off;
        } else {
            o[0] = 0; o[1] = 0; o[2] = 0; o[3] = 0; o[4] = 0; o[5] = 0;
        }
    }
}

// Picks the fast or the generic form for this wave's jobs (uniform branch).
template <int TIER>
__device__ __forceinline__ void ext_wave(const ExtDev &d, const ExtParams &p,
                                         const uint32_t *idx, int count,
                                         typename St<TIER>::W *col, const int8_t *smat,
                                         uint32_t *qlds)
{
    if (TIER != TIER_I32 && p.std_mat) {
        const int lane = threadIdx.x & 63;
        const bool valid = lane < count;
        const int job = valid ? (int)idx[lane] : 0;
        const bool has_n = valid && (d.qn_word[job] != kNoN || d.tn_word[job] != kNoN);
        if (__ballot(has_n) == 0) {
            ext_wave_fast<TIER>(d, p, idx, count, col, qlds);
            return;
        }
    }
    ext_wave_generic<TIER>(d, p, idx, count, col, smat);
}

// Per-wave state pointer: LDS when the columns fit, else a per-block slice of
// a global scratch buffer (cols * 64 words per block).
__device__ __forceinline__ int query_blocks(int cols) { return (cols - 1 + 15) >> 4; }

template <int TIER, bool GLOBAL>
__device__ __forceinline__ typename St<TIER>::W *state_base(void *gscratch, int cols)
{
    using W = typename St<TIER>::W;
    if (GLOBAL) {
        return reinterpret_cast<W *>(gscratch) + (size_t)blockIdx.x * cols * 64 + threadIdx.x;
    } else {
        extern __shared__ __align__(16) unsigned char lds_raw[];
        return reinterpret_cast<W *>(lds_raw + 32 + (size_t)query_blocks(cols) * 256) + threadIdx.x;
    }
}

__device__ __forceinline__ uint32_t *query_lds(int cols)
{
    (void)cols;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    return reinterpret_cast<uint32_t *>(lds_raw + 32);
}

__device__ __forceinline__ const int8_t *load_smat(const ExtParams &p)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    int8_t *smat = reinterpret_cast<int8_t *>(lds_raw);
    if (threadIdx.x < 32) smat[threadIdx.x] = p.mat[threadIdx.x];
    __syncthreads();
    return smat;
}

// Plain form: one launch over the jobs in input order, 64 consecutive jobs per
// wave.  With global state the grid is capped and waves stride over groups.
template <int TIER, bool GLOBAL>
__global__ __launch_bounds__(64) void ext_plain_kernel(ExtDev d, ExtParams p,
                                                       const uint32_t *order, int n, int cols,
                                                       void *gscratch)
{
    const int8_t *smat = load_smat(p);
    typename St<TIER>::W *col = state_base<TIER, GLOBAL>(gscratch, cols);
    __shared__ uint32_t idx[64];
    const int lane = threadIdx.x;
    const int groups = (n + 63) / 64;
    for (int g = blockIdx.x; g < groups; g += gridDim.x) {
        const int base = g * 64;
        const int count = min(64, n - base);
        if (lane < count) idx[lane] = order ? order[base + lane] : (uint32_t)(base + lane);
        __syncthreads();
        ext_wave<TIER>(d, p, idx, count, col, smat, query_lds(cols));
        __syncthreads();
    }
}

// Persistent work-stealing form: the grid is sized to the machine; each wave
// takes the next WorkItem from a global counter until the list is drained.
// Items are ordered by decreasing DP area, so the longest groups start first.
template <int TIER, bool GLOBAL>
__global__ __launch_bounds__(64) void ext_persistent_kernel(ExtDev d, ExtParams p,
                                                            const uint32_t *order,
                                                            const WorkItem *items, int n_items,
                                                            uint32_t *counter, int cols,
                                                            void *gscratch)
{
    const int8_t *smat = load_smat(p);
    typename St<TIER>::W *col = state_base<TIER, GLOBAL>(gscratch, cols);
    __shared__ uint32_t idx[64];
    __shared__ int item_sh;
    for (;;) {
        if (threadIdx.x == 0) item_sh = (int)atomicAdd(counter, 1u);
        __syncthreads();
        const int item = item_sh;
        if (item >= n_items) break;
        const WorkItem wi = items[item];
        if (threadIdx.x < wi.count) idx[threadIdx.x] = order[wi.start + threadIdx.x];
        __syncthreads();
        ext_wave<TIER>(d, p, idx, (int)wi.count, col, smat, query_lds(cols));
        __syncthreads();
    }
}

__global__ void zero_u32_kernel(uint32_t *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

hipError_t launch_zero_u32(uint32_t *p, int n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(zero_u32_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

size_t ext_lds_bytes(int tier, bool global, int cols)
{
    const size_t qblocks = (size_t)(cols - 1 + 15) / 16;
    return 32 + qblocks * 256 + (global ? 0 : (size_t)cols * 64 * ext_state_bytes(tier));
}
static size_t lds_bytes(int tier, bool global, int cols) { return ext_lds_bytes(tier, global, cols); }

template <int TIER, bool GLOBAL>
static hipError_t plain_t(const ExtDev &d, const ExtParams &p, const uint32_t *order, int n,
                          int cols, int grid, void *gs, hipStream_t s)
{
    hipLaunchKernelGGL((ext_plain_kernel<TIER, GLOBAL>), dim3(grid), dim3(64),
                       lds_bytes(TIER, GLOBAL, cols), s, d, p, order, n, cols, gs);
    return hipGetLastError();
}

template <int TIER, bool GLOBAL>
static hipError_t persist_t(const ExtDev &d, const ExtParams &p, const uint32_t *order,
                            const WorkItem *items, int n_items, uint32_t *counter, int cols,
                            int grid, void *gs, hipStream_t s)
{
    hipLaunchKernelGGL((ext_persistent_kernel<TIER, GLOBAL>), dim3(grid), dim3(64),
                       lds_bytes(TIER, GLOBAL, cols), s, d, p, order, items, n_items, counter,
                       cols, gs);
    return hipGetLastError();
}

hipError_t launch_ext_plain(int tier, bool global_state, const ExtDev &d, const ExtParams &p,
                            const uint32_t *order, int n, int cols, int grid, void *gscratch,
                            hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    switch (tier * 2 + (global_state ? 1 : 0)) {
    case 0: return plain_t<TIER_U8, false>(d, p, order, n, cols, grid, gscratch, s);
    case 1: return plain_t<TIER_U8, true>(d, p, order, n, cols, grid, gscratch, s);
    case 2: return plain_t<TIER_U16, false>(d, p, order, n, cols, grid, gscratch, s);
    case 3: return plain_t<TIER_U16, true>(d, p, order, n, cols, grid, gscratch, s);
    case 4: return plain_t<TIER_I32, false>(d, p, order, n, cols, grid, gscratch, s);
    default: return plain_t<TIER_I32, true>(d, p, order, n, cols, grid, gscratch, s);
    }
}

hipError_t launch_ext_persistent(int tier, bool global_state, const ExtDev &d,
                                 const ExtParams &p, const uint32_t *order,
                                 const WorkItem *items, int n_items, uint32_t *counter,
                                 int cols, int grid, void *gscratch, hipStream_t s)
{
    if (n_items <= 0) return hipSuccess;
    switch (tier * 2 + (global_state ? 1 : 0)) {
    case 0: return persist_t<TIER_U8, false>(d, p, order, items, n_items, counter, cols, grid, gscratch, s);
    case 1: return persist_t<TIER_U8, true>(d, p, order, items, n_items, counter, cols, grid, gscratch, s);
    case 2: return persist_t<TIER_U16, false>(d, p, order, items, n_items, counter, cols, grid, gscratch, s);
    case 3: return persist_t<TIER_U16, true>(d, p, order, items, n_items, counter, cols, grid, gscratch, s);
    case 4: return persist_t<TIER_I32, false>(d, p, order, items, n_items, counter, cols, grid, gscratch, s);
    default: return persist_t<TIER_I32, true>(d, p, order, items, n_items, counter, cols, grid, gscratch, s);
    }
}


// ---------------------------------------------------------------------------
// Packed form: two jobs per lane, 8-bit tier.
//
// Lane l carries job idx[l] in the low 16-bit half and job idx[l+64] in the
// high half of every 32-bit register, and each cell step is a packed 16-bit
// op (v_pk_*) on both.  The column state word holds (hA, eA, hB, eB) as bytes
// 0..3, so a column read yields both jobs' H and E and a job's own (h, e) pair
// is one 16-bit half (its band-end store is one ds_write_b16).  All values of
// the 8-bit tier are <= 255, so every packed intermediate (M + A, H * 256 + j)
// fits in 16 bits and the saturating 16-bit subtract is the max(x - c, 0) of
// the scalar recurrence.
//
// Band edges.  Each row sweeps whole 16-column blocks from the wave's lowest
// band start to its highest band end.
//  * Left of a job's band (j < beg) the job computes without masks: ksw_extend2
//    never reads h/e entries below beg again (beg only grows), so the kernel
//    keeps every such entry at zero -- the narrowing step (:175-177) leaves
//    zeros behind, and the one entry the w-clamp (:96) drops per row is
//    cleared -- and a cell with M = E = F = 0 and h1 = 0 (the row's h1 when
//    beg > 0) yields H = 0 and writes zeros again.  Its (H = 0, j) row-max
//    candidates never win: a row whose maximum is 0 ends the job (:155).
//  * Right of a job's band (j >= end) entries are live (the band end may grow
//    by up to two columns per row), so a block that crosses some job's end
//    commits the state word, h1 and the row-max candidate under a per-half
//    mask; blocks left of every live end run unmasked.
// ---------------------------------------------------------------------------
namespace {

typedef unsigned short us2 __attribute__((ext_vector_type(2)));
typedef short ss2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ us2 u2(uint32_t v) { return __builtin_bit_cast(us2, v); }
__device__ __forceinline__ uint32_t w32(us2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ us2 splat(int v) { return us2{(unsigned short)v, (unsigned short)v}; }
__device__ __forceinline__ us2 vmax(us2 a, us2 b) { return __builtin_elementwise_max(a, b); }
__device__ __forceinline__ us2 vmin(us2 a, us2 b) { return __builtin_elementwise_min(a, b); }
__device__ __forceinline__ us2 vsubs(us2 a, us2 b) { return __builtin_elementwise_sub_sat(a, b); }
// signed 16-bit max of a possibly negative value and a non-negative one
__device__ __forceinline__ us2 smax(ss2 a, us2 b)
{
    return __builtin_bit_cast(us2, __builtin_elementwise_max(a, __builtin_bit_cast(ss2, b)));
}
// A constant the compiler cannot see through: keeps min(M, 1) a packed min
// instead of per-half compares and selects.
__device__ __forceinline__ uint32_t opaque(uint32_t v)
{
    asm volatile("" : "+s"(v));
    return v;
}

// Wave-uniform min / max of a per-lane int (all 64 lanes active): DPP steps
// reduce each 16-lane row, lane reads combine the four rows in scalar registers.
__device__ __forceinline__ int row_reduce_min(int v)
{
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xF, 0xF, false));    // quad_perm 1,0,3,2
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xF, 0xF, false));    // quad_perm 2,3,0,1
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x141, 0xF, 0xF, false));   // row_half_mirror
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x140, 0xF, 0xF, false));   // row_mirror
    return v;
}
__device__ __forceinline__ int uniform_min(int v)
{
    v = row_reduce_min(v);
    return min(min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
               min(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}
__device__ __forceinline__ int uniform_max(int v) { return -uniform_min(-v); }

// 16 match/mismatch bits of a block, compacted: bit k = base k differs.
__device__ __forceinline__ uint32_t mismatch16(uint32_t qw, uint32_t tpat)
{
    uint32_t x = qw ^ tpat;
    x = (x | (x >> 1)) & 0x55555555u;
    x = (x | (x >> 1)) & 0x33333333u;
    x = (x | (x >> 2)) & 0x0F0F0F0Fu;
    x = (x | (x >> 4)) & 0x00FF00FFu;
    x = (x | (x >> 8)) & 0x0000FFFFu;
    return x;
}

// Column state layout: groups of four columns, lane-major inside a group, so
// one 128-bit LDS access moves four columns of a lane (a 16-column block is
// four reads and four writes).  Word (j, lane) is u32 number
// (j / 4) * 256 + lane * 4 + j % 4.  A job's (h, e) entry of column j sits at
// cpos<T>(j) elements from the job's pointer: T = u16 for the packed form (one
// 16-bit half of the word), T = u32 for the one-job-per-lane form.
template <typename T> __device__ __forceinline__ int cpos(int j);
template <> __device__ __forceinline__ int cpos<uint16_t>(int j) { return (j >> 2) * 512 + (j & 3) * 2; }
template <> __device__ __forceinline__ int cpos<uint32_t>(int j) { return (j >> 2) * 256 + (j & 3); }

// Hybrid column state: a kernel instance keeps its first LB 16-column blocks
// (LC = 16 * LB columns) in LDS and the rest in the wave's own HBM region (same
// group layout, L2-resident), so that a wave's LDS share, and with it the waves
// a CU holds, does not grow with the query: 150-bp queries held whole take
// 40 KiB (four waves per CU), held as 96 + 64 columns 24 KiB (six).  ColRef is a
// job's (h, e) entry of column j, wherever it lives; LC = 0: all in LDS.
// u32 words of one wave's HBM region (blocks LB..NQ-1, 64 lanes)
__host__ __device__ constexpr size_t ext_hyb_words(int nq, int lb) { return nq > lb ? (size_t)(nq - lb) * 1024 : 0; }
template <typename T, int LC> struct ColRef {
    T *l, *g;
    __device__ __forceinline__ T &operator[](int j) const
    {
        if (LC > 0 && j >= LC) return g[cpos<T>(j - LC)];
        return l[cpos<T>(j)];
    }
};

// Per-job row-level state of one half.
struct PkJob {
    bool valid, nonempty, alive;
    int job, qlen, tlen, h0, w, beg, end, maxv, max_i, max_j, gsc, gtle, max_off, tnw;
    uint32_t qw0, tw0, twe, two, tb, ncells;   // twe / two: target words of even / odd 16-row blocks
};

__device__ __forceinline__ void pk_setup(PkJob &J, const ExtDev &d, const ExtParams &p,
                                         const PkMeta *meta, int slot, int count)
{
    J.valid = slot < count;
    const PkMeta m = meta[J.valid ? slot : 0];
    J.job = (int)m.job;
    J.qlen = J.valid ? m.q_len : 0;
    J.tlen = J.valid ? m.t_len : 0;
    J.h0 = J.valid ? m.h0 : 0;
    int w = J.valid ? m.w : 0;
    const int eb = J.valid ? m.eb : 0;
    J.qw0 = J.valid ? m.q_word : 0;
    J.tw0 = J.valid ? m.t_word : 0;
    J.nonempty = J.valid && J.qlen > 0 && J.tlen > 0;        // :33-35
    J.alive = J.nonempty;
    if (J.alive) {                                             // w cap, :52-62
        int mi = (int)((double)(J.qlen * p.max_mat + eb - p.o_ins) / (double)p.e_ins + 1.0);
        w = min(w, max(1, mi));
        int md = (int)((double)(J.qlen * p.max_mat + eb - p.o_del) / (double)p.e_del + 1.0);
        w = min(w, max(1, md));
    }
    J.w = w;
    J.beg = 0; J.end = J.qlen;                                 // :79-87
    J.maxv = J.h0; J.max_i = -1; J.max_j = -1;
    J.gsc = -1; J.gtle = -1; J.max_off = 0;
    // target words of rows 0-15 and 16-31 (the second is a prefetch); word
    // indexes stop at the job's last word + 1, inside the padded arena
    J.tnw = (J.tlen + 15) >> 4;
    J.twe = d.seq[J.tw0];
    J.two = d.seq[J.tw0 + min(1, J.tnw)];
    J.tb = 0; J.ncells = 0;
}

// Row start of one job (:92-107); returns h1.  `c16` is the job's column
// state ((h, e) of column j at c16[cpos<T>(j)]).  The caller keeps the target
// words of the current and the next 16-row block loaded.
template <typename T, int LC>
__device__ __forceinline__ int pk_row_start(PkJob &J, const ExtParams &p, int i,
                                            const ColRef<T, LC> &c16)
{
    J.alive = J.alive && i < J.tlen;
    J.tb = (((i >> 4) & 1 ? J.two : J.twe) >> ((i & 15) * 2)) & 3u;   // this row's target base
    const int wb = i - J.w;
    if (J.alive && J.beg < wb) {
        // beg = i - w (:96) drops column beg: with w >= 0, beg >= i - 1 - w
        // after the previous row, so exactly one column per row; clear it
        c16[J.beg] = 0;
        J.beg = wb;
    }
    J.end = min(min(J.end, i + J.w + 1), J.qlen);             // :97-98
    J.ncells += (J.alive && J.end > J.beg) ? (uint32_t)(J.end - J.beg) : 0u;
    return (J.alive && J.beg == 0) ? max(0, J.h0 - (p.o_del + p.e_del * (i + 1))) : 0;
}

// Row end of one half (:145-172): band-end entry, global-score column, row
// maximum and z-drop.  A finished job's half is never read again, so its
// band-end store needs no guard (end <= qlen keeps it in bounds).
template <typename T, int LC>
__device__ __forceinline__ void pk_row_end(PkJob &J, const ExtParams &p, int i, int h1, int key,
                                           const ColRef<T, LC> &c16)
{
    c16[J.end] = (T)h1;                                        // h[end] = h1, e[end] = 0
    const int rm = key >> 8, rj = key & 0xFF;
    const bool at_q = J.alive && J.end == J.qlen;              // :149-152
    J.gtle = (at_q && !(J.gsc > h1)) ? i : J.gtle;
    J.gsc = at_q ? max(J.gsc, h1) : J.gsc;
    const bool up = J.alive && rm > J.maxv;                    // :158-163
    const int di = i - J.max_i, dj = rj - J.max_j;             // :164-172
    const int drop = di > dj ? J.maxv - rm - (di - dj) * p.e_del : J.maxv - rm - (dj - di) * p.e_ins;
    const bool zd = p.zdrop > 0 && drop > p.zdrop;
    J.max_off = up ? max(J.max_off, abs(rj - i)) : J.max_off;
    J.max_i = up ? i : J.max_i;
    J.max_j = up ? rj : J.max_j;
    J.maxv = up ? rm : J.maxv;
    J.alive = J.alive && rm != 0 && (up || !zd);               // :155, :171
}

// Band narrowing of both halves (:175-180).  Each trip reads four columns per
// half at once (one LDS round trip); more trips only while all four are zero.
template <typename T, int LC>
__device__ __forceinline__ void pk_narrow(PkJob &A, PkJob &B, int h1a, int h1b,
                                          const ColRef<T, LC> &ca, const ColRef<T, LC> &cb)
{
    // first column in [beg, end) whose (h, e) is non-zero, else end
    int ja = A.beg, jb = B.beg;
    bool ga = A.alive && ja < A.end, gb = B.alive && jb < B.end;
    while (ga || gb) {
        const int ea = A.end, eb = B.end;
        uint32_t va[4], vb[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            va[k] = ca[min(ja + k, ea)];
            vb[k] = cb[min(jb + k, eb)];
        }
        bool za = ga && va[0] == 0, zb = gb && vb[0] == 0;
        int na = za, nb = zb;
#pragma unroll
        for (int k = 1; k < 4; ++k) {
            za = za && ja + k < ea && va[k] == 0;
            zb = zb && jb + k < eb && vb[k] == 0;
            na += za;
            nb += zb;
        }
        ja += na;
        jb += nb;
        ga = za && ja < ea;
        gb = zb && jb < eb;
    }
    if (A.alive) A.beg = ja;
    if (B.alive) B.beg = jb;
    // last column in [beg, end] whose (h, e) is non-zero; column end holds
    // (h1, 0), so only a zero h1 starts the backward scan
    int ka = A.end, kb = B.end;
    bool sa = A.alive && ka >= A.beg && h1a == 0, sb = B.alive && kb >= B.beg && h1b == 0;
    ka -= sa;
    kb -= sb;
    sa = sa && ka >= A.beg;
    sb = sb && kb >= B.beg;
    while (sa || sb) {
        uint32_t va[4], vb[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            va[k] = ca[max(ka - k, 0)];
            vb[k] = cb[max(kb - k, 0)];
        }
        bool za = sa && va[0] == 0, zb = sb && vb[0] == 0;
        int na = za, nb = zb;
#pragma unroll
        for (int k = 1; k < 4; ++k) {
            za = za && ka - k >= A.beg && va[k] == 0;
            zb = zb && kb - k >= B.beg && vb[k] == 0;
            na += za;
            nb += zb;
        }
        ka -= na;
        kb -= nb;
        sa = za && ka >= A.beg;
        sb = zb && kb >= B.beg;
    }
    if (A.alive) A.end = min(ka + 2, A.qlen);
    if (B.alive) B.end = min(kb + 2, B.qlen);
}

// The 16 column words of block b of this lane: four 128-bit accesses to LDS,
// or to the wave's HBM region for blocks from LB on (HYB).  The pointers
// carry their address space so that each path keeps its own instructions
// (ds_*_b128 / global_*_dwordx4) instead of generic flat accesses.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) u32x4 lds_u4;
typedef __attribute__((address_space(1))) u32x4 glb_u4;

template <bool HYB, int LB>
__device__ __forceinline__ void load_block(const uint32_t *col, const uint32_t *gcol, int b,
                                           uint32_t (&wv)[16])
{
    if (!HYB || b < LB) {
        const lds_u4 *cb = (const lds_u4 *)(col + b * 1024);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const u32x4 v = cb[q * 64];
            wv[4 * q] = v.x; wv[4 * q + 1] = v.y; wv[4 * q + 2] = v.z; wv[4 * q + 3] = v.w;
        }
    } else {
        const glb_u4 *cb = (const glb_u4 *)(gcol + (b - LB) * 1024);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const u32x4 v = cb[q * 64];
            wv[4 * q] = v.x; wv[4 * q + 1] = v.y; wv[4 * q + 2] = v.z; wv[4 * q + 3] = v.w;
        }
    }
}

template <bool HYB, int LB>
__device__ __forceinline__ void store_block(uint32_t *col, uint32_t *gcol, int b,
                                            const uint32_t (&wv)[16])
{
    if (!HYB || b < LB) {
        lds_u4 *cb = (lds_u4 *)(col + b * 1024);
#pragma unroll
        for (int q = 0; q < 4; ++q)
            cb[q * 64] = u32x4{wv[4 * q], wv[4 * q + 1], wv[4 * q + 2], wv[4 * q + 3]};
    } else {
        glb_u4 *cb = (glb_u4 *)(gcol + (b - LB) * 1024);
#pragma unroll
        for (int q = 0; q < 4; ++q)
            cb[q * 64] = u32x4{wv[4 * q], wv[4 * q + 1], wv[4 * q + 2], wv[4 * q + 3]};
    }
}

// Groups g .. g + 3 of this lane (4-column groups, 128-bit accesses); each
// group index is clamped to glast (a read past the row's last group repeats it:
// in bounds, unused; g <= glast).  Groups below GL live in LDS, the rest in the
// wave's HBM region (GL = 0: all in LDS).  At most one block per row straddles
// the two; it takes a uniform branch per group.
template <int GL>
__device__ __forceinline__ void load_groups(const uint32_t *col, const uint32_t *gcol, int g,
                                            int glast, uint32_t (&wv)[16])
{
    const lds_u4 *lc = (const lds_u4 *)col;
    const glb_u4 *gc = (const glb_u4 *)gcol;
    int gq[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) gq[q] = min(g + q, glast);
    u32x4 v[4];
    if (GL == 0 || gq[3] < GL) {
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = lc[gq[q] * 64];
    } else if (g >= GL) {
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = gc[(gq[q] - GL) * 64];
    } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (gq[q] < GL) v[q] = lc[gq[q] * 64];
            else v[q] = gc[(gq[q] - GL) * 64];
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        wv[4 * q] = v[q].x; wv[4 * q + 1] = v[q].y; wv[4 * q + 2] = v[q].z; wv[4 * q + 3] = v[q].w;
    }
}

// Stores groups g .. g + NG - 1 (all inside the row's sweep).
template <int GL, int NG>
__device__ __forceinline__ void store_groups(uint32_t *col, uint32_t *gcol, int g,
                                             const uint32_t (&wv)[16])
{
    lds_u4 *lc = (lds_u4 *)col;
    glb_u4 *gc = (glb_u4 *)gcol;
    if (GL == 0 || g + NG - 1 < GL) {
#pragma unroll
        for (int q = 0; q < NG; ++q)
            lc[(g + q) * 64] = u32x4{wv[4 * q], wv[4 * q + 1], wv[4 * q + 2], wv[4 * q + 3]};
    } else if (g >= GL) {
#pragma unroll
        for (int q = 0; q < NG; ++q)
            gc[(g + q - GL) * 64] = u32x4{wv[4 * q], wv[4 * q + 1], wv[4 * q + 2], wv[4 * q + 3]};
    } else {
#pragma unroll
        for (int q = 0; q < NG; ++q) {
            const u32x4 v{wv[4 * q], wv[4 * q + 1], wv[4 * q + 2], wv[4 * q + 3]};
            if (g + q < GL) lc[(g + q) * 64] = v;
            else gc[(g + q - GL) * 64] = v;
        }
    }
}

// Groups g .. g + 3 (clamped to gmax) of one memory only: KIND 0 = LDS
// (groups below GL), KIND 2 = the wave's HBM region (groups GL and up).  A
// sweep phase that touches one memory only lets the compiler wait for exactly
// the block it is about to use while the next block's reads stay in flight;
// behind a per-block LDS/HBM branch it has to wait for every outstanding load
// and store.
template <int KIND, int GL>
__device__ __forceinline__ void load_groups_k(const uint32_t *col, const uint32_t *gcol, int g,
                                              int gmax, uint32_t (&wv)[16])
{
    static_assert(KIND == 0 || KIND == 2, "one memory per phase");
    u32x4 v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int gq = min(g + q, gmax);
        if (KIND == 0) v[q] = ((const lds_u4 *)col)[gq * 64];
        else v[q] = ((const glb_u4 *)gcol)[(gq - GL) * 64];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        wv[4 * q] = v[q].x; wv[4 * q + 1] = v[q].y; wv[4 * q + 2] = v[q].z; wv[4 * q + 3] = v[q].w;
    }
}

template <int KIND, int GL, int NG>
__device__ __forceinline__ void store_groups_k(uint32_t *col, uint32_t *gcol, int g,
                                               const uint32_t (&wv)[16])
{
    static_assert(KIND == 0 || KIND == 2, "one memory per phase");
#pragma unroll
    for (int q = 0; q < NG; ++q) {
        const u32x4 v{wv[4 * q], wv[4 * q + 1], wv[4 * q + 2], wv[4 * q + 3]};
        if (KIND == 0) ((lds_u4 *)col)[(g + q) * 64] = v;
        else ((glb_u4 *)gcol)[(g + q - GL) * 64] = v;
    }
}

__device__ __forceinline__ void pk_finish(const PkJob &J, const ExtDev &d)
{
    if (!J.valid) return;                                      // :183-190
    int32_t *o = d.out + (size_t)J.job * 6;
    if (J.nonempty) {
        o[0] = J.maxv; o[1] = J.max_j + 1; o[2] = J.max_i + 1;
        o[3] = J.gtle + 1; o[4] = J.gsc; o[5] = J.max_off;
    } else {
        o[0] = 0; o[1] = 0; o[2] = 0; o[3] = 0; o[4] = 0; o[5] = 0;
    }
}

// ---- row-level helpers of the packed form ---------------------------------

// Wave-uniform minima of three per-lane u32 values (all 64 lanes active).
// Passing the identity as the DPP "old" value lets the compiler fuse each
// step into one v_min_u32_dpp; the three chains are interleaved so no step
// waits on the previous one.  Rows are reduced by quad_perm / mirror steps,
// then row_bcast:15 and row_bcast:31 carry rows 0-2 into row 3; lane 63
// holds the result.
__device__ __forceinline__ void wave_min3(uint32_t &a, uint32_t &b, uint32_t &c)
{
#define HIPSW_MIN3_STEP(CTRL, RM)                                                             \
    a = min(a, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)a, CTRL, RM, 0xF, false));      \
    b = min(b, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)b, CTRL, RM, 0xF, false));      \
    c = min(c, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)c, CTRL, RM, 0xF, false));
    HIPSW_MIN3_STEP(0xB1, 0xF)      // quad_perm 1,0,3,2
    HIPSW_MIN3_STEP(0x4E, 0xF)      // quad_perm 2,3,0,1
    HIPSW_MIN3_STEP(0x141, 0xF)     // row_half_mirror
    HIPSW_MIN3_STEP(0x140, 0xF)     // row_mirror
    HIPSW_MIN3_STEP(0x142, 0xA)     // row_bcast:15 into rows 1, 3
    HIPSW_MIN3_STEP(0x143, 0xC)     // row_bcast:31 into rows 2, 3
#undef HIPSW_MIN3_STEP
    a = __builtin_amdgcn_readlane(a, 63);
    b = __builtin_amdgcn_readlane(b, 63);
    c = __builtin_amdgcn_readlane(c, 63);
}

// Stores the 16-bit entry of column j, half hs, with an LDS or a global store
// chosen per lane.  A generic (flat) store would leave a flat access pending
// into the next sweep, and with one pending the compiler waits for every LDS
// access at each block.
template <int LC>
__device__ __forceinline__ void store16(uint32_t *col, uint32_t *gcol, int j, int hs, uint16_t v)
{
    typedef __attribute__((address_space(3))) uint16_t lds_u16;
    typedef __attribute__((address_space(1))) uint16_t glb_u16;
    if (LC == 0 || j < LC) ((lds_u16 *)col)[cpos<uint16_t>(j) + hs] = v;
    else ((glb_u16 *)gcol)[cpos<uint16_t>(j - LC) + hs] = v;
}

// Group g (columns 4g..4g+3, both halves) of this lane's column state.
// An LDS or a global read chosen per lane (no generic access: see store16).
template <int GL>
__device__ __forceinline__ u32x4 group_rd(const uint32_t *col, const uint32_t *gcol, int g)
{
    if (GL == 0 || g < GL) return *(const lds_u4 *)(col + g * 256);
    return *(const glb_u4 *)(gcol + (g - GL) * 256);
}

// Non-zero columns of two groups: bit c (c < 4) = column c of va's low half
// (job A) is non-zero, bit 16 + c = column c of vb's high half (job B).
__device__ __forceinline__ uint32_t nz_bits(const u32x4 &va, const u32x4 &vb)
{
    const us2 one = u2(opaque(0x00010001u));   // keeps min(x, 1) one packed op
    const uint32_t m0 = (va.x & 0xFFFFu) | (vb.x & 0xFFFF0000u);
    const uint32_t m1 = (va.y & 0xFFFFu) | (vb.y & 0xFFFF0000u);
    const uint32_t m2 = (va.z & 0xFFFFu) | (vb.z & 0xFFFF0000u);
    const uint32_t m3 = (va.w & 0xFFFFu) | (vb.w & 0xFFFF0000u);
    return w32(vmin(u2(m0), one)) | (w32(vmin(u2(m1), one)) << 1) |
           (w32(vmin(u2(m2), one)) << 2) | (w32(vmin(u2(m3), one)) << 3);
}

// Row-level state of the packed form.  Each field pair holds job A in the low
// and job B in the high 16-bit half, so the per-row bookkeeping of both jobs
// is one packed instruction per step.  All values fit 16 bits: qlen <= 255,
// every score of the 8-bit tier <= 255, rows and t_len <= 65535; -1 (max_i,
// max_j, gscore, gtle before their first update) is 0xFFFF.
struct PkRow {
    uint32_t alive;                        // 0xFFFF per live half
    us2 qlen, tlen, h0, w, beg, end;       // w capped (:52-62), at most 65535
    us2 maxv, maxi, maxj, gsc, gtle, maxoff;
    us2 nc16;                              // cells of this and earlier rows, < 65536
    uint32_t ncells;                       // both jobs' cells flushed from nc16
    uint32_t twe[2], two[2];               // target words of even / odd 16-row blocks
    uint32_t tw0[2], tnw[2], qw0[2];
};

__device__ __forceinline__ us2 bfi16(uint32_t m, us2 a, us2 b) { return u2((m & w32(a)) | (~m & w32(b))); }
// 0xFFFF in each half where a < b (signed halves, |a - b| < 32768)
__device__ __forceinline__ uint32_t lt16(us2 a, us2 b)
{
    return w32(__builtin_bit_cast(
        us2, (__builtin_bit_cast(ss2, a) - __builtin_bit_cast(ss2, b)) >> ss2{15, 15}));
}
// 0xFFFF in each half where a != 0 (one = the opaque constant 0x00010001)
__device__ __forceinline__ uint32_t nz16(us2 a, us2 one) { return w32(splat(0) - vmin(a, one)); }
// both halves of v set to min(v.lo, v.hi)
__device__ __forceinline__ uint32_t hmin16(uint32_t v)
{
    return w32(vmin(u2(v), u2(__builtin_amdgcn_alignbit(v, v, 16))));
}

__device__ __forceinline__ void pkr_setup(PkRow &R, const ExtDev &d, const ExtParams &p,
                                          const PkMeta *meta, int lane, int count)
{
    unsigned short ql[2], tl[2], h0[2], wc[2];
    uint32_t al = 0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int slot = lane + 64 * h;
        const bool valid = slot < count;
        const PkMeta m = meta[valid ? slot : 0];
        const int qlen = valid ? m.q_len : 0, tlen = valid ? m.t_len : 0;
        int w = valid ? m.w : 0;
        const int eb = valid ? m.eb : 0;
        const bool nonempty = valid && qlen > 0 && tlen > 0;    // :33-35
        if (nonempty) {                                          // w cap, :52-62
            int mi = (int)((double)(qlen * p.max_mat + eb - p.o_ins) / (double)p.e_ins + 1.0);
            w = min(w, max(1, mi));
            int md = (int)((double)(qlen * p.max_mat + eb - p.o_del) / (double)p.e_del + 1.0);
            w = min(w, max(1, md));
            al |= 0xFFFFu << (16 * h);
        }
        // a band wider than 65535 clamps nothing (i - w < 0 and i + w + 1 > qlen)
        ql[h] = (unsigned short)qlen;
        tl[h] = (unsigned short)tlen;
        h0[h] = (unsigned short)(valid ? m.h0 : 0);
        wc[h] = (unsigned short)min(max(w, 0), 65535);
        R.qw0[h] = valid ? m.q_word : 0u;
        R.tw0[h] = valid ? m.t_word : 0u;
        // target words of rows 0-15 and 16-31 (the second is a prefetch);
        // word indexes stop at the job's last word + 1, inside the padded arena
        R.tnw[h] = (uint32_t)(tlen + 15) >> 4;
        R.twe[h] = d.seq[R.tw0[h]];
        R.two[h] = d.seq[R.tw0[h] + min(1u, R.tnw[h])];
    }
    R.alive = al;
    R.qlen = us2{ql[0], ql[1]};
    R.tlen = us2{tl[0], tl[1]};
    R.h0 = us2{h0[0], h0[1]};
    R.w = us2{wc[0], wc[1]};
    R.beg = splat(0);                                            // :79-87
    R.end = R.qlen;
    R.maxv = R.h0;
    R.maxi = R.maxj = R.gsc = R.gtle = splat(0xFFFF);
    R.maxoff = splat(0);
    R.nc16 = splat(0);
    R.ncells = 0;
}

// SWResult of half h (:183-190); an empty or absent job writes zeros.
__device__ __forceinline__ void pkr_finish(const PkRow &R, const ExtDev &d, const PkMeta *meta,
                                           int lane, int count, int h)
{
    const int slot = lane + 64 * h;
    if (slot >= count) return;
    const int sh = 16 * h;
    auto s16 = [&](us2 v) { return (int)(short)(w32(v) >> sh); };     // -1 kept
    auto u16 = [&](us2 v) { return (int)((w32(v) >> sh) & 0xFFFF); };
    int32_t *o = d.out + (size_t)meta[slot].job * 6;
    if (u16(R.qlen) > 0 && u16(R.tlen) > 0) {
        const int mi = u16(R.maxi), gt = u16(R.gtle);
        o[0] = u16(R.maxv);
        o[1] = s16(R.maxj) + 1;
        o[2] = mi == 0xFFFF ? 0 : mi + 1;
        o[3] = gt == 0xFFFF ? 0 : gt + 1;
        o[4] = s16(R.gsc);
        o[5] = u16(R.maxoff);
    } else {
        o[0] = 0; o[1] = 0; o[2] = 0; o[3] = 0; o[4] = 0; o[5] = 0;
    }
}

// Column state of the packed forms.  8-bit tier: one u32 per (column, lane)
// holding the bytes (hA, eA, hB, eB), a job's entry the 16-bit half
// h | e << 8.  16-bit tier (W16): two planes, h words (hA | hB << 16) and e
// words (eA | eB << 16), the e plane after the h plane in LDS and in the
// wave's HBM region.  Stores (h, e = 0) into half hs of column j.
template <int LC, bool W16>
__device__ __forceinline__ void store_he(uint32_t *col, uint32_t *gcol, uint32_t *colE, uint32_t *gcolE,
                                         int j, int hs, uint16_t h)
{
    store16<LC>(col, gcol, j, hs, h);
    if (W16) store16<LC>(colE, gcolE, j, hs, 0);
}

// Row start of both jobs (:92-98; the h1 of :99-107 is made by the caller):
// band clamp -- the column the w-clamp drops is cleared, which keeps every
// entry below beg at zero -- band end, cell count.
template <int LC, bool W16>
__device__ __forceinline__ void pkr_row_start(PkRow &R, int i, uint32_t *col, uint32_t *gcol,
                                              uint32_t *colE, uint32_t *gcolE)
{
    // beg = i - w (:96) drops column beg: with w >= 0, beg >= i - 1 - w after
    // the previous row, so exactly one column per row
    const us2 nb = vmax(R.beg, vsubs(splat(i), R.w));
    const uint32_t mv = (w32(nb) ^ w32(R.beg)) & R.alive;
    if (mv & 0xFFFFu) store_he<LC, W16>(col, gcol, colE, gcolE, (int)(w32(R.beg) & 0xFFFF), 0, 0);
    if (mv >> 16) store_he<LC, W16>(col, gcol, colE, gcolE, (int)(w32(R.beg) >> 16), 1, 0);
    R.beg = bfi16(R.alive, nb, R.beg);
    R.end = vmin(R.end, __builtin_elementwise_add_sat(R.w, splat(i + 1)));   // :97-98 (end <= qlen holds)
    R.nc16 = R.nc16 + u2(w32(vsubs(R.end, R.beg)) & R.alive);
}

// Row end of both jobs (:145-172): band-end entry, global-score column, row
// maximum, z-drop, and the last-row test of the next row (:90).  rm / rj:
// the row maximum per half and its (last) column; rm = 0 when the row's
// maximum is 0.  Scores are below 2^8 (8-bit tier) or 2^11 (W16).
template <bool SYM, int LC, bool W16>
__device__ __forceinline__ void pkr_row_end(PkRow &R, const ExtParams &p, int i, us2 h1, us2 rm,
                                            us2 rj, us2 one, uint32_t *col, uint32_t *gcol,
                                            uint32_t *colE, uint32_t *gcolE)
{
    // h[end] = h1, e[end] = 0; a finished job's half is never read again, so
    // its store needs no guard (end <= qlen keeps it in bounds)
    store_he<LC, W16>(col, gcol, colE, gcolE, (int)(w32(R.end) & 0xFFFF), 0, (uint16_t)(w32(h1) & 0xFFFF));
    store_he<LC, W16>(col, gcol, colE, gcolE, (int)(w32(R.end) >> 16), 1, (uint16_t)(w32(h1) >> 16));
    // score bound of the tier minus one: every maxv - rm - zdrop <= KCL - 1
    constexpr int KCL = W16 ? 2047 : 255;
    const us2 si = splat(i);
    // :149-152: at the query end, gtle = i and gscore = h1 unless gscore > h1
    const uint32_t atq = R.alive & ~nz16(u2(w32(R.end) ^ w32(R.qlen)), one);
    const uint32_t c = atq & ~lt16(h1, R.gsc);
    R.gtle = bfi16(c, si, R.gtle);
    R.gsc = bfi16(c, h1, R.gsc);
    // :158-163: a new maximum
    const uint32_t up = R.alive & lt16(R.maxv, rm);
    // :164-172: z-drop against the current maximum.  drop > zdrop is
    // tested as dd * e < maxv - rm - zdrop with dd = |di - dj|; di saturates
    // at 2 * KCL + 256 and dd at KCL, which keeps each sign and each outcome
    // (maxv - rm - zdrop < KCL), and dd * e within 16 bits: e <= 255 in the
    // 8-bit tier, e <= 31 for W16 (the host's condition for that form).
    const us2 di = vmin(si - R.maxi, splat(2 * KCL + 256));        // i - max_i >= 1
    const ss2 dd = __builtin_bit_cast(ss2, di) - __builtin_bit_cast(ss2, rj - R.maxj);
    const ss2 nd = ss2{0, 0} - dd;
    const us2 ad = vmin(__builtin_bit_cast(us2, __builtin_elementwise_max(dd, nd)), splat(KCL));
    us2 ev;
    if (SYM) {
        ev = splat(min(p.e_del, 255));
    } else {
        const uint32_t pos = w32(__builtin_bit_cast(us2, nd >> ss2{15, 15}));   // di > dj
        ev = bfi16(pos, splat(min(p.e_del, 255)), splat(min(p.e_ins, 255)));
    }
    const int zd_lim = p.zdrop > 0 ? p.zdrop : 0x7FFF;         // zdrop <= 0: never
    const ss2 r = __builtin_elementwise_max(
        __builtin_bit_cast(ss2, R.maxv) - __builtin_bit_cast(ss2, rm) - ss2{(short)zd_lim, (short)zd_lim},
        ss2{0, 0});
    const uint32_t zd = nz16(vsubs(__builtin_bit_cast(us2, r), ad * ev), one);
    // :158-163 updates
    const us2 off = u2(w32(vsubs(si, rj)) | w32(vsubs(rj, si)));   // |j - i|
    R.maxoff = bfi16(up, vmax(R.maxoff, off), R.maxoff);
    R.maxi = bfi16(up, si, R.maxi);
    R.maxj = bfi16(up, rj, R.maxj);
    R.maxv = bfi16(up, rm, R.maxv);
    // :155 (a zero row maximum ends the job), :171 (z-drop), :90 (last row)
    const uint32_t keep = nz16(vmin(rm, vsubs(R.tlen, splat(i + 1))), one);
    R.alive = R.alive & keep & (up | ~zd);
}

// Band narrowing of both halves (:175-180), a 4-column group per trip and
// half.  Forward: first column in [beg, end) whose (h, e) is non-zero, else
// end.  Backward: last column in [beg, end] whose (h, e) is non-zero --
// column end holds (h1, 0), so only a zero h1 searches -- then
// end = min(that + 2, qlen); none found: beg - 1.
template <int LC, bool W16>
__device__ __forceinline__ void pkr_narrow(PkRow &R, us2 h1, const uint32_t *col, const uint32_t *gcol,
                                           const uint32_t *colE, const uint32_t *gcolE)
{
    constexpr int GL = LC / 4;
    // group g of this lane, both planes OR-ed for W16 (a column is non-zero
    // when its h or its e is)
    auto grp = [&](int g) {
        u32x4 v = group_rd<GL>(col, gcol, g);
        if (W16) v |= group_rd<GL>(colE, gcolE, g);
        return v;
    };
    const bool la = (R.alive & 0xFFFFu) != 0, lb = (R.alive >> 16) != 0;
    int begA = (int)(w32(R.beg) & 0xFFFF), begB = (int)(w32(R.beg) >> 16);
    const int endA = (int)(w32(R.end) & 0xFFFF), endB = (int)(w32(R.end) >> 16);
    bool fa = la && begA < endA, fb = lb && begB < endB;
    int ga = begA >> 2, gb = begB >> 2;
    uint32_t ma = 0xFu << (begA & 3), mb = 0xFu << (begB & 3);   // columns >= beg
    while (__builtin_amdgcn_ballot_w64(fa || fb)) {
        // a lane done searching re-reads group 0 (its own index may have run
        // past the state)
        const uint32_t z = nz_bits(grp(fa ? ga : 0), grp(fb ? gb : 0));
        const uint32_t za = z & ma, zb = (z >> 16) & mb;
        if (fa) {
            if (za) { begA = min(4 * ga + (int)__builtin_ctz(za), endA); fa = false; }
            else { ++ga; fa = 4 * ga < endA; if (!fa) begA = endA; }
        }
        if (fb) {
            if (zb) { begB = min(4 * gb + (int)__builtin_ctz(zb), endB); fb = false; }
            else { ++gb; fb = 4 * gb < endB; if (!fb) begB = endB; }
        }
        ma = mb = 0xFu;
    }
    // backward: search from column end - 1 down to beg when h1 == 0
    int ka = endA, kb = endB;                                   // h1 != 0: k = end
    bool sa = la && (w32(h1) & 0xFFFFu) == 0, sb = lb && (w32(h1) >> 16) == 0;
    ka -= sa;
    kb -= sb;
    sa = sa && ka >= begA;
    sb = sb && kb >= begB;
    int gka = max(ka, 0) >> 2, gkb = max(kb, 0) >> 2;
    while (__builtin_amdgcn_ballot_w64(sa || sb)) {
        const uint32_t z = nz_bits(grp(sa ? gka : 0), grp(sb ? gkb : 0));
        // columns of the group in [beg, k]
        const int xa = begA - 4 * gka, xb = begB - 4 * gkb;
        const uint32_t ra = (0xFu >> (3 - min(ka - 4 * gka, 3))) & (0xFu << max(min(xa, 4), 0));
        const uint32_t rb = (0xFu >> (3 - min(kb - 4 * gkb, 3))) & (0xFu << max(min(xb, 4), 0));
        const uint32_t za = z & ra & 0xFu, zb = (z >> 16) & rb & 0xFu;
        if (sa) {
            if (za) { ka = 4 * gka + 31 - (int)__builtin_clz(za); sa = false; }
            else { --gka; ka = 4 * gka + 3; sa = ka >= begA; if (!sa) ka = begA - 1; }
        }
        if (sb) {
            if (zb) { kb = 4 * gkb + 31 - (int)__builtin_clz(zb); sb = false; }
            else { --gkb; kb = 4 * gkb + 3; sb = kb >= begB; if (!sb) kb = begB - 1; }
        }
    }
    const int qa = (int)(w32(R.qlen) & 0xFFFF), qb = (int)(w32(R.qlen) >> 16);
    const us2 nbeg = us2{(unsigned short)begA, (unsigned short)begB};
    const us2 nend = us2{(unsigned short)min(ka + 2, qa), (unsigned short)min(kb + 2, qb)};
    R.beg = bfi16(R.alive, nbeg, R.beg);
    R.end = bfi16(R.alive, nend, R.end);
}

}  // namespace

// Extends jobs idx[0..count) (count <= 128) with one wave, two per lane.
// Requirements (checked by the caller): 8-bit tier for every job, qlen <= 255,
// no N base, the standard A/-B matrix, LDS for 16 * ceil((qmax + 1) / 16)
// columns (at most LB blocks), qmax <= 16 * NQ - 1.  SYM: o_del == o_ins and
// e_del == e_ins (one shared gap-open term).  `col` is this lane's word of
// column 0 (group layout, see cpos) and `gcol` its word of column 16 * LB in
// the wave's HBM region (used when LB < NQ).
template <bool SYM, int NQ, int LB, bool W16>
__device__ void ext_wave_pk(const ExtDev &d, const ExtParams &p, const PkMeta *meta, int count,
                            uint32_t *col, uint32_t *gcol)
{
    constexpr bool HYB = LB < NQ;
    constexpr int LC = HYB ? LB * 16 : 0;                      // ColRef: columns in LDS
    constexpr int GL = LC / 4;                                 // groups in LDS (0: all)
    // W16: the e plane follows the h plane (LDS: after the LDS groups; HBM:
    // after the wave's h region)
    uint32_t *colE = W16 ? col + (size_t)(HYB ? LB : NQ) * 1024 : nullptr;
    uint32_t *gcolE = W16 && HYB ? gcol + ext_hyb_words(NQ, LB) : nullptr;
    const int lane = threadIdx.x & 63;
    PkRow R;
    pkr_setup(R, d, p, meta, lane, count);
    const int qlA = (int)(w32(R.qlen) & 0xFFFF), qlB = (int)(w32(R.qlen) >> 16);
    const int h0A = (int)(w32(R.h0) & 0xFFFF), h0B = (int)(w32(R.h0) >> 16);
    const bool liveA = (R.alive & 0xFFFFu) != 0, liveB = (R.alive >> 16) != 0;
    // Per 16-base query word, the "differs from base c" bit masks of every
    // c, held in registers for the whole item: N01 = ne0 | ne1 << 16 and
    // N23 = ne2 | ne3 << 16.  A row picks its target base's mask with one
    // select and one shift per job and block.
    uint32_t NA01[NQ], NA23[NQ], NB01[NQ], NB23[NQ];
#pragma unroll
    for (int b = 0; b < NQ; ++b) {
        const uint32_t qa = b * 16 < qlA ? d.seq[R.qw0[0] + b] : 0u;
        const uint32_t qb = b * 16 < qlB ? d.seq[R.qw0[1] + b] : 0u;
        NA01[b] = mismatch16(qa, 0u) | (mismatch16(qa, 0x55555555u) << 16);
        NA23[b] = mismatch16(qa, 0xAAAAAAAAu) | (mismatch16(qa, 0xFFFFFFFFu) << 16);
        NB01[b] = mismatch16(qb, 0u) | (mismatch16(qb, 0x55555555u) << 16);
        NB23[b] = mismatch16(qb, 0xAAAAAAAAu) | (mismatch16(qb, 0xFFFFFFFFu) << 16);
    }
    const int qmax = uniform_max(max(liveA ? qlA : 0, liveB ? qlB : 0));
    for (int j = 0; j <= qmax; ++j) {                          // first row, :68-77
        const int ha = (liveA && j <= qlA) ? (j == 0 ? h0A : max(0, h0A - p.o_ins - p.e_ins * j)) : 0;
        const int hb = (liveB && j <= qlB) ? (j == 0 ? h0B : max(0, h0B - p.o_ins - p.e_ins * j)) : 0;
        // 8-bit: bytes (hA, eA = 0, hB, eB = 0); W16: the h plane word
        const uint32_t v = (uint32_t)ha | ((uint32_t)hb << 16);
        // j is wave-uniform: one LDS or one global store (no generic access)
        if (!HYB || j < LC) {
            ((__attribute__((address_space(3))) uint32_t *)col)[cpos<uint32_t>(j)] = v;
            if (W16) ((__attribute__((address_space(3))) uint32_t *)colE)[cpos<uint32_t>(j)] = 0u;
        } else {
            ((__attribute__((address_space(1))) uint32_t *)gcol)[cpos<uint32_t>(j - LC)] = v;
            if (W16) ((__attribute__((address_space(1))) uint32_t *)gcolE)[cpos<uint32_t>(j - LC)] = 0u;
        }
    }
    const ss2 Ssc = ss2{(short)p.match_sc, (short)p.match_sc};
    const ss2 nAB = ss2{(short)(p.mis_sc - p.match_sc), (short)(p.mis_sc - p.match_sc)};
    const us2 eD = splat(p.e_del), eI = splat(p.e_ins);
    const ss2 soeD = ss2{(short)p.oe_del, (short)p.oe_del}, soeI = ss2{(short)p.oe_ins, (short)p.oe_ins};
    const uint32_t LOW = 0x00FF00FFu;
    const us2 ONE = u2(opaque(0x00010001u));
    const us2 C256 = u2(opaque(0x01000100u));   // H * 256 + k: one packed mad
    const us2 C16 = u2(opaque(0x00100010u));    // W16: H * 16 + k
    // Rows whose band-start h1 (:101-104) can be non-zero for some job of the
    // wave: h0 - o_del - e_del * (i + 1) > 0 needs i < ceil((h0 - o_del) / e_del).
    int h1rows;
    {
        const int ra = liveA ? max(0, (h0A - p.o_del + p.e_del - 1) / p.e_del) : 0;
        const int rb = liveB ? max(0, (h0B - p.o_del + p.e_del - 1) / p.e_del) : 0;
        uint32_t x = ~(uint32_t)max(ra, rb), y = ~0u, z = ~0u;
        wave_min3(x, y, z);
        h1rows = (int)~x;
    }

    for (int i = 0;; ++i) {
        if (i > 0 && (i & 15) == 0) {
            // 16-row block i/16 starts: fetch block i/16 + 1 into the register
            // of block i/16 - 1 (it is first read 16 rows from now)
            const int nb = (i >> 4) + 1;
            if (nb & 1) {
                R.two[0] = d.seq[R.tw0[0] + min((uint32_t)nb, R.tnw[0])];
                R.two[1] = d.seq[R.tw0[1] + min((uint32_t)nb, R.tnw[1])];
            } else {
                R.twe[0] = d.seq[R.tw0[0] + min((uint32_t)nb, R.tnw[0])];
                R.twe[1] = d.seq[R.tw0[1] + min((uint32_t)nb, R.tnw[1])];
            }
        }
        pkr_row_start<LC, W16>(R, i, col, gcol, colE, gcolE);
        // wave band: lowest live start, highest and lowest live end (each
        // lane's two halves folded first); no live job leaves the first at
        // 0xFFFF
        uint32_t rlo = hmin16(w32(R.beg) | ~R.alive);
        uint32_t rhi = hmin16(~(w32(R.end) & R.alive));
        uint32_t rmh = hmin16(w32(R.end) | ~R.alive);
        wave_min3(rlo, rhi, rmh);
        if ((rlo & 0xFFFFu) == 0xFFFFu) break;                 // every job has ended
        const int jlo = (int)(rlo & 0xFFFFu), jhi = (int)(~rhi & 0xFFFFu), mhi = (int)(rmh & 0xFFFFu);
        // this row's target base of each job: mask register (N23 for G/T) and shift
        const uint32_t tbA = (((i >> 4) & 1 ? R.two[0] : R.twe[0]) >> ((i & 15) * 2)) & 3u;
        const uint32_t tbB = (((i >> 4) & 1 ? R.two[1] : R.twe[1]) >> ((i & 15) * 2)) & 3u;
        const bool hiA = (tbA & 2u) != 0, hiB = (tbB & 2u) != 0;
        const uint32_t shA = (tbA & 1u) * 16, shB = (tbB & 1u) * 16;
        us2 h1 = splat(0);
        if (i < h1rows) {                                      // :99-107, beg == 0 only
            const us2 t = vsubs(R.h0, splat(min(p.o_del + p.e_del * (i + 1), 65535)));
            h1 = u2(w32(t) & R.alive & ~nz16(R.beg, ONE));
        }
        us2 f = splat(0), key = splat(0);
        uint32_t keyA = 0, keyB = 0;                           // W16: H * 256 + j per job
        // per-half band end (columns >= it commit nothing); 0 for a finished job
        const us2 L = u2(w32(R.end) & R.alive);

        // Each row sweeps the 4-column groups from the wave's lowest band start
        // to its highest band end in blocks of 16 columns (the last one 4, 8
        // or 12 wide).  A block starts at column 16 q + o of query word q, o in
        // {0, 4, 8, 12} the same for the whole row, so its 32 mismatch bits
        // (16 per job) come from the masks of words q and q + 1.
        auto qmask = [&](int b) {   // packed "differs" bits of query word b, this row's target bases
            const uint32_t a01 = NA01[b], a23 = NA23[b], b01 = NB01[b], b23 = NB23[b];
            const uint32_t ma = (hiA ? a23 : a01) >> shA;
            const uint32_t mb = (hiB ? b23 : b01) >> shB;
            return __builtin_amdgcn_perm(mb, ma, 0x05040100u);   // ma.lo16 | mb.lo16 << 16
        };
        // One block of N columns from c0, in three passes so that the packed
        // ops of one cell need not wait on each other: (1) everything a cell
        // derives from its own state word and base pair -- diag (:117-119),
        // max(diag, E) and the new E (:128-129); (2) the running F and h1
        // (:121-123, :130-131); (3) the row-maximum candidates H * 256 + k
        // (:124-125).  MASKED: commit only left of each half's band end.
        auto block = [&](auto NC, auto MK, auto KD, uint32_t (&wv)[16], uint32_t (&we)[16], int c0,
                         uint32_t mm) {
            constexpr int N = decltype(NC)::value;
            constexpr bool MASKED = decltype(MK)::value;
            constexpr int KIND = decltype(KD)::value;   // 0 LDS, 2 HBM, 1 either (per group)
            us2 dgE[16], tf[16], ne[16], Hs[16], M[16], E[16], s[16];
#pragma unroll
            for (int k = 0; k < N; ++k) {
                const uint32_t ab = (mm >> k) & 0x00010001u;          // per-half mismatch flag
                s[k] = __builtin_bit_cast(us2, __builtin_bit_cast(ss2, ab) * nAB + Ssc);
                if (W16) {
                    M[k] = u2(wv[k]);
                    E[k] = u2(we[k]);
                } else {
                    M[k] = u2(wv[k] & LOW);
                    E[k] = u2(__builtin_amdgcn_perm(0u, wv[k], 0x0C030C01u));   // bytes 1, 3
                }
            }
#pragma unroll
            for (int k = 0; k < N; ++k) s[k] = s[k] * vmin(M[k], ONE) + M[k];   // M ? M + s : 0
#pragma unroll
            for (int k = 0; k < N; ++k) {
                // M' = M ? M + s : 0 may be negative here (a mismatch after a
                // small M); ksw_extend2's max(M', 0) is left out because every
                // use of M' is a max with a non-negative term: H = max(M', E, F),
                // E' = max(E - e_del, M' - oe_del, 0) with E - e_del saturating
                // at 0, likewise F'.  So M' - oe stays signed and the maxes are
                // signed 16-bit ones.
                const ss2 dg = __builtin_bit_cast(ss2, s[k]);
                dgE[k] = smax(dg, E[k]);
                const ss2 t = dg - soeD;
                ne[k] = smax(t, vsubs(E[k], eD));
                tf[k] = __builtin_bit_cast(us2, SYM ? t : dg - soeI);
            }
            if (!MASKED) {
#pragma unroll
                for (int k = 0; k < N; ++k) {
                    const us2 H = vmax(dgE[k], f);
                    f = smax(__builtin_bit_cast(ss2, tf[k]), vsubs(f, eI));
                    if (W16) {
                        wv[k] = w32(h1);
                        we[k] = w32(ne[k]);
                    } else {
                        wv[k] = __builtin_amdgcn_perm(w32(ne[k]), w32(h1), 0x06020400u);
                    }
                    h1 = H;
                    Hs[k] = H * (W16 ? C16 : C256) + splat(k);
                }
            } else {
                const us2 Lb = vsubs(L, splat(c0));
                uint32_t hms[16];
#pragma unroll
                for (int k = 0; k < N; ++k) {
                    hms[k] = w32(__builtin_bit_cast(
                        us2, __builtin_bit_cast(ss2, splat(k) - Lb) >> ss2{15, 15}));
                    asm volatile("" : "+v"(hms[k]));   // keep the selects bitwise
                }
#pragma unroll
                for (int k = 0; k < N; ++k) {
                    const uint32_t hm = hms[k];
                    const us2 H = vmax(dgE[k], f);
                    f = smax(__builtin_bit_cast(ss2, tf[k]), vsubs(f, eI));
                    if (W16) {
                        wv[k] = (w32(h1) & hm) | (wv[k] & ~hm);
                        we[k] = (w32(ne[k]) & hm) | (we[k] & ~hm);
                    } else {
                        const uint32_t nw = __builtin_amdgcn_perm(w32(ne[k]), w32(h1), 0x06020400u);
                        wv[k] = (nw & hm) | (wv[k] & ~hm);
                    }
                    h1 = u2((w32(H) & hm) | (w32(h1) & ~hm));
                    Hs[k] = u2(w32(H * (W16 ? C16 : C256) + splat(k)) & hm);
                }
            }
            if constexpr (KIND == 1) {
                store_groups<GL, N / 4>(col, gcol, c0 >> 2, wv);
                if (W16) store_groups<GL, N / 4>(colE, gcolE, c0 >> 2, we);
            } else {
                store_groups_k<KIND, GL, N / 4>(col, gcol, c0 >> 2, wv);
                if (W16) store_groups_k<KIND, GL, N / 4>(colE, gcolE, c0 >> 2, we);
            }
#pragma unroll
            for (int st = 8; st >= 1; st >>= 1)
#pragma unroll
                for (int k = 0; k < st; ++k)
                    if (k + st < N) Hs[k] = vmax(Hs[k], Hs[k + st]);
            if (W16) {
                // the block's H * 16 + k per job, as H * 256 + (c0 + k)
                const uint32_t bk = w32(Hs[0]);
                const uint32_t ka = ((bk & 0xFFF0u) << 4) | (c0 + (bk & 15u));
                const uint32_t kb = ((bk >> 12) & 0xFFF00u) | (c0 + ((bk >> 16) & 15u));
                keyA = max(keyA, ka);
                keyB = max(keyB, kb);
            } else {
                key = vmax(key, Hs[0] + splat(c0));
            }
        };
        using N16 = std::integral_constant<int, 16>;
        using Unmasked = std::integral_constant<bool, false>;
        using Masked = std::integral_constant<bool, true>;
        using KAny = std::integral_constant<int, 1>;
        // a block's column words: the state words (8-bit) or the h plane
        // (W16) in h, the e plane (W16 only) in e
        struct WSet {
            uint32_t h[16], e[16];       // e unused (and dropped) in the 8-bit tier
        };
        auto full = [&](auto KD, WSet &w, int c0, uint32_t mm) {
            if (c0 + 16 <= mhi) block(N16{}, Unmasked{}, KD, w.h, w.e, c0, mm);   // left of every live band end
            else block(N16{}, Masked{}, KD, w.h, w.e, c0, mm);
        };
        // partial block of r = 1..3 groups (always at a band end or a memory boundary)
        auto part = [&](auto KD, int r, WSet &w, int c0, uint32_t mm) {
            if (r == 1) block(std::integral_constant<int, 4>{}, Masked{}, KD, w.h, w.e, c0, mm);
            else if (r == 2) block(std::integral_constant<int, 8>{}, Masked{}, KD, w.h, w.e, c0, mm);
            else block(std::integral_constant<int, 12>{}, Masked{}, KD, w.h, w.e, c0, mm);
        };
        // groups g .. g + 3 (clamped to glast) of both planes; KIND as in block
        auto ld = [&](auto KD, int g, int glast, WSet &w) {
            constexpr int KIND = decltype(KD)::value;
            if constexpr (KIND == 1) {
                load_groups<GL>(col, gcol, g, glast, w.h);
                if constexpr (W16) load_groups<GL>(colE, gcolE, g, glast, w.e);
            } else {
                load_groups_k<KIND, GL>(col, gcol, g, glast, w.h);
                if constexpr (W16) load_groups_k<KIND, GL>(colE, gcolE, g, glast, w.e);
            }
        };
        if (jlo < jhi) {
            const int g0 = jlo >> 2, gend = (jhi + 3) >> 2;
            int o = (g0 & 3) * 4;
            us2 osh = splat(o), ish = splat((16 - o) & 15);
            uint32_t ohi = o ? 0xFFFFFFFFu : 0u;
            int q = g0 >> 2;
            uint32_t mcur = qmask(q);
            // mismatch bits of the block at column 16 q + o; moves to the next word
            auto next_mm = [&]() {
                const uint32_t mnext = qmask(min(q + 1, NQ - 1));
                const uint32_t m = w32(u2(mcur) >> osh) | (w32(u2(mnext) << ish) & ohi);
                mcur = mnext;
                ++q;
                return m;
            };
            // Blocks of four groups from g to gstop in one memory: two word
            // sets in turn, the next block's reads in flight while the current
            // one computes (each read clamped to gstop - 1, so a read past the
            // phase repeats its last group: same memory, unused); then the
            // partial block of the remaining 1-3 groups, whose words the last
            // trip has already read.  W0 holds the first block on entry.
            auto phase = [&](auto KD, int &g, int gstop, WSet &W0, WSet &W1) {
                const int gl = gstop - 1;
                const int nf = (gstop - g) >> 2, r = (gstop - g) & 3;
                bool in_w1 = false;
                for (int t = 0; t < nf;) {
                    ld(KD, min(g + 4, gl), gl, W1);
                    full(KD, W0, 4 * g, next_mm());
                    g += 4;
                    if (++t >= nf) {
                        in_w1 = true;
                        break;
                    }
                    ld(KD, min(g + 4, gl), gl, W0);
                    full(KD, W1, 4 * g, next_mm());
                    g += 4;
                    ++t;
                }
                if (r) {
                    if (in_w1) W0 = W1;
                    part(KD, r, W0, 4 * g, next_mm());
                    g += r;
                }
            };
            WSet W0, W1;
            int g = g0;
            if constexpr (!HYB) {
                ld(KAny{}, g, gend - 1, W0);
                phase(KAny{}, g, gend, W0, W1);
            } else {
                // LDS groups [g0, min(GL, gend)) first, then HBM groups
                // [max(g0, GL), gend); the HBM phase's first block is read
                // into W2 before the LDS phase so its latency hides there.
                using KL = std::integral_constant<int, 0>;
                using KH = std::integral_constant<int, 2>;
                WSet W2;
                const int gh = max(g0, GL);
                // Retire the previous row's global accesses (its last HBM trip
                // read a block it did not use) so that W2's reads are the only
                // ones in flight; otherwise the LDS phase's first overwrite of
                // such a register waits for W2 as well (vmcnt counts in order).
                // The row-end narrowing has waited on its own reads already,
                // so this wait is normally free.  0x0F70: vmcnt(0), others
                // unchanged.
                __builtin_amdgcn_s_waitcnt(0x0F70);
                if (gend > GL) ld(KH{}, gh, gend - 1, W2);
                if (g0 < GL) {
                    const int gs = min(GL, gend);
                    ld(KL{}, g, gs - 1, W0);
                    phase(KL{}, g, gs, W0, W1);
                }
                if (gend > GL) {
                    if (g0 < GL) {
                        // g == GL: 16-column aligned, the mask offset is 0
                        o = 0;
                        osh = splat(0); ish = splat(0);
                        ohi = 0u;
                        q = GL >> 2;
                        mcur = qmask(q);
                    }
                    W0 = W2;
                    phase(KH{}, g, gend, W0, W1);
                }
            }
        }

        us2 rm, rj;
        if (W16) {
            rm = us2{(unsigned short)(keyA >> 8), (unsigned short)(keyB >> 8)};
            rj = us2{(unsigned short)(keyA & 0xFFu), (unsigned short)(keyB & 0xFFu)};
        } else {
            rm = key >> splat(8);
            rj = u2(w32(key) & 0x00FF00FFu);
        }
        pkr_row_end<SYM, LC, W16>(R, p, i, h1, rm, rj, ONE, col, gcol, colE, gcolE);
        pkr_narrow<LC, W16>(R, h1, col, gcol, colE, gcolE);
        if ((i & 255) == 255) {                                // keep each half of nc16 below 65536
            R.ncells += (w32(R.nc16) & 0xFFFFu) + (w32(R.nc16) >> 16);
            R.nc16 = splat(0);
        }
    }
    if (d.cells) {
        unsigned long long c = R.ncells + (w32(R.nc16) & 0xFFFFu) + (w32(R.nc16) >> 16);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
        if (lane == 0) atomicAdd(d.cells, c);
    }
    pkr_finish(R, d, meta, lane, count, 0);
    pkr_finish(R, d, meta, lane, count, 1);
}

// Extends jobs meta[0..count) (count <= 64) with one wave, one per lane, for
// the 8- and 16-bit tiers: the column word is (h | e << 16) in 16-bit fields
// and the arithmetic 32-bit.  Same structure as the packed form (band edges,
// query masks in registers, three passes per block, 4-column LDS groups, two
// word sets in turn); requirements as there with qmax <= 16 * NQ - 1 and
// scores below 65536.
template <bool SYM, int NQ, int LB>
__device__ void ext_wave_ln(const ExtDev &d, const ExtParams &p, const PkMeta *meta, int count,
                            uint32_t *col, uint32_t *gcol)
{
    constexpr bool HYB = LB < NQ;
    constexpr int LC = HYB ? LB * 16 : 0;
    const ColRef<uint32_t, LC> cw{col, gcol};
    const int lane = threadIdx.x & 63;
    PkJob A, B;
    pk_setup(A, d, p, meta, lane, count);
    B.valid = B.nonempty = B.alive = false;                    // no second job: narrowing skips it
    B.beg = B.end = B.qlen = 0;
    uint32_t NA01[NQ], NA23[NQ];
#pragma unroll
    for (int b = 0; b < NQ; ++b) {
        const uint32_t qa = b * 16 < A.qlen ? d.seq[A.qw0 + b] : 0u;
        NA01[b] = mismatch16(qa, 0u) | (mismatch16(qa, 0x55555555u) << 16);
        NA23[b] = mismatch16(qa, 0xAAAAAAAAu) | (mismatch16(qa, 0xFFFFFFFFu) << 16);
    }
    const int qmax = uniform_max(A.alive ? A.qlen : 0);
    for (int j = 0; j <= qmax; ++j)                            // first row, :68-77
        cw[j] = (A.alive && j <= A.qlen)
                                     ? (uint32_t)(j == 0 ? A.h0 : max(0, A.h0 - p.o_ins - p.e_ins * j))
                                     : 0u;
    const int sA = p.match_sc, nAB = p.mis_sc - p.match_sc;
    const uint32_t eD = p.e_del, eI = p.e_ins, oeD = p.oe_del, oeI = p.oe_ins;
    const uint32_t ONE = opaque(1u);
    const int BIG = 0x7fffffff;

    for (int i = 0;; ++i) {
        if (i > 0 && (i & 15) == 0) {
            const int nb = (i >> 4) + 1;
            if (nb & 1) A.two = d.seq[A.tw0 + min(nb, A.tnw)];
            else A.twe = d.seq[A.tw0 + min(nb, A.tnw)];
        }
        uint32_t h1 = (uint32_t)pk_row_start(A, p, i, cw);
        if (__ballot(A.alive) == 0) break;
        uint32_t f = 0, key = 0;
        const int jlo = uniform_min(A.alive ? A.beg : BIG);
        const int jhi = uniform_max(A.alive ? A.end : 0);
        const int mhi = uniform_min(A.alive ? A.end : BIG);
        const int L = A.alive ? A.end : 0;
        const bool hiA = (A.tb & 2u) != 0;
        const uint32_t shA = (A.tb & 1u) * 16;

        auto block = [&](uint32_t (&wv)[16], int b) {
            const uint32_t a01 = NA01[b], a23 = NA23[b];
            const uint32_t mm = (hiA ? a23 : a01) >> shA;          // bit k: base c0 + k differs
            const int c0 = b * 16;
            uint32_t dgE[16], tf[16], ne[16], Hs[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const uint32_t M = wv[k] & 0xFFFFu, E = wv[k] >> 16;
                const int s = (int)((mm >> k) & 1u) * nAB + sA;
                const uint32_t dg = (uint32_t)max(s * (int)min(M, ONE) + (int)M, 0);   // :117-119
                dgE[k] = max(dg, E);
                const uint32_t t = __builtin_elementwise_sub_sat(dg, oeD);
                ne[k] = max(__builtin_elementwise_sub_sat(E, eD), t);
                tf[k] = SYM ? t : __builtin_elementwise_sub_sat(dg, oeI);
            }
            if (c0 + 16 <= mhi) {
#pragma unroll
                for (int k = 0; k < 16; ++k) {
                    const uint32_t H = max(dgE[k], f);
                    f = max(__builtin_elementwise_sub_sat(f, eI), tf[k]);
                    wv[k] = h1 | (ne[k] << 16);
                    h1 = H;
                    Hs[k] = (H << 8) | (uint32_t)k;
                }
            } else {
                const int lb = L - c0;                           // cells k < lb commit
#pragma unroll
                for (int k = 0; k < 16; ++k) {
                    const bool in = k < lb;
                    const uint32_t H = max(dgE[k], f);
                    f = max(__builtin_elementwise_sub_sat(f, eI), tf[k]);
                    wv[k] = in ? (h1 | (ne[k] << 16)) : wv[k];
                    h1 = in ? H : h1;
                    Hs[k] = in ? ((H << 8) | (uint32_t)k) : 0u;
                }
            }
            store_block<HYB, LB>(col, gcol, b, wv);
#pragma unroll
            for (int k = 0; k < 8; ++k) Hs[k] = max(Hs[k], Hs[k + 8]);
#pragma unroll
            for (int k = 0; k < 4; ++k) Hs[k] = max(Hs[k], Hs[k + 4]);
            key = max(key, max(max(Hs[0], Hs[2]), max(Hs[1], Hs[3])) + (uint32_t)c0);
        };
        auto load = [&](uint32_t (&wv)[16], int b) { load_block<HYB, LB>(col, gcol, b, wv); };
        if (jlo < jhi) {
            const int bb = jlo >> 4, be = (jhi - 1) >> 4;
            uint32_t W0[16], W1[16];
            load(W0, bb);
            for (int b = bb;; b += 2) {
                load(W1, min(b + 1, be));
                block(W0, b);
                if (b + 1 > be) break;
                load(W0, min(b + 2, be));
                block(W1, b + 1);
                if (b + 2 > be) break;
            }
        }
        pk_row_end(A, p, i, (int)h1, (int)key, cw);
        pk_narrow(A, B, (int)h1, 1, cw, cw);
    }
    if (d.cells) {
        unsigned long long c = A.ncells;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
        if (lane == 0) atomicAdd(d.cells, c);
    }
    pk_finish(A, d);
}

// Persistent kernel of the one-job-per-lane form: items of up to 64 jobs, no
// N base (the planner bins N-holding jobs apart).  With a null counter every
// block runs the one item of its index (the input-order form).
template <bool SYM, int NQ, int LB>
__global__ __launch_bounds__(64) void ext_persistent_ln_kernel(ExtDev d, ExtParams p,
                                                               const WorkItem *items, int n_items,
                                                               uint32_t *counter, uint32_t *gscratch)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    uint32_t *col = reinterpret_cast<uint32_t *>(lds_raw) + threadIdx.x * 4;
    uint32_t *gcol = gscratch ? gscratch + (size_t)blockIdx.x * ext_hyb_words(NQ, LB) + threadIdx.x * 4
                              : nullptr;
    for (int it = 0;; ++it) {
        int item = 0;
        if (counter) {                 // work stealing
            if (threadIdx.x == 0) item = (int)atomicAdd(counter, 1u);
            item = __builtin_amdgcn_readfirstlane(item);
        } else {                       // static: block b runs item b only
            item = it == 0 ? (int)blockIdx.x : n_items;
        }
        if (item >= n_items) break;
        const WorkItem wi = items[item];
        ext_wave_ln<SYM, NQ, LB>(d, p, d.pk + wi.start, (int)wi.count, col, gcol);
    }
}

// Persistent kernel of the packed form: items of up to 128 jobs, no N base
// (the planner bins N-holding jobs apart).  The block's LDS is exactly the
// column state (no static shared memory), so a CU holds as many waves as their
// LDS shares allow (150-bp queries: four held whole, six as 96 + 64 columns).
// The 8-bit form is held to 256 registers, two waves per SIMD (its 16-query-
// word instances would otherwise take 257 and run one); the 16-bit form needs
// 294-316 and runs one.
template <bool SYM, int NQ, int LB, bool W16>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(!W16 && NQ == 16 ? 2 : 1)))
void ext_persistent_pk_kernel(ExtDev d, ExtParams p,
                                                               const WorkItem *items, int n_items,
                                                               uint32_t *counter, uint32_t *gscratch)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    uint32_t *col = reinterpret_cast<uint32_t *>(lds_raw) + threadIdx.x * 4;
    uint32_t *gcol = gscratch ? gscratch + (size_t)blockIdx.x * ext_hyb_words(NQ, LB) * (W16 ? 2 : 1) + threadIdx.x * 4
                              : nullptr;
    for (int it = 0;; ++it) {
        int item = 0;
        if (counter) {                 // work stealing
            if (threadIdx.x == 0) item = (int)atomicAdd(counter, 1u);
            item = __builtin_amdgcn_readfirstlane(item);
        } else {                       // static: block b runs item b only
            item = it == 0 ? (int)blockIdx.x : n_items;
        }
        if (item >= n_items) break;
        const WorkItem wi = items[item];
        ext_wave_pk<SYM, NQ, LB, W16>(d, p, d.pk + wi.start, (int)wi.count, col, gcol);
    }
}

void build_pk_meta(int n, const uint32_t *order, const uint32_t *q_word, const uint32_t *t_word,
                   const uint32_t *qn_word, const uint32_t *tn_word, const int32_t *q_len,
                   const int32_t *t_len, const int32_t *h0, const int32_t *w, const int32_t *eb,
                   PkMeta *out)
{
    for (int s = 0; s < n; ++s) {
        const uint32_t k = order[s];
        out[s] = PkMeta{k, q_word[k], t_word[k], (uint16_t)q_len[k], (uint16_t)t_len[k],
                        h0[k], w[k], eb[k], (uint32_t)(qn_word[k] != kNoN || tn_word[k] != kNoN)};
    }
}

// Query words of the kernel instance that runs a bin whose longest query needs
// `cols` columns: 4, 10 or 16.
int ext_pk_nq(int cols)
{
    const int nq = (cols + 15) / 16;
    return nq <= 4 ? 4 : nq <= 10 ? 10 : 16;
}

int ext_pk_default_lb(int cols) { return std::min(ext_pk_nq(cols), 10); }

// The 16-bit packed form keeps two planes (8 KiB of LDS per 16-column
// block), so its default holds at most 5 blocks in LDS.
int ext_pk_default_lb_k(int kind, int cols)
{
    return kind == EXT_PACKED16 ? std::min(ext_pk_nq(cols), 5) : ext_pk_default_lb(cols);
}

// Instances built: blocks held in LDS per query-word class.
bool ext_pk_lb_supported(int kind, int cols, int lb)
{
    if (cols < 1 || cols > 256) return false;
    const int nq = ext_pk_nq(cols);
    if (kind == EXT_PACKED16)
        return (nq == 4 && (lb == 4 || lb == 2)) || (nq > 4 && (lb == 5 || lb == 3 || lb == 2));
    if (lb == ext_pk_default_lb(cols)) return true;
    if (kind != EXT_PACKED && kind != EXT_LANE) return false;
    return (nq == 10 && (lb == 5 || lb == 6)) || (nq == 16 && (lb == 6 || lb == 8));
}

int ext_pk_lb_candidates(int kind, int cols, int *out)
{
    static const int cand[] = {10, 8, 6, 5, 4, 3, 2};
    int n = 0;
    for (int lb : cand)
        if (n < 5 && ext_pk_lb_supported(kind, cols, lb)) out[n++] = lb;
    return n;
}

size_t ext_pk_lds_bytes(int cols, int lb)
{
    return (size_t)std::min((cols + 15) / 16 * 16, 16 * lb) * 64 * 4;
}

size_t ext_pk_hbm_bytes_per_wave(int cols, int lb)
{
    return (cols + 15) / 16 > lb ? ext_hyb_words(ext_pk_nq(cols), lb) * 4 : 0;
}

size_t ext_pk_lds_bytes_k(int kind, int cols, int lb)
{
    // the 16-bit form's e plane starts a whole h plane (min(NQ, LB) blocks of
    // 4 KiB) after the h plane, whatever the bin's longest query
    if (kind == EXT_PACKED16) return (size_t)2 * std::min(ext_pk_nq(cols), lb) * 4096;
    return ext_pk_lds_bytes(cols, lb);
}

size_t ext_pk_hbm_bytes_per_wave_k(int kind, int cols, int lb)
{
    return ext_pk_hbm_bytes_per_wave(cols, lb) * (kind == EXT_PACKED16 ? 2 : 1);
}

#define HIPSW_EXT_LAUNCH(KER, S, Q, B, ...)                                                     \
    hipLaunchKernelGGL((KER<S, Q, B __VA_ARGS__>), dim3(grid), dim3(64), lds, s, d, p, items,   \
                       n_items, counter, gs)
#define HIPSW_EXT_SYM(KER, Q, B, ...)                                                           \
    do {                                                                                        \
        if (sym) HIPSW_EXT_LAUNCH(KER, true, Q, B, __VA_ARGS__);                               \
        else HIPSW_EXT_LAUNCH(KER, false, Q, B, __VA_ARGS__);                                  \
    } while (0)

hipError_t launch_ext_persistent_ln(const ExtDev &d, const ExtParams &p, const WorkItem *items,
                                    int n_items, uint32_t *counter, int cols, int lb, int grid,
                                    void *gscratch, hipStream_t s)
{
    if (n_items <= 0) return hipSuccess;
    if (!ext_pk_lb_supported(EXT_LANE, cols, lb)) return hipErrorInvalidValue;
    if (ext_pk_hbm_bytes_per_wave(cols, lb) && !gscratch) return hipErrorInvalidValue;
    const bool sym = p.o_del == p.o_ins && p.e_del == p.e_ins;
    const int nq = ext_pk_nq(cols);
    const size_t lds = ext_pk_lds_bytes(cols, lb);
    uint32_t *gs = static_cast<uint32_t *>(gscratch);
    if (nq == 4) HIPSW_EXT_SYM(ext_persistent_ln_kernel, 4, 4);
    else if (nq == 10 && lb == 10) HIPSW_EXT_SYM(ext_persistent_ln_kernel, 10, 10);
    else if (nq == 10 && lb == 6) HIPSW_EXT_SYM(ext_persistent_ln_kernel, 10, 6);
    else if (nq == 10) HIPSW_EXT_SYM(ext_persistent_ln_kernel, 10, 5);
    else if (lb == 10) HIPSW_EXT_SYM(ext_persistent_ln_kernel, 16, 10);
    else if (lb == 8) HIPSW_EXT_SYM(ext_persistent_ln_kernel, 16, 8);
    else HIPSW_EXT_SYM(ext_persistent_ln_kernel, 16, 6);
    return hipGetLastError();
}

hipError_t launch_ext_persistent_pk(const ExtDev &d, const ExtParams &p, const WorkItem *items,
                                    int n_items, uint32_t *counter, int cols, int lb, int grid,
                                    void *gscratch, hipStream_t s)
{
    if (n_items <= 0) return hipSuccess;
    if (!ext_pk_lb_supported(EXT_PACKED, cols, lb)) return hipErrorInvalidValue;
    if (ext_pk_hbm_bytes_per_wave(cols, lb) && !gscratch) return hipErrorInvalidValue;
    uint32_t *gs = static_cast<uint32_t *>(gscratch);
    const bool sym = p.o_del == p.o_ins && p.e_del == p.e_ins;
    const int nq = ext_pk_nq(cols);             // query words of the longest query
    const size_t lds = ext_pk_lds_bytes(cols, lb);
    if (nq == 4) HIPSW_EXT_SYM(ext_persistent_pk_kernel, 4, 4, , false);
    else if (nq == 10 && lb == 10) HIPSW_EXT_SYM(ext_persistent_pk_kernel, 10, 10, , false);
    else if (nq == 10 && lb == 6) HIPSW_EXT_SYM(ext_persistent_pk_kernel, 10, 6, , false);
    else if (nq == 10) HIPSW_EXT_SYM(ext_persistent_pk_kernel, 10, 5, , false);
    else if (lb == 10) HIPSW_EXT_SYM(ext_persistent_pk_kernel, 16, 10, , false);
    else if (lb == 8) HIPSW_EXT_SYM(ext_persistent_pk_kernel, 16, 8, , false);
    else HIPSW_EXT_SYM(ext_persistent_pk_kernel, 16, 6, , false);
    return hipGetLastError();
}

hipError_t launch_ext_persistent_pk16(const ExtDev &d, const ExtParams &p, const WorkItem *items,
                                      int n_items, uint32_t *counter, int cols, int lb, int grid,
                                      void *gscratch, hipStream_t s)
{
    if (n_items <= 0) return hipSuccess;
    if (!ext_pk_lb_supported(EXT_PACKED16, cols, lb)) return hipErrorInvalidValue;
    if (ext_pk_hbm_bytes_per_wave_k(EXT_PACKED16, cols, lb) && !gscratch) return hipErrorInvalidValue;
    uint32_t *gs = static_cast<uint32_t *>(gscratch);
    const bool sym = p.o_del == p.o_ins && p.e_del == p.e_ins;
    const int nq = ext_pk_nq(cols);
    const size_t lds = ext_pk_lds_bytes_k(EXT_PACKED16, cols, lb);
    if (nq == 4 && lb == 4) HIPSW_EXT_SYM(ext_persistent_pk_kernel, 4, 4, , true);
    else if (nq == 4) HIPSW_EXT_SYM(ext_persistent_pk_kernel, 4, 2, , true);
    else if (nq == 10 && lb == 5) HIPSW_EXT_SYM(ext_persistent_pk_kernel, 10, 5, , true);
    else if (nq == 10 && lb == 3) HIPSW_EXT_SYM(ext_persistent_pk_kernel, 10, 3, , true);
    else if (nq == 10) HIPSW_EXT_SYM(ext_persistent_pk_kernel, 10, 2, , true);
    else if (lb == 5) HIPSW_EXT_SYM(ext_persistent_pk_kernel, 16, 5, , true);
    else if (lb == 3) HIPSW_EXT_SYM(ext_persistent_pk_kernel, 16, 3, , true);
    else HIPSW_EXT_SYM(ext_persistent_pk_kernel, 16, 2, , true);
    return hipGetLastError();
}
#undef HIPSW_EXT_SYM
#undef HIPSW_EXT_LAUNCH

}  // namespace hipsw
