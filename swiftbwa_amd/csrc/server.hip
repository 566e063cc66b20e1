// Serving loop: continuous batching of read chunks over a resident reference.
//
// A chunk is the hot-path work of a batch of read pairs: banded extensions,
// mate-rescue local SW and CIGAR NW, whose targets are windows of the reference
// `.pac` kept resident in HBM.  Chunks are submitted as they arrive and return
// a ticket.  Each of the server's S slots owns a HIP stream and a worker thread;
// a slot takes the chunk at the head of the FIFO as soon as the chunk fits the
// HBM budget next to the chunks already running, so work of different chunks
// overlaps on the device.  A chunk's device memory is stream-ordered
// (hipMallocAsync/hipFreeAsync on its slot stream) and counted against the
// budget from admission to completion.
//
// Window fetch follows FMIndex.getReference (Sources/FMIndex/FMIndex.swift:25-65):
// positions in [0, gl) read the forward strand, positions in [gl, 2gl) the
// reverse complement of [2gl-pos-len, 2gl-pos), both clipped at the strand
// boundary; the .pac is 2 bits per base, high bits first
// (Sources/FMIndex/PackedReference.swift:4-5, :34-38).  Left-extension windows
// are then reversed, as ExtensionAligner does (ExtensionAligner.swift:219-225).
#include "engine_internal.h"
#include "local_internal.h"
#include "nw_internal.h"

#include <atomic>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <thread>

using namespace hipsw;

namespace hipsw {

struct Window {
    int64_t pos;        // BWT-space start
    uint32_t word;      // destination word in the chunk's 2-bit arena
    int32_t len;        // bases actually fetched (host-computed clip)
    int32_t rev;        // reverse after fetching (left extensions)
};

// Sixteen lanes per window, one 16-base output word per lane in turn: fetch,
// strand-complement, optionally reverse, and pack 16 bases per word into the
// chunk arena.
__global__ void fetch_windows_kernel(const uint8_t *pac, int64_t gl, const Window *win, int n,
                                     uint32_t *seq)
{
    const int k = blockIdx.x * (blockDim.x >> 4) + (threadIdx.x >> 4);
    if (k >= n) return;
    const Window w = win[k];
    if (w.len <= 0) return;
    const bool rc = w.pos >= gl;
    // forward-strand start of the fetched bases (FMIndex.swift:27-41)
    const int64_t fstart = rc ? (2 * gl - w.pos) - w.len : w.pos;
    uint32_t *dst = seq + w.word;
    const int words = (w.len + 15) >> 4;
    for (int wd = threadIdx.x & 15; wd < words; wd += 16) {
        uint32_t v = 0;
        const int lo = wd * 16, hi = min(w.len, lo + 16);
        for (int i = lo; i < hi; ++i) {
            // index within the strand-oriented window, then optional reversal
            const int o = w.rev ? w.len - 1 - i : i;
            const int64_t fp = rc ? fstart + (w.len - 1 - o) : fstart + o;
            uint32_t b = (pac[fp >> 2] >> ((3 - (int)(fp & 3)) << 1)) & 3u;
            if (rc) b = 3u - b;
            v |= b << ((i - lo) * 2);
        }
        dst[wd] = v;
    }
}

}  // namespace hipsw

// A per-job host array without value-initialisation: a chunk's arrays are
// filled in parallel, so zeroing them first on one thread would cost more
// than the fill (tens of MB per large chunk).
template <typename T> struct RawVec {
    std::unique_ptr<T[]> p;
    RawVec() = default;
    explicit RawVec(size_t n) : p(new T[n > 0 ? n : 1]) {}
    void reset(size_t n) { p.reset(new T[n > 0 ? n : 1]); }
    T *data() { return p.get(); }
    const T *data() const { return p.get(); }
    T &operator[](size_t i) { return p[i]; }
    const T &operator[](size_t i) const { return p[i]; }
};

struct ChunkReq {
    int64_t ticket = 0;
    double t_submit = 0, t_start = 0, t_done = 0;
    size_t est_bytes = 0;
    int status = HIPSW_OK;
    bool done = false;
    // inputs (owned copies), per job in the order extensions, local SW, NW.
    // The queries are packed at submit, in the device arena's layout: job k's
    // 2-bit words follow the words of the jobs before it (ceil(len / 16)
    // each), and its N mask, if it holds an N, follows the masks of the
    // earlier jobs holding one (ceil(len / 32) words each).
    int n_ext = 0, n_loc = 0, n_nw = 0;
    RawVec<uint16_t> q_len;
    RawVec<uint8_t> qhas_n;
    RawVec<uint32_t> qwords, nmask;
    uint64_t n_qwords = 0, n_nmask = 0;
    RawVec<int64_t> t_pos;
    RawVec<uint16_t> t_len_req;
    RawVec<uint8_t> t_rev;
    // sums the admission estimate needs, taken at submit
    uint64_t est_words = 0;                      // query + requested target words
    int est_loc_q = 1, est_loc_t = 1, est_nw_q = 1, est_nw_t = 1;
    std::vector<int32_t> h0, w, eb, nw_w, nw_ts;
    int32_t loc_min_sub = 0;
    std::vector<int64_t> loc_win;                // clamped window start per local job (-1: none)
    // outputs
    std::vector<int32_t> ext_out, loc_out, loc_ok, nw_score, nw_ncig, nw_flags;
    std::vector<uint32_t> nw_cig;                // the used ops of every CIGAR, back to back
    std::vector<uint32_t> nw_cig_off;            // n_nw + 1 offsets into nw_cig
    int64_t cells = 0;
    double host_ms = 0;                          // host packing and planning of the chunk
    int64_t dev_peak = 0;                        // largest device footprint of the run
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    // small NW bins run beside the large one on `side` (fork / join events)
    hipStream_t side = nullptr;
    hipEvent_t evf = nullptr, evj = nullptr;
    uint32_t *counters = nullptr;
    // pinned host staging of the slot's chunk (query words, extension
    // descriptors, result rows), grown on demand and reused: copies from it
    // run at link speed instead of through the driver's pageable staging.
    // A chunk's copies finish before run_chunk returns, so the next chunk of
    // the slot may overwrite it.
    char *pin = nullptr;
    size_t pin_cap = 0;
    std::thread th;
};


// A view of an array inside a slot's pinned staging.
template <typename T> struct PinArr {
    T *p;
    T *data() const { return p; }
    T &operator[](size_t i) const { return p[i]; }
};

// Grows a slot's pinned staging to at least `bytes` (contents not kept).
static bool pin_reserve(Slot &sl, size_t bytes)
{
    if (bytes <= sl.pin_cap) return true;
    if (sl.pin) (void)hipHostFree(sl.pin);
    sl.pin = nullptr;
    sl.pin_cap = 0;
    const size_t cap = std::max(bytes + bytes / 4, (size_t)1 << 20);
    if (hipHostMalloc((void **)&sl.pin, cap, hipHostMallocDefault) != hipSuccess) {
        sl.pin = nullptr;
        return false;
    }
    sl.pin_cap = cap;
    return true;
}

struct hipsw_server {
    hipsw_engine *eng = nullptr;
    int64_t budget = 0;
    int nw_cap = 512;
    hipsw_scoring sc;
    int32_t match_score = 1, mismatch = 4;
    uint8_t *d_pac = nullptr;
    int64_t gl = 0;                              // genome length (one strand)
    // reference sequences (ReferenceMetadata.annotations): offsets into the
    // forward strand and lengths, tiling [0, gl)
    std::vector<int64_t> ctg_off, ctg_len;
    std::vector<std::unique_ptr<Slot>> slots;
    std::mutex mu;
    std::condition_variable cv, done_cv;
    std::deque<std::shared_ptr<ChunkReq>> queue;
    std::map<int64_t, std::shared_ptr<ChunkReq>> live;
    int64_t next_ticket = 1;
    int64_t in_use = 0;
    int running = 0;
    int64_t peak_in_use = 0;
    bool stop = false;
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

// Bases getReference(pos, len) returns (FMIndex.swift:27-50).
int fetched_len(int64_t pos, int len, int64_t gl)
{
    if (pos >= gl) {
        const int64_t fend = 2 * gl - pos, fstart = fend - len;
        const int64_t s = std::max<int64_t>(0, fstart), e = std::min<int64_t>(fend, gl);
        return e > s ? (int)(e - s) : 0;
    }
    if (pos < 0) return 0;
    return (int)std::max<int64_t>(0, std::min<int64_t>(len, gl - pos));
}

// PackedReference.fetchSequence (Sources/FMIndex/PackedReference.swift:72-121):
// the BWT-space window [rb, re) clamped to the reference sequence that holds
// its start, on its strand.  Returns false for an empty clamp; otherwise the
// clamped BWT start and length and the sequence id.
bool fetch_sequence(const hipsw_server *srv, int64_t rb, int64_t re, int64_t *begin, int32_t *len,
                    int32_t *rid)
{
    if (!(rb < re)) return false;
    const int64_t gl = srv->gl;
    const bool fwd = rb < gl;
    const int64_t fb = fwd ? rb : 2 * gl - re, fe = fwd ? re : 2 * gl - rb;
    // ReferenceMetadata.sequenceID (ReferenceMetadata.swift:77-89)
    int lo = 0, hi = (int)srv->ctg_off.size() - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (srv->ctg_off[mid] <= fb) lo = mid;
        else hi = mid - 1;
    }
    const int64_t cb = std::max(fb, srv->ctg_off[lo]);
    const int64_t ce = std::min(fe, srv->ctg_off[lo] + srv->ctg_len[lo]);
    if (!(cb < ce)) return false;
    *begin = fwd ? cb : 2 * gl - ce;
    *len = (int32_t)(ce - cb);
    *rid = lo;
    return true;
}

double now_ms_s() { return hipsw::now_ms(); }

// Copies the used ops of each job's CIGAR row (cap ops per row) back to back:
// job k's n[k] ops go to dst[off[k] ..).  Only these bytes cross PCIe.
__global__ void compact_cigars_kernel(const uint32_t *src, int cap, const uint32_t *off, int n,
                                      uint32_t *dst)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t o = off[k], used = off[k + 1] - o;
        const uint32_t *row = src + (size_t)k * cap;
        for (uint32_t j = 0; j < used; ++j) dst[o + j] = row[j];
    }
}

// Runs one chunk on one slot.  Returns a status.
int run_chunk(hipsw_server *srv, Slot &sl, ChunkReq &r)
{
    hipsw_engine *eng = srv->eng;
    const int n = r.n_ext + r.n_loc + r.n_nw;
    if (n == 0) return HIPSW_OK;
    hipStream_t st = sl.stream;
    const double t_host0 = now_ms_s();
    // --- host packing: queries from the request, targets as fetch windows ---
    // Sequence arena: every query's words, then every target's words (written
    // by the window fetch on the device), then 64 zero words of padding, so
    // the host packs and uploads the queries only.
    const uint8_t *qhas_n = r.qhas_n.data();
    // word offsets: per-block sums in parallel, a scan over the blocks, then
    // each block's offsets in parallel.  The blocks are parallel_for's own
    // chunks of the jobs (one thread each; one call covers them all when the
    // chunk is too small to split).
    const int nth = host_threads();
    const int64_t bs = std::max<int64_t>(1, (n + nth - 1) / nth), nb = (n + bs - 1) / bs;
    auto blocks = [&](const std::function<void(int64_t, int64_t, int64_t)> &body) {
        parallel_for(n, nth, [&](int64_t lo, int64_t hi) {
            for (int64_t b = lo / bs; b * bs < hi; ++b)
                body(b, std::max(lo, b * bs), std::min(hi, (b + 1) * bs));
        });
    };
    std::vector<uint64_t> bq(nb + 1, 0), bt(nb + 1, 0), bn(nb + 1, 0);
    blocks([&](int64_t b, int64_t k0, int64_t k1) {
        uint64_t sq = 0, stt = 0, sn = 0;
        for (int64_t k = k0; k < k1; ++k) {
            const int q = r.q_len[k], t = fetched_len(r.t_pos[k], r.t_len_req[k], srv->gl);
            sq += (q + 15) / 16;
            stt += (t + 15) / 16;
            if (qhas_n[k]) sn += (q + 31) / 32;
        }
        bq[b + 1] = sq; bt[b + 1] = stt; bn[b + 1] = sn;
    });
    for (int64_t b = 0; b < nb; ++b) { bq[b + 1] += bq[b]; bt[b + 1] += bt[b]; bn[b + 1] += bn[b]; }
    const uint64_t qwords = bq[nb], twords = bt[nb];
    uint64_t nwords = bn[nb];
    const uint64_t words = qwords + twords + 64;
    if (words > 0xFFFFFFF0ull) return HIPSW_E_INVALID;
    nwords += 64;
    // Pinned staging of the chunk: the per-job arrays in the device arena's
    // layout (one copy moves them all) | extension descriptors | result rows |
    // query words.
    const size_t meta_b = 4 * span<uint32_t>(n) + 5 * span<int32_t>(n) + span<Window>(n);
    const size_t pin_pk = span<PkMeta>((size_t)r.n_ext);
    const size_t pin_out = span<int32_t>((size_t)n * 6 + 4 * (size_t)n);
    const size_t pin_q = span<uint32_t>(qwords);
    if (!pin_reserve(sl, meta_b + pin_pk + pin_out + pin_q)) return HIPSW_E_DEVICE;
    char *pc = sl.pin;
    const PinArr<uint32_t> qw{carve<uint32_t>(pc, n)}, tw{carve<uint32_t>(pc, n)};
    const PinArr<uint32_t> qnw{carve<uint32_t>(pc, n)}, tnw{carve<uint32_t>(pc, n)};
    const PinArr<int32_t> ql{carve<int32_t>(pc, n)}, tl{carve<int32_t>(pc, n)};
    const PinArr<int32_t> h0v{carve<int32_t>(pc, n)}, wv{carve<int32_t>(pc, n)}, ebv{carve<int32_t>(pc, n)};
    const PinArr<Window> win{carve<Window>(pc, n)};
    PkMeta *hpk = carve<PkMeta>(pc, (size_t)r.n_ext);
    int32_t *hout6 = carve<int32_t>(pc, (size_t)n * 6 + 4 * (size_t)n);
    uint32_t *hq = carve<uint32_t>(pc, qwords);
    int32_t *hstv = hout6 + (size_t)n * 6, *hsc = hstv + n, *hnc = hsc + n, *hfl = hnc + n;
    blocks([&](int64_t b, int64_t k0, int64_t k1) {
        uint64_t q = bq[b], t = qwords + bt[b], m = bn[b];
        for (int64_t k = k0; k < k1; ++k) {
            ql[k] = r.q_len[k];
            tl[k] = fetched_len(r.t_pos[k], r.t_len_req[k], srv->gl);
            qw[k] = (uint32_t)q;
            q += (ql[k] + 15) / 16;
            tw[k] = (uint32_t)t;
            t += (tl[k] + 15) / 16;
            qnw[k] = qhas_n[k] ? (uint32_t)m : kNoN;
            if (qhas_n[k]) m += (ql[k] + 31) / 32;
            tnw[k] = kNoN;                                // windows come from the .pac
            win[k] = Window{r.t_pos[k], tw[k], tl[k], (int32_t)r.t_rev[k]};
        }
    });
    // the query words, packed at submit in this layout, into the pinned staging
    if (qwords != r.n_qwords || nwords - 64 != r.n_nmask) return HIPSW_E_INVALID;
    {
        const int64_t cw = 1 << 18;                            // 1 MiB pieces
        parallel_for((int64_t)((qwords + cw - 1) / cw), host_threads(), [&](int64_t lo, int64_t hi) {
            for (int64_t p = lo; p < hi; ++p) {
                const uint64_t a = (uint64_t)p * cw, e = std::min<uint64_t>(qwords, a + cw);
                memcpy(hq + a, r.qwords.data() + a, (e - a) * 4);
            }
        });
    }
    // extension plan (jobs [0, n_ext))
    const ExtParams ep = make_params(&srv->sc);
    std::vector<uint32_t> ext_order;
    std::vector<WorkItem> ext_items;
    std::vector<ExtLaunch> ext_launch;
    size_t ext_scratch = 0;
    if (r.n_ext) {
        // windows come from the .pac (no N); queries may hold N
        std::vector<uint8_t> any_n(r.n_ext);
        for (int k = 0; k < r.n_ext; ++k) any_n[k] = qnw[k] != kNoN || tnw[k] != kNoN;
        ext_scratch = plan_ext_batched(eng, eng->tune_snapshot(), r.n_ext, ql.data(), tl.data(),
                                       r.h0.data(), r.w.data(), any_n.data(), ep.max_mat,
                                       ep.std_mat != 0, std::max(ep.e_del, ep.e_ins), ext_order,
                                       ext_items, ext_launch);
    }
    // local SW split
    std::vector<uint32_t> loc_s, loc_c;
    int loc_tmax = 1;
    for (int k = r.n_ext; k < r.n_ext + r.n_loc; ++k) {
        if (ql[k] <= 0 || tl[k] <= 0) continue;
        loc_tmax = std::max(loc_tmax, tl[k]);
        if ((int64_t)ql[k] * srv->match_score <= 250 && ql[k] <= 256) {
            loc_s.push_back((uint32_t)k);
        } else {
            loc_c.push_back((uint32_t)k);
        }
    }
    // NW first pass
    const int nw0 = r.n_ext + r.n_loc;
    std::vector<int32_t> wcur(n);
    parallel_for(n, host_threads(), [&](int64_t lo, int64_t hi) {
        for (int64_t k = lo; k < hi; ++k) wcur[k] = k >= nw0 ? std::max(1, r.nw_w[k - nw0]) : 1;
    });
    int mat_abs = 0;
    for (int k = 0; k < 25; ++k) mat_abs = std::max(mat_abs, std::abs((int)srv->sc.mat[k]));
    const int max_pen = std::max(mat_abs, std::max(srv->sc.o_del + srv->sc.e_del, srv->sc.o_ins + srv->sc.e_ins));

    // --- device arena ---
    const size_t colmax_b = r.loc_min_sub > 0 ? loc_s.size() * (size_t)loc_tmax : 1;
    const int cap = srv->nw_cap;
    const size_t bytes = span<uint32_t>(words) + span<uint32_t>(nwords) + 4 * span<uint32_t>(n) +
                         5 * span<int32_t>(n) + span<Window>(n) + span<int32_t>((size_t)n * 6) +
                         span<int32_t>(n) + span<uint32_t>((size_t)r.n_nw * cap + 1) +
                         3 * span<int32_t>(n) + 2 * span<uint32_t>(n) +
                         span<WorkItem>(n / 64 + 64) + span<uint8_t>(colmax_b) +
                         span<PkMeta>(r.n_ext) +
                         span<unsigned long long>(2);
    r.host_ms = now_ms_s() - t_host0;
    void *dmem = nullptr;
    if (eng->inject(HIPSW_FAULT_ALLOC) || hipMallocAsync(&dmem, bytes, st) != hipSuccess)
        return HIPSW_E_DEVICE;
    // the phases' scratch buffers are freed in stream order before the next
    // phase allocates, so the footprint is the arena plus the largest of them
    auto note_phase = [&](size_t b) { r.dev_peak = std::max<int64_t>(r.dev_peak, (int64_t)(bytes + b)); };
    note_phase(0);
    char *cur = reinterpret_cast<char *>(dmem);
    uint32_t *d_seq = carve<uint32_t>(cur, words), *d_nmask = carve<uint32_t>(cur, nwords);
    uint32_t *d_qw = carve<uint32_t>(cur, n), *d_tw = carve<uint32_t>(cur, n);
    uint32_t *d_qn = carve<uint32_t>(cur, n), *d_tn = carve<uint32_t>(cur, n);
    int32_t *d_ql = carve<int32_t>(cur, n), *d_tl = carve<int32_t>(cur, n);
    int32_t *d_h0 = carve<int32_t>(cur, n), *d_w = carve<int32_t>(cur, n), *d_eb = carve<int32_t>(cur, n);
    Window *d_win = carve<Window>(cur, n);
    int32_t *d_out6 = carve<int32_t>(cur, (size_t)n * 6);
    int32_t *d_status = carve<int32_t>(cur, n);
    uint32_t *d_cig = carve<uint32_t>(cur, (size_t)r.n_nw * cap + 1);
    int32_t *d_score = carve<int32_t>(cur, n), *d_ncig = carve<int32_t>(cur, n), *d_flags = carve<int32_t>(cur, n);
    uint32_t *d_order = carve<uint32_t>(cur, n), *d_order2 = carve<uint32_t>(cur, n);
    WorkItem *d_items = carve<WorkItem>(cur, n / 64 + 64);
    uint8_t *d_colmax = carve<uint8_t>(cur, colmax_b);
    unsigned long long *d_cells = carve<unsigned long long>(cur, 2);
    PkMeta *d_pk = carve<PkMeta>(cur, r.n_ext);
    parallel_for(n, host_threads(), [&](int64_t lo, int64_t hi) {
        for (int64_t k = lo; k < hi; ++k) {
            const bool e = k < r.n_ext;
            h0v[k] = e ? r.h0[k] : 0; wv[k] = e ? r.w[k] : 0; ebv[k] = e ? r.eb[k] : 0;
        }
    });
    bool ok = true;
    auto up = [&](void *dst, const void *src, size_t b) {
        if (ok && b) ok = hipMemcpyAsync(dst, src, b, hipMemcpyHostToDevice, st) == hipSuccess;
    };
    up(d_seq, hq, (size_t)qwords * 4);
    if (ok) ok = launch_zero_u32(d_seq + qwords + twords, 64, st) == hipSuccess;
    up(d_nmask, r.nmask.data(), (nwords - 64) * 4);
    if (ok) ok = launch_zero_u32(d_nmask + (nwords - 64), 64, st) == hipSuccess;
    // qw .. win: one copy of the pinned block (same carve order and spans as
    // the arena, d_qw .. d_win)
    up(d_qw, qw.data(), meta_b);
    if (ok) ok = launch_zero_u32(reinterpret_cast<uint32_t *>(d_cells), 4, st) == hipSuccess;
    if (ok) {
        hipLaunchKernelGGL(fetch_windows_kernel, dim3((n + 7) / 8), dim3(128), 0, st,
                           srv->d_pac, srv->gl, d_win, n, d_seq);   // 8 windows per block
        ok = hipGetLastError() == hipSuccess;
    }
    if (ok && eng->inject(HIPSW_FAULT_LAUNCH)) ok = false;
    // --- extension ---
    if (ok && r.n_ext) {
        up(d_order, ext_order.data(), ext_order.size() * 4);
        up(d_items, ext_items.data(), ext_items.size() * sizeof(WorkItem));
        parallel_for(r.n_ext, host_threads(), [&](int64_t lo, int64_t hi) {
            build_pk_meta((int)(hi - lo), ext_order.data() + lo, qw.data(), tw.data(), qnw.data(),
                          tnw.data(), ql.data(), tl.data(), h0v.data(), wv.data(), ebv.data(),
                          hpk + lo);
        });
        up(d_pk, hpk, (size_t)r.n_ext * sizeof(PkMeta));
        void *gs = nullptr;
        if (ok && ext_scratch) ok = hipMallocAsync(&gs, ext_scratch, st) == hipSuccess;
        note_phase(ext_scratch);
        if (ok) ok = launch_zero_u32(sl.counters, 64, st) == hipSuccess;
        const ExtDev ed{d_seq, d_nmask, d_qw, d_tw, d_qn, d_tn, d_ql, d_tl, d_h0, d_w, d_eb, d_out6,
                        d_cells, d_pk};
        int k = 0;
        for (const ExtLaunch &L : ext_launch) {
            if (!ok || k >= 64) { ok = false; break; }
            ok = launch_ext_bin(L, ed, ep, d_order, d_items, sl.counters + k, gs, st) == hipSuccess;
            ++k;
        }
        if (gs) hipFreeAsync(gs, st);
    }
    // --- local SW ---
    if (ok && r.n_loc) {
        LocalParams lp;
        memset(&lp, 0, sizeof lp);
        int max_u8 = 0, max_i = 0;
        for (int k = 0; k < 25; ++k) {
            lp.mat[k] = srv->sc.mat[k];
            max_u8 = std::max(max_u8, std::min(std::max((int)srv->sc.mat[k] + srv->mismatch, 0), 255));
            max_i = std::max(max_i, (int)srv->sc.mat[k]);
        }
        lp.bias = srv->mismatch;
        lp.gap_oe = srv->sc.o_ins + srv->sc.e_ins;
        lp.gap_e = srv->sc.e_ins;
        lp.min_sub = std::min(std::max(r.loc_min_sub, 0), 255);
        lp.max_base_u8 = max_u8 > srv->mismatch ? max_u8 - srv->mismatch : 0;
        lp.max_base_i = max_i;
        const LocalDev ld{d_seq, d_nmask, d_qw, d_tw, d_qn, d_tn, d_ql, d_tl, d_out6, d_status};
        if (ok) ok = hipMemsetAsync(d_status, 0, (size_t)n * 4, st) == hipSuccess;
        if (ok && !loc_s.empty()) {
            std::vector<uint32_t> grouped(loc_s.size());
            int32_t grp_off[17];
            group_by_stripes(loc_s.data(), loc_s.size(), ql.data(), grouped.data(), grp_off);
            up(d_order2, grouped.data(), grouped.size() * 4);
            if (ok) ok = launch_local_striped(ld, lp, d_order2, grp_off, d_colmax,
                                              lp.min_sub > 0 ? loc_tmax : 0, st) == hipSuccess;
            if (ok) ok = hipMemcpyAsync(hstv + r.n_ext, d_status + r.n_ext, (size_t)r.n_loc * 4,
                                        hipMemcpyDeviceToHost, st) == hipSuccess;   // local rows
            if (ok) ok = hipStreamSynchronize(st) == hipSuccess;
            if (ok)
                for (uint32_t k : loc_s)
                    if (hstv[k] == LOCAL_NEEDS_SCALAR) loc_c.push_back(k);
        }
        if (ok && !loc_c.empty()) {
            int qc = 1, tc = 1;
            for (uint32_t k : loc_c) { qc = std::max(qc, ql[k]); tc = std::max(tc, tl[k]); }
            const int blocks = (int)((loc_c.size() + 63) / 64);
            void *scr = nullptr;
            up(d_order2, loc_c.data(), loc_c.size() * 4);
            const size_t scr_b = (size_t)blocks * 64 * (2 * (tc + 1) + qc) * 4;
            if (ok) ok = hipMallocAsync(&scr, scr_b, st) == hipSuccess;
            note_phase(scr_b);
            if (ok) ok = launch_local_scalar(ld, lp, d_order2, (int)loc_c.size(), (int *)scr, tc, qc, st) == hipSuccess;
            if (scr) hipFreeAsync(scr, st);
        }
    }
    // --- NW with the CPU retry loop ---
    if (ok && r.n_nw) {
        NwParams np;
        memset(&np, 0, sizeof np);
        for (int k = 0; k < 25; ++k) np.mat[k] = srv->sc.mat[k];
        np.o_del = srv->sc.o_del; np.e_del = srv->sc.e_del;
        np.o_ins = srv->sc.o_ins; np.e_ins = srv->sc.e_ins;
        np.oe_del = np.o_del + np.e_del; np.oe_ins = np.o_ins + np.e_ins;
        // NW outputs are indexed by job id; cigar rows by (job - nw0).
        const NwDev nd{d_seq, d_nmask, d_qw, d_tw, d_qn, d_tn, d_ql, d_tl, d_w, d_score,
                       d_cig - (size_t)nw0 * cap, d_ncig, d_flags, d_cells + 1, cap};
        std::vector<uint32_t> active;
        for (int k = 0; k < r.n_nw; ++k) active.push_back((uint32_t)(nw0 + k));
        std::vector<int> tries(r.n_nw, 0);                     // indexed by job - nw0
        std::vector<uint32_t> order;
        std::vector<WorkItem> items;
        std::vector<NwLaunch> launches;
        for (int pass = 0; ok && !active.empty(); ++pass) {
            // small bins and the retry passes one wave per job, as the batched
            // form of hipsw_nw_run does
            nw_plan_pass(eng, ql.data(), tl.data(), max_pen, active, wcur,
                         pass == 0 ? HIPSW_MODE_BATCHED : HIPSW_MODE_PLAIN, order, items, launches,
                         np.o_ins >= 0 ? (pass == 0 ? 1 : 2) : 0);
            size_t zb = 0, zs = 0, gb = 0;
            for (const NwLaunch &L : launches) {
                if (L.wide) {
                    if (L.zmax > 49152) zb = std::max(zb, (size_t)L.grid * L.zmax);
                    continue;
                }
                if (nw_side_launch(L, launches)) zs = std::max(zs, nw_launch_zbytes(L));
                else zb = std::max(zb, nw_launch_zbytes(L));
                if (L.global_state) gb = std::max(gb, (size_t)L.grid * L.S * 64 * (L.tier == NW_I16 ? 4 : 8));
            }
            zb = (zb + 255) & ~(size_t)255;
            zs = (zs + 255) & ~(size_t)255;
            void *scr = nullptr;
            up(d_w + nw0, wcur.data() + nw0, (size_t)r.n_nw * 4);   // NW rows only
            up(d_order, order.data(), order.size() * 4);
            up(d_items, items.data(), items.size() * sizeof(WorkItem));
            if (ok) ok = hipMallocAsync(&scr, zb + zs + gb + 512, st) == hipSuccess;
            note_phase(zb + zs + gb + 512);
            if (ok) ok = launch_zero_u32(sl.counters, 64, st) == hipSuccess;
            if (launches.size() > 64) ok = false;
            // as hipsw_nw_run: small bins in the wave-per-job form with their
            // flags in LDS run on the side stream, queued first
            auto side_ok = [&](const NwLaunch &L) { return nw_side_launch(L, launches); };
            auto run_wide = [&](const NwLaunch &L, int k, hipStream_t ls) {
                const WorkItem &first = items[L.item_begin];
                const WorkItem &last = items[L.item_begin + L.item_count - 1];
                return launch_nw_wide(nd, np, d_order + first.start,
                                      (int)(last.start + last.count - first.start), sl.counters + k,
                                      L.qmax, (uint8_t *)scr, L.zmax, L.zmax <= 49152, L.grid,
                                      ls) == hipSuccess;
            };
            bool forked = false;
            for (size_t li = 0; ok && li < launches.size(); ++li) {
                if (!side_ok(launches[li])) continue;
                if (!forked) {
                    ok = hipEventRecord(sl.evf, st) == hipSuccess &&
                         hipStreamWaitEvent(sl.side, sl.evf, 0) == hipSuccess;
                    forked = true;
                }
                const NwLaunch &L = launches[li];
                if (ok && L.wide) ok = run_wide(L, (int)li, sl.side);
                else if (ok)
                    ok = launch_nw(L.tier, false, L.reg, nd, np, d_order, d_items + L.item_begin,
                                   L.item_count, sl.counters + li, 0, (uint32_t *)((char *)scr + zb),
                                   L.zrows, (L.ncol_max + 7) / 8, L.grid, pass == 0, nullptr, L.qw8,
                                   sl.side) == hipSuccess;
            }
            int k = -1;
            for (const NwLaunch &L : launches) {
                ++k;
                if (!ok) break;
                if (side_ok(L)) continue;
                if (L.wide) {
                    ok = run_wide(L, k, st);
                    continue;
                }
                ok = launch_nw(L.tier, L.global_state, L.reg, nd, np, d_order, d_items + L.item_begin,
                               L.item_count, sl.counters + k, L.S, (uint32_t *)scr, L.zrows,
                               (L.ncol_max + 7) / 8, L.grid, pass == 0, (char *)scr + zb + zs,
                               L.qw8, st) == hipSuccess;
            }
            if (forked) {
                // joined even after a failed launch: scr is freed in stream order
                const bool j = hipEventRecord(sl.evj, sl.side) == hipSuccess &&
                               hipStreamWaitEvent(st, sl.evj, 0) == hipSuccess;
                ok = ok && j;
            }
            if (scr) hipFreeAsync(scr, st);
            // the NW scores only, into the slot's pinned score row
            if (ok) ok = hipMemcpyAsync(hsc + nw0, d_score + nw0, (size_t)r.n_nw * 4,
                                        hipMemcpyDeviceToHost, st) == hipSuccess;
            if (ok) ok = hipStreamSynchronize(st) == hipSuccess;
            std::vector<uint32_t> next;
            for (uint32_t j : active) {
                if (hsc[j] < r.nw_ts[j - nw0] && tries[j - nw0] < 3) {
                    wcur[j] = wcur[j] < (1 << 20) ? wcur[j] * 2 : wcur[j];
                    tries[j - nw0]++;
                    next.push_back(j);
                }
            }
            active.swap(next);
        }
    }
    // --- results ---
    int32_t *out6 = hout6, *stv = hstv, *sc = hsc, *nc = hnc, *fl = hfl;
    unsigned long long cells[2] = {0, 0};
    auto down = [&](void *dst, const void *src, size_t b) {
        if (ok && b) ok = hipMemcpyAsync(dst, src, b, hipMemcpyDeviceToHost, st) == hipSuccess;
    };
    if (ok && eng->inject(HIPSW_FAULT_COPY)) ok = false;
    // only the rows each kind reads: extension and local results, local
    // status, NW score / op count / flags
    down(out6, d_out6, (size_t)nw0 * 6 * 4);
    down(stv + r.n_ext, d_status + r.n_ext, (size_t)r.n_loc * 4);
    down(sc + nw0, d_score + nw0, (size_t)r.n_nw * 4);
    down(nc + nw0, d_ncig + nw0, (size_t)r.n_nw * 4);
    down(fl + nw0, d_flags + nw0, (size_t)r.n_nw * 4);
    down(cells, d_cells, sizeof cells);
    if (ok) ok = hipStreamSynchronize(st) == hipSuccess;
    // CIGARs: only the used ops of each row come back (rows hold cap ops)
    r.nw_cig_off.assign((size_t)r.n_nw + 1, 0);
    for (int k = 0; k < r.n_nw; ++k)
        r.nw_cig_off[k + 1] = r.nw_cig_off[k] + (uint32_t)std::min(std::max(nc[nw0 + k], 0), cap);
    r.nw_cig.assign(r.n_nw ? (size_t)r.nw_cig_off[r.n_nw] : 0, 0u);
    if (ok && r.n_nw && !r.nw_cig.empty()) {
        void *cm = nullptr;
        const size_t off_b = ((size_t)(r.n_nw + 1) * 4 + 255) & ~(size_t)255;
        ok = hipMallocAsync(&cm, off_b + r.nw_cig.size() * 4, st) == hipSuccess;
        note_phase(off_b + r.nw_cig.size() * 4);
        uint32_t *d_off = static_cast<uint32_t *>(cm);
        uint32_t *d_dense = reinterpret_cast<uint32_t *>(static_cast<char *>(cm) + off_b);
        if (ok) ok = hipMemcpyAsync(d_off, r.nw_cig_off.data(), (size_t)(r.n_nw + 1) * 4,
                                    hipMemcpyHostToDevice, st) == hipSuccess;
        if (ok) {
            hipLaunchKernelGGL(compact_cigars_kernel, dim3(std::min((r.n_nw + 255) / 256, 4096)),
                               dim3(256), 0, st, d_cig, cap, d_off, r.n_nw, d_dense);
            ok = hipGetLastError() == hipSuccess;
        }
        down(r.nw_cig.data(), d_dense, r.nw_cig.size() * 4);
        if (cm) hipFreeAsync(cm, st);
    }
    hipFreeAsync(dmem, st);
    // drained on every path: the next chunk of this slot reuses the pinned
    // staging that queued copies may still read or write
    const bool drained = hipStreamSynchronize(st) == hipSuccess;
    if (!ok || !drained) return HIPSW_E_DEVICE;
    // result rows out of the pinned staging, in parallel
    r.ext_out.resize((size_t)r.n_ext * 6);
    parallel_for(r.n_ext, host_threads(), [&](int64_t lo, int64_t hi) {
        memcpy(r.ext_out.data() + lo * 6, out6 + lo * 6, (size_t)(hi - lo) * 6 * 4);
    });
    r.loc_out.assign((size_t)r.n_loc * 6, 0);
    r.loc_ok.assign(r.n_loc, 0);
    parallel_for(r.n_loc, host_threads(), [&](int64_t lo, int64_t hi) {
        for (int64_t k = lo; k < hi; ++k) {
            const int64_t j = r.n_ext + k;
            r.loc_ok[k] = stv[j] == LOCAL_OK ? 1 : 0;
            if (r.loc_ok[k]) memcpy(&r.loc_out[(size_t)k * 6], &out6[(size_t)j * 6], 6 * 4);
        }
    });
    r.nw_score.assign(sc + nw0, sc + nw0 + r.n_nw);
    r.nw_ncig.assign(nc + nw0, nc + nw0 + r.n_nw);
    r.nw_flags.assign(fl + nw0, fl + nw0 + r.n_nw);
    r.cells = (int64_t)(cells[0] + cells[1]);
    return HIPSW_OK;
}

// Device bytes a chunk can hold at once, as run_chunk allocates them: its
// arena (every term of run_chunk's arena size, with the requested window
// lengths as the fetched ones' upper bound) plus the largest phase scratch --
// the phases free theirs in stream order before the next one allocates.  The
// phase bounds hold for any plan: the extension bound considers every built
// LDS/HBM split (ext_scratch_bound), local SW assumes every task may fall
// back to the scalar kernel, and each NW pass is bounded with the longest
// query's column count and the longest target's rows for every wave.
size_t estimate_bytes(const hipsw_server *srv, const ChunkReq &r)
{
    const size_t n = (size_t)(r.n_ext + r.n_loc + r.n_nw);
    const int cap = srv->nw_cap;
    const uint64_t words = 64 + r.est_words, nwords = 64 + r.n_nmask;
    const int loc_q = r.est_loc_q, loc_t = r.est_loc_t, nw_q = r.est_nw_q, nw_t = r.est_nw_t;
    const size_t colmax_b = r.loc_min_sub > 0 ? (size_t)r.n_loc * loc_t : 1;
    const size_t arena = span<uint32_t>(words) + span<uint32_t>(nwords) + 4 * span<uint32_t>(n) +
                         5 * span<int32_t>(n) + span<Window>(n) + span<int32_t>(n * 6) +
                         span<int32_t>(n) + span<uint32_t>((size_t)r.n_nw * cap + 1) +
                         3 * span<int32_t>(n) + 2 * span<uint32_t>(n) +
                         span<WorkItem>(n / 64 + 64) + span<uint8_t>(colmax_b) +
                         span<PkMeta>(r.n_ext) + span<unsigned long long>(2);
    size_t phase = 0;
    if (r.n_ext) {
        const ExtParams ep = make_params(&srv->sc);
        RawVec<int32_t> ql(r.n_ext);
        parallel_for(r.n_ext, host_threads(), [&](int64_t lo, int64_t hi) {
            for (int64_t k = lo; k < hi; ++k) ql[k] = r.q_len[k];
        });
        phase = std::max(phase, ext_scratch_bound(srv->eng, r.n_ext, ql.data(), r.h0.data(),
                                                  r.qhas_n.data(), ep.max_mat, ep.std_mat != 0));
    }
    if (r.n_loc) {
        const size_t blocks = ((size_t)r.n_loc + 63) / 64;
        phase = std::max(phase, blocks * 64 * (2 * ((size_t)loc_t + 1) + loc_q) * 4);
    }
    if (r.n_nw) {
        // lane-per-job launches: per wave, tlen rows of ceil(ncol/8) dwords
        // per lane (ncol <= qlen), plus HBM row state of S <= qlen + 10 slots;
        // wave-per-job launches: a tlen x ncol byte matrix per wave
        const size_t waves = ((size_t)r.n_nw + 63) / 64 + 16;
        const size_t lane = waves * (size_t)nw_t * ((nw_q + 7) / 8) * 64 * 4 +
                            waves * ((size_t)nw_q + 10) * 64 * 8 + 1024;
        const size_t wide = std::min<size_t>((size_t)r.n_nw, (size_t)srv->eng->n_cu * 8) *
                            (size_t)nw_t * nw_q + 1024;
        phase = std::max(phase, std::max(lane, wide));
        phase = std::max(phase, ((size_t)(r.n_nw + 1) * 4 + 255) + (size_t)r.n_nw * cap * 4);
    }
    return arena + phase;
}

void slot_loop(hipsw_server *srv, Slot *sl)
{
    hipSetDevice(srv->eng->device);
    for (;;) {
        std::shared_ptr<ChunkReq> req;
        {
            std::unique_lock<std::mutex> lk(srv->mu);
            srv->cv.wait(lk, [&] {
                if (srv->stop) return true;
                if (srv->queue.empty()) return false;
                const auto &f = srv->queue.front();
                return srv->running == 0 || srv->in_use + (int64_t)f->est_bytes <= srv->budget;
            });
            if (srv->stop) return;
            req = srv->queue.front();
            srv->queue.pop_front();
            srv->in_use += (int64_t)req->est_bytes;
            srv->peak_in_use = std::max(srv->peak_in_use, srv->in_use);
            srv->running++;
            req->t_start = now_ms_s();
        }
        const int rc = run_chunk(srv, *sl, *req);
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            req->status = rc;
            req->t_done = now_ms_s();
            req->done = true;
            srv->in_use -= (int64_t)req->est_bytes;
            srv->running--;
        }
        srv->done_cv.notify_all();
        srv->cv.notify_all();
    }
}

}  // namespace

extern "C" hipsw_server *hipsw_server_create(hipsw_engine *eng, int32_t n_slots, int64_t hbm_budget,
                                            const hipsw_scoring *sc, int32_t match_score,
                                            int32_t mismatch_penalty, int32_t nw_cap)
{
    if (!eng || n_slots < 1 || n_slots > 16 || hbm_budget <= 0 || !scoring_valid(sc) || nw_cap < 1)
        return nullptr;
    hipSetDevice(eng->device);
    auto *srv = new hipsw_server();
    srv->eng = eng;
    srv->budget = hbm_budget;
    srv->sc = *sc;
    srv->match_score = match_score;
    srv->mismatch = mismatch_penalty;
    srv->nw_cap = nw_cap;
    // Chunk arenas come from the device's default stream-ordered pool: keep
    // up to the budget cached there, so that a chunk reuses the memory its
    // predecessors freed instead of mapping it again.
    hipMemPool_t pool = nullptr;
    if (hipDeviceGetDefaultMemPool(&pool, eng->device) == hipSuccess && pool) {
        uint64_t keep = 0;
        if (hipMemPoolGetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep) == hipSuccess &&
            keep < (uint64_t)hbm_budget) {
            keep = (uint64_t)hbm_budget;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        }
    }
    for (int s = 0; s < n_slots; ++s) {
        auto sl = std::make_unique<Slot>();
        if (hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking) != hipSuccess ||
            hipStreamCreateWithFlags(&sl->side, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl->evf, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&sl->evj, hipEventDisableTiming) != hipSuccess ||
            hipMalloc((void **)&sl->counters, 64 * sizeof(uint32_t)) != hipSuccess) {
            delete srv;
            return nullptr;
        }
        srv->slots.push_back(std::move(sl));
    }
    for (auto &sl : srv->slots) sl->th = std::thread(slot_loop, srv, sl.get());
    return srv;
}

extern "C" int hipsw_server_set_reference(hipsw_server *srv, const uint8_t *pac, int64_t n_bases)
{
    if (!srv || !pac || n_bases <= 0) return HIPSW_E_INVALID;
    std::unique_lock<std::mutex> lk(srv->mu);
    if (srv->running || !srv->queue.empty()) return HIPSW_E_PENDING;
    hipSetDevice(srv->eng->device);
    // the new reference goes to a buffer of its own; the server's reference
    // changes only once it is complete, and a failure leaves no reference
    // (submits are then refused) rather than a stale length over new memory
    const size_t bytes = (size_t)(n_bases + 3) / 4;
    uint8_t *fresh = nullptr;
    const bool ok = hipMalloc((void **)&fresh, bytes) == hipSuccess &&
                    hipMemcpy(fresh, pac, bytes, hipMemcpyHostToDevice) == hipSuccess;
    if (srv->d_pac) hipFree(srv->d_pac);
    srv->d_pac = nullptr;
    srv->gl = 0;
    if (!ok) {
        if (fresh) hipFree(fresh);
        return HIPSW_E_DEVICE;
    }
    srv->d_pac = fresh;
    srv->gl = n_bases;
    srv->ctg_off.assign(1, 0);
    srv->ctg_len.assign(1, n_bases);
    return HIPSW_OK;
}

extern "C" int hipsw_server_set_contigs(hipsw_server *srv, int32_t n, const int64_t *off,
                                        const int64_t *len)
{
    if (!srv || n < 1 || !off || !len) return HIPSW_E_INVALID;
    std::unique_lock<std::mutex> lk(srv->mu);
    if (srv->running || !srv->queue.empty()) return HIPSW_E_PENDING;
    if (!srv->d_pac) return HIPSW_E_INVALID;
    int64_t at = 0;
    for (int k = 0; k < n; ++k) {
        if (off[k] != at || len[k] <= 0) return HIPSW_E_INVALID;
        at += len[k];
    }
    if (at != srv->gl) return HIPSW_E_INVALID;
    srv->ctg_off.assign(off, off + n);
    srv->ctg_len.assign(len, len + n);
    return HIPSW_OK;
}

extern "C" int64_t hipsw_server_submit(hipsw_server *srv, const hipsw_chunk *c)
{
    if (!srv || !c) return HIPSW_E_INVALID;
    if (!srv->d_pac) return HIPSW_E_INVALID;
    auto r = std::make_shared<ChunkReq>();
    r->t_submit = now_ms_s();
    r->n_ext = c->n_ext; r->n_loc = c->n_loc; r->n_nw = c->n_nw;
    if (r->n_ext < 0 || r->n_loc < 0 || r->n_nw < 0) return HIPSW_E_INVALID;
    r->loc_min_sub = c->loc_min_sub;
    // Jobs in order: extensions, local SW, NW.  The query offsets come from
    // one scan over the lengths; every job's fields and query bytes are then
    // copied in parallel (the byte arena is not zero-filled first).
    struct Src {
        int cnt;
        const uint8_t *q; const uint32_t *qo; const uint16_t *ql;
        const int64_t *tp; const uint16_t *tl; const uint8_t *tr;
    };
    const Src src[3] = {
        {c->n_ext, c->ext_q, c->ext_q_off, c->ext_q_len, c->ext_t_pos, c->ext_t_len, c->ext_t_rev},
        {c->n_loc, c->loc_q, c->loc_q_off, c->loc_q_len, c->loc_t_pos, c->loc_t_len, nullptr},
        {c->n_nw, c->nw_q, c->nw_q_off, c->nw_q_len, c->nw_t_pos, c->nw_t_len, nullptr}};
    for (const Src &sr : src)
        if (sr.cnt && (!sr.q || !sr.qo || !sr.ql || !sr.tp || !sr.tl)) return HIPSW_E_INVALID;
    const int64_t n = (int64_t)r->n_ext + r->n_loc + r->n_nw;
    const int64_t b1 = r->n_ext, b2 = b1 + r->n_loc;
    r->q_len.reset(n);
    r->qhas_n.reset(n);
    r->t_pos.reset(n);
    r->t_len_req.reset(n);
    r->t_rev.reset(n);
    // Pass 1 over blocks of jobs: the fields, the N test and each block's
    // word counts; then offsets by a scan over the blocks; pass 2 packs each
    // query from the caller's bytes straight into the chunk's words.
    const int nth = host_threads();
    const int64_t bs = std::max<int64_t>(1, (n + nth - 1) / nth), nb = n > 0 ? (n + bs - 1) / bs : 0;
    auto blocks = [&](const std::function<void(int64_t, int64_t, int64_t)> &body) {
        parallel_for(n, nth, [&](int64_t lo, int64_t hi) {
            for (int64_t b = lo / bs; b * bs < hi; ++b)
                body(b, std::max(lo, b * bs), std::min(hi, (b + 1) * bs));
        });
    };
    struct BlockSums {
        uint64_t q = 0, nm = 0, t = 0;
        int loc_q = 1, loc_t = 1, nw_q = 1, nw_t = 1;
    };
    std::vector<BlockSums> bsum(nb + 1);
    auto src_of = [&](int64_t j, int64_t &k) -> const Src & {
        const int kind = j < b1 ? 0 : j < b2 ? 1 : 2;
        k = j - (kind == 0 ? 0 : kind == 1 ? b1 : b2);
        return src[kind];
    };
    blocks([&](int64_t b, int64_t j0, int64_t j1) {
        BlockSums &S = bsum[b + 1];
        for (int64_t j = j0; j < j1; ++j) {
            int64_t k;
            const Src &sr = src_of(j, k);
            const int q = sr.ql[k];
            r->q_len[j] = (uint16_t)q;
            r->t_pos[j] = sr.tp[k];
            r->t_len_req[j] = sr.tl[k];
            r->t_rev[j] = sr.tr ? sr.tr[k] : 0;
            r->qhas_n[j] = has_n(sr.q + sr.qo[k], q);
            S.q += (q + 15) / 16;
            S.t += (sr.tl[k] + 15) / 16;
            if (r->qhas_n[j]) S.nm += (q + 31) / 32;
            if (j >= b1 && j < b2) {
                S.loc_q = std::max(S.loc_q, q);
                S.loc_t = std::max(S.loc_t, (int)sr.tl[k]);
            } else if (j >= b2) {
                S.nw_q = std::max(S.nw_q, q);
                S.nw_t = std::max(S.nw_t, (int)sr.tl[k]);
            }
        }
    });
    for (int64_t b = 0; b < nb; ++b) {
        BlockSums &A = bsum[b + 1];
        const BlockSums &P = bsum[b];
        A.q += P.q; A.nm += P.nm; A.t += P.t;
        A.loc_q = std::max(A.loc_q, P.loc_q); A.loc_t = std::max(A.loc_t, P.loc_t);
        A.nw_q = std::max(A.nw_q, P.nw_q); A.nw_t = std::max(A.nw_t, P.nw_t);
    }
    const BlockSums &T = bsum[nb];
    if (T.q + T.t > 0xFFFFFF00ull) return HIPSW_E_INVALID;
    r->n_qwords = T.q;
    r->n_nmask = T.nm;
    r->est_words = T.q + T.t;
    r->est_loc_q = T.loc_q; r->est_loc_t = T.loc_t; r->est_nw_q = T.nw_q; r->est_nw_t = T.nw_t;
    r->qwords.reset(T.q);
    r->nmask.reset(T.nm);
    std::atomic<bool> packed_ok{true};
    blocks([&](int64_t b, int64_t j0, int64_t j1) {
        uint64_t q = bsum[b].q, m = bsum[b].nm;
        for (int64_t j = j0; j < j1; ++j) {
            int64_t k;
            const Src &sr = src_of(j, k);
            const int len = r->q_len[j];
            if (!pack2(sr.q + sr.qo[k], len, r->qwords.data() + q,
                       r->qhas_n[j] ? r->nmask.data() + m : nullptr))
                packed_ok = false;
            q += (len + 15) / 16;
            if (r->qhas_n[j]) m += (len + 31) / 32;
        }
    });
    if (!packed_ok) return HIPSW_E_INVALID;                 // a query byte above 4
    if (c->loc_rid) {
        // mate-rescue windows: fetchSequence clamping, and the template's
        // sequence must hold the window (MateRescue.swift:105-111, :252-256)
        std::lock_guard<std::mutex> lk(srv->mu);
        r->loc_win.assign(r->n_loc, -1);
        parallel_for(r->n_loc, nth, [&](int64_t lo, int64_t hi) {
            for (int64_t k = lo; k < hi; ++k) {
                const int64_t j = b1 + k;
                int64_t begin;
                int32_t len, rid;
                if (fetch_sequence(srv, r->t_pos[j], r->t_pos[j] + r->t_len_req[j], &begin, &len,
                                   &rid) &&
                    rid == c->loc_rid[k]) {
                    r->loc_win[k] = begin;
                    r->t_pos[j] = begin;
                    r->t_len_req[j] = (uint16_t)len;
                } else {
                    r->t_len_req[j] = 0;                          // skipped: ok = 0
                }
            }
        });
    }
    if (c->n_ext && (!c->ext_h0 || !c->ext_w || !c->ext_eb)) return HIPSW_E_INVALID;
    for (int k = 0; k < c->n_ext; ++k)
        if (c->ext_w[k] < 0) return HIPSW_E_INVALID;   // band width must be >= 0
    if (c->n_nw && (!c->nw_w || !c->nw_true_score)) return HIPSW_E_INVALID;
    r->h0.assign(c->ext_h0, c->ext_h0 + c->n_ext);
    r->w.assign(c->ext_w, c->ext_w + c->n_ext);
    r->eb.assign(c->ext_eb, c->ext_eb + c->n_ext);
    r->nw_w.assign(c->nw_w, c->nw_w + c->n_nw);
    r->nw_ts.assign(c->nw_true_score, c->nw_true_score + c->n_nw);
    r->est_bytes = estimate_bytes(srv, *r);
    std::lock_guard<std::mutex> lk(srv->mu);
    if ((int64_t)r->est_bytes > srv->budget) return HIPSW_E_BUDGET;
    r->ticket = srv->next_ticket++;
    srv->live[r->ticket] = r;
    srv->queue.push_back(r);
    srv->cv.notify_all();
    return r->ticket;
}

extern "C" int hipsw_server_poll(hipsw_server *srv, int64_t ticket)
{
    if (!srv) return HIPSW_E_INVALID;
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->live.find(ticket);
    if (it == srv->live.end()) return HIPSW_E_INVALID;
    return it->second->done ? HIPSW_OK : HIPSW_E_PENDING;
}

extern "C" int hipsw_server_wait(hipsw_server *srv, int64_t ticket, hipsw_chunk_result *res)
{
    if (!srv) return HIPSW_E_INVALID;
    std::shared_ptr<ChunkReq> r;
    {
        std::unique_lock<std::mutex> lk(srv->mu);
        auto it = srv->live.find(ticket);
        if (it == srv->live.end()) return HIPSW_E_INVALID;
        r = it->second;
        srv->done_cv.wait(lk, [&] { return r->done; });
        srv->live.erase(it);
    }
    if (res) {
        res->latency_ms = r->t_done - r->t_submit;
        res->queue_ms = r->t_start - r->t_submit;
        res->run_ms = r->t_done - r->t_start;
        res->host_ms = r->host_ms;
        res->cells = r->cells;
        res->est_bytes = (int64_t)r->est_bytes;
        res->dev_peak_bytes = r->dev_peak;
        if (r->status == HIPSW_OK) {
            const int cap = srv->nw_cap;
            if (res->ext_out && r->n_ext) memcpy(res->ext_out, r->ext_out.data(), (size_t)r->n_ext * 24);
            if (res->loc_out && r->n_loc) memcpy(res->loc_out, r->loc_out.data(), (size_t)r->n_loc * 24);
            if (res->loc_ok && r->n_loc) memcpy(res->loc_ok, r->loc_ok.data(), (size_t)r->n_loc * 4);
            if (res->loc_win && r->n_loc && !r->loc_win.empty())
                memcpy(res->loc_win, r->loc_win.data(), (size_t)r->n_loc * 8);
            if (res->nw_score && r->n_nw) memcpy(res->nw_score, r->nw_score.data(), (size_t)r->n_nw * 4);
            if (res->nw_n_cigar && r->n_nw) memcpy(res->nw_n_cigar, r->nw_ncig.data(), (size_t)r->n_nw * 4);
            if (res->nw_flags && r->n_nw) memcpy(res->nw_flags, r->nw_flags.data(), (size_t)r->n_nw * 4);
            if (res->nw_cigar && r->n_nw && res->nw_cigar_off) {
                // compact: the used ops back to back, with their offsets
                memcpy(res->nw_cigar_off, r->nw_cig_off.data(), ((size_t)r->n_nw + 1) * 4);
                const ChunkReq *rq = r.get();
                parallel_for((int64_t)rq->nw_cig.size(), host_threads(), [&](int64_t lo, int64_t hi) {
                    memcpy(res->nw_cigar + lo, rq->nw_cig.data() + lo, (size_t)(hi - lo) * 4);
                });
            } else if (res->nw_cigar && r->n_nw) {
                // rows of cap ops: the used ones, then zeros
                const ChunkReq *rq = r.get();
                parallel_for(rq->n_nw, host_threads(), [&](int64_t lo, int64_t hi) {
                    for (int64_t k = lo; k < hi; ++k) {
                        const uint32_t o = rq->nw_cig_off[k], used = rq->nw_cig_off[k + 1] - o;
                        uint32_t *row = res->nw_cigar + (size_t)k * cap;
                        if (used) memcpy(row, rq->nw_cig.data() + o, (size_t)used * 4);
                        memset(row + used, 0, (size_t)(cap - used) * 4);
                    }
                });
            }
        }
    }
    return r->status;
}

extern "C" int hipsw_server_stats(hipsw_server *srv, int64_t *in_use, int64_t *peak_in_use,
                                  int32_t *running, int32_t *queued)
{
    if (!srv) return HIPSW_E_INVALID;
    std::lock_guard<std::mutex> lk(srv->mu);
    if (in_use) *in_use = srv->in_use;
    if (peak_in_use) *peak_in_use = srv->peak_in_use;
    if (running) *running = srv->running;
    if (queued) *queued = (int32_t)srv->queue.size();
    return HIPSW_OK;
}

extern "C" void hipsw_server_destroy(hipsw_server *srv)
{
    if (!srv) return;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
    }
    srv->cv.notify_all();
    for (auto &sl : srv->slots)
        if (sl->th.joinable()) sl->th.join();
    hipSetDevice(srv->eng->device);
    for (auto &sl : srv->slots) {
        if (sl->stream) {
            hipStreamSynchronize(sl->stream);
            hipStreamDestroy(sl->stream);
        }
        if (sl->side) {
            hipStreamSynchronize(sl->side);
            hipStreamDestroy(sl->side);
        }
        if (sl->evf) hipEventDestroy(sl->evf);
        if (sl->evj) hipEventDestroy(sl->evj);
        if (sl->counters) hipFree(sl->counters);
        if (sl->pin) (void)hipHostFree(sl->pin);
    }
    if (srv->d_pac) hipFree(srv->d_pac);
    delete srv;
}
