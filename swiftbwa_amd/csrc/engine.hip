// Host side of libhipsw: the C ABI declared in include/hipsw.h.
//
// Replaces MetalSW's engine and dispatchers (Sources/MetalSW/MetalSWEngine.swift,
// BandedSWDispatcher.swift, GlobalSWDispatcher.swift, LocalSWDispatcher.swift)
// with a HIP engine: a pool of stream contexts for concurrent one-shot calls,
// grow-only device arenas, 2-bit sequence packing on the host, and
// length-binned persistent launches.
#include "hipsw.h"
#include "hipsw_internal.h"
#include "engine_internal.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

using namespace hipsw;

namespace {
// Error text per (thread, engine): concurrent calls on one engine from
// different threads each keep their own message, as errno does.
thread_local std::unordered_map<const hipsw_engine *, std::string> tl_errors;
}  // namespace

void hipsw_engine::set_error(const std::string &msg) const { tl_errors[this] = msg; }

bool hipsw_engine::fail(const char *what, hipError_t e) const
{
    char buf[256];
    snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    set_error(buf);
    return false;
}

bool hipsw_engine::check(hipError_t e, const char *what) const
{
    if (e == hipSuccess) return true;
    return fail(what, e);
}

bool hipsw_engine::inject(int site)
{
    std::lock_guard<std::mutex> lk(fault_mu);
    if (fault_site == 0 || fault_site != site) return false;
    if (--fault_countdown > 0) return false;
    fault_site = 0;
    return true;
}

TuneTable hipsw_engine::tune_snapshot() const
{
    std::lock_guard<std::mutex> lk(tune_mu);
    return tune;
}

namespace hipsw {

const char *thread_error(const hipsw_engine *eng)
{
    auto it = tl_errors.find(eng);
    return it == tl_errors.end() ? "" : it->second.c_str();
}

void *DevBuf::get(size_t bytes)
{
    if (bytes <= cap && p) return p;
    release();
    if (hipMalloc(&p, std::max<size_t>(bytes, 256)) != hipSuccess) {
        p = nullptr;
        return nullptr;
    }
    cap = std::max<size_t>(bytes, 256);
    return p;
}

void DevBuf::release()
{
    if (p) hipFree(p);
    p = nullptr;
    cap = 0;
}

bool ExecCtx::init()
{
    return hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) == hipSuccess &&
           hipStreamCreateWithFlags(&side, hipStreamNonBlocking) == hipSuccess &&
           hipEventCreate(&ev0) == hipSuccess && hipEventCreate(&ev1) == hipSuccess &&
           hipEventCreateWithFlags(&evf, hipEventDisableTiming) == hipSuccess &&
           hipEventCreateWithFlags(&evj, hipEventDisableTiming) == hipSuccess &&
           hipMalloc((void **)&counters, 64 * sizeof(uint32_t)) == hipSuccess;
}

void ExecCtx::destroy()
{
    if (stream) hipStreamSynchronize(stream);
    if (side) hipStreamSynchronize(side);
    arena.release();
    scratch.release();
    nw.release();
    if (counters) hipFree(counters);
    if (ev0) hipEventDestroy(ev0);
    if (ev1) hipEventDestroy(ev1);
    if (evf) hipEventDestroy(evf);
    if (evj) hipEventDestroy(evj);
    if (side) hipStreamDestroy(side);
    if (stream) hipStreamDestroy(stream);
    counters = nullptr;
    ev0 = ev1 = evf = evj = nullptr;
    stream = side = nullptr;
}

SlotLease::SlotLease(hipsw_engine *eng) : eng_(eng), idx_(-1)
{
    std::unique_lock<std::mutex> lk(eng->slot_mu);
    eng->slot_cv.wait(lk, [&] { return !eng->free_slots.empty(); });
    idx_ = eng->free_slots.back();
    eng->free_slots.pop_back();
}

SlotLease::~SlotLease()
{
    {
        std::lock_guard<std::mutex> lk(eng_->slot_mu);
        eng_->free_slots.push_back(idx_);
    }
    eng_->slot_cv.notify_one();
}

}  // namespace hipsw

namespace hipsw {

double now_ms()
{
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// Runs fn(lo, hi) over [0, n) on up to `threads` host threads.
void parallel_for(int64_t n, int threads, const std::function<void(int64_t, int64_t)> &fn)
{
    if (threads <= 1 || n < 4096) {
        fn(0, n);
        return;
    }
    std::vector<std::thread> th;
    const int64_t chunk = (n + threads - 1) / threads;
    for (int k = 0; k < threads; ++k) {
        const int64_t lo = k * chunk, hi = std::min(n, lo + chunk);
        if (lo >= hi) break;
        th.emplace_back([&fn, lo, hi] { fn(lo, hi); });
    }
    for (auto &t : th) t.join();
}

void parallel_chunks(int chunks, const std::function<void(int)> &fn)
{
    if (chunks <= 1) {
        if (chunks == 1) fn(0);
        return;
    }
    std::vector<std::thread> th;
    for (int c = 0; c < chunks; ++c) th.emplace_back([&fn, c] { fn(c); });
    for (auto &t : th) t.join();
}

int host_threads()
{
    unsigned h = std::thread::hardware_concurrency();
    return (int)std::min(16u, std::max(1u, h));
}

bool scoring_valid(const hipsw_scoring *sc)
{
    return sc && sc->e_ins > 0 && sc->e_del > 0 && sc->o_ins >= 0 && sc->o_del >= 0;
}

ExtParams make_params(const hipsw_scoring *sc)
{
    ExtParams p;
    memset(&p, 0, sizeof p);
    int max_mat = 0;                       // BandedSWScalar.swift:54-58 (starts at 0)
    for (int k = 0; k < 25; ++k) {
        p.mat[k] = sc->mat[k];
        max_mat = std::max(max_mat, (int)sc->mat[k]);
    }
    p.o_del = sc->o_del; p.e_del = sc->e_del;
    p.o_ins = sc->o_ins; p.e_ins = sc->e_ins;
    p.zdrop = sc->zdrop;
    p.oe_del = sc->o_del + sc->e_del;
    p.oe_ins = sc->o_ins + sc->e_ins;
    p.max_mat = max_mat;
    p.match_sc = sc->mat[0];
    p.mis_sc = sc->mat[1];
    p.std_mat = 1;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            if (sc->mat[i * 5 + j] != (i == j ? p.match_sc : p.mis_sc)) p.std_mat = 0;
    return p;
}

namespace {

constexpr uint64_t kHigh6 = 0xFCFCFCFCFCFCFCFCull;   // any byte >= 4

inline uint64_t load8(const uint8_t *p)
{
    uint64_t x;
    memcpy(&x, p, 8);
    return x;
}

// Eight codes 0..3, one per byte, to 16 bits (byte k at bits 2k..2k+1).
inline uint32_t gather8(uint64_t x)
{
    x = (x | (x >> 6)) & 0x000F000F000F000Full;
    x = (x | (x >> 12)) & 0x000000FF000000FFull;
    return (uint32_t)((x | (x >> 24)) & 0xFFFFu);
}

}  // namespace

// 2-bit packing of one sequence into `dst` (16 bases per word, base k at bits
// 2k..2k+1; N -> code 0 plus an N bit).  Returns false on a code > 4.  Whole
// words of codes 0..3 go eight bytes at a time; a word holding an N (or a bad
// code) and the last partial word go byte by byte.
bool pack2(const uint8_t *src, int len, uint32_t *dst, uint32_t *ndst)
{
    const int words = (len + 15) / 16;
    for (int wdx = 0; wdx < words; ++wdx) {
        const int lo = wdx * 16, hi = std::min(len, lo + 16);
        if (hi - lo == 16) {
            const uint64_t a = load8(src + lo), b = load8(src + lo + 8);
            if (((a | b) & kHigh6) == 0) {
                dst[wdx] = gather8(a) | (gather8(b) << 16);
                continue;
            }
        }
        uint32_t v = 0;
        for (int k = lo; k < hi; ++k) {
            const uint8_t c = src[k];
            if (c > 4) return false;
            v |= (uint32_t)(c & 3) << ((k - lo) * 2);
        }
        dst[wdx] = v;
    }
    if (ndst) {
        const int nw = (len + 31) / 32;
        for (int wdx = 0; wdx < nw; ++wdx) {
            uint32_t v = 0;
            const int lo = wdx * 32, hi = std::min(len, lo + 32);
            for (int k = lo; k < hi; ++k)
                if (src[k] == 4) v |= 1u << (k - lo);
            ndst[wdx] = v;
        }
    }
    return true;
}

int pack2_scan(const uint8_t *src, int len, uint32_t *dst)
{
    const int words = (len + 15) / 16;
    int any_n = 0;
    for (int wdx = 0; wdx < words; ++wdx) {
        const int lo = wdx * 16, hi = std::min(len, lo + 16);
        if (hi - lo == 16) {
            const uint64_t a = load8(src + lo), b = load8(src + lo + 8);
            if (((a | b) & kHigh6) == 0) {
                dst[wdx] = gather8(a) | (gather8(b) << 16);
                continue;
            }
        }
        uint32_t v = 0;
        for (int k = lo; k < hi; ++k) {
            const uint8_t c = src[k];
            if (c > 4) return -1;
            any_n |= c == 4;
            v |= (uint32_t)(c & 3) << ((k - lo) * 2);
        }
        dst[wdx] = v;
    }
    return any_n;
}

void nmask_fill(const uint8_t *src, int len, uint32_t *ndst)
{
    const int nw = (len + 31) / 32;
    for (int wdx = 0; wdx < nw; ++wdx) {
        uint32_t v = 0;
        const int lo = wdx * 32, hi = std::min(len, lo + 32);
        for (int k = lo; k < hi; ++k)
            if (src[k] == 4) v |= 1u << (k - lo);
        ndst[wdx] = v;
    }
}

bool has_n(const uint8_t *src, int len)
{
    int k = 0;
    for (; k + 8 <= len; k += 8)
        if (load8(src + k) & kHigh6) return true;
    for (; k < len; ++k)
        if (src[k] >= 4) return true;
    return false;
}

bool pack_jobs(int n, const uint8_t *queries, const uint32_t *q_off, const uint16_t *q_len,
               const uint8_t *targets, const uint32_t *t_off, const uint16_t *t_len,
               PackedJobs &pj, std::string &err)
{
    const int threads = host_threads();
    pj.qw.resize(n); pj.tw.resize(n); pj.qnw.resize(n); pj.tnw.resize(n);
    pj.ql.resize(n); pj.tl.resize(n);
    std::vector<uint8_t> qn_flag(n), tn_flag(n);
    // sequences back to back from word boundaries: chunks of jobs sum their
    // words, the chunk totals are scanned, then each chunk writes its offsets
    const int nch = (threads > 1 && n >= 4096) ? threads : 1;
    const int64_t chunk = ((int64_t)n + nch - 1) / nch;
    std::vector<uint64_t> cw(nch + 1, 0), cn(nch + 1, 0);
    parallel_chunks(nch, [&](int c) {
        uint64_t a = 0, b = 0;
        for (int64_t k = c * chunk, e = std::min<int64_t>(n, (c + 1) * chunk); k < e; ++k) {
            qn_flag[k] = has_n(queries + q_off[k], q_len[k]);
            tn_flag[k] = has_n(targets + t_off[k], t_len[k]);
            a += (q_len[k] + 15) / 16 + (t_len[k] + 15) / 16;
            b += (qn_flag[k] ? (q_len[k] + 31) / 32 : 0) + (tn_flag[k] ? (t_len[k] + 31) / 32 : 0);
        }
        cw[c + 1] = a;
        cn[c + 1] = b;
    });
    for (int c = 0; c < nch; ++c) {
        cw[c + 1] += cw[c];
        cn[c + 1] += cn[c];
    }
    uint64_t words = cw[nch], nwords = cn[nch];
    if (words + 64 > 0xFFFFFFF0ull) {
        err = "job set exceeds 2^32 sequence words";
        return false;
    }
    parallel_chunks(nch, [&](int c) {
        uint64_t a = cw[c], b = cn[c];
        for (int64_t k = c * chunk, e = std::min<int64_t>(n, (c + 1) * chunk); k < e; ++k) {
            pj.ql[k] = q_len[k];
            pj.tl[k] = t_len[k];
            pj.qw[k] = (uint32_t)a;
            a += (q_len[k] + 15) / 16;
            pj.tw[k] = (uint32_t)a;
            a += (t_len[k] + 15) / 16;
            pj.qnw[k] = qn_flag[k] ? (uint32_t)b : kNoN;
            if (qn_flag[k]) b += (q_len[k] + 31) / 32;
            pj.tnw[k] = tn_flag[k] ? (uint32_t)b : kNoN;
            if (tn_flag[k]) b += (t_len[k] + 31) / 32;
        }
    });
    words += 64;
    nwords += 64;
    // every sequence word is written by the packing below; only the padding
    // is cleared here
    pj.seq.resize(words);
    std::fill(pj.seq.end() - 64, pj.seq.end(), 0u);
    pj.nmask.assign(nwords, 0);
    std::atomic<bool> bad{false};
    parallel_for(n, threads, [&](int64_t lo, int64_t hi) {
        for (int64_t k = lo; k < hi; ++k) {
            if (!pack2(queries + q_off[k], q_len[k], pj.seq.data() + pj.qw[k],
                       qn_flag[k] ? pj.nmask.data() + pj.qnw[k] : nullptr))
                bad = true;
            if (!pack2(targets + t_off[k], t_len[k], pj.seq.data() + pj.tw[k],
                       tn_flag[k] ? pj.nmask.data() + pj.tnw[k] : nullptr))
                bad = true;
        }
    });
    if (bad) {
        err = "base code > 4";
        return false;
    }
    return true;
}

void *pinned_or_heap_alloc(size_t bytes)
{
    constexpr size_t kHdr = 64, kPinMax = size_t(1) << 30;
    void *raw = nullptr;
    uint32_t pinned = 0;
    if (bytes <= kPinMax && hipHostMalloc(&raw, bytes + kHdr, hipHostMallocDefault) == hipSuccess) {
        pinned = 1;
    } else {
        (void)hipGetLastError();                               // clear a failed pin
        raw = ::operator new(bytes + kHdr);
    }
    *static_cast<uint32_t *>(raw) = pinned;
    return static_cast<char *>(raw) + kHdr;
}

void pinned_or_heap_free(void *p) noexcept
{
    if (!p) return;
    void *raw = static_cast<char *>(p) - 64;
    if (*static_cast<uint32_t *>(raw)) (void)hipHostFree(raw);
    else ::operator delete(raw);
}

void radix_sort_by_key(std::vector<uint32_t> &keys, std::vector<uint32_t> &idx)
{
    // Stable LSD sort, four 8-bit digits.  The jobs are cut into the chunks
    // parallel_for hands its threads; per pass each chunk counts its digits,
    // the per-chunk offsets follow digit by digit and chunk by chunk, and each
    // chunk scatters its own elements in order, so the result is the serial
    // stable sort's.
    const int64_t n = (int64_t)keys.size();
    const int threads = host_threads();
    const bool par = threads > 1 && n >= 4096;
    const int64_t chunk = par ? (n + threads - 1) / threads : std::max<int64_t>(n, 1);
    const int nc = (int)((n + chunk - 1) / chunk);
    std::vector<uint32_t> k2(n), i2(n);
    std::vector<uint32_t> hist((size_t)std::max(nc, 1) * 256);
    for (int shift = 0; shift < 32; shift += 8) {
        std::fill(hist.begin(), hist.end(), 0u);
        parallel_for(n, threads, [&](int64_t lo, int64_t hi) {
            uint32_t *h = hist.data() + (size_t)(lo / chunk) * 256;
            for (int64_t a = lo; a < hi; ++a) h[(keys[a] >> shift) & 0xFFu]++;
        });
        uint32_t run = 0;
        for (int d = 0; d < 256; ++d)
            for (int c = 0; c < nc; ++c) {
                const uint32_t v = hist[(size_t)c * 256 + d];
                hist[(size_t)c * 256 + d] = run;
                run += v;
            }
        parallel_for(n, threads, [&](int64_t lo, int64_t hi) {
            uint32_t *h = hist.data() + (size_t)(lo / chunk) * 256;
            for (int64_t a = lo; a < hi; ++a) {
                const uint32_t pos = h[(keys[a] >> shift) & 0xFFu]++;
                k2[pos] = keys[a];
                i2[pos] = idx[a];
            }
        });
        keys.swap(k2);
        idx.swap(i2);
    }
}

}  // namespace hipsw

// ------------------------------------------------------------------------
// Engine lifetime
// ------------------------------------------------------------------------
extern "C" hipsw_engine *hipsw_engine_create(int device)
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0 || device < 0 || device >= ndev)
        return nullptr;
    if (hipSetDevice(device) != hipSuccess) return nullptr;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return nullptr;
    auto *eng = new hipsw_engine();
    eng->device = device;
    eng->n_cu = prop.multiProcessorCount;
    eng->lds_per_cu = (int)prop.maxSharedMemoryPerMultiProcessor;
    // Dynamic LDS per block is kept within 64 KiB, the size every launch may
    // request without an attribute change.
    eng->lds_per_block = (int)std::min<size_t>(prop.sharedMemPerBlock, 65536);
    eng->name = prop.name;
    eng->arch = prop.gcnArchName;
    if (!eng->main.init()) {
        delete eng;
        return nullptr;
    }
    for (int k = 0; k < kEngineSlots; ++k) {
        auto c = std::make_unique<ExecCtx>();
        const bool ok = c->init();
        eng->slots.push_back(std::move(c));
        if (!ok) {
            delete eng;
            return nullptr;
        }
        eng->free_slots.push_back(k);
    }
    eng->noalloc_scratch = noalloc_scratch_address(device);
    if (!eng->noalloc_scratch) {
        delete eng;
        return nullptr;
    }
    return eng;
}

hipsw_engine::~hipsw_engine()
{
    hipSetDevice(device);
    for (auto &c : slots) c->destroy();
    main.destroy();
}

extern "C" void hipsw_engine_destroy(hipsw_engine *eng) { delete eng; }

extern "C" const char *hipsw_last_error(const hipsw_engine *eng)
{
    return eng ? thread_error(eng) : "no engine";
}

extern "C" int32_t hipsw_min_batch_size(void) { return 32; }

extern "C" int hipsw_fault_inject(hipsw_engine *eng, int32_t site, int32_t countdown)
{
    if (!eng || site < HIPSW_FAULT_NONE || site > HIPSW_FAULT_COPY || countdown < 0)
        return HIPSW_E_INVALID;
    std::lock_guard<std::mutex> lk(eng->fault_mu);
    eng->fault_site = countdown > 0 ? site : HIPSW_FAULT_NONE;
    eng->fault_countdown = countdown;
    return HIPSW_OK;
}

extern "C" int hipsw_device_info(const hipsw_engine *eng, char *name, int name_cap, int32_t *n_cu)
{
    if (!eng) return HIPSW_E_INVALID;
    if (name && name_cap > 0) {
        std::string s = eng->name + " (" + eng->arch + ")";
        snprintf(name, (size_t)name_cap, "%s", s.c_str());
    }
    if (n_cu) *n_cu = eng->n_cu;
    return HIPSW_OK;
}
