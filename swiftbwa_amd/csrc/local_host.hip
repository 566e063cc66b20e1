// Host side of the mate-rescue local-SW entry point (include/hipsw.h).
//
// Replaces LocalSWDispatcher.dispatchBatch (Sources/MetalSW/LocalSWDispatcher.swift:
// 31-95) with LocalSWAligner.align semantics (LocalSWAligner.swift:42-79): the
// striped 8-bit pass for tasks with qLen*A <= 250, and the scalar int32 pass for
// the rest and for every striped task that overflowed.
#include "engine_internal.h"
#include "local_internal.h"

#include <algorithm>
#include <cstring>
#include <vector>

using namespace hipsw;

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
}  // namespace

extern "C" int hipsw_local_sw(hipsw_engine *eng, int32_t n, const uint8_t *queries,
                              const uint32_t *q_off, const uint16_t *q_len,
                              const uint8_t *targets, const uint32_t *t_off,
                              const uint16_t *t_len, int32_t min_sub_score, int32_t match_score,
                              int32_t mismatch_penalty, const hipsw_scoring *sc, int32_t *out6,
                              int32_t *ok, hipsw_stats *stats)
{
    if (!eng) return HIPSW_E_INVALID;
    if (n == 0) return HIPSW_OK;
    if (n < 0 || !out6 || !ok || !queries || !q_off || !q_len || !targets || !t_off || !t_len ||
        !sc) {
        eng->set_error("hipsw_local_sw: invalid argument");
        return HIPSW_E_INVALID;
    }
    if (mismatch_penalty < 0 || mismatch_penalty > 255 || sc->o_ins + sc->e_ins < 0 ||
        sc->o_ins + sc->e_ins > 255 || sc->e_ins < 0 || sc->e_ins > 255) {
        // LocalSWAligner converts bias, gapOE and gapE to UInt8 (:53, :176-177)
        eng->set_error("hipsw_local_sw: penalties must fit in 8 bits");
        return HIPSW_E_UNSUPPORTED;
    }
    const double t0 = now_ms();
    PackedJobs pj;
    std::string err;
    if (!pack_jobs(n, queries, q_off, q_len, targets, t_off, t_len, pj, err)) {
        eng->set_error("hipsw_local_sw: " + err);
        return HIPSW_E_INVALID;
    }
    LocalParams p;
    memset(&p, 0, sizeof p);
    int max_u8 = 0, max_i = 0;
    for (int k = 0; k < 25; ++k) {
        p.mat[k] = sc->mat[k];
        max_u8 = std::max(max_u8, std::min(std::max((int)sc->mat[k] + mismatch_penalty, 0), 255));
        max_i = std::max(max_i, (int)sc->mat[k]);
    }
    p.bias = mismatch_penalty;
    p.gap_oe = sc->o_ins + sc->e_ins;
    p.gap_e = sc->e_ins;
    p.min_sub = std::min(std::max(min_sub_score, 0), 255);
    p.max_base_u8 = max_u8 > mismatch_penalty ? max_u8 - mismatch_penalty : 0;
    p.max_base_i = max_i;

    std::vector<uint32_t> striped, scalar;
    int tmax = 0;
    for (int k = 0; k < n; ++k) {
        const int q = pj.ql[k], t = pj.tl[k];
        if (q <= 0 || t <= 0) continue;                        // :48, nil
        tmax = std::max(tmax, t);
        if ((int64_t)q * match_score <= 250 && q <= 256) {
            striped.push_back((uint32_t)k);
        } else {
            scalar.push_back((uint32_t)k);
        }
    }
    const size_t words = pj.seq.size(), nwords = pj.nmask.size();
    const size_t colmax_bytes = p.min_sub > 0 ? (size_t)striped.size() * tmax : 1;
    const size_t bytes = span<uint32_t>(words) + span<uint32_t>(nwords) + 4 * span<uint32_t>(n) +
                         2 * span<int32_t>(n) + span<int32_t>((size_t)n * 6) + span<int32_t>(n) +
                         2 * span<uint32_t>(n) + span<uint8_t>(colmax_bytes);
    SlotLease lease(eng);
    ExecCtx &c = lease.ctx();
    hipSetDevice(eng->device);
    void *dmem = eng->inject(HIPSW_FAULT_ALLOC) ? nullptr : c.arena.get(bytes);
    if (!dmem) {
        eng->set_error("hipsw_local_sw: arena allocation failed (injected or out of memory)");
        return HIPSW_E_DEVICE;
    }
    char *cur = reinterpret_cast<char *>(dmem);
    uint32_t *d_seq = carve<uint32_t>(cur, words);
    uint32_t *d_nmask = carve<uint32_t>(cur, nwords);
    uint32_t *d_qw = carve<uint32_t>(cur, n), *d_tw = carve<uint32_t>(cur, n);
    uint32_t *d_qn = carve<uint32_t>(cur, n), *d_tn = carve<uint32_t>(cur, n);
    int32_t *d_ql = carve<int32_t>(cur, n), *d_tl = carve<int32_t>(cur, n);
    int32_t *d_out = carve<int32_t>(cur, (size_t)n * 6);
    int32_t *d_status = carve<int32_t>(cur, n);
    uint32_t *d_ord_s = carve<uint32_t>(cur, n), *d_ord_c = carve<uint32_t>(cur, n);
    uint8_t *d_colmax = carve<uint8_t>(cur, colmax_bytes);
    bool good = true;
    if (eng->inject(HIPSW_FAULT_COPY)) {
        eng->set_error("hipsw_local_sw: upload failed (injected)");
        good = false;
    }
    auto up = [&](void *dst, const void *src, size_t b) {
        if (good && b) good = eng->check(hipMemcpyAsync(dst, src, b, hipMemcpyHostToDevice, c.stream), "upload");
    };
    up(d_seq, pj.seq.data(), words * 4);
    up(d_nmask, pj.nmask.data(), nwords * 4);
    up(d_qw, pj.qw.data(), (size_t)n * 4);
    up(d_tw, pj.tw.data(), (size_t)n * 4);
    up(d_qn, pj.qnw.data(), (size_t)n * 4);
    up(d_tn, pj.tnw.data(), (size_t)n * 4);
    up(d_ql, pj.ql.data(), (size_t)n * 4);
    up(d_tl, pj.tl.data(), (size_t)n * 4);
    std::vector<uint32_t> grouped(striped.size());
    int32_t grp_off[17];
    group_by_stripes(striped.data(), striped.size(), pj.ql.data(), grouped.data(), grp_off);
    up(d_ord_s, grouped.data(), grouped.size() * 4);
    const LocalDev dev{d_seq, d_nmask, d_qw, d_tw, d_qn, d_tn, d_ql, d_tl, d_out, d_status};
    double kernel_ms = 0;
    int64_t launches = 0;
    std::vector<int32_t> status(n, LOCAL_NIL);
    if (good) good = eng->check(hipMemsetAsync(d_status, 0, (size_t)n * 4, c.stream), "memset");
    if (good && eng->inject(HIPSW_FAULT_LAUNCH)) {
        eng->set_error("hipsw_local_sw: kernel launch failed (injected)");
        good = false;
    }
    if (good && !striped.empty()) {
        good = eng->check(hipEventRecord(c.ev0, c.stream), "event") &&
               eng->check(launch_local_striped(dev, p, d_ord_s, grp_off, d_colmax,
                                               p.min_sub > 0 ? tmax : 0, c.stream),
                          "launch local_striped") &&
               eng->check(hipEventRecord(c.ev1, c.stream), "event") &&
               eng->check(hipEventSynchronize(c.ev1), "local_striped");
        if (good) {
            float ms = 0;
            hipEventElapsedTime(&ms, c.ev0, c.ev1);
            kernel_ms += ms;
            ++launches;
            good = eng->check(hipMemcpyAsync(status.data(), d_status, (size_t)n * 4,
                                             hipMemcpyDeviceToHost, c.stream), "download status") &&
                   eng->check(hipStreamSynchronize(c.stream), "download status");
        }
        if (good)
            for (uint32_t k : striped)
                if (status[k] == LOCAL_NEEDS_SCALAR) scalar.push_back(k);
    }
    if (good && !scalar.empty()) {
        int qc = 1, tc = 1;
        for (uint32_t k : scalar) { qc = std::max(qc, pj.ql[k]); tc = std::max(tc, pj.tl[k]); }
        const int blocks = (int)((scalar.size() + 63) / 64);
        const size_t sbytes = (size_t)blocks * 64 * (2 * (tc + 1) + qc) * sizeof(int);
        int *scr = eng->inject(HIPSW_FAULT_ALLOC) ? nullptr : reinterpret_cast<int *>(c.scratch.get(sbytes));
        good = scr != nullptr;
        if (!good) eng->set_error("hipsw_local_sw: scratch allocation failed (injected or out of memory)");
        if (good) good = eng->check(hipMemcpyAsync(d_ord_c, scalar.data(), scalar.size() * 4,
                                                   hipMemcpyHostToDevice, c.stream), "upload");
        if (good)
            good = eng->check(hipEventRecord(c.ev0, c.stream), "event") &&
                   eng->check(launch_local_scalar(dev, p, d_ord_c, (int)scalar.size(), scr, tc, qc,
                                                  c.stream), "launch local_scalar") &&
                   eng->check(hipEventRecord(c.ev1, c.stream), "event") &&
                   eng->check(hipEventSynchronize(c.ev1), "local_scalar");
        if (good) {
            float ms = 0;
            hipEventElapsedTime(&ms, c.ev0, c.ev1);
            kernel_ms += ms;
            ++launches;
        }
    }
    if (good && eng->inject(HIPSW_FAULT_COPY)) {
        eng->set_error("hipsw_local_sw: download failed (injected)");
        good = false;
    }
    if (good) good = eng->check(hipMemcpyAsync(out6, d_out, (size_t)n * 6 * 4, hipMemcpyDeviceToHost,
                                               c.stream), "download");
    if (good) good = eng->check(hipMemcpyAsync(status.data(), d_status, (size_t)n * 4,
                                               hipMemcpyDeviceToHost, c.stream), "download");
    if (good) good = eng->check(hipStreamSynchronize(c.stream), "download");
    if (!good) {
        (void)hipStreamSynchronize(c.stream);   // queued copies read host arrays freed on return
        return HIPSW_E_DEVICE;
    }
    for (int k = 0; k < n; ++k) {
        ok[k] = status[k] == LOCAL_OK ? 1 : 0;
        if (!ok[k]) memset(out6 + (size_t)k * 6, 0, 6 * sizeof(int32_t));
    }
    if (stats) {
        // Cells as LocalSWAligner's loops count them (LocalSWAligner.swift:209,
        // :385): the forward pass's qlen x tlen, again for a task redone by the
        // scalar pass, plus the reverse pass over the (qe+1) x (te+1) prefixes
        // of a found alignment.  A striped pass that stops at an overflow is
        // counted whole.
        int64_t cells = 0;
        for (uint32_t k : striped) cells += (int64_t)pj.ql[k] * pj.tl[k];
        for (uint32_t k : scalar) cells += (int64_t)pj.ql[k] * pj.tl[k];
        for (int k = 0; k < n; ++k)
            if (ok[k]) cells += (int64_t)(out6[(size_t)k * 6 + 2] + 1) * (out6[(size_t)k * 6 + 4] + 1);
        memset(stats, 0, sizeof *stats);
        stats->kernel_ms = kernel_ms;
        stats->total_ms = now_ms() - t0;
        stats->n_jobs = n;
        stats->launches = launches;
        stats->promoted = (int64_t)scalar.size();
        stats->cells = cells;
    }
    return HIPSW_OK;
}
