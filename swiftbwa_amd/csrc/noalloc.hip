// Allocation-free banded extension (hipsw_banded_extend_noalloc, include/hipsw.h).
//
// The same work as hipsw_banded_extend -- BandedSWScalar.align per task
// (Sources/Alignment/BandedSWScalar.swift:21-191), plain or batched -- for
// callers that must not allocate device memory.  Everything the kernels need
// beyond the caller's inputs and outputs (the packed job arrays, the batched
// order and work items, the packed-form descriptors, the cell counter, the
// work-queue counters and the HBM column state of the hybrid layouts) lives in
// one static device array of kScratchBytes.  A call whose layout does not fit
// is refused with HIPSW_E_BUDGET before anything runs.
//
// Each device holds its own copy of the array for the life of the process.
// Calls on one device share it, so the library serialises them with a mutex
// per device; a call's kernels clear the counters they use before running.
#include "engine_internal.h"

#include <cstdio>
#include <mutex>

using namespace hipsw;

namespace {

constexpr size_t kScratchBytes = (size_t)64 << 20;
constexpr int kMaxDevices = 64;

std::mutex g_addr_mu;
void *g_addr[kMaxDevices] = {};          // device address of `scratch`, per device
std::mutex g_run_mu[kMaxDevices];         // one call at a time per device

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace hipsw {

// Called by hipsw_engine_create with `device` current.  The host-side name of
// the array is not a device address; every use goes through the address
// returned here.
void *noalloc_scratch_address(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_addr_mu);
    if (!g_addr[device]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
        g_addr[device] = p;
    }
    return g_addr[device];
}

}  // namespace hipsw

extern "C" int hipsw_banded_extend_noalloc(hipsw_engine *eng, int32_t n, const uint8_t *queries,
                                           const uint32_t *q_off, const uint16_t *q_len,
                                           const uint8_t *targets, const uint32_t *t_off,
                                           const uint16_t *t_len, const int32_t *h0,
                                           const int32_t *w, const int32_t *end_bonus,
                                           const hipsw_scoring *sc, int32_t mode, int32_t *out,
                                           hipsw_stats *stats)
{
    if (!eng) return HIPSW_E_INVALID;
    if (n == 0) return HIPSW_OK;
    if (n < 0 || !out || (mode != HIPSW_MODE_PLAIN && mode != HIPSW_MODE_BATCHED))
        return HIPSW_E_INVALID;
    if (!eng->noalloc_scratch || eng->device < 0 || eng->device >= kMaxDevices) {
        eng->set_error("hipsw_banded_extend_noalloc: static scratch not available on this device");
        return HIPSW_E_DEVICE;
    }
    const double t0 = now_ms();
    PrepLease prep(eng);
    ExtPrep &P = prep.get();
    std::string err;
    const int rc = ext_prepare(eng, eng->tune_snapshot(), n, queries, q_off, q_len, targets, t_off,
                               t_len, h0, w, end_bonus, sc, P, err);
    if (rc != HIPSW_OK) {
        eng->set_error("hipsw_banded_extend_noalloc: " + err);
        return rc;
    }
    // scratch layout: job arrays | HBM column state | 64 work-queue counters
    const hipsw_ext_set &S = P.shape;
    const size_t state = mode == HIPSW_MODE_PLAIN ? (S.plain_global ? S.plain_scratch : 0)
                                                  : S.batched_scratch;
    const size_t arena_b = align256(P.arena_bytes), state_b = align256(state);
    const size_t total = arena_b + state_b + 256;
    if (total > kScratchBytes) {
        char buf[192];
        snprintf(buf, sizeof buf,
                 "hipsw_banded_extend_noalloc: %d jobs need %zu bytes of device scratch; "
                 "the static scratch holds %zu",
                 n, total, kScratchBytes);
        eng->set_error(buf);
        return HIPSW_E_BUDGET;
    }
    char *base = static_cast<char *>(eng->noalloc_scratch);
    std::lock_guard<std::mutex> run_lock(g_run_mu[eng->device]);
    SlotLease lease(eng);
    hipSetDevice(eng->device);
    return ext_run_once(eng, lease.ctx(), P, mode, base, state ? base + arena_b : nullptr,
                        reinterpret_cast<uint32_t *>(base + arena_b + state_b), out, stats, t0);
}
