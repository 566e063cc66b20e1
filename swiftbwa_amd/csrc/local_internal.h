// Internal definitions of the mate-rescue local-SW path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hipsw_internal.h"

namespace hipsw {

enum LocalStatus : int32_t { LOCAL_OK = 1, LOCAL_NIL = 0, LOCAL_NEEDS_SCALAR = 2 };

struct LocalDev {
    const uint32_t *seq, *nmask;
    const uint32_t *q_word, *t_word, *qn_word, *tn_word;
    const int32_t *q_len, *t_len;
    int32_t *out;       // 6 per task: score, qb, qe, tb, te, score2
    int32_t *status;    // LocalStatus per task
};

struct LocalParams {
    int8_t mat[32];
    int32_t bias;        // mismatchPenalty (LocalSWAligner.swift:53)
    int32_t gap_oe;      // gapOpenPenalty + gapExtendPenalty (:51)
    int32_t gap_e;       // gapExtendPenalty (:52)
    int32_t min_sub;     // minSubScore (KSW_XSUBO threshold), 0 = off
    int32_t max_base_u8; // :191-199
    int32_t max_base_i;  // :406-412
};

// Stable counting sort of the task ids ids[0..n) by stripe count
// sn = ceil(q_len/16), 1..16 (every q_len in 1..256): out[off[sn-1]..off[sn])
// is group sn.
void group_by_stripes(const uint32_t *ids, size_t n, const int32_t *q_len, uint32_t *out,
                      int32_t off[17]);
// Striped launches over the grouped id list `order` on the device, one per
// group of at least a thousand tasks (smaller groups join the next larger
// one); colmax holds tcap bytes per position of order.
hipError_t launch_local_striped(const LocalDev &d, const LocalParams &p, const uint32_t *order,
                                const int32_t off[17], uint8_t *colmax, int tcap, hipStream_t s);
hipError_t launch_local_scalar(const LocalDev &d, const LocalParams &p, const uint32_t *order,
                               int n, int *scratch, int tcap, int qcap, hipStream_t s);

}  // namespace hipsw
