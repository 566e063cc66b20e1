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

hipError_t launch_local_striped(int snmax, const LocalDev &d, const LocalParams &p,
                                const uint32_t *order, int n, uint8_t *colmax, int tcap,
                                hipStream_t s);
hipError_t launch_local_scalar(const LocalDev &d, const LocalParams &p, const uint32_t *order,
                               int n, int *scratch, int tcap, int qcap, hipStream_t s);

}  // namespace hipsw
