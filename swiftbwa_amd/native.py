"""ctypes binding of libhipsw.so, the C ABI declared in include/hipsw.h.

The library is built in-tree by ``swiftbwa_amd.build``.  Loading fails loudly
when it is missing: there is no CPU fallback inside this package (callers that
want one keep their own CPU path, as swiftbwa does when MetalSW is nil).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhipsw.so")

HIPSW_OK = 0
HIPSW_E_INVALID = -1
HIPSW_E_DEVICE = -2
HIPSW_E_UNSUPPORTED = -3
HIPSW_E_BUDGET = -4
HIPSW_E_PENDING = -5

FAULT_NONE, FAULT_ALLOC, FAULT_LAUNCH, FAULT_COPY = 0, 1, 2, 3

MODE_PLAIN = 0
MODE_BATCHED = 1
MODE_INPUT_ORDER = 2      # hipsw_ext_run only: batched kernel, input order, no binning


class HipswScoring(C.Structure):
    _fields_ = [("mat", C.c_int8 * 25), ("o_ins", C.c_int32), ("e_ins", C.c_int32),
                ("o_del", C.c_int32), ("e_del", C.c_int32), ("zdrop", C.c_int32)]


class HipswStats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("total_ms", C.c_double), ("n_jobs", C.c_int64),
                ("launches", C.c_int64), ("promoted", C.c_int64), ("tier_jobs", C.c_int64 * 3),
                ("cells", C.c_int64)]

    def as_dict(self):
        return {"kernel_ms": self.kernel_ms, "total_ms": self.total_ms, "n_jobs": self.n_jobs,
                "launches": self.launches, "promoted": self.promoted,
                "tier_jobs": list(self.tier_jobs), "cells": self.cells}


class HipswRegion(C.Structure):
    _fields_ = [("rb", C.c_int64), ("re", C.c_int64), ("qb", C.c_int32), ("qe", C.c_int32),
                ("score", C.c_int32), ("true_score", C.c_int32), ("sub", C.c_int32),
                ("sub_n", C.c_int32), ("w", C.c_int32), ("seed_cov", C.c_int32),
                ("seed_len0", C.c_int32), ("reserved", C.c_int32)]


class HipswExtOpts(C.Structure):
    _fields_ = [("match_score", C.c_int32), ("mismatch_penalty", C.c_int32),
                ("band_width", C.c_int32), ("pen_clip5", C.c_int32), ("pen_clip3", C.c_int32),
                ("min_seed_len", C.c_int32)]


class HipswPacRef(C.Structure):
    _fields_ = [("pac", C.c_void_p), ("genome_len", C.c_int64)]


class HipswTuneEntry(C.Structure):
    _fields_ = [("tier", C.c_int32), ("length_class", C.c_int32), ("jobs", C.c_int32),
                ("kind", C.c_int32), ("lds_blocks", C.c_int32), ("chosen", C.c_int32),
                ("ms", C.c_double)]


class HipswChunk(C.Structure):
    _P = C.c_void_p
    _fields_ = [("n_ext", C.c_int32), ("ext_q", _P), ("ext_q_off", _P), ("ext_q_len", _P),
                ("ext_t_pos", _P), ("ext_t_len", _P), ("ext_t_rev", _P),
                ("ext_h0", _P), ("ext_w", _P), ("ext_eb", _P),
                ("n_loc", C.c_int32), ("loc_q", _P), ("loc_q_off", _P), ("loc_q_len", _P),
                ("loc_t_pos", _P), ("loc_t_len", _P), ("loc_min_sub", C.c_int32),
                ("n_nw", C.c_int32), ("nw_q", _P), ("nw_q_off", _P), ("nw_q_len", _P),
                ("nw_t_pos", _P), ("nw_t_len", _P), ("nw_w", _P), ("nw_true_score", _P),
                ("loc_rid", _P)]


class HipswChunkResult(C.Structure):
    _P = C.c_void_p
    _fields_ = [("ext_out", _P), ("loc_out", _P), ("loc_ok", _P), ("nw_score", _P),
                ("nw_cigar", _P), ("nw_n_cigar", _P), ("nw_flags", _P),
                ("latency_ms", C.c_double), ("queue_ms", C.c_double), ("run_ms", C.c_double),
                ("cells", C.c_int64), ("host_ms", C.c_double), ("nw_cigar_off", _P),
                ("est_bytes", C.c_int64), ("dev_peak_bytes", C.c_int64), ("loc_win", _P)]


class HipswChainPass(C.Structure):
    _P = C.c_void_p
    _fields_ = [("index", C.c_int32), ("n", C.c_int32), ("q", _P), ("q_off", _P),
                ("q_len", _P), ("t_pos", _P), ("t_len", _P), ("t_rev", _P), ("h0", _P),
                ("w", _P), ("eb", _P)]


class HipswPlanEntry(C.Structure):
    _fields_ = [("tier", C.c_int32), ("length_class", C.c_int32), ("has_n", C.c_int32),
                ("jobs", C.c_int32), ("kind", C.c_int32), ("lds_blocks", C.c_int32),
                ("items", C.c_int32), ("grid", C.c_int32)]


# Every symbol include/hipsw.h declares, with (restype, argtypes).
_P = C.c_void_p
_I32P = C.POINTER(C.c_int32)
_SIGS = {
    "hipsw_engine_create": (_P, [C.c_int]),
    "hipsw_engine_destroy": (None, [_P]),
    "hipsw_last_error": (C.c_char_p, [_P]),
    "hipsw_min_batch_size": (C.c_int32, []),
    "hipsw_fault_inject": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "hipsw_device_info": (C.c_int, [_P, C.c_char_p, C.c_int, _I32P]),
    "hipsw_banded_extend": (C.c_int, [_P, C.c_int32, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                      C.POINTER(HipswScoring), C.c_int32, _P,
                                      C.POINTER(HipswStats)]),
    "hipsw_banded_extend_noalloc": (C.c_int, [_P, C.c_int32, _P, _P, _P, _P, _P, _P, _P, _P,
                                              _P, C.POINTER(HipswScoring), C.c_int32, _P,
                                              C.POINTER(HipswStats)]),
    "hipsw_ext_upload": (_P, [_P, C.c_int32, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                              C.POINTER(HipswScoring)]),
    "hipsw_ext_run": (C.c_int, [_P, _P, C.c_int32, C.POINTER(HipswStats)]),
    "hipsw_ext_plan": (C.c_int, [_P, _P, _P, C.c_int32, _P]),
    "hipsw_ext_download": (C.c_int, [_P, _P, _P]),
    "hipsw_ext_free": (None, [_P, _P]),
    "hipsw_global_nw": (C.c_int, [_P, C.c_int32, _P, _P, _P, _P, _P, _P, _P, _P, C.c_int32,
                                  C.c_int32, C.POINTER(HipswScoring), C.c_int32, _P, _P, _P,
                                  _P, _P, C.POINTER(HipswStats)]),
    "hipsw_cigar_generate": (C.c_int, [_P, C.c_int32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                       _P, _P, _P, C.POINTER(HipswScoring), C.c_int32, C.c_int32,
                                       C.c_int32, C.c_int32, _P, _P, _P, _P, _P, _P, _P,
                                       C.POINTER(HipswStats)]),
    "hipsw_cigar_route": (C.c_int, [C.c_int32, _P, _P, _P, _P, _P, _P, C.POINTER(HipswScoring),
                                    C.c_int32, C.c_int32, _P, _P]),
    "hipsw_cigar_finish": (C.c_int, [C.c_int32] + [_P] * 18 + [C.c_int32, C.c_int32, C.c_int32]
                           + [_P] * 7),
    "hipsw_get_reference_pac": (C.c_int32, [_P, C.c_int64, C.c_int32, _P]),
    "hipsw_chain_begin": (_P, [C.c_int32, _P, _P, _P, _P, _P, _P, C.c_int32, _P, _P, _P,
                               C.POINTER(HipswScoring), C.POINTER(HipswExtOpts)]),
    "hipsw_chain_next": (C.c_int, [_P, C.POINTER(HipswChainPass)]),
    "hipsw_chain_feed": (C.c_int, [_P, _P]),
    "hipsw_chain_finish": (C.c_int, [_P, _P, _P]),
    "hipsw_chain_free": (None, [_P]),
    "hipsw_extend_chains": (C.c_int, [_P, C.c_int32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                      _P, C.POINTER(HipswScoring), C.POINTER(HipswExtOpts), _P,
                                      _P, C.POINTER(HipswStats)]),
    "hipsw_ext_autotune": (C.c_int, [_P, _P, C.c_int32, _P, C.c_int32, _P]),
    "hipsw_ext_set_layout": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32]),
    "hipsw_ext_set_layout16": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "hipsw_ext_set_layout16_kind": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32]),
    "hipsw_nw_upload": (_P, [_P, C.c_int32, _P, _P, _P, _P, _P, _P, _P, _P, C.c_int32,
                             C.c_int32, C.POINTER(HipswScoring)]),
    "hipsw_nw_run": (C.c_int, [_P, _P, C.c_int32, C.POINTER(HipswStats)]),
    "hipsw_nw_download": (C.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    "hipsw_nw_free": (None, [_P, _P]),
    "hipsw_nw_set_state": (C.c_int, [_P, C.c_int32]),
    "hipsw_local_sw": (C.c_int, [_P, C.c_int32, _P, _P, _P, _P, _P, _P, C.c_int32, C.c_int32,
                                 C.c_int32, C.POINTER(HipswScoring), _P, _P,
                                 C.POINTER(HipswStats)]),
    "hipsw_server_create": (_P, [_P, C.c_int32, C.c_int64, C.POINTER(HipswScoring),
                                 C.c_int32, C.c_int32, C.c_int32]),
    "hipsw_server_set_reference": (C.c_int, [_P, _P, C.c_int64]),
    "hipsw_server_set_contigs": (C.c_int, [_P, C.c_int32, _P, _P]),
    "hipsw_server_submit": (C.c_int64, [_P, C.POINTER(HipswChunk)]),
    "hipsw_server_poll": (C.c_int, [_P, C.c_int64]),
    "hipsw_server_wait": (C.c_int, [_P, C.c_int64, C.POINTER(HipswChunkResult)]),
    "hipsw_server_stats": (C.c_int, [_P, _P, _P, _P, _P]),
    "hipsw_server_destroy": (None, [_P]),
}

_lib = None


def load(path=LIB_PATH):
    """Loads libhipsw.so and binds every declared symbol; raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            "libhipsw.so is not built (%s); run `python -m swiftbwa_amd.build`" % path)
    lib = C.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def declared_symbols():
    return sorted(_SIGS)


def ptr(a):
    """Address of a contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


class HipswError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("hipsw status %d: %s" % (status, msg))
        self.status = status
