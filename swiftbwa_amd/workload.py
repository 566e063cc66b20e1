"""Deterministic synthetic job sets (libswwork.so, csrc/workload.c) as numpy arrays."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libswwork.so")

EXT_UNIFORM = 2   # 150-bp queries, 0..100-bp flanks, w=100, h0 ~ U[19,60]
NW = 3            # banded global NW pairs with a planted 0..20-bp indel
EXT_MIXED = 4     # uneven left/right extension jobs, w in {20,50,100}
LOCAL = 5         # 150-bp mates against 600-bp rescue windows

_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libswwork.so is not built; run `python -m swiftbwa_amd.build`")
        lib = C.CDLL(LIB_PATH)
        lib.wl_generate.restype = C.c_void_p
        lib.wl_generate.argtypes = [C.c_int, C.c_uint64, C.c_int64]
        lib.wl_field.restype = C.c_void_p
        lib.wl_field.argtypes = [C.c_void_p, C.c_int]
        lib.wl_count.restype = C.c_int64
        lib.wl_count.argtypes = [C.c_void_p, C.c_int]
        lib.wl_free.restype = None
        lib.wl_free.argtypes = [C.c_void_p]
        _lib = lib
    return _lib


_FIELDS = [("queries", np.uint8), ("q_off", np.uint32), ("q_len", np.uint16),
           ("targets", np.uint8), ("t_off", np.uint32), ("t_len", np.uint16),
           ("h0", np.int32), ("w", np.int32), ("end_bonus", np.int32),
           ("true_score", np.int32), ("direction", np.int32)]


def generate(kind, seed, n):
    """Returns a dict of numpy arrays describing n jobs of the given kind."""
    lib = _load()
    h = lib.wl_generate(int(kind), int(seed), int(n))
    if not h:
        raise ValueError("bad workload request")
    try:
        out = {}
        for k, (name, dt) in enumerate(_FIELDS):
            cnt = lib.wl_count(h, k)
            addr = lib.wl_field(h, k)
            nbytes = int(cnt) * np.dtype(dt).itemsize
            buf = (C.c_char * max(nbytes, 1)).from_address(addr)
            out[name] = np.frombuffer(buf, dtype=dt, count=int(cnt)).copy()
        out["n"] = int(n)
        return out
    finally:
        lib.wl_free(h)


def slice_jobs(js, idx):
    """Sub-set of a job dict (re-packed arenas) for the given job indices."""
    idx = np.asarray(idx, dtype=np.int64)
    res = {"n": len(idx)}
    for side, off, ln in (("queries", "q_off", "q_len"), ("targets", "t_off", "t_len")):
        lens = js[ln][idx].astype(np.int64)
        offs = js[off][idx].astype(np.int64)
        new_off = np.zeros(len(idx), dtype=np.int64)
        if len(idx):
            new_off[1:] = np.cumsum(lens)[:-1]
        total = int(lens.sum())
        arena = np.zeros(max(total, 1), dtype=np.uint8)
        for k in range(len(idx)):
            arena[new_off[k]:new_off[k] + lens[k]] = js[side][offs[k]:offs[k] + lens[k]]
        res[side] = arena
        res[off] = new_off.astype(np.uint32)
        res[ln] = lens.astype(np.uint16)
    for name in ("h0", "w", "end_bonus", "true_score", "direction"):
        res[name] = js[name][idx].copy()
    return res


def se_reads(ref, seed, n, length=150, p_sub=0.005, p_indel=0.0005):
    """Single-end reads of `length` bases from both strands of `ref` (bytes
    0..3): (bases n x length, 0-based source starts, reverse flags)."""
    lib = _load()
    if not hasattr(lib, "_se_bound"):
        lib.wl_se_reads.restype = C.c_int
        lib.wl_se_reads.argtypes = [C.c_void_p, C.c_int64, C.c_uint64, C.c_int64, C.c_int,
                                    C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]
        lib._se_bound = True
    ref = np.ascontiguousarray(ref, np.uint8)
    bases = np.zeros((n, length), np.uint8)
    pos = np.zeros(n, np.int64)
    rev = np.zeros(n, np.uint8)
    rc = lib.wl_se_reads(ref.ctypes.data, len(ref), int(seed), int(n), int(length), float(p_sub),
                         float(p_indel), bases.ctypes.data, pos.ctypes.data, rev.ctypes.data)
    if rc != 0:
        raise ValueError("reference too short for the read length")
    return bases, pos, rev


def pe_reads(ref, seed, n, length=150, ins_mean=400.0, ins_sd=50.0, p_sub=0.005, p_indel=0.0005,
             hard=0.0):
    """FR read pairs from `ref` (bytes 0..3): (read 1 bases n x length, read 2
    bases, fragment starts, read-1-reverse flags); a fraction `hard` of pairs
    carry a substitution in read 2 every 15 bases."""
    lib = _load()
    if not hasattr(lib, "_pe_bound"):
        lib.wl_pe_reads.restype = C.c_int
        lib.wl_pe_reads.argtypes = [C.c_void_p, C.c_int64, C.c_uint64, C.c_int64, C.c_int,
                                    C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        lib._pe_bound = True
    ref = np.ascontiguousarray(ref, np.uint8)
    b1 = np.zeros((n, length), np.uint8)
    b2 = np.zeros((n, length), np.uint8)
    pos = np.zeros(n, np.int64)
    rev = np.zeros(n, np.uint8)
    rc = lib.wl_pe_reads(ref.ctypes.data, len(ref), int(seed), int(n), int(length),
                         float(ins_mean), float(ins_sd), float(p_sub), float(p_indel), float(hard),
                         b1.ctypes.data, b2.ctypes.data, pos.ctypes.data, rev.ctypes.data)
    if rc != 0:
        raise ValueError("reference too short for the fragment lengths")
    return b1, b2, pos, rev
