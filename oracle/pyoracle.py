"""ctypes binding of the CPU oracle (oracle/_build/libswor.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the swiftbwa_amd package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libswor.so")


class Scoring(C.Structure):
    _fields_ = [("a", C.c_int32), ("b", C.c_int32), ("o_ins", C.c_int32), ("e_ins", C.c_int32),
                ("o_del", C.c_int32), ("e_del", C.c_int32), ("w", C.c_int32),
                ("zdrop", C.c_int32), ("pen_clip5", C.c_int32), ("pen_clip3", C.c_int32),
                ("min_seed_len", C.c_int32), ("mat", C.c_int8 * 25)]


class SWRes(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("score", "qle", "tle", "gtle", "gscore", "max_off")]

    def tuple(self):
        return (self.score, self.qle, self.tle, self.gtle, self.gscore, self.max_off)


class CigarRes(C.Structure):
    _fields_ = [("score", C.c_int32), ("nm", C.c_int32), ("pos", C.c_int64),
                ("n_cigar", C.c_int32), ("nw_calls", C.c_int32), ("cells", C.c_int64)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None
_lib_path = LIB_PATH


def use_library(path):
    """Loads the oracle from another build of the same sources (bench.py's
    CPU baseline compiles one with -march=native on the measuring host).
    Must be called before the first oracle call."""
    global _lib_path
    if _lib is not None and path != _lib_path:
        raise RuntimeError("oracle library already loaded from %s" % _lib_path)
    _lib_path = path


def build_native(out_dir):
    """Compiles the oracle sources with -O3 -march=native into out_dir and
    returns the library path (None when no C compiler is available)."""
    out = os.path.join(out_dir, "libswor_native.so")
    cmd = ["gcc", "-O3", "-march=native", "-std=c11", "-fPIC", "-shared", "-o", out,
           os.path.join(HERE, "sw_oracle.c"), os.path.join(HERE, "e2e_oracle.c"), "-lpthread", "-lm"]
    try:
        subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    except (OSError, subprocess.CalledProcessError):
        return None
    return out


def lib():
    global _lib
    if _lib is None:
        if _lib_path == LIB_PATH and not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(_lib_path)
        P = C.c_void_p
        L.or_scoring_default.argtypes = [C.POINTER(Scoring)]
        L.or_scoring_fill_matrix.argtypes = [C.POINTER(Scoring)]
        L.or_ksw_extend.restype = C.c_int64
        L.or_ksw_extend.argtypes = [C.c_int, P, C.c_int, P, C.POINTER(Scoring), C.c_int32,
                                    C.c_int32, C.c_int32, C.POINTER(SWRes)]
        L.or_ksw_global.restype = C.c_int
        L.or_ksw_global.argtypes = [C.c_int, P, C.c_int, P, C.POINTER(Scoring), C.c_int32,
                                    C.POINTER(C.c_int32), P, C.c_int, C.POINTER(C.c_int64)]
        L.or_cigar_generate.restype = C.c_int
        L.or_cigar_generate.argtypes = [P, C.c_int, P, C.c_int, C.c_int32, C.c_int32, C.c_int32,
                                        C.c_int, C.c_int32, C.c_int32, C.POINTER(Scoring),
                                        C.c_int64, P, C.c_int, C.c_char_p, C.c_int,
                                        C.POINTER(CigarRes)]
        L.or_global_with_retry.restype = C.c_int
        L.or_global_with_retry.argtypes = [P, C.c_int, P, C.c_int, C.c_int32, C.c_int32, C.c_int,
                                           C.POINTER(Scoring), C.POINTER(C.c_int32), P, C.c_int,
                                           C.POINTER(C.c_int32), C.POINTER(C.c_int64)]
        L.or_local_sw.restype = C.c_int
        L.or_local_sw.argtypes = [P, C.c_int, P, C.c_int, C.POINTER(Scoring), C.c_int, P,
                                  C.POINTER(C.c_int64)]
        L.or_extend_chain.restype = C.c_int
        L.or_extend_chain.argtypes = [C.c_int, P, P, P, P, P, C.c_int, P, C.c_int64,
                                      C.POINTER(Scoring), P, C.c_int, C.POINTER(C.c_int64),
                                      C.POINTER(C.c_int64)]
        L.or_ksw_extend_batch.restype = C.c_int64
        L.or_ksw_extend_batch.argtypes = [C.c_int, P, P, P, P, P, P, P, P, C.POINTER(Scoring),
                                          P, C.c_int]
        L.or_global_batch.restype = C.c_int64
        L.or_global_batch.argtypes = [C.c_int, P, P, P, P, P, P, P, C.c_int, C.POINTER(Scoring),
                                      P, P, C.c_int, P, C.c_int]
        L.or_local_batch.restype = C.c_int64
        L.or_local_batch.argtypes = [C.c_int, P, P, P, P, P, C.c_int, C.POINTER(Scoring), P, P,
                                     C.c_int]
        _lib = L
    return _lib


def scoring(**kw):
    """Default ScoringParameters (ScoringParameters.swift:2-45) with overrides."""
    s = Scoring()
    lib().or_scoring_default(C.byref(s))
    mat_override = kw.pop("mat", None)
    for k, v in kw.items():
        setattr(s, k, v)
    lib().or_scoring_fill_matrix(C.byref(s))
    if mat_override is not None:
        for i, v in enumerate(mat_override):
            s.mat[i] = int(v)
    return s


def _u8(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint8))


def ksw_extend(query, target, sc=None, w=100, h0=1, end_bonus=0):
    """BandedSWScalar.align -> ((score,qle,tle,gtle,gscore,max_off), cells)."""
    sc = sc or scoring()
    q, t = _u8(query), _u8(target)
    r = SWRes()
    cells = lib().or_ksw_extend(len(q), q.ctypes.data, len(t), t.ctypes.data, C.byref(sc),
                                int(w), int(h0), int(end_bonus), C.byref(r))
    return r.tuple(), cells


def ksw_global(query, target, sc=None, w=100):
    """GlobalAligner.align -> (score, [packed cigar], cells)."""
    sc = sc or scoring()
    q, t = _u8(query), _u8(target)
    cap = len(q) + len(t) + 2
    cig = np.zeros(cap, dtype=np.uint32)
    score = C.c_int32()
    cells = C.c_int64()
    n = lib().or_ksw_global(len(q), q.ctypes.data, len(t), t.ctypes.data, C.byref(sc), int(w),
                            C.byref(score), cig.ctypes.data, cap, C.byref(cells))
    return score.value, [int(x) for x in cig[:n]], cells.value


def global_with_retry(query, target, w0, true_score, max_retries=3, sc=None):
    sc = sc or scoring()
    q, t = _u8(query), _u8(target)
    cap = len(q) + len(t) + 2
    cig = np.zeros(cap, dtype=np.uint32)
    score = C.c_int32()
    fw = C.c_int32()
    cells = C.c_int64()
    n = lib().or_global_with_retry(q.ctypes.data, len(q), t.ctypes.data, len(t), int(w0),
                                   int(true_score), int(max_retries), C.byref(sc),
                                   C.byref(score), cig.ctypes.data, cap, C.byref(fw),
                                   C.byref(cells))
    return score.value, [int(x) for x in cig[:n]], fw.value, cells.value


def cigar_generate(qseg, rseg, qb, qe, read_len, is_reverse, true_score, initial_w,
                   sc=None, ref_pos=0):
    """CIGARGenerator.generate -> dict(cigar, nm, md, score, pos, nw_calls)."""
    sc = sc or scoring()
    q, t = _u8(qseg), _u8(rseg)
    cap = len(q) + len(t) + 8
    cig = np.zeros(cap, dtype=np.uint32)
    md = C.create_string_buffer(4 * (len(q) + len(t)) + 64)
    res = CigarRes()
    n = lib().or_cigar_generate(q.ctypes.data, len(q), t.ctypes.data, len(t), int(qb), int(qe),
                                int(read_len), int(bool(is_reverse)), int(true_score),
                                int(initial_w), C.byref(sc), int(ref_pos), cig.ctypes.data, cap,
                                md, len(md), C.byref(res))
    return {"cigar": [int(x) for x in cig[:n]], "nm": res.nm, "md": md.value.decode(),
            "score": res.score, "pos": res.pos, "nw_calls": res.nw_calls}


def local_sw(query, target, sc=None, min_sub_score=0):
    """LocalSWAligner.align -> (score, qb, qe, tb, te, score2) or None."""
    sc = sc or scoring()
    q, t = _u8(query), _u8(target)
    out = np.zeros(6, dtype=np.int32)
    cells = C.c_int64()
    ok = lib().or_local_sw(q.ctypes.data, len(q), t.ctypes.data, len(t), C.byref(sc),
                           int(min_sub_score), out.ctypes.data, C.byref(cells))
    return tuple(int(x) for x in out) if ok else None


def extend_chain(seeds, read, ref, sc=None):
    """ExtensionAligner.extend for one chain.  seeds = [(rbeg, qbeg, len, score)]."""
    sc = sc or scoring()
    sq = np.array([s[1] for s in seeds], dtype=np.int32)
    sl = np.array([s[2] for s in seeds], dtype=np.int32)
    sr = np.array([s[0] for s in seeds], dtype=np.int64)
    ss = np.array([s[3] for s in seeds], dtype=np.int32)
    rd, rf = _u8(read), _u8(ref)
    out = np.zeros((max(len(seeds), 1), 10), dtype=np.int64)
    calls = C.c_int64()
    cells = C.c_int64()
    n = lib().or_extend_chain(len(seeds), sq.ctypes.data, sl.ctypes.data, sr.ctypes.data,
                              ss.ctypes.data, rd.ctypes.data, len(rd), rf.ctypes.data, len(rf),
                              C.byref(sc), out.ctypes.data, len(out), C.byref(calls),
                              C.byref(cells))
    keys = ("qb", "qe", "rb", "re", "score", "true_score", "sub", "sub_n", "w", "seed_cov")
    return [dict(zip(keys, (int(v) for v in out[k]))) for k in range(n)]


def _arenas(js):
    """One byte arena holding queries then targets, with 64-bit offsets."""
    q, t = js["queries"], js["targets"]
    seqs = np.concatenate([q, t]).astype(np.uint8)
    q_off = js["q_off"].astype(np.uint64)
    t_off = js["t_off"].astype(np.uint64) + np.uint64(len(q))
    return seqs, q_off, js["q_len"].astype(np.int32), t_off, js["t_len"].astype(np.int32)


def ksw_extend_batch(js, sc=None, nthreads=1):
    """Oracle over a job dict (swiftbwa_amd.workload layout) -> (out[n,6], cells)."""
    sc = sc or scoring()
    seqs, qo, ql, to, tl = _arenas(js)
    n = int(js["n"])
    out = np.zeros((n, 6), dtype=np.int32)
    h0 = np.ascontiguousarray(js["h0"], dtype=np.int32)
    w = np.ascontiguousarray(js["w"], dtype=np.int32)
    eb = np.ascontiguousarray(js["end_bonus"], dtype=np.int32)
    cells = lib().or_ksw_extend_batch(n, seqs.ctypes.data, qo.ctypes.data, ql.ctypes.data,
                                      to.ctypes.data, tl.ctypes.data, h0.ctypes.data,
                                      w.ctypes.data, eb.ctypes.data, C.byref(sc),
                                      out.ctypes.data, int(nthreads))
    return out, cells


def global_batch(js, sc=None, max_retries=3, cap=512, nthreads=1):
    sc = sc or scoring()
    seqs, qo, ql, to, tl = _arenas(js)
    n = int(js["n"])
    w = np.ascontiguousarray(js["w"], dtype=np.int32)
    ts = np.ascontiguousarray(js["true_score"], dtype=np.int32)
    score = np.zeros(n, dtype=np.int32)
    cig = np.zeros((n, cap), dtype=np.uint32)
    ncig = np.zeros(n, dtype=np.int32)
    cells = lib().or_global_batch(n, seqs.ctypes.data, qo.ctypes.data, ql.ctypes.data,
                                  to.ctypes.data, tl.ctypes.data, w.ctypes.data, ts.ctypes.data,
                                  int(max_retries), C.byref(sc), score.ctypes.data,
                                  cig.ctypes.data, cap, ncig.ctypes.data, int(nthreads))
    return score, cig, ncig, cells


def local_batch(js, sc=None, min_sub_score=0, nthreads=1):
    sc = sc or scoring()
    seqs, qo, ql, to, tl = _arenas(js)
    n = int(js["n"])
    out = np.zeros((n, 6), dtype=np.int32)
    ok = np.zeros(n, dtype=np.int32)
    cells = lib().or_local_batch(n, seqs.ctypes.data, qo.ctypes.data, ql.ctypes.data,
                                 to.ctypes.data, tl.ctypes.data, int(min_sub_score), C.byref(sc),
                                 out.ctypes.data, ok.ctypes.data, int(nthreads))
    return out, ok, cells


def get_reference(ref, gl, pos, length):
    """FMIndex.getReference (Sources/FMIndex/FMIndex.swift:25-65) over a forward
    reference given as bytes 0..3: positions >= gl read the reverse complement."""
    if pos >= gl:
        fend = 2 * gl - pos
        fstart = fend - length
        s, e = max(0, fstart), min(fend, gl)
        if e <= s:
            return np.zeros(0, np.uint8)
        return (3 - ref[s:e][::-1]).astype(np.uint8)
    n = min(length, gl - pos)
    if n <= 0 or pos < 0:
        return np.zeros(0, np.uint8)
    return ref[pos:pos + n].astype(np.uint8)


def chunk_to_job_sets(ref, chunk):
    """Materialises a serving chunk's windows (test side) into the job dicts the
    batch oracles take: extension (with left-window reversal), local SW, NW."""
    gl = len(ref)
    out = {}
    for kind in ("ext", "loc", "nw"):
        j = chunk[kind]
        n = j["n"]
        ts, tl = [], []
        for k in range(n):
            t = get_reference(ref, gl, int(j["t_pos"][k]), int(j["t_len"][k]))
            if kind == "ext" and j["t_rev"][k]:
                t = t[::-1]
            ts.append(t)
            tl.append(len(t))
        t_off = np.zeros(n, np.uint32)
        if n:
            t_off[1:] = np.cumsum(tl)[:-1]
        d = {"n": n, "queries": j["q"], "q_off": j["q_off"], "q_len": j["q_len"],
             "targets": np.concatenate(ts) if sum(tl) else np.zeros(1, np.uint8),
             "t_off": t_off, "t_len": np.array(tl, np.uint16)}
        if kind == "ext":
            d.update(h0=j["a"], w=j["b"], end_bonus=j["c"])
        if kind == "nw":
            d.update(w=j["a"], true_score=j["b"])
        out[kind] = d
    return out
