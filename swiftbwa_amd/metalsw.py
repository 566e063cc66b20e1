"""Host mirror of swiftbwa's MetalSW call surface, backed by libhipsw.

Names, argument meaning and failure signalling follow the Swift types:
  ScoringParameters   Sources/BWACore/ScoringParameters.swift:2-73
  SWResult            Sources/BWACore/SWResult.swift:2-25
  BandedSWTask        Sources/MetalSW/BandedSWDispatcher.swift:6-20 (+ endBonus)
  BandedSWDispatcher  Sources/MetalSW/BandedSWDispatcher.swift:23-158
  GlobalSWTask/...    Sources/MetalSW/GlobalSWDispatcher.swift:6-118
  LocalSWTask/...     Sources/MetalSW/LocalSWDispatcher.swift:6-95
  HipSWEngine.shared  Sources/MetalSW/MetalSWEngine.swift:8-17 (None = no GPU)
Failure is signalled the way MetalSW signals it: dispatchBatch16 and the
global/local dispatchers return [] and the caller keeps its CPU path.
"""
import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import native


@dataclass
class ScoringParameters:
    matchScore: int = 1
    mismatchPenalty: int = 4
    gapOpenPenalty: int = 6
    gapExtendPenalty: int = 1
    gapOpenPenaltyDeletion: int = 6
    gapExtendPenaltyDeletion: int = 1
    bandWidth: int = 100
    zDrop: int = 100
    minSeedLength: int = 19
    penClip5: int = 5
    penClip3: int = 5
    maxMatesw: int = 50

    def scoringMatrix(self) -> List[int]:
        """ScoringParameters.swift:65-73: A on the diagonal, -B off it, -1 for N."""
        mat = [-1] * 25
        for i in range(4):
            for j in range(4):
                mat[i * 5 + j] = self.matchScore if i == j else -self.mismatchPenalty
        return mat

    def to_c(self) -> native.HipswScoring:
        s = native.HipswScoring()
        for i, v in enumerate(self.scoringMatrix()):
            s.mat[i] = v
        s.o_ins, s.e_ins = self.gapOpenPenalty, self.gapExtendPenalty
        s.o_del, s.e_del = self.gapOpenPenaltyDeletion, self.gapExtendPenaltyDeletion
        s.zdrop = self.zDrop
        return s


@dataclass(frozen=True)
class SWResult:
    score: int = 0
    queryEnd: int = 0
    targetEnd: int = 0
    globalTargetEnd: int = 0
    globalScore: int = 0
    maxOff: int = 0


@dataclass
class BandedSWTask:
    query: bytes
    target: bytes
    h0: int
    w: int
    scoring: ScoringParameters = field(default_factory=ScoringParameters)
    endBonus: int = 0


class HipSWEngine:
    """Singleton GPU engine; ``HipSWEngine.shared()`` is None without a device."""
    minBatchSize = 32
    _shared = None
    _tried = False

    def __init__(self, device=0):
        self.lib = native.load()
        self.handle = self.lib.hipsw_engine_create(int(device))
        if not self.handle:
            raise RuntimeError("hipsw_engine_create failed (no HIP device %d)" % device)

    @classmethod
    def shared(cls) -> Optional["HipSWEngine"]:
        if not cls._tried:
            cls._tried = True
            try:
                cls._shared = HipSWEngine(0)
            except RuntimeError:
                cls._shared = None
        return cls._shared

    def fault_inject(self, site, countdown=1):
        """Arms hipsw_fault_inject: the countdown-th allocation / launch / copy
        of a later call fails (a test hook for callers' fallbacks)."""
        if self.lib.hipsw_fault_inject(self.handle, int(site), int(countdown)):
            raise ValueError("bad fault site")

    def set_nw_state(self, form):
        """Row-state form of the NW fill's 16-bit tier: 0 registers where the
        band fits, two jobs per lane when the bands of at most 64 columns, but
        not those of at most 32, hold all but 1/32 of those cells (default); 1
        the LDS ring throughout; 2
        registers, one job per lane; 3 two jobs per lane for every band of at
        most 64 columns."""
        if self.lib.hipsw_nw_set_state(self.handle, int(form)):
            raise ValueError("hipsw_nw_set_state: invalid form %r" % (form,))

    def last_error(self) -> str:
        return (self.lib.hipsw_last_error(self.handle) or b"").decode()

    def device_info(self):
        buf = C.create_string_buffer(256)
        ncu = C.c_int32()
        self.lib.hipsw_device_info(self.handle, buf, 256, C.byref(ncu))
        return buf.value.decode(), ncu.value

    def close(self):
        if self.handle:
            self.lib.hipsw_engine_destroy(self.handle)
            self.handle = None

    # ---- array-level entry points (what the Swift shim calls through the C ABI) ----
    def banded_extend(self, jobs, scoring: ScoringParameters, mode=native.MODE_BATCHED,
                      stats=None):
        """Runs hipsw_banded_extend on a job dict; returns int32[n, 6]."""
        n = int(jobs["n"])
        out = np.zeros((n, 6), dtype=np.int32)
        if n == 0:
            return out
        a = _ext_arrays(jobs)
        sc = scoring.to_c()
        st = native.HipswStats()
        rc = self.lib.hipsw_banded_extend(
            self.handle, n, native.ptr(a["queries"]), native.ptr(a["q_off"]),
            native.ptr(a["q_len"]), native.ptr(a["targets"]), native.ptr(a["t_off"]),
            native.ptr(a["t_len"]), native.ptr(a["h0"]), native.ptr(a["w"]),
            native.ptr(a["end_bonus"]), C.byref(sc), int(mode), native.ptr(out), C.byref(st))
        if rc != native.HIPSW_OK:
            raise native.HipswError(rc, self.last_error())
        if stats is not None:
            stats.update(st.as_dict())
        return out


    def global_nw(self, jobs, scoring: ScoringParameters, max_retries=3, cap=512,
                  mode=native.MODE_BATCHED, stats=None):
        """hipsw_global_nw on a job dict (w = starting band, true_score = retry
        threshold).  Returns (score[n], cigar[n, cap], n_cigar[n], flags[n], final_w[n])."""
        n = int(jobs["n"])
        score = np.zeros(n, np.int32)
        cig = np.zeros((n, cap), np.uint32)
        ncig = np.zeros(n, np.int32)
        flags = np.zeros(n, np.int32)
        fw = np.zeros(n, np.int32)
        if n == 0:
            return score, cig, ncig, flags, fw
        a = _ext_arrays(jobs)
        ts = np.ascontiguousarray(jobs["true_score"], dtype=np.int32)
        sc = scoring.to_c()
        st = native.HipswStats()
        rc = self.lib.hipsw_global_nw(
            self.handle, n, native.ptr(a["queries"]), native.ptr(a["q_off"]),
            native.ptr(a["q_len"]), native.ptr(a["targets"]), native.ptr(a["t_off"]),
            native.ptr(a["t_len"]), native.ptr(a["w"]), native.ptr(ts), int(max_retries),
            int(cap), C.byref(sc), int(mode), native.ptr(score), native.ptr(cig),
            native.ptr(ncig), native.ptr(flags), native.ptr(fw), C.byref(st))
        if rc != native.HIPSW_OK:
            raise native.HipswError(rc, self.last_error())
        if stats is not None:
            stats.update(st.as_dict())
        return score, cig, ncig, flags, fw


    def local_sw(self, jobs, scoring: ScoringParameters, min_sub_score=0, stats=None):
        """hipsw_local_sw on a job dict; returns (out[n, 6], ok[n])."""
        n = int(jobs["n"])
        out = np.zeros((n, 6), np.int32)
        ok = np.zeros(n, np.int32)
        if n == 0:
            return out, ok
        a = _ext_arrays(jobs)
        sc = scoring.to_c()
        st = native.HipswStats()
        rc = self.lib.hipsw_local_sw(
            self.handle, n, native.ptr(a["queries"]), native.ptr(a["q_off"]),
            native.ptr(a["q_len"]), native.ptr(a["targets"]), native.ptr(a["t_off"]),
            native.ptr(a["t_len"]), int(min_sub_score), int(scoring.matchScore),
            int(scoring.mismatchPenalty), C.byref(sc), native.ptr(out), native.ptr(ok),
            C.byref(st))
        if rc != native.HIPSW_OK:
            raise native.HipswError(rc, self.last_error())
        if stats is not None:
            stats.update(st.as_dict())
        return out, ok


def _ext_arrays(jobs):
    def c(name, dt):
        return np.ascontiguousarray(jobs[name], dtype=dt)
    a = {"queries": c("queries", np.uint8), "q_off": c("q_off", np.uint32),
         "q_len": c("q_len", np.uint16), "targets": c("targets", np.uint8),
         "t_off": c("t_off", np.uint32), "t_len": c("t_len", np.uint16),
         "w": c("w", np.int32) if "w" in jobs else np.zeros(int(jobs["n"]), np.int32)}
    n = int(jobs["n"])
    for name in ("h0", "end_bonus"):
        a[name] = c(name, np.int32) if name in jobs else np.zeros(n, np.int32)
    if a["queries"].size == 0:
        a["queries"] = np.zeros(1, np.uint8)
    if a["targets"].size == 0:
        a["targets"] = np.zeros(1, np.uint8)
    return a


def tasks_to_jobs(tasks):
    """Packs BandedSWTask objects the way QueryTargetPacker does
    (MetalDispatchHelper.swift:30-46)."""
    n = len(tasks)
    ql = np.array([len(t.query) for t in tasks], dtype=np.uint16)
    tl = np.array([len(t.target) for t in tasks], dtype=np.uint16)
    qo = np.zeros(n, np.uint32)
    to = np.zeros(n, np.uint32)
    if n:
        qo[1:] = np.cumsum(ql.astype(np.int64))[:-1]
        to[1:] = np.cumsum(tl.astype(np.int64))[:-1]
    q = np.frombuffer(b"".join(bytes(t.query) for t in tasks), dtype=np.uint8)
    t = np.frombuffer(b"".join(bytes(t.target) for t in tasks), dtype=np.uint8)
    return {"n": n, "queries": q, "q_off": qo, "q_len": ql, "targets": t, "t_off": to,
            "t_len": tl, "h0": np.array([t.h0 for t in tasks], np.int32),
            "w": np.array([t.w for t in tasks], np.int32),
            "end_bonus": np.array([t.endBonus for t in tasks], np.int32)}


class BandedSWDispatcher:
    """BandedSWDispatcher.swift:23-158.  Both entry points return exact
    BandedSWScalar results for every task; the engine widens the score type
    itself, so dispatchBatch8 never reports an 8-bit overflow (nil)."""

    @staticmethod
    def dispatchBatch8(tasks: List[BandedSWTask], engine: HipSWEngine) -> List[Optional[SWResult]]:
        res = BandedSWDispatcher.dispatchBatch16(tasks, engine)
        if not res and tasks:
            return [None] * len(tasks)
        return res

    @staticmethod
    def dispatchBatch16(tasks: List[BandedSWTask], engine: HipSWEngine) -> List[SWResult]:
        if not tasks:
            return []
        sc = tasks[0].scoring
        if any(t.scoring != sc for t in tasks):
            return []      # one scoring per batch, as the Metal kernels take it
        try:
            out = engine.banded_extend(tasks_to_jobs(tasks), sc)
        except native.HipswError:
            return []
        return [SWResult(*map(int, row)) for row in out]


@dataclass
class GlobalSWTask:
    """GlobalSWDispatcher.swift:6-20."""
    query: bytes
    target: bytes
    trueScore: int
    initialW: int
    scoring: ScoringParameters = field(default_factory=ScoringParameters)


@dataclass
class GlobalSWResult:
    """GlobalSWDispatcher.swift:23-31: score plus packed CIGAR (len << 4 | op)."""
    score: int
    cigar: List[int]


class GlobalSWDispatcher:
    """GlobalSWDispatcher.swift:37-118.  Per task: NW at initialW, then w
    doubling while score < trueScore, at most 3 retries (the CPU loop,
    CIGARGenerator.swift:88-96).  Returns [] on failure; a task whose CIGAR
    exceeds MAX_CIGAR_OPS also fails the batch, as the caller must redo it."""
    MAX_CIGAR_OPS = 512

    @staticmethod
    def dispatchBatch(tasks: List[GlobalSWTask], engine: HipSWEngine) -> List[GlobalSWResult]:
        if not tasks:
            return []
        sc = tasks[0].scoring
        if any(t.scoring != sc for t in tasks):
            return []
        jobs = tasks_to_jobs([BandedSWTask(t.query, t.target, 0, t.initialW, sc) for t in tasks])
        jobs["true_score"] = np.array([t.trueScore for t in tasks], np.int32)
        try:
            score, cig, ncig, flags, _ = engine.global_nw(
                jobs, sc, max_retries=3, cap=GlobalSWDispatcher.MAX_CIGAR_OPS)
        except native.HipswError:
            return []
        if (flags != 0).any():
            return []
        return [GlobalSWResult(int(score[k]), [int(x) for x in cig[k, :ncig[k]]])
                for k in range(len(tasks))]


@dataclass
class LocalSWTask:
    """LocalSWDispatcher.swift:6-16."""
    query: bytes
    target: bytes
    scoring: ScoringParameters = field(default_factory=ScoringParameters)


@dataclass
class LocalSWResult:
    """LocalSWAligner.swift:4-23 (ends inclusive); score2 included."""
    score: int
    queryBegin: int
    queryEnd: int
    targetBegin: int
    targetEnd: int
    score2: int = 0


class LocalSWDispatcher:
    """LocalSWDispatcher.swift:31-95 with LocalSWAligner semantics: one result
    per task, None where the CPU aligner returns nil.  Returns [] on failure."""

    @staticmethod
    def dispatchBatch(tasks: List[LocalSWTask], engine: HipSWEngine,
                      minSubScore: int = 0) -> List[Optional[LocalSWResult]]:
        if not tasks:
            return []
        sc = tasks[0].scoring
        if any(t.scoring != sc for t in tasks):
            return []
        jobs = tasks_to_jobs([BandedSWTask(t.query, t.target, 0, 0, sc) for t in tasks])
        try:
            out, ok = engine.local_sw(jobs, sc, min_sub_score=minSubScore)
        except native.HipswError:
            return []
        return [LocalSWResult(*map(int, out[k])) if ok[k] else None for k in range(len(tasks))]


@dataclass
class CIGARRegion:
    """The arguments of CIGARGenerator.generate (CIGARGenerator.swift:30-42)."""
    querySegment: bytes
    refSegment: bytes
    qb: int
    qe: int
    readLength: int
    isReverse: bool
    trueScore: int
    initialW: int
    refPos: int = 0


@dataclass
class CIGARResult:
    """CIGARResult: packed CIGAR (len << 4 | op, M0 I1 D2 S4), NM, MD, score, pos."""
    cigar: List[int]
    nm: int
    md: str
    score: int
    pos: int


class CIGARGenerator:
    """CIGARGenerator.generate for a batch of regions through
    hipsw_cigar_generate.  A region whose CIGAR or MD does not fit the caps
    comes back as None (the caller generates it on the CPU); a failed call
    returns []."""
    MAX_CIGAR_OPS = 512
    MAX_MD = 1024

    @staticmethod
    def generateBatch(regions: List[CIGARRegion], scoring: ScoringParameters,
                      engine: HipSWEngine, stats=None) -> List[Optional[CIGARResult]]:
        if not regions:
            return []
        n = len(regions)
        jobs = tasks_to_jobs([BandedSWTask(r.querySegment, r.refSegment, 0, 0, scoring)
                              for r in regions])
        a = _ext_arrays(jobs)
        i32 = lambda xs: np.ascontiguousarray(xs, np.int32)
        qb, qe = i32([r.qb for r in regions]), i32([r.qe for r in regions])
        rl, ts = i32([r.readLength for r in regions]), i32([r.trueScore for r in regions])
        iw = i32([r.initialW for r in regions])
        rev = np.ascontiguousarray([1 if r.isReverse else 0 for r in regions], np.uint8)
        pos = np.ascontiguousarray([r.refPos for r in regions], np.int64)
        cap, mcap = CIGARGenerator.MAX_CIGAR_OPS, CIGARGenerator.MAX_MD
        o_score, o_nm, o_ncig, o_flags = (np.zeros(n, np.int32) for _ in range(4))
        o_pos = np.zeros(n, np.int64)
        o_cig = np.zeros((n, cap), np.uint32)
        o_md = np.zeros((n, mcap), np.uint8)
        sc = scoring.to_c()
        st = native.HipswStats()
        rc = engine.lib.hipsw_cigar_generate(
            engine.handle, n, native.ptr(a["queries"]), native.ptr(a["q_off"]),
            native.ptr(a["q_len"]), native.ptr(a["targets"]), native.ptr(a["t_off"]),
            native.ptr(a["t_len"]), native.ptr(qb), native.ptr(qe), native.ptr(rl),
            native.ptr(rev), native.ptr(ts), native.ptr(iw), native.ptr(pos), C.byref(sc),
            scoring.matchScore, scoring.mismatchPenalty, cap, mcap, native.ptr(o_score),
            native.ptr(o_nm), native.ptr(o_pos), native.ptr(o_cig), native.ptr(o_ncig),
            native.ptr(o_md), native.ptr(o_flags), C.byref(st))
        if rc != native.HIPSW_OK:
            return []
        if stats is not None:
            stats.update(st.as_dict())
        out = []
        for k in range(n):
            if o_flags[k]:
                out.append(None)
                continue
            md = bytes(o_md[k]).split(b"\0", 1)[0].decode()
            out.append(CIGARResult([int(x) for x in o_cig[k, :o_ncig[k]]], int(o_nm[k]), md,
                                   int(o_score[k]), int(o_pos[k])))
        return out


class ExtensionAligner:
    """ExtensionAligner.extend (ExtensionAligner.swift:114-346) for a batch of
    chains through hipsw_extend_chains, reference windows read from a resident
    .pac with FMIndex.getReference semantics.  chains: list of (read index,
    [(rbeg, qbeg, len, score), ...]); reads: list of base-code byte strings.
    Returns, per chain, its regions as dicts in creation order, or [] when the
    call fails (the caller keeps its CPU path)."""
    KEYS = ("qb", "qe", "rb", "re", "score", "true_score", "sub", "sub_n", "w", "seed_cov",
            "seed_len0")

    @staticmethod
    def extendChains(chains, reads, pac, genome_len, scoring: ScoringParameters,
                     engine: HipSWEngine, stats=None):
        nc = len(chains)
        if nc == 0:
            return []
        lib = engine.lib
        off = np.zeros(nc + 1, np.uint32)
        off[1:] = np.cumsum([len(s) for _, s in chains])
        ns = int(off[-1])
        flat = [x for _, s in chains for x in s]
        rbeg = np.ascontiguousarray([x[0] for x in flat], np.int64)
        qbeg = np.ascontiguousarray([x[1] for x in flat], np.int32)
        slen = np.ascontiguousarray([x[2] for x in flat], np.int32)
        sscore = np.ascontiguousarray([x[3] for x in flat], np.int32)
        cread = np.ascontiguousarray([r for r, _ in chains], np.uint32)
        rl = np.ascontiguousarray([len(r) for r in reads], np.int32)
        roff = np.zeros(len(reads), np.uint32)
        if len(reads) > 1:
            roff[1:] = np.cumsum(rl.astype(np.int64))[:-1]
        rbytes = np.frombuffer(b"".join(bytes(r) for r in reads) or b"\0", np.uint8).copy()
        pac = np.ascontiguousarray(pac, np.uint8)
        ref = native.HipswPacRef(pac.ctypes.data, int(genome_len))
        opt = native.HipswExtOpts(scoring.matchScore, scoring.mismatchPenalty, scoring.bandWidth,
                                  scoring.penClip5, scoring.penClip3, scoring.minSeedLength)
        out = (native.HipswRegion * max(ns, 1))()
        out_n = np.zeros(nc, np.int32)
        sc = scoring.to_c()
        st = native.HipswStats()
        fn = C.cast(lib.hipsw_get_reference_pac, C.c_void_p)
        rc = lib.hipsw_extend_chains(engine.handle, nc, native.ptr(off), native.ptr(rbeg),
                                     native.ptr(qbeg), native.ptr(slen), native.ptr(sscore),
                                     native.ptr(cread), native.ptr(rbytes), native.ptr(roff),
                                     native.ptr(rl), fn, C.byref(ref), C.byref(sc), C.byref(opt),
                                     C.cast(out, C.c_void_p), native.ptr(out_n), C.byref(st))
        if rc != native.HIPSW_OK:
            return []
        if stats is not None:
            stats.update(st.as_dict())
        res = []
        for c in range(nc):
            base = int(off[c])
            res.append([{k: int(getattr(out[base + r], k)) for k in ExtensionAligner.KEYS}
                        for r in range(int(out_n[c]))])
        return res
