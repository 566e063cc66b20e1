"""Serving loop binding (hipsw_server_*) and the read-chunk workload.

A chunk is the hot-path work of a batch of read pairs -- banded extensions,
mate-rescue local SW and CIGAR NW -- whose targets are windows of the reference
`.pac`, which the server keeps resident in HBM.  See include/hipsw.h.
"""
import ctypes as C
import os
import threading
import time

import numpy as np

from . import native, workload as wl
from .metalsw import HipSWEngine, ScoringParameters

_P = C.c_void_p


HipswChunk = native.HipswChunk
HipswChunkResult = native.HipswChunkResult
_SERVER_SIGS = {k: v for k, v in native._SIGS.items() if k.startswith("hipsw_server_")}


def _wl():
    lib = wl._load()
    if not hasattr(lib, "_serving_bound"):
        lib.wl_reference.restype = C.c_void_p
        lib.wl_reference.argtypes = [C.c_uint64, C.c_int64, C.c_int, C.c_double]
        lib.wl_pack_pac.restype = None
        lib.wl_pack_pac.argtypes = [_P, C.c_int64, _P]
        lib.wl_free_bytes.restype = None
        lib.wl_free_bytes.argtypes = [_P]
        lib.wl_chunk.restype = C.c_void_p
        lib.wl_chunk.argtypes = [_P, C.c_int64, C.c_uint64, C.c_int64]
        lib.wl_chunk_field.restype = C.c_void_p
        lib.wl_chunk_field.argtypes = [_P, C.c_int, C.c_int]
        lib.wl_chunk_count.restype = C.c_int64
        lib.wl_chunk_count.argtypes = [_P, C.c_int, C.c_int]
        lib.wl_chunk_free.restype = None
        lib.wl_chunk_free.argtypes = [_P]
        lib._serving_bound = True
    return lib


def make_reference(seed, n_bases, copies=1000, divergence=0.10):
    """Random ACGT reference with planted 300-bp and 6-kb families (bytes 0..3)."""
    lib = _wl()
    h = lib.wl_reference(int(seed), int(n_bases), int(copies), float(divergence))
    try:
        return np.frombuffer((C.c_uint8 * n_bases).from_address(h), dtype=np.uint8).copy()
    finally:
        lib.wl_free_bytes(h)


def contig_lengths(n_bases, n_contigs):
    """Lengths of n_contigs reference sequences tiling n_bases: decreasing
    sizes (as a genome's chromosomes) and a last short one of 16,569 bases when
    it fits (a mitochondrion-sized sequence, so windows hit its ends)."""
    if n_contigs <= 1:
        return [int(n_bases)]
    short = 16569 if n_bases > 16569 * n_contigs * 4 else max(1, n_bases // (4 * n_contigs))
    rest = n_bases - short
    wts = np.array([1.0 / np.sqrt(k + 1) for k in range(n_contigs - 1)])
    lens = np.floor(rest * wts / wts.sum()).astype(np.int64)
    lens[0] += rest - lens.sum()
    return [int(x) for x in lens] + [int(short)]


def pack_pac(ref):
    """.pac bytes: 4 bases per byte, first base in bits 7-6."""
    ref = np.ascontiguousarray(ref, np.uint8)
    pac = np.zeros((len(ref) + 3) // 4, np.uint8)
    _wl().wl_pack_pac(ref.ctypes.data, len(ref), pac.ctypes.data)
    return pac


_JOB_FIELDS = [("q", np.uint8), ("q_off", np.uint32), ("q_len", np.uint16), ("t_pos", np.int64),
               ("t_len", np.uint16), ("t_rev", np.uint8), ("a", np.int32), ("b", np.int32),
               ("c", np.int32), ("r", np.int32)]


def make_chunk(ref, seed, pairs):
    """Jobs of `pairs` read pairs: dict with 'ext', 'loc', 'nw' job dicts."""
    lib = _wl()
    h = lib.wl_chunk(ref.ctypes.data, len(ref), int(seed), int(pairs))
    if not h:
        raise ValueError("bad chunk request")
    try:
        out = {"pairs": int(pairs)}
        for kind, name in ((0, "ext"), (1, "loc"), (2, "nw")):
            d = {}
            for k, (f, dt) in enumerate(_JOB_FIELDS):
                cnt = int(lib.wl_chunk_count(h, kind, k))
                addr = lib.wl_chunk_field(h, kind, k)
                nb = cnt * np.dtype(dt).itemsize
                d[f] = np.frombuffer((C.c_char * max(nb, 1)).from_address(addr), dtype=dt,
                                     count=cnt).copy()
            d["n"] = int(lib.wl_chunk_count(h, kind, 1))
            out[name] = d
        return out
    finally:
        lib.wl_chunk_free(h)


class Server:
    """Continuous-batching server over a resident reference (hipsw_server_*)."""

    def __init__(self, engine: HipSWEngine, slots=4, hbm_budget=32 << 30,
                 scoring: ScoringParameters = None, nw_cap=512, loc_min_sub=19):
        self.eng = engine
        self.lib = native.load()
        for name, (res, args) in _SERVER_SIGS.items():
            fn = getattr(self.lib, name)
            fn.restype = res
            fn.argtypes = args
        self.scoring = scoring or ScoringParameters()
        self.nw_cap = nw_cap
        self.loc_min_sub = loc_min_sub
        sc = self.scoring.to_c()
        self.h = self.lib.hipsw_server_create(engine.handle, int(slots), int(hbm_budget), C.byref(sc),
                                              self.scoring.matchScore, self.scoring.mismatchPenalty,
                                              int(nw_cap))
        if not self.h:
            raise RuntimeError("hipsw_server_create failed")
        self._keep = {}

    def set_reference(self, pac, n_bases):
        pac = np.ascontiguousarray(pac, np.uint8)
        rc = self.lib.hipsw_server_set_reference(self.h, pac.ctypes.data, int(n_bases))
        if rc:
            raise native.HipswError(rc, "set_reference")

    def set_contigs(self, lengths):
        """Splits the reference into sequences of these lengths (in order), the
        index's annotations; mate-rescue windows are clamped to them."""
        lens = np.ascontiguousarray(lengths, np.int64)
        offs = np.ascontiguousarray(np.concatenate([[0], np.cumsum(lens)[:-1]]), np.int64)
        rc = self.lib.hipsw_server_set_contigs(self.h, len(lens), offs.ctypes.data, lens.ctypes.data)
        if rc:
            raise native.HipswError(rc, "set_contigs")

    def submit(self, chunk):
        """Queues a chunk: dict of 'ext', 'loc', 'nw' job dicts (missing kinds
        are empty).  A 'loc' dict with a 'rid' array holds raw mate-rescue
        windows, clamped on the server to the sequence loc['rid'][k]."""
        empty = {"n": 0, "q": np.zeros(1, np.uint8), "q_off": np.zeros(1, np.uint32),
                 "q_len": np.zeros(1, np.uint16), "t_pos": np.zeros(1, np.int64),
                 "t_len": np.zeros(1, np.uint16), "t_rev": np.zeros(1, np.uint8),
                 "a": np.zeros(1, np.int32), "b": np.zeros(1, np.int32), "c": np.zeros(1, np.int32)}
        e, l, n = (chunk.get(k) or empty for k in ("ext", "loc", "nw"))
        if e["n"] and np.any(np.asarray(e["t_rev"]) > 1):
            raise ValueError("t_rev must be 0 or 1")
        arrs = [np.ascontiguousarray(x) for x in (
            e["q"], e["q_off"], e["q_len"], e["t_pos"], e["t_len"], e["t_rev"], e["a"], e["b"], e["c"],
            l["q"], l["q_off"], l["q_len"], l["t_pos"], l["t_len"],
            n["q"], n["q_off"], n["q_len"], n["t_pos"], n["t_len"], n["a"], n["b"])]
        rid = np.ascontiguousarray(l["rid"], np.int32) if "rid" in l else None
        p = [a.ctypes.data for a in arrs]
        c = HipswChunk(e["n"], *p[0:9], l["n"], *p[9:14], self.loc_min_sub, n["n"], *p[14:21],
                       rid.ctypes.data if rid is not None else None)
        t = self.lib.hipsw_server_submit(self.h, C.byref(c))
        if t < 0:
            raise native.HipswError(int(t), "submit")
        self._keep[int(t)] = (e["n"], l["n"], n["n"])
        return int(t)

    def poll(self, ticket):
        return self.lib.hipsw_server_poll(self.h, int(ticket)) == native.HIPSW_OK

    def wait(self, ticket, compact=False):
        """Results of a ticket.  CIGARs as rows of nw_cap ops ("nw_cigar"), or
        with compact=True as the used ops back to back ("nw_cigar_ops", job k's
        at nw_cigar_off[k] .. nw_cigar_off[k + 1])."""
        ne, nl, nn = self._keep.pop(int(ticket))
        ext = np.zeros((ne, 6), np.int32)
        loc = np.zeros((nl, 6), np.int32)
        lok = np.zeros(nl, np.int32)
        nws = np.zeros(nn, np.int32)
        # compact: the buffer keeps its n_nw x nw_cap bound, but only the used
        # prefix is written (the rest is never touched)
        nwc = np.empty(nn * self.nw_cap, np.uint32) if compact else np.zeros((nn, self.nw_cap), np.uint32)
        off = np.zeros(nn + 1, np.uint32) if compact else None
        nwn = np.zeros(nn, np.int32)
        nwf = np.zeros(nn, np.int32)
        win = np.full(nl, -1, np.int64)
        r = HipswChunkResult(ext.ctypes.data, loc.ctypes.data, lok.ctypes.data, nws.ctypes.data,
                             nwc.ctypes.data, nwn.ctypes.data, nwf.ctypes.data, 0, 0, 0, 0, 0,
                             off.ctypes.data if compact else None, 0, 0, win.ctypes.data)
        rc = self.lib.hipsw_server_wait(self.h, int(ticket), C.byref(r))
        if rc:
            raise native.HipswError(rc, "chunk failed")
        out = {"ext": ext, "loc": loc, "loc_ok": lok, "nw_score": nws,
               "nw_n_cigar": nwn, "nw_flags": nwf, "latency_ms": r.latency_ms,
               "queue_ms": r.queue_ms, "run_ms": r.run_ms, "cells": r.cells,
               "host_ms": r.host_ms, "est_bytes": r.est_bytes,
               "dev_peak_bytes": r.dev_peak_bytes, "loc_win": win}
        if compact:
            out["nw_cigar_ops"] = nwc[:int(off[-1])]
            out["nw_cigar_off"] = off
        else:
            out["nw_cigar"] = nwc
        return out

    def stats(self):
        a, b = C.c_int64(), C.c_int64()
        c, d = C.c_int32(), C.c_int32()
        self.lib.hipsw_server_stats(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(d))
        return {"in_use": a.value, "peak_in_use": b.value, "running": c.value, "queued": d.value}

    def close(self):
        if self.h:
            self.lib.hipsw_server_destroy(self.h)
            self.h = None


def run_poisson(server, chunks, rate_per_s, seed=0, on_result=None):
    """Submits `chunks` at Poisson arrival times (rate chunks/s) from one thread
    and collects them from another; returns per-chunk latency records.
    on_result(k, results) sees each chunk's results as they are collected."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_per_s, len(chunks)) if rate_per_s > 0 else np.zeros(len(chunks))
    tickets = [None] * len(chunks)
    ready = threading.Semaphore(0)
    t0 = time.perf_counter()

    def producer():
        t = t0
        for k, ch in enumerate(chunks):
            t += gaps[k]
            d = t - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets[k] = server.submit(ch)
            ready.release()

    th = threading.Thread(target=producer)
    th.start()
    out = []
    for k in range(len(chunks)):
        ready.acquire()
        r = server.wait(tickets[k], compact=True)
        if on_result is not None:
            on_result(k, r)
        out.append({"pairs": chunks[k]["pairs"], "latency_ms": r["latency_ms"],
                    "queue_ms": r["queue_ms"], "run_ms": r["run_ms"], "cells": r["cells"],
                    "host_ms": r["host_ms"]})
    th.join()
    wall = time.perf_counter() - t0
    return out, wall


class ChainPasses:
    """The chain extension driver as four passes (hipsw_chain_*): each pass's
    jobs become the 'ext' kind of a server chunk, with windows the server
    fetches from the resident reference; finish() gives the regions
    hipsw_extend_chains would (ExtensionAligner.extend, ExtensionAligner.swift:
    114-346).  Chains: dict with 'off' (n_chains + 1 seed offsets), seed arrays
    'rbeg', 'qbeg', 'len', 'score', and 'read' (read of each chain); reads as
    one byte arena with offsets and lengths."""

    def __init__(self, chains, reads, read_off, read_len, scoring: ScoringParameters = None):
        self.lib = native.load()
        sp = scoring or ScoringParameters()
        self._sc = sp.to_c()
        self._opt = native.HipswExtOpts(sp.matchScore, sp.mismatchPenalty, sp.bandWidth,
                                        sp.penClip5, sp.penClip3, sp.minSeedLength)
        self.off = np.ascontiguousarray(chains["off"], np.uint32)
        self.n_chains = len(self.off) - 1
        arr = [np.ascontiguousarray(chains[k], dt) for k, dt in
               (("rbeg", np.int64), ("qbeg", np.int32), ("len", np.int32), ("score", np.int32),
                ("read", np.uint32))]
        self._reads = [np.ascontiguousarray(reads, np.uint8), np.ascontiguousarray(read_off, np.uint32),
                       np.ascontiguousarray(read_len, np.int32)]
        self.h = self.lib.hipsw_chain_begin(self.n_chains, self.off.ctypes.data,
                                            *[a.ctypes.data for a in arr], len(self._reads[2]),
                                            *[a.ctypes.data for a in self._reads],
                                            C.byref(self._sc), C.byref(self._opt))
        if not self.h:
            raise native.HipswError(native.HIPSW_E_INVALID, "hipsw_chain_begin")
        self._keep = arr

    def next_chunk(self):
        """The next pass's jobs as a chunk (its 'ext' kind)."""
        p = native.HipswChainPass()
        rc = self.lib.hipsw_chain_next(self.h, C.byref(p))
        if rc:
            raise native.HipswError(rc, "hipsw_chain_next")
        n = p.n

        def view(addr, dt, cnt):
            if cnt == 0 or not addr:
                return np.zeros(0, dt)
            nb = cnt * np.dtype(dt).itemsize
            return np.frombuffer((C.c_char * nb).from_address(addr), dtype=dt, count=cnt).copy()

        q_off = view(p.q_off, np.uint32, n)
        q_len = view(p.q_len, np.uint16, n)
        qbytes = int(q_off[-1]) + int(q_len[-1]) if n else 0
        ext = {"n": n, "q": view(p.q, np.uint8, max(qbytes, 0)) if n else np.zeros(1, np.uint8),
               "q_off": q_off, "q_len": q_len, "t_pos": view(p.t_pos, np.int64, n),
               "t_len": view(p.t_len, np.int32, n).astype(np.uint16),
               "t_rev": view(p.t_rev, np.uint8, n), "a": view(p.h0, np.int32, n),
               "b": view(p.w, np.int32, n), "c": view(p.eb, np.int32, n)}
        if n and len(ext["q"]) == 0:
            ext["q"] = np.zeros(1, np.uint8)
        return {"ext": ext, "pass": p.index}

    def feed(self, out6):
        out6 = np.ascontiguousarray(out6, np.int32)
        rc = self.lib.hipsw_chain_feed(self.h, out6.ctypes.data if out6.size else None)
        if rc:
            raise native.HipswError(rc, "hipsw_chain_feed")

    def finish(self):
        """(regions as a structured array over the seed slots, count per chain)."""
        n_seeds = int(self.off[-1]) if self.n_chains else 0
        out = (native.HipswRegion * max(n_seeds, 1))()
        out_n = np.zeros(max(self.n_chains, 1), np.int32)
        rc = self.lib.hipsw_chain_finish(self.h, out, out_n.ctypes.data)
        if rc:
            raise native.HipswError(rc, "hipsw_chain_finish")
        return out, out_n[:self.n_chains]

    def close(self):
        if self.h:
            self.lib.hipsw_chain_free(self.h)
            self.h = None

    def __del__(self):
        self.close()


def cigar_route(jobs, scoring: ScoringParameters = None):
    """Band and route of each CIGAR region (hipsw_cigar_route): (w, route) with
    route 0 the fast path, 1 NW, 2 an empty segment."""
    lib = native.load()
    sp = scoring or ScoringParameters()
    sc = sp.to_c()
    n = len(jobs["q_len"])
    w = np.zeros(n, np.int32)
    route = np.zeros(n, np.uint8)
    a = [np.ascontiguousarray(jobs[k], dt) for k, dt in
         (("q_len", np.uint16), ("r_len", np.uint16), ("qb", np.int32), ("qe", np.int32),
          ("true_score", np.int32), ("w", np.int32))]
    rc = lib.hipsw_cigar_route(n, *[x.ctypes.data for x in a], C.byref(sc), sp.matchScore,
                               sp.mismatchPenalty, w.ctypes.data, route.ctypes.data)
    if rc:
        raise native.HipswError(rc, "hipsw_cigar_route")
    return w, route


def cigar_finish(jobs, route, nw, nw_cap, cap=4096, md_cap=8192):
    """postProcess of every region given the NW results of the route-1 regions
    in region order (nw: 'score', 'ops', 'off', 'n_cigar', 'flags').  Returns
    per-region score, nm, pos, CIGAR op lists, MD strings, flags."""
    lib = native.load()
    n = len(jobs["q_len"])
    a = [np.ascontiguousarray(jobs[k], dt) for k, dt in
         (("qseg", np.uint8), ("q_off", np.uint32), ("q_len", np.uint16), ("rseg", np.uint8),
          ("r_off", np.uint32), ("r_len", np.uint16), ("qb", np.int32), ("qe", np.int32),
          ("read_len", np.int32), ("rev", np.uint8), ("true_score", np.int32),
          ("ref_pos", np.int64))]
    route = np.ascontiguousarray(route, np.uint8)
    b = [np.ascontiguousarray(nw[k], dt) for k, dt in
         (("score", np.int32), ("ops", np.uint32), ("off", np.uint32), ("n_cigar", np.int32),
          ("flags", np.int32))]
    b = [x if x.size else np.zeros(1, x.dtype) for x in b]
    a = [x if x.size else np.zeros(1, x.dtype) for x in a]
    score = np.zeros(n, np.int32)
    nm = np.zeros(n, np.int32)
    pos = np.zeros(n, np.int64)
    cig = np.zeros((n, cap), np.uint32)
    ncig = np.zeros(n, np.int32)
    md = np.zeros((n, md_cap), np.uint8)
    flags = np.zeros(n, np.int32)
    rc = lib.hipsw_cigar_finish(n, *[x.ctypes.data for x in a], route.ctypes.data,
                                *[x.ctypes.data for x in b], int(nw_cap), cap, md_cap,
                                score.ctypes.data, nm.ctypes.data, pos.ctypes.data,
                                cig.ctypes.data, ncig.ctypes.data, md.ctypes.data, flags.ctypes.data)
    if rc:
        raise native.HipswError(rc, "hipsw_cigar_finish")
    out = []
    for k in range(n):
        m = bytes(md[k])
        out.append({"score": int(score[k]), "nm": int(nm[k]), "pos": int(pos[k]),
                    "cigar": cig[k, :min(int(ncig[k]), cap)].copy(),
                    "md": m[:m.index(0)].decode() if 0 in m else m.decode(),
                    "flags": int(flags[k])})
    return out
