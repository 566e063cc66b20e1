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


def pack_pac(ref):
    """.pac bytes: 4 bases per byte, first base in bits 7-6."""
    ref = np.ascontiguousarray(ref, np.uint8)
    pac = np.zeros((len(ref) + 3) // 4, np.uint8)
    _wl().wl_pack_pac(ref.ctypes.data, len(ref), pac.ctypes.data)
    return pac


_JOB_FIELDS = [("q", np.uint8), ("q_off", np.uint32), ("q_len", np.uint16), ("t_pos", np.int64),
               ("t_len", np.uint16), ("t_rev", np.uint8), ("a", np.int32), ("b", np.int32),
               ("c", np.int32)]


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

    def submit(self, chunk):
        e, l, n = chunk["ext"], chunk["loc"], chunk["nw"]
        arrs = [np.ascontiguousarray(x) for x in (
            e["q"], e["q_off"], e["q_len"], e["t_pos"], e["t_len"], e["t_rev"], e["a"], e["b"], e["c"],
            l["q"], l["q_off"], l["q_len"], l["t_pos"], l["t_len"],
            n["q"], n["q_off"], n["q_len"], n["t_pos"], n["t_len"], n["a"], n["b"])]
        p = [a.ctypes.data for a in arrs]
        c = HipswChunk(e["n"], *p[0:9], l["n"], *p[9:14], self.loc_min_sub, n["n"], *p[14:21])
        t = self.lib.hipsw_server_submit(self.h, C.byref(c))
        if t < 0:
            raise native.HipswError(int(t), "submit")
        self._keep[int(t)] = (chunk["ext"]["n"], chunk["loc"]["n"], chunk["nw"]["n"])
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
        r = HipswChunkResult(ext.ctypes.data, loc.ctypes.data, lok.ctypes.data, nws.ctypes.data,
                             nwc.ctypes.data, nwn.ctypes.data, nwf.ctypes.data, 0, 0, 0, 0, 0,
                             off.ctypes.data if compact else None)
        rc = self.lib.hipsw_server_wait(self.h, int(ticket), C.byref(r))
        if rc:
            raise native.HipswError(rc, "chunk failed")
        out = {"ext": ext, "loc": loc, "loc_ok": lok, "nw_score": nws,
               "nw_n_cigar": nwn, "nw_flags": nwf, "latency_ms": r.latency_ms,
               "queue_ms": r.queue_ms, "run_ms": r.run_ms, "cells": r.cells,
               "host_ms": r.host_ms, "est_bytes": r.est_bytes,
               "dev_peak_bytes": r.dev_peak_bytes}
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


def run_poisson(server, chunks, rate_per_s, seed=0):
    """Submits `chunks` at Poisson arrival times (rate chunks/s) from one thread
    and collects them from another; returns per-chunk latency records."""
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
        out.append({"pairs": chunks[k]["pairs"], "latency_ms": r["latency_ms"],
                    "queue_ms": r["queue_ms"], "run_ms": r["run_ms"], "cells": r["cells"],
                    "host_ms": r["host_ms"]})
    th.join()
    wall = time.perf_counter() - t0
    return out, wall
