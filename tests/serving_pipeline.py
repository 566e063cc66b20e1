"""Paired-end read chunks through the serving loop (test helper, not collected).

One read chunk is one batch of the reference's paired-end path
(BWAMemAligner.processFullBatch, CPU order, BWAMemAligner.swift:1857-2128):
the oracle's harness seeds, chains and filters (oracle/e2e_oracle.c); the
chains' four extension passes (hipsw_chain_*, ExtensionAligner.extend), both
mate-rescue directions (MateRescue.collectCandidates with fetchSequence
windows clamped on the server) and the CIGAR NW of the regions that leave the
fast path (hipsw_cigar_route / hipsw_cigar_finish) go to the server as chunks;
the harness then pairs, rescues, re-marks and writes SAM.  Several read chunks
run at once, each on its own thread and harness fork, so their device work
overlaps on the server's slots.

CpuServer is an oracle-backed stand-in for serving.Server (same submit/wait),
with fetchSequence restated below, so the plumbing runs on the CPU too.
"""
import threading
import time

import numpy as np

from oracle import pyoracle as orc
from swiftbwa_amd import serving
from swiftbwa_amd.metalsw import ExtensionAligner, ScoringParameters


def chain_arrays(chains, reads):
    """Harness chains -> the arrays of hipsw_chain_begin."""
    nc = len(chains)
    off = np.zeros(nc + 1, np.uint32)
    if nc:
        off[1:] = np.cumsum([len(s) for _, s in chains])
    flat = [x for _, s in chains for x in s]
    ch = {"off": off,
          "rbeg": np.array([x[0] for x in flat], np.int64),
          "qbeg": np.array([x[1] for x in flat], np.int32),
          "len": np.array([x[2] for x in flat], np.int32),
          "score": np.array([x[3] for x in flat], np.int32),
          "read": np.array([r for r, _ in chains], np.uint32)}
    rl = np.array([len(r) for r in reads], np.int32)
    roff = np.zeros(len(reads), np.uint32)
    if len(reads) > 1:
        roff[1:] = np.cumsum(rl.astype(np.int64))[:-1]
    arena = np.concatenate(reads).astype(np.uint8) if len(reads) and rl.sum() else np.zeros(1, np.uint8)
    return ch, arena, roff, rl


def _wait(server, chunk, log, stage, **kw):
    t = server.submit(chunk)
    r = server.wait(t, **kw)
    log.append((stage, r.get("latency_ms", 0.0), r.get("cells", 0)))
    return r


def run_pe_served(E, server, names, r1, r2, q1, q2, sp: ScoringParameters, log=None):
    """SAM text of one read-pair chunk with its hot path on `server`."""
    log = [] if log is None else log
    chains = E.seed_pairs(r1, r2)
    ch, arena, roff, rl = chain_arrays(chains, E.reads)
    drv = serving.ChainPasses(ch, arena, roff, rl, sp)
    try:
        for _ in range(4):
            c = drv.next_chunk()
            n = c["ext"]["n"]
            out = _wait(server, c, log, "ext%d" % c["pass"])["ext"] if n else np.zeros((0, 6), np.int32)
            drv.feed(out)
        regs, cnt = drv.finish()
    finally:
        drv.close()
    per_chain = [[{k: int(getattr(regs[int(ch["off"][c]) + r], k)) for k in ExtensionAligner.KEYS}
                  for r in range(int(cnt[c]))] for c in range(len(chains))]
    E.set_regions(per_chain)
    if E.pe_finish():
        for direction in (1, 2):
            js = E.rescue_windows(direction)
            if js["n"]:
                span = js["re"] - js["rb"]
                assert span.max() < 65536
                loc = {"n": js["n"], "q": js["queries"], "q_off": js["q_off"], "q_len": js["q_len"],
                       "t_pos": js["rb"], "t_len": span.astype(np.uint16), "rid": js["rid"]}
                r = _wait(server, {"loc": loc}, log, "rescue%d" % direction)
                E.set_rescue_windows(r["loc_win"])
                E.set_rescue(r["loc"], r["loc_ok"])
            E.rescue_apply()
    jobs = E.pe_cigar_plan()
    if jobs:
        cj = cigar_jobs(jobs)
        w, route = serving.cigar_route(cj, sp)
        sel = np.flatnonzero(route == 1)
        nwres = {"score": np.zeros(0, np.int32), "ops": np.zeros(0, np.uint32),
                 "off": np.zeros(1, np.uint32), "n_cigar": np.zeros(0, np.int32),
                 "flags": np.zeros(0, np.int32)}
        if len(sel):
            q_len = cj["q_len"][sel]
            q_off = np.zeros(len(sel), np.uint32)
            q_off[1:] = np.cumsum(q_len.astype(np.int64))[:-1]
            q = np.concatenate([cj["qseg"][cj["q_off"][k]:cj["q_off"][k] + cj["q_len"][k]] for k in sel])
            nw = {"n": len(sel), "q": q.astype(np.uint8), "q_off": q_off, "q_len": q_len,
                  "t_pos": cj["ref_pos"][sel], "t_len": cj["r_len"][sel], "a": w[sel],
                  "b": cj["true_score"][sel]}
            r = _wait(server, {"nw": nw}, log, "cigar", compact=True)
            nwres = {"score": r["nw_score"], "ops": r["nw_cigar_ops"], "off": r["nw_cigar_off"],
                     "n_cigar": r["nw_n_cigar"], "flags": r["nw_flags"]}
        res = serving.cigar_finish(cj, route, nwres, server.nw_cap)
        for j, g in enumerate(res):
            assert g["flags"] == 0, "CIGAR %d over the caps" % j
            E.set_cigar(j, g["cigar"], g["nm"], g["md"], g["pos"])
    return E.pe_sam(names, list(q1) + list(q2))


def cigar_jobs(jobs):
    """Harness CIGAR jobs (list of dicts) -> the arrays of hipsw_cigar_route/finish."""
    n = len(jobs)
    ql = np.array([len(j["qseg"]) for j in jobs], np.uint16)
    rl = np.array([len(j["rseg"]) for j in jobs], np.uint16)
    qo = np.zeros(n, np.uint32)
    ro = np.zeros(n, np.uint32)
    if n > 1:
        qo[1:] = np.cumsum(ql.astype(np.int64))[:-1]
        ro[1:] = np.cumsum(rl.astype(np.int64))[:-1]
    cat = lambda xs: np.concatenate(xs).astype(np.uint8) if sum(len(x) for x in xs) else np.zeros(1, np.uint8)
    return {"qseg": cat([j["qseg"] for j in jobs]), "q_off": qo, "q_len": ql,
            "rseg": cat([j["rseg"] for j in jobs]), "r_off": ro, "r_len": rl,
            "qb": np.array([j["qb"] for j in jobs], np.int32),
            "qe": np.array([j["qe"] for j in jobs], np.int32),
            "read_len": np.array([j["read_len"] for j in jobs], np.int32),
            "rev": np.array([j["rev"] for j in jobs], np.uint8),
            "true_score": np.array([j["true_score"] for j in jobs], np.int32),
            "w": np.array([j["w"] for j in jobs], np.int32),
            "ref_pos": np.array([j["ref_pos"] for j in jobs], np.int64)}


def split_pairs(n, sizes):
    """Consecutive [lo, hi) ranges of the given sizes covering n pairs."""
    out, lo = [], 0
    for s in sizes:
        if lo >= n:
            break
        out.append((lo, min(n, lo + s)))
        lo += s
    if lo < n:
        out.append((lo, n))
    return out


def run_chunks(E, server, names, r1, r2, q, ranges, sp, concurrent=True):
    """SAM text per read chunk (ranges of pairs), several chunks in flight."""
    sams = [None] * len(ranges)
    logs = [[] for _ in ranges]
    errs = []
    walls = [0.0] * len(ranges)

    def one(i):
        lo, hi = ranges[i]
        F = E.fork()
        try:
            t0 = time.perf_counter()
            sams[i] = run_pe_served(F, server, names[lo:hi], r1[lo:hi], r2[lo:hi], q[lo:hi],
                                    q[lo:hi], sp, logs[i])
            walls[i] = time.perf_counter() - t0
        except Exception as ex:      # reported by the caller
            errs.append((i, ex))
        finally:
            F.close()

    if concurrent:
        ths = [threading.Thread(target=one, args=(i,)) for i in range(len(ranges))]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
    else:
        for i in range(len(ranges)):
            one(i)
    if errs:
        raise errs[0][1]
    return sams, logs, walls


# ---- CPU stand-in for serving.Server -------------------------------------------

def fetch_sequence(contigs, gl, rb, re):
    """PackedReference.fetchSequence (PackedReference.swift:72-121) over
    (offset, length) sequences: (clamped BWT begin, length, rid) or None."""
    if not rb < re:
        return None
    fwd = rb < gl
    fb, fe = (rb, re) if fwd else (2 * gl - re, 2 * gl - rb)
    rid = 0
    for k, (off, _) in enumerate(contigs):                  # ReferenceMetadata.sequenceID
        if off <= fb:
            rid = k
    off, ln = contigs[rid]
    cb, ce = max(fb, off), min(fe, off + ln)
    if not cb < ce:
        return None
    return (cb if fwd else 2 * gl - ce), ce - cb, rid


class CpuServer:
    """submit/wait of serving.Server computed by the oracle on the CPU."""

    def __init__(self, ref, contig_lengths=None, nw_cap=512, loc_min_sub=19):
        self.ref = np.ascontiguousarray(ref, np.uint8)
        self.gl = len(ref)
        lens = list(contig_lengths or [self.gl])
        offs = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64)
        self.contigs = [(int(o), int(n)) for o, n in zip(offs, lens)]
        self.nw_cap = nw_cap
        self.loc_min_sub = loc_min_sub
        self._done = {}
        self._next = 1
        self._mu = threading.Lock()

    def submit(self, chunk):
        empty = {"n": 0}
        ch = {k: chunk.get(k) or empty for k in ("ext", "loc", "nw")}
        out = {"latency_ms": 0.0, "cells": 0}
        e = ch["ext"]
        if e["n"]:
            js = orc.chunk_to_job_sets(self.ref, {"ext": e, "loc": _none(), "nw": _none()})["ext"]
            out["ext"], _ = orc.ksw_extend_batch(js, nthreads=4)
        l = ch["loc"]
        if l["n"]:
            loc = dict(l)
            win = np.full(l["n"], -1, np.int64)
            if "rid" in l:
                tp, tl = np.array(l["t_pos"], np.int64), np.zeros(l["n"], np.uint16)
                for k in range(l["n"]):
                    f = fetch_sequence(self.contigs, self.gl, int(l["t_pos"][k]),
                                       int(l["t_pos"][k]) + int(l["t_len"][k]))
                    if f is not None and f[2] == int(l["rid"][k]):
                        tp[k], tl[k], win[k] = f[0], f[1], f[0]
                loc["t_pos"], loc["t_len"] = tp, tl
            js = orc.chunk_to_job_sets(self.ref, {"ext": _none(), "loc": loc, "nw": _none()})["loc"]
            o, ok, _ = orc.local_batch(js, min_sub_score=self.loc_min_sub, nthreads=4)
            o[ok == 0] = 0
            out.update(loc=o, loc_ok=ok, loc_win=win)
        n = ch["nw"]
        if n["n"]:
            js = orc.chunk_to_job_sets(self.ref, {"ext": _none(), "loc": _none(), "nw": n})["nw"]
            score, cig, ncig, _ = orc.global_batch(js, max_retries=3, cap=self.nw_cap, nthreads=4)
            used = np.minimum(np.maximum(ncig, 0), self.nw_cap)
            off = np.zeros(n["n"] + 1, np.uint32)
            off[1:] = np.cumsum(used)
            ops = np.concatenate([cig[k, :used[k]] for k in range(n["n"])]) if off[-1] else np.zeros(0, np.uint32)
            out.update(nw_score=score, nw_cigar_ops=ops.astype(np.uint32), nw_cigar_off=off,
                       nw_n_cigar=ncig, nw_flags=(ncig > self.nw_cap).astype(np.int32))
        with self._mu:
            t = self._next
            self._next += 1
            self._done[t] = out
        return t

    def wait(self, ticket, compact=False):
        with self._mu:
            return self._done.pop(ticket)


def _none():
    z = np.zeros(0)
    return {"n": 0, "q": z, "q_off": z, "q_len": z, "t_pos": z, "t_len": z, "t_rev": z,
            "a": z, "b": z, "c": z}
