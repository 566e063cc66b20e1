"""Concurrent one-shot calls on one engine.

swiftbwa shares one MetalSW engine between up to `-t` concurrent
processFullBatch tasks (BWAMemAligner.swift:1506-1523, MetalBufferPool.swift:
10-23).  libhipsw's one-shot entry points each lease a stream context from the
engine's pool, so calls from several host threads overlap.  Each thread's
results must equal the oracle's, whatever the interleaving.  ctypes releases
the GIL around foreign calls, so the threads below do run their calls at the
same time."""
import threading

import numpy as np
import pytest

from oracle import pyoracle as orc
from swiftbwa_amd import native, workload as wl
from swiftbwa_amd.metalsw import ScoringParameters

pytestmark = pytest.mark.gpu


def _oracle_sc(sp):
    return orc.scoring(a=sp.matchScore, b=sp.mismatchPenalty, o_ins=sp.gapOpenPenalty,
                       e_ins=sp.gapExtendPenalty, o_del=sp.gapOpenPenaltyDeletion,
                       e_del=sp.gapExtendPenaltyDeletion, zdrop=sp.zDrop)


def test_threads_share_one_engine(engine):
    sp = ScoringParameters()
    osc = _oracle_sc(sp)
    ext_sets = [wl.generate(wl.EXT_UNIFORM, 11, 6000), wl.generate(wl.EXT_MIXED, 12, 8000)]
    ext_ref = [orc.ksw_extend_batch(js, osc, nthreads=8)[0] for js in ext_sets]
    loc = wl.generate(wl.LOCAL, 13, 3000)
    loc_ref, loc_ok, _ = orc.local_batch(loc, osc, min_sub_score=19, nthreads=8)
    loc_ref[loc_ok == 0] = 0
    nw = wl.generate(wl.NW, 14, 4000)
    nw_ref = orc.global_batch(nw, osc, max_retries=3, cap=512, nthreads=8)

    errors = []

    def ext_worker(k, mode):
        try:
            for _ in range(3):
                got = engine.banded_extend(ext_sets[k], sp, mode=mode)
                if not np.array_equal(got, ext_ref[k]):
                    errors.append("extension set %d mode %d differs" % (k, mode))
        except Exception as e:   # noqa: BLE001 -- reported below
            errors.append("extension: %r" % e)

    def loc_worker():
        try:
            for _ in range(3):
                got, ok = engine.local_sw(loc, sp, min_sub_score=19)
                if not (np.array_equal(ok, loc_ok) and np.array_equal(got, loc_ref)):
                    errors.append("local SW differs")
        except Exception as e:   # noqa: BLE001
            errors.append("local: %r" % e)

    def nw_worker(mode):
        try:
            for _ in range(3):
                score, cig, ncig, flags, _fw = engine.global_nw(nw, sp, mode=mode)
                if not np.array_equal(score, nw_ref[0]) or not np.array_equal(ncig, nw_ref[2]):
                    errors.append("NW mode %d differs" % mode)
                    continue
                for j in range(nw["n"]):
                    m = min(int(ncig[j]), 512)
                    if not np.array_equal(cig[j, :m], nw_ref[1][j, :m]):
                        errors.append("NW CIGAR %d differs" % j)
                        break
        except Exception as e:   # noqa: BLE001
            errors.append("nw: %r" % e)

    threads = [threading.Thread(target=ext_worker, args=(0, native.MODE_BATCHED)),
               threading.Thread(target=ext_worker, args=(1, native.MODE_BATCHED)),
               threading.Thread(target=ext_worker, args=(0, native.MODE_PLAIN)),
               threading.Thread(target=ext_worker, args=(1, native.MODE_PLAIN)),
               threading.Thread(target=loc_worker), threading.Thread(target=loc_worker),
               threading.Thread(target=nw_worker, args=(native.MODE_BATCHED,)),
               threading.Thread(target=nw_worker, args=(native.MODE_PLAIN,))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=100)
    assert not any(t.is_alive() for t in threads), "a worker did not finish"
    assert not errors, errors[:5]


def test_thread_local_errors(engine):
    """A failing call reports its own message on its own thread."""
    sp = ScoringParameters()
    js = wl.generate(wl.EXT_UNIFORM, 15, 64)
    bad = dict(js)
    bad["w"] = js["w"].copy()
    bad["w"][3] = -1
    msgs = {}

    def fail():
        try:
            engine.banded_extend(bad, sp)
        except native.HipswError as e:
            msgs["fail"] = (e.status, engine.last_error())

    def good():
        engine.banded_extend(js, sp)
        msgs["good"] = engine.last_error()

    t1, t2 = threading.Thread(target=fail), threading.Thread(target=good)
    t1.start(); t2.start(); t1.join(); t2.join()
    assert msgs["fail"][0] == native.HIPSW_E_INVALID
    assert "band width" in msgs["fail"][1]
    assert msgs["good"] == ""
