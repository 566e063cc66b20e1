"""hipsw_banded_extend_noalloc against hipsw_banded_extend.

The allocation-free entry point runs the same kernels with every device
buffer in the library's static scratch array.  Two calls in a row -- the
second one smaller, so it runs over what the first left in the array -- must
give exactly the results and cell counts of the allocating entry point (and of
the oracle), in both dispatch modes; a set too large for the array is refused
before anything runs."""
import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as orc
from swiftbwa_amd import native, workload as wl
from swiftbwa_amd.metalsw import ScoringParameters, _ext_arrays

pytestmark = pytest.mark.gpu


def _noalloc(engine, js, sp, mode):
    lib = engine.lib
    n = int(js["n"])
    out = np.full((n, 6), -7, np.int32)
    a = _ext_arrays(js)
    sc = sp.to_c()
    st = native.HipswStats()
    rc = lib.hipsw_banded_extend_noalloc(
        engine.handle, n, native.ptr(a["queries"]), native.ptr(a["q_off"]), native.ptr(a["q_len"]),
        native.ptr(a["targets"]), native.ptr(a["t_off"]), native.ptr(a["t_len"]),
        native.ptr(a["h0"]), native.ptr(a["w"]), native.ptr(a["end_bonus"]), C.byref(sc), int(mode),
        native.ptr(out), C.byref(st))
    return rc, out, st.cells


def _long_queries(seed, n):
    """Queries of 256..320 bases: the plain form keeps their column state in
    HBM, so that region comes from the static scratch too."""
    rng = np.random.default_rng(seed)
    qs, ts = [], []
    for _ in range(n):
        L = int(rng.integers(256, 321))
        q = rng.integers(0, 4, L).astype(np.uint8)
        t = np.concatenate([q, rng.integers(0, 4, int(rng.integers(0, 80))).astype(np.uint8)])
        flip = rng.random(len(t)) < 0.02
        t[flip] = (t[flip] + 1) % 4
        qs.append(q); ts.append(t)
    ql = np.array([len(x) for x in qs], np.uint16); tl = np.array([len(x) for x in ts], np.uint16)
    qo = np.zeros(n, np.uint32); to = np.zeros(n, np.uint32)
    qo[1:] = np.cumsum(ql.astype(np.int64))[:-1]; to[1:] = np.cumsum(tl.astype(np.int64))[:-1]
    return {"n": n, "queries": np.concatenate(qs), "q_off": qo, "q_len": ql,
            "targets": np.concatenate(ts), "t_off": to, "t_len": tl,
            "h0": rng.integers(10, 80, n).astype(np.int32), "w": np.full(n, 100, np.int32),
            "end_bonus": np.full(n, 5, np.int32)}


@pytest.mark.parametrize("mode", [native.MODE_BATCHED, native.MODE_PLAIN])
def test_noalloc_matches_allocating_entry_point(engine, mode):
    sp = ScoringParameters()
    osc = orc.scoring()
    first = wl.generate(wl.EXT_MIXED, 31, 20000)            # fits the plain form's HBM state too
    second = wl.generate(wl.EXT_UNIFORM, 32, 5000)          # smaller, over the first's leftovers
    third = _long_queries(33, 700)
    for js in (first, second, third):
        print("noalloc mode %d: %d jobs" % (mode, js["n"]), flush=True)
        rc, got, cells = _noalloc(engine, js, sp, mode)
        assert rc == native.HIPSW_OK, engine.last_error()
        st = {}
        want = engine.banded_extend(js, sp, mode=mode, stats=st)
        assert np.array_equal(got, want)
        assert cells == st["cells"]
        ref, ref_cells = orc.ksw_extend_batch(js, osc, nthreads=8)
        assert np.array_equal(got, ref)
        assert cells == ref_cells


def test_noalloc_refuses_sets_beyond_its_scratch(engine):
    sp = ScoringParameters()
    big = wl.generate(wl.EXT_UNIFORM, 34, 400000)            # ~85 MB of job arrays
    rc, out, _ = _noalloc(engine, big, sp, native.MODE_BATCHED)
    assert rc == native.HIPSW_E_BUDGET
    assert "static scratch holds" in engine.last_error()
    assert (out == -7).all()                                 # nothing ran
    # the entry point still works after a refusal
    small = wl.generate(wl.EXT_UNIFORM, 35, 2000)
    rc, got, _ = _noalloc(engine, small, sp, native.MODE_BATCHED)
    assert rc == native.HIPSW_OK
    assert np.array_equal(got, engine.banded_extend(small, sp))
