"""GPU parity of the banded-extension kernels against the CPU oracle
(BandedSWScalar.align restated in oracle/sw_oracle.c).  Bit-exact on all six
SWResult fields and on the DP cell count."""
import numpy as np
import pytest

from oracle import pyoracle as orc
from swiftbwa_amd import native, workload as wl
from swiftbwa_amd.metalsw import (BandedSWDispatcher, BandedSWTask, ScoringParameters,
                                  SWResult)

pytestmark = pytest.mark.gpu

MODES = [native.MODE_PLAIN, native.MODE_BATCHED]


def oracle_sc(sp: ScoringParameters):
    return orc.scoring(a=sp.matchScore, b=sp.mismatchPenalty, o_ins=sp.gapOpenPenalty,
                       e_ins=sp.gapExtendPenalty, o_del=sp.gapOpenPenaltyDeletion,
                       e_del=sp.gapExtendPenaltyDeletion, zdrop=sp.zDrop)


def check(engine, js, sp=None, modes=MODES):
    sp = sp or ScoringParameters()
    ref, cells = orc.ksw_extend_batch(js, oracle_sc(sp), nthreads=8)
    for mode in modes:
        st = {}
        got = engine.banded_extend(js, sp, mode=mode, stats=st)
        bad = np.nonzero((got != ref).any(axis=1))[0]
        assert len(bad) == 0, "mode %d: %d mismatches, first job %d: gpu %s oracle %s" % (
            mode, len(bad), bad[0], got[bad[0]].tolist(), ref[bad[0]].tolist())
        assert st["cells"] == cells, (mode, st["cells"], cells)
    return ref


def test_known_answers_through_dispatcher(engine):
    # AlignmentTests.swift:48-63, :82-96, :115-136, :208-237 (closed forms)
    sp = ScoringParameters()
    tasks = [BandedSWTask(bytes([0, 1, 2, 3, 0, 1]), bytes([0, 1, 2, 3, 0, 1]), 1, 10, sp),
             BandedSWTask(bytes([0, 1, 2]), bytes([0, 1, 2]), 5, 10, sp),
             BandedSWTask(bytes([0, 1, 2, 3] * 2), bytes([0, 1, 2, 3] * 2), 1, 10, sp),
             BandedSWTask(bytes([0, 1, 2, 3, 0, 1, 2, 3, 0, 1]),
                          bytes([0, 1, 2, 3, 3, 1, 2, 3, 0, 1]), 1, 10, sp)]
    res = BandedSWDispatcher.dispatchBatch16(tasks, engine)
    assert [r.score for r in res] == [7, 8, 9, 6]
    # z-drop case (AlignmentTests.swift:239-283): score 6, targetEnd < 50
    sz = ScoringParameters(zDrop=10)
    t = BandedSWTask(bytes([0, 1, 2, 3, 0] + [0] * 100), bytes([0, 1, 2, 3, 0] + [3] * 100), 1, 5, sz)
    r = BandedSWDispatcher.dispatchBatch8([t], engine)[0]
    assert r.score == 6 and r.targetEnd < 50


def test_empty_tasks(engine):
    sp = ScoringParameters()
    tasks = [BandedSWTask(b"", bytes([0, 1]), 5, 10, sp), BandedSWTask(bytes([0, 1]), b"", 5, 10, sp),
             BandedSWTask(bytes([0, 1]), bytes([0, 1]), 5, 10, sp)]
    res = BandedSWDispatcher.dispatchBatch16(tasks, engine)
    assert res[0] == SWResult() and res[1] == SWResult()
    assert res[2].score == 7


def test_uniform_subset(engine):
    check(engine, wl.generate(wl.EXT_UNIFORM, 2, 20000))


def test_mixed_subset(engine):
    check(engine, wl.generate(wl.EXT_MIXED, 4, 30000))


def _random_jobs(rng, n, qmax, tmax, n_rate=0.0, h0_range=(1, 60), w_choices=(1, 5, 20, 100)):
    qs, ts, ql, tl, h0, w, eb = [], [], [], [], [], [], []
    for _ in range(n):
        q = rng.integers(0, 4, rng.integers(0, qmax + 1)).astype(np.uint8)
        t = rng.integers(0, 4, rng.integers(0, tmax + 1)).astype(np.uint8)
        # make targets similar to queries half of the time
        if len(q) and len(t) and rng.random() < 0.5:
            m = min(len(q), len(t))
            t[:m] = q[:m]
            flip = rng.random(m) < 0.05
            t[:m][flip] = rng.integers(0, 4, flip.sum())
        if n_rate:
            q[rng.random(len(q)) < n_rate] = 4
            t[rng.random(len(t)) < n_rate] = 4
        qs.append(q); ts.append(t)
        ql.append(len(q)); tl.append(len(t))
        h0.append(int(rng.integers(h0_range[0], h0_range[1] + 1)))
        w.append(int(rng.choice(w_choices)))
        eb.append(int(rng.integers(0, 8)))
    qo = np.zeros(n, np.uint32); to = np.zeros(n, np.uint32)
    qo[1:] = np.cumsum(ql)[:-1]; to[1:] = np.cumsum(tl)[:-1]
    cat = lambda xs: np.concatenate(xs) if sum(map(len, xs)) else np.zeros(1, np.uint8)
    return {"n": n, "queries": cat(qs), "targets": cat(ts), "q_off": qo, "t_off": to,
            "q_len": np.array(ql, np.uint16), "t_len": np.array(tl, np.uint16),
            "h0": np.array(h0, np.int32), "w": np.array(w, np.int32),
            "end_bonus": np.array(eb, np.int32)}


def test_random_small_with_n(engine):
    rng = np.random.default_rng(11)
    check(engine, _random_jobs(rng, 3000, 40, 60, n_rate=0.05))


def test_one_shot_reuses_host_buffers(engine):
    """hipsw_banded_extend packs into host buffers the engine keeps between
    calls and does not clear them: a large set, then a smaller one of other
    shapes (N bases, empty jobs, longer queries), then the large one again,
    each equal to the oracle, so no word of an earlier call leaks into a
    later one."""
    big = wl.generate(wl.EXT_MIXED, 21, 40000)
    rng = np.random.default_rng(22)
    small = _random_jobs(rng, 2500, 200, 260, n_rate=0.03)
    for js in (big, small, big, small):
        check(engine, js, modes=[native.MODE_BATCHED])


def test_int32_tier_and_negative_h0(engine):
    rng = np.random.default_rng(12)
    js = _random_jobs(rng, 1500, 80, 120, h0_range=(-20, 70000))
    check(engine, js)


def test_long_queries_global_state(engine):
    # qlen up to 700 pushes the u16 state past the per-block LDS limit
    rng = np.random.default_rng(13)
    check(engine, _random_jobs(rng, 300, 700, 900, h0_range=(20, 300), w_choices=(50, 100, 400)))


def test_scoring_variants(engine):
    rng = np.random.default_rng(14)
    js = _random_jobs(rng, 2000, 100, 150)
    for sp in (ScoringParameters(zDrop=0), ScoringParameters(matchScore=2, mismatchPenalty=3),
               ScoringParameters(gapOpenPenalty=3, gapExtendPenalty=2, gapOpenPenaltyDeletion=8,
                                 gapExtendPenaltyDeletion=3, zDrop=30)):
        check(engine, js, sp)


def test_full_size_uniform_properties(engine):
    # 10^6 jobs: invariants that hold job by job at full size, plus
    # plain == batched (two different launch shapes and orders).
    js = wl.generate(wl.EXT_UNIFORM, 2, 1000000)
    sp = ScoringParameters()
    st_a, st_b = {}, {}
    a = engine.banded_extend(js, sp, mode=native.MODE_PLAIN, stats=st_a)
    b = engine.banded_extend(js, sp, mode=native.MODE_BATCHED, stats=st_b)
    assert np.array_equal(a, b)
    assert st_a["cells"] == st_b["cells"]
    ql = js["q_len"].astype(np.int32); tl = js["t_len"].astype(np.int32)
    assert (a[:, 0] >= js["h0"]).all()
    assert ((a[:, 1] >= 0) & (a[:, 1] <= ql)).all()
    assert ((a[:, 2] >= 0) & (a[:, 2] <= tl)).all()
    assert ((a[:, 4] == -1) == (a[:, 3] == 0)).all()
    # a random sample checked against the oracle
    idx = np.random.default_rng(5).choice(js["n"], 5000, replace=False)
    sub = wl.slice_jobs(js, np.sort(idx))
    ref, _ = orc.ksw_extend_batch(sub, nthreads=8)
    assert np.array_equal(a[np.sort(idx)], ref)


def test_full_size_uneven_4M_every_row(engine):
    """The whole uneven 4,000,000-job set (the benchmarked one): every row and
    the cell count against the oracle, batched and plain."""
    import os
    js = wl.generate(wl.EXT_MIXED, 4, 4000000)
    ref, cells = orc.ksw_extend_batch(js, nthreads=min(16, os.cpu_count() or 1))
    sp = ScoringParameters()
    for mode in MODES:
        print("uneven 4M mode %d" % mode, flush=True)
        st = {}
        got = engine.banded_extend(js, sp, mode=mode, stats=st)
        bad = np.nonzero((got != ref).any(axis=1))[0]
        assert len(bad) == 0, (mode, len(bad), bad[:5])
        assert st["cells"] == cells


def test_autotune_both_layouts_exact(engine):
    """The autotuner times the packed and the one-job-per-lane layouts on each
    8-bit bin; the lane layout runs last, so the sampled jobs' results on the
    device right after tuning are its results.  Both must equal the oracle, and
    the re-planned set must give the same results again."""
    import ctypes as C
    from swiftbwa_amd.metalsw import _ext_arrays
    js = wl.generate(wl.EXT_MIXED, 7, 40000)
    sp = ScoringParameters()
    ref, _ = orc.ksw_extend_batch(js, oracle_sc(sp), nthreads=8)
    lib = engine.lib
    a = _ext_arrays(js)
    sc = sp.to_c()
    st = lib.hipsw_ext_upload(engine.handle, js["n"], native.ptr(a["queries"]),
                              native.ptr(a["q_off"]), native.ptr(a["q_len"]),
                              native.ptr(a["targets"]), native.ptr(a["t_off"]),
                              native.ptr(a["t_len"]), native.ptr(a["h0"]), native.ptr(a["w"]),
                              native.ptr(a["end_bonus"]), C.byref(sc))
    assert st
    try:
        s = native.HipswStats()
        out = np.zeros((js["n"], 6), np.int32)
        assert lib.hipsw_ext_run(engine.handle, st, native.MODE_BATCHED, C.byref(s)) == 0
        assert lib.hipsw_ext_download(engine.handle, st, native.ptr(out)) == 0
        assert np.array_equal(out, ref)
        rep = (native.HipswTuneEntry * 32)()
        nrep = C.c_int32()
        assert lib.hipsw_ext_autotune(engine.handle, st, 6000, rep, 32, C.byref(nrep)) == 0
        ents = rep[:nrep.value]
        bins = sorted({(e.tier, e.length_class) for e in ents})   # (tier, class)
        assert 1 <= len([b for b in bins if b[0] == 0]) <= 3
        for tier, c in bins:
            es = [e for e in ents if e.tier == tier and e.length_class == c]
            assert sum(e.chosen for e in es) == 1
            for e in es:
                assert e.jobs > 0 and e.ms > 0
                assert 4 <= e.lds_blocks <= 10 if e.kind != 3 else 2 <= e.lds_blocks <= 5
            if tier == 0:
                assert es[-1].kind == 2                      # one job per lane runs last
                assert {e.kind for e in es} == {1, 2}
            else:                                            # 16-bit tier: both forms (scores < 2048)
                assert tier == 1 and {e.kind for e in es} == {2, 3}
        # every LDS/HBM split of the 64-159 and 160-255 classes was timed
        t8 = [e for e in ents if e.tier == 0 and e.kind == 1]
        assert {e.lds_blocks for e in t8 if e.length_class == 1} == {10, 6, 5}
        assert {e.lds_blocks for e in t8 if e.length_class == 2} == {10, 8, 6}
        t16 = [e for e in ents if e.tier == 1 and e.kind == 2]
        for c, splits in ((1, {10, 6, 5}), (2, {10, 8, 6})):
            if any(e.length_class == c for e in t16):
                assert {e.lds_blocks for e in t16 if e.length_class == c} == splits
        p16 = [e for e in ents if e.tier == 1 and e.kind == 3]
        for c, splits in ((0, {4, 2}), (1, {5, 3, 2}), (2, {5, 3, 2})):
            if any(e.length_class == c for e in p16):
                assert {e.lds_blocks for e in p16 if e.length_class == c} == splits
        assert lib.hipsw_ext_download(engine.handle, st, native.ptr(out)) == 0
        assert np.array_equal(out, ref)                       # lane-layout results
        assert lib.hipsw_ext_run(engine.handle, st, native.MODE_BATCHED, C.byref(s)) == 0
        assert lib.hipsw_ext_download(engine.handle, st, native.ptr(out)) == 0
        assert np.array_equal(out, ref)                       # re-planned set
    finally:
        lib.hipsw_ext_free(engine.handle, st)


def plan_of(engine, js, sp=None):
    """The batched plan hipsw_ext_plan reports for js: one dict per bin."""
    import ctypes as C
    from swiftbwa_amd.metalsw import _ext_arrays
    lib = engine.lib
    a = _ext_arrays(js)
    sc = (sp or ScoringParameters()).to_c()
    st = lib.hipsw_ext_upload(engine.handle, js["n"], native.ptr(a["queries"]),
                              native.ptr(a["q_off"]), native.ptr(a["q_len"]),
                              native.ptr(a["targets"]), native.ptr(a["t_off"]),
                              native.ptr(a["t_len"]), native.ptr(a["h0"]), native.ptr(a["w"]),
                              native.ptr(a["end_bonus"]), C.byref(sc))
    assert st
    try:
        ents = (native.HipswPlanEntry * 64)()
        k = C.c_int32()
        assert lib.hipsw_ext_plan(engine.handle, st, ents, 64, C.byref(k)) == 0
        return [{f: getattr(e, f) for f, _ in native.HipswPlanEntry._fields_} for e in ents[:k.value]]
    finally:
        lib.hipsw_ext_free(engine.handle, st)


def test_every_layout_exact(engine):
    """Every built layout of the register-query kernels -- the packed form at
    each LDS/HBM split of the column state and one job per lane -- pinned per
    length class through hipsw_ext_set_layout, on a set that fills every class
    (queries at the split edges 79..81 and 95..97 included)."""
    js = wl.generate(wl.EXT_MIXED, 11, 12000)
    rng = np.random.default_rng(3)
    extra = []
    for ql in (79, 80, 81, 95, 96, 97, 127, 128, 129, 175, 176, 177, 250):
        for rep in range(40):
            q = rng.integers(0, 4, ql).astype(np.uint8)
            t = q.copy()
            flip = rng.random(ql) < 0.03
            t[flip] = (t[flip] + 1) % 4
            extra.append((q, np.concatenate([t, rng.integers(0, 4, rep * 3).astype(np.uint8)])))
    qs = [js["queries"][js["q_off"][k]:js["q_off"][k] + js["q_len"][k]] for k in range(js["n"])]
    ts = [js["targets"][js["t_off"][k]:js["t_off"][k] + js["t_len"][k]] for k in range(js["n"])]
    qs += [q for q, _ in extra]
    ts += [t for _, t in extra]
    n = len(qs)
    ql = np.array([len(x) for x in qs], np.uint16); tl = np.array([len(x) for x in ts], np.uint16)
    qo = np.zeros(n, np.uint32); to = np.zeros(n, np.uint32)
    qo[1:] = np.cumsum(ql.astype(np.int64))[:-1]; to[1:] = np.cumsum(tl.astype(np.int64))[:-1]
    h0 = np.concatenate([js["h0"], rng.integers(1, 60, len(extra)).astype(np.int32)])
    w = np.concatenate([js["w"], np.full(len(extra), 100, np.int32)])
    eb = np.concatenate([js["end_bonus"], np.full(len(extra), 5, np.int32)])
    big = {"n": n, "queries": np.concatenate(qs), "q_off": qo, "q_len": ql,
           "targets": np.concatenate(ts), "t_off": to, "t_len": tl, "h0": h0, "w": w,
           "end_bonus": eb}
    lib = engine.lib
    assert lib.hipsw_ext_set_layout(engine.handle, 1, 1, 7) == native.HIPSW_E_INVALID
    assert lib.hipsw_ext_set_layout(engine.handle, 1, 2, 7) == native.HIPSW_E_INVALID
    assert lib.hipsw_ext_set_layout(engine.handle, 3, 1, 10) == native.HIPSW_E_INVALID
    combos = [((1, 4), (1, 10), (1, 10)), ((2, 4), (1, 6), (1, 8)),
              ((1, 4), (1, 5), (1, 6)), ((2, 4), (2, 10), (2, 10)),
              ((2, 4), (2, 5), (2, 6)), ((2, 4), (2, 6), (2, 8))]
    try:
        for combo in combos:
            for cls, (kind, lb) in enumerate(combo):
                assert lib.hipsw_ext_set_layout(engine.handle, cls, kind, lb) == 0
            # the pinned layout reached every 8-bit register-query bin
            seen = set()
            for b in plan_of(engine, big):
                if b["tier"] == 0 and b["kind"] in (1, 2) and not b["has_n"]:
                    assert (b["kind"], b["lds_blocks"]) == combo[b["length_class"]], (b, combo)
                    seen.add(b["length_class"])
            assert seen == {0, 1, 2}
            check(engine, big, modes=(native.MODE_BATCHED,))
    finally:
        for cls in range(3):
            lib.hipsw_ext_set_layout(engine.handle, cls, 0, 0)


def test_every_layout16_exact(engine):
    """Every built layout of the 16-bit tier's one-job-per-lane kernel, pinned
    per length class through hipsw_ext_set_layout16, on jobs whose seed score
    puts them in the 16-bit tier in every class (split edges included)."""
    rng = np.random.default_rng(23)
    qs, ts = [], []
    for ql in (5, 40, 63, 64, 79, 80, 81, 95, 96, 97, 120, 159, 160, 175, 176, 177, 200, 250):
        for rep in range(30):
            q = rng.integers(0, 4, ql).astype(np.uint8)
            t = q.copy()
            flip = rng.random(ql) < 0.03
            t[flip] = (t[flip] + 1) % 4
            qs.append(q)
            ts.append(np.concatenate([t, rng.integers(0, 4, rep * 4).astype(np.uint8)]))
    n = len(qs)
    ql = np.array([len(x) for x in qs], np.uint16); tl = np.array([len(x) for x in ts], np.uint16)
    qo = np.zeros(n, np.uint32); to = np.zeros(n, np.uint32)
    qo[1:] = np.cumsum(ql.astype(np.int64))[:-1]; to[1:] = np.cumsum(tl.astype(np.int64))[:-1]
    js = {"n": n, "queries": np.concatenate(qs), "q_off": qo, "q_len": ql,
          "targets": np.concatenate(ts), "t_off": to, "t_len": tl,
          "h0": rng.integers(200, 1000, n).astype(np.int32), "w": np.full(n, 100, np.int32),
          "end_bonus": np.full(n, 5, np.int32)}
    lib = engine.lib
    assert lib.hipsw_ext_set_layout16(engine.handle, 1, 7) == native.HIPSW_E_INVALID
    assert lib.hipsw_ext_set_layout16(engine.handle, 3, 10) == native.HIPSW_E_INVALID
    try:
        for combo in ((4, 10, 10), (4, 6, 8), (4, 5, 6)):
            for cls, lb in enumerate(combo):
                assert lib.hipsw_ext_set_layout16(engine.handle, cls, lb) == 0
            # the pinned split reached every 16-bit bin (one job per lane)
            bins = [b for b in plan_of(engine, js) if b["tier"] == 1]
            assert {b["length_class"] for b in bins} == {0, 1, 2}
            for b in bins:
                assert b["kind"] == 2 and b["lds_blocks"] == combo[b["length_class"]], (b, combo)
            check(engine, js, modes=(native.MODE_BATCHED,))
    finally:
        for cls in range(3):
            lib.hipsw_ext_set_layout16(engine.handle, cls, 0)


def test_every_layout16_packed_exact(engine):
    """The 16-bit tier's two-jobs-per-lane kernel (h and e planes of 16-bit
    halves) at every built split, pinned per class through
    hipsw_ext_set_layout16_kind; bins whose score bound reaches 2048, or a
    scoring with a gap extension above 31, keep the one-job-per-lane kernel."""
    rng = np.random.default_rng(24)
    qs, ts = [], []
    for ql in (5, 40, 63, 64, 79, 80, 81, 95, 96, 97, 120, 159, 160, 175, 176, 177, 200, 250):
        for rep in range(30):
            q = rng.integers(0, 4, ql).astype(np.uint8)
            t = q.copy()
            flip = rng.random(ql) < 0.03
            t[flip] = (t[flip] + 1) % 4
            qs.append(q)
            ts.append(np.concatenate([t, rng.integers(0, 4, rep * 4).astype(np.uint8)]))
    n = len(qs)
    ql = np.array([len(x) for x in qs], np.uint16); tl = np.array([len(x) for x in ts], np.uint16)
    qo = np.zeros(n, np.uint32); to = np.zeros(n, np.uint32)
    qo[1:] = np.cumsum(ql.astype(np.int64))[:-1]; to[1:] = np.cumsum(tl.astype(np.int64))[:-1]
    # 16-bit tier in every class, bounds h0 + qlen up to 1,999 (below 2,048)
    h0 = np.maximum(rng.integers(200, 1750, n), 256 - ql.astype(np.int64)).astype(np.int32)
    js = {"n": n, "queries": np.concatenate(qs), "q_off": qo, "q_len": ql,
          "targets": np.concatenate(ts), "t_off": to, "t_len": tl,
          "h0": h0, "w": rng.choice(np.array([10, 50, 100, 300], np.int32), n),
          "end_bonus": np.full(n, 5, np.int32)}
    lib = engine.lib
    assert lib.hipsw_ext_set_layout16_kind(engine.handle, 1, 3, 4) == native.HIPSW_E_INVALID
    assert lib.hipsw_ext_set_layout16_kind(engine.handle, 0, 3, 5) == native.HIPSW_E_INVALID
    assert lib.hipsw_ext_set_layout16_kind(engine.handle, 1, 1, 5) == native.HIPSW_E_INVALID
    try:
        for combo in ((4, 5, 5), (2, 3, 3), (4, 2, 2), (0, 0, 0)):
            for cls, lb in enumerate(combo):
                assert lib.hipsw_ext_set_layout16_kind(engine.handle, cls, 3, lb) == 0
            bins = [b for b in plan_of(engine, js) if b["tier"] == 1]
            assert {b["length_class"] for b in bins} == {0, 1, 2}
            for b in bins:
                want = combo[b["length_class"]] or (4 if b["length_class"] == 0 else 5)
                assert b["kind"] == 3 and b["lds_blocks"] == want, (b, combo)
            check(engine, js, modes=(native.MODE_BATCHED,))
            # other scorings: a mismatch-heavy matrix and asymmetric gaps
            check(engine, js, ScoringParameters(matchScore=2, mismatchPenalty=3, zDrop=40),
                  modes=(native.MODE_BATCHED,))
            check(engine, js, ScoringParameters(gapOpenPenalty=3, gapExtendPenalty=2,
                                                gapOpenPenaltyDeletion=8, gapExtendPenaltyDeletion=3,
                                                zDrop=30), modes=(native.MODE_BATCHED,))
        # a bound of 2,048 or more in a class keeps that class on one job per lane
        big = dict(js)
        big["h0"] = js["h0"].copy()
        big["h0"][ql.astype(np.int64) >= 160] = 3000
        for b in [b for b in plan_of(engine, big) if b["tier"] == 1]:
            assert b["kind"] == (2 if b["length_class"] == 2 else 3), b
        check(engine, big, modes=(native.MODE_BATCHED,))
        # so does a gap extension above 31
        sp_e = ScoringParameters(gapExtendPenalty=40, gapExtendPenaltyDeletion=40)
        for b in [b for b in plan_of(engine, js, sp_e) if b["tier"] == 1]:
            assert b["kind"] == 2, b
    finally:
        for cls in range(3):
            lib.hipsw_ext_set_layout16(engine.handle, cls, 0)


def test_size_and_tier_boundaries(engine):
    """Query lengths at every block / LDS / class edge (15..17, 63/64, 159..161,
    255/256), targets up to the driver's 10,000-base window, bands 0..1000, and
    h0 placed exactly at the 8/16-bit score-tier bounds."""
    rng = np.random.default_rng(21)
    qls = [0, 1, 15, 16, 17, 63, 64, 65, 159, 160, 161, 191, 254, 255, 256, 300]
    qs, ts, h0, w, eb = [], [], [], [], []
    for ql in qls:
        for rep in range(24):
            q = rng.integers(0, 4, ql).astype(np.uint8)
            tl = int(rng.choice([0, 1, ql, ql + 50, 2 * ql + 200, 10000 if rep == 0 else 600]))
            t = rng.integers(0, 4, tl).astype(np.uint8)
            m = min(ql, tl)
            if m and rep % 3:                                   # similar prefix: long alignments
                t[:m] = q[:m]
                flip = rng.random(m) < 0.03
                t[:m][flip] = (t[:m][flip] + 1) % 4
            qs.append(q); ts.append(t)
            bound = [255 - ql, 256 - ql, 65535 - ql, 65536 - ql, 20, 0][rep % 6]
            h0.append(max(0, bound))
            w.append(int([0, 1, 2, 100, 1000][rep % 5]))
            eb.append(int(rep % 7))
    n = len(qs)
    ql = np.array([len(x) for x in qs], np.uint16); tl = np.array([len(x) for x in ts], np.uint16)
    qo = np.zeros(n, np.uint32); to = np.zeros(n, np.uint32)
    qo[1:] = np.cumsum(ql.astype(np.int64))[:-1]; to[1:] = np.cumsum(tl.astype(np.int64))[:-1]
    js = {"n": n, "queries": np.concatenate(qs), "q_off": qo, "q_len": ql,
          "targets": np.concatenate(ts), "t_off": to, "t_len": tl,
          "h0": np.array(h0, np.int32), "w": np.array(w, np.int32),
          "end_bonus": np.array(eb, np.int32)}
    check(engine, js)
    check(engine, js, ScoringParameters(gapOpenPenalty=4, gapExtendPenalty=2, zDrop=40))


def test_batched_graph_replays(engine):
    """The batched form runs as a captured graph (counter resets, then the
    bins).  Replays interleaved with eager work on the same
    engine -- the plain form of the same set, and another set whose larger
    scratch need moves the engine's scratch region (which forces a re-capture)
    -- each give the oracle's results and cell count."""
    import ctypes as C
    from swiftbwa_amd.metalsw import _ext_arrays
    sp = ScoringParameters()
    sc = sp.to_c()
    lib = engine.lib

    def upload(js):
        a = _ext_arrays(js)
        st = lib.hipsw_ext_upload(engine.handle, js["n"], native.ptr(a["queries"]),
                                  native.ptr(a["q_off"]), native.ptr(a["q_len"]),
                                  native.ptr(a["targets"]), native.ptr(a["t_off"]),
                                  native.ptr(a["t_len"]), native.ptr(a["h0"]), native.ptr(a["w"]),
                                  native.ptr(a["end_bonus"]), C.byref(sc))
        assert st
        return st

    js1 = wl.generate(wl.EXT_MIXED, 11, 30000)
    js2 = wl.generate(wl.EXT_MIXED, 12, 60000)
    ref1, cells1 = orc.ksw_extend_batch(js1, oracle_sc(sp), nthreads=8)
    ref2, cells2 = orc.ksw_extend_batch(js2, oracle_sc(sp), nthreads=8)
    s1, s2 = upload(js1), upload(js2)
    try:
        out1 = np.zeros((js1["n"], 6), np.int32)
        out2 = np.zeros((js2["n"], 6), np.int32)

        def run(st, mode, out, ref, cells):
            s = native.HipswStats()
            assert lib.hipsw_ext_run(engine.handle, st, mode, C.byref(s)) == 0
            assert lib.hipsw_ext_download(engine.handle, st, native.ptr(out)) == 0
            assert np.array_equal(out, ref)
            assert s.cells == cells

        for rep in range(4):
            run(s1, native.MODE_BATCHED, out1, ref1, cells1)
            run(s1, native.MODE_PLAIN, out1, ref1, cells1)
            run(s2, native.MODE_BATCHED, out2, ref2, cells2)
        run(s1, native.MODE_BATCHED, out1, ref1, cells1)
    finally:
        lib.hipsw_ext_free(engine.handle, s1)
        lib.hipsw_ext_free(engine.handle, s2)


def test_input_order_form_exact(engine):
    """HIPSW_MODE_INPUT_ORDER (the batched form's own kernel over the jobs in
    input order, one work item per block): oracle results and cell counts on
    an 8-bit set (packed kernel) and on a set with 16-bit jobs (one job per
    lane); refused for a set that needs the generic kernel."""
    import ctypes as C
    from swiftbwa_amd.metalsw import _ext_arrays
    sp = ScoringParameters()
    sc = sp.to_c()
    lib = engine.lib
    for js in (wl.generate(wl.EXT_UNIFORM, 41, 20000), wl.generate(wl.EXT_MIXED, 42, 30000)):
        ref, cells = orc.ksw_extend_batch(js, oracle_sc(sp), nthreads=8)
        a = _ext_arrays(js)
        st = lib.hipsw_ext_upload(engine.handle, js["n"], native.ptr(a["queries"]),
                                  native.ptr(a["q_off"]), native.ptr(a["q_len"]),
                                  native.ptr(a["targets"]), native.ptr(a["t_off"]),
                                  native.ptr(a["t_len"]), native.ptr(a["h0"]), native.ptr(a["w"]),
                                  native.ptr(a["end_bonus"]), C.byref(sc))
        assert st
        try:
            out = np.zeros((js["n"], 6), np.int32)
            for rep in range(2):
                s = native.HipswStats()
                assert lib.hipsw_ext_run(engine.handle, st, 2, C.byref(s)) == 0, engine.last_error()
                assert lib.hipsw_ext_download(engine.handle, st, native.ptr(out)) == 0
                assert np.array_equal(out, ref)
                assert s.cells == cells and s.launches == 1
        finally:
            lib.hipsw_ext_free(engine.handle, st)
    bad = wl.generate(wl.EXT_UNIFORM, 43, 500)
    bad["queries"] = bad["queries"].copy()
    bad["queries"][5] = 4                                   # an N: generic kernel only
    a = _ext_arrays(bad)
    st = lib.hipsw_ext_upload(engine.handle, bad["n"], native.ptr(a["queries"]),
                              native.ptr(a["q_off"]), native.ptr(a["q_len"]),
                              native.ptr(a["targets"]), native.ptr(a["t_off"]),
                              native.ptr(a["t_len"]), native.ptr(a["h0"]), native.ptr(a["w"]),
                              native.ptr(a["end_bonus"]), C.byref(sc))
    try:
        assert lib.hipsw_ext_run(engine.handle, st, 2, None) == native.HIPSW_E_UNSUPPORTED
    finally:
        lib.hipsw_ext_free(engine.handle, st)
