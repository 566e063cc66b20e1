"""GPU parity of the banded global NW + traceback kernel against the CPU oracle
(GlobalAligner.align + traceback + CIGARGenerator's retry loop, restated in
oracle/sw_oracle.c).  Bit-exact on score, CIGAR, op count and final band."""
import numpy as np
import pytest

from oracle import pyoracle as orc
from swiftbwa_amd import native, workload as wl
from swiftbwa_amd.metalsw import GlobalSWDispatcher, GlobalSWTask, ScoringParameters

pytestmark = pytest.mark.gpu
MODES = [native.MODE_PLAIN, native.MODE_BATCHED]


def oracle_sc(sp):
    return orc.scoring(a=sp.matchScore, b=sp.mismatchPenalty, o_ins=sp.gapOpenPenalty,
                       e_ins=sp.gapExtendPenalty, o_del=sp.gapOpenPenaltyDeletion,
                       e_del=sp.gapExtendPenaltyDeletion, zdrop=sp.zDrop)


def check(engine, js, sp=None, max_retries=3, cap=512, forms=(0, 1, 2, 3)):
    """Every mode, and the 16-bit tier's row state in registers (form 0 chooses;
    form 2 one job per lane; form 3 two jobs per lane in int16 pairs wherever
    the band is at most 64 columns) and in the LDS ring (form 1)."""
    sp = sp or ScoringParameters()
    rs, rc, rn, cells = orc.global_batch(js, oracle_sc(sp), max_retries=max_retries, cap=cap,
                                         nthreads=8)
    for form, mode in [(f, m) for f in forms for m in MODES]:
        engine.set_nw_state(form)
        st = {}
        try:
            s, c, n, fl, fw = engine.global_nw(js, sp, max_retries=max_retries, cap=cap,
                                               mode=mode, stats=st)
        finally:
            engine.set_nw_state(0)
        bad = np.nonzero(s != rs)[0]
        assert len(bad) == 0, "mode %d score mismatch on %d jobs, first %d: %d vs %d" % (
            mode, len(bad), bad[0], s[bad[0]], rs[bad[0]])
        assert np.array_equal(n, rn), mode
        assert np.array_equal(fl, (rn > cap).astype(np.int32)), mode
        for k in range(js["n"]):
            m = min(n[k], cap)
            assert np.array_equal(c[k, :m], rc[k, :m]), (mode, k)
        assert st["cells"] >= cells  # retries re-run only the retried jobs, as the oracle does
        assert st["cells"] == cells, (mode, st["cells"], cells)


def test_known_answers(engine):
    # AlignmentTests.swift:313-446
    sp = ScoringParameters()
    T = lambda q, t, w: GlobalSWTask(bytes(q), bytes(t), -(1 << 30), w, sp)
    res = GlobalSWDispatcher.dispatchBatch([
        T([0, 1, 2, 3, 0, 1, 2, 3, 0, 1], [0, 1, 2, 3, 0, 1, 2, 3, 0, 1], 10),
        T([0, 1, 2, 3, 0, 1, 2, 3, 0, 1], [0, 1, 2, 3, 0, 3, 2, 3, 0, 1], 10),
        T([0, 1, 2, 3, 2, 0, 1], [0, 1, 2, 3, 0, 1], 10),
        T([0, 1, 2, 3, 0, 1], [0, 1, 2, 3, 2, 0, 1], 10),
        T([0, 1, 2, 3] * 2, [0, 1, 2, 3] * 2, 1),
        T([0, 1, 2, 3] * 2, [0, 1, 2, 3] * 2, 50)], engine)
    assert res[0].score == 10 and res[0].cigar == [10 << 4]
    assert res[1].score == 5 and res[1].cigar == [10 << 4]
    ops = lambda cg: [(c & 0xF, c >> 4) for c in cg]
    q_ins = sum(l for o, l in ops(res[2].cigar) if o in (0, 1))
    t_ins = sum(l for o, l in ops(res[2].cigar) if o in (0, 2))
    assert any(o == 1 for o, _ in ops(res[2].cigar)) and (q_ins, t_ins) == (7, 6)
    q_d = sum(l for o, l in ops(res[3].cigar) if o in (0, 1))
    t_d = sum(l for o, l in ops(res[3].cigar) if o in (0, 2))
    assert any(o == 2 for o, _ in ops(res[3].cigar)) and (q_d, t_d) == (6, 7)
    assert res[4].score == res[5].score and res[4].cigar == res[5].cigar


def test_nw_subset(engine):
    check(engine, wl.generate(wl.NW, 3, 20000))


def test_narrow_bands_two_jobs_per_lane(engine):
    """Pairs whose bands are at most 64 columns: the default form fills them
    two jobs per lane (one int16 pair per column and row); every form equal
    to the oracle, with scorings whose E, F and H flags tie differently."""
    js = wl.generate(wl.NW, 5, 30000)
    q = js["q_len"].astype(int); t = js["t_len"].astype(int)
    wp = np.maximum(np.minimum(js["w"].astype(int), 1 << 20), np.abs(q - t))
    js = wl.slice_jobs(js, np.nonzero(np.minimum(q, 2 * wp + 1) <= 64)[0])
    check(engine, js)
    check(engine, js, ScoringParameters(gapOpenPenalty=0, gapExtendPenalty=2, gapOpenPenaltyDeletion=3,
                                        gapExtendPenaltyDeletion=1, mismatchPenalty=3))
    check(engine, js, ScoringParameters(matchScore=2, mismatchPenalty=2, gapOpenPenalty=1,
                                        gapExtendPenalty=1, gapOpenPenaltyDeletion=1,
                                        gapExtendPenaltyDeletion=1))


def test_forced_retries(engine):
    js = wl.generate(wl.NW, 7, 3000)
    js["w"] = np.maximum(1, js["w"] // 8).astype(np.int32)      # too narrow: retries fire
    js["true_score"] = js["true_score"] + 2
    check(engine, js)


def _random(rng, n, qmax, tmax, n_rate=0.0):
    qs, ts, ql, tl, w, tsc = [], [], [], [], [], []
    for _ in range(n):
        q = rng.integers(0, 4, rng.integers(1, qmax + 1)).astype(np.uint8)
        t = q.copy()
        t[rng.random(len(t)) < 0.05] = rng.integers(0, 4)
        cut = int(rng.integers(0, max(1, len(t) // 4)))
        t = np.concatenate([t[:len(t) // 2], rng.integers(0, 4, int(rng.integers(0, 6))).astype(np.uint8),
                            t[len(t) // 2 + cut:]])[:tmax]
        if len(t) == 0:
            t = np.zeros(1, np.uint8)
        if n_rate:
            q[rng.random(len(q)) < n_rate] = 4
            t[rng.random(len(t)) < n_rate] = 4
        qs.append(q); ts.append(t); ql.append(len(q)); tl.append(len(t))
        w.append(int(rng.choice([1, 2, 5, 20, 100])))
        tsc.append(int(rng.integers(-50, len(q))))
    qo = np.zeros(n, np.uint32); to = np.zeros(n, np.uint32)
    qo[1:] = np.cumsum(ql)[:-1]; to[1:] = np.cumsum(tl)[:-1]
    return {"n": n, "queries": np.concatenate(qs), "targets": np.concatenate(ts), "q_off": qo,
            "t_off": to, "q_len": np.array(ql, np.uint16), "t_len": np.array(tl, np.uint16),
            "w": np.array(w, np.int32), "true_score": np.array(tsc, np.int32)}


def test_random_with_n(engine):
    check(engine, _random(np.random.default_rng(21), 2000, 120, 140, n_rate=0.03))


def test_long_and_int32_tier(engine):
    # long pairs exceed the int16 bound and the LDS budget
    sp = ScoringParameters(gapOpenPenalty=20, gapExtendPenalty=3, gapOpenPenaltyDeletion=20,
                           gapExtendPenaltyDeletion=3, mismatchPenalty=9)
    check(engine, _random(np.random.default_rng(22), 200, 900, 1000), sp)


def test_int16_bound_edge(engine):
    """Pairs at the 16-bit tier's edge, (qlen + tlen) * max_pen + 64 <= 8000:
    with max_pen = 23 (o = 20, e = 3, mismatch 9), qlen + tlen = 345 is the
    largest 16-bit pair and 346 goes to int32.  Identical, random and indel
    pairs with bands of at most 64 columns, so the int16-pair form (form 3)
    fills them where the sentinel's drift and the score range are widest."""
    rng = np.random.default_rng(31)
    sp = ScoringParameters(gapOpenPenalty=20, gapExtendPenalty=3, gapOpenPenaltyDeletion=20,
                           gapExtendPenaltyDeletion=3, mismatchPenalty=9)
    qs, ts, w = [], [], []
    for total in (345, 346):
        for kind in range(4):
            for band in (1, 3, 10, 31):
                for dl in (0, 1, 5):
                    ql = (total - dl) // 2
                    tl = total - ql
                    q = rng.integers(0, 4, ql).astype(np.uint8)
                    if kind == 0:                                   # identical prefix
                        t = np.concatenate([q, rng.integers(0, 4, tl - ql)]).astype(np.uint8)
                    elif kind == 1:                                 # unrelated
                        t = rng.integers(0, 4, tl).astype(np.uint8)
                    elif kind == 2:                                 # an insertion in the middle
                        t = np.concatenate([q[:ql // 2], rng.integers(0, 4, tl - ql), q[ql // 2:]]).astype(np.uint8)
                    else:                                           # 5% substitutions, Ns
                        t = np.concatenate([q, rng.integers(0, 4, tl - ql)]).astype(np.uint8)
                        t[rng.random(tl) < 0.05] = rng.integers(0, 4)
                        q[rng.random(ql) < 0.02] = 4
                    qs.append(q); ts.append(t); w.append(band)
    n = len(qs)
    ql = np.array([len(x) for x in qs]); tl = np.array([len(x) for x in ts])
    qo = np.zeros(n, np.uint32); to = np.zeros(n, np.uint32)
    qo[1:] = np.cumsum(ql)[:-1]; to[1:] = np.cumsum(tl)[:-1]
    js = {"n": n, "queries": np.concatenate(qs), "targets": np.concatenate(ts), "q_off": qo,
          "t_off": to, "q_len": ql.astype(np.uint16), "t_len": tl.astype(np.uint16),
          "w": np.array(w, np.int32), "true_score": np.full(n, -(1 << 20), np.int32)}
    check(engine, js, sp)
    # many copies: full waves in every form, both halves of the pairs used
    rep = 64
    big = {k: (np.tile(v, rep) if isinstance(v, np.ndarray) else v) for k, v in js.items()}
    big["n"] = n * rep
    big["q_off"] = (qo[None, :].astype(np.int64) + len(js["queries"]) * np.arange(rep)[:, None]).ravel().astype(np.uint32)
    big["t_off"] = (to[None, :].astype(np.int64) + len(js["targets"]) * np.arange(rep)[:, None]).ravel().astype(np.uint32)
    check(engine, big, sp, forms=(0, 2, 3))


def test_small_cap_flags(engine):
    check(engine, _random(np.random.default_rng(23), 500, 60, 70), cap=2)


def _edges(rng):
    """Pairs whose query lengths sit at the 64-column chunk edges of the
    wave-per-job form (64, 65, 128, 129, 256, 257, 512, 513), with bands from
    1 to wider than the query, targets shorter and longer than the query, and
    N bases; few jobs per bin, so the batched form runs them one wave each."""
    qs, ts, ql, tl, w, tsc = [], [], [], [], [], []
    for q_len in (1, 2, 63, 64, 65, 127, 128, 129, 255, 256, 257, 511, 512, 513):
        for t_off in (-30, -1, 0, 1, 40):
            t_len = max(1, q_len + t_off)
            for band in (1, 7, 64, 300):
                q = rng.integers(0, 4, q_len).astype(np.uint8)
                t = np.resize(q, t_len).copy()
                t[rng.random(t_len) < 0.05] = rng.integers(0, 4)
                if rng.random() < 0.3:
                    q[rng.random(q_len) < 0.02] = 4
                    t[rng.random(t_len) < 0.02] = 4
                qs.append(q); ts.append(t); ql.append(q_len); tl.append(t_len)
                w.append(band); tsc.append(int(rng.integers(-20, q_len + 1)))
    n = len(ql)
    qo = np.zeros(n, np.uint32); to = np.zeros(n, np.uint32)
    qo[1:] = np.cumsum(ql)[:-1]; to[1:] = np.cumsum(tl)[:-1]
    return {"n": n, "queries": np.concatenate(qs), "targets": np.concatenate(ts), "q_off": qo,
            "t_off": to, "q_len": np.array(ql, np.uint16), "t_len": np.array(tl, np.uint16),
            "w": np.array(w, np.int32), "true_score": np.array(tsc, np.int32)}


def _band_edges(rng):
    """Bands of 1..130 columns around the register-state instances' limits
    (32, 64, 96, 128 columns, and the 8-column block edges): odd widths
    2w+1 on 150-base queries, even and odd widths set by short queries under
    wide bands, targets shorter and longer than the query (w' = max(w, |qlen -
    tlen|)), N bases in a few; one bin holds several widths, so lanes of one
    wave have different bands, band ramps and band ends."""
    qs, ts, ql, tl, w, tsc = [], [], [], [], [], []
    widths = (1, 2, 7, 8, 9, 15, 16, 17, 30, 31, 32, 33, 34, 62, 63, 64, 65, 66, 94, 95, 96, 97,
              98, 126, 127, 128, 129, 130)
    for c in widths:
        for kind in range(6):
            if kind < 3:                     # ncol = 2w + 1 (or qlen when shorter)
                q_len, band = 150, max(1, (c - 1) // 2)
                t_len = q_len + (0, -min(band, 20), min(band, 20))[kind]
            else:                            # ncol = qlen
                q_len, band = c, c + 3
                t_len = max(1, q_len + (0, -5, 9)[kind - 3])
            q = rng.integers(0, 4, q_len).astype(np.uint8)
            t = np.resize(q, t_len).copy()
            t[rng.random(t_len) < 0.06] = rng.integers(0, 4)
            if rng.random() < 0.2:
                q[rng.random(q_len) < 0.02] = 4
                t[rng.random(t_len) < 0.02] = 4
            qs.append(q); ts.append(t); ql.append(q_len); tl.append(t_len)
            w.append(band); tsc.append(int(rng.integers(-20, q_len + 1)))
    n = len(ql)
    qo = np.zeros(n, np.uint32); to = np.zeros(n, np.uint32)
    qo[1:] = np.cumsum(ql)[:-1]; to[1:] = np.cumsum(tl)[:-1]
    return {"n": n, "queries": np.concatenate(qs), "targets": np.concatenate(ts), "q_off": qo,
            "t_off": to, "q_len": np.array(ql, np.uint16), "t_len": np.array(tl, np.uint16),
            "w": np.array(w, np.int32), "true_score": np.array(tsc, np.int32)}


def test_register_state_band_edges(engine):
    js = _band_edges(np.random.default_rng(26))
    check(engine, js)
    # the same pairs many times over: full bins, so every register-state
    # instance runs whole waves of mixed widths in the batched form
    rep = 40
    big = {"n": js["n"] * rep,
           "queries": np.tile(js["queries"], rep), "targets": np.tile(js["targets"], rep),
           "q_off": (js["q_off"][None, :].astype(np.int64) +
                     len(js["queries"]) * np.arange(rep)[:, None]).ravel().astype(np.uint32),
           "t_off": (js["t_off"][None, :].astype(np.int64) +
                     len(js["targets"]) * np.arange(rep)[:, None]).ravel().astype(np.uint32),
           "q_len": np.tile(js["q_len"], rep), "t_len": np.tile(js["t_len"], rep),
           "w": np.tile(js["w"], rep), "true_score": np.tile(js["true_score"], rep)}
    check(engine, big, forms=(0, 2, 3))


def test_wave_per_job_edges(engine):
    check(engine, _edges(np.random.default_rng(24)))
    sp = ScoringParameters(gapOpenPenalty=0, gapExtendPenalty=2, gapOpenPenaltyDeletion=3,
                           gapExtendPenaltyDeletion=1, mismatchPenalty=3)
    check(engine, _edges(np.random.default_rng(25)), sp)


def test_pinned_tie_vectors(engine):
    """Equal-score path choices pinned in tests/oracle_pins.py: the lambda
    fixture's deletion inside AAC AAC (leftmost, 20M3D20M), homopolymer
    deletion / insertion (1D4M / 1I4M), and binary pairs with many equal
    paths under cheap gaps -- each against the oracle, op for op."""
    import itertools
    from tests import oracle_pins as P
    from tests.ref_restatements import global_tb_full
    seq = "".join(l.strip() for l in open(P.os.path.join(P.HERE, "golden", "lambda", "lambda.fa"))
                  if not l.startswith(">"))
    enc = {"A": 0, "C": 1, "G": 2, "T": 3}
    ref = [enc[c] for c in seq[30030:30073]]
    read = ref[:20] + ref[23:]
    sp = ScoringParameters()
    T = lambda q, t, w, s=sp: GlobalSWTask(bytes(q), bytes(t), -(1 << 30), w, s)
    res = GlobalSWDispatcher.dispatchBatch([T(read, ref, 10), T([0, 0, 0, 1], [0, 0, 0, 0, 1], 3),
                                            T([0, 0, 0, 0, 1], [0, 0, 0, 1], 3)], engine)
    assert res[0].cigar == [20 << 4, 3 << 4 | 2, 20 << 4] and res[0].score == 31
    assert res[1].cigar == [1 << 4 | 2, 4 << 4]
    assert res[2].cigar == [1 << 4 | 1, 4 << 4]
    cheap = ScoringParameters(mismatchPenalty=1, gapOpenPenalty=1, gapExtendPenalty=1,
                              gapOpenPenaltyDeletion=1, gapExtendPenaltyDeletion=1)
    osc = oracle_sc(cheap)
    seqs = [list(s) for n in range(1, 5) for s in itertools.product((0, 1), repeat=n)]
    pairs = [(q, t) for q in seqs for t in seqs]
    res = GlobalSWDispatcher.dispatchBatch([T(q, t, 2, cheap) for q, t in pairs], engine)
    assert len(res) == len(pairs)
    for (q, t), r in zip(pairs, res):
        want = global_tb_full(q, t, list(osc.mat), 1, 1, 1, 1, 2)
        assert (r.score, r.cigar) == want, (q, t, r, want)


def test_full_size_500k_every_row(engine):
    """The whole 500,000-pair NW set (the benchmarked one): every pair's
    score, op count, CIGAR ops and the cell count against the oracle, in the
    batched and the plain form."""
    import os
    js = wl.generate(wl.NW, 3, 500000)
    sp = ScoringParameters()
    cap = 512
    rs, rc, rn, cells = orc.global_batch(js, oracle_sc(sp), max_retries=3, cap=cap,
                                         nthreads=min(16, os.cpu_count() or 1))
    used = np.arange(cap)[None, :] < np.minimum(rn, cap)[:, None]
    for mode in MODES:
        print("NW 500k mode %d" % mode, flush=True)
        st = {}
        s, c, n, fl, _ = engine.global_nw(js, sp, max_retries=3, cap=cap, mode=mode, stats=st)
        assert np.array_equal(s, rs), mode
        assert np.array_equal(n, rn), mode
        assert not ((c != rc) & used).any(), mode
        assert st["cells"] == cells, (mode, st["cells"], cells)
