"""GPU parity of the mate-rescue local-SW kernels against the CPU oracle
(LocalSWAligner.align restated in oracle/sw_oracle.c): all six fields and the
nil/non-nil outcome, bit-exact."""
import numpy as np
import pytest

from oracle import pyoracle as orc
from swiftbwa_amd import workload as wl
from swiftbwa_amd.metalsw import LocalSWDispatcher, LocalSWTask, ScoringParameters

pytestmark = pytest.mark.gpu


def oracle_sc(sp):
    return orc.scoring(a=sp.matchScore, b=sp.mismatchPenalty, o_ins=sp.gapOpenPenalty,
                       e_ins=sp.gapExtendPenalty, o_del=sp.gapOpenPenaltyDeletion,
                       e_del=sp.gapExtendPenaltyDeletion)


def check(engine, js, sp=None, min_sub=0):
    sp = sp or ScoringParameters()
    ref, rok, _ = orc.local_batch(js, oracle_sc(sp), min_sub_score=min_sub, nthreads=8)
    ref[rok == 0] = 0
    got, ok = engine.local_sw(js, sp, min_sub_score=min_sub)
    assert np.array_equal(ok, rok), np.nonzero(ok != rok)[0][:5]
    bad = np.nonzero((got != ref).any(axis=1))[0]
    assert len(bad) == 0, "%d mismatches, first %d: gpu %s oracle %s" % (
        len(bad), bad[0], got[bad[0]].tolist(), ref[bad[0]].tolist())


def test_known_answers(engine):
    # AlignmentTests.swift:715-780
    sp = ScoringParameters()
    T = lambda q, t: LocalSWTask(bytes(q), bytes(t), sp)
    r = LocalSWDispatcher.dispatchBatch([
        T([0, 1, 2, 3, 0, 1], [0, 1, 2, 3, 0, 1]), T([0, 1, 2, 3], [3, 3, 0, 1, 2, 3, 3, 3]),
        T([0, 0, 0, 0], [3, 3, 3, 3]), T([], [0, 1, 2])], engine)
    assert (r[0].score, r[0].queryBegin, r[0].queryEnd, r[0].targetBegin, r[0].targetEnd) == (6, 0, 5, 0, 5)
    assert (r[1].score, r[1].queryBegin, r[1].queryEnd, r[1].targetBegin, r[1].targetEnd) == (4, 0, 3, 2, 5)
    assert r[2] is None and r[3] is None


def test_rescue_windows(engine):
    check(engine, wl.generate(wl.LOCAL, 5, 20000), min_sub=19)


def test_cell_count(engine):
    """hipsw_local_sw counts the cells the reference's loops visit (forward
    pass, scalar redo, reverse pass), equal to the oracle's count on tasks
    that do not overflow the striped pass."""
    js = wl.generate(wl.LOCAL, 7, 4000)
    _, _, cells = orc.local_batch(js, oracle_sc(ScoringParameters()), min_sub_score=19, nthreads=8)
    st = {}
    engine.local_sw(js, ScoringParameters(), min_sub_score=19, stats=st)
    assert st["cells"] == cells


def test_rescue_windows_no_score2(engine):
    check(engine, wl.generate(wl.LOCAL, 6, 5000), min_sub=0)


def _random(rng, n, qmax, tmax, n_rate=0.0):
    qs, ts = [], []
    for _ in range(n):
        q = rng.integers(0, 4, rng.integers(0, qmax + 1)).astype(np.uint8)
        t = rng.integers(0, 4, rng.integers(0, tmax + 1)).astype(np.uint8)
        if len(q) > 4 and len(t) > len(q) and rng.random() < 0.7:
            p = int(rng.integers(0, len(t) - len(q) + 1))
            t[p:p + len(q)] = q
            if rng.random() < 0.3 and len(t) >= 2 * len(q) + p:           # tandem copy
                p2 = int(rng.integers(p + len(q), len(t) - len(q) + 1))
                t[p2:p2 + len(q)] = q
            flip = rng.random(len(t)) < 0.03
            t[flip] = rng.integers(0, 4, int(flip.sum()))
        if n_rate:
            q[rng.random(len(q)) < n_rate] = 4
            t[rng.random(len(t)) < n_rate] = 4
        qs.append(q); ts.append(t)
    ql = np.array([len(x) for x in qs], np.uint16); tl = np.array([len(x) for x in ts], np.uint16)
    qo = np.zeros(n, np.uint32); to = np.zeros(n, np.uint32)
    qo[1:] = np.cumsum(ql.astype(np.int64))[:-1]; to[1:] = np.cumsum(tl.astype(np.int64))[:-1]
    cat = lambda xs: np.concatenate(xs) if sum(map(len, xs)) else np.zeros(1, np.uint8)
    return {"n": n, "queries": cat(qs), "targets": cat(ts), "q_off": qo, "t_off": to,
            "q_len": ql, "t_len": tl}


def test_random_small_with_n(engine):
    check(engine, _random(np.random.default_rng(31), 4000, 60, 120, n_rate=0.03), min_sub=10)


def test_long_queries_scalar_path(engine):
    # qLen*A > 250 -> scalar int32 pass (LocalSWAligner.swift:56-78)
    check(engine, _random(np.random.default_rng(32), 300, 400, 700), min_sub=19)


def test_scoring_variants(engine):
    js = _random(np.random.default_rng(33), 2000, 150, 300)
    for sp in (ScoringParameters(matchScore=2, mismatchPenalty=3),
               ScoringParameters(gapOpenPenalty=3, gapExtendPenalty=2),
               ScoringParameters(matchScore=1, mismatchPenalty=1, gapOpenPenalty=1,
                                 gapExtendPenalty=1),
               # free gaps: padding cells can hold a row's maximum, so the
               # kernel's direct best-cell search falls back to the stripe walk
               ScoringParameters(gapOpenPenalty=0, gapExtendPenalty=0)):
        check(engine, js, sp, min_sub=15)


def test_all_stripe_counts(engine):
    # queries of 1..250 bases: every stripe count 1..16, small groups merged
    # into larger ones (padded waves) and full groups (exact waves)
    check(engine, _random(np.random.default_rng(35), 6000, 250, 320), min_sub=15)



def test_pinned_vectors(engine):
    """The hand-derived vectors that pin the oracle where the reference's
    striped pass differs from textbook SW (tests/oracle_pins.py): one-pass
    lazy-F, its stop on a tie, stripe-then-lane ties, a padding-lane
    maximum under free gaps, and the score2 exclusion radius."""
    from tests import oracle_pins as P
    sp = ScoringParameters()
    free = ScoringParameters(gapOpenPenalty=0, gapExtendPenalty=0, gapOpenPenaltyDeletion=0,
                             gapExtendPenaltyDeletion=0)
    q22 = [0, 1, 1, 1, 1, 0, 0, 0, 1, 1, 1, 0, 0, 1, 0, 0, 1, 0, 1, 1, 1, 0]
    t20 = [0, 1, 1, 1, 1, 0, 0, 0, 1, 1, 0, 1, 0, 0, 1, 0, 1, 1, 1, 0]
    tie = [1] * 20
    tie[3] = tie[4] = 0
    cases = [(P.X20 + [3] * 4 + P.Y20, [3] * 5 + P.X20 + P.Y20 + [3] * 5, sp, 0, 29),
             (P.X20 + [3] * 5 + P.Y20, [3] * 5 + P.X20 + P.Y20 + [3] * 5, sp, 0, 23),
             (q22, t20, sp, 0, (10, 0, 9, 0, 9, 0)),
             (tie, [0], sp, 0, (1, 4, 4, 0, 0, 0)),
             ([1] * 17 + [0], [1, 0], free, 0, (2, 0, 0, 0, 1, 0))]
    for g, s2 in ((5, 0), (6, 25)):
        cases.append((P.Q30, [3] * 12 + P.Q30 + [3] * g + P.Q30[:25] + [3] * 12, sp, 19,
                      (30, 0, 29, 12, 41, s2)))
    for q, t, sc, ms, want in cases:
        r = LocalSWDispatcher.dispatchBatch([LocalSWTask(bytes(q), bytes(t), sc)], engine,
                                            minSubScore=ms)[0]
        got = (r.score, r.queryBegin, r.queryEnd, r.targetBegin, r.targetEnd, r.score2)
        if isinstance(want, int):
            assert got[0] == want, (q, t, got)
        else:
            assert got == want, (q, t, got, want)
