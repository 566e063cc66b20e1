"""Pins of the local-SW oracle (LocalSWAligner.align)."""
import numpy as np
import pytest

from oracle import pyoracle as orc
from tests.ref_restatements import local_affine_brute

SC = orc.scoring()


def test_known_answers():                                   # AlignmentTests.swift:715-780
    assert orc.local_sw([0, 1, 2, 3, 0, 1], [0, 1, 2, 3, 0, 1])[:5] == (6, 0, 5, 0, 5)
    assert orc.local_sw([0, 1, 2, 3], [3, 3, 0, 1, 2, 3, 3, 3])[:5] == (4, 0, 3, 2, 5)
    assert orc.local_sw([0, 1, 2, 3] * 2, [0, 1, 2, 3, 3, 1, 2, 3])[0] >= 4
    assert orc.local_sw([0, 0, 0, 0], [3, 3, 3, 3]) is None
    assert orc.local_sw([], [0, 1, 2]) is None


@pytest.mark.parametrize("seed", range(4))
def test_score_vs_brute_force(seed):
    # textbook local SW with the striped pass' gap model; the single-pass lazy-F
    # of the striped code can only miss an F update, never invent one, so its
    # score is <= the brute force and equal on the vast majority of inputs.
    rng = np.random.default_rng(500 + seed)
    eq = tot = 0
    for _ in range(120):
        lq, lt = int(rng.integers(1, 70)), int(rng.integers(1, 90))
        q = rng.integers(0, 4, lq)
        t = rng.integers(0, 4, lt)
        if rng.random() < 0.6 and lt > 10:
            p = int(rng.integers(0, lt - 5))
            m = min(lq, lt - p)
            t[p:p + m] = q[:m]
        r = orc.local_sw(q, t)
        b = local_affine_brute(list(q), list(t), list(SC.mat), SC.o_ins, SC.e_ins)
        s = 0 if r is None else r[0]
        assert s <= b
        eq += s == b
        tot += 1
    assert eq >= 0.9 * tot


def test_start_end_rescoring():
    # the reported [qb,qe] x [tb,te] box contains a local alignment of the score
    rng = np.random.default_rng(7)
    for _ in range(60):
        q = rng.integers(0, 4, 30)
        t = np.concatenate([rng.integers(0, 4, 20), q, rng.integers(0, 4, 20)])
        t[25] = (t[25] + 1) % 4
        r = orc.local_sw(q, t)
        assert r is not None
        s, qb, qe, tb, te, _ = r
        sub = orc.local_sw(q[qb:qe + 1], t[tb:te + 1])
        assert sub is not None and sub[0] == s


def test_scalar_fallback_used_for_long_queries():
    # qlen * A > 250 takes the scalar path (LocalSWAligner.swift:56-78)
    rng = np.random.default_rng(8)
    q = rng.integers(0, 4, 300)
    t = np.concatenate([rng.integers(0, 4, 50), q, rng.integers(0, 4, 50)])
    r = orc.local_sw(q, t)
    assert r[0] == 300 and (r[1], r[2], r[3], r[4]) == (0, 299, 50, 349)


def test_score2_tandem():
    # two copies of the query in the window: score2 reports the second peak
    rng = np.random.default_rng(10)
    q = rng.integers(0, 4, 40)
    t = np.concatenate([rng.integers(0, 4, 30), q, rng.integers(0, 4, 60), q, rng.integers(0, 4, 30)])
    r = orc.local_sw(q, t, min_sub_score=19)
    assert r[0] == 40 and r[5] == 40
