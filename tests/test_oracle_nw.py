"""Pins of the NW oracle (GlobalAligner.align + traceback) and of the CIGAR
generator oracle (CIGARGenerator.generate)."""
import numpy as np
import pytest

from oracle import pyoracle as orc
from tests.ref_restatements import gotoh_global, rescore_cigar

SC = orc.scoring()
MAT = list(SC.mat)


def ops(cig):
    return [(c & 0xF, c >> 4) for c in cig]


# ---- reference known answers (AlignmentTests.swift) ----
def test_perfect_match():                                   # :313-329
    s, cig, _ = orc.ksw_global([0, 1, 2, 3, 0, 1, 2, 3, 0, 1], [0, 1, 2, 3, 0, 1, 2, 3, 0, 1], w=10)
    assert s == 10 and cig == [10 << 4]


def test_single_mismatch():                                 # :331-354
    q, t = [0, 1, 2, 3, 0, 1, 2, 3, 0, 1], [0, 1, 2, 3, 0, 3, 2, 3, 0, 1]
    s, cig, _ = orc.ksw_global(q, t, w=10)
    assert s == 5 and cig == [10 << 4]
    r = orc.cigar_generate(q, t, 0, 10, 10, False, 5, 10)
    assert r["nm"] == 1


def test_insertion_and_deletion():                          # :356-424
    _, cig, _ = orc.ksw_global([0, 1, 2, 3, 2, 0, 1], [0, 1, 2, 3, 0, 1], w=10)
    o = ops(cig)
    assert any(x == 1 for x, _ in o)
    assert sum(l for x, l in o if x in (0, 1)) == 7 and sum(l for x, l in o if x in (0, 2)) == 6
    _, cig, _ = orc.ksw_global([0, 1, 2, 3, 0, 1], [0, 1, 2, 3, 2, 0, 1], w=10)
    o = ops(cig)
    assert any(x == 2 for x, _ in o)
    assert sum(l for x, l in o if x in (0, 1)) == 6 and sum(l for x, l in o if x in (0, 2)) == 7


def test_narrow_equals_wide():                              # :426-446
    a = orc.ksw_global([0, 1, 2, 3] * 2, [0, 1, 2, 3] * 2, w=1)
    b = orc.ksw_global([0, 1, 2, 3] * 2, [0, 1, 2, 3] * 2, w=50)
    assert a[0] == b[0] and a[1] == b[1]


def test_generator_known_answers():                         # :450-688
    r = orc.cigar_generate([0, 1, 2, 3] * 2, [0, 1, 2, 3] * 2, 0, 8, 8, False, 8, 10)
    assert all((c & 0xF) != 4 for c in r["cigar"]) and r["nm"] == 0 and r["md"] == "8"
    r = orc.cigar_generate([0, 1, 2, 3, 0], [0, 1, 2, 3, 0], 3, 8, 12, False, 5, 10, ref_pos=100)
    assert ops(r["cigar"]) == [(4, 3), (0, 5), (4, 4)]
    r = orc.cigar_generate([0, 1, 2, 3, 0], [0, 1, 2, 3, 0], 3, 8, 12, True, 5, 10, ref_pos=100)
    assert [l for _, l in ops(r["cigar"])] == [4, 5, 3]
    r = orc.cigar_generate([0, 1, 2, 3, 2, 2, 0, 1, 2, 3], [0, 1, 2, 3, 0, 1, 2, 3], 0, 10, 10,
                           False, 0, 1)
    o = ops(r["cigar"])
    assert sum(l for x, l in o if x in (0, 1)) == 10 and sum(l for x, l in o if x in (0, 2)) == 8
    r = orc.cigar_generate([0, 1, 2, 3] * 2, [0, 1, 2, 3, 3, 1, 2, 3], 0, 8, 8, False, 3, 10)
    assert r["md"] == "4T3" and r["nm"] == 1
    r = orc.cigar_generate([0, 1, 2, 3, 0, 1], [0, 1, 2, 3, 2, 0, 1], 0, 6, 6, False, -1, 10)
    assert "^" in r["md"]
    r = orc.cigar_generate([0, 1, 2, 3, 2, 0, 1], [0, 1, 2, 3, 0, 1], 0, 7, 7, False, -1, 10)
    assert "^" not in r["md"]
    r = orc.cigar_generate([0, 1, 2, 3, 0, 1, 2, 3, 0, 1], [0, 3, 2, 3, 0, 1, 2, 0, 0, 1], 0, 10,
                           10, False, 0, 10)
    assert r["md"] == "1T5A2" and r["nm"] == 2


# ---- brute force: full-matrix Gotoh score when the band cannot bind ----
@pytest.mark.parametrize("seed", range(4))
def test_unbinding_band_equals_gotoh(seed):
    rng = np.random.default_rng(300 + seed)
    scs = [orc.scoring(), orc.scoring(a=2, b=3, o_ins=4, e_ins=2, o_del=5, e_del=1)]
    for _ in range(80):
        lq, lt = int(rng.integers(1, 25)), int(rng.integers(1, 25))
        q = rng.integers(0, 4, lq)
        t = rng.integers(0, 4, lt)
        if rng.random() < 0.5:
            m = min(lq, lt)
            t[:m] = q[:m]
        sc = scs[int(rng.integers(0, 2))]
        w = max(lq, lt) + 1
        s, cig, _ = orc.ksw_global(q, t, sc=sc, w=w)
        g = gotoh_global(list(q), list(t), list(sc.mat), sc.o_del, sc.e_del, sc.o_ins, sc.e_ins)
        assert s == g
        # the traceback path consumes both sequences and re-scores to the score
        rs, qi, ti = rescore_cigar(cig, list(q), list(t), list(sc.mat), sc.o_del, sc.e_del,
                                   sc.o_ins, sc.e_ins)
        assert (qi, ti) == (lq, lt)
        assert rs == s


def test_band_binds_lowers_or_keeps_score():
    rng = np.random.default_rng(9)
    for _ in range(100):
        q = rng.integers(0, 4, int(rng.integers(5, 40)))
        t = np.concatenate([q[:10], rng.integers(0, 4, 6), q[10:]])
        s_narrow = orc.ksw_global(q, t, w=1)[0]
        s_wide = orc.ksw_global(q, t, w=100)[0]
        assert s_narrow <= s_wide


def test_retry_doubles_until_true_score():
    # deficit needs a band wider than 1: retries double w (1,2,4,8)
    q = [0, 1, 2, 3] * 5
    t = q[:8] + [3, 3, 3, 3, 3] + q[8:]
    s, cig, fw, _ = orc.global_with_retry(q, t, 1, 10 ** 6, max_retries=3)
    assert fw == 8
    s0 = orc.global_with_retry(q, t, 1, -10 ** 6, max_retries=3)
    assert s0[2] == 1
