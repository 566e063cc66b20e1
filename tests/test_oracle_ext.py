"""Pins of the extension oracle (or_ksw_extend = BandedSWScalar.align)."""
import itertools

import numpy as np
import pytest

from oracle import pyoracle as orc
from tests.ref_restatements import ksw_extend2_bwa

SC = orc.scoring()
MAT = list(SC.mat)


def ext(q, t, w=100, h0=1, eb=0, sc=SC):
    return orc.ksw_extend(q, t, sc=sc, w=w, h0=h0, end_bonus=eb)[0]


# ---- reference known answers (Tests/AlignmentTests/AlignmentTests.swift) ----
def test_perfect_match_known_answer():        # :48-63
    assert ext([0, 1, 2, 3, 0, 1], [0, 1, 2, 3, 0, 1], w=10, h0=1)[0] == 7


def test_h0_known_answer():                   # :82-96 (asserts >= 8; closed form 8)
    assert ext([0, 1, 2], [0, 1, 2], w=10, h0=5)[0] == 8


def test_acgt_x2_known_answer():              # :115-136
    assert ext([0, 1, 2, 3] * 2, [0, 1, 2, 3] * 2, w=10, h0=1)[0] == 9


def test_global_score_positive():             # :140-162
    r = ext([0, 1, 2, 3, 0, 1], [0, 1, 2, 3, 0, 1], w=10, h0=1)
    assert r[4] > 0


def test_single_mismatch_closed_form():       # :208-237: 1 + 4 - 4 + 5 = 6
    r = ext([0, 1, 2, 3, 0, 1, 2, 3, 0, 1], [0, 1, 2, 3, 3, 1, 2, 3, 0, 1], w=10, h0=1)
    assert r[0] == 6


def test_zdrop_closed_form():                 # :239-283 (closed form applied to the scalar path)
    sc = orc.scoring(zdrop=10)
    r = ext([0, 1, 2, 3, 0] + [0] * 100, [0, 1, 2, 3, 0] + [3] * 100, w=5, h0=1, sc=sc)
    assert r[0] == 6 and r[2] < 50


# ---- closed forms ----
@pytest.mark.parametrize("L,h0", [(1, 1), (7, 3), (40, 25), (150, 60)])
def test_perfect_match_closed_form(L, h0):
    rng = np.random.default_rng(L)
    s = rng.integers(0, 4, L)
    r = ext(s, s, w=100, h0=h0)
    assert r[0] == h0 + L and r[1] == L and r[2] == L
    assert r[4] == h0 + L and r[3] == L and r[5] == 0


def test_all_mismatch_closed_form():
    r = ext([0] * 20, [3] * 20, w=100, h0=10)
    assert r[0] == 10 and r[1] == 0 and r[2] == 0


def test_empty_inputs():                      # BandedSWScalar.swift:33-35
    assert ext([], [0, 1]) == (0, 0, 0, 0, 0, 0)
    assert ext([0, 1], []) == (0, 0, 0, 0, 0, 0)


@pytest.mark.parametrize("L,eb", [(5, 0), (12, 0), (12, 5)])
def test_cell_count_closed_form(L, eb):
    # perfect match with a large h0: no H/E value reaches 0, so the band is the
    # static band of the capped w (BandedSWScalar.swift:52-62):
    # w = min(100, max(1, int((L*A + eb - O)/E + 1))), row i spans
    # [max(0, i-w), min(L, i+w+1)).
    s = [0, 1, 2, 3] * 4
    s = s[:L]
    r, cells = orc.ksw_extend(s, s, w=100, h0=200, end_bonus=eb)
    w = min(100, max(1, int((L + eb - 6) / 1 + 1)))
    assert cells == sum(min(L, i + w + 1) - max(0, i - w) for i in range(L))


# ---- independent restatement of published bwa ksw_extend2 ----
def _cmp(q, t, w, h0, eb, sc):
    a = ext(q, t, w=w, h0=h0, eb=eb, sc=sc)
    b = ksw_extend2_bwa(list(q), list(t), list(sc.mat), sc.o_del, sc.e_del, sc.o_ins,
                        sc.e_ins, w, eb, sc.zdrop, h0)
    assert a == b, (q, t, w, h0, eb, a, b)


def test_exhaustive_tiny_vs_bwa_form():
    alphabet = [0, 1, 3]
    sc = orc.scoring()
    for lq in range(1, 4):
        for lt in range(1, 4):
            for q in itertools.product(alphabet, repeat=lq):
                for t in itertools.product(alphabet, repeat=lt):
                    for h0 in (1, 8):
                        _cmp(q, t, 100, h0, 0, sc)


@pytest.mark.parametrize("seed", range(6))
def test_random_vs_bwa_form(seed):
    rng = np.random.default_rng(100 + seed)
    scs = [orc.scoring(), orc.scoring(zdrop=5), orc.scoring(a=2, b=3, o_ins=3, e_ins=2, o_del=5,
                                                           e_del=1, zdrop=20)]
    for _ in range(150):
        lq, lt = int(rng.integers(1, 60)), int(rng.integers(1, 80))
        q = rng.integers(0, 5 if rng.random() < 0.2 else 4, lq)
        t = rng.integers(0, 4, lt)
        if rng.random() < 0.6:
            m = min(lq, lt)
            t[:m] = np.where(rng.random(m) < 0.1, rng.integers(0, 4, m), np.minimum(q[:m], 3))
        w = int(rng.choice([1, 2, 5, 10, 30, 100]))
        h0 = int(rng.integers(1, 60))
        eb = int(rng.integers(0, 6))
        _cmp(q, t, w, h0, eb, scs[int(rng.integers(0, 3))])


def test_invariants_on_generated_jobs():
    from swiftbwa_amd import workload as wl
    js = wl.generate(wl.EXT_MIXED, 4, 3000)
    out, _ = orc.ksw_extend_batch(js, nthreads=4)
    ql, tl = js["q_len"].astype(int), js["t_len"].astype(int)
    nz = (ql > 0) & (tl > 0)
    assert (out[nz, 0] >= js["h0"][nz]).all()
    assert ((out[:, 1] >= 0) & (out[:, 1] <= ql)).all()
    assert ((out[:, 2] >= 0) & (out[:, 2] <= tl)).all()
    assert ((out[nz, 4] == -1) == (out[nz, 3] == 0)).all()
    assert (out[~nz] == 0).all()
