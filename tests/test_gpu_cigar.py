"""GPU parity of hipsw_cigar_generate (CIGARGenerator.generate for a batch of
regions: band choice, fast path, NW with retries on the device, squeeze, NM,
MD, soft clips) against the oracle's restatement.  Bit-exact on every field."""
import numpy as np
import pytest

from oracle import pyoracle as orc
from swiftbwa_amd import workload as wl
from swiftbwa_amd.metalsw import CIGARGenerator, CIGARRegion, ScoringParameters

pytestmark = pytest.mark.gpu


def oracle_sc(sp):
    return orc.scoring(a=sp.matchScore, b=sp.mismatchPenalty, o_ins=sp.gapOpenPenalty,
                       e_ins=sp.gapExtendPenalty, o_del=sp.gapOpenPenaltyDeletion,
                       e_del=sp.gapExtendPenaltyDeletion, zdrop=sp.zDrop)


def check(engine, regions, sp=None):
    sp = sp or ScoringParameters()
    st = {}
    got = CIGARGenerator.generateBatch(regions, sp, engine, stats=st)
    assert len(got) == len(regions)
    osc = oracle_sc(sp)
    for k, r in enumerate(regions):
        ref = orc.cigar_generate(list(r.querySegment), list(r.refSegment), r.qb, r.qe,
                                 r.readLength, r.isReverse, r.trueScore, r.initialW, osc,
                                 ref_pos=r.refPos)
        g = got[k]
        assert g is not None, k
        assert (g.cigar, g.nm, g.md, g.score, g.pos) == (
            ref["cigar"], ref["nm"], ref["md"], ref["score"], ref["pos"]), (k, g, ref)
    return st


def R(q, t, qb, qe, rl, rev, ts, w, pos=0):
    return CIGARRegion(bytes(q), bytes(t), qb, qe, rl, rev, ts, w, pos)


def test_known_answers(engine):
    # AlignmentTests.swift:450-688 (the oracle's own test pins the same values)
    regions = [
        R([0, 1, 2, 3] * 2, [0, 1, 2, 3] * 2, 0, 8, 8, False, 8, 10),
        R([0, 1, 2, 3, 0], [0, 1, 2, 3, 0], 3, 8, 12, False, 5, 10, 100),
        R([0, 1, 2, 3, 0], [0, 1, 2, 3, 0], 3, 8, 12, True, 5, 10, 100),
        R([0, 1, 2, 3, 2, 2, 0, 1, 2, 3], [0, 1, 2, 3, 0, 1, 2, 3], 0, 10, 10, False, 0, 1),
        R([0, 1, 2, 3] * 2, [0, 1, 2, 3, 3, 1, 2, 3], 0, 8, 8, False, 3, 10),
        R([0, 1, 2, 3, 0, 1], [0, 1, 2, 3, 2, 0, 1], 0, 6, 6, False, -1, 10),
        R([0, 1, 2, 3, 2, 0, 1], [0, 1, 2, 3, 0, 1], 0, 7, 7, False, -1, 10),
        R([0, 1, 2, 3, 0, 1, 2, 3, 0, 1], [0, 3, 2, 3, 0, 1, 2, 0, 0, 1], 0, 10, 10, False, 0,
          10),
    ]
    sp = ScoringParameters()
    got = CIGARGenerator.generateBatch(regions, sp, engine)
    ops = lambda cg: [(c & 0xF, c >> 4) for c in cg]
    assert got[0].nm == 0 and got[0].md == "8" and all(o != 4 for o, _ in ops(got[0].cigar))
    assert ops(got[1].cigar) == [(4, 3), (0, 5), (4, 4)]
    assert [l for _, l in ops(got[2].cigar)] == [4, 5, 3]
    o = ops(got[3].cigar)
    assert sum(l for x, l in o if x in (0, 1)) == 10 and sum(l for x, l in o if x in (0, 2)) == 8
    assert got[4].md == "4T3" and got[4].nm == 1
    assert "^" in got[5].md and "^" not in got[6].md
    assert got[7].md == "1T5A2" and got[7].nm == 2
    check(engine, regions)


def test_edge_cases(engine):
    regions = [
        R([], [0, 1, 2], 5, 5, 10, False, 0, 5, 7),           # empty query segment
        R([0, 1, 2], [], 0, 3, 3, True, 0, 5, 7),             # empty reference segment
        R([0], [0], 0, 1, 1, False, 1, 1),                    # one base, fast path
        R([4, 0, 1], [0, 4, 1], 2, 5, 9, True, -3, 2, 11),    # N bases in both
        R([0, 1, 2, 3] * 4, [2] * 3 + [0, 1, 2, 3] * 4 + [2] * 2, 1, 17, 20, False, 10, 1, 50),
    ]
    check(engine, regions)


@pytest.mark.parametrize("seed", [1, 2])
def test_random_regions(engine, seed):
    """NW-workload pairs as regions: half with the planted true score (retries
    stay rare), a quarter forced to retry (true score above any alignment), and
    equal-length substitution-only pairs for the fast path; random strand,
    clips and positions."""
    js = wl.generate(wl.NW, 40 + seed, 3000)
    rng = np.random.default_rng(seed)
    sp = ScoringParameters()
    regions = []
    for k in range(js["n"]):
        q = js["queries"][js["q_off"][k]:js["q_off"][k] + js["q_len"][k]]
        t = js["targets"][js["t_off"][k]:js["t_off"][k] + js["t_len"][k]]
        u = rng.random()
        if u < 0.25:
            t = q.copy()
            sub = rng.random(len(t)) < 0.02
            t[sub] = (t[sub] + rng.integers(1, 4, sub.sum())) % 4
            ts = int(len(q) - 5 * sub.sum())
        elif u < 0.5:
            ts = int(len(q) + 5)                             # unreachable: all retries run
        else:
            ts = int(js["true_score"][k])
        lead, trail = int(rng.integers(0, 30)), int(rng.integers(0, 30))
        qb, qe = lead, lead + len(q)
        regions.append(R(q.tolist(), t.tolist(), qb, qe, qe + trail, bool(rng.random() < 0.5),
                         ts, int(js["w"][k]), int(rng.integers(0, 1 << 40))))
    st = check(engine, regions, sp)
    assert st["n_jobs"] == len(regions)


def test_cap_overflow_falls_back(engine):
    # alternating indels make a CIGAR longer than MAX_CIGAR_OPS: that region is None
    q, t = [], []
    for k in range(700):
        q += [0, 1, 2, 3]
        t += [0, 1, 2, 3] + ([1] if k % 2 else [])
    regions = [R(q, t, 0, len(q), len(q), False, len(q) * 2, 5),
               R([0, 1, 2, 3], [0, 1, 2, 3], 0, 4, 4, False, 4, 5)]
    got = CIGARGenerator.generateBatch(regions, ScoringParameters(), engine)
    ref = orc.cigar_generate(q, t, 0, len(q), len(q), False, len(q) * 2, 5)
    if len(ref["cigar"]) > CIGARGenerator.MAX_CIGAR_OPS:
        assert got[0] is None
    else:
        assert got[0] is not None and got[0].cigar == ref["cigar"]
    assert got[1] is not None and got[1].cigar == [4 << 4]
