"""Pins of the CPU oracle's local-SW and NW restatements where textbook
algorithms and the reference differ (collected by tests/test_oracle_pins.py,
and run against deliberately broken builds of the oracle by its mutation test).

Every check returns a list of failure strings; an empty list means pinned.
The expected values below are derived in each docstring from the reference's
source (LocalSWAligner.swift, GlobalAligner.swift) or taken from the
reference's own fixture; the Python restatements in tests/ref_restatements.py
are a second, separately written transcription used on random inputs.

Default scoring (ScoringParameters.swift): A=1, B=4, O=6, E=1, so the striped
pass uses bias 4, gapOE 7, gapE 1, and a profile byte of 5 for a match, 0 for
a mismatch.  With 16 lanes and sn = ceil(qlen/16) stripes, query position j
sits in stripe j % sn, lane j // sn (LocalSWAligner.swift:152-167).
"""
import itertools
import os

import numpy as np

from oracle import pyoracle as orc
from tests.ref_restatements import global_tb_full, local_affine_brute, local_sw_swift

HERE = os.path.dirname(os.path.abspath(__file__))

# fixed sequences over A, C, G (code 3 = T never occurs in them)
X20 = [1, 2, 1, 1, 1, 1, 0, 0, 0, 0, 2, 1, 0, 1, 0, 1, 2, 0, 1, 1]
Y20 = [2, 1, 0, 0, 1, 1, 0, 1, 1, 0, 2, 0, 2, 2, 2, 2, 1, 1, 0, 1]
Q30 = [2, 0, 0, 1, 1, 1, 0, 1, 1, 1, 2, 2, 2, 2, 2, 0, 2, 1, 0, 0, 0, 2, 2, 2, 0, 1, 1, 2, 0, 1]


def _expect(fails, name, got, want):
    if got != want:
        fails.append("%s: got %r, expected %r" % (name, got, want))


def local_single_pass_lazy_f():
    """LocalSWAligner.swift:234-242 runs the lazy-F correction once: F leaves
    the last stripe of lane l for the first stripe of lane l+1 (one shiftRight)
    and the walk stops at the first stripe it does not raise.  A horizontal
    gap (query bases against no target base) can therefore cross one lane
    boundary per row, not two.

    Query X20 + T*g + Y20 against target T*5 + X20 + Y20 + T*5 (T occurs only
    in the gap and the flanks).  Textbook local SW aligns X, a g-base gap, Y:
    40 - (7 + g), a gap run costing o + e * (len + 1) in the striped pass'
    model (open: satsub(H, gapOE), then satsub(., gapE)).  With g = 4, qlen = 44, sn = 3: the gap is query 20..23,
    lanes 6, 7, 7, 7 -- one boundary (20 -> 21), found: 29 = textbook.  With
    g = 5, qlen = 45, sn = 3: query 20..24 sits in lanes 6, 7, 7, 7, 8 -- two
    boundaries (20 -> 21, 23 -> 24), so the one-gap path is not found and the
    reference scores below the textbook 28; both transcriptions of the
    reference give 23."""
    fails = []
    for g, want_score, textbook in ((4, 29, 29), (5, 23, 28)):
        q = X20 + [3] * g + Y20
        t = [3] * 5 + X20 + Y20 + [3] * 5
        r = orc.local_sw(q, t)
        _expect(fails, "lazy-F g=%d score" % g, None if r is None else r[0], want_score)
        sc = orc.scoring()
        _expect(fails, "lazy-F g=%d textbook" % g, local_affine_brute(q, t, list(sc.mat), 6, 1), textbook)
    return fails


def local_lazy_f_stops_on_tie():
    """The lazy-F walk (LocalSWAligner.swift:234-242) stops at the first
    stripe it does not raise, even when its F would still raise a later
    stripe.  Query ACCCCAAACCCAACAACACCCA (22 bases, sn 2) against target
    ACCCCAAACCACAACACCCA (20): the first 10 bases match, the query then has
    two extra bases (query 10, 11), and its last 10 match the target's last
    10.  Row 9 (the 10th target base): H(9,9) = 10; in the striped gap model
    F entering query 10 is satsub(10, 7) - 1 = 2, carried by the correction
    into stripe 0, lane 5.  That cell's own value is also 2 (a diagonal
    match from H(8,9) = 1), so stripe 0 is unchanged and the walk stops; the
    correction would have set query 11 (stripe 1, lane 5) to
    max(satsub(2, 7), 2) - 1 = 1, from which the last 10 matches reach 11
    (the textbook score).  Without it query 11 stays 0, both halves score 10
    alone, and the first to reach 10 is reported: (10, 0, 9, 0, 9, 0)."""
    fails = []
    q = [0, 1, 1, 1, 1, 0, 0, 0, 1, 1, 1, 0, 0, 1, 0, 0, 1, 0, 1, 1, 1, 0]
    t = [0, 1, 1, 1, 1, 0, 0, 0, 1, 1, 0, 1, 0, 0, 1, 0, 1, 1, 1, 0]
    _expect(fails, "lazy-F stop on tie", orc.local_sw(q, t), (10, 0, 9, 0, 9, 0))
    sc = orc.scoring()
    _expect(fails, "lazy-F stop on tie, textbook", local_affine_brute(q, t, list(sc.mat), 6, 1), 11)
    return fails


def local_tie_across_stripes():
    """Best-cell ties (LocalSWAligner.swift:245-263): stripes are scanned in
    ascending order, lanes inside a stripe in ascending order, and only a
    strictly higher stripe maximum moves the best cell.  So among equal row
    maxima the winner is the smallest stripe s = j % sn, then the smallest
    lane, not the smallest query position.

    Query of 20 bases (sn = 2) all C except A at 3 and 4; target "A".  Row 0
    scores 1 at query 3 (stripe 1, lane 1) and at query 4 (stripe 0, lane 2):
    stripe 0 is scanned first, so qEnd = 4 (textbook: 3).  The reverse pass
    (query 4..0 reversed starts with A) gives qb = 4: (1, 4, 4, 0, 0, 0).
    Same-stripe ties go to the lower lane: A at query 2 and 4 (both stripe
    0, lanes 1 and 2) gives qEnd = 2: (1, 2, 2, 0, 0, 0)."""
    fails = []
    q = [1] * 20
    q[3] = q[4] = 0
    _expect(fails, "cross-stripe tie", orc.local_sw(q, [0]), (1, 4, 4, 0, 0, 0))
    q = [1] * 20
    q[2] = q[4] = 0
    _expect(fails, "same-stripe tie", orc.local_sw(q, [0]), (1, 2, 2, 0, 0, 0))
    return fails


def local_padding_lane_max():
    """Padding lanes (j >= qlen) take part in the stripe maximum but not in
    the best-column search (LocalSWAligner.swift:250-262): when only a padding
    lane holds a new maximum, score and row move and qEnd keeps its old value.

    Free gaps (O = E = 0: gapOE = gapE = 0), query C*17 + A (qlen 18, sn 2,
    padding lanes 9..15), target "CA".  Row 0: every real column scores 1
    (C matches or F carries 1), best (1, row 0, query 0).  Row 1 (A): query 17
    scores 2 (diagonal from query 16); the lazy-F step carries that 2 from
    lane 8 (stripe 1) into padding lane 9 of stripe 0 (query 18 >= qlen).
    Stripe 0 is scanned first: its maximum 2 is in padding only, so the score
    becomes 2 at row 1 and qEnd stays 0; stripe 1's 2 at query 17 is not
    strictly higher.  The reverse pass on query[0..0], target[1..0] reversed
    ("C" against "AC") ends at (0, 1): (2, 0, 0, 0, 1, 0)."""
    fails = []
    fg = orc.scoring(o_ins=0, e_ins=0, o_del=0, e_del=0)
    _expect(fails, "padding-lane max", orc.local_sw([1] * 17 + [0], [1, 0], fg), (2, 0, 0, 0, 1, 0))
    return fails


def local_score2_radius():
    """score2 (LocalSWAligner.swift:266-303): one peak per run of consecutive
    rows whose column maximum is >= minSubScore, at the run's first row of its
    highest value; score2 is the best peak with column < maxI - r or
    > maxI + r, r = ceil(score / maxBase) (maxBase = A = 1 here, so r = score).

    Query Q30 against T*12 + Q30 + T*g + Q30[:25] + T*12: the full copy scores
    30 ending at row 41 (r = 30, so the zone is rows 11..71); the 25-base copy
    peaks at its last row 12 + 30 + g + 24.  g = 5 puts that peak at row 71,
    inside the zone: score2 = 0.  g = 6 puts it at row 72, just outside:
    score2 = 25.  The runs are separate: the T rows drive the column maximum
    below 19 between the copies."""
    fails = []
    for g, want in ((5, 0), (6, 25)):
        t = [3] * 12 + Q30 + [3] * g + Q30[:25] + [3] * 12
        _expect(fails, "score2 radius g=%d" % g, orc.local_sw(Q30, t, min_sub_score=19),
                (30, 0, 29, 12, 41, want))
    return fails


def local_vs_swift_transcription(n=400, seed=9):
    """The C oracle against the Python transcription of LocalSWAligner on
    inputs built to exercise long horizontal gaps, ties, tandem copies
    (score2), free gaps and the scalar path."""
    fails = []
    rng = np.random.default_rng(seed)
    for k in range(n):
        lq = int(rng.integers(1, 60)) if k % 7 else int(rng.integers(200, 300))
        lt = int(rng.integers(1, 120))
        q = [int(x) for x in rng.integers(0, 4 if k % 3 else 2, lq)]
        t = [int(x) for x in rng.integers(0, 4 if k % 3 else 2, lt)]
        if lt > 10 and k % 2:
            p = int(rng.integers(0, lt - 5))
            m = min(lq, lt - p)
            t[p:p + m] = q[:m]
            if m > 12:
                c = int(rng.integers(2, m - 2))
                t = t[:p + c] + t[p + c + int(rng.integers(1, 9)):]
        free = k % 11 == 0
        sc = orc.scoring(o_ins=0, e_ins=0, o_del=0, e_del=0) if free else orc.scoring()
        ms = 19 if k % 4 else 0
        got = orc.local_sw(q, t, sc, min_sub_score=ms)
        want = local_sw_swift(q, t, list(sc.mat), sc.o_ins, sc.e_ins, sc.a, sc.b, ms)
        if got != want:
            fails.append("local transcription %d: %r vs %r" % (k, got, want))
            if len(fails) > 3:
                break
    return fails


def nw_lambda_fixture_deletion():
    """The reference's fixture TestData/e2e_expected_se.sam aligns read
    del3_30000 as 50M3D50M with MD 50^AAC50: the reference there reads
    ...GCG AAC AAC AAG..., so five placements of the 3-base deletion score the
    same and the expected record has the leftmost.  NW of the read's 20 bases
    either side of the deletion against the 43 reference bases must place it
    the same way: 20M3D20M."""
    fails = []
    seq = "".join(l.strip() for l in open(os.path.join(HERE, "golden", "lambda", "lambda.fa"))
                  if not l.startswith(">"))
    enc = {"A": 0, "C": 1, "G": 2, "T": 3}
    pos = 30000                                   # 0-based start of the read (POS 30001)
    ref = [enc[c] for c in seq[pos + 30:pos + 73]]
    read = ref[:20] + ref[23:]
    score, cig, _ = orc.ksw_global(read, ref, w=10)
    _expect(fails, "lambda deletion CIGAR", list(cig), [20 << 4 | 0, 3 << 4 | 2, 20 << 4 | 0])
    _expect(fails, "lambda deletion score", score, 40 - 9)
    return fails


def nw_homopolymer_placement():
    """Equal-score gap placement in a homopolymer, read off the strict
    comparisons of GlobalAligner.swift:114-136: the traceback starts at the
    last cell and takes the diagonal whenever H came from it (ties favour the
    diagonal), so the gap is pushed to the left end of the run.
    AAAC vs AAAAC (one deleted A): 1D4M.  AAAAC vs AAAC (one inserted A): 1I4M."""
    fails = []
    _, cig, _ = orc.ksw_global([0, 0, 0, 1], [0, 0, 0, 0, 1], w=3)
    _expect(fails, "homopolymer deletion", list(cig), [1 << 4 | 2, 4 << 4 | 0])
    _, cig, _ = orc.ksw_global([0, 0, 0, 0, 1], [0, 0, 0, 1], w=3)
    _expect(fails, "homopolymer insertion", list(cig), [1 << 4 | 1, 4 << 4 | 0])
    return fails


def nw_vs_full_matrix(max_len=4):
    """GlobalAligner (fill + traceback) against the full-matrix restatement
    with the tie policy stated declaratively, on every pair of binary
    sequences up to max_len bases (many equal-score paths), bands 1, 2 and 4,
    under three scorings: the default, cheap gaps (ties between gap and
    mismatch paths), and asymmetric insertion/deletion penalties."""
    fails = []
    scorings = [orc.scoring(), orc.scoring(b=1, o_ins=1, e_ins=1, o_del=1, e_del=1),
                orc.scoring(b=2, o_ins=1, e_ins=2, o_del=3, e_del=1)]
    seqs = [list(s) for n in range(1, max_len + 1) for s in itertools.product((0, 1), repeat=n)]
    for sc in scorings:
        mat = list(sc.mat)
        for q in seqs:
            for t in seqs:
                for w in (1, 2, 4):
                    s, cig, _ = orc.ksw_global(q, t, sc, w=w)
                    ws, wc = global_tb_full(q, t, mat, sc.o_del, sc.e_del, sc.o_ins, sc.e_ins, w)
                    if (s, list(cig)) != (ws, wc):
                        fails.append("NW %r %r w=%d: oracle (%d, %r) restatement (%d, %r)" %
                                     (q, t, w, s, list(cig), ws, wc))
                        if len(fails) > 3:
                            return fails
    return fails


def nw_vs_full_matrix_random(n=300, seed=4):
    fails = []
    rng = np.random.default_rng(seed)
    sc = orc.scoring()
    mat = list(sc.mat)
    for k in range(n):
        lq = int(rng.integers(1, 40))
        lt = max(1, lq + int(rng.integers(-8, 9)))
        alpha = 2 if k % 2 else 4
        q = [int(x) for x in rng.integers(0, alpha, lq)]
        t = [int(x) for x in rng.integers(0, alpha, lt)]
        w = int(rng.integers(1, 12))
        s, cig, _ = orc.ksw_global(q, t, sc, w=w)
        ws, wc = global_tb_full(q, t, mat, sc.o_del, sc.e_del, sc.o_ins, sc.e_ins, w)
        if (s, list(cig)) != (ws, wc):
            fails.append("NW random %d: (%d, %r) vs (%d, %r)" % (k, s, list(cig), ws, wc))
            if len(fails) > 3:
                break
    return fails


LOCAL_PINS = [local_single_pass_lazy_f, local_lazy_f_stops_on_tie, local_tie_across_stripes, local_padding_lane_max,
              local_score2_radius, local_vs_swift_transcription]
NW_PINS = [nw_lambda_fixture_deletion, nw_homopolymer_placement, nw_vs_full_matrix,
           nw_vs_full_matrix_random]


def run_all():
    """Failures of every pin, as one list (used by the mutation test)."""
    out = []
    for fn in LOCAL_PINS + NW_PINS:
        try:
            out += fn()
        except Exception as e:   # noqa: BLE001 -- a crash in a mutant counts as a failure
            out.append("%s raised %r" % (fn.__name__, e))
    return out


if __name__ == "__main__":
    import sys
    lib = os.environ.get("ORACLE_LIB")
    if lib:
        orc.use_library(lib)
    f = run_all()
    for line in f:
        print(line)
    print("FAILURES %d" % len(f))
    sys.exit(0)
