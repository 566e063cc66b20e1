"""The end-to-end single-end harness of the oracle (oracle/e2e_oracle.c), pinned
on the CPU:

* index: SA rows are the sorted suffixes of fwd + revcomp + '$', the BWT and
  occ / count tables follow from them (IndexBuilder.swift:20-110,
  FMIndexWriter.swift:18-60, BWT.swift:58-77);
* SMEMs: against a brute-force SMEM enumeration on the text, with every SA
  interval equal to the substring's occurrence set (SMEMFinder.swift);
* the whole pipeline: the reference's lambda fixture
  (TestData/e2e_expected_se.sam, checked field by field as
  IntegrationTests.swift:319-358 does: POS, FLAG bits, CIGAR, MAPQ within 5,
  NM, MD);
* the config-1 read set: every read maps, at its simulated locus in >= 99.5%
  of cases, and the call set is non-empty;
* paired end: the reference's lambda PE fixture (TestData/e2e_expected_pe.sam,
  IntegrationTests.swift:430-479: POS and CIGAR of mapped reads, FLAG bits,
  MAPQ, NM, MD; four pairs, so the insert-size estimate fails and no rescue
  runs), and a simulated set whose FR distribution, rescue and proper pairs
  follow from how it was made.
"""
import os

import numpy as np

from oracle import e2e
from swiftbwa_amd import serving, workload as wl

GOLD = os.path.join(os.path.dirname(__file__), "golden", "lambda")


def _text(fwd):
    fwd = np.asarray(fwd, np.uint8)
    return bytes(np.concatenate([fwd, 3 - fwd[::-1]]).tolist())


def test_index_small():
    rng = np.random.default_rng(3)
    fwd = rng.integers(0, 4, 300).astype(np.uint8)
    E = e2e.E2E(fwd, "t")
    T = _text(fwd)
    n = len(T)
    sa = [E.sa(i) for i in range(n + 1)]
    assert sa[0] == n and sorted(sa) == list(range(n + 1))
    sfx = [T[p:] for p in sa]                  # '$' (end of text) sorts first
    assert all(sfx[i] < sfx[i + 1] for i in range(n))
    bwt = [E.bwt(i) for i in range(n + 1)]
    assert bwt == [T[p - 1] if p > 0 else 4 for p in sa]
    cnt = E.counts()
    assert cnt[0] == 1 and cnt[4] == n + 1
    for c in range(4):
        assert cnt[c + 1] - cnt[c] == T.count(bytes([c]))
    for pos in (0, 1, 7, n // 2, n, n + 1):
        for c in range(4):
            assert E.occ(pos, c) == sum(1 for x in bwt[:pos] if x == c)


def _occurrences(T, s):
    out, i = [], T.find(s)
    while i >= 0:
        out.append(i)
        i = T.find(s, i + 1)
    return out


def _brute_smems(T, q, min_len):
    """SMEMs with >= 1 occurrence: for each start i the longest match L(i);
    (i, i + L(i)) is super-maximal when no earlier start reaches as far."""
    ends = []
    for i in range(len(q)):
        lo, hi = 0, len(q) - i
        while lo < hi:                          # longest occurring prefix of q[i:]
            mid = (lo + hi + 1) // 2
            if T.find(bytes(q[i:i + mid])) >= 0:
                lo = mid
            else:
                hi = mid - 1
        ends.append(i + lo)
    out, best = set(), -1
    for i in range(len(q)):
        if ends[i] > best and ends[i] - i >= min_len:
            out.add((i, ends[i]))
        best = max(best, ends[i])
    return out


def test_smems_brute_force():
    rng = np.random.default_rng(5)
    fwd = rng.integers(0, 4, 3000).astype(np.uint8)
    fwd[1000:1100] = fwd[2000:2100]             # a repeat: multi-occurrence SMEMs
    E = e2e.E2E(fwd, "t")
    T = _text(fwd)
    sa = np.array([E.sa(i) for i in range(len(T) + 1)])
    for r in range(40):
        p = int(rng.integers(0, 2800))
        q = np.frombuffer(T[p:p + 120] if r % 2 else T[3000 + p:3000 + p + 120], np.uint8).copy()
        q[rng.random(len(q)) < 0.03] = rng.integers(0, 4)
        if r % 7 == 0:
            q[40:60] = rng.integers(0, 4, 20)
        got = E.smems(q, 19, 1)
        spans = [(qb, qe) for _, _, qb, qe in got]
        assert set(spans) == _brute_smems(T, q, 19), r
        for k, l, qb, qe in got:
            occ = sorted(sa[k:l + 1].tolist())
            assert occ == _occurrences(T, bytes(q[qb:qe])), (r, qb, qe)


def _lambda():
    name, fwd = e2e.read_fasta(os.path.join(GOLD, "lambda.fa"))
    reads = e2e.read_fastq(os.path.join(GOLD, "e2e_reads_se.fq"))
    with open(os.path.join(GOLD, "e2e_expected_se.sam")) as f:
        expected = {k: v[0] for k, v in e2e.parse_sam(f.read()).items()}
    return name, fwd, reads, expected


FLAG_MASK = 0x4 | 0x10 | 0x1 | 0x2 | 0x8 | 0x20 | 0x40 | 0x80 | 0x100 | 0x800


def check_lambda_sam(sam, expected):
    """IntegrationTests.swift:319-358 (flagBitsMatch :189-192)."""
    got = e2e.parse_sam(sam)
    for name, exp in expected.items():
        act = got[name][0]
        assert act["pos"] == exp["pos"], name
        assert act["flag"] & FLAG_MASK == exp["flag"] & FLAG_MASK, name
        if not exp["flag"] & 4:
            assert act["cigar"] == exp["cigar"], name
        assert abs(act["mapq"] - exp["mapq"]) <= 5, name
        if "NM" in exp["tags"]:
            assert act["tags"]["NM"] == exp["tags"]["NM"], name
        if "MD" in exp["tags"]:
            assert act["tags"]["MD"] == exp["tags"]["MD"], name


def test_lambda_fixture():
    name, fwd, reads, expected = _lambda()
    E = e2e.E2E(fwd, name)
    sam = E.run_cpu([r[0] for r in reads], [r[1] for r in reads], [r[2] for r in reads])
    assert len(sam.splitlines()) == len(reads)
    check_lambda_sam(sam, expected)


def config1(n=10000):
    """Config 1: 48,502-bp uniform random reference (seed 0), 150-bp SE reads
    (seed 1) at 0.5% substitutions and 0.05% indels from both strands."""
    ref = serving.make_reference(0, 48502, copies=0)
    bases, pos, rev = wl.se_reads(ref, 1, n)
    names = ["r%d" % i for i in range(n)]
    quals = [b"I" * bases.shape[1]] * n
    return ref, names, list(bases), quals, pos, rev


def test_config1_cpu():
    ref, names, reads, quals, pos, rev = config1(3000)
    E = e2e.E2E(ref, "config1")
    sam = E.run_cpu(names, reads, quals)
    recs = e2e.parse_sam(sam)
    hits = 0
    for i, nm in enumerate(names):
        r = recs[nm][0]
        assert not r["flag"] & 4
        hits += abs(r["pos"] - 1 - pos[i]) <= 5 and bool(r["flag"] & 16) == bool(rev[i])
    assert hits >= 0.995 * len(names)
    js, out = E.calls()
    assert js["n"] > 0 and np.all(out[:, 0] >= js["h0"])


# ---- paired end ---------------------------------------------------------------

def _lambda_pe():
    name, fwd = e2e.read_fasta(os.path.join(GOLD, "lambda.fa"))
    r1 = e2e.read_fastq(os.path.join(GOLD, "e2e_reads_r1.fq"))
    r2 = e2e.read_fastq(os.path.join(GOLD, "e2e_reads_r2.fq"))
    with open(os.path.join(GOLD, "e2e_expected_pe.sam")) as f:
        exp = e2e.parse_sam(f.read())
    names = [a[0] for a in r1]
    assert names == [b[0] for b in r2]
    return name, fwd, names, r1, r2, exp


def check_lambda_pe_sam(sam, exp):
    """IntegrationTests.swift:430-479: POS and CIGAR for mapped reads, FLAG bits,
    MAPQ exact when 60 (else within 5), NM, MD; records keyed by name and mate."""
    got = e2e.parse_sam(sam)
    n_checked = 0
    for name, recs in exp.items():
        for er in recs:
            mate = er["flag"] & 0xC0
            ar = [r for r in got[name] if r["flag"] & 0xC0 == mate]
            assert len(ar) == 1, (name, mate)
            ar = ar[0]
            if not er["flag"] & 4:
                assert ar["pos"] == er["pos"], name
                assert ar["cigar"] == er["cigar"], name
            assert ar["flag"] & FLAG_MASK == er["flag"] & FLAG_MASK, (name, ar["flag"], er["flag"])
            if er["mapq"] == 60:
                assert ar["mapq"] == 60, name
            elif er["mapq"] > 0:
                assert abs(ar["mapq"] - er["mapq"]) <= 5, name
            for t in ("NM", "MD"):
                if t in er["tags"]:
                    assert ar["tags"][t] == er["tags"][t], (name, t)
            n_checked += 1
    assert n_checked == 8


def test_lambda_pe_fixture():
    name, fwd, names, r1, r2, exp = _lambda_pe()
    E = e2e.E2E(fwd, name)
    sam = E.run_pe_cpu(names, [a[1] for a in r1], [b[1] for b in r2], [a[2] for a in r1],
                       [b[2] for b in r2])
    check_lambda_pe_sam(sam, exp)
    d, _ = E.pe_dist()
    assert np.all(d[:, 0] == 1)           # four pairs: too few for the insert-size estimate


def pe_set(n=10000):
    """Paired-end set: 1-Mbp reference with planted 300-bp / 6-kb families (20
    copies each at 2% divergence), FR pairs of 2 x 150 bp with insert ~ N(400, 50),
    0.5% substitutions and 0.05% indels; 5% of pairs have a read 2 that seeding
    cannot place (a substitution every 15 bases), so mate rescue has to."""
    ref = serving.make_reference(5, 1_000_000, copies=20, divergence=0.02)
    b1, b2, pos, rev = wl.pe_reads(ref, 5, n, hard=0.05)
    names = ["p%d" % i for i in range(n)]
    q = [b"I" * b1.shape[1]] * n
    return ref, names, list(b1), list(b2), q, pos, rev


def test_pe_cpu():
    ref, names, r1, r2, q, pos, rev = pe_set(3000)
    E = e2e.E2E(ref, "chr5")
    E.seed_pairs(r1, r2)
    E.extend_cpu()
    assert E.pe_finish()
    d, pri = E.pe_dist()
    assert pri == 1 and d[1, 0] == 0       # FR estimated
    assert abs(d[1, 3] - 400) < 10 and abs(d[1, 4] - 50) < 10
    added = 0
    for direction in (1, 2):
        E.rescue_jobs(direction)
        E.rescue_cpu()
        added += E.rescue_apply()
    assert added > 0.03 * len(names)
    E.pe_cigar_plan()
    E.cigar_cpu()
    recs = e2e.parse_sam(E.pe_sam(names, q + q))
    proper = sum(all(r["flag"] & 0x2 for r in recs[n]) for n in names)
    assert proper >= 0.98 * len(names)
