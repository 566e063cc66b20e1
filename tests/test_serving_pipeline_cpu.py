"""Read chunks through the serving pipeline with the oracle standing in for the
device (tests/serving_pipeline.py CpuServer), on a multi-sequence reference:
the chain passes (hipsw_chain_*), raw rescue windows clamped per
PackedReference.fetchSequence, and the CIGAR route / finish steps must give
SAM byte-identical to the harness's all-CPU run of each chunk.  Also the
multi-sequence harness itself: ReferenceMetadata.sequenceID vectors
(BWACoreTests.swift:97-127), per-sequence SAM fields, and alignments that never
cross a sequence end."""
import re

import numpy as np
import pytest

from oracle import e2e
from swiftbwa_amd import serving, workload as wl
from swiftbwa_amd.metalsw import ScoringParameters
import serving_pipeline as SP


def test_sequence_id_reference_vectors():
    ref = serving.make_reference(3, 450, copies=0)
    E = e2e.E2E(ref, contigs=[("chr1", 100), ("chr2", 200), ("chr3", 150)])
    # BWACoreTests.swift:107-112
    for pos, rid in ((0, 0), (99, 0), (100, 1), (299, 1), (300, 2), (449, 2)):
        assert E.seq_id(pos) == rid
    E2 = e2e.E2E(ref[:300], contigs=[("chr1", 100), ("chr2", 200)])
    assert E2.seq_id(150) == 1                   # decodePosition(150) = (1, 50), :124-126
    with pytest.raises(ValueError):
        e2e.E2E(ref, contigs=[("a", 100), ("b", 100)])


def _multi(total=600_000, n_ctg=6, pairs=1500, seed=11):
    ref = serving.make_reference(seed, total, copies=6, divergence=0.10)
    lens = serving.contig_lengths(total, n_ctg)
    contigs = [("chr%d" % (k + 1), n) for k, n in enumerate(lens)]
    b1, b2, pos, rev = wl.pe_reads(ref, seed + 1, pairs, hard=0.05)
    names = ["p%d" % i for i in range(pairs)]
    q = [b"I" * b1.shape[1]] * pairs
    return ref, contigs, names, list(b1), list(b2), q


def test_multi_contig_sam_fields():
    """Every mapped record lies inside its sequence; RNAME/POS are per sequence;
    translating them back gives the single-sequence run's position for pairs
    far from any sequence end."""
    ref, contigs, names, r1, r2, q = _multi()
    sam = e2e.E2E(ref, contigs=contigs).run_pe_cpu(names, r1, r2, q, q)
    one = e2e.E2E(ref, "all").run_pe_cpu(names, r1, r2, q, q)
    lens = {n: l for n, l in contigs}
    offs = dict(zip([n for n, _ in contigs], np.concatenate([[0], np.cumsum([l for _, l in contigs])[:-1]])))
    recs = e2e.parse_sam(sam)
    recs1 = e2e.parse_sam(one)
    same = far = 0
    for name, rs in recs.items():
        for r in rs:
            if r["flag"] & 4:
                continue
            rname, pos, cig = r["rname"], r["pos"], r["cigar"]
            assert rname in lens
            span = sum(int(n) for n, _ in re.findall(r"(\d+)([MDN=X])", cig))
            assert 1 <= pos and pos - 1 + span <= lens[rname], (name, rname, pos, cig)
            if r["flag"] & 0x900:
                continue
            gpos = offs[rname] + pos
            inside = all(abs(gpos - o) > 2000 for o in offs.values()) and \
                all(abs(gpos - (offs[n] + lens[n])) > 2000 for n in lens)
            if inside:
                far += 1
                m = [x for x in recs1[name] if x["flag"] & 0xC0 == r["flag"] & 0xC0 and not x["flag"] & 0x900]
                same += len(m) == 1 and m[0]["pos"] == gpos
    assert far > 1000 and same >= 0.99 * far, (same, far)


def test_pipeline_cpu_server_identical():
    ref, contigs, names, r1, r2, q = _multi(pairs=1200)
    sp = ScoringParameters()
    E = e2e.E2E(ref, contigs=contigs)
    ranges = SP.split_pairs(len(names), [300, 500, 400])
    cpu = []
    for lo, hi in ranges:
        F = E.fork()
        cpu.append(F.run_pe_cpu(names[lo:hi], r1[lo:hi], r2[lo:hi], q[lo:hi], q[lo:hi]))
        F.close()
    srv = SP.CpuServer(ref, [n for _, n in contigs])
    sams, logs, _ = SP.run_chunks(E, srv, names, r1, r2, q, ranges, sp, concurrent=True)
    for k in range(len(ranges)):
        assert sams[k] == cpu[k], k
    stages = {s for lg in logs for s, _, _ in lg}
    assert {"ext0", "ext2", "rescue1", "cigar"} <= stages, stages
