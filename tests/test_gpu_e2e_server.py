"""Paired-end reads -> serving loop -> SAM, on a multi-sequence repeat reference.

A 3-Mbp reference of random ACGT with planted 300-bp and 6-kb families (12
copies each at 10% divergence, the serving workload's density) split into
eight sequences, one of them 16,569 bases; 6,000 read pairs (2 x 150 bp,
insert N(400, 50), 0.5% substitutions, 0.05% indels, 5% of read 2s only
rescue can place) in read chunks of log-uniform sizes, all in flight at once
on a 4-slot server.  Each chunk's chain extension passes, rescue local SW
(fetchSequence-clamped windows) and CIGAR NW run on the device through the
server; the SAM of every chunk must be byte-identical to the harness's
all-CPU run of the same chunk (each chunk is one batch, as the reference's
processFullBatch sees it)."""
import time

import numpy as np
import pytest

from oracle import e2e
from swiftbwa_amd import serving, workload as wl
from swiftbwa_amd.metalsw import ScoringParameters
import serving_pipeline as SP

pytestmark = pytest.mark.gpu


def _set(total=3_000_000, n_ctg=8, pairs=6000, seed=21):
    ref = serving.make_reference(seed, total, copies=12, divergence=0.10)
    lens = serving.contig_lengths(total, n_ctg)
    contigs = [("chr%d" % (k + 1), n) for k, n in enumerate(lens)]
    b1, b2, _, _ = wl.pe_reads(ref, seed + 1, pairs, hard=0.05)
    names = ["q%d" % i for i in range(pairs)]
    q = [b"I" * b1.shape[1]] * pairs
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(200), np.log(2500), 64)).astype(int)
    return ref, contigs, names, list(b1), list(b2), q, SP.split_pairs(pairs, sizes)


def test_pe_reads_through_server_sam_identical(engine):
    ref, contigs, names, r1, r2, q, ranges = _set()
    E = e2e.E2E(ref, contigs=contigs)
    sp = ScoringParameters()
    srv = serving.Server(engine, slots=4, hbm_budget=4 << 30, scoring=sp)
    try:
        srv.set_reference(serving.pack_pac(ref), len(ref))
        srv.set_contigs([n for _, n in contigs])
        print("read chunks:", [hi - lo for lo, hi in ranges], flush=True)
        t0 = time.perf_counter()
        sams, logs, walls = SP.run_chunks(E, srv, names, r1, r2, q, ranges, sp, concurrent=True)
        wall = time.perf_counter() - t0
    finally:
        srv.close()
    stages = [lat for lg in logs for _, lat, _ in lg]
    print("pairs %d in %d chunks, %.2f s (%.0f reads/s incl. CPU seeding); chunk wall p50 %.0f ms "
          "p99 %.0f ms; server stage latency p50 %.1f ms p99 %.1f ms"
          % (len(names), len(ranges), wall, 2 * len(names) / wall, np.percentile(walls, 50) * 1e3,
             np.percentile(walls, 99) * 1e3, np.percentile(stages, 50), np.percentile(stages, 99)),
          flush=True)
    for k, (lo, hi) in enumerate(ranges):
        F = E.fork()
        cpu = F.run_pe_cpu(names[lo:hi], r1[lo:hi], r2[lo:hi], q[lo:hi], q[lo:hi])
        F.close()
        assert sams[k] == cpu, "chunk %d (pairs %d..%d)" % (k, lo, hi)
    recs = e2e.parse_sam("".join(sams))
    assert {r["rname"] for rs in recs.values() for r in rs if not r["flag"] & 4} >= {"chr1", "chr8"}
