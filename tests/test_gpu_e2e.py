"""End-to-end single-end SAM through the GPU library.

The oracle's harness (oracle/e2e_oracle.c) runs seeding, chaining and chain
filtering; the chains are extended by hipsw_extend_chains
(ExtensionAligner.extendChains) and the output regions' CIGARs come from
hipsw_cigar_generate (CIGARGenerator.generateBatch); the harness then does
dedup, primary marking, MAPQ and SAM text.  The SAM must be byte-identical to
the all-CPU run of the same harness, on the reference's lambda fixture (also
checked against TestData/e2e_expected_se.sam) and on config 1 (10,000
150-bp reads on a 48,502-bp random reference).  Config 1's call set (every
BandedSWScalar.align call of the CPU run) goes through hipsw_banded_extend in
both forms, bit-exact against the oracle's logged results.
"""
import time

import numpy as np
import pytest

from oracle import e2e
from swiftbwa_amd import native, serving
from swiftbwa_amd.metalsw import CIGARGenerator, CIGARRegion, ExtensionAligner, ScoringParameters
from test_oracle_e2e import _lambda, check_lambda_sam, config1

pytestmark = pytest.mark.gpu


def run_gpu(E, names, reads, quals, engine, sp, times=None):
    t0 = time.perf_counter()
    chains = E.seed(reads)
    t1 = time.perf_counter()
    pac = serving.pack_pac(E.fwd)
    regs = ExtensionAligner.extendChains(chains, [bytes(r.tolist()) for r in E.reads], pac, E.gl,
                                         sp, engine)
    assert len(regs) == len(chains), "hipsw_extend_chains failed: " + engine.last_error()
    t2 = time.perf_counter()
    E.set_regions(regs)
    jobs = E.finish()
    t3 = time.perf_counter()
    regions = [CIGARRegion(bytes(j["qseg"].tolist()), bytes(j["rseg"].tolist()), j["qb"], j["qe"],
                           j["read_len"], j["rev"], j["true_score"], j["w"], j["ref_pos"])
               for j in jobs]
    res = CIGARGenerator.generateBatch(regions, sp, engine) if regions else []
    assert len(res) == len(regions), "hipsw_cigar_generate failed: " + engine.last_error()
    for j, r in enumerate(res):
        assert r is not None, "CIGAR %d over the caps" % j
        E.set_cigar(j, r.cigar, r.nm, r.md, r.pos)
    t4 = time.perf_counter()
    sam = E.sam(names, quals)
    if times is not None:
        times.update(seed=t1 - t0, extend=t2 - t1, finish=t3 - t2, cigar=t4 - t3,
                     sam=time.perf_counter() - t4)
    return sam


def test_lambda_gpu(engine):
    name, fwd, reads, expected = _lambda()
    names, bases, quals = [r[0] for r in reads], [r[1] for r in reads], [r[2] for r in reads]
    cpu = e2e.E2E(fwd, name).run_cpu(names, bases, quals)
    gpu = run_gpu(e2e.E2E(fwd, name), names, bases, quals, engine, ScoringParameters())
    assert gpu == cpu
    check_lambda_sam(gpu, expected)


def test_config1_gpu(engine):
    ref, names, reads, quals, pos, rev = config1(10000)
    E = e2e.E2E(ref, "config1")
    t0 = time.perf_counter()
    cpu = E.run_cpu(names, reads, quals)
    t_cpu = time.perf_counter() - t0
    times = {}
    gpu = run_gpu(e2e.E2E(ref, "config1"), names, reads, quals, engine, ScoringParameters(), times)
    print("config 1: CPU harness %.3f s; GPU run %s" % (
        t_cpu, " ".join("%s %.3f s" % kv for kv in times.items())), flush=True)
    cl, gl = cpu.splitlines(), gpu.splitlines()
    assert len(cl) == len(gl) == len(names)
    diff = [k for k in range(len(cl)) if cl[k] != gl[k]]
    assert not diff, "first differing record:\n%s\n%s" % (cl[diff[0]], gl[diff[0]])


def test_config1_call_set(engine):
    ref, names, reads, quals, pos, rev = config1(10000)
    E = e2e.E2E(ref, "config1")
    E.seed(reads)
    E.extend_cpu()
    js, want = E.calls()
    assert js["n"] > 1000
    sp = ScoringParameters()
    for mode in (native.MODE_PLAIN, native.MODE_BATCHED):
        got = engine.banded_extend(js, sp, mode=mode)
        bad = np.nonzero((got != want).any(axis=1))[0]
        assert len(bad) == 0, "mode %d: %d of %d calls differ, first %d: gpu %s oracle %s" % (
            mode, len(bad), js["n"], bad[0], got[bad[0]].tolist(), want[bad[0]].tolist())
