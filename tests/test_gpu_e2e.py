"""End-to-end SAM through the GPU library, single-end and paired-end.

The oracle's harness (oracle/e2e_oracle.c) runs seeding, chaining and chain
filtering; the chains are extended by hipsw_extend_chains
(ExtensionAligner.extendChains) and the output regions' CIGARs come from
hipsw_cigar_generate (CIGARGenerator.generateBatch); the harness then does
dedup, primary marking, MAPQ and SAM text.  The SAM must be byte-identical to
the all-CPU run of the same harness, on the reference's lambda fixture (also
checked against TestData/e2e_expected_se.sam) and on config 1 (10,000
150-bp reads on a 48,502-bp random reference).  Config 1's call set (every
BandedSWScalar.align call of the CPU run) goes through hipsw_banded_extend in
both forms, bit-exact against the oracle's logged results.  Paired end: the
mate-rescue local alignments (both directions, in the CPU path's order) run
through hipsw_local_sw as well; SAM byte-identical to the CPU run on the
reference's lambda PE fixture and on a simulated 10,000-pair set with repeats
and rescue-only mates.
"""
import time

import numpy as np
import pytest

from oracle import e2e
from swiftbwa_amd import native, serving
from swiftbwa_amd.metalsw import CIGARGenerator, CIGARRegion, ExtensionAligner, ScoringParameters
from test_oracle_e2e import (_lambda, _lambda_pe, check_lambda_pe_sam, check_lambda_sam, config1,
                             pe_set)

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


def run_pe_gpu(E, names, r1, r2, q1, q2, engine, sp, stats=None):
    chains = E.seed_pairs(r1, r2)
    pac = serving.pack_pac(E.fwd)
    regs = ExtensionAligner.extendChains(chains, [bytes(r.tolist()) for r in E.reads], pac, E.gl,
                                         sp, engine)
    assert len(regs) == len(chains), "hipsw_extend_chains failed: " + engine.last_error()
    E.set_regions(regs)
    if E.pe_finish():
        for direction in (1, 2):
            js = E.rescue_jobs(direction)
            if js["n"]:
                out, ok = engine.local_sw(js, sp, min_sub_score=sp.minSeedLength * sp.matchScore)
                E.set_rescue(out, ok)
                if stats is not None:
                    stats.setdefault("rescue_jobs", []).append(js["n"])
            E.rescue_apply()
    jobs = E.pe_cigar_plan()
    regions = [CIGARRegion(bytes(j["qseg"].tolist()), bytes(j["rseg"].tolist()), j["qb"], j["qe"],
                           j["read_len"], j["rev"], j["true_score"], j["w"], j["ref_pos"])
               for j in jobs]
    res = CIGARGenerator.generateBatch(regions, sp, engine) if regions else []
    assert len(res) == len(regions), "hipsw_cigar_generate failed: " + engine.last_error()
    for j, r in enumerate(res):
        assert r is not None, "CIGAR %d over the caps" % j
        E.set_cigar(j, r.cigar, r.nm, r.md, r.pos)
    return E.pe_sam(names, list(q1) + list(q2))


def test_lambda_pe_gpu(engine):
    name, fwd, names, r1, r2, exp = _lambda_pe()
    args = (names, [a[1] for a in r1], [b[1] for b in r2], [a[2] for a in r1], [b[2] for b in r2])
    cpu = e2e.E2E(fwd, name).run_pe_cpu(*args)
    gpu = run_pe_gpu(e2e.E2E(fwd, name), *args, engine, ScoringParameters())
    assert gpu == cpu
    check_lambda_pe_sam(gpu, exp)


def test_pe_gpu(engine):
    ref, names, r1, r2, q, pos, rev = pe_set(10000)
    cpu = e2e.E2E(ref, "chr5").run_pe_cpu(names, r1, r2, q, q)
    st = {}
    t0 = time.perf_counter()
    gpu = run_pe_gpu(e2e.E2E(ref, "chr5"), names, r1, r2, q, q, engine, ScoringParameters(), st)
    print("paired end: GPU run %.3f s, rescue jobs per direction %s" % (
        time.perf_counter() - t0, st.get("rescue_jobs")), flush=True)
    assert st.get("rescue_jobs") and st["rescue_jobs"][0] > 100
    cl, gl = cpu.splitlines(), gpu.splitlines()
    assert len(cl) == len(gl) >= 2 * len(names)
    diff = [k for k in range(len(cl)) if cl[k] != gl[k]]
    assert not diff, "first differing record:\n%s\n%s" % (cl[diff[0]], gl[diff[0]])
