"""Failure paths: an injected allocation, launch or copy failure makes the
entry point return HIPSW_E_DEVICE (the Swift shim's nil / []), the engine stays
usable, and the caller's CPU path then gives exactly the results the GPU would
have given -- swiftbwa's fallbacks (BWAMemAligner.swift:1941-1949, :2961-3019).
The faults are forced from the host (hipsw_fault_inject): every kernel that
does run still exits normally."""
import numpy as np
import pytest

from oracle import e2e, pyoracle as orc
from swiftbwa_amd import native, serving, workload as wl
from swiftbwa_amd.metalsw import (CIGARGenerator, CIGARRegion, ExtensionAligner,
                                  ScoringParameters)
from test_oracle_e2e import pe_set

pytestmark = pytest.mark.gpu
SITES = [native.FAULT_ALLOC, native.FAULT_LAUNCH, native.FAULT_COPY]


def _calls(engine, sp):
    ext = wl.generate(wl.EXT_MIXED, 51, 3000)
    nw = wl.generate(wl.NW, 52, 2000)
    loc = wl.generate(wl.LOCAL, 53, 1500)
    osc = orc.scoring()
    ref_ext, _ = orc.ksw_extend_batch(ext, osc, nthreads=8)
    ref_nw = orc.global_batch(nw, osc, max_retries=3, cap=512, nthreads=8)
    ref_loc, ref_ok, _ = orc.local_batch(loc, osc, min_sub_score=19, nthreads=8)
    ref_loc[ref_ok == 0] = 0
    return [
        ("extension", lambda: engine.banded_extend(ext, sp),
         lambda got: np.array_equal(got, ref_ext)),
        ("nw", lambda: engine.global_nw(nw, sp),
         lambda got: np.array_equal(got[0], ref_nw[0]) and np.array_equal(got[2], ref_nw[2])),
        ("local", lambda: engine.local_sw(loc, sp, min_sub_score=19),
         lambda got: np.array_equal(got[0], ref_loc) and np.array_equal(got[1], ref_ok)),
    ]


@pytest.mark.parametrize("site", SITES)
def test_one_shot_entry_points_fail_cleanly(engine, site):
    sp = ScoringParameters()
    for name, call, ok in _calls(engine, sp):
        for countdown in (1, 2):
            print("fault site %d countdown %d on %s" % (site, countdown, name), flush=True)
            engine.fault_inject(site, countdown)
            try:
                got = call()
                # fewer steps of this site than the countdown: the call ran clean
                assert ok(got), (name, site, countdown)
            except native.HipswError as e:
                assert e.status == native.HIPSW_E_DEVICE
                assert "injected" in engine.last_error(), engine.last_error()
            finally:
                engine.fault_inject(native.FAULT_NONE, 0)
            assert ok(call()), (name, site, countdown)          # the engine is still usable


def _pe_run_with_faults(E, names, r1, r2, q1, q2, engine, sp, fault_phase, site):
    """run_pe_gpu of test_gpu_e2e with the caller's fallback: a phase whose GPU
    call fails runs the harness's CPU code for that phase instead."""
    fell_back = []

    def armed(phase):
        if phase == fault_phase:
            engine.fault_inject(site, 1)

    chains = E.seed_pairs(r1, r2)
    pac = serving.pack_pac(E.fwd)
    armed("extension")
    regs = ExtensionAligner.extendChains(chains, [bytes(r.tolist()) for r in E.reads], pac, E.gl,
                                         sp, engine)
    engine.fault_inject(native.FAULT_NONE, 0)
    if len(regs) != len(chains):
        fell_back.append("extension")
        E.extend_cpu()
    else:
        E.set_regions(regs)
    if E.pe_finish():
        for direction in (1, 2):
            js = E.rescue_jobs(direction)
            if js["n"]:
                armed("rescue")
                try:
                    out, ok = engine.local_sw(js, sp, min_sub_score=sp.minSeedLength * sp.matchScore)
                    E.set_rescue(out, ok)
                except native.HipswError:
                    fell_back.append("rescue")
                    E.rescue_cpu()
                engine.fault_inject(native.FAULT_NONE, 0)
            E.rescue_apply()
    jobs = E.pe_cigar_plan()
    regions = [CIGARRegion(bytes(j["qseg"].tolist()), bytes(j["rseg"].tolist()), j["qb"], j["qe"],
                           j["read_len"], j["rev"], j["true_score"], j["w"], j["ref_pos"])
               for j in jobs]
    armed("cigar")
    res = CIGARGenerator.generateBatch(regions, sp, engine) if regions else []
    engine.fault_inject(native.FAULT_NONE, 0)
    if len(res) != len(regions):
        fell_back.append("cigar")
        E.cigar_cpu()
    else:
        for j, r in enumerate(res):
            E.set_cigar(j, r.cigar, r.nm, r.md, r.pos)
    return E.pe_sam(names, list(q1) + list(q2)), fell_back


@pytest.mark.parametrize("phase", ["extension", "rescue", "cigar"])
def test_pe_sam_identical_after_fallback(engine, phase):
    """Paired-end SAM with one GPU phase failing (each site in turn) and the
    caller falling back to its CPU code for that phase: byte-identical to the
    all-CPU run."""
    ref, names, r1, r2, q, _, _ = pe_set(3000)
    cpu = e2e.E2E(ref, "chr5").run_pe_cpu(names, r1, r2, q, q)
    for site in SITES:
        print("pe fallback: phase %s site %d" % (phase, site), flush=True)
        sam, fell = _pe_run_with_faults(e2e.E2E(ref, "chr5"), names, r1, r2, q, q, engine,
                                        ScoringParameters(), phase, site)
        assert phase in fell, (phase, site, fell)
        assert sam == cpu, (phase, site)


def test_server_chunk_fault(engine):
    """A chunk whose run fails returns a negative status from wait; the next
    chunk on the same server runs clean and equals the oracle."""
    ref = serving.make_reference(9, 1000000, copies=100)
    srv = serving.Server(engine, slots=2, hbm_budget=4 << 30)
    try:
        srv.set_reference(serving.pack_pac(ref), len(ref))
        ch = serving.make_chunk(ref, 61, 800)
        for site in SITES:
            engine.fault_inject(site, 1)
            t = srv.submit(ch)
            with pytest.raises(native.HipswError):
                srv.wait(t)
            engine.fault_inject(native.FAULT_NONE, 0)
        res = srv.wait(srv.submit(ch))
        sets = orc.chunk_to_job_sets(ref, ch)
        ext_ref, _ = orc.ksw_extend_batch(sets["ext"], nthreads=8)
        assert np.array_equal(res["ext"], ext_ref)
    finally:
        srv.close()
