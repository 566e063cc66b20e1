"""Benchmark of the banded-extension hot path on one MI355X.

Headline workload: 1,000,000 independent banded-extension pairs (uniform, int16
score range): 150-bp queries, targets = source window + 0..100-bp flank mutated
at 1% substitutions / 0.1% indels, h0 ~ U[19,60], w = 100, z-drop 100, end
bonus 5, 2-bit packed and resident in HBM before timing starts.

A step is one pass of the extension kernels over the whole job set.  The
headline value is GCUPS of the batched persistent form; the plain one-launch
form (generic kernel, input order) and the input-order form (the batched
form's own kernel, input order, one work item per block) are timed beside it.
Cells are counted exactly as the CPU loop counts them (one per inner iteration
of BandedSWScalar.swift:112).

Other legs: the uneven 4,000,000-job set (binning + autotuner), banded global
NW + CIGAR on 500,000 pairs, mate-rescue local SW on 200,000 rescue windows,
the one-shot entry point with host buffers (PCIe included) and the serving
loop.  Every leg's results are downloaded: the forms of one leg must agree on
every row, and the cpu_baseline leg runs the CPU oracle over the whole job
sets (timing it) and compares every row with the GPU's.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
       python bench.py --profile-leg ext|mixed|nw|local|copy   (one leg, for rocprofv3)
"""
import argparse
import ctypes as C
import json
import os
import re
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

OPS_EXT = 11               # algorithmic VALU ops per extension cell
OPS_NW = 14                # per NW fill cell
OPS_LOC = 11               # per local-SW cell
PEAK_PK16_TOPS = 78.6      # 256 CU x 64 lanes x 2.4 GHz x 2 (packed int16)
PEAK_I32_TOPS = 39.3       # 32-bit VALU lane-ops
HBM_PEAK_BPS = 8.0e12
PMC_SUMMARY = os.path.join(ROOT, "profiles", "ext_uniform_1M_pmc.md")


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def pmc_traffic(lds_blocks):
    """HBM bytes per launch of the headline kernel from the committed
    rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE passes, each scaled by
    what the counter reads for a copy of known size), when that summary was
    taken with the layout the autotuner chose in this run; else None."""
    try:
        text = open(PMC_SUMMARY).read()
    except OSError:
        return None, "no pmc summary"
    t = re.search(r"traffic_bytes_per_launch:\s*(\d+)", text)
    lb = re.search(r"profiled_lds_blocks:\s*(\d+)", text)
    if not t or not lb:
        return None, "pmc summary has no traffic line"
    if int(lb.group(1)) != lds_blocks:
        return None, "pmc summary profiled %d LDS blocks, this run chose %d" % (int(lb.group(1)),
                                                                               lds_blocks)
    return int(t.group(1)), "profiles/ext_uniform_1M_pmc.md"


# ---------------------------------------------------------------------------
# device-resident extension sets
# ---------------------------------------------------------------------------
def ext_upload(eng, js, sp):
    from swiftbwa_amd import native
    from swiftbwa_amd.metalsw import _ext_arrays
    a = _ext_arrays(js)
    sc = sp.to_c()
    st = eng.lib.hipsw_ext_upload(eng.handle, js["n"], native.ptr(a["queries"]),
                                  native.ptr(a["q_off"]), native.ptr(a["q_len"]),
                                  native.ptr(a["targets"]), native.ptr(a["t_off"]),
                                  native.ptr(a["t_len"]), native.ptr(a["h0"]), native.ptr(a["w"]),
                                  native.ptr(a["end_bonus"]), C.byref(sc))
    if not st:
        raise RuntimeError(eng.last_error())
    return st


def ext_time(eng, st, mode, steps, warmup):
    """(wall ms/step, mean kernel ms, last stats) of `steps` timed passes."""
    from swiftbwa_amd import native
    s = native.HipswStats()
    for _ in range(warmup):
        if eng.lib.hipsw_ext_run(eng.handle, st, mode, C.byref(s)):
            raise RuntimeError(eng.last_error())
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        if eng.lib.hipsw_ext_run(eng.handle, st, mode, C.byref(s)):
            raise RuntimeError(eng.last_error())
        kms.append(s.kernel_ms)
    wall = (time.perf_counter() - t0) * 1000.0 / steps
    return wall, float(np.mean(kms)), s


def ext_download(eng, st, n):
    from swiftbwa_amd import native
    out = np.zeros((n, 6), np.int32)
    if eng.lib.hipsw_ext_download(eng.handle, st, native.ptr(out)):
        raise RuntimeError(eng.last_error())
    return out


def ext_properties(js, out):
    """Row invariants of BandedSWScalar.align on every row: 0 <= qle <= qlen,
    0 <= tle <= tlen, score >= h0 for non-empty jobs, gtle = 0 when gscore = -1,
    max_off >= 0.  Returns the number of violating rows."""
    ql = js["q_len"].astype(np.int64)
    tl = js["t_len"].astype(np.int64)
    nonempty = (ql > 0) & (tl > 0)
    bad = (out[:, 1] < 0) | (out[:, 1] > ql) | (out[:, 2] < 0) | (out[:, 2] > tl)
    bad |= nonempty & (out[:, 0] < js["h0"])
    bad |= (out[:, 4] == -1) & (out[:, 3] != 0)
    bad |= out[:, 5] < 0
    return int(bad.sum())


def autotune(eng, st, max_jobs):
    """hipsw_ext_autotune on an uploaded set: per 8-bit bin, the packed layout
    at each LDS/HBM split of the column state and the one-job-per-lane layout;
    per 16-bit bin, the one-job-per-lane layout at each split; timed on the
    bin's first max_jobs jobs."""
    from swiftbwa_amd import native
    rep = (native.HipswTuneEntry * 32)()
    nrep = C.c_int32()
    if eng.lib.hipsw_ext_autotune(eng.handle, st, int(max_jobs), rep, 32, C.byref(nrep)):
        raise RuntimeError(eng.last_error())
    classes = ["0-63", "64-159", "160-255", "256+"]
    kinds = {1: "packed (2 jobs/lane, int16x2)", 2: "one job/lane (int32)",
             3: "packed (2 jobs/lane, int16x2, 16-bit h/e planes)"}
    out = []
    tiers = {0: "8-bit", 1: "16-bit", 2: "32-bit"}
    for e in rep[:nrep.value]:
        tier = tiers.get(e.tier, str(e.tier))
        if not out or out[-1]["length_class"] != classes[e.length_class] or out[-1]["tier"] != tier:
            out.append({"tier": tier, "length_class": classes[e.length_class],
                        "jobs_timed": e.jobs, "candidates": []})
        c = {"kind": kinds.get(e.kind, str(e.kind)), "lds_columns": 16 * e.lds_blocks,
             "ms": round(e.ms, 3), "kind_id": int(e.kind), "lds_blocks": int(e.lds_blocks)}
        out[-1]["candidates"].append(c)
        if e.chosen:
            out[-1]["chosen"] = c
    for b in out:
        log("autotune %s %s: %s -> %s, %d LDS columns" % (
            b["tier"], b["length_class"],
            ", ".join("%s/%d %.3f ms" % (c["kind"].split(" ")[0], c["lds_columns"], c["ms"])
                      for c in b["candidates"]),
            b["chosen"]["kind"], b["chosen"]["lds_columns"]))
    return out


def ext_forms(eng, js, sp, args, tune_jobs, name):
    """Uploads js, autotunes, times plain / input-order / batched, downloads
    each form's rows and checks that they agree on every row."""
    from swiftbwa_amd import native
    t0 = time.perf_counter()
    st = ext_upload(eng, js, sp)
    upload_s = time.perf_counter() - t0
    tune = autotune(eng, st, tune_jobs)
    res = {"jobs": js["n"], "upload_s_incl_pack": round(upload_s, 3), "autotune": tune}
    outs = {}
    for form, mode in (("plain", native.MODE_PLAIN), ("input_order", native.MODE_INPUT_ORDER),
                       ("batched", native.MODE_BATCHED)):
        wall, kms, s = ext_time(eng, st, mode, args.steps, args.warmup)
        res[form] = {"ms_per_step": round(wall, 3), "kernel_ms": round(kms, 3),
                     "gcups": round(s.cells / (wall * 1e-3) / 1e9, 2),
                     "kernel_gcups": round(s.cells / (kms * 1e-3) / 1e9, 2),
                     "ext_per_s": round(js["n"] / (wall * 1e-3), 1), "launches": int(s.launches)}
        res.setdefault("cells", int(s.cells))
        if int(s.cells) != res["cells"]:
            raise RuntimeError("%s: cell counts differ between forms" % name)
        res["tier_jobs"] = list(s.tier_jobs)
        outs[form] = ext_download(eng, st, js["n"])
        log("%s %s: %.2f ms/step (kernel %.2f ms)" % (name, form, wall, kms))
    eng.lib.hipsw_ext_free(eng.handle, st)
    res["batched_over_plain"] = round(res["plain"]["ms_per_step"] / res["batched"]["ms_per_step"], 3)
    res["batched_over_input_order"] = round(
        res["input_order"]["ms_per_step"] / res["batched"]["ms_per_step"], 3)
    diff = sum(int((outs[f] != outs["batched"]).any(axis=1).sum()) for f in ("plain", "input_order"))
    res["rows_differing_between_forms"] = diff
    res["rows_violating_invariants"] = ext_properties(js, outs["batched"])
    return res, outs["batched"]


# ---------------------------------------------------------------------------
# legs
# ---------------------------------------------------------------------------
def leg_one_shot(eng, js, sp, ref_out, steps):
    """hipsw_banded_extend on host buffers: packing, upload, run and download
    inside the timed call (the PCIe-inclusive rate)."""
    from swiftbwa_amd import native
    st = {}
    got = eng.banded_extend(js, sp, mode=native.MODE_BATCHED, stats=st)     # warm
    walls = []
    for _ in range(max(1, steps)):
        t0 = time.perf_counter()
        got = eng.banded_extend(js, sp, mode=native.MODE_BATCHED, stats=st)
        walls.append(time.perf_counter() - t0)
    wall = float(np.median(walls))
    log("one-shot hipsw_banded_extend: %.1f ms per call (kernel %.2f ms)" % (wall * 1e3, st["kernel_ms"]))
    return {"calls_timed": len(walls), "ms_per_call": round(wall * 1e3, 2),
            "gcups_pcie_inclusive": round(st["cells"] / wall / 1e9, 2),
            "kernel_ms": round(st["kernel_ms"], 3),
            "rows_differing_from_resident_run": int((got != ref_out).any(axis=1).sum())}


def leg_nw(eng, args, modes=None):
    """Banded global NW + traceback: 500,000 pairs, qlen U[50,250], target =
    query +- a planted 0..20-bp indel, 1% subs / 0.2% indels; w = min(100,
    2|dlen|+3); score + CIGAR per pair, retries while score < planted score."""
    from swiftbwa_amd import native, workload as wl
    from swiftbwa_amd.metalsw import ScoringParameters, _ext_arrays
    lib = eng.lib
    js = wl.generate(wl.NW, 3, args.nw_n)
    a = _ext_arrays(js)
    ts = np.ascontiguousarray(js["true_score"], np.int32)
    sc = ScoringParameters().to_c()
    cap = 512
    st = lib.hipsw_nw_upload(eng.handle, js["n"], native.ptr(a["queries"]), native.ptr(a["q_off"]),
                             native.ptr(a["q_len"]), native.ptr(a["targets"]),
                             native.ptr(a["t_off"]), native.ptr(a["t_len"]), native.ptr(a["w"]),
                             native.ptr(ts), 3, cap, C.byref(sc))
    if not st:
        raise RuntimeError(eng.last_error())
    res, outs = {}, {}
    for name, mode in modes or (("plain", native.MODE_PLAIN), ("batched", native.MODE_BATCHED)):
        s = native.HipswStats()
        for _ in range(args.warmup):
            if lib.hipsw_nw_run(eng.handle, st, mode, C.byref(s)):
                raise RuntimeError(eng.last_error())
        kms = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            if lib.hipsw_nw_run(eng.handle, st, mode, C.byref(s)):
                raise RuntimeError(eng.last_error())
            kms.append(s.kernel_ms)
        wall = (time.perf_counter() - t0) * 1000.0 / args.steps
        res[name] = {"ms_per_step": round(wall, 3), "kernel_ms": round(float(np.mean(kms)), 3),
                     "gcups": round(s.cells / (wall * 1e-3) / 1e9, 2),
                     "kernel_gcups": round(s.cells / (np.mean(kms) * 1e-3) / 1e9, 2),
                     "launches": int(s.launches)}
        res["cells"] = int(s.cells)
        n = js["n"]
        o = {"score": np.zeros(n, np.int32), "cigar": np.zeros((n, cap), np.uint32),
             "n_cigar": np.zeros(n, np.int32), "flags": np.zeros(n, np.int32)}
        if lib.hipsw_nw_download(eng.handle, st, native.ptr(o["score"]), native.ptr(o["cigar"]),
                                 native.ptr(o["n_cigar"]), native.ptr(o["flags"]), None):
            raise RuntimeError(eng.last_error())
        outs[name] = o
        log("nw %s: %.2f ms/step (kernel %.2f ms)" % (name, wall, np.mean(kms)))
    lib.hipsw_nw_free(eng.handle, st)
    res["pairs"] = js["n"]
    if "plain" in res:
        res["batched_over_plain"] = round(res["plain"]["ms_per_step"] / res["batched"]["ms_per_step"], 3)
        p, b = outs["plain"], outs["batched"]
        res["rows_differing_between_forms"] = int(
            ((p["score"] != b["score"]) | (p["n_cigar"] != b["n_cigar"]) |
             (p["cigar"] != b["cigar"]).any(axis=1)).sum())
        del outs["plain"]
    return res, js, outs.get("batched")


def nw_narrow_set(js):
    """The pairs of an NW set whose band is at most 64 columns (GlobalAligner.swift:47-48)."""
    from swiftbwa_amd import workload as wl
    q = js["q_len"].astype(np.int64)
    t = js["t_len"].astype(np.int64)
    wp = np.maximum(np.minimum(js["w"].astype(np.int64), 1 << 20), np.abs(q - t))
    return wl.slice_jobs(js, np.nonzero(np.minimum(q, 2 * wp + 1) <= 64)[0])


def leg_nw_narrow(eng, args, js, forms=((0, "two_jobs_per_lane"), (2, "one_job_per_lane"))):
    """NW pairs with bands of at most 64 columns (367,300 of the 500,000): the
    default row-state form (two jobs per lane in int16 pairs) against one job
    per lane, batched, with identical results required."""
    from swiftbwa_amd import native
    from swiftbwa_amd.metalsw import ScoringParameters, _ext_arrays
    lib = eng.lib
    a = _ext_arrays(js)
    ts = np.ascontiguousarray(js["true_score"], np.int32)
    sc = ScoringParameters().to_c()
    cap = 512
    st = lib.hipsw_nw_upload(eng.handle, js["n"], native.ptr(a["queries"]), native.ptr(a["q_off"]),
                             native.ptr(a["q_len"]), native.ptr(a["targets"]),
                             native.ptr(a["t_off"]), native.ptr(a["t_len"]), native.ptr(a["w"]),
                             native.ptr(ts), 3, cap, C.byref(sc))
    if not st:
        raise RuntimeError(eng.last_error())
    res, outs = {"pairs": js["n"]}, {}
    try:
        for form, name in forms:
            eng.set_nw_state(form)
            s = native.HipswStats()
            kms = []
            for k in range(args.warmup + args.steps):
                if lib.hipsw_nw_run(eng.handle, st, native.MODE_BATCHED, C.byref(s)):
                    raise RuntimeError(eng.last_error())
                if k >= args.warmup:
                    kms.append(s.kernel_ms)
            n = js["n"]
            o = {"score": np.zeros(n, np.int32), "cigar": np.zeros((n, cap), np.uint32),
                 "n_cigar": np.zeros(n, np.int32), "flags": np.zeros(n, np.int32)}
            if lib.hipsw_nw_download(eng.handle, st, native.ptr(o["score"]), native.ptr(o["cigar"]),
                                     native.ptr(o["n_cigar"]), native.ptr(o["flags"]), None):
                raise RuntimeError(eng.last_error())
            outs[name] = o
            res["cells"] = int(s.cells)
            res[name] = {"form": form, "kernel_ms": round(float(np.mean(kms)), 3),
                         "kernel_gcups": round(s.cells / (np.mean(kms) * 1e-3) / 1e9, 2)}
            log("nw narrow (form %d): kernel %.3f ms" % (form, np.mean(kms)))
    finally:
        eng.set_nw_state(0)
        lib.hipsw_nw_free(eng.handle, st)
    if len(outs) == 2:
        x, y = list(outs.values())
        res["rows_differing_between_forms"] = int(
            ((x["score"] != y["score"]) | (x["n_cigar"] != y["n_cigar"]) |
             (x["cigar"] != y["cigar"]).any(axis=1)).sum())
        n0, n1 = [f[1] for f in forms]
        res["speedup"] = round(res[n1]["kernel_ms"] / res[n0]["kernel_ms"], 3)
    return res


def leg_local(eng, args):
    """Mate-rescue local SW: 200,000 150-bp mates against 600-bp windows
    (score2 on, minSubScore 19), one-shot hipsw_local_sw calls."""
    from swiftbwa_amd import workload as wl
    from swiftbwa_amd.metalsw import ScoringParameters
    js = wl.generate(wl.LOCAL, 5, args.local_n)
    sp = ScoringParameters()
    st = {}
    out, ok = eng.local_sw(js, sp, min_sub_score=19, stats=st)            # warm
    walls, kms = [], []
    for _ in range(max(1, args.steps // 4)):
        t0 = time.perf_counter()
        out, ok = eng.local_sw(js, sp, min_sub_score=19, stats=st)
        walls.append(time.perf_counter() - t0)
        kms.append(st["kernel_ms"])
    wall, km = float(np.median(walls)), float(np.mean(kms))
    log("local SW: %.1f ms per call (kernel %.2f ms)" % (wall * 1e3, km))
    return ({"tasks": js["n"], "ms_per_call_pcie_inclusive": round(wall * 1e3, 2),
             "kernel_ms": round(km, 3), "cells": int(st["cells"]),
             "kernel_gcups": round(st["cells"] / (km * 1e-3) / 1e9, 2),
             "scalar_fallbacks": int(st["promoted"])}, js, (out, ok))


def leg_serving(eng, args):
    """Serving loop: 250 Mbp reference with planted repeat families resident in
    HBM, split into 24 sequences; chunks of 2x150-bp read pairs (log-uniform
    1k..200k pairs) carrying their extension, mate-rescue (raw windows the
    server clamps to the template's sequence, fetchSequence) and CIGAR jobs;
    Poisson arrivals at 25/50/75/90% of the measured saturation throughput;
    p50/p99 of arrival -> results ready.  reads/s counts aligned reads: those
    whose extensions reach the minimum output score (30) from the device's
    results."""
    from swiftbwa_amd import serving
    t0 = time.perf_counter()
    ref = serving.make_reference(5, args.serve_ref, copies=1000, divergence=0.10)
    pac = serving.pack_pac(ref)
    lens = serving.contig_lengths(len(ref), 24)
    offs = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64)
    rng = np.random.default_rng(5)
    sizes = np.exp(rng.uniform(np.log(1000), np.log(200000), args.serve_chunks)).astype(int)
    pool = [serving.make_chunk(ref, 1000 + k, int(n)) for k, n in enumerate(sizes)]
    for ch in pool:                       # the template's sequence: that of the window middle
        loc = ch["loc"]
        loc["rid"] = (np.searchsorted(offs, loc["t_pos"] + loc["t_len"] // 2, side="right") - 1).astype(np.int32)
    log("serving: reference %d bp in %d sequences, %d chunks (%d pairs) generated in %.1f s" % (
        len(ref), len(lens), len(pool), int(sizes.sum()), time.perf_counter() - t0))
    aligned = np.zeros(len(pool), np.int64)
    clamped = np.zeros(len(pool), np.int64)

    def count_aligned(k, r):
        ext = pool[k]["ext"]
        score = r["ext"][:, 0]
        ok = (score >= 30) & (score >= ext["a"][:ext["n"]])   # reached minOutputScore, kept h0
        per_read = np.zeros(2 * pool[k]["pairs"], bool)
        per_read[ext["r"][:ext["n"]][ok]] = True
        aligned[k] = int(per_read.sum())
        clamped[k] = int((r["loc_win"] != pool[k]["loc"]["t_pos"]).sum())

    srv = serving.Server(eng, slots=args.serve_slots, hbm_budget=args.serve_budget_gb << 30)
    try:
        srv.set_reference(pac, len(ref))
        srv.set_contigs(lens)
        serving.run_poisson(srv, pool, 0.0)       # warm: pinned staging and pool at size
        recs, wall = serving.run_poisson(srv, pool, 0.0, on_result=count_aligned)
        sat = sizes.sum() / wall
        frac_aligned = aligned.sum() / (2 * sizes.sum())
        log("serving saturation: %.0f pairs/s, %.4f of reads aligned, %d rescue windows clamped or "
            "skipped" % (sat, frac_aligned, int(clamped.sum())))
        mean_pairs = sizes.mean()
        loads = []
        for rho in (0.25, 0.5, 0.75, 0.9):
            rate = rho * sat / mean_pairs
            recs, wall = serving.run_poisson(srv, pool, rate, seed=int(rho * 100))
            lat = np.array([r["latency_ms"] for r in recs])
            loads.append({"offered_load": rho, "chunks_per_s": round(rate, 3),
                          "p50_ms": round(float(np.percentile(lat, 50)), 2),
                          "p99_ms": round(float(np.percentile(lat, 99)), 2),
                          "aligned_reads_per_s": round(float(aligned.sum()) / wall, 1),
                          "mean_queue_ms": round(float(np.mean([r["queue_ms"] for r in recs])), 2)})
            log("serving load %.2f: p50 %.1f ms p99 %.1f ms" % (rho, loads[-1]["p50_ms"], loads[-1]["p99_ms"]))
        st = srv.stats()
    finally:
        srv.close()
    return {"workload": "synthetic job chunks (4 extensions, 2 NW, 1 rescue per pair)",
            "reference_bp": len(ref), "sequences": len(lens), "chunks": len(pool),
            "pairs_per_pass": int(sizes.sum()), "aligned_fraction": round(float(frac_aligned), 5),
            "rescue_windows_clamped_or_skipped": int(clamped.sum()),
            "slots": args.serve_slots, "hbm_budget_gb": args.serve_budget_gb,
            "peak_admitted_bytes": int(st["peak_in_use"]),
            "saturation_pairs_per_s": round(float(sat), 1),
            "saturation_aligned_reads_per_s": round(float(aligned.sum()) / (sizes.sum() / sat), 1),
            "loads": loads}


def cpu_baseline(sets, threads):
    """The C oracle (test infrastructure, a restatement of the reference's CPU
    loops) built -O3 -march=native on this host, timed on this GPU's share of
    host cores over the whole job sets of the legs, and used as the checker:
    every GPU row of every leg is compared with the oracle's row."""
    from oracle import pyoracle as orc
    from swiftbwa_amd import workload as wl
    tmp = tempfile.mkdtemp(prefix="swor_native_")
    path = orc.build_native(tmp)
    build = "gcc -O3 -march=native (built on this host)"
    if path:
        orc.use_library(path)
    else:
        build = "prebuilt -O3 -march=x86-64-v3 (no C compiler on this host)"
    res = {"kind": "port", "cores": threads, "build": build, "cpu": cpu_model(),
           "host_cpus_visible": os.cpu_count()}
    checks = {}
    js, gpu = sets["uniform"]
    t0 = time.perf_counter()
    ref, cells = orc.ksw_extend_batch(js, nthreads=threads)
    dt = time.perf_counter() - t0
    checks["uniform_1M"] = {"rows": js["n"], "mismatches": int((ref != gpu).any(axis=1).sum())}
    res.update(value=round(cells / dt / 1e9, 4), unit="GCUPS",
               sample="all %d jobs of the 1,000,000-pair set (%d cells) on %d threads, %.2f s"
                      % (js["n"], cells, threads, dt),
               ext_per_s=round(js["n"] / dt, 1))
    idx1 = np.arange(min(50000, js["n"]))
    t1 = time.perf_counter()
    _, cells1 = orc.ksw_extend_batch(wl.slice_jobs(js, idx1), nthreads=1)
    dt1 = time.perf_counter() - t1
    res["one_thread"] = {"value": round(cells1 / dt1 / 1e9, 4), "unit": "GCUPS", "cores": 1,
                         "sample": "first %d jobs (%d cells), %.2f s" % (len(idx1), cells1, dt1)}
    legs = {}
    if "mixed" in sets:
        js, gpu = sets["mixed"]
        t0 = time.perf_counter()
        ref, cells = orc.ksw_extend_batch(js, nthreads=threads)
        dt = time.perf_counter() - t0
        checks["uneven_4M"] = {"rows": js["n"], "mismatches": int((ref != gpu).any(axis=1).sum())}
        legs["uneven_4M"] = {"gcups": round(cells / dt / 1e9, 4), "s": round(dt, 2)}
    if "nw" in sets:
        js, gpu = sets["nw"]
        t0 = time.perf_counter()
        score, cig, ncig, cells = orc.global_batch(js, max_retries=3, cap=512, nthreads=threads)
        dt = time.perf_counter() - t0
        used = np.arange(cig.shape[1])[None, :] < np.minimum(ncig, cig.shape[1])[:, None]
        bad = (score != gpu["score"]) | (ncig != gpu["n_cigar"]) | ((cig != gpu["cigar"]) & used).any(axis=1)
        checks["nw_500k"] = {"rows": js["n"], "mismatches": int(bad.sum())}
        legs["nw_500k"] = {"gcups": round(cells / dt / 1e9, 4), "s": round(dt, 2)}
    if "local" in sets:
        js, (gout, gok) = sets["local"]
        t0 = time.perf_counter()
        out, ok, cells = orc.local_batch(js, min_sub_score=19, nthreads=threads)
        dt = time.perf_counter() - t0
        out[ok == 0] = 0
        bad = (ok != gok) | (out != gout).any(axis=1)
        checks["local_200k"] = {"rows": js["n"], "mismatches": int(bad.sum())}
        legs["local_200k"] = {"gcups": round(cells / dt / 1e9, 4), "s": round(dt, 2)}
    res["other_legs"] = legs
    res["checks"] = checks
    for k, v in checks.items():
        log("oracle check %s: %d rows, %d mismatches" % (k, v["rows"], v["mismatches"]))
    return res


def profile_leg(args):
    """Runs one leg alone so a profiler sees only its kernels.

    ext:   the headline job set, batched form, `warmup` + `steps` passes.
    mixed: the uneven-size job set (4,000,000 jobs), batched form, likewise.
    nw:    the NW + CIGAR set (500,000 pairs), batched form, likewise.
    nw_narrow: its pairs with bands of at most 64 columns, default form, likewise.
    local: the 200,000 rescue windows, `steps` one-shot calls.
    copy:  a 256 MiB device-to-device copy, `steps` times -- the known-size
           reference that scales the FETCH_SIZE / WRITE_SIZE counters."""
    from swiftbwa_amd import native, workload as wl
    from swiftbwa_amd.metalsw import HipSWEngine, ScoringParameters
    eng = HipSWEngine(0)
    lib = eng.lib
    if args.pk_lds_blocks:
        # pin the packed layout of the 64-159 class (what the autotuner picked
        # in the benchmark run) so the profiled kernel is the benchmarked one
        if lib.hipsw_ext_set_layout(eng.handle, 1, 1, args.pk_lds_blocks):
            raise RuntimeError(eng.last_error())
    if args.profile_leg == "nw":
        leg_nw(eng, args, modes=(("batched", native.MODE_BATCHED),))
    elif args.profile_leg == "nw_narrow":
        leg_nw_narrow(eng, args, nw_narrow_set(wl.generate(wl.NW, 3, args.nw_n)),
                      forms=((0, "two_jobs_per_lane"),))
    elif args.profile_leg == "local":
        js = wl.generate(wl.LOCAL, 5, args.local_n)
        for k in range(args.steps):
            st = {}
            eng.local_sw(js, ScoringParameters(), min_sub_score=19, stats=st)
            log("local call %d: kernel %.3f ms" % (k, st["kernel_ms"]))
    elif args.profile_leg in ("ext", "mixed"):
        js = (wl.generate(wl.EXT_UNIFORM, 2, args.n) if args.profile_leg == "ext"
              else wl.generate(wl.EXT_MIXED, 4, args.mixed_n))
        st = ext_upload(eng, js, ScoringParameters())
        if args.profile_leg == "mixed" and args.tune:
            autotune(eng, st, 100000)
        s = native.HipswStats()
        for k in range(args.warmup + args.steps):
            if lib.hipsw_ext_run(eng.handle, st, native.MODE_BATCHED, C.byref(s)):
                raise RuntimeError(eng.last_error())
            log("%s batched pass %d: kernel %.3f ms, %d cells" % (args.profile_leg, k, s.kernel_ms,
                                                                  s.cells))
        lib.hipsw_ext_free(eng.handle, st)
    else:
        hip = C.CDLL("libamdhip64.so")
        nb = 256 << 20
        src, dst = C.c_void_p(), C.c_void_p()
        if hip.hipMalloc(C.byref(src), C.c_size_t(nb)) or hip.hipMalloc(C.byref(dst), C.c_size_t(nb)):
            raise RuntimeError("hipMalloc failed")
        hip.hipMemset(src, 1, C.c_size_t(nb))
        for k in range(args.steps):
            if hip.hipMemcpy(dst, src, C.c_size_t(nb), 3):     # device to device
                raise RuntimeError("hipMemcpy failed")
        hip.hipDeviceSynchronize()
        log("copy: %d x %d bytes" % (args.steps, nb))
        hip.hipFree(src)
        hip.hipFree(dst)
    eng.close()


def roofline(cells, kernel_ms, ops, peak, unit="TOP/s", **extra):
    achieved = cells * ops / (kernel_ms * 1e-3) / 1e12
    d = {"bound": "valu", "achieved": round(achieved, 3), "peak": peak, "unit": unit,
         "frac": round(achieved / peak, 4), "ops_per_cell": ops}
    d.update(extra)
    return d


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1000000)
    ap.add_argument("--cpu-threads", type=int, default=16,
                    help="host threads of the CPU baseline (this GPU's share of the host)")
    ap.add_argument("--nw-n", type=int, default=500000)
    ap.add_argument("--local-n", type=int, default=200000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-ref", type=int, default=250000000)
    ap.add_argument("--serve-chunks", type=int, default=24)
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-budget-gb", type=int, default=64)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--mixed-n", type=int, default=4000000)
    ap.add_argument("--profile-leg", choices=["ext", "mixed", "nw", "nw_narrow", "local", "copy"],
                    default=None)
    ap.add_argument("--tune", action="store_true", help="profile leg mixed: autotune first")
    ap.add_argument("--pk-lds-blocks", type=int, default=0,
                    help="profile leg: 16-column blocks of packed column state in LDS for "
                         "64-159-base queries (10, 6 or 5; 0 = default)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: measures one GPU only (each job set runs on one "
              "device; a multi-GPU run would shard the job set by read chunk, one engine per "
              "GPU). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.profile_leg:
        profile_leg(args)
        return

    from swiftbwa_amd import workload as wl
    from swiftbwa_amd.metalsw import HipSWEngine, ScoringParameters

    log("generating %d jobs" % args.n)
    js = wl.generate(wl.EXT_UNIFORM, 2, args.n)
    eng = HipSWEngine(0)
    dev_name, n_cu = eng.device_info()
    sp = ScoringParameters()
    log("uniform leg on %s" % dev_name)
    uni, uni_out = ext_forms(eng, js, sp, args, 200000, "uniform")
    cells = uni["cells"]
    b_wall, b_kms = uni["batched"]["ms_per_step"], uni["batched"]["kernel_ms"]
    one_shot = leg_one_shot(eng, js, sp, uni_out, max(1, args.steps // 4))

    sets = {"uniform": (js, uni_out)}
    nw = mixed = local = None
    if args.nw_n > 0:
        nw, nw_js, nw_out = leg_nw(eng, args)
        sets["nw"] = (nw_js, nw_out)
        nw["narrow_bands_le64"] = leg_nw_narrow(eng, args, nw_narrow_set(nw_js))
    if args.mixed_n > 0:
        mjs = wl.generate(wl.EXT_MIXED, 4, args.mixed_n)
        mixed, mixed_out = ext_forms(eng, mjs, sp, args, 100000, "mixed")
        sets["mixed"] = (mjs, mixed_out)
    if args.local_n > 0:
        local, ljs, lout = leg_local(eng, args)
        sets["local"] = (ljs, lout)
    serve = None if args.no_serve else leg_serving(eng, args)

    # headline kernel and its layout (the autotuner's choice for 64-159 bases)
    chosen = next((b["chosen"] for b in uni["autotune"] if b["length_class"] == "64-159"), None)
    lb = chosen["lds_blocks"] if chosen else 10
    kname = ("ext_persistent_pk_kernel<SYM, NQ=10, LB=%d>" % lb if not chosen or chosen["kind_id"] == 1
             else "ext_persistent_ln_kernel<SYM, NQ=10, LB=%d>" % lb)
    traffic, traffic_src = pmc_traffic(lb)
    qb = (js["q_len"].astype(np.int64) + 15) // 16 * 4
    tb = (js["t_len"].astype(np.int64) + 15) // 16 * 4
    algo_bytes = int(qb.sum() + tb.sum()) + js["n"] * (32 + 24)
    head = roofline(cells, b_kms, OPS_EXT, PEAK_PK16_TOPS, peak_int32=PEAK_I32_TOPS,
                    kernel=kname, traffic=traffic, traffic_source=traffic_src,
                    algorithmic_bytes=algo_bytes,
                    hbm_frac_algorithmic=round(algo_bytes / (b_kms * 1e-3) / HBM_PEAK_BPS, 6))
    head["frac_int32"] = round(head["achieved"] / PEAK_I32_TOPS, 4)
    if traffic:
        head["hbm_frac_measured"] = round(traffic / (b_kms * 1e-3) / HBM_PEAK_BPS, 4)
    head["note"] = ("VALU-bound integer DP: cells x 11 algorithmic ops over the packed-int16 peak; "
                    "traffic = measured HBM bytes per launch (pmc), hbm_frac_measured its share "
                    "of 8 TB/s over the kernel time")
    others = {}
    if mixed:
        others["uneven_4M_batched"] = roofline(mixed["cells"], mixed["batched"]["kernel_ms"], OPS_EXT,
                                               PEAK_PK16_TOPS, kernel="binned persistent bins")
    if nw:
        others["nw_500k_batched"] = roofline(nw["cells"], nw["batched"]["kernel_ms"], OPS_NW,
                                             PEAK_I32_TOPS,
                                             kernel="nw_reg_kernel<96> + nw_wide_kernel (retry passes, wide bands)",
                                             note="int32 VALU peak: the fill runs 32-bit arithmetic on "
                                                  "16-bit state")
        nn = nw.get("narrow_bands_le64")
        if nn and "two_jobs_per_lane" in nn:
            others["nw_narrow_bands_two_jobs_per_lane"] = roofline(
                nn["cells"], nn["two_jobs_per_lane"]["kernel_ms"], OPS_NW, PEAK_PK16_TOPS,
                kernel="nw_pk_kernel<64> + nw_wide_kernel (retry passes, wide bands)",
                note="int16-pair arithmetic: packed-int16 VALU peak")
    if local:
        others["local_sw_200k"] = roofline(local["cells"], local["kernel_ms"], OPS_LOC, PEAK_I32_TOPS,
                                           kernel="local_striped_kernel<10> (+ scalar fallback)",
                                           note="u8 cells in 32-bit lanes: int32 VALU peak")
    rec = {
        "metric": "banded-SW GCUPS (batched persistent form, 1 GPU)",
        "value": round(cells / (b_wall * 1e-3) / 1e9, 2),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(b_wall, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16x2 (two jobs per lane in packed 16-bit halves; 8-bit h/e state)",
        "data": "synthetic",
        "config": {
            "workload": "uniform banded extension, int16 range: 1,000,000 pairs, q=150 bp, "
                        "t=150..250 bp (+indels), h0~U[19,60], w=100, zdrop=100, end bonus 5",
            "jobs": js["n"], "cells": int(cells),
            "ext_per_s": round(js["n"] / (b_wall * 1e-3), 1),
            "kernel_ms": b_kms,
            "uniform_1M": uni,
            "one_shot_host_buffers": one_shot,
            "device": dev_name, "n_cu": n_cu,
            "nw_cigar": nw,
            "uneven_batching": mixed,
            "local_sw": local,
            "serving": serve,
        },
        "roofline": head,
        "roofline_other_kernels": others,
    }
    ok = uni["rows_differing_between_forms"] == 0 and uni["rows_violating_invariants"] == 0
    ok &= one_shot["rows_differing_from_resident_run"] == 0
    if mixed:
        ok &= mixed["rows_differing_between_forms"] == 0 and mixed["rows_violating_invariants"] == 0
    if nw:
        ok &= nw.get("rows_differing_between_forms", 0) == 0
        ok &= nw["narrow_bands_le64"].get("rows_differing_between_forms", 0) == 0
    if not args.no_cpu:
        log("cpu baseline and oracle checks (%d threads)" % args.cpu_threads)
        rec["cpu_baseline"] = cpu_baseline(sets, args.cpu_threads)
        ok &= all(v["mismatches"] == 0 for v in rec["cpu_baseline"]["checks"].values())
    rec["verified"] = bool(ok)
    eng.close()
    print(json.dumps(rec), flush=True)
    if not ok:
        log("RESULT CHECK FAILED: see rows_differing / mismatches fields")
        sys.exit(1)


if __name__ == "__main__":
    main()
