"""Config 5 end to end: paired reads -> serving loop -> SAM on the 250-Mbp
repeat reference (a measurement script under tests/, since the seeding,
chaining, pairing and SAM steps are the oracle's harness; not collected).

    python -m tests.e2e_config5 [--ref-bp 250000000] [--chunks 12] [--out file]

Reference: random ACGT with planted 300-bp and 6-kb families (1,000 copies
each at 10% divergence), split into 24 sequences (serving.contig_lengths),
indexed on the host by the harness.  Read chunks: 2 x 150-bp pairs, insert
N(400, 50), 0.5% substitutions, 0.05% indels, 5% of read 2s placeable only by
rescue; sizes log-uniform in 1k..200k pairs.  Each chunk is one batch of the
reference's paired-end path; its chain-extension passes, rescue local SW
(windows clamped on the server) and CIGAR NW run on the GPU through one
4-slot server, all chunks in flight together (saturation) and then as a
Poisson stream at half the saturation rate.  The SAM of the --check smallest
chunks (all of them by default) is compared byte for byte with the harness's
all-CPU run.
"""
import argparse
import json
import sys
import threading
import time

import numpy as np

from oracle import e2e
from swiftbwa_amd import serving, workload as wl
from swiftbwa_amd.metalsw import HipSWEngine, ScoringParameters
from tests import serving_pipeline as SP


def log(msg):
    print("[config5] " + msg, file=sys.stderr, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref-bp", type=int, default=250_000_000)
    ap.add_argument("--sequences", type=int, default=24)
    ap.add_argument("--chunks", type=int, default=12)
    ap.add_argument("--check", type=int, default=12)
    ap.add_argument("--load", type=float, default=0.5)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    t0 = time.perf_counter()
    ref = serving.make_reference(5, a.ref_bp, copies=1000, divergence=0.10)
    lens = serving.contig_lengths(a.ref_bp, a.sequences)
    contigs = [("chr%d" % (k + 1), n) for k, n in enumerate(lens)]
    log("reference %d bp in %d sequences (%.1f s)" % (a.ref_bp, len(lens), time.perf_counter() - t0))
    t1 = time.perf_counter()
    E = e2e.E2E(ref, contigs=contigs)
    t_index = time.perf_counter() - t1
    log("index built in %.1f s" % t_index)
    rng = np.random.default_rng(5)
    sizes = np.exp(rng.uniform(np.log(1000), np.log(200000), a.chunks)).astype(int)
    chunks = []
    for k, n in enumerate(sizes):
        b1, b2, _, _ = wl.pe_reads(ref, 100 + k, int(n), hard=0.05)
        chunks.append((["c%dp%d" % (k, i) for i in range(n)], list(b1), list(b2),
                       [b"I" * b1.shape[1]] * int(n)))
    log("%d chunks, %d pairs" % (len(chunks), int(sizes.sum())))
    sp = ScoringParameters()
    eng = HipSWEngine(0)
    srv = serving.Server(eng, slots=4, hbm_budget=32 << 30, scoring=sp)
    srv.set_reference(serving.pack_pac(ref), len(ref))
    srv.set_contigs(lens)
    del ref

    def run(arrivals):
        sams = [None] * len(chunks)
        lat = [0.0] * len(chunks)
        stage_logs = [[] for _ in chunks]
        errs = []
        start = time.perf_counter()

        def one(k):
            d = start + arrivals[k] - time.perf_counter()
            if d > 0:
                time.sleep(d)
            names, r1, r2, q = chunks[k]
            F = E.fork()
            try:
                ts = time.perf_counter()
                sams[k] = SP.run_pe_served(F, srv, names, r1, r2, q, q, sp, stage_logs[k])
                lat[k] = time.perf_counter() - ts
                log("chunk %d (%d pairs) done in %.2f s" % (k, len(names), lat[k]))
            except Exception as ex:
                errs.append(ex)
            finally:
                F.close()

        ths = [threading.Thread(target=one, args=(k,)) for k in range(len(chunks))]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        if errs:
            raise errs[0]
        return sams, np.array(lat), time.perf_counter() - start, stage_logs

    res = {"reference_bp": a.ref_bp, "sequences": len(lens), "index_s": round(t_index, 1),
           "chunks": len(chunks), "pairs": int(sizes.sum()), "chunk_pairs": [int(x) for x in sizes]}
    sams, lat, wall, stage_logs = run([0.0] * len(chunks))
    aligned = sum(sum(1 for line in s.splitlines() if not int(line.split("\t")[1]) & 0x904)
                  for s in sams)
    gpu_stage = [l for lg in stage_logs for _, l, _ in lg]
    res["saturation"] = {"wall_s": round(wall, 2), "reads_per_s": round(2 * sizes.sum() / wall, 1),
                         "aligned_reads_per_s": round(aligned / wall, 1),
                         "aligned_fraction": round(aligned / (2 * sizes.sum()), 5),
                         "chunk_p50_s": round(float(np.percentile(lat, 50)), 3),
                         "chunk_p99_s": round(float(np.percentile(lat, 99)), 3),
                         "server_stage_p50_ms": round(float(np.percentile(gpu_stage, 50)), 2),
                         "server_stage_p99_ms": round(float(np.percentile(gpu_stage, 99)), 2),
                         "server_cells": int(sum(c for lg in stage_logs for _, _, c in lg))}
    log("saturation: %s" % json.dumps(res["saturation"]))
    rate = a.load * len(chunks) / wall
    arrivals = np.cumsum(rng.exponential(1.0 / rate, len(chunks)))
    arrivals -= arrivals[0]
    sams2, lat2, wall2, _ = run(list(arrivals))
    res["poisson"] = {"offered_load": a.load, "chunks_per_s": round(rate, 3),
                      "chunk_p50_s": round(float(np.percentile(lat2, 50)), 3),
                      "chunk_p99_s": round(float(np.percentile(lat2, 99)), 3),
                      "wall_s": round(wall2, 2)}
    log("poisson: %s" % json.dumps(res["poisson"]))
    srv.close()
    same_runs = all(x == y for x, y in zip(sams, sams2))
    checked = []
    for k in np.argsort(sizes)[:a.check]:
        names, r1, r2, q = chunks[k]
        F = E.fork()
        t2 = time.perf_counter()
        cpu = F.run_pe_cpu(names, r1, r2, q, q)
        F.close()
        checked.append({"chunk": int(k), "pairs": int(sizes[k]), "identical": cpu == sams[k],
                        "cpu_s": round(time.perf_counter() - t2, 2)})
        log("check chunk %d (%d pairs): %s" % (k, sizes[k], "identical" if cpu == sams[k] else "DIFFERS"))
    res["sam_checks"] = checked
    res["saturation_equals_poisson_run"] = bool(same_runs)
    out = json.dumps(res)
    print(out, flush=True)
    if a.out:
        with open(a.out, "w") as f:
            f.write(out + "\n")
    return 0 if same_runs and all(c["identical"] for c in checked) else 1


if __name__ == "__main__":
    sys.exit(main())
