"""Serving loop on the GPU: results of a chunk are the same alone and in a mix of
concurrently running chunks, equal to the oracle, and admission honours the
HBM budget."""
import numpy as np
import pytest

from oracle import pyoracle as orc
from swiftbwa_amd import native, serving

pytestmark = pytest.mark.gpu
KEYS = ("ext", "loc", "loc_ok", "nw_score", "nw_n_cigar", "nw_flags", "nw_cigar")


@pytest.fixture(scope="module")
def ref():
    return serving.make_reference(5, 2000000, copies=200)


def test_alone_vs_mix_and_oracle(engine, ref):
    srv = serving.Server(engine, slots=4, hbm_budget=8 << 30)
    try:
        srv.set_reference(serving.pack_pac(ref), len(ref))
        target = serving.make_chunk(ref, 100, 3000)
        others = [serving.make_chunk(ref, 200 + k, n) for k, n in enumerate([5000, 1000, 8000, 2000, 4000])]
        alone = srv.wait(srv.submit(target))
        tickets = [srv.submit(o) for o in others[:3]] + [srv.submit(target)] + \
                  [srv.submit(o) for o in others[3:]]
        res = [srv.wait(t) for t in tickets]
        mix = res[3]
        for k in KEYS:
            assert np.array_equal(alone[k], mix[k]), k
        sets = orc.chunk_to_job_sets(ref, target)
        ext_ref, _ = orc.ksw_extend_batch(sets["ext"], nthreads=8)
        assert np.array_equal(alone["ext"], ext_ref)
        loc_ref, lok, _ = orc.local_batch(sets["loc"], min_sub_score=19, nthreads=8)
        loc_ref[lok == 0] = 0
        assert np.array_equal(alone["loc_ok"], lok)
        assert np.array_equal(alone["loc"], loc_ref)
        rs, rc, rn, _ = orc.global_batch(sets["nw"], max_retries=3, cap=512, nthreads=8)
        assert np.array_equal(alone["nw_score"], rs)
        assert np.array_equal(alone["nw_n_cigar"], rn)
        for k in range(sets["nw"]["n"]):
            m = min(rn[k], 512)
            assert np.array_equal(alone["nw_cigar"][k, :m], rc[k, :m]), k
        # the compact CIGAR form: the same used ops back to back, with offsets
        comp = srv.wait(srv.submit(target), compact=True)
        for k in KEYS[:-1]:
            assert np.array_equal(comp[k], alone[k]), k
        off = comp["nw_cigar_off"]
        assert off[0] == 0 and len(comp["nw_cigar_ops"]) == off[-1]
        for k in range(sets["nw"]["n"]):
            m = min(rn[k], 512)
            assert off[k + 1] - off[k] == m, k
            assert np.array_equal(comp["nw_cigar_ops"][off[k]:off[k + 1]], rc[k, :m]), k
        # every chunk of the mix also matches its own oracle on extension
        for o, r in zip(others[:3] + [target] + others[3:], res):
            s = orc.chunk_to_job_sets(ref, o)
            e, _ = orc.ksw_extend_batch(s["ext"], nthreads=8)
            assert np.array_equal(r["ext"], e)
    finally:
        srv.close()


def test_budget_admission(engine, ref):
    budget = 200 << 20
    srv = serving.Server(engine, slots=4, hbm_budget=budget)
    try:
        srv.set_reference(serving.pack_pac(ref), len(ref))
        chunks = [serving.make_chunk(ref, 300 + k, 1500) for k in range(6)]
        ts = [srv.submit(c) for c in chunks]
        res = [srv.wait(t) for t in ts]
        st = srv.stats()
        assert st["peak_in_use"] <= budget
        assert st["running"] == 0 and st["queued"] == 0
        alone = srv.wait(srv.submit(chunks[2]))
        for k in KEYS:
            assert np.array_equal(alone[k], res[2][k]), k
        with pytest.raises(native.HipswError):
            srv.submit(serving.make_chunk(ref, 999, 20000))
    finally:
        srv.close()


def test_queries_with_n(engine, ref):
    # ambiguous query bases: the chunk's N masks, the extension kernels' N
    # path and the masked local-SW / NW reads, against the oracle
    srv = serving.Server(engine, slots=2, hbm_budget=4 << 30)
    try:
        srv.set_reference(serving.pack_pac(ref), len(ref))
        ch = serving.make_chunk(ref, 400, 2000)
        rng = np.random.default_rng(4)
        for kind in ("ext", "loc", "nw"):
            q = ch[kind]["q"]
            q[rng.random(len(q)) < 0.01] = 4
        res = srv.wait(srv.submit(ch))
        sets = orc.chunk_to_job_sets(ref, ch)
        ext_ref, _ = orc.ksw_extend_batch(sets["ext"], nthreads=8)
        assert np.array_equal(res["ext"], ext_ref)
        loc_ref, lok, _ = orc.local_batch(sets["loc"], min_sub_score=19, nthreads=8)
        loc_ref[lok == 0] = 0
        assert np.array_equal(res["loc_ok"], lok)
        assert np.array_equal(res["loc"], loc_ref)
        rs, rc, rn, _ = orc.global_batch(sets["nw"], max_retries=3, cap=512, nthreads=8)
        assert np.array_equal(res["nw_score"], rs)
        assert np.array_equal(res["nw_n_cigar"], rn)
        for k in range(sets["nw"]["n"]):
            m = min(rn[k], 512)
            assert np.array_equal(res["nw_cigar"][k, :m], rc[k, :m]), k
    finally:
        srv.close()
