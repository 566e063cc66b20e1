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


def _long_query_chunk(ref, seed, n_ext, n_loc):
    """A chunk whose extensions have 160-255-base queries, half of them with a
    seed score that puts them in the 16-bit tier, plus 150-bp rescue tasks in
    600-bp windows: the bins with the largest HBM column state."""
    rng = np.random.default_rng(seed)
    gl = len(ref)
    ext = {k: [] for k in ("q", "q_len", "t_pos", "t_len", "t_rev", "a", "b", "c")}
    for k in range(n_ext):
        L = int(rng.integers(160, 256))
        p = int(rng.integers(1000, gl - 2000))
        q = ref[p:p + L].copy()
        flip = rng.random(L) < 0.02
        q[flip] = (q[flip] + 1) % 4
        ext["q"].append(q)
        ext["q_len"].append(L)
        ext["t_pos"].append(p)
        ext["t_len"].append(L + int(rng.integers(0, 100)))
        ext["t_rev"].append(k & 1)
        ext["a"].append(int(rng.integers(19, 60)) if k % 2 else int(rng.integers(300, 900)))
        ext["b"].append(100)
        ext["c"].append(5)
    loc = {k: [] for k in ("q", "q_len", "t_pos", "t_len")}
    for k in range(n_loc):
        p = int(rng.integers(1000, gl - 2000))
        q = ref[p + 200:p + 350].copy()
        loc["q"].append(q)
        loc["q_len"].append(150)
        loc["t_pos"].append(p)
        loc["t_len"].append(600)

    def pack(d, n):
        out = {"n": n, "q": np.concatenate(d["q"]) if n else np.zeros(1, np.uint8)}
        ql = np.array(d["q_len"], np.uint16)
        off = np.zeros(n, np.uint32)
        if n:
            off[1:] = np.cumsum(ql.astype(np.int64))[:-1]
        out["q_off"], out["q_len"] = off, ql
        out["t_pos"] = np.array(d["t_pos"], np.int64)
        out["t_len"] = np.array(d["t_len"], np.uint16)
        for f, dt in (("t_rev", np.uint8), ("a", np.int32), ("b", np.int32), ("c", np.int32)):
            out[f] = np.array(d.get(f, [0] * n), dt)
        return out

    nw = serving.make_chunk(ref, seed + 1, 200)["nw"]
    return {"pairs": n_ext, "ext": pack(ext, n_ext), "loc": pack(loc, n_loc), "nw": nw}


def test_budget_estimate_bounds_peak(engine, ref):
    """Each chunk's admission estimate bounds the device memory its run
    holds at once (arena plus the largest phase scratch), also when the
    160-255-base bins of both score tiers run the layouts with the most HBM
    column state; results stay the oracle's."""
    lib = engine.lib
    srv = serving.Server(engine, slots=2, hbm_budget=16 << 30)
    try:
        srv.set_reference(serving.pack_pac(ref), len(ref))
        assert lib.hipsw_ext_set_layout(engine.handle, 2, 1, 6) == 0
        assert lib.hipsw_ext_set_layout16(engine.handle, 2, 6) == 0
        chunks = [_long_query_chunk(ref, 700 + k, n, m)
                  for k, (n, m) in enumerate([(300, 50), (4000, 2000), (20000, 6000)])]
        chunks.append(serving.make_chunk(ref, 710, 3000))
        for ch in chunks:
            r = srv.wait(srv.submit(ch))
            assert 0 < r["dev_peak_bytes"] <= r["est_bytes"], (r["dev_peak_bytes"], r["est_bytes"])
            sets = orc.chunk_to_job_sets(ref, ch)
            ext_ref, _ = orc.ksw_extend_batch(sets["ext"], nthreads=8)
            assert np.array_equal(r["ext"], ext_ref)
    finally:
        lib.hipsw_ext_set_layout(engine.handle, 2, 0, 0)
        lib.hipsw_ext_set_layout16(engine.handle, 2, 0)
        srv.close()


def test_rescue_windows_clamped_to_sequences(engine, ref):
    """Raw mate-rescue windows (loc_rid set) on a reference split into
    sequences: the server's fetchSequence clamp and sequence check, then local
    SW, equal the restatement in tests/serving_pipeline.py followed by the
    oracle -- clamped starts, skipped windows and every result row."""
    import serving_pipeline as SP
    lens = serving.contig_lengths(len(ref), 9)
    offs = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64)
    ch = serving.make_chunk(ref, 77, 3000)
    loc = ch["loc"]
    rng = np.random.default_rng(7)
    # push a third of the windows onto sequence ends, both strands
    k = rng.choice(loc["n"], loc["n"] // 3, replace=False)
    ends = np.concatenate([offs[1:], [len(ref)]])
    loc["t_pos"][k] = ends[rng.integers(0, len(ends), len(k))] - rng.integers(0, 600, len(k))
    flip = k[::2]
    loc["t_pos"][flip] = 2 * len(ref) - (loc["t_pos"][flip] + loc["t_len"][flip])
    fwd_start = np.where(loc["t_pos"] < len(ref), loc["t_pos"], 2 * len(ref) - loc["t_pos"] - loc["t_len"])
    loc["rid"] = (np.searchsorted(offs, fwd_start + 300, side="right") - 1).astype(np.int32)
    only = {"ext": None, "loc": loc, "nw": None}
    srv = serving.Server(engine, slots=2, hbm_budget=2 << 30)
    try:
        srv.set_reference(serving.pack_pac(ref), len(ref))
        srv.set_contigs(lens)
        got = srv.wait(srv.submit(only))
    finally:
        srv.close()
    exp = SP.CpuServer(ref, lens)
    want = exp.wait(exp.submit(only))
    assert np.array_equal(got["loc_win"], want["loc_win"])
    assert (want["loc_win"] == -1).sum() > 10 and (want["loc_win"] != loc["t_pos"]).sum() > 100
    assert np.array_equal(got["loc_ok"], want["loc_ok"])
    assert np.array_equal(got["loc"], want["loc"])
