"""CPU checks of the serving workload and of the window-fetch restatement."""
import numpy as np

from oracle import pyoracle as orc
from swiftbwa_amd import serving


def test_reference_and_pac_roundtrip():
    ref = serving.make_reference(3, 200000, copies=20)
    assert ref.max() <= 3 and len(ref) == 200000
    pac = serving.pack_pac(ref)
    # PackedReference.base(at:) (PackedReference.swift:34-38)
    idx = np.arange(0, 200000, 997)
    got = (pac[idx >> 2] >> ((3 - (idx & 3)) << 1)) & 3
    assert np.array_equal(got, ref[idx])


def test_get_reference_strands():
    ref = np.array([0, 1, 2, 3, 3, 2, 1, 0, 0, 2], np.uint8)
    gl = len(ref)
    assert list(orc.get_reference(ref, gl, 2, 3)) == [2, 3, 3]
    assert list(orc.get_reference(ref, gl, 8, 5)) == [0, 2]            # clipped at gl
    # reverse strand: pos in [gl, 2gl) -> RC of forward [2gl-pos-len, 2gl-pos)
    assert list(orc.get_reference(ref, gl, gl, 3)) == [3 - 2, 3 - 0, 3 - 0]
    assert list(orc.get_reference(ref, gl, 2 * gl - 2, 5)) == [3 - 1, 3 - 0]
    assert len(orc.get_reference(ref, gl, -1, 5)) == 0


def test_chunk_jobs_are_well_formed():
    ref = serving.make_reference(4, 300000, copies=30)
    ch = serving.make_chunk(ref, 9, 300)
    gl = len(ref)
    assert ch["loc"]["n"] == 300 and ch["nw"]["n"] == 600
    assert 0 < ch["ext"]["n"] <= 1200
    for kind in ("ext", "loc", "nw"):
        j = ch[kind]
        assert (j["t_pos"] >= 0).all() and (j["t_pos"] < 2 * gl).all()
    sets = orc.chunk_to_job_sets(ref, ch)
    out, _ = orc.ksw_extend_batch(sets["ext"], nthreads=4)
    # most extensions of error-light reads reach the end of the query
    assert np.mean(out[:, 1] == sets["ext"]["q_len"]) > 0.8
