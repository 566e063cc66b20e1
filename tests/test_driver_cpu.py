"""Host-side parts of the chain-extension driver that need no device:
hipsw_get_reference_pac (FMIndex.getReference over a .pac) against the
oracle's restatement, on both strands and across every boundary."""
import ctypes as C

import numpy as np

from oracle import pyoracle as orc
from swiftbwa_amd import native, serving


def test_get_reference_pac_matches_fmindex():
    lib = native.load()
    rng = np.random.default_rng(3)
    gl = 5003
    ref = rng.integers(0, 4, gl).astype(np.uint8)
    pac = serving.pack_pac(ref)
    ctx = native.HipswPacRef(pac.ctypes.data, gl)
    buf = np.zeros(300, np.uint8)
    cases = [(0, 10), (gl - 5, 10), (gl, 7), (2 * gl - 3, 10), (2 * gl, 5), (-4, 9),
             (gl - 1, 1), (2 * gl - 1, 1), (123, 0)]
    cases += [(int(rng.integers(-20, 2 * gl + 20)), int(rng.integers(0, 300))) for _ in range(400)]
    for pos, ln in cases:
        n = lib.hipsw_get_reference_pac(C.byref(ctx), pos, ln, native.ptr(buf))
        want = orc.get_reference(ref, gl, pos, ln)
        assert n == len(want), (pos, ln)
        assert np.array_equal(buf[:n], want), (pos, ln)
