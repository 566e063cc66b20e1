"""swiftbwa_amd -- MI355X (gfx950) engine for swiftbwa's Smith-Waterman hot path.

The compute lives in libhipsw.so (hand-written HIP + a C ABI, include/hipsw.h);
this package holds the ctypes binding, the host mirror of swiftbwa's MetalSW
call surface, and the synthetic workload generator.
"""
from . import native
from .metalsw import (BandedSWDispatcher, BandedSWTask, HipSWEngine, ScoringParameters,
                      SWResult)

__all__ = ["native", "BandedSWDispatcher", "BandedSWTask", "HipSWEngine",
           "ScoringParameters", "SWResult"]
