"""amca3d_amd: MI355X (gfx950) implementation of AMCA3D's cellular-automaton step.

The compute path is libamca3d.so (hand-written HIP kernels behind a C ABI,
include/amca3d.h) with a C++ host layer that reads the reference's YAML input
decks.  This Python package is plumbing for tests and benchmarks: it loads the
library with ctypes and has no compute path of its own.
"""
from ._lib import AmcaError, load  # noqa: F401
from .sim import Domain, OwnedDomain, Simulation, envelope_selftest  # noqa: F401
