"""nbody_amd -- an MI355X (gfx950) gravity solver for the stett/nbody simulation API.

Layout:
  csrc/      HIP kernels and the C ABI of libnbody_hip.so (declared in include/nbody_hip.h)
  host/      C++ `Solver` subclass over the C ABI (the drop-in for reference source/solver.h)
  _lib.py    ctypes binding of the C ABI for Python callers (tests, bench, serving loop)
  ics.py     seeded initial conditions (Plummer, two-Plummer, disk + bulge, cube)
  build.py   build recipe (hipcc for gfx950, g++ for the host layer)
"""
from . import ics  # noqa: F401
from ._lib import Context, HipError, Params, load, params, probe  # noqa: F401
