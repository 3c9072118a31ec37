"""dopamine_amd: DOPAMINE's RK3W fractional-step hot path on MI355X (gfx950).

Native library: libdopamine_hip.so (HIP kernels, mesh setup, C ABI; include/*.h).
Python side: mesh generators + builder front end (mesh), solver parameters (params), workloads
(cases) and the FVSolver mirror of the reference solver type (solver).
"""
from ._lib import lib as _load_native

_load_native()  # fail loudly at import when the native library is missing

from . import cases, mesh, params  # noqa: E402,F401
from .solver import FVSolver  # noqa: E402,F401
