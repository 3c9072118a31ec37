"""The C-ABI library: it loads, exports every function include/*.h declares, and refuses to run
without a HIP device instead of falling back to the CPU. No compute calls here."""
import ctypes
import os
import re
import subprocess
import sys

import pytest

import dopamine_amd._lib as _lib
from dopamine_amd.params import SolverParams

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = ["dopamine_hip.h", "dopamine_mesh.h", "fvm_pressure_backend.h", "dopamine_mpnn.h"]


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\(", text, flags=re.M)
    return sorted({n for n in names if n not in ("if", "sizeof")})


def exported():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], check=True, capture_output=True, text=True)
    return {line.split()[-1] for line in out.stdout.splitlines() if " T " in line}


@pytest.mark.parametrize("header", HEADERS)
def test_every_declared_function_is_exported(header):
    names = declared(header)
    assert len(names) >= 4
    missing = sorted(set(names) - exported())
    assert not missing, missing


def test_declared_sets_are_what_the_reference_binds():
    """The pressure backend carries exactly the symbols of src/solver/pressure_backend_petsc.c,
    src/mpi_gpu_support.c and src/io/nvtx_shim.c that the Fortran side binds."""
    assert declared("fvm_pressure_backend.h") == sorted([
        "fvm_petsc_backend_available", "fvm_petsc_backend_initialize", "fvm_petsc_backend_finalize",
        "fvm_petsc_set_pressure_solver", "fvm_petsc_insert_option", "fvm_petsc_set_device_local_rank",
        "fvm_petsc_device_io_supported", "fvm_petsc_set_io_on_device", "fvm_petsc_solve_pressure",
        "dopamine_mpi_gpu_aware", "dopamine_nvtx_push", "dopamine_nvtx_pop"])


def test_gpu_aware_env_override():
    code = ("import ctypes,sys; L=ctypes.CDLL(sys.argv[1]); print(L.dopamine_mpi_gpu_aware())")
    for env, want in (("1", "1"), ("0", "0"), (None, "0")):
        e = dict(os.environ)
        e.pop("DOPAMINE_GPU_AWARE_MPI", None)
        if env is not None:
            e["DOPAMINE_GPU_AWARE_MPI"] = env
        r = subprocess.run([sys.executable, "-c", code, _lib.LIB_PATH], env=e, capture_output=True, text=True)
        assert r.stdout.strip() == want


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(_lib, "LIB_PATH", "/nonexistent/libdopamine_hip.so")
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(ImportError):
        _lib.lib()


def test_params_refuse_out_of_scope_options():
    SolverParams(cell_projection="EXACT").validate()
    with pytest.raises(ValueError):
        SolverParams(cell_projection="MIXED").validate()
    with pytest.raises(ValueError):
        SolverParams(cell_projection="EXACT", cell_projection_tol=0.0).validate()
    with pytest.raises(ValueError):
        SolverParams(cell_projection="LS", advection_scheme="UPWIND").validate()
    with pytest.raises(ValueError):
        SolverParams(cell_projection="LS", adapt_dt=True).validate()


def test_solver_create_without_device_is_an_error():
    """On a machine without a HIP device the solver is not created (no CPU fallback)."""
    L = _lib.lib()
    L.dopamine_device_count.restype = ctypes.c_int
    if L.dopamine_device_count() > 0:
        pytest.skip("a HIP device is present")
    from dopamine_amd import cases
    from dopamine_amd.solver import FVSolver
    case = cases.regression_cases()["hydro"]
    with pytest.raises(RuntimeError, match="no HIP device"):
        FVSolver(case.make_mesh(), case.params)
