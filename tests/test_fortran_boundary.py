"""The Fortran boundary on the CPU: the drop-in modules export the reference's routine names
(src/numerics/mod_gradient.f90, mod_advection.f90, mod_viscous.f90, src/sgs/mod_vreman.f90,
mod_amd.f90, mod_wall_model.f90, FVSolver's type-bound procedures), the harness library loads, and
the Fortran caller of the pressure backend runs to its end without a GPU, reporting the backend
unavailable and the solve refused (no compute call reaches a device)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLIB = os.path.join(ROOT, "dopamine_amd", "fortran", "libdopamine_fortran.so")
FTEST = os.path.join(ROOT, "tests", "fortran", "libdopamine_ftest.so")
DROPIN = os.path.join(ROOT, "tests", "fortran", "pressure_dropin.x")

MODULE_ROUTINES = {
    "mod_gradient": ["ls_gradient_vector", "ls_gradient_scalar", "green_gauss_gradient_vector",
                     "green_gauss_gradient_scalar"],
    "mod_advection": ["advection_rhs_vector", "advection_rhs_scalar"],
    "mod_viscous": ["viscous_rhs_vector", "viscous_rhs_scalar"],
    "mod_vreman": ["vreman_sgs_viscosity", "smagorinsky_sgs"],
    "mod_amd": ["amd_sgs_viscosity"],
    "mod_dynamic_smagorinsky": ["dynamic_smagorinsky_sgs"],
    "mod_wall_model": ["apply_wall_model_bc"],
    "mod_fv_solver": ["solver_advance", "solver_project_once", "solver_compute_dt", "solver_state_is_finite"],
}


def _exports(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_fortran_modules_export_reference_routines():
    syms = _exports(FLIB)
    for mod, names in MODULE_ROUTINES.items():
        for n in names:
            assert f"_QM{mod}P{n}" in syms, (mod, n)
    # the scheme ids of src/mod_parameters.f90:54-60 as module constants
    for c in ("scheme_vanleer_tvd", "scheme_koren_tvd", "scheme_cds", "scheme_kep", "scheme_quick"):
        assert f"_QMmod_meshEC{c}" in syms


def test_fortran_harness_loads():
    L = ctypes.CDLL(FTEST)
    for n in ("ft_gradient_vector", "ft_gradient_scalar", "ft_advection_vector", "ft_viscous_vector",
              "ft_advection_scalar", "ft_viscous_scalar", "ft_sgs", "ft_wall_model", "ft_solver_run"):
        assert hasattr(L, n)


@pytest.mark.skipif(os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK),
                    reason="a GPU is present: the GPU test runs the solve")
def test_fortran_pressure_caller_without_gpu(tmp_path):
    n, nf = 4, 3
    i32 = lambda a: np.asarray(a, dtype=np.int32).tobytes()  # noqa: E731
    f64 = lambda a: np.asarray(a, dtype=np.float64).tobytes()  # noqa: E731
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as fh:
        fh.write(i32([n, nf, 0, 1, 100]) + f64([1e-8, 1e-12]))
        fh.write(i32([6]) + b"jacobi" + i32([0]))
        fh.write(i32([1, 2, 3, 4]) + i32([1, 2, 3]) + i32([2, 3, 4]) + i32([2, 3, 4]) + i32([0, 0, 0]))
        fh.write(f64([1, 2, 2, 1]) + f64([1, 0, 0, -1]) + f64([1, 1, 1]))
        fh.write(i32([0]) + f64([0]) + f64([0, 0, 0, 0]))
    proc = subprocess.run([DROPIN, str(inp), str(out)], capture_output=True, text=True, timeout=120,
                          env={**os.environ, "HIP_VISIBLE_DEVICES": ""})
    assert proc.returncode == 0, proc.stderr
    raw = open(out, "rb").read()
    avail, init, rc, its, fin = np.frombuffer(raw[:20], dtype=np.int32)
    assert avail == 0 and init < 0 and rc < 0
