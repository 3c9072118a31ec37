"""FVSolver: the host-side mirror of the reference's solver type, running on the MI355X.

Method names follow the FVSolver type-bound procedures of src/solver/mod_fv_solver.f90:270-285
(init / project_once / advance / compute_dt / divergence_metrics / sync_host); every call goes
through the C ABI of libdopamine_hip.so (include/dopamine_hip.h). There is no CPU fallback: without
the library or a HIP device, construction raises.
"""
import ctypes

import numpy as np

from ._lib import lib
from . import params as P
from .params import WALL_MODEL_IDS

_P = ctypes.POINTER
_d, _i32, _i64 = ctypes.c_double, ctypes.c_int32, ctypes.c_int64

MESH_FIELDS = [("cell_centroid", _d), ("cell_volume", _d), ("cell_face_offset", _i32), ("cell_faces", _i32),
               ("face_owner", _i32), ("face_neighbour", _i32), ("face_patch", _i32),
               ("face_area", _d), ("face_centre", _d), ("face_normal", _d), ("face_sf", _d), ("face_dcf", _d),
               ("face_gc", _d), ("face_skew", _d), ("face_eps", _d),
               ("patch_bc", _i32), ("patch_start", _i32), ("patch_nfaces", _i32),
               ("periodic_pair", _i32), ("periodic_pair_dcf", _d), ("periodic_gc", _d),
               ("ls_offset", _i32), ("ls_cell", _i32), ("ls_face", _i32), ("ls_weight", _d)]


class MeshDesc(ctypes.Structure):
    _fields_ = [("n_cells", _i32), ("n_faces", _i32), ("n_int_faces", _i32), ("n_bnd_faces", _i32),
                ("n_patches", _i32)] + [(n, _P(t)) for n, t in MESH_FIELDS]


class Params(ctypes.Structure):
    _fields_ = [("nu", _d), ("f_const", _d * 3), ("scheme", _i32), ("gradient", _i32), ("gg_iterations", _i32),
                ("sgs", _i32), ("c_vreman", _d), ("cs_smag", _d), ("wall_model", _i32), ("wall_model_height", _d),
                ("n_nonortho_corr", _i32), ("min_cos_theta", _d), ("rhie_chow_dual", _i32),
                ("nonortho_check", _i32), ("rotational", _i32), ("p_tol", _d), ("p_max_iter", _i32),
                ("cfl_max", _d), ("dt", _d), ("patch_p_val", _P(_d)), ("patch_U_val", _P(_d)),
                ("precision", _i32), ("cg_variant", _i32), ("c_amd", _d), ("amd_wall_gradient", _i32),
                ("boussinesq", _i32), ("alpha_T", _d), ("beta", _d), ("T_ref", _d), ("g", _d * 3), ("Pr_t", _d),
                ("t_scheme", _i32), ("patch_T_val", _P(_d)), ("precond", _i32), ("learn_collect", _i32),
                ("learn_capacity", _i32), ("learn_iters", _i32), ("learn_batch", _i32), ("learn_seed", _i32),
                ("learn_lr", _d), ("most_z0_m", _d), ("most_z0_h", _d), ("T_surface", _d), ("L_obukhov", _d),
                ("compute_L", _i32), ("cell_projection", _i32), ("cell_projection_tol", _d),
                ("use_rotation", _i32), ("omega_rot", _d * 3), ("use_harmonic", _i32), ("forcing_amp", _d),
                ("forcing_omega", _d), ("forcing_phase", _d), ("forcing_dir", _d * 3), ("mass_flux", _i32),
                ("U_bulk_target", _d * 3), ("n_scalars", _i32), ("scalar_scheme", _i32),
                ("scalar_diffusivity", _d * 4), ("scalar_pr_t", _d * 4), ("scalar_bc_mode", _P(_i32)),
                ("scalar_bc_value", _P(_d)), ("scalar_bc_flux", _P(_d)), ("scalar_bc_omega", _P(_d)),
                ("scalar_bc_phase", _P(_d)), ("patch_outlet_conv", _P(_i32)), ("patch_backflow_beta", _P(_d)),
                ("patch_z0_m", _P(_d)), ("patch_z0_h", _P(_d)), ("wall_model_filter_time", _d),
                ("wall_model_filter_tf_fac", _d)]


STATS_COLUMNS = ("E", "eps", "Umax", "Ub", "tauw_mean", "nu_sgs_mean", "max_abs_divU", "volavg_abs_divU")


def build_abi(mesh, params, precision=0):
    """ctypes mesh and params structures (plus the arrays they point into, kept alive)."""
    keep = []

    def arr(a, t):
        a = np.ascontiguousarray(a, dtype=np.float64 if t is _d else np.int32)
        keep.append(a)
        return a.ctypes.data_as(_P(t))

    m = mesh
    md = MeshDesc(m.n_cells, m.n_faces, m.n_int_faces, m.n_bnd_faces, m.n_patches,
                  *[arr(getattr(m, n), t) for n, t in MESH_FIELDS])
    pv, uv, tv = params.patch_arrays(mesh, with_T=True)
    sm, sv, sfl, so, sph = params.scalar_arrays(mesh)
    oc, ob, zm, zh = params.bc_ext_arrays(mesh)
    pr = Params(params.nu, (_d * 3)(*params.f_const), params.scheme_id(),
                0 if params.gradient_scheme.upper() == "LEAST_SQUARES" else 1, params.gg_iterations,
                params.sgs_id(), params.c_vreman, params.cs_smagorinsky, WALL_MODEL_IDS[params.wall_model],
                params.wall_model_height, params.n_nonortho_corr, params.pressure_min_cos_theta,
                int(params.rhie_chow_dual), int(params.nonortho_check), int(params.use_rotational_projection),
                params.p_tol, params.p_max_iter, params.cfl_max, params.dt, arr(pv, _d), arr(uv.reshape(-1), _d),
                int(precision), int(params.cg_variant), params.c_amd, int(params.amd_wall_gradient),
                int(params.boussinesq), params.alpha_T, params.beta, params.T_ref, (_d * 3)(*params.g), params.Pr_t,
                params.t_scheme_id(), arr(tv, _d), P.PRECOND_IDS[params.precond.upper()],
                params.learn_collect, params.learn_capacity, params.learn_iters, params.learn_batch,
                params.learn_seed, params.learn_lr, params.z0_m, params.z0_h, params.T_surface, params.L_obukhov,
                int(params.compute_L), params.projection_id(), params.cell_projection_tol, *params.forcing_args(),
                params.n_scalars, params.scalar_scheme_id(), (_d * 4)(*params.scalar_diffusivity),
                (_d * 4)(*params.scalar_pr_t), arr(sm, _i32), arr(sv, _d), arr(sfl, _d), arr(so, _d), arr(sph, _d),
                arr(oc, _i32), arr(ob, _d), arr(zm, _d), arr(zh, _d), params.wall_model_filter_time,
                params.wall_model_filter_tf_fac)
    return md, pr, keep


def _bind(L):
    L.dopamine_solver_create.restype = ctypes.c_void_p
    L.dopamine_solver_create.argtypes = [_P(MeshDesc), _P(Params)]
    L.dopamine_solver_destroy.argtypes = [ctypes.c_void_p]
    L.dopamine_last_error.restype = ctypes.c_char_p
    L.dopamine_solver_set_field.argtypes = [ctypes.c_void_p, ctypes.c_char_p, _P(_d)]
    L.dopamine_solver_get_field.argtypes = [ctypes.c_void_p, ctypes.c_char_p, _P(_d), _i64]
    L.dopamine_solver_project_once.argtypes = [ctypes.c_void_p, _d]
    L.dopamine_solver_advance.argtypes = [ctypes.c_void_p, _d, _P(_d)]
    L.dopamine_solver_compute_dt.restype = _d
    L.dopamine_solver_compute_dt.argtypes = [ctypes.c_void_p]
    L.dopamine_solver_stats.argtypes = [ctypes.c_void_p, _d * 8]
    L.dopamine_solver_time.restype = _d
    L.dopamine_solver_time.argtypes = [ctypes.c_void_p]
    L.dopamine_solver_counters.argtypes = [ctypes.c_void_p, _i64 * 4]
    L.dopamine_solver_nonortho_info.argtypes = [ctypes.c_void_p, _i32 * 3]
    L.dopamine_solver_sync.argtypes = [ctypes.c_void_p]
    L.dopamine_solver_kernel_timing.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    L.dopamine_solver_kernel_time.argtypes = [ctypes.c_void_p, ctypes.c_int, _P(_i64), _P(_d)]
    L.dopamine_solver_learned_info.argtypes = [ctypes.c_void_p, _d * 8]
    L.dopamine_solver_learned_precond_info.argtypes = [ctypes.c_void_p, _i64 * 2]
    L.dopamine_solver_projection_info.argtypes = [ctypes.c_void_p, _i64 * 2]
    L.dopamine_solver_multigrid_info.argtypes = [ctypes.c_void_p, _P(_i64), ctypes.c_int]
    L.dopamine_solver_cell_divergence.argtypes = [ctypes.c_void_p, _d * 2]
    L.dopamine_solver_write_checkpoint.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
    L.dopamine_solver_read_checkpoint.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
    L.dopamine_solver_stats_csv.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, _d, _P(_d), ctypes.c_int]
    L.dopamine_solver_learned_loss.argtypes = [ctypes.c_void_p, _P(_d), ctypes.c_int]
    L.dopamine_solver_state_is_finite.argtypes = [ctypes.c_void_p]
    L.dopamine_solver_cell_net_flux.argtypes = [ctypes.c_void_p, _P(_d)]
    L.dopamine_solver_retries.restype = _i64
    L.dopamine_solver_retries.argtypes = [ctypes.c_void_p]
    return L


# Kernel classes with device timing (include/dopamine_hip.h).
KERNELS = {"cg_dir": 0, "cg_upd": 1, "explicit": 2, "ls_grad": 3, "mg_down": 4, "mg_up": 5}


def check_precision(params, precision):
    """The pressure paths that exist: precision 0 (FP64) with any preconditioner, precision 1 (FP32
    sweeps in FP64 refinement) with the Jacobi CG."""
    if precision not in (0, 1):
        raise ValueError(f"precision must be 0 (FP64) or 1 (mixed), got {precision}")
    if precision == 1 and params.precond.upper() == "LEARNED":
        raise ValueError("the learned preconditioner runs in the FP64 flexible CG only (precision 0)")


class FVSolver:
    """One RK3W solver on the current HIP device."""

    def __init__(self, mesh, params, precision=0):
        params.validate()
        check_precision(params, precision)
        self.L = _bind(lib())
        self.mesh = mesh
        self.params = params
        md, pr, keep = build_abi(mesh, params, precision)
        self.h = self.L.dopamine_solver_create(ctypes.byref(md), ctypes.byref(pr))
        if not self.h:
            raise RuntimeError("dopamine_solver_create failed: " + self.L.dopamine_last_error().decode())
        self._sizes = {"U": 3 * mesh.n_cells, "p": mesh.n_cells, "nu_sgs": mesh.n_cells,
                       "U_bnd": 3 * max(1, mesh.n_bnd_faces), "p_bnd": max(1, mesh.n_bnd_faces),
                       "grad_U": 9 * mesh.n_cells, "U_face": mesh.n_faces, "tau_w_bnd": 3 * max(1, mesh.n_bnd_faces),
                       "rhs_p": mesh.n_cells, "H": 3 * mesh.n_cells, "w_gp": 3 * mesh.n_cells,
                       "T": mesh.n_cells, "T_bnd": max(1, mesh.n_bnd_faces), "grad_T": 3 * mesh.n_cells,
                       "q_w_bnd": max(1, mesh.n_bnd_faces), "L_bnd": max(1, mesh.n_bnd_faces),
                       "outlet_pb": max(1, mesh.n_bnd_faces), "wm_filt": 3 * max(1, mesh.n_bnd_faces)}
        for k in range(params.n_scalars):
            self._sizes.update({f"C{k}": mesh.n_cells, f"C{k}_bnd": max(1, mesh.n_bnd_faces),
                                f"grad_C{k}": 3 * mesh.n_cells})

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc}): {self.L.dopamine_last_error().decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.L.dopamine_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def get(self, name):
        n = self._sizes[name]
        out = np.empty(n)
        self._check(self.L.dopamine_solver_get_field(self.h, name.encode(), out.ctypes.data_as(_P(_d)), n), "get")
        return out

    def set(self, name, values):
        a = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        if a.size != self._sizes[name]:
            raise ValueError(f"{name}: expected {self._sizes[name]} values, got {a.size}")
        self._check(self.L.dopamine_solver_set_field(self.h, name.encode(), a.ctypes.data_as(_P(_d))), "set")

    @property
    def U(self):
        return self.get("U").reshape(-1, 3)

    @property
    def p(self):
        return self.get("p")

    def set_velocity(self, U):
        self.set("U", U)

    def set_temperature(self, T):
        self.set("T", T)

    def set_scalar(self, k, C):
        """Passive scalar k (cell values); its boundary values are refreshed."""
        self.set(f"C{k}", C)

    def scalar(self, k):
        return self.get(f"C{k}")

    def project_once(self, dt):
        self._check(self.L.dopamine_solver_project_once(self.h, dt), "project_once")

    def advance(self, dt=None):
        used = _d(0.0)
        self._check(self.L.dopamine_solver_advance(self.h, self.params.dt if dt is None else dt, ctypes.byref(used)),
                    "advance")
        return used.value

    def compute_dt(self):
        return self.L.dopamine_solver_compute_dt(self.h)

    def state_is_finite(self):
        """solver_state_is_finite (src/solver/mod_fv_solver.f90:3925-3945): U and p hold no NaN/Inf."""
        rc = self.L.dopamine_solver_state_is_finite(self.h)
        if rc < 0:
            raise RuntimeError(f"state_is_finite failed ({rc}): {self.L.dopamine_last_error().decode()}")
        return rc == 1

    def run(self, n_steps, dt=None, check_every=10):
        """n_steps of advance(); like the reference's time loop (src/main.f90:434-442) the state is
        checked for NaN/Inf every `check_every` steps and a diverged run stops with an error."""
        for step in range(1, n_steps + 1):
            self.advance(dt)
            if check_every and step % check_every == 0 and not self.state_is_finite():
                raise FloatingPointError(f"non-finite velocity/pressure detected at step {step}, t = {self.time:.4e}")

    def cell_net_flux_max(self):
        """max over the cells of |sum of the projected face fluxes| (m^3/s)."""
        out = _d(0.0)
        self._check(self.L.dopamine_solver_cell_net_flux(self.h, ctypes.byref(out)), "cell_net_flux")
        return out.value

    @property
    def retries(self):
        """Pressure solves that were repeated from p = 0 with four times the iteration limit."""
        return self.L.dopamine_solver_retries(self.h)

    @property
    def time(self):
        return self.L.dopamine_solver_time(self.h)

    def stats(self):
        out = (_d * 8)()
        self._check(self.L.dopamine_solver_stats(self.h, out), "stats")
        return dict(zip(STATS_COLUMNS, list(out)))

    def counters(self):
        out = (_i64 * 4)()
        self.L.dopamine_solver_counters(self.h, out)
        return {"solves": out[0], "iterations": out[1], "last": out[2], "sweeps": out[3]}

    def nonortho_info(self):
        out = (_i32 * 3)()
        self.L.dopamine_solver_nonortho_info(self.h, out)
        return {"skip_nonortho": out[0], "skip_grad_p": out[1], "n_nonortho_corr": out[2]}

    def sync(self):
        self._check(self.L.dopamine_solver_sync(self.h), "sync")

    def kernel_timing(self, kernel, enable=True):
        """Start (clearing the totals) or stop HIP-event timing of a kernel class."""
        self._check(self.L.dopamine_solver_kernel_timing(self.h, KERNELS[kernel], int(enable)), "kernel_timing")

    def write_checkpoint(self, directory, step, rank=0):
        """The reference's per-rank checkpoint (dir/processor<rank>/checkpoint_<step>.bin)."""
        self._check(self.L.dopamine_solver_write_checkpoint(self.h, str(directory).encode(), step, rank),
                    "write_checkpoint")

    def read_checkpoint(self, directory, step, rank=0):
        """Restore from a checkpoint (then the restart refresh); returns its step."""
        rc = self.L.dopamine_solver_read_checkpoint(self.h, str(directory).encode(), step, rank)
        if rc < 0:
            raise RuntimeError(f"read_checkpoint failed ({rc}): {self.L.dopamine_last_error().decode()}")
        return rc

    def cell_divergence(self):
        out = (_d * 2)()
        self._check(self.L.dopamine_solver_cell_divergence(self.h, out), "cell_divergence")
        return {"max_abs_divU_cell": out[0], "volavg_abs_divU_cell": out[1]}

    def stats_csv(self, path, step, dt, E0=0.0, append=True):
        """Append a row of the reference's stats CSV; returns the reference energy E0 used."""
        e0 = _d(E0)
        self._check(self.L.dopamine_solver_stats_csv(self.h, str(path).encode(), step, dt, ctypes.byref(e0),
                                                     int(append)), "stats_csv")
        return e0.value

    LEARNED_INFO = ("phase", "collected", "samples", "learned_solves", "learned_iterations", "fallbacks",
                    "jacobi_iterations_mean", "train_ms")

    def learned_info(self):
        """State of the learned preconditioner (precond LEARNED), or None."""
        out = (_d * 8)()
        if self.L.dopamine_solver_learned_info(self.h, out) != 0:
            return None
        d = dict(zip(self.LEARNED_INFO, list(out)))
        d["phase"] = ("collecting", "in use", "off")[int(d["phase"])]
        return d

    def learned_precond_info(self):
        """Exact cell projection with precond LEARNED: its preconditioner solves that ran the learned
        flexible CG and their iterations, or None."""
        out = (_i64 * 2)()
        if self.L.dopamine_solver_learned_precond_info(self.h, out) != 0:
            return None
        return {"solves": out[0], "iterations": out[1]}

    def proj_counters(self):
        """Exact cell projection: projections and flexible-PCG iterations so far."""
        out = (_i64 * 2)()
        self._check(self.L.dopamine_solver_projection_info(self.h, out), "projection_info")
        return {"projections": out[0], "iterations": out[1]}

    def multigrid_info(self):
        """Hierarchy of the multigrid pressure preconditioner (None before its first solve or when
        the preconditioner is not multigrid): cells and stored off-diagonals per level."""
        out = (_i64 * 68)()
        if self.L.dopamine_solver_multigrid_info(self.h, out, 68) != 0:
            return None
        nl = int(out[0])
        return {"levels": nl, "aggressive": bool(out[1]), "cells": [int(out[2 + i]) for i in range(nl)],
                "nnz": [int(out[2 + nl + i]) for i in range(nl)],
                "stored": [int(out[2 + 2 * nl + i]) for i in range(nl)]}

    def learned_loss(self, max_steps=100000):
        out = np.zeros(max_steps)
        m = self.L.dopamine_solver_learned_loss(self.h, out.ctypes.data_as(_P(_d)), max_steps)
        return out[:max(0, m)]

    def kernel_time(self, kernel):
        """(launches, summed device milliseconds) since timing of the class was enabled."""
        n, ms = _i64(0), _d(0.0)
        self._check(self.L.dopamine_solver_kernel_time(self.h, KERNELS[kernel], ctypes.byref(n), ctypes.byref(ms)),
                    "kernel_time")
        return n.value, ms.value
