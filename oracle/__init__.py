"""ORACLE — test infrastructure only.

Python binding of liboracle.so, the FP64 CPU restatement of the reference's RK3W hot path
(dopamine_oracle.c). It is the parity checker for the HIP product path: only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import it. Nothing it does is ever
used by dopamine_amd.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_P = ctypes.POINTER
_d, _i32, _i64 = ctypes.c_double, ctypes.c_int32, ctypes.c_int64


class _Mesh(ctypes.Structure):
    _fields_ = [("n_cells", _i32), ("n_faces", _i32), ("n_int_faces", _i32), ("n_bnd_faces", _i32),
                ("n_patches", _i32)] + [(n, _P(t)) for n, t in [
                    ("cell_centroid", _d), ("cell_volume", _d), ("cell_face_offset", _i32), ("cell_faces", _i32),
                    ("face_owner", _i32), ("face_neighbour", _i32), ("face_patch", _i32),
                    ("face_area", _d), ("face_centre", _d), ("face_normal", _d), ("face_sf", _d), ("face_dcf", _d),
                    ("face_gc", _d), ("face_skew", _d), ("face_eps", _d),
                    ("patch_bc", _i32), ("patch_start", _i32), ("patch_nfaces", _i32),
                    ("periodic_pair", _i32), ("periodic_pair_dcf", _d), ("periodic_gc", _d),
                    ("ls_offset", _i32), ("ls_cell", _i32), ("ls_face", _i32), ("ls_weight", _d)]]


class _Params(ctypes.Structure):
    _fields_ = [("nu", _d), ("f_const", _d * 3), ("scheme", ctypes.c_int), ("gradient", ctypes.c_int),
                ("gg_iterations", ctypes.c_int), ("sgs", ctypes.c_int), ("c_vreman", _d), ("cs_smag", _d),
                ("wall_model", ctypes.c_int), ("wall_model_height", _d), ("n_nonortho_corr", ctypes.c_int),
                ("min_cos_theta", _d), ("rhie_chow_dual", ctypes.c_int), ("nonortho_check", ctypes.c_int),
                ("rotational", ctypes.c_int), ("p_tol", _d), ("p_max_iter", ctypes.c_int), ("cfl_max", _d),
                ("dt", _d), ("patch_p_val", _P(_d)), ("patch_U_val", _P(_d)), ("c_amd", _d),
                ("amd_wall_gradient", ctypes.c_int), ("boussinesq", ctypes.c_int), ("alpha_T", _d), ("beta", _d),
                ("T_ref", _d), ("g", _d * 3), ("Pr_t", _d), ("t_scheme", ctypes.c_int), ("patch_T_val", _P(_d)),
                ("most_z0_m", _d), ("most_z0_h", _d), ("T_surface", _d), ("L_obukhov", _d), ("compute_L", ctypes.c_int),
                ("cell_projection", ctypes.c_int), ("cell_projection_tol", _d),
                ("use_rotation", ctypes.c_int), ("omega_rot", _d * 3), ("use_harmonic", ctypes.c_int),
                ("forcing_amp", _d), ("forcing_omega", _d), ("forcing_phase", _d), ("forcing_dir", _d * 3),
                ("mass_flux", ctypes.c_int), ("U_bulk_target", _d * 3),
                ("n_scalars", ctypes.c_int), ("scalar_scheme", ctypes.c_int), ("scalar_diffusivity", _d * 4),
                ("scalar_pr_t", _d * 4), ("scalar_bc_mode", _P(ctypes.c_int)), ("scalar_bc_value", _P(_d)),
                ("scalar_bc_flux", _P(_d)), ("scalar_bc_omega", _P(_d)), ("scalar_bc_phase", _P(_d)),
                ("patch_outlet_conv", _P(ctypes.c_int)), ("patch_backflow_beta", _P(_d)), ("patch_z0_m", _P(_d)),
                ("patch_z0_h", _P(_d)), ("wall_model_filter_time", _d), ("wall_model_filter_tf_fac", _d)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        L.om_create.restype = ctypes.c_void_p
        L.om_create.argtypes = [_P(_Mesh), _P(_Params)]
        L.om_set_threads.argtypes = [ctypes.c_int]
        L.om_get_threads.restype = ctypes.c_int
        L.om_free.argtypes = [ctypes.c_void_p]
        L.om_array.restype = _P(_d)
        L.om_array.argtypes = [ctypes.c_void_p, ctypes.c_char_p, _P(_i64)]
        L.om_apply_all_bc.argtypes = [ctypes.c_void_p]
        L.om_project_once.argtypes = [ctypes.c_void_p, _d]
        L.om_advance.argtypes = [ctypes.c_void_p, _d, _P(_d)]
        L.om_stats.argtypes = [ctypes.c_void_p, _d * 8]
        L.om_divergence_cell.argtypes = [ctypes.c_void_p, _d * 2]
        L.om_int_array.restype = _P(_i32)
        L.om_int_array.argtypes = [ctypes.c_void_p, ctypes.c_char_p, _P(_i64)]
        L.om_time.restype = _d
        L.om_time.argtypes = [ctypes.c_void_p]
        L.om_eval_closures.argtypes = [ctypes.c_void_p, _d]
        L.om_wall_tau.restype = _d
        L.om_wall_tau.argtypes = [_d, _d, _d]
        L.om_amd_point.restype = _d
        L.om_amd_point.argtypes = [_d * 9, _d * 9, _d]
        L.om_counters.argtypes = [ctypes.c_void_p, _i64 * 3]
        L.om_proj_counters.argtypes = [ctypes.c_void_p, _i64 * 2]
        L.om_proj_divergence.argtypes = [ctypes.c_void_p, _P(_d), _P(_d)]
        L.om_proj_adjoint.argtypes = [ctypes.c_void_p, _P(_d), _P(_d)]
        L.om_proj_project.argtypes = [ctypes.c_void_p, _P(_d), _P(_d), ctypes.c_int, _d, ctypes.c_int]
        L.om_most_iterate.argtypes = [_d, _d, _d, _d, _d, _d, _d, _d * 3]
        L.om_psi_m.restype = _d
        L.om_psi_m.argtypes = [_d]
        L.om_psi_h.restype = _d
        L.om_psi_h.argtypes = [_d]
        L.om_nonortho_info.argtypes = [ctypes.c_void_p, ctypes.c_int * 3]
        vp, ci = ctypes.c_void_p, ctypes.c_int
        L.om_op_gradient_vector.argtypes = [vp, ci, ci, _P(_d)]
        L.om_op_gradient_scalar.argtypes = [vp, ci, ci, _P(_d), _P(_d), _P(_d)]
        L.om_op_advection_vector.argtypes = [vp, ci, _P(_d)]
        L.om_op_viscous_vector.argtypes = [vp, _P(_d)]
        L.om_op_advection_scalar.argtypes = [vp, ci, _P(_d), _P(_d), _P(_d), _P(_d)]
        L.om_op_viscous_scalar.argtypes = [vp, _P(_d), _P(_d), _P(_d), _d, _d, _P(_d)]
        L.om_op_sgs.argtypes = [vp, ci, _d, ci, _d, _P(_d)]
        L.om_op_wall_model.argtypes = [vp, _d, _P(_d)]
        L.om_outlet_dn_u.restype = _d
        L.om_outlet_dn_u.argtypes = [_d] * 6
        L.om_wall_sample.restype = ci
        L.om_wall_sample.argtypes = [vp, ci, _P(_d)]
        L.om_wall_filter_state.restype = ci
        L.om_wall_filter_state.argtypes = [vp, _P(_d)]
        L.om_outlet_speeds.argtypes = [vp, _P(_d)]
        L.om_cg_jacobi.argtypes = [ctypes.c_int, ctypes.c_int, _P(_i32), _P(_i32), _P(_d), _P(_d), _P(_d), _P(_d),
                                   _d, _d, ctypes.c_int, ctypes.c_int, _P(ctypes.c_int), _P(_d)]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(_P(t))


STATS_COLUMNS = ("E", "eps", "Umax", "Ub", "tauw_mean", "nu_sgs_mean", "max_abs_divU", "volavg_abs_divU")


class Solver:
    """Oracle solver on a dopamine_amd.mesh.Mesh with dopamine_amd.params.SolverParams."""

    def __init__(self, mesh, params):
        params.validate()
        L = lib()
        self.mesh = mesh
        self._keep = []

        def arr(a, t):
            a = np.ascontiguousarray(a)
            self._keep.append(a)
            return _p(a, t)

        m = mesh
        self._m = _Mesh(m.n_cells, m.n_faces, m.n_int_faces, m.n_bnd_faces, m.n_patches,
                        arr(m.cell_centroid, _d), arr(m.cell_volume, _d), arr(m.cell_face_offset, _i32),
                        arr(m.cell_faces, _i32), arr(m.face_owner, _i32), arr(m.face_neighbour, _i32),
                        arr(m.face_patch, _i32), arr(m.face_area, _d), arr(m.face_centre, _d),
                        arr(m.face_normal, _d), arr(m.face_sf, _d), arr(m.face_dcf, _d), arr(m.face_gc, _d),
                        arr(m.face_skew, _d), arr(m.face_eps, _d), arr(m.patch_bc, _i32), arr(m.patch_start, _i32),
                        arr(m.patch_nfaces, _i32), arr(m.periodic_pair, _i32), arr(m.periodic_pair_dcf, _d),
                        arr(m.periodic_gc, _d), arr(m.ls_offset, _i32), arr(m.ls_cell, _i32), arr(m.ls_face, _i32),
                        arr(m.ls_weight, _d))
        pv, uv, tv = params.patch_arrays(mesh, with_T=True)
        sm, sv, sfl, so, sph = params.scalar_arrays(mesh)
        scalar_ptrs = (arr(sm, ctypes.c_int), arr(sv, _d), arr(sfl, _d), arr(so, _d), arr(sph, _d))
        oc, ob, zm, zh = params.bc_ext_arrays(mesh)
        ext = (arr(oc, ctypes.c_int), arr(ob, _d), arr(zm, _d), arr(zh, _d), params.wall_model_filter_time,
               params.wall_model_filter_tf_fac)
        from dopamine_amd.params import WALL_MODEL_IDS
        self._pr = _Params(params.nu, (_d * 3)(*params.f_const), params.scheme_id(),
                           0 if params.gradient_scheme.upper() == "LEAST_SQUARES" else 1, params.gg_iterations,
                           params.sgs_id(), params.c_vreman, params.cs_smagorinsky,
                           WALL_MODEL_IDS[params.wall_model], params.wall_model_height, params.n_nonortho_corr,
                           params.pressure_min_cos_theta, int(params.rhie_chow_dual), int(params.nonortho_check),
                           int(params.use_rotational_projection), params.p_tol, params.p_max_iter, params.cfl_max,
                           params.dt, arr(pv, _d), arr(uv.reshape(-1), _d), params.c_amd,
                           int(params.amd_wall_gradient), int(params.boussinesq), params.alpha_T, params.beta,
                           params.T_ref, (_d * 3)(*params.g), params.Pr_t, params.t_scheme_id(), arr(tv, _d),
                           params.z0_m, params.z0_h, params.T_surface, params.L_obukhov, int(params.compute_L),
                           params.projection_id(), params.cell_projection_tol, *params.forcing_args(),
                           params.n_scalars, params.scalar_scheme_id(), (_d * 4)(*params.scalar_diffusivity),
                           (_d * 4)(*params.scalar_pr_t), *scalar_ptrs, *ext)
        self.params = params
        self.mesh_nc = mesh.n_cells
        self.h = L.om_create(ctypes.byref(self._m), ctypes.byref(self._pr))
        if not self.h:
            raise ValueError("oracle om_create refused the set-up (MOST: first cell inside the roughness layer)")

    def __del__(self):
        if getattr(self, "h", None):
            lib().om_free(self.h)
            self.h = None

    def array(self, name):
        n = _i64(0)
        ptr = lib().om_array(self.h, name.encode(), ctypes.byref(n))
        if not ptr:
            raise KeyError(name)
        return np.ctypeslib.as_array(ptr, shape=(n.value,))

    def int_array(self, name):
        n = _i64(0)
        ptr = lib().om_int_array(self.h, name.encode(), ctypes.byref(n))
        if not ptr:
            raise KeyError(name)
        return np.ctypeslib.as_array(ptr, shape=(n.value,))

    @property
    def U(self):
        return self.array("U").reshape(-1, 3)

    @property
    def p(self):
        return self.array("p")

    def set_temperature(self, T):
        self.array("T")[:] = np.asarray(T, dtype=np.float64).reshape(-1)
        lib().om_apply_all_bc(self.h)

    def set_scalar(self, k, C):
        self.array(f"C{k}")[:] = np.asarray(C, dtype=np.float64).reshape(-1)
        lib().om_apply_all_bc(self.h)

    def scalar(self, k):
        return self.array(f"C{k}").copy()

    def set_velocity(self, U):
        self.array("U")[:] = np.asarray(U, dtype=np.float64).reshape(-1)
        lib().om_apply_all_bc(self.h)

    def eval_closures(self, dt_stab):
        """Gradients, SGS viscosity and wall traction of the current U (no time advance)."""
        lib().om_eval_closures(self.h, dt_stab)

    def project_once(self, dt):
        return lib().om_project_once(self.h, dt)

    def advance(self, dt=None):
        used = _d(0.0)
        rc = lib().om_advance(self.h, self.params.dt if dt is None else dt, ctypes.byref(used))
        if rc != 0:
            raise RuntimeError(f"oracle advance failed ({rc})")
        return used.value

    @property
    def time(self):
        return lib().om_time(self.h)

    def stats(self):
        out = (_d * 8)()
        lib().om_stats(self.h, out)
        return dict(zip(STATS_COLUMNS, list(out)))

    def cell_divergence(self):
        out = (_d * 2)()
        lib().om_divergence_cell(self.h, out)
        return {"max_abs_divU_cell": out[0], "volavg_abs_divU_cell": out[1]}

    def counters(self):
        out = (_i64 * 3)()
        lib().om_counters(self.h, out)
        return {"solves": out[0], "iterations": out[1], "last": out[2]}

    def proj_counters(self):
        out = (_i64 * 2)()
        lib().om_proj_counters(self.h, out)
        return {"projections": out[0], "iterations": out[1]}

    def proj_divergence(self, u):
        """B u on the solver's projection faces (cp_divergence)."""
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        out = np.zeros(self.mesh_nc)
        lib().om_proj_divergence(self.h, _p(u, _d), _p(out, _d))
        return out

    def proj_adjoint(self, chi):
        """V^-1 B^T chi (cp_adjoint_gradient), shape (nc, 3)."""
        chi = np.ascontiguousarray(chi, dtype=np.float64)
        v = np.zeros(3 * self.mesh_nc)
        lib().om_proj_adjoint(self.h, _p(chi, _d), _p(v, _d))
        return v.reshape(-1, 3)

    def proj_project(self, u, bconst=None, precond="pressure", tol=1e-1, max_it=100):
        """cp_project on a copy of u; returns (projected u, iterations)."""
        u = np.array(u, dtype=np.float64).reshape(-1)
        bc = np.zeros(self.mesh_nc) if bconst is None else np.ascontiguousarray(bconst, dtype=np.float64)
        its = lib().om_proj_project(self.h, _p(u, _d), _p(bc, _d), 0 if precond == "identity" else 1, tol, max_it)
        return u.reshape(-1, 3), its

    # single operator routines (src/numerics/mod_gradient.f90, mod_advection.f90, mod_viscous.f90,
    # src/sgs/mod_vreman.f90, mod_amd.f90, mod_wall_model.f90); the momentum ones read the solver's
    # U, U_bnd, U_face, grad_U and nu_sgs
    def op_gradient_vector(self, gg=False, iterations=1):
        out = np.zeros(9 * self.mesh_nc)
        lib().om_op_gradient_vector(self.h, int(gg), iterations, _p(out, _d))
        return out

    def op_gradient_scalar(self, phi, phi_bnd, gg=False, iterations=1):
        phi, pb = (np.ascontiguousarray(a, dtype=np.float64) for a in (phi, phi_bnd))
        out = np.zeros(3 * self.mesh_nc)
        lib().om_op_gradient_scalar(self.h, int(gg), iterations, _p(phi, _d), _p(pb, _d), _p(out, _d))
        return out

    def op_advection_vector(self, scheme):
        out = np.zeros(3 * self.mesh_nc)
        lib().om_op_advection_vector(self.h, scheme, _p(out, _d))
        return out

    def op_viscous_vector(self):
        out = np.zeros(3 * self.mesh_nc)
        lib().om_op_viscous_vector(self.h, _p(out, _d))
        return out

    def op_advection_scalar(self, scheme, phi, phi_bnd, grad_phi):
        phi, pb, g = (np.ascontiguousarray(a, dtype=np.float64) for a in (phi, phi_bnd, grad_phi))
        out = np.zeros(self.mesh_nc)
        lib().om_op_advection_scalar(self.h, scheme, _p(phi, _d), _p(pb, _d), _p(g, _d), _p(out, _d))
        return out

    def op_viscous_scalar(self, phi, phi_bnd, grad_phi, alpha, pr_t):
        phi, pb, g = (np.ascontiguousarray(a, dtype=np.float64) for a in (phi, phi_bnd, grad_phi))
        out = np.zeros(self.mesh_nc)
        lib().om_op_viscous_scalar(self.h, _p(phi, _d), _p(pb, _d), _p(g, _d), alpha, pr_t, _p(out, _d))
        return out

    def op_sgs(self, model, coef, wall_gradient, dt_stab):
        """model 1 Vreman, 2 Smagorinsky, 3 AMD on the solver's grad_U."""
        out = np.zeros(self.mesh_nc)
        lib().om_op_sgs(self.h, model, coef, int(wall_gradient), dt_stab, _p(out, _d))
        return out

    def op_wall_model(self, dt=0.0):
        out = np.zeros(self.array("tau_w_bnd").size)
        lib().om_op_wall_model(self.h, dt, _p(out, _d))
        return out

    def wall_sample(self, b):
        """Sampling cell and wall distance of boundary face b (the log-law wall model)."""
        y = _d(0.0)
        c = lib().om_wall_sample(self.h, b, ctypes.byref(y))
        return c, y.value

    def wall_filter_state(self):
        """(seeded, filtered wall-parallel velocity per boundary face (nb, 3))."""
        out = np.zeros(self.array("tau_w_bnd").size)
        ready = lib().om_wall_filter_state(self.h, _p(out, _d))
        return bool(ready), out.reshape(-1, 3)

    def outlet_speeds(self):
        """Per patch (convection speed U_c, reference speed U_ref) of the last stage."""
        out = np.zeros(2 * max(1, self.mesh.n_patches))
        lib().om_outlet_speeds(self.h, _p(out, _d))
        return out.reshape(-1, 2)

    def nonortho_info(self):
        out = (ctypes.c_int * 3)()
        lib().om_nonortho_info(self.h, out)
        return {"skip_nonortho": out[0], "skip_grad_p": out[1], "n_nonortho_corr": out[2]}


def outlet_dn_u(u_b, u_o, gdd, gt, len_d, nd):
    """outlet_dn_u (src/bc/mod_bc.f90:775-786) of the oracle."""
    return lib().om_outlet_dn_u(u_b, u_o, gdd, gt, len_d, nd)


def wall_tau(u_par, y_w, nu):
    """Reichardt-law wall shear stress u_tau^2 for speed u_par at wall distance y_w."""
    return lib().om_wall_tau(u_par, y_w, nu)


def amd_point(alpha, G, c_amd):
    """AMD closure at one point; alpha[i][j] = du_j/dx_i, G[i][j] (no cap)."""
    a = np.asarray(alpha, dtype=np.float64)
    g = np.asarray(G, dtype=np.float64)
    return lib().om_amd_point((_d * 9)(*a.T.reshape(-1)), (_d * 9)(*g.T.reshape(-1)), c_amd)


def most_iterate(u_par, T_o, T_surface, y_w, z0_m, z0_h, L_init):
    """MOST surface-layer solve at one face: (u_star, T_star, L)."""
    out = (_d * 3)()
    lib().om_most_iterate(u_par, T_o, T_surface, y_w, z0_m, z0_h, L_init, out)
    return tuple(out)


def psi_m(zeta):
    return lib().om_psi_m(zeta)


def psi_h(zeta):
    return lib().om_psi_h(zeta)


def set_threads(n):
    """OpenMP threads of the oracle's Jacobi CG (1: the sequential face loops; more: the same CG as
    row-parallel loops, used for the CPU baseline of the measurement)."""
    lib().om_set_threads(int(n))


def cg_jacobi(owner, neighbour, coeff, diag, b, x0, rtol, atol, max_it, nullspace):
    """Standalone Jacobi CG with the reference operator convention (0-based faces)."""
    x = np.array(x0, dtype=np.float64, copy=True)
    its = ctypes.c_int(0)
    res = _d(0.0)
    own = np.ascontiguousarray(owner, dtype=np.int32)
    nbr = np.ascontiguousarray(neighbour, dtype=np.int32)
    co = np.ascontiguousarray(coeff, dtype=np.float64)
    dg = np.ascontiguousarray(diag, dtype=np.float64)
    bb = np.ascontiguousarray(b, dtype=np.float64)
    rc = lib().om_cg_jacobi(len(bb), len(own), _p(own, _i32), _p(nbr, _i32), _p(co, _d), _p(dg, _d), _p(bb, _d),
                            _p(x, _d), rtol, atol, max_it, int(nullspace), ctypes.byref(its), ctypes.byref(res))
    return x, its.value, res.value, rc
