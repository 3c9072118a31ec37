"""Solver parameters: the fields of the reference's solver_params_t that the hot path reads.

Defaults are the reference's &dopamine_numerics / &dopamine_physics defaults
(src/io/mod_input_parser.f90:178-211, src/solver/mod_fv_solver.f90:41-110).
"""
from dataclasses import dataclass, field
import math

import numpy as np

SCHEME_IDS = {"VANLEER_TVD": 1, "VANLEER": 1, "KOREN_TVD": 2, "KOREN": 2, "CDS": 3, "CENTRAL": 3, "KEP": 4,
              "SKEW": 4, "QUICK": 5}
SGS_IDS = {"NONE": 0, "VREMAN": 1, "SMAGORINSKY": 2, "AMD": 3, "DYNAMIC_SMAG": 4}
PRECOND_IDS = {"JACOBI": 0, "LEARNED": 1, "MULTIGRID": 2}
WALL_MODEL_IDS = {"loglaw": 0, "none": 1, "most": 2}

MAX_SCALARS = 4
SCALAR_DIRICHLET, SCALAR_NEUMANN, SCALAR_ZERO_GRADIENT = 1, 2, 3  # src/mod_scalar.f90:13-15

# Wray RK3 coefficients (src/mod_parameters.f90:64-71)
RK3W_GAMMA1 = 8.0 / 15.0


@dataclass
class SolverParams:
    nu: float = 1.0e-5
    f_const: tuple = (0.0, 0.0, 0.0)
    advection_scheme: str = "CDS"            # parser default 'CENTRAL'
    gradient_scheme: str = "LEAST_SQUARES"   # or GREEN_GAUSS
    gg_iterations: int = 2
    use_sgs: bool = True
    sgs_model: str = "VREMAN"
    c_vreman: float = 0.025                  # C_VREMAN, src/mod_parameters.f90:20
    cs_smagorinsky: float = 0.1
    c_amd: float = 1.0 / 3.0                 # AMD_C_DEFAULT, src/mod_parameters.f90:23
    amd_wall_gradient: bool = True
    wall_model: str = "loglaw"
    wall_model_height: float = 0.0
    # MOST surface layer (wall_model "most", src/abl/mod_abl_types.f90:22-27): roughness lengths
    # (z0_h <= 0: z0_m exp(-2), the patch default), surface temperature, initial Obukhov length
    # (0 = neutral), domain-mean L update at the first stage
    z0_m: float = 0.1
    z0_h: float = 0.0
    T_surface: float = 300.0
    L_obukhov: float = 0.0
    compute_L: bool = True
    n_nonortho_corr: int = 0                 # 0 = auto
    pressure_min_cos_theta: float = 0.0
    rhie_chow_dual: bool = True
    nonortho_check: bool = False
    # cell velocity correction: "LS" (correct_velocity_psolver_dual) or "EXACT" (V-orthogonal
    # projection by flexible PCG with the pressure solve as preconditioner); the reference's
    # parser default is EXACT (src/io/mod_input_parser.f90:200)
    cell_projection: str = "EXACT"
    cell_projection_tol: float = 1.0e-1
    use_rotational_projection: bool = True
    p_tol: float = 1.0e-10
    p_max_iter: int = 1000
    cfl_max: float = 0.4
    dt: float = 1.0e-3
    adapt_dt: bool = False
    # pressure CG recurrence: 0 two passes per iteration (KSPCG's), 1 single reduction
    cg_variant: int = 0
    # temperature / Boussinesq (src/solver/mod_fv_solver.f90:41-110 defaults)
    boussinesq: bool = False
    alpha_T: float = 1.4e-5
    beta: float = 3.4e-3
    T_ref: float = 300.0
    g: tuple = (0.0, 0.0, 0.0)
    Pr_t: float = 0.85
    temperature_scheme: str = "VANLEER_TVD"
    patch_T: dict = field(default_factory=dict)
    # pressure preconditioner: "JACOBI", "LEARNED" (flexible CG with the message-passing network,
    # trained on residuals of the first learn_collect Jacobi solves; Jacobi CG is the fallback) or
    # "MULTIGRID" (flexible CG with the FP32 aggregation-multigrid V-cycle; in mixed precision the
    # FP32 sweeps are that flexible CG)
    precond: str = "JACOBI"
    learn_collect: int = 600
    learn_capacity: int = 128
    learn_iters: int = 300
    learn_batch: int = 4
    learn_lr: float = 3.0e-3
    learn_seed: int = 1
    # momentum forcing (&dopamine_physics, src/io/mod_input_parser.f90:392-438): Coriolis -2 Omega x U,
    # harmonic forcing amp cos(omega t + phase) along forcing_dir, constant mass flux (the uniform
    # force (U_bulk_target - U_bulk) / dt of each step replaces f_const, which must then be 0)
    use_rotation: bool = False
    omega_rot: tuple = (0.0, 0.0, 0.0)
    use_harmonic_forcing: bool = False
    forcing_amp: float = 0.0
    forcing_omega: float = 0.0
    forcing_phase: float = 0.0
    forcing_dir: tuple = (1.0, 0.0, 0.0)
    constant_mass_flux_forcing: bool = False
    U_bulk_target: tuple = (0.0, 0.0, 0.0)
    # passive scalars (&dopamine_scalars; src/solver/mod_fv_solver.f90:56-78): count (at most 4),
    # advection scheme, molecular diffusivity and turbulent Schmidt number per scalar; per patch
    # name a list with one entry per scalar: {"mode": "dirichlet"|"neumann"|"zero_gradient",
    # "value", "flux", "omega", "phase"} (patch_scalar_*, src/io/mod_input_parser.f90:465-546)
    n_scalars: int = 0
    scalar_scheme: str = "VANLEER_TVD"
    scalar_diffusivity: tuple = (0.0, 0.0, 0.0, 0.0)
    scalar_pr_t: tuple = (1.0, 1.0, 1.0, 1.0)
    patch_scalar: dict = field(default_factory=dict)
    # per-patch values, by patch name: p_val (outlet), U_val (inlet / moving wall)
    patch_p: dict = field(default_factory=dict)
    patch_U: dict = field(default_factory=dict)
    # outlet patches (src/io/mod_input_parser.f90:481-521): 'extrapolate' (default) or 'convective'
    # (du_b/dt = -U_c n.grad u, src/bc/mod_bc.f90:681-786); the backflow-traction coefficient
    # (< 0: 1 for a convective outlet, else 0; src/bc/mod_bc.f90:789-818)
    patch_outlet_mode: dict = field(default_factory=dict)
    patch_backflow_beta: dict = field(default_factory=dict)
    # MOST roughness per wall patch (0 inherits z0_m / z0_h; a negative z0_h derives it from z0_m)
    patch_z0_m: dict = field(default_factory=dict)
    patch_z0_h: dict = field(default_factory=dict)
    # log-law wall model time filter (src/sgs/mod_wall_model.f90:176-313): the exponential filter's
    # time constant (<= 0 off) and the adaptive factor f of T_f = f y / (kappa u_tau) (<= 0 off)
    wall_model_filter_time: float = 0.0
    wall_model_filter_tf_fac: float = 0.0

    def validate(self):
        if self.cell_projection.upper() not in ("LS", "EXACT"):
            raise ValueError(f"cell_projection {self.cell_projection!r} not supported (LS, EXACT)")
        if not self.cell_projection_tol > 0.0:
            raise ValueError("cell_projection_tol must be positive")
        if self.advection_scheme.upper() not in SCHEME_IDS:
            raise ValueError(f"advection_scheme {self.advection_scheme!r} not supported "
                             "(KOREN_TVD, VANLEER_TVD, QUICK, KEP, CDS)")
        if self.gradient_scheme.upper() not in ("LEAST_SQUARES", "GREEN_GAUSS"):
            raise ValueError("gradient_scheme must be LEAST_SQUARES or GREEN_GAUSS")
        if self.use_sgs and self.sgs_model.upper() not in SGS_IDS:
            raise ValueError(f"sgs_model {self.sgs_model!r} not supported (VREMAN, SMAGORINSKY, DYNAMIC_SMAG, AMD)")
        if self.boussinesq and self.temperature_scheme.upper() not in ("CDS", "CENTRAL", "QUICK", "VANLEER_TVD",
                                                                         "VANLEER", "KOREN_TVD", "KOREN"):
            raise ValueError(f"temperature_scheme {self.temperature_scheme!r} not supported")
        if self.wall_model not in WALL_MODEL_IDS:
            raise ValueError(f"wall_model {self.wall_model!r} not supported (loglaw, most, none)")
        if self.precond.upper() not in PRECOND_IDS:
            raise ValueError(f"precond {self.precond!r} not supported (JACOBI, LEARNED, MULTIGRID)")
        if self.constant_mass_flux_forcing and any(f != 0.0 for f in self.f_const):
            raise ValueError("f_const must be 0 with constant_mass_flux_forcing")
        if not 0 <= self.n_scalars <= MAX_SCALARS:
            raise ValueError(f"n_scalars must be 0..{MAX_SCALARS}")
        if self.n_scalars and self.scalar_scheme.upper() not in ("CDS", "CENTRAL", "QUICK", "VANLEER_TVD", "VANLEER",
                                                                 "KOREN_TVD", "KOREN"):
            raise ValueError(f"scalar_scheme {self.scalar_scheme!r} not supported")
        if self.adapt_dt:
            raise ValueError("adapt_dt is not supported; the fixed-dt CFL clamp of solver_advance applies")
        for name, mode in self.patch_outlet_mode.items():
            if str(mode).strip().lower() not in ("extrapolate", "convective"):
                raise ValueError(f"unrecognised patch_outlet_mode {mode!r} for patch {name!r}")

    def bc_ext_arrays(self, mesh):
        """Per-patch convective-outlet flag (int32), backflow-traction coefficient and MOST roughness
        lengths in the mesh's patch order, as read_bc resolves them (mod_input_parser.f90:504-529)."""
        n = max(1, mesh.n_patches)
        conv = np.zeros(n, dtype=np.int32)
        beta, z0m, z0h = np.zeros(n), np.zeros(n), np.zeros(n)
        for i, name in enumerate(mesh.patch_names[:mesh.n_patches]):
            conv[i] = 1 if str(self.patch_outlet_mode.get(name, "extrapolate")).strip().lower() == "convective" else 0
            b = float(self.patch_backflow_beta.get(name, -1.0))
            beta[i] = (1.0 if conv[i] else 0.0) if b < 0.0 else b
            z0m[i] = float(self.patch_z0_m.get(name, 0.0))
            z0h[i] = float(self.patch_z0_h.get(name, 0.0))
        return conv, beta, z0m, z0h

    def scalar_arrays(self, mesh):
        """Per-patch scalar BC arrays [MAX_SCALARS * patch + scalar]: mode (int32), value, flux,
        omega, phase. Modes by name as scalar_bc_mode_from_name (src/mod_scalar.f90:63-73)."""
        n = max(1, mesh.n_patches) * MAX_SCALARS
        mode = np.full(n, SCALAR_ZERO_GRADIENT, dtype=np.int32)
        vals = {k: np.zeros(n) for k in ("value", "flux", "omega", "phase")}
        for p, name in enumerate(mesh.patch_names[:mesh.n_patches]):
            for k, spec in enumerate(self.patch_scalar.get(name, [])[:MAX_SCALARS]):
                q = MAX_SCALARS * p + k
                m = str(spec.get("mode", "zero_gradient")).strip().lower()
                mode[q] = {"dirichlet": SCALAR_DIRICHLET, "fixed": SCALAR_DIRICHLET, "fixedvalue": SCALAR_DIRICHLET,
                           "neumann": SCALAR_NEUMANN, "flux": SCALAR_NEUMANN}.get(m, SCALAR_ZERO_GRADIENT)
                for key in vals:
                    vals[key][q] = float(spec.get(key, 0.0))
        return mode, vals["value"], vals["flux"], vals["omega"], vals["phase"]

    def forcing_args(self):
        """The forcing fields in the C ABI's order (after cell_projection_tol)."""
        import ctypes
        d3 = ctypes.c_double * 3
        return (int(self.use_rotation), d3(*self.omega_rot), int(self.use_harmonic_forcing), self.forcing_amp,
                self.forcing_omega, self.forcing_phase, d3(*self.forcing_dir), int(self.constant_mass_flux_forcing),
                d3(*self.U_bulk_target))

    def projection_id(self):
        return 1 if self.cell_projection.upper() == "EXACT" else 0

    def scalar_scheme_id(self):
        return SCHEME_IDS[self.scalar_scheme.upper()]

    def scheme_id(self):
        return SCHEME_IDS[self.advection_scheme.upper()]

    def sgs_id(self):
        return SGS_IDS[self.sgs_model.upper()] if self.use_sgs else 0

    def t_scheme_id(self):
        return SCHEME_IDS[self.temperature_scheme.upper()]

    def patch_arrays(self, mesh, with_T=False):
        """Per-patch outlet pressure, inlet velocity (and fixed temperature) arrays in the mesh's
        patch order."""
        pv = np.zeros(mesh.n_patches)
        uv = np.zeros((mesh.n_patches, 3))
        tv = np.full(mesh.n_patches, self.T_ref)
        for i, name in enumerate(mesh.patch_names[:mesh.n_patches]):
            pv[i] = self.patch_p.get(name, 0.0)
            uv[i] = self.patch_U.get(name, (0.0, 0.0, 0.0))
            tv[i] = self.patch_T.get(name, self.T_ref)
        return (pv, uv, tv) if with_T else (pv, uv)


def taylor_green_ic(mesh, uref=1.0):
    """Taylor-Green initial velocity on the mesh's bounding box (src/init/mod_initconditions.f90:276-334)."""
    lo, hi = mesh.bounds_min, mesh.bounds_max
    L = hi - lo
    L[L <= 0.0] = 1.0
    k = 2.0 * math.pi / L
    x = np.mod(mesh.cell_centroid - lo, L)
    tx, ty, tz = k[0] * x[:, 0], k[1] * x[:, 1], k[2] * x[:, 2]
    U = np.zeros((mesh.n_cells, 3))
    U[:, 0] = uref * np.sin(tx) * np.cos(ty) * np.cos(tz)
    U[:, 1] = -uref * np.cos(tx) * np.sin(ty) * np.cos(tz)
    return U


def channel_ic(mesh, uref=1.0, wall_dir=1, noise=0.0, seed=395):
    """Laminar parabolic profile u = 4 uref z (1 - z) across the wall-normal direction
    (src/init/mod_initconditions.f90:114-196), with multiplicative uniform noise
    u_k *= 1 + noise (2 r - 1) as add_random_velocity_noise applies it (:200-250). The random
    stream is numpy's seeded PCG64, not gfortran's random_number."""
    lo, hi = mesh.bounds_min, mesh.bounds_max
    H = max(hi[wall_dir] - lo[wall_dir], 1e-30)
    stream = 0 if wall_dir != 0 else 1
    zeta = np.clip((mesh.cell_centroid[:, wall_dir] - lo[wall_dir]) / H, 0.0, 1.0)
    U = np.zeros((mesh.n_cells, 3))
    U[:, stream] = 4.0 * uref * zeta * (1.0 - zeta)
    if noise > 0.0:
        r = np.random.default_rng(seed).random((mesh.n_cells, 3))
        U = U * (1.0 + noise * (2.0 * r - 1.0))
    return U


def initial_velocity(case, mesh):
    """Initial cell velocity of a case, and whether a second projection pass follows (noise > 0)."""
    noise = case.extra.get("noise", 0.0)
    if case.ic == "taylor_green":
        U = taylor_green_ic(mesh, case.ic_uref)
        if noise > 0.0:  # additive seeded Gaussian perturbation (benchmark workload definition)
            U = U + noise * np.random.default_rng(case.extra.get("seed", 1600)).standard_normal(U.shape)
    elif case.ic == "channel_parabolic":
        U = channel_ic(mesh, case.ic_uref, case.extra.get("wall_dir", 1), noise, case.extra.get("seed", 395))
    else:
        U = np.zeros((mesh.n_cells, 3))
        U[:, 0] = case.ic_uref
    return U, noise > 0.0


def scalar_advection_ic(mesh, kind):
    """initialize_scalar_advection_ic (src/init/mod_initconditions.f90:336-360): with s the cell's
    x position as a fraction of the domain length (periodic), a Gaussian exp(-d^2 / (2 0.05^2)) at
    s = 0.25 (d the periodic distance), plus 1 on 0.55 <= s < 0.75 for 'scalar_pulse'. The domain
    bounds are the mesh vertices' (compute_domain_bounds, :252-274)."""
    xmin, xmax = mesh.bounds_min[0], mesh.bounds_max[0]
    lx = max(xmax - xmin, 1e-30)
    s = np.mod(mesh.cell_centroid[:, 0] - xmin, lx) / lx
    d = s - 0.25
    d = d - np.where(d >= 0.0, np.floor(d + 0.5), np.ceil(d - 0.5))  # Fortran anint: halves away from zero
    f = np.exp(-0.5 * (d / 0.05) ** 2)
    if kind == "scalar_pulse":
        f = f + ((s >= 0.55) & (s < 0.75))
    return f


def initialize(solver, case, mesh):
    """apply_initial_conditions (src/init/mod_initconditions.f90:39-105): set U, then one projection
    with dt = gamma_1 dt (two when noise was added)."""
    U, twice = initial_velocity(case, mesh)
    if case.ic in ("scalar_gauss", "scalar_pulse"):
        solver.set_temperature(case.params.T_ref + scalar_advection_ic(mesh, case.ic))
    elif case.params.boussinesq:
        # temperature: T_ref plus a seeded Gaussian perturbation of amplitude t_noise
        T = np.full(mesh.n_cells, case.params.T_ref)
        tn = case.extra.get("t_noise", 0.0)
        if tn > 0.0:
            T = T + tn * np.random.default_rng(case.extra.get("seed", 1700)).standard_normal(mesh.n_cells)
        solver.set_temperature(T)
    solver.set_velocity(U)
    dtp = RK3W_GAMMA1 * case.params.dt
    if twice:
        solver.project_once(dtp)
    solver.project_once(dtp)
