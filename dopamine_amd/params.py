"""Solver parameters: the fields of the reference's solver_params_t that the hot path reads.

Defaults are the reference's &dopamine_numerics / &dopamine_physics defaults
(src/io/mod_input_parser.f90:178-211, src/solver/mod_fv_solver.f90:41-110).
"""
from dataclasses import dataclass, field
import math

import numpy as np

SCHEME_IDS = {"CDS": 3, "CENTRAL": 3, "KEP": 4, "SKEW": 4}
SGS_IDS = {"NONE": 0, "VREMAN": 1, "SMAGORINSKY": 2}
WALL_MODEL_IDS = {"loglaw": 0, "none": 1}

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
    wall_model: str = "loglaw"
    wall_model_height: float = 0.0
    n_nonortho_corr: int = 0                 # 0 = auto
    pressure_min_cos_theta: float = 0.0
    rhie_chow_dual: bool = True
    nonortho_check: bool = False
    cell_projection: str = "EXACT"
    use_rotational_projection: bool = True
    p_tol: float = 1.0e-10
    p_max_iter: int = 1000
    cfl_max: float = 0.4
    dt: float = 1.0e-3
    adapt_dt: bool = False
    # per-patch values, by patch name: p_val (outlet), U_val (inlet / moving wall)
    patch_p: dict = field(default_factory=dict)
    patch_U: dict = field(default_factory=dict)

    def validate(self):
        if self.cell_projection.upper() != "LS":
            raise ValueError("cell_projection must be 'LS' (the EXACT V-orthogonal projection is not part of this "
                             "hot path)")
        if self.advection_scheme.upper() not in SCHEME_IDS:
            raise ValueError(f"advection_scheme {self.advection_scheme!r} not supported (KEP, CDS)")
        if self.gradient_scheme.upper() not in ("LEAST_SQUARES", "GREEN_GAUSS"):
            raise ValueError("gradient_scheme must be LEAST_SQUARES or GREEN_GAUSS")
        if self.use_sgs and self.sgs_model.upper() not in ("VREMAN", "SMAGORINSKY"):
            raise ValueError(f"sgs_model {self.sgs_model!r} not supported (VREMAN, SMAGORINSKY)")
        if self.adapt_dt:
            raise ValueError("adapt_dt is not supported; the fixed-dt CFL clamp of solver_advance applies")

    def scheme_id(self):
        return SCHEME_IDS[self.advection_scheme.upper()]

    def sgs_id(self):
        return SGS_IDS[self.sgs_model.upper()] if self.use_sgs else 0

    def patch_arrays(self, mesh):
        """Per-patch outlet pressure and inlet velocity arrays in the mesh's patch order."""
        pv = np.zeros(mesh.n_patches)
        uv = np.zeros((mesh.n_patches, 3))
        for i, name in enumerate(mesh.patch_names[:mesh.n_patches]):
            pv[i] = self.patch_p.get(name, 0.0)
            uv[i] = self.patch_U.get(name, (0.0, 0.0, 0.0))
        return pv, uv


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
