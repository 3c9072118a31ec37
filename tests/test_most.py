"""MOST surface-layer wall model of the oracle (src/abl/mod_abl_wall.f90, mod_abl_profiles.f90):
the Businger-Dyer corrections against their defining integrals, and the Newton solve against the
similarity profiles it inverts."""
import numpy as np
import pytest
from scipy.integrate import quad

import oracle

KAPPA, G = 0.41, 9.81


@pytest.mark.parametrize("zeta", [-2.0, -0.5, -0.05, 0.01, 0.3, 1.0])
def test_psi_functions_are_the_integrals_of_their_gradient_functions(zeta):
    """psi(zeta) = integral_0^zeta (1 - phi(z)) / z dz with phi_m = (1 - 16 z)^(-1/4),
    phi_h = (1 - 16 z)^(-1/2) when unstable and phi = 1 + 5 z when stable."""
    if zeta < 0:
        phi_m = lambda z: (1.0 - 16.0 * z) ** -0.25
        phi_h = lambda z: (1.0 - 16.0 * z) ** -0.5
    else:
        phi_m = phi_h = lambda z: 1.0 + 5.0 * z
    im = quad(lambda z: (1.0 - phi_m(z)) / z, 0.0, zeta, epsabs=1e-13, epsrel=1e-12)[0]
    ih = quad(lambda z: (1.0 - phi_h(z)) / z, 0.0, zeta, epsabs=1e-13, epsrel=1e-12)[0]
    assert abs(oracle.psi_m(zeta) - im) < 1e-9
    assert abs(oracle.psi_h(zeta) - ih) < 1e-9


def test_psi_is_continuous_at_neutral():
    for f in (oracle.psi_m, oracle.psi_h):
        assert f(0.0) == 0.0
        assert abs(f(-2e-6)) < 1e-4 and abs(f(2e-6)) < 1e-4


def _profile_misfit(u_par, T_o, T_s, y, z0m, z0h, us, ts, L):
    dm = np.log(y / z0m) - oracle.psi_m(y / L) + oracle.psi_m(z0m / L)
    dh = np.log(y / z0h) - oracle.psi_h(y / L) + oracle.psi_h(z0h / L)
    return us / KAPPA * dm - u_par, ts / KAPPA * dh - (T_o - T_s)


@pytest.mark.parametrize("dT", [-3.0, -0.5, 0.5, 3.0])
def test_most_solve_reproduces_the_similarity_profiles(dT):
    """The returned (u*, T*, L) put the sampled speed and temperature on the MOST profiles, with
    L = u*^2 T / (kappa g T*)."""
    u_par, T_s, y, z0m = 6.0, 300.0, 10.0, 0.1
    z0h = z0m * np.exp(-2.0)
    T_o = T_s + dT
    us, ts, L = oracle.most_iterate(u_par, T_o, T_s, y, z0m, z0h, 0.0)
    eu, eT = _profile_misfit(u_par, T_o, T_s, y, z0m, z0h, us, ts, L)
    assert abs(eu) < 1e-4 * u_par and abs(eT) < 1e-4 * abs(dT)
    assert abs(L - us * us * T_o / (KAPPA * G * ts)) < 1e-3 * abs(L)
    # stable (warm air over a cold surface) damps the momentum flux, unstable enhances it
    neutral = KAPPA * u_par / np.log(y / z0m)
    assert (us < neutral) if dT > 0 else (us > neutral)
    assert np.sign(ts) == np.sign(dT) and np.sign(L) == np.sign(dT)


def test_most_neutral_limit_is_the_rough_log_law():
    u_par, y, z0m = 5.0, 2.0, 0.01
    us, ts, L = oracle.most_iterate(u_par, 300.0, 300.0, y, z0m, z0m * np.exp(-2.0), 0.0)
    assert ts == 0.0 and L == 5000.0
    want = KAPPA * u_par / (np.log(y / z0m) - oracle.psi_m(y / 5000.0) + oracle.psi_m(z0m / 5000.0))
    assert abs(us - want) < 1e-7 * want  # relaxed Newton, |Fu| < 1e-8
