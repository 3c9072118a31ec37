"""The cos(theta) floor of the over-relaxed non-orthogonal k-vector (nonortho_kfactor_limited,
src/numerics/mod_nonortho.f90:65-98), with the known answers of the reference's own unit test
(tests/unit/test_nonortho_limiter.f90:14-41), for the oracle and for the product's set-up routine
(host-check build). The over-relaxed vector (Jasak) is k = Sf - (|Sf|^2 / Sf.d) d."""
import ctypes

import numpy as np
import pytest

import oracle
import hostcheck

_d3 = ctypes.c_double * 3


def _oracle(sf, d, cmin):
    L = oracle.lib()
    L.om_kfactor_limited.argtypes = [_d3, _d3, ctypes.c_double, _d3]
    k = _d3()
    L.om_kfactor_limited(_d3(*sf), _d3(*d), cmin, k)
    return np.array(k[:])


def _product(sf, d, cmin):
    L = hostcheck.lib()
    L.hc_kfactor_limited.argtypes = [_d3, _d3, ctypes.c_double, _d3]
    k = _d3()
    L.hc_kfactor_limited(_d3(*sf), _d3(*d), cmin, k)
    return np.array(k[:])


@pytest.mark.parametrize("kf", [_oracle, _product])
def test_reference_known_answers(kf):
    # orthogonal face: floor inactive, k = 0
    assert np.abs(kf([2.0, 0.0, 0.0], [0.5, 0.0, 0.0], 0.3)).max() == 0.0
    # floor 0 is the over-relaxed vector, even at 89.4 degrees
    sf, d = np.array([1.0, 0.0, 0.0]), np.array([0.01, 1.0, 0.0])
    k0 = sf - (sf @ sf) / (sf @ d) * d
    np.testing.assert_allclose(kf(sf, d, 0.0), k0, rtol=1e-12, atol=1e-12)
    # a floor below the face's cos(theta) changes nothing
    np.testing.assert_allclose(kf(sf, d, 0.005), k0, rtol=1e-12, atol=1e-12)
    # active floor: the implicit coefficient a = (Sf - k).d / (d.d) is bounded by |Sf| / (|d| cmin)
    a0 = (sf[0] - k0[0]) / d[0]
    k1 = kf(sf, d, 0.2)
    a = (sf[0] - k1[0]) / d[0]
    assert a0 > 50.0
    assert a == pytest.approx(1.0 / (np.linalg.norm(d) * 0.2), rel=1e-12)
    # linear fields stay exact: a (d.g) + k.g = Sf.g
    g = np.array([0.7, -1.3, 2.1])
    assert a * (d @ g) + k1 @ g == pytest.approx(sf @ g, rel=1e-12)
