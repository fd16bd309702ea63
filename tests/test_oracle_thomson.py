"""Pin the oracle's Thomson-scattering factors (row A5, fomo-thomson.cpp:39-92).

The van de Hulst (1950) coefficients for a finite solar disk (see also Billings 1966) are
evaluated here independently with mpmath at 50 digits, and their point-source limit
(A ~ s^2, B ~ 2s^2/3, C ~ s^2, D ~ 2s^2/3 for s = Rsun/r -> 0) is checked in closed form.
"""
import mpmath as mp
import numpy as np
import pytest

from oracle import oracle as O

RSUN = 695.7
U = 0.58
COEFF = 0.5 * mp.pi * mp.mpf("2.8179e-13") ** 2 * mp.mpf("2.8e10")


def vdh(r):
    mp.mp.dps = 50
    s = mp.mpf(RSUN) / mp.mpf(r)
    c = mp.sqrt(1 - s * s)
    L = mp.log((1 + s) / c)
    A = c * s * s
    B = -mp.mpf(1) / 8 * (1 - 3 * s * s - (c * c / s) * (1 + 3 * s * s) * L)
    C = mp.mpf(4) / 3 - c - c ** 3 / 3
    D = mp.mpf(1) / 8 * (5 + s * s - (c * c / s) * (5 - s * s) * L)
    k = COEFF / (1 - mp.mpf(U) / 3)
    pol = k * (A + U * (B - A))
    tot = k * 2 * (C + U * (D - C))
    return float(tot), float(pol)


@pytest.mark.parametrize("r_rsun", [1.001, 1.05, 1.5, 2.0, 3.0, 10.0])
def test_factors_match_mpmath(r_rsun):
    r = r_rsun * RSUN
    t1, p1 = O.thomson_factor(1, r)
    t0, p0 = O.thomson_factor(0, r)
    T, P = vdh(r)
    assert abs(t1 / T - 1) < 1e-11 and abs(p1 / P - 1) < 1e-11
    # the reference's float pipeline: float rounding of r, sin(omega) and the coefficients
    assert abs(t0 / T - 1) < 2e-5 and abs(p0 / P - 1) < 2e-5


def test_point_source_limit():
    r = 1000 * RSUN
    s2 = (RSUN / r) ** 2
    t, p = O.thomson_factor(1, r)
    coeff = float(COEFF)
    # pol -> Coeff s^2, tot -> 2 Coeff s^2 (relative corrections O(s^2) ~ 1e-6)
    assert abs(p / (coeff * s2) - 1) < 1e-5
    assert abs(t / (2 * coeff * s2) - 1) < 1e-5


def test_inside_sun_is_nan():
    t, p = O.thomson_factor(1, 0.5 * RSUN)
    assert np.isnan(t) and np.isnan(p)


def test_emission_multiplies_density():
    x = np.array([1.5 * RSUN, 0.0, 0.0], np.float32)
    y = np.array([0.0, 2.0 * RSUN, 0.0], np.float32)
    z = np.array([0.0, 0.0, -2.5 * RSUN], np.float32)
    n = np.array([1e8, 3e7, 1e6], np.float32)
    tot, pol = O.thomson_emission(1, x, y, z, n)
    for i, r in enumerate([1.5, 2.0, 2.5]):
        T, P = vdh(np.float64(np.float32(r * RSUN)))
        nn = float(n[i])
        assert abs(tot[i] / (nn * T) - 1) < 1e-10 and abs(pol[i] / (nn * P) - 1) < 1e-10
