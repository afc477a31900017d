"""Pins the oracle's Ewald coefficient and correction tables.

No reference fixture covers Ewald, so these follow what the mathematics
fixes: erfc(beta rc) = rtol (ewald_utils.cpp:46-74), the closed forms the
reference's SIMD test uses for the analytic correction
(simd/tests/simd_math.cpp:938-977), the table's stated error bounds
(tables/forcetable.cpp:182-186) and F = -dV/dr.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O


@pytest.mark.parametrize("rc,expected", [(1.0, 3.123413), (1.2, 2.602844)])
def test_ewald_coeff_solves_erfc(rc, expected):
    for prec in ("f32", "f64"):
        beta = O.ewald_coeff(rc, 1e-5, prec)
        assert abs(beta - expected) < 2e-6
    beta = O.ewald_coeff(rc, 1e-5, "f64")
    assert math.erfc(beta * rc) == pytest.approx(1e-5, rel=1e-9)


def test_ewald_coeff_other_tolerances():
    for rtol in (1e-3, 1e-4, 1e-6):
        beta = O.ewald_coeff(0.9, rtol, "f64")
        assert math.erfc(beta * 0.9) == pytest.approx(rtol, rel=1e-9)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_table_scale_and_size(prec):
    beta = O.ewald_coeff(1.0, 1e-5, prec)
    scale, F, V, FDV0 = O.ewald_tables(1.0, beta, prec)
    assert scale == pytest.approx(1071.77, abs=0.01)
    assert len(F) == int(1.0 * scale) + 2 == 1073
    beta = O.ewald_coeff(1.2, 1e-5, prec)
    scale, F, V, FDV0 = O.ewald_tables(1.2, beta, prec)
    assert len(F) == 1176


def test_table_values_and_error_bounds():
    beta = O.ewald_coeff(1.0, 1e-5, "f64")
    scale, F, V, FDV0 = O.ewald_tables(1.0, beta, "f64")
    dx = 1.0 / scale
    r = np.arange(len(V)) * dx
    v_exact = np.array([O.ewald_v_exact(beta, ri) for ri in r])
    np.testing.assert_allclose(V, v_exact, rtol=1e-15, atol=0)
    # FDV0 quadruplets {F_i, F_{i+1} - F_i, V_i, 0}
    q = FDV0.reshape(-1, 4)
    np.testing.assert_array_equal(q[:, 0], F)
    np.testing.assert_array_equal(q[:-1, 1], F[1:] - F[:-1])
    np.testing.assert_array_equal(q[:, 2], V)
    np.testing.assert_array_equal(q[:, 3], 0)
    # linear interpolation of F against -dV/dr: the table comment states a
    # maximum force error of V'''/(6*4) dx^2 with max |(erf(x)/x)'''| = 1.0522
    # (scaled by beta^4); the averaged node derivatives reach twice that at
    # r ~ 0.2 nm.  A mis-stated spline step would miss by ~F' dx ~ 1e-2.
    bound = 2.0 * 1.0522 * beta ** 4 / 24 * dx * dx
    rs = np.linspace(0.001, 0.999, 4001)
    t = rs * scale
    i = t.astype(int)
    frac = t - i
    f_tab = (1 - frac) * F[i] + frac * F[i + 1]
    f_ex = np.array([O.ewald_f_exact(beta, x) for x in rs])
    assert np.max(np.abs(f_tab - f_ex)) <= 1.5 * bound
    # energy interpolation, as the kernels use it
    v_tab = V[i] - 0.5 / scale * frac * (F[i] + f_tab)
    v_ex = np.array([O.ewald_v_exact(beta, x) for x in rs])
    ebound = 1.0522 * beta ** 4 / (6 * 12 * math.sqrt(3)) * dx ** 3
    assert np.max(np.abs(v_tab - v_ex)) <= 50 * ebound + 1e-12


def test_exact_force_is_minus_dv_dr():
    beta = 3.123413
    h = 1e-6
    for r in np.linspace(0.01, 1.2, 61):
        fd = -(O.ewald_v_exact(beta, r + h) - O.ewald_v_exact(beta, r - h)) / (2 * h)
        assert O.ewald_f_exact(beta, r) == pytest.approx(fd, rel=1e-7, abs=1e-9)
    # small-r series branch continues the closed form
    r = 0.9e-3 / beta
    fd = -(O.ewald_v_exact(beta, r + 1e-9) - O.ewald_v_exact(beta, r - 1e-9)) / 2e-9
    assert O.ewald_f_exact(beta, r) == pytest.approx(fd, rel=1e-5)
    assert O.ewald_v_exact(beta, 0) == pytest.approx(2 * beta / math.sqrt(math.pi), rel=1e-15)


def test_closed_forms_of_simd_math_test():
    # pmeForceCorrection(x) = 2 exp(-x) / (sqrt(pi) x) - erf(sqrt x) / x^1.5
    # pmePotentialCorrection(x) = erf(sqrt x) / sqrt x, x in [0.15, 4]
    beta = 2.0
    for x in np.linspace(0.15, 4.0, 40):
        z = math.sqrt(x)
        r = z / beta
        force_corr = 2 * math.exp(-x) / (math.sqrt(math.pi) * x) - math.erf(z) / x ** 1.5
        pot_corr = math.erf(z) / z
        assert O.ewald_f_exact(beta, r) / r == pytest.approx(-beta ** 3 * force_corr, rel=1e-12)
        assert O.ewald_v_exact(beta, r) == pytest.approx(beta * pot_corr, rel=1e-14)


def test_one_4pi_eps0():
    assert O.one_4pi_eps0() == pytest.approx(138.935457, rel=1e-8)
