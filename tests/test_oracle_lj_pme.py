"""Pins the oracle's LJ-PME real-space flavour (kernel_ref_inner.h:207-248,
self term kernel_ref_outer.h:315-322, constants forcerec.cpp:703-717,
grid parameters atomdata.cpp:288-330) with closed forms computed here in
numpy, independently of the oracle's constants:
  * beta_lj solves g(beta rc) = rtol, g(x) = exp(-x^2) (1 + x^2 + x^4 / 2);
  * for two atoms the energy minus the plain cut-off's is
    C6g ((1 - g(beta r)) / r^6 - (1 - g(beta rc)) / rc^6) plus the two self
    terms, and the force difference is minus its derivative, with
    C6g = sqrt(C6_i C6_j) (geometric) or 4 sqrt(eps_i eps_j)
    ((sigma_i + sigma_j) / 2)^6 (Lorentz-Berthelot);
  * an excluded pair keeps the grid correction without the shift;
  * each atom's self energy is half the r -> 0 limit of C6 (1 - g(beta r)) / r^6
    (found here by extrapolation, not from the series);
  * forces are minus the energy gradient on a heterogeneous system.
"""
import dataclasses

import numpy as np
import pytest

from gmxapi_amd import systems as S
from oracle import oracle as O
from tests.helpers import oracle_for

RC = 1.0
RTOL = 1e-3


def _g(x):
    x2 = x * x
    return np.exp(-x2) * (1 + x2 + 0.5 * x2 * x2)


def test_beta_solves_the_tail_condition():
    o = O.Oracle("f64", "rf", rlist=RC, rcoulomb=RC, lj_pme="geometric", ewald_rtol_lj=RTOL)
    beta = o.lj_pme_beta()
    assert _g(beta * RC) == pytest.approx(RTOL, rel=1e-9)
    assert O._lib("f64").orc_ewald_coeff_lj(1.2, 1e-4) > O._lib("f64").orc_ewald_coeff_lj(1.2, 1e-3)


SIG = np.array([0.30, 0.38])
EPS = np.array([0.40, 0.90])


def _two_atoms(r, lj_pme, excluded=False, types=(0, 1)):
    c6 = 4 * EPS * SIG**6
    c12 = 4 * EPS * SIG**12
    o = O.Oracle("f64", "rf", rlist=RC, rcoulomb=RC, lj_pme=lj_pme, ewald_rtol_lj=RTOL)
    box = np.array([5.0, 5.0, 5.0])
    x = np.array([[1.0, 2.5, 2.5], [1.0 + r, 2.5, 2.5]])
    # geometric C6/C12 matrices (the kernel's pair parameters)
    m6 = np.sqrt(np.outer(c6, c6))
    m12 = np.sqrt(np.outer(c12, c12))
    nbfp = np.stack([6 * m6, 12 * m12], axis=-1).reshape(-1)
    if excluded:
        off, idx = np.array([0, 2, 4], np.int32), np.array([0, 1, 0, 1], np.int32)
    else:
        off, idx = np.array([0, 1, 2], np.int32), np.array([0, 1], np.int32)
    o.setup(box, x, np.array(types, np.int32), np.zeros(2), nbfp, 2, off, idx)
    res = o.compute(x)
    return res["Vvdw"], -res["f"][0, 0], o.lj_pme_beta(), c6


def _c6g(rule, i=0, j=1):
    if rule == "geometric":
        c6 = 4 * EPS * SIG**6
        return np.sqrt(c6[i] * c6[j])
    return 4 * np.sqrt(EPS[i] * EPS[j]) * (0.5 * (SIG[i] + SIG[j])) ** 6


def _grid(r, beta, rc=RC):
    return (1 - _g(beta * r)) / r**6 - (1 - _g(beta * rc)) / rc**6


def _dgrid(r, beta, h=1e-6):
    return (_grid(r + h, beta) - _grid(r - h, beta)) / (2 * h)


@pytest.mark.parametrize("rule", ["geometric", "lorentz-berthelot"])
@pytest.mark.parametrize("r", [0.32, 0.55, 0.9])
def test_two_atom_closed_form(rule, r):
    v0, f0, _, _ = _two_atoms(r, None)
    v, f, beta, c6 = _two_atoms(r, rule)
    # self energies of the two atoms (half the r -> 0 limit, pinned below)
    vself = 0.5 * (c6[0] + c6[1]) * beta**6 / 6
    assert v - v0 - vself == pytest.approx(_c6g(rule) * _grid(r, beta), rel=1e-9)
    # force along +x on atom 0 is minus dV/dr of the pair energy
    assert f - f0 == pytest.approx(-_c6g(rule) * _dgrid(r, beta), rel=1e-6)


def test_excluded_pair_keeps_the_grid_correction():
    r = 0.25
    v, f, beta, c6 = _two_atoms(r, "geometric", excluded=True)
    vself = 0.5 * (c6[0] + c6[1]) * beta**6 / 6
    corr = (1 - _g(beta * r)) / r**6
    assert v - vself == pytest.approx(_c6g("geometric") * corr, rel=1e-9)
    h = 1e-6
    dcorr = ((1 - _g(beta * (r + h))) / (r + h) ** 6 - (1 - _g(beta * (r - h))) / (r - h) ** 6) / (2 * h)
    assert f == pytest.approx(-_c6g("geometric") * dcorr, rel=1e-6)


def test_self_energy_is_half_the_zero_distance_limit():
    # two atoms of type 0 beyond the cut-off: only the self terms remain
    v, _, beta, c6 = _two_atoms(1.5, "geometric", types=(0, 0))
    # lim_{r->0} (1 - g(beta r)) / r^6 by a polynomial fit in r^2
    rr = np.linspace(0.02, 0.06, 9)
    h = (1 - _g(beta * rr)) / rr**6
    lim = np.polyval(np.polyfit(rr**2, h, 4), 0.0)
    assert v == pytest.approx(2 * 0.5 * c6[0] * lim, rel=1e-6)


@pytest.mark.parametrize("rule", ["geometric", "lorentz-berthelot"])
def test_forces_are_minus_the_energy_gradient(rule):
    # the LJ-PME energy minus the plain cut-off's, against the same
    # difference of forces (the grid part is smooth; the close contacts'
    # r^-12 forces cancel in the difference)
    s = S.hetero_chains(600, 1.9, ntypes=5, seed=4, rc=0.9, rlist=1.0)
    s = dataclasses.replace(s, charges=np.zeros_like(s.charges))
    o = oracle_for(s, "f64", lj_pme=rule)
    base = oracle_for(s, "f64")

    def de(x):
        return o.compute(x)["Vvdw"] - base.compute(x)["Vvdw"]

    df = o.compute(s.x)["f"] - base.compute(s.x)["f"]
    dmax = np.abs(df).max()
    assert dmax > 0.1
    rng = np.random.default_rng(6)
    h = 1e-5
    for a in rng.choice(len(s.x), 10, replace=False):
        for d in range(3):
            xp = s.x.copy()
            xm = s.x.copy()
            xp[a, d] += h
            xm[a, d] -= h
            g = (de(xp) - de(xm)) / (2 * h)
            assert -g == pytest.approx(df[a, d], abs=2e-3 * dmax), (a, d)


def test_lj_pme_set_up_errors():
    with pytest.raises(ValueError):
        O.Oracle("f64", "rf", vdw_modifier="force-switch", rvdw_switch=0.5, lj_pme="geometric")
    with pytest.raises(ValueError):
        O.Oracle("f64", "rf", lj_pme="geometric", ewald_rtol_lj=0.0)
