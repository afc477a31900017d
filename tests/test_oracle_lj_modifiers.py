"""Pins the oracle's LJ force-switch and potential-switch flavours
(kernel_ref_inner.h:155-205 with the constants of forcerec.cpp:780-818).

No reference fixture covers them, so each is held to what the mathematics
fixes, independently of the constants the oracle computes:
  * forces are minus the gradient of the energy (central differences over a
    heterogeneous system, charges off);
  * potential and force reach zero at the cut-off;
  * below the switch radius the force is the plain LJ force, the force
    switch's potential is the plain one plus a constant, and the potential
    switch's is the unshifted C12/r^12 - C6/r^6;
  * at the middle of the switching range the potential switch's smoothstep
    is exactly 1/2.
"""
import dataclasses

import numpy as np
import pytest

from gmxapi_amd import systems as S
from oracle import oracle as O
from tests.helpers import oracle_for

RSW = 0.6


def _uncharged(s):
    return dataclasses.replace(s, charges=np.zeros_like(s.charges))


@pytest.mark.parametrize("modifier", ["force-switch", "potential-switch"])
def test_forces_are_minus_the_energy_gradient(modifier):
    # central differences of the modifier's energy minus the potential
    # shift's, against the same difference of forces: below the switch
    # radius both forces are the plain LJ force, so the difference isolates
    # the switching range from the close contacts' r^-12 forces (~1e6 here)
    s = _uncharged(S.hetero_chains(600, 1.9, ntypes=5, seed=4, rc=0.9, rlist=1.0))
    o = oracle_for(s, "f64", vdw=(modifier, RSW))
    base = oracle_for(s, "f64")

    def de(x):
        return o.compute(x)["Vvdw"] - base.compute(x)["Vvdw"]

    df = o.compute(s.x)["f"] - base.compute(s.x)["f"]
    dmax = np.abs(df).max()
    assert dmax > 0.1
    rng = np.random.default_rng(5)
    h = 1e-5
    for a in rng.choice(len(s.x), 12, replace=False):
        for d in range(3):
            xp = s.x.copy()
            xm = s.x.copy()
            xp[a, d] += h
            xm[a, d] -= h
            g = (de(xp) - de(xm)) / (2 * h)
            assert -g == pytest.approx(df[a, d], abs=2e-3 * dmax), (a, d)


def _pair(modifier, r, c6=2.6e-3, c12=2.6e-6, rc=1.0):
    """Energy and the force on atom 0 (along x) of two uncharged atoms at
    distance r."""
    o = O.Oracle("f64", "rf", rlist=rc, rcoulomb=rc, vdw_modifier=modifier, rvdw_switch=RSW)
    box = np.array([5.0, 5.0, 5.0])
    x = np.array([[1.0, 2.5, 2.5], [1.0 + r, 2.5, 2.5]])
    nbfp = O.nblib_nbfp([c6], [c12], "f64")
    o.setup(box, x, np.zeros(2, np.int32), np.zeros(2), nbfp, 1, np.array([0, 1, 2], np.int32),
            np.array([0, 1], np.int32))
    res = o.compute(x)
    return res["Vvdw"], -res["f"][0, 0]


def _plain(r, c6=2.6e-3, c12=2.6e-6):
    v = c12 / r**12 - c6 / r**6
    f = 12 * c12 / r**13 - 6 * c6 / r**7
    return v, f


@pytest.mark.parametrize("modifier", ["force-switch", "potential-switch"])
def test_zero_at_the_cut_off(modifier):
    v, f = _pair(modifier, 1.0 - 1e-7)
    v0, f0 = _plain(0.8)
    assert abs(v) <= 1e-9 * abs(v0)
    assert abs(f) <= 1e-6 * abs(f0)


@pytest.mark.parametrize("r", [0.35, 0.45, 0.59])
def test_below_the_switch_radius(r):
    vp, fp = _plain(r)
    vs, fs = _pair("potential-shift", r)
    vf, ff = _pair("force-switch", r)
    vw, fw = _pair("potential-switch", r)
    # forces: the plain LJ force (the oracle's sign: repulsive positive
    # along the separation)
    for got in (fs, ff, fw):
        assert got == pytest.approx(fp, rel=1e-12)
    # potential switch: unshifted; potential shift: shifted by -V(rc)
    assert vw == pytest.approx(vp, rel=1e-12, abs=1e-15)
    assert vs == pytest.approx(vp - _plain(1.0)[0], rel=1e-12, abs=1e-15)
    # force switch: the plain potential plus one constant for every r
    vf2, _ = _pair("force-switch", 0.5)
    assert vf - vp == pytest.approx(vf2 - _plain(0.5)[0], rel=1e-10, abs=1e-15)


def test_potential_switch_is_one_half_mid_range():
    r = 0.5 * (RSW + 1.0)
    vw, _ = _pair("potential-switch", r)
    assert vw == pytest.approx(0.5 * _plain(r)[0], rel=1e-12)


def test_bad_modifier_is_refused():
    with pytest.raises(ValueError):
        O.Oracle("f64", "rf", vdw_modifier="force-switch", rvdw_switch=1.0)
    with pytest.raises(KeyError):
        O.Oracle("f64", "rf", vdw_modifier="exact-cutoff")
