"""Pins the oracle's twin cut-off (rvdw < rcoulomb: the TAB_TWIN / EWALD_TWIN
kernels, kerneldispatch.cpp:175-192, with VDW_CUTOFF_CHECK,
kernel_ref_inner.h:254-271, and the LJ potential shift at rvdw,
forcerec.cpp:848-873).

No reference fixture covers it, so it is held to what the mathematics fixes:
  * the twin calculator's forces and energies are the sum of two independent
    FP64 direct pair sums: LJ alone with cut-off and shift at rvdw (charges
    off) and Ewald Coulomb alone at rcoulomb (LJ off);
  * two atoms between rvdw and rcoulomb have no LJ energy or force, and
    closer than rvdw the LJ energy is the potential shifted at rvdw;
  * the switches act at rvdw: zero potential and force there;
  * grompp's restriction (Ewald with an LJ cut-off only) is enforced.
A mask at rcoulomb instead of rvdw, or the shift taken at rcoulomb, planted
in the oracle, fails the first two checks.
"""
import numpy as np
import pytest

from gmxapi_amd import systems as S
from oracle import oracle as O
from tests.helpers import nbfp_f64, oracle_for

RVDW = 0.7


def _brute(s, rc, q, nbfp, beta):
    return O.brute_forces(s.x, s.box, s.types, q, nbfp, s.ntypes, s.excl_off, s.excl_idx, "ewald", rc, beta,
                          O.one_4pi_eps0())


@pytest.mark.parametrize("name", ["water", "chains"])
def test_twin_is_lj_at_rvdw_plus_coulomb_at_rc(name):
    if name == "water":
        s = S.water_box(300, 2.2, seed=5, rc=0.95, rlist=1.0)
    else:
        s = S.hetero_chains(900, 2.1, ntypes=5, seed=6, rc=0.95, rlist=1.0)
    o = oracle_for(s, "f64", ewald_exact=True, rvdw=RVDW)
    r = o.compute(s.x)
    beta = o.constants()["beta"]
    nbfp = nbfp_f64(s)
    f_lj, (v_lj, _), _ = _brute(s, RVDW, np.zeros(s.natoms), nbfp, beta)
    f_c, (_, v_c), _ = _brute(s, s.rcoulomb, s.charges, np.zeros_like(nbfp), beta)
    fref = f_lj + f_c
    scale = np.abs(fref).max()
    np.testing.assert_allclose(r["f"], fref, rtol=0, atol=1e-9 * scale)
    assert r["Vvdw"] == pytest.approx(v_lj, rel=1e-10)
    assert r["Vc"] == pytest.approx(v_c, rel=1e-10)
    # and it differs from the single cut-off
    single = oracle_for(s, "f64", ewald_exact=True).compute(s.x)
    assert abs(single["Vvdw"] - r["Vvdw"]) > 1e-5 * abs(r["Vvdw"])
    assert single["Vc"] == pytest.approx(r["Vc"], rel=1e-12)


def _pair(r, modifier="potential-shift", rsw=0.0, c6=2.6e-3, c12=2.6e-6, q=(0.0, 0.0)):
    o = O.Oracle("f64", "ewald", ewald_exact=True, rlist=1.0, rcoulomb=1.0, rvdw=RVDW, vdw_modifier=modifier,
                 rvdw_switch=rsw)
    box = np.array([5.0, 5.0, 5.0])
    x = np.array([[1.0, 2.5, 2.5], [1.0 + r, 2.5, 2.5]])
    o.setup(box, x, np.zeros(2, np.int32), np.array(q), O.nblib_nbfp([c6], [c12], "f64"), 1,
            np.array([0, 1, 2], np.int32), np.array([0, 1], np.int32))
    res = o.compute(x)
    return res["Vvdw"], -res["f"][0, 0]


def _plain(r, c6=2.6e-3, c12=2.6e-6):
    return c12 / r**12 - c6 / r**6, 12 * c12 / r**13 - 6 * c6 / r**7


@pytest.mark.parametrize("r", [0.71, 0.85, 0.99])
def test_no_lj_between_rvdw_and_rcoulomb(r):
    v, f = _pair(r)
    assert v == 0.0 and f == 0.0


@pytest.mark.parametrize("r", [0.35, 0.5, 0.69])
def test_lj_shifted_at_rvdw(r):
    v, f = _pair(r)
    vp, fp = _plain(r)
    assert f == pytest.approx(fp, rel=1e-12)
    assert v == pytest.approx(vp - _plain(RVDW)[0], rel=1e-12, abs=1e-15)


@pytest.mark.parametrize("modifier", ["force-switch", "potential-switch"])
def test_switches_end_at_rvdw(modifier):
    v, f = _pair(RVDW - 1e-7, modifier, rsw=0.5)
    v0, f0 = _plain(0.6)
    assert abs(v) <= 1e-9 * abs(v0)
    assert abs(f) <= 1e-6 * abs(f0)
    # below the switch radius the force is the plain one
    assert _pair(0.45, modifier, rsw=0.5)[1] == pytest.approx(_plain(0.45)[1], rel=1e-12)


def test_twin_restrictions():
    with pytest.raises(ValueError):
        O.Oracle("f64", "rf", rlist=1.0, rcoulomb=1.0, rvdw=0.8)
    with pytest.raises(ValueError):
        O.Oracle("f64", "ewald", rlist=1.0, rcoulomb=1.0, rvdw=0.8, lj_pme="geometric")
    with pytest.raises(ValueError):
        O.Oracle("f64", "ewald", rlist=1.0, rcoulomb=1.0, rvdw=1.1)
    with pytest.raises(ValueError):
        O.Oracle("f64", "ewald", rlist=1.0, rcoulomb=1.0, rvdw=0.8, vdw_modifier="force-switch", rvdw_switch=0.8)
    # rvdw = rcoulomb is the single cut-off
    s = S.small_water_box(216, seed=11)
    a = oracle_for(s, "f64", rvdw=s.rcoulomb).compute(s.x)
    b = oracle_for(s, "f64").compute(s.x)
    np.testing.assert_array_equal(a["f"], b["f"])
    assert a["Vvdw"] == b["Vvdw"]


def test_charges_do_not_change_the_lj_part():
    v0, _f0 = _pair(0.5)
    v1, _f1 = _pair(0.5, q=(0.4, -0.4))
    assert v1 == pytest.approx(v0, rel=1e-13)
