"""CPU checks that pin the FP64 reference gradient (oracle/adjoint.py) used
by the GPU adjoint tests: its forces are the FP64 oracle's (exact Ewald), and
its gradients of a force-matching loss match central finite differences of
the FP64 oracle's forces with respect to charges, per-type C6/C12 and
positions."""
import dataclasses

import numpy as np
import pytest

from gmxapi_amd import systems as S
from oracle import adjoint as A
from tests.helpers import oracle_for, rms

H = 1e-4  # relative finite-difference step


@pytest.fixture(scope="module")
def case():
    s = S.hetero_chains(1200, 2.3, ntypes=5, seed=3, rc=0.9, rlist=1.0)
    rng = np.random.default_rng(5)
    # the target: the same system with LJ parameters and charges perturbed by 5 %
    t = dataclasses.replace(s, c6=s.c6 * (1 + 0.05 * rng.normal(size=s.ntypes)),
                            c12=s.c12 * (1 + 0.05 * rng.normal(size=s.ntypes)),
                            charges=s.charges * (1 + 0.05 * rng.normal(size=s.natoms)))
    f_t = oracle_for(t, "f64", ewald_exact=True).compute(t.x)["f"]
    f = oracle_for(s, "f64", ewald_exact=True).compute(s.x)["f"]
    g = 2.0 * (f - f_t)
    ref = A.forces_and_gradient(s, g, s.c6_matrix(), s.c12_matrix())
    return s, f, f_t, ref


def _loss(s, f_t, x=None, o=None):
    o = o or oracle_for(s, "f64", ewald_exact=True)
    f = o.compute(s.x if x is None else x)["f"]
    return float(np.sum((f - f_t) ** 2))


def test_reference_forces_are_the_oracle_forces(case):
    s, f, _, ref = case
    assert rms(ref["f"] - f) <= 1e-10 * rms(f)


def test_charge_gradient_matches_finite_differences(case):
    s, _, f_t, ref = case
    for k in (0, 517, 1111):
        h = H * abs(s.charges[k])
        qp, qm = s.charges.copy(), s.charges.copy()
        qp[k] += h
        qm[k] -= h
        fd = (_loss(dataclasses.replace(s, charges=qp), f_t) - _loss(dataclasses.replace(s, charges=qm), f_t)) / (2 * h)
        assert abs(fd - ref["grad_q"][k]) <= 1e-6 * max(abs(fd), 1e-3 * np.abs(ref["grad_q"]).max()), (k, fd)


def test_lj_gradient_matches_finite_differences(case):
    # per-type C6_t / C12_t through nblib's geometric rule C_ab = sqrt(C_a C_b)
    s, _, f_t, ref = case
    for key, grad in (("c6", ref["grad_c6"]), ("c12", ref["grad_c12"])):
        base = getattr(s, key)
        for tt in range(s.ntypes):
            cp, cm = base.copy(), base.copy()
            h = H * base[tt]
            cp[tt] += h
            cm[tt] -= h
            fd = (_loss(dataclasses.replace(s, **{key: cp}), f_t) -
                  _loss(dataclasses.replace(s, **{key: cm}), f_t)) / (2 * h)
            # dC_ab/dC_t = (delta_at C_b + delta_bt C_a) / (2 sqrt(C_a C_b))
            dm = np.zeros_like(grad)
            for a in range(s.ntypes):
                for b in range(s.ntypes):
                    num = (base[b] if a == tt else 0.0) + (base[a] if b == tt else 0.0)
                    dm[a, b] = num / (2.0 * np.sqrt(base[a] * base[b]))
            an = float(np.sum(grad * dm))
            assert abs(fd - an) <= 1e-6 * abs(fd), (key, tt, fd, an)


def test_position_gradient_matches_finite_differences(case):
    # the Hessian-vector product: coordinates whose step moves no pair across
    # the cut-off (the forces jump there).  The step is 1e-6 nm, not relative
    # to the coordinate: bonded neighbours sit at 0.15 nm, where r^-14 terms
    # make a 1e-4 nm step's truncation error about 0.5 %.
    s, _, f_t, ref = case
    o = oracle_for(s, "f64", ewald_exact=True)
    box = np.asarray(s.box)
    checked = 0
    for k in range(0, s.natoms, 97):
        d = s.x - s.x[k]
        d -= box * np.round(d / box)
        r = np.sqrt(np.sum(d * d, axis=1))
        h = 1e-6
        if np.any(np.abs(r - s.rcoulomb) < 4 * h):
            continue
        for dim in range(3):
            xp, xm = s.x.copy(), s.x.copy()
            xp[k, dim] += h
            xm[k, dim] -= h
            fd = (_loss(s, f_t, xp, o) - _loss(s, f_t, xm, o)) / (2 * h)
            an = ref["grad_x"][k, dim]
            assert abs(fd - an) <= 1e-5 * max(abs(fd), 1e-3 * np.abs(ref["grad_x"]).max()), (k, dim, fd, an)
        checked += 1
        if checked == 4:
            break
    assert checked == 4
