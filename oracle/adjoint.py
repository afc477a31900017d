"""FP64 reference gradient of the pair forces (test infrastructure only).

The reference's nonbonded kernels are forward only, so this restates no
reference function: it is the hand derivative of the forward pair sum that
the FP64 oracle computes (the plain-C kernel kernel_ref_inner.h:44-375 with
the exact Ewald closed form), written with numpy over all atom pairs within
the cut-off (O(N^2) enumeration, minimum image).  For a loss L(F) with
g_i = dL/dF_i, each pair adds f_ij = s(r^2) d_ij to atom i (and -f_ij to j),
d_ij = x_i - x_j (minimum image), and with u = (g_i - g_j) . d_ij:

    dL/dx_i += s (g_i - g_j) + 2 s' u d_ij      (dL/dx_j gets the negative)
    dL/dq_i += u epsfac q_j h,  dL/dq_j += u epsfac q_i h
    dL/dC6[t_i][t_j]  += -6 u r^-8 (non-excluded pairs)
    dL/dC12[t_i][t_j] += 12 u r^-14 (non-excluded pairs)

s = 12 C12 r^-14 - 6 C6 r^-8 + epsfac q_i q_j h(r) (LJ for non-excluded
pairs only), s' = ds/d(r^2), h = r^-3 - corr(r) for non-excluded pairs and
-corr(r) for excluded ones, corr(r) = (erf(beta r) - 2 beta r exp(-beta^2 r^2)
/ sqrt(pi)) / r^3 for Ewald (the real-space correction the kernel applies to
every pair inside the cut-off) or 2 k_rf for reaction field.

Pinned by central finite differences of the FP64 oracle's forces
(tests/test_adjoint_reference.py).  Only tests/ import it.
"""
from __future__ import annotations

import math

import numpy as np

from . import oracle as O


def _erf(z: np.ndarray) -> np.ndarray:
    from scipy.special import erf

    return erf(z)


def pair_table(x, box, rc: float):
    """(i, j, d) for all pairs i < j with minimum-image distance below rc, in
    float64 (d = x_i - x_j)."""
    x = np.asarray(x, np.float64)
    box = np.asarray(box, np.float64)
    keys = O.brute_pairs(x, box, rc, "f64")
    i = (keys >> 32).astype(np.int64)
    j = (keys & 0xFFFFFFFF).astype(np.int64)
    d = x[i] - x[j]
    d -= box * np.round(d / box)
    return i, j, d


def excluded_mask(i, j, excl_off, excl_idx) -> np.ndarray:
    """True where the pair (i, j) is in the exclusion lists."""
    excl_off = np.asarray(excl_off, np.int64)
    excl_idx = np.asarray(excl_idx, np.int64)
    n = len(excl_off) - 1
    a = np.repeat(np.arange(n, dtype=np.int64), np.diff(excl_off))
    b = excl_idx
    keep = a != b
    ekeys = np.unique(np.minimum(a, b)[keep] * n + np.maximum(a, b)[keep])
    i = np.asarray(i, np.int64)
    j = np.asarray(j, np.int64)
    return np.isin(np.minimum(i, j) * n + np.maximum(i, j), ekeys)


def forces_and_gradient(system, g, c6m, c12m, charges=None, coulomb="ewald", beta=None, k_rf=0.0, pairs=None):
    """FP64 pair forces F and, for the cotangent g (N, 3), the gradients
    dict(f, grad_x (N, 3), grad_q (N,), grad_c6 (T, T), grad_c12 (T, T)).

    c6m, c12m are the unscaled C6/C12 type-pair matrices; charges default to
    system.charges; beta defaults to the oracle's calc_ewaldcoeff_q value for
    rtol 1e-5 at system.rcoulomb."""
    rc = system.rcoulomb
    q = np.asarray(system.charges if charges is None else charges, np.float64)
    t = np.asarray(system.types, np.int64)
    g = np.asarray(g, np.float64).reshape(-1, 3)
    n = len(q)
    if beta is None:
        beta = O.ewald_coeff(rc, 1e-5, "f64")
    epsfac = O.one_4pi_eps0()
    i, j, d = pairs if pairs is not None else pair_table(system.x, system.box, rc)
    excl = excluded_mask(i, j, system.excl_off, system.excl_idx)
    inter = (~excl).astype(np.float64)
    r2 = np.einsum("ij,ij->i", d, d)
    r = np.sqrt(r2)
    rinv2 = 1.0 / r2
    rinv = 1.0 / r
    rinv3 = rinv * rinv2
    c6 = np.asarray(c6m, np.float64)[t[i], t[j]]
    c12 = np.asarray(c12m, np.float64)[t[i], t[j]]
    r8 = inter * rinv2 ** 4
    r14 = inter * rinv2 ** 7
    s_lj = 12.0 * c12 * r14 - 6.0 * c6 * r8
    ds_lj = -84.0 * c12 * r14 * rinv2 + 24.0 * c6 * r8 * rinv2
    if coulomb == "ewald":
        br = beta * r
        ex = np.exp(-br * br)
        tp = 2.0 * beta / math.sqrt(math.pi)
        erf_br = _erf(br)
        corr = erf_br * rinv3 - tp * ex * rinv2
        dcorr_dr = tp * ex * (3.0 * rinv3 + 2.0 * beta * beta * rinv) - 3.0 * erf_br * rinv2 * rinv2
        dcorr = dcorr_dr * 0.5 * rinv
    else:
        corr = np.full_like(r, 2.0 * k_rf)
        dcorr = np.zeros_like(r)
    h = inter * rinv3 - corr
    dh = -1.5 * inter * rinv3 * rinv2 - dcorr
    qq = epsfac * q[i] * q[j]
    s = s_lj + qq * h
    sp = ds_lj + qq * dh

    fij = s[:, None] * d
    f = np.zeros((n, 3))
    np.add.at(f, i, fij)
    np.add.at(f, j, -fij)

    gd = g[i] - g[j]
    u = np.einsum("ij,ij->i", gd, d)
    w = s[:, None] * gd + (2.0 * sp * u)[:, None] * d
    gx = np.zeros((n, 3))
    np.add.at(gx, i, w)
    np.add.at(gx, j, -w)
    gq = np.zeros(n)
    np.add.at(gq, i, u * epsfac * q[j] * h)
    np.add.at(gq, j, u * epsfac * q[i] * h)
    nt = np.asarray(c6m).shape[0]
    g6 = np.zeros((nt, nt))
    g12 = np.zeros((nt, nt))
    np.add.at(g6, (t[i], t[j]), -6.0 * u * r8)
    np.add.at(g12, (t[i], t[j]), 12.0 * u * r14)
    return dict(f=f, grad_x=gx, grad_q=gq, grad_c6=g6, grad_c12=g12)


def symmetrized(m: np.ndarray) -> np.ndarray:
    """Type-pair gradient summed over both orientations (i, j) and (j, i):
    what a symmetric parametrization sees, independent of which atom of a
    pair a kernel treats as i."""
    return m + m.T - np.diag(np.diag(m))
