"""Pins the analytical Ewald correction used by the HIP kernel
(NBX_EWALD_ANALYTICAL): the rational fits P(s)/Q(s), s = (beta r)^2, whose
coefficients are in gmxapi_amd/csrc/kernel_common.h, evaluated in float32 the way
the kernel does (Horner with fused multiply-adds, then a reciprocal), against
the closed forms the reference's SIMD kernels approximate
(src/gromacs/simd/tests/simd_math.cpp:938-977):
    force      (erf z - 2 z exp(-z^2)/sqrt(pi)) / z^3, fitted in s = z^2
    energy     erfc(z) exp(z^2), fitted in z (the kernel multiplies by
               exp(-z^2)/r: erfc(beta r)/r without cancellation)
over the fitted range z <= 3.6 (beta rc = 3.123 at ewald_rtol 1e-5)."""
import math
import os
import re

import numpy as np

SRC = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "gmxapi_amd", "csrc", "kernel_common.h")


def _coefficients():
    text = open(SRC).read()
    vals = {m.group(1): float(m.group(2)) for m in re.finditer(r"\b(k[GE][pq]\d)\s*=\s*([-+0-9.eE]+)f", text)}
    gp = [vals[f"kGp{i}"] for i in range(6)]
    gq = [1.0] + [vals[f"kGq{i}"] for i in range(1, 6)]
    ep = [vals[f"kEp{i}"] for i in range(5)]
    eq = [1.0] + [vals[f"kEq{i}"] for i in range(1, 5)]
    return gp, gq, ep, eq


def _horner_f32(c, s):
    # fmaf(acc, s, c_k) with one rounding per step, as the kernel's fmaf chain
    acc = np.float32(c[-1])
    for ck in c[-2::-1]:
        acc = np.float32(np.float64(acc) * np.float64(s) + np.float64(np.float32(ck)))
    return acc


def _g(z):
    if z == 0:
        return 4 / (3 * math.sqrt(math.pi))
    if z < 1e-3:
        return 4 / (3 * math.sqrt(math.pi)) * (1 - 3 * z * z / 5)
    return (math.erf(z) - 2 * z / math.sqrt(math.pi) * math.exp(-z * z)) / z ** 3


def _e(z):
    # erfc(z) exp(z^2) without overflow or cancellation for the tested range
    return math.erfc(z) * math.exp(z * z)


def test_rational_fits_match_closed_forms():
    gp, gq, ep, eq = _coefficients()
    worst_g = worst_e = 0.0
    for z in np.linspace(0.0, 3.6, 4001):
        s = np.float32(z * z)
        g = np.float64(_horner_f32(gp, s)) / np.float64(_horner_f32(gq, s))
        e = np.float64(_horner_f32(ep, np.float32(z))) / np.float64(_horner_f32(eq, np.float32(z)))
        worst_g = max(worst_g, abs(g / _g(z) - 1))
        worst_e = max(worst_e, abs(e / _e(float(np.float32(z))) - 1))
    # well inside the 5e-4 force bar; a few ulp in single precision
    assert worst_g < 5e-7, worst_g
    assert worst_e < 5e-7, worst_e


def test_fit_values_at_zero():
    # g(0) = 4/(3 sqrt(pi)) (the force correction's r -> 0 limit) and
    # erfc(0) exp(0) = 1
    gp, gq, ep, eq = _coefficients()
    assert abs(gp[0] - 4 / (3 * math.sqrt(math.pi))) < 2e-7
    assert ep[0] == 1.0
