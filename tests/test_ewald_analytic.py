"""Pins the analytical Ewald correction used by the HIP kernel
(NBX_EWALD_ANALYTICAL): the rational fits P(s)/Q(s), s = (beta r)^2, whose
coefficients are in gmxapi_amd/csrc/kernel.hip, evaluated in float32 the way
the kernel does (Horner with fused multiply-adds, then a reciprocal), against
the closed forms the reference's SIMD kernels approximate
(src/gromacs/simd/tests/simd_math.cpp:938-977):
    force      (erf z - 2 z exp(-z^2)/sqrt(pi)) / z^3
    potential  erf(z) / z
over the fitted range z <= 3.6 (beta rc = 3.123 at ewald_rtol 1e-5)."""
import math
import os
import re

import numpy as np

SRC = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "gmxapi_amd", "csrc", "kernel.hip")


def _coefficients():
    text = open(SRC).read()
    vals = {m.group(1): float(m.group(2)) for m in re.finditer(r"\b(k[GH][pq]\d)\s*=\s*([-+0-9.eE]+)f", text)}
    gp = [vals[f"kGp{i}"] for i in range(6)]
    gq = [1.0] + [vals[f"kGq{i}"] for i in range(1, 6)]
    hp = [vals[f"kHp{i}"] for i in range(6)]
    hq = [1.0] + [vals[f"kHq{i}"] for i in range(1, 5)]
    return gp, gq, hp, hq


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


def _h(z):
    return 2 / math.sqrt(math.pi) if z == 0 else math.erf(z) / z


def test_rational_fits_match_closed_forms():
    gp, gq, hp, hq = _coefficients()
    worst_g = worst_h = 0.0
    for z in np.linspace(0.0, 3.6, 4001):
        s = np.float32(z * z)
        g = np.float64(_horner_f32(gp, s)) / np.float64(_horner_f32(gq, s))
        h = np.float64(_horner_f32(hp, s)) / np.float64(_horner_f32(hq, s))
        worst_g = max(worst_g, abs(g / _g(z) - 1))
        worst_h = max(worst_h, abs(h / _h(z) - 1))
    # well inside the 5e-4 force bar; about 4 ulp in single precision
    assert worst_g < 5e-7, worst_g
    assert worst_h < 5e-7, worst_h


def test_fit_values_at_zero():
    # g(0) = 4/(3 sqrt(pi)), h(0) = 2/sqrt(pi): the self-energy limit
    gp, gq, hp, hq = _coefficients()
    assert abs(gp[0] - 4 / (3 * math.sqrt(math.pi))) < 2e-7
    assert abs(hp[0] - 2 / math.sqrt(math.pi)) < 2e-7
