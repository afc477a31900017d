"""Generates the rational approximations of the analytical Ewald
correction used by the pair kernels (coefficients in kernel_common.h; NBX_EWALD_ANALYTICAL).

With z = beta r and s = z^2, the kernel needs
    g(z) = (erf z - 2 z / sqrt(pi) exp(-z^2)) / z^3   (force: F_corr / r = beta^3 g)
    E(z) = erfc(z) exp(z^2)                           (energy: erfc(beta r) / r
                                                       = exp(-s) E(z) / r)
for z up to 3.6, which covers beta rc for ewald_rtol down to about 3e-7 (the
library rejects the analytical flavour beyond it).  The energy uses erfc
directly rather than 1/r - erf(beta r)/r, which would cancel.  g is even
and fitted as P(s)/Q(s); E is fitted as P(z)/Q(z); both with Q(0) = 1, by
iteratively reweighted linear least squares on the relative error (Loeb's
linearisation, then Lawson-style reweighting toward equal ripple), against
40-digit values from mpmath.  Run: python3 fit_ewald_correction.py
"""
import mpmath as mp
import numpy as np

mp.mp.dps = 40
S_MAX = 3.6 ** 2


def g(s):
    z = mp.sqrt(mp.mpf(s))
    if z == 0:
        return 4 / (3 * mp.sqrt(mp.pi))
    return (mp.erf(z) - 2 * z / mp.sqrt(mp.pi) * mp.exp(-z * z)) / z ** 3


def e_of_s(s):
    # E as a function of the fit variable; fitted in z = sqrt(s) below
    z = mp.mpf(s)
    return mp.erfc(z) * mp.exp(z * z)


def fit(fun, m, n, iters=40, nodes=3000, x_max=S_MAX):
    s = (x_max / 2) * (1 - np.cos(np.pi * (np.arange(nodes) + 0.5) / nodes))
    f = np.array([float(fun(v)) for v in s])
    w = np.ones(nodes)
    for it in range(iters):
        A = np.hstack([np.vander(s, m + 1, increasing=True),
                       -f[:, None] * np.vander(s, n + 1, increasing=True)[:, 1:]])
        A = A / f[:, None] * w[:, None]
        sc = np.abs(A).max(axis=0)
        x, *_ = np.linalg.lstsq(A / sc, w, rcond=None)
        x /= sc
        p, q = x[:m + 1], np.concatenate([[1.0], x[m + 1:]])
        qs = np.polyval(q[::-1], s)
        err = (np.polyval(p[::-1], s) / qs - f) / f
        w = (w * np.abs(err) ** 0.5) if it > 2 else 1 / np.abs(qs)
        w /= w.max()
    return p, q, np.abs(err).max()


if __name__ == "__main__":
    for name, fun, (m, n), x_max in (("g(s)", g, (5, 5), S_MAX), ("E(z)", e_of_s, (4, 4), 3.6)):
        p, q, e = fit(fun, m, n, x_max=x_max)
        print(f"{name}: P degree {m}, Q degree {n}, max relative error {e:.2e} (exact arithmetic)")
        print("  P:", ", ".join(f"{float(np.float32(v))!r}f" for v in p))
        print("  Q:", ", ".join(f"{float(np.float32(v))!r}f" for v in q[1:]))
