"""Small grids with closed-form fields for the parity and physics tests."""
from __future__ import annotations

import numpy as np


def make_grid(nx, ny, nz, h, rf=None, ac=None, dc1=None, mask=None, axis=None, z_inject=None,
              z_return=-1.0, z_exit=1e9, z_return_legacy=-1.0, z_exit_legacy=1e9, crash=(-1e9, 1e9)):
    """Grid dict with fields given as callables f(X, Y, Z) -> (Ex, Ey, Ez) per volt
    (X, Y, Z absolute node coordinates in m) or None (zero)."""
    z = np.arange(nz) * h
    y = np.arange(ny) * h
    x = np.arange(nx) * h
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    N = nx * ny * nz

    def build(fn):
        out = np.zeros((nz, ny, nx, 3))
        if fn is not None:
            ex, ey, ez = fn(X, Y, Z)
            out[..., 0], out[..., 1], out[..., 2] = ex, ey, ez
        return out.reshape(3 * N)

    ax, ay = axis if axis is not None else ((nx // 2) * h, (ny // 2) * h)
    m = np.zeros(N, np.uint8) if mask is None else np.asarray(mask(X, Y, Z), np.uint8).reshape(N)
    return {
        "nx": nx, "ny": ny, "nz": nz, "h": h, "axis_x": ax, "axis_y": ay,
        "z_inject": (nz // 2) * h if z_inject is None else z_inject,
        "z_return": z_return, "z_exit": z_exit, "z_return_legacy": z_return_legacy, "z_exit_legacy": z_exit_legacy,
        "crash_z0_legacy": crash[0], "crash_z1_legacy": crash[1],
        "rf": build(rf), "ac": build(ac), "dc": [None, build(dc1), None, None], "mask": m,
    }


def binom_pull(k1, n1, k2, n2):
    """Pull of two binomial fractions: (p1 - p2) / sqrt(se1^2 + se2^2) (validate.rs:114-117)."""
    p1, p2 = k1 / n1, k2 / n2
    se = np.sqrt(max(p1 * (1 - p1) / n1 + p2 * (1 - p2) / n2, 1e-30))
    # a zero-variance pair (both 0 or both 1) is in agreement only when equal
    if p1 * (1 - p1) == 0 and p2 * (1 - p2) == 0:
        return 0.0 if k1 * n2 == k2 * n1 else np.inf
    return (p1 - p2) / se


def mathieu_beta(a: float, q: float, terms: int = 30) -> float:
    """Characteristic exponent beta of the Mathieu equation from the standard
    continued fraction (March, Quadrupole Ion Trap Mass Spectrometry, 2005),
    solved by fixed-point iteration."""
    beta = np.sqrt(max(a + q * q / 2.0, 1e-12))
    for _ in range(200):
        def cf(b, sgn):
            v = 0.0
            for k in range(terms, 0, -1):
                m = b + sgn * 2 * k
                v = q * q / (m * m - a - v)
            return v
        b2 = a + cf(beta, +1) + cf(beta, -1)
        nb = np.sqrt(max(b2, 1e-12))
        if abs(nb - beta) < 1e-13:
            break
        beta = nb
    return beta
