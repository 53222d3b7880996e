"""Collapsed Gauss-Jacobi quadrature on the reference simplex.

The reference asks basix for ``make_quadrature(cell, m, gauss_jacobi)`` with
``m = 2 * max(k) + 1`` over the fields of a product space
(``src/spaces/product_space.py:23-50``). basix's collapsed Gauss-Jacobi scheme
uses ``np = (m + 2) // 2`` points per direction (Karniadakis & Sherwin,
Appendix B): Gauss-Jacobi(alpha, 0) rules with alpha = 0, 1, 2 on [-1, 1]
combined by the Duffy collapse, points ascending per direction, the first
direction outermost.

Pinned by exactness on every monomial of total degree <= m:
``int_T x^a y^b = a! b! / (a + b + 2)!`` and the tetrahedral analogue
(tests/test_oracle_quadrature.py). The point order only affects rounding.
"""

from __future__ import annotations

import functools

import numpy as np
from scipy.special import roots_jacobi


def _gj(alpha, n):
    x, w = roots_jacobi(n, alpha, 0.0)
    order = np.argsort(x)
    return x[order], w[order]


@functools.lru_cache(maxsize=None)
def gauss_jacobi(dim, m):
    """Return (points (nq, dim), weights (nq,)) on the reference simplex of ``dim``."""
    if dim == 0:
        return np.array([[1.0]])[:, :0], np.array([1.0])
    n = (m + 2) // 2
    if dim == 1:
        x, w = _gj(0.0, n)
        return (0.5 + 0.5 * x)[:, None], 0.5 * w
    if dim == 2:
        px, wx = _gj(0.0, n)
        py, wy = _gj(1.0, n)
        pts = np.zeros((n * n, 2))
        wts = np.zeros(n * n)
        for i in range(n):
            for j in range(n):
                c = i * n + j
                pts[c, 0] = 0.25 * (1.0 + px[i]) * (1.0 - py[j])
                pts[c, 1] = 0.5 * (1.0 + py[j])
                wts[c] = wx[i] * wy[j] * 0.125
        return pts, wts
    if dim == 3:
        px, wx = _gj(0.0, n)
        py, wy = _gj(1.0, n)
        pz, wz = _gj(2.0, n)
        pts = np.zeros((n**3, 3))
        wts = np.zeros(n**3)
        for i in range(n):
            for j in range(n):
                for k in range(n):
                    c = i * n * n + j * n + k
                    pts[c, 0] = 0.125 * (1.0 + px[i]) * (1.0 - py[j]) * (1.0 - pz[k])
                    pts[c, 1] = 0.25 * (1.0 + py[j]) * (1.0 - pz[k])
                    pts[c, 2] = 0.5 * (1.0 + pz[k])
                    wts[c] = wx[i] * wy[j] * wz[k] * 0.125 * 0.125
        return pts, wts
    raise ValueError(dim)
