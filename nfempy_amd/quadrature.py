"""Collapsed Gauss-Jacobi quadrature on the reference simplex.

Stands in for ``basix.make_quadrature(cell, m, QuadratureType.gauss_jacobi)`` as
called by ``ProductSpace._define_integration_order_and_quadrature``
(``src/spaces/product_space.py:23-50``): ``(m + 2) // 2`` points per direction,
Gauss-Jacobi(alpha, 0) rules (alpha = 0, 1, 2) combined by the Duffy collapse,
points ascending per direction with the first direction outermost.

The 1-D rules come from the Golub-Welsch eigenvalue problem of the Jacobi
recurrence.
"""

from __future__ import annotations

import functools
import math

import numpy as np


def _golub_welsch(n: int, alpha: float):
    """Nodes/weights of int_{-1}^{1} (1-x)^alpha f(x) dx, exact to degree 2n-1."""
    a, b = float(alpha), 0.0
    diag = np.zeros(n)
    off = np.zeros(max(n - 1, 0))
    for k in range(n):
        s = 2 * k + a + b
        if k == 0:
            diag[k] = (b - a) / (s + 2)
        else:
            diag[k] = (b * b - a * a) / (s * (s + 2))
        if k >= 1:
            kk = k
            s2 = 2 * kk + a + b
            off[k - 1] = math.sqrt(
                4 * kk * (kk + a) * (kk + b) * (kk + a + b) / (s2 * s2 * (s2 + 1) * (s2 - 1))
            )
    T = np.diag(diag) + np.diag(off, 1) + np.diag(off, -1)
    x, V = np.linalg.eigh(T)
    mu0 = 2.0 ** (a + b + 1) * math.gamma(a + 1) * math.gamma(b + 1) / math.gamma(a + b + 2)
    w = mu0 * V[0, :] ** 2
    return x, w


@functools.lru_cache(maxsize=None)
def make_quadrature(dim: int, m: int):
    """(points (nq, dim), weights (nq,)) on the reference simplex of ``dim``."""
    if dim == 0:
        return np.zeros((1, 0)), np.ones(1)
    n = (m + 2) // 2
    x0, w0 = _golub_welsch(n, 0.0)
    if dim == 1:
        return (0.5 * (x0 + 1.0))[:, None], 0.5 * w0
    x1, w1 = _golub_welsch(n, 1.0)
    if dim == 2:
        I, J = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        I, J = I.ravel(), J.ravel()
        pts = np.stack([0.25 * (1 + x0[I]) * (1 - x1[J]), 0.5 * (1 + x1[J])], axis=1)
        return pts, w0[I] * w1[J] * 0.125
    x2, w2 = _golub_welsch(n, 2.0)
    I, J, K = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    I, J, K = I.ravel(), J.ravel(), K.ravel()
    pts = np.stack(
        [
            0.125 * (1 + x0[I]) * (1 - x1[J]) * (1 - x2[K]),
            0.25 * (1 + x1[J]) * (1 - x2[K]),
            0.5 * (1 + x2[K]),
        ],
        axis=1,
    )
    return pts, w0[I] * w1[J] * w2[K] * 0.015625
