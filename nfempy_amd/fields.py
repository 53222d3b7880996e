"""Synthetic coefficient fields for the benchmark configurations.

Log-normal permeability / Lame fields: ``log K = g`` with ``g`` a Gaussian random
field of squared-exponential covariance (correlation length ``ell``, standard
deviation ``sigma``), sampled at element centroids through M random Fourier
features, ``g(x) = sigma sqrt(2/M) sum_m cos(w_m . x + b_m)``,
``w_m ~ N(0, ell^-2 I)``, ``b_m ~ U(0, 2 pi)``. The features are drawn on the host
from ``numpy.random.default_rng(seed)`` so the same field can be evaluated on the
device (benchmark) and on the host (tests) from the same seed.
"""

from __future__ import annotations

import numpy as np


def rff_features(seed: int, ell: float, n_features: int = 2048, dim: int = 3):
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((n_features, dim)) / ell
    b = rng.uniform(0.0, 2.0 * np.pi, n_features)
    return w, b


def grf_host(points, seed, ell, sigma, n_features=2048):
    """FP64 host evaluation of g at points (npts, 3)."""
    w, b = rff_features(seed, ell, n_features, points.shape[1])
    out = np.empty(points.shape[0])
    for s in range(0, points.shape[0], 8192):
        ph = points[s:s + 8192] @ w.T + b
        out[s:s + 8192] = np.cos(ph).sum(axis=1)
    return sigma * np.sqrt(2.0 / n_features) * out


def grf_device(points, seed, ell, sigma, n_features=2048, chunk=16384):
    """Device evaluation (torch, FP32 phases accumulated in FP64) of g at points."""
    import torch

    w, b = rff_features(seed, ell, n_features, points.shape[1])
    wt = torch.as_tensor(w, dtype=torch.float32, device=points.device)
    bt = torch.as_tensor(b, dtype=torch.float32, device=points.device)
    pts = points.to(torch.float32)
    out = torch.empty(points.shape[0], dtype=torch.float64, device=points.device)
    for s in range(0, points.shape[0], chunk):
        ph = pts[s:s + chunk] @ wt.T + bt
        out[s:s + chunk] = torch.cos(ph).sum(dim=1, dtype=torch.float64)
    return sigma * np.sqrt(2.0 / n_features) * out


def centroids(points, cells):
    return points[cells].mean(axis=1)


def lame_lognormal(n_elements, seed=2024, sigma=0.5):
    """Per-element Lame parameters of the 3-D elasticity workload: ``lambda_e, mu_e =
    exp(N(0, sigma^2))``, independent, drawn in that order from ``default_rng(seed)``."""
    rng = np.random.default_rng(seed)
    lam = np.exp(sigma * rng.standard_normal(n_elements))
    mu = np.exp(sigma * rng.standard_normal(n_elements))
    return lam, mu


def elasticity_body_force(x, y, z):
    """Constant body force (1, 1/2, 1/4) on the displacement rows, none on the rotation rows."""
    one = np.ones_like(np.asarray(x, dtype=np.float64))
    zero = np.zeros_like(one)
    return np.array([one, 0.5 * one, 0.25 * one, zero, zero, zero])
