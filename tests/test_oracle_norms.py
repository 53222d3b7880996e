"""Closed-form pins of the oracle's error norms ``l2_error`` and ``div_error``
(reference ``src/postprocess/l2_error_post_processor.py:7-63,185-237``).

On an affinely mapped cube (3-D) and square (2-D), with discrete fields built exactly
(RT1 contains u(x) = a + b x, whose divergence is d b; P0 holds one value per cell) and
exact fields that differ from them by a linear function, each squared error is an exact
simplex moment: for linear g on a simplex K with centroid c_K and vertex covariance
C_K = sum_i (v_i - c_K)(v_i - c_K)^T / ((d + 1)(d + 2)),
``int_K g^2 = |K| (g(c_K)^2 + grad g^T C_K grad g)``. A missing square root, a missing
|det J| in the measure, or a missing 1/|det J| in the divergence fails these checks.
"""

import numpy as np
import pytest

import oracle.mesh as om
import oracle.spaces as osp
from oracle.norms import div_error, l2_error
from oracle.projector import l2_projector

RNG = np.random.default_rng(7)


def _setup(dim, n):
    pts, bl = om.structured_cube(n) if dim == 3 else om.structured_square(n)
    M = np.eye(3)
    M[:dim, :dim] = np.array([[1.3, 0.2, -0.1], [0.25, 0.8, 0.3], [-0.2, 0.1, 1.1]])[:dim, :dim]
    pts = pts @ M.T + np.array([0.2, -0.4, 0.6 if dim == 3 else 0.0])
    m = om.ConformalMesh(dim, pts, bl)
    ps = osp.ProductSpace({"u": (dim, 1, "RT", 1, m), "p": (dim, 1, "Lagrange", 0, m)}, {"p": True})
    ps.build_structures({"u": [1], "p": [1]}, {})
    pu = osp.ProductSpace({"u": (dim, 1, "RT", 1, m)})
    pu.build_structures({"u": [1]}, {})
    cells = np.array([m.cells[c].node_tags for c in ps.spaces["u"].element_ids])
    X = m.points[cells][:, :, :dim]
    fact = 6.0 if dim == 3 else 2.0
    vol = np.abs(np.linalg.det(X[:, 1:] - X[:, :1])) / fact
    cen = X.mean(axis=1)
    dev = X - cen[:, None, :]
    C = np.einsum("kai,kaj->kij", dev, dev) / ((dim + 1) * (dim + 2))
    return ps, pu, vol, cen, C


def _moment(vol, cen, C, grad, val0):
    """sum_K int_K (grad . x + val0 - c_K)^2 with c_K folded into val0 (per cell)."""
    return float(np.sum(vol * ((cen @ grad + val0) ** 2 + np.einsum("i,kij,j->k", grad, C, grad))))


@pytest.mark.parametrize("dim", [2, 3])
def test_l2_and_div_error_closed_forms(dim):
    ps, pu, vol, cen, C = _setup(dim, 2)
    E = len(vol)
    a = RNG.standard_normal(dim)
    b = RNG.standard_normal()
    pad = lambda v: np.concatenate([v, np.zeros(3 - dim)])

    def u_h(x, y, z):
        P = np.stack([np.atleast_1d(x), np.atleast_1d(y), np.atleast_1d(z)])[:dim]
        return (pad(a)[:, None] + b * np.concatenate([P, np.zeros((3 - dim, P.shape[1]))]))[None, :, :]

    alpha_u = l2_projector(pu, {"u": u_h})
    pc = RNG.standard_normal(E)
    alpha = np.zeros(ps.n_dof)
    alpha[: len(alpha_u)] = alpha_u
    for i in range(E):
        alpha[ps.destination_indexes(i, "p")] = pc[i]
    # exact fields in the discrete space: zero error
    G = RNG.standard_normal((dim, dim))
    g = RNG.standard_normal(dim)
    g0 = RNG.standard_normal()

    def u_ex(extra):
        def f(x, y, z):
            xv = np.array([x, y, z], dtype=float)[:dim]
            v = pad(a) + b * pad(xv) + (pad(G @ xv) if extra else 0.0)
            return v[None, :]
        return f

    def p_ex(extra):
        def f(x, y, z):
            X = np.stack([np.atleast_1d(x), np.atleast_1d(y), np.atleast_1d(z)])[:dim]
            return (g @ X + g0)[None, :] if extra else None
        return f

    # pressure: the exact field is linear; the discrete one is pc per cell
    e_u0, _ = l2_error(ps, {"u": u_ex(False), "p": p_ex(True)}, alpha)
    assert e_u0 < 1e-12
    e_u, e_p = l2_error(ps, {"u": u_ex(True), "p": p_ex(True)}, alpha)
    # |G x|^2 summed over components: each component is a linear function
    expect_u = sum(_moment(vol, cen, C, G[r], 0.0) for r in range(dim))
    expect_p = _moment(vol, cen, C, g, g0 - pc)
    assert abs(e_u - np.sqrt(expect_u)) < 1e-12 * np.sqrt(expect_u), (e_u, np.sqrt(expect_u))
    assert abs(e_p - np.sqrt(expect_p)) < 1e-12 * np.sqrt(expect_p), (e_p, np.sqrt(expect_p))
    # divergence: div u_h = dim b; exact divergence dim b (zero) or a linear function
    (e_d0,) = div_error(ps, {"div_u": lambda x, y, z: np.array([dim * b])}, alpha)
    assert e_d0 < 1e-12
    (e_d,) = div_error(ps, {"div_u": lambda x, y, z: np.array([g @ np.array([x, y, z])[:dim] + g0])}, alpha)
    expect_d = _moment(vol, cen, C, g, g0 - dim * b)
    assert abs(e_d - np.sqrt(expect_d)) < 1e-12 * np.sqrt(expect_d), (e_d, np.sqrt(expect_d))
