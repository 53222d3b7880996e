"""Pin the oracle's element kernels and assembled systems.

- one-element RT1-P0 Darcy matrices derived by hand on the reference triangle,
  and by exact symbolic integration of the textbook RT1 basis on the reference
  tetrahedron (independent of the oracle's dual-basis construction);
- Piola invariance: a rotated and scaled copy of the element gives the same matrix;
- Dirichlet facet term: |r| = |p_D| x (unit flux of the facet DoF);
- manufactured-solution convergence (first order for RT1-P0 flux and pressure,
  second order for RT2-P1) and symmetry / definiteness of the elasticity operators.
"""

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle.mesh as om
import oracle.spaces as osp
from oracle.assemble import assemble_darcy, assemble_elasticity, symmetric_from_upper
from oracle.forms import darcy_element
from oracle.norms import l2_error
from tests._cases import darcy_functions, oracle_darcy_space


def one_element_space(dim, X):
    if dim == 2:
        blocks = [om.Block(2, [[0, 1, 2]], [1])]
    else:
        blocks = [om.Block(3, [[0, 1, 2, 3]], [1])]
    m = om.ConformalMesh(dim, X, blocks)
    ps = osp.ProductSpace({"u": (dim, 1, "RT", 1, m), "p": (dim, 1, "Lagrange", 0, m)}, {"p": True})
    ps.build_structures({"u": [1], "p": [1]}, {})
    return ps


ONE = {"rhs": lambda x, y, z: np.array([np.ones_like(x)]), "kappa": lambda x, y, z: 1.0}


def test_reference_triangle_by_hand():
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], dtype=float)
    r, J = darcy_element(one_element_space(2, X), 0, ONE)
    M = np.array([[1 / 6, 0, 0], [0, 1 / 3, 1 / 6], [0, 1 / 6, 1 / 3]])
    B = np.array([-1.0, 1.0, -1.0])  # int div phi_j over the triangle
    expect = np.zeros((4, 4))
    expect[:3, :3] = M
    expect[:3, 3] = -B
    expect[3, :3] = B
    assert np.allclose(J, expect, atol=1e-14)
    assert np.allclose(r, [0, 0, 0, -0.5], atol=1e-15)


def test_reference_tet_by_symbolic_integration():
    import sympy as sp

    x, y, z = sp.symbols("x y z")
    phi = [sp.Matrix([2 * x, 2 * y, 2 * z]), sp.Matrix([2 - 2 * x, -2 * y, -2 * z]),
           sp.Matrix([2 * x, 2 * y - 2, 2 * z]), sp.Matrix([-2 * x, -2 * y, 2 - 2 * z])]

    def integ(e):
        return sp.integrate(sp.integrate(sp.integrate(e, (z, 0, 1 - x - y)), (y, 0, 1 - x)), (x, 0, 1))

    M = np.array([[float(integ(phi[i].dot(phi[j]))) for j in range(4)] for i in range(4)])
    B = np.array([float(integ(sum(sp.diff(phi[j][c], v) for c, v in enumerate((x, y, z))))) for j in range(4)])
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=float)
    r, J = darcy_element(one_element_space(3, X), 0, ONE)
    assert np.allclose(J[:4, :4], M, atol=1e-14)
    assert np.allclose(J[4, :4], B, atol=1e-14)
    assert np.allclose(J[:4, 4], -B, atol=1e-14)
    assert abs(r[4] + 1 / 6) < 1e-15


@pytest.mark.parametrize("dim", [2, 3])
def test_piola_invariance(dim):
    rng = np.random.default_rng(3)
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=float)[: dim + 1]
    A, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    if np.linalg.det(A) < 0:
        A[:, 0] *= -1
    if dim == 2:
        # rotation inside the plane keeps z = 0 for the planar mapping
        th = 0.7
        A = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    Y = 2.5 * X @ A.T + np.array([0.3, -1.0, 0.0 if dim == 2 else 2.0])
    _, J0 = darcy_element(one_element_space(dim, X), 0, ONE)
    _, J1 = darcy_element(one_element_space(dim, Y), 0, ONE)
    nb = dim + 1
    # mass block: int phi^T J^T J phi / detJ scales as s^(2 - dim); div blocks are invariant
    assert np.allclose(J1[:nb, :nb], J0[:nb, :nb] * 2.5 ** (2 - dim), atol=1e-12)
    assert np.allclose(J1[nb:, :], J0[nb:, :], atol=1e-12)
    assert np.allclose(J1[:, nb:], J0[:, nb:], atol=1e-12)


def test_dirichlet_term_unit_flux():
    from oracle.forms import darcy_bc_facet

    ps = oracle_darcy_space(2, 2, 1)
    pd = {"p": lambda x, y, z: np.array([3.0])}
    for j in range(len(ps.spaces["u"].bc_element_ids)):
        r, _ = darcy_bc_facet(ps, j, pd)
        assert abs(abs(r[0]) - 3.0) < 1e-13


def _darcy_errors(dim, n, k):
    ps = oracle_darcy_space(dim, n, k)
    fun, ex = darcy_functions(dim)
    A, rg = assemble_darcy(ps, fun, ex)
    x = spla.spsolve(A.tocsc(), -rg)
    return l2_error(ps, ex, x)


@pytest.mark.parametrize("dim,k,ns,rate", [(2, 1, (4, 8), 1.0), (2, 2, (4, 8), 2.0), (3, 1, (2, 4), 1.0)])
def test_darcy_convergence(dim, k, ns, rate):
    e0 = _darcy_errors(dim, ns[0], k)
    e1 = _darcy_errors(dim, ns[1], k)
    for a, b in zip(e0, e1):
        assert np.log2(a / b) > rate - 0.25, (e0, e1)


def test_elasticity_operators_symmetric_definite():
    pts, bl = om.structured_cube(1)
    m = om.ConformalMesh(3, pts, bl)
    ps = osp.ProductSpace({"s": (3, 3, "BDM", 1, m), "u": (3, 3, "Lagrange", 0, m), "t": (3, 3, "Lagrange", 0, m)},
                          {"u": True, "t": True})
    ps.build_structures({"s": [1], "u": [1], "t": [1]}, {"s": [2, 3, 4, 5, 6, 7]})
    rng = np.random.default_rng(0)
    fun = {"rhs": lambda x, y, z: np.zeros((6,) + np.shape(x)), "lambda": np.exp(0.5 * rng.standard_normal(6)),
           "mu": np.exp(0.5 * rng.standard_normal(6))}
    A, P, rg = assemble_elasticity(ps, fun)
    Af = symmetric_from_upper(A).toarray()
    Pf = symmetric_from_upper(P).toarray()
    assert np.all(np.linalg.eigvalsh(Pf) > 0)
    ev = np.linalg.eigvalsh(Af)
    ns = ps.n_dofs["s"]
    # saddle point: as many negative eigenvalues as multiplier DoFs (u, t)
    assert np.sum(ev < -1e-12) == ps.n_dof - ns
    assert np.all(np.abs(ev) > 1e-10)
