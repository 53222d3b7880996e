"""Pins of the oracle's dual-mixed elasticity forms (``elasticity_element``,
``riesz_element``, ``elasticity_bc_facet``; reference ``le_dual_weak_form.py:354-578``,
``le_riesz_map_weak_form.py:258-360``) that do not restate them:

1. Closed-form bilinear identities on an affinely mapped 6-tet cube with per-element
   Lame parameters. Every linear tensor field lies in BDM1 (rows are full linear vectors),
   so the L2 projection of a linear or constant field reproduces it exactly; P0 fields
   are one value per cell. Then, with ``a`` the coefficient vectors,
   - ``a_tau^T A a_sigma = sum_K |K| / 2 mu_K (sigma:tau - lam_K / (2 mu_K + 3 lam_K) tr sigma tr tau)``
     for constant sigma, tau (compliance and volumetric factor);
   - ``a_v^T A a_sigma = sum_K |K| v_K . div sigma`` (the (u, s) block and its sign);
   - ``a_w^T A a_sigma = sum_K |K| w_K . (s_yz - s_zy, s_zx - s_xz, s_xy - s_yx)(c_K)``
     (the (t, s) skew coupling: which stress entries each rotation component pairs with,
     and the sign);
   - residual rows ``-|K| f_c``; the Dirichlet term ``a_sigma . r = -int_dOmega (G x) . sigma n
     = -|Omega| sigma : G`` (divergence theorem, u_D = G x);
   - the Riesz map: ``int A sigma : tau + int div sigma . div tau`` for linear fields (exact
     simplex moments), and unit-mass blocks on u and t.
   Planted mistakes in the oracle (a volumetric factor with dim = 2, a swapped or
   sign-flipped skew map) fail these identities.
2. Convergence on manufactured solutions: O(h) for sigma, u, t and div sigma, in 3-D
   (1^3, 2^3, 4^3 cubes) and in 2-D with the displacement of the reference's own example
   (``examples/cauchy_elasticity/strong_solution_elasticity_example_1.py``), for
   lambda = 1 and lambda = 100.
"""

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle.forms as of
import oracle.mesh as om
import oracle.spaces as osp
from oracle.norms import div_error, l2_error
from oracle.projector import l2_projector
from tests._cases import elasticity_functions

RNG = np.random.default_rng(20261203)
MAP = np.array([[1.2, 0.3, -0.2], [0.1, 0.9, 0.25], [-0.15, 0.2, 1.1]])
SHIFT = np.array([0.3, -0.5, 0.7])


def _space(dim, n, btags, mapped=True):
    pts, bl = om.structured_cube(n) if dim == 3 else om.structured_square(n)
    if mapped and dim == 3:
        pts = pts @ MAP.T + SHIFT
    m = om.ConformalMesh(dim, pts, bl)
    tc = 3 if dim == 3 else 1
    ps = osp.ProductSpace({"s": (dim, dim, "BDM", 1, m), "u": (dim, dim, "Lagrange", 0, m),
                           "t": (dim, tc, "Lagrange", 0, m)}, {"u": True, "t": True})
    ps.build_structures({"s": [1], "u": [1], "t": [1]}, {"s": btags})
    sp_s = osp.ProductSpace({"s": (dim, dim, "BDM", 1, m)})
    sp_s.build_structures({"s": [1]}, {})
    return ps, sp_s, m


def _dense(ps, fn, fun):
    n = ps.n_dof
    A = np.zeros((n, n))
    r = np.zeros(n)
    for i in range(len(ps.spaces["s"].element_ids)):
        d = ps.destination_indexes(i)
        ri, Ji = fn(ps, i, fun)
        A[np.ix_(d, d)] += Ji
        r[d] += ri
    return A, r


def _cells(ps, m):
    cells = np.array([m.cells[c].node_tags for c in ps.spaces["s"].element_ids])
    X = m.points[cells]
    vol = np.abs(np.linalg.det(X[:, 1:] - X[:, :1])) / 6.0
    return X, vol


def _project(sp_s, ps, field):
    a_s = l2_projector(sp_s, {"s": field})
    a = np.zeros(ps.n_dof)
    for i in range(len(ps.spaces["s"].element_ids)):
        assert np.array_equal(sp_s.destination_indexes(i, "s"), ps.destination_indexes(i, "s"))
    a[: len(a_s)] = a_s
    return a


def _cell_field(ps, name, values):
    a = np.zeros(ps.n_dof)
    for i, v in enumerate(values):
        a[ps.destination_indexes(i, name)] = v
    return a


def _const(S):
    return lambda x, y, z: np.repeat(S[:, :, None], np.size(x), axis=2)


def _linear(S0, S):
    """sigma(x) = S0 + sum_k S[k] x_k as (3, 3, nq)."""
    def f(x, y, z):
        pts = np.stack([np.atleast_1d(x), np.atleast_1d(y), np.atleast_1d(z)])
        return S0[:, :, None] + np.einsum("kij,kq->ijq", S, pts)
    return f


def _skewvec(s):
    return np.array([s[1, 2] - s[2, 1], s[2, 0] - s[0, 2], s[0, 1] - s[1, 0]])


def _simplex_product(fv, gv, vol):
    """Exact integral of the product of two linear functions on a tet from vertex values."""
    M = (np.ones((4, 4)) + np.eye(4)) / 20.0
    return vol * fv @ M @ gv


def identity_errors():
    """Relative errors of every closed-form identity (all ~1e-14 for a correct oracle)."""
    ps, sp_s, m = _space(3, 1, [2, 3, 4, 5, 6, 7])
    X, vol = _cells(ps, m)
    E = len(vol)
    lam = np.exp(0.5 * RNG.standard_normal(E))
    mu = np.exp(0.5 * RNG.standard_normal(E))
    f0 = RNG.standard_normal(6)
    fun = {"lambda": lam, "mu": mu, "rhs": lambda x, y, z: np.repeat(f0[:, None], np.size(x), axis=1)}
    A, r = _dense(ps, of.elasticity_element, fun)
    P, _ = _dense(ps, of.riesz_element, fun)
    err = {}
    S1, S2 = RNG.standard_normal((3, 3)), RNG.standard_normal((3, 3))
    a1, a2 = _project(sp_s, ps, _const(S1)), _project(sp_s, ps, _const(S2))
    got = a2 @ A @ a1
    expect = np.sum(vol / (2 * mu) * (np.sum(S1 * S2) - lam / (2 * mu + 3 * lam) * np.trace(S1) * np.trace(S2)))
    err["ss_volumetric"] = abs(got - expect) / abs(expect)
    L0, L = RNG.standard_normal((3, 3)), RNG.standard_normal((3, 3, 3))
    al = _project(sp_s, ps, _linear(L0, L))
    divs = np.array([sum(L[j][c, j] for j in range(3)) for c in range(3)])
    v = RNG.standard_normal((E, 3))
    av = _cell_field(ps, "u", v)
    expect = np.sum(vol * (v @ divs))
    err["us"] = abs(av @ A @ al - expect) / abs(expect)
    w = RNG.standard_normal((E, 3))
    aw = _cell_field(ps, "t", w)
    cen = X.mean(axis=1)
    expect = sum(vol[K] * w[K] @ _skewvec(L0 + np.einsum("kij,k->ij", L, cen[K])) for K in range(E))
    err["ts"] = abs(aw @ A @ al - expect) / abs(expect)
    ru = np.array([r[ps.destination_indexes(i, "u")] for i in range(E)])
    rt = np.array([r[ps.destination_indexes(i, "t")] for i in range(E)])
    err["rhs"] = max(np.abs(ru + vol[:, None] * f0[None, :3]).max(), np.abs(rt + vol[:, None] * f0[None, 3:]).max())
    # Dirichlet term with u_D = G x + g0 on every boundary face
    G, g0 = RNG.standard_normal((3, 3)), RNG.standard_normal(3)
    uD = {"u": lambda x, y, z: G @ np.array([x, y, z]).reshape(3, -1) + g0[:, None]}
    rb = np.zeros(ps.n_dof)
    for j in range(len(ps.spaces["s"].bc_element_ids)):
        rj, _ = of.elasticity_bc_facet(ps, j, uD)
        rb[ps.bc_destination_indexes(j)] += rj
    Vol = vol.sum()
    expect = -Vol * np.sum(S1 * G)
    err["bc"] = abs(a1 @ rb - expect) / abs(expect)
    # Riesz map: compliance + div-div on two linear fields, unit mass on u and t
    M0, Mk = RNG.standard_normal((3, 3)), RNG.standard_normal((3, 3, 3))
    am = _project(sp_s, ps, _linear(M0, Mk))
    divm = np.array([sum(Mk[j][c, j] for j in range(3)) for c in range(3)])
    expect = 0.0
    for K in range(E):
        sl = np.array([L0 + np.einsum("kij,k->ij", L, X[K, a]) for a in range(4)])  # vertex values
        sm = np.array([M0 + np.einsum("kij,k->ij", Mk, X[K, a]) for a in range(4)])
        ab = lam[K] / (2 * mu[K] + 3 * lam[K])
        comp = sum(_simplex_product(sl[:, i, j], sm[:, i, j], vol[K]) for i in range(3) for j in range(3))
        tr = _simplex_product(np.trace(sl, axis1=1, axis2=2), np.trace(sm, axis1=1, axis2=2), vol[K])
        expect += (comp - ab * tr) / (2 * mu[K]) + vol[K] * divs @ divm
    err["riesz"] = abs(am @ P @ al - expect) / abs(expect)
    err["riesz_mass"] = max(abs(av @ P @ av - np.sum(vol[:, None] * v * v)) / np.sum(vol[:, None] * v * v),
                            abs(aw @ P @ aw - np.sum(vol[:, None] * w * w)) / np.sum(vol[:, None] * w * w))
    return err


def test_closed_form_identities():
    err = identity_errors()
    assert max(err.values()) < 1e-12, err


@pytest.mark.parametrize("mistake", ["volumetric_dim2", "skew_swapped", "skew_sign"])
def test_planted_mistakes_fail(monkeypatch, mistake):
    if mistake == "volumetric_dim2":
        monkeypatch.setattr(of, "volumetric_factor", lambda lam, mu, dim: (1 / (2 * mu)) * lam / (2 * mu + 2 * lam))
        key = "ss_volumetric"
    elif mistake == "skew_swapped":
        sk = {3: dict(of.SKEW_MAP[3]), 2: of.SKEW_MAP[2]}
        sk[3][0], sk[3][1] = of.SKEW_MAP[3][1], of.SKEW_MAP[3][0]
        monkeypatch.setattr(of, "SKEW_MAP", sk)
        key = "ts"
    else:
        sk = {3: {c: [(r, d, -s) for r, d, s in v] for c, v in of.SKEW_MAP[3].items()}, 2: of.SKEW_MAP[2]}
        monkeypatch.setattr(of, "SKEW_MAP", sk)
        key = "ts"
    err = identity_errors()
    assert err[key] > 1e-2, (mistake, err)


def _convergence(dim, ns, lam, mu):
    out = []
    btags = [2, 3, 4, 5, 6, 7] if dim == 3 else [2, 3, 4, 5]
    fun, ex = elasticity_functions(lam=lam, mu=mu, dim=dim)
    from oracle.assemble import assemble_elasticity, symmetric_from_upper

    for n in ns:
        ps, _, _ = _space(dim, n, btags, mapped=False)
        A, _, rg = assemble_elasticity(ps, fun, {"u": ex["u"]})
        x = spla.spsolve(symmetric_from_upper(A).tocsc(), -rg)
        es, eu, et = l2_error(ps, ex, x)
        (ed,) = div_error(ps, ex, x)
        out.append((es, eu, et, ed))
    return np.array(out)


@pytest.mark.parametrize("lam", [1.0, 100.0])
def test_convergence_3d(lam):
    e = _convergence(3, (1, 2, 4), lam, 1.0)
    rates = np.log2(e[:-1] / e[1:])
    assert np.all(rates[-1] > 0.75), (rates, e)


@pytest.mark.parametrize("lam", [1.0, 100.0])
def test_convergence_2d_reference_solution(lam):
    e = _convergence(2, (4, 8, 16), lam, 1.0)
    rates = np.log2(e[:-1] / e[1:])
    assert np.all(rates[-1] > 0.85), (rates, e)


def test_rotation_sign_is_pinned():
    """Against the opposite rotation convention (t = -skew) the rotation error does not
    decrease: the (t, s) coupling fixes the sign of t."""
    fun, ex = elasticity_functions(dim=2)
    ex_wrong = dict(ex)
    ex_wrong["t"] = lambda x, y, z: -ex["t"](x, y, z)
    from oracle.assemble import assemble_elasticity, symmetric_from_upper

    et = []
    for n in (8, 16):
        ps, _, _ = _space(2, n, [2, 3, 4, 5], mapped=False)
        A, _, rg = assemble_elasticity(ps, fun, {"u": ex["u"]})
        x = spla.spsolve(symmetric_from_upper(A).tocsc(), -rg)
        et.append((l2_error(ps, ex, x)[2], l2_error(ps, ex_wrong, x)[2]))
    assert et[1][0] < 0.6 * et[0][0]
    assert et[1][1] > 0.9 * et[0][1]
