"""Pins of the learned-preconditioner oracle (CPU, FP64).

- ``schur_ell`` equals a dense ``B diag(M)^-1 B^T`` of an assembled RT1-P0 system;
- ``backward`` equals central finite differences of ``twogrid_loss(forward(theta))``
  for parameters of every group (no rounding);
- ``twogrid_loss``'s gamma equals finite differences with respect to symmetric
  perturbations of the smoother entries;
- ``twogrid_apply`` is the symmetric two-grid operator: ``I - M S`` equals
  ``(I - W S)(I - P Ac^-1 P^T S)(I - W S)`` and ``M`` is symmetric;
- a planted mistake (message sum without the coupling weights) fails the
  finite-difference check of the forward/backward pair;
- host and oracle agree on the parameter layout and on the aggregation rule.
"""

import numpy as np
import pytest

from tests._cases import oracle_darcy_space

_SYS = {}


def small_system(n=2, seed=0):
    key = (n, seed)
    if key not in _SYS:
        from oracle.assemble import assemble_darcy
        from oracle.learned import aggregate_cells, schur_ell

        ps = oracle_darcy_space(3, n, 1)
        E = len(ps.spaces["u"].element_ids)
        rng = np.random.default_rng(seed)
        kappa = np.exp(1.5 * rng.standard_normal(E))
        A, rg = assemble_darcy(ps, {"rhs": lambda x, y, z: 1.0 + 0 * x, "kappa": kappa})
        n_u = ps.n_dofs["u"]
        nbr, sd, so = schur_ell(A.indptr, A.indices, A.data, n_u)
        cent = np.array([ps.spaces["u"].mesh.points[c].mean(axis=0) for c in _cells(ps)])
        prow = np.array([ps.destination_indexes(i)[-1] - n_u for i in range(E)])
        agg_e, n_agg = aggregate_cells(cent, 1.0 / n * (2 if n > 2 else 1))
        agg = np.empty(E, dtype=np.int64)
        agg[prow] = agg_e
        _SYS[key] = (A, n_u, nbr, sd, so, agg, n_agg)
    return _SYS[key]


def _cells(ps):
    sp = ps.spaces["u"]
    m = sp.mesh
    return [m.cells[c].node_tags for c in sp.element_ids]


def test_schur_ell_dense():
    from oracle.learned import ell_to_csr

    A, n_u, nbr, sd, so, _, _ = small_system()
    Ad = A.toarray()
    M, B = Ad[:n_u, :n_u], Ad[n_u:, :n_u]
    S = B @ np.diag(1 / np.diag(M)) @ B.T
    got = ell_to_csr(nbr, sd, so).toarray()
    assert np.allclose(got, S, rtol=1e-13, atol=1e-13 * abs(S).max())
    # every nonzero of S is in the ELL pattern, and the pattern is symmetric
    assert ((np.abs(S) > 1e-14 * abs(S).max()) <= (got != 0)).all()
    assert (nbr >= -1).all() and ((nbr >= 0).sum(1) <= 4).all()


def _theta(seed=1):
    from oracle.learned import param_layout

    lay, n = param_layout()
    rng = np.random.default_rng(seed)
    th = rng.standard_normal(n) * 0.15
    return th


def _loss(theta, sysd, e, fwd=None):
    from oracle.learned import forward, twogrid_loss

    A, n_u, nbr, sd, so, agg, n_agg = sysd
    Wd, Wo, cache = (fwd or forward)(theta, nbr, sd, so, emulate_bf16=False)
    return twogrid_loss(Wd, Wo, nbr, sd, so, agg, n_agg, e), cache


def test_backward_matches_finite_differences():
    from oracle.learned import backward, param_layout

    sysd = small_system()
    E = len(sysd[3])
    e = np.random.default_rng(2).standard_normal((E, 6))
    th = _theta()
    (L, gd, go), cache = _loss(th, sysd, e)
    g = backward(th, cache, gd, go)
    lay, n = param_layout()
    rng = np.random.default_rng(3)
    idx = []
    for name, (o, shape) in lay.items():
        size = int(np.prod(shape))
        idx += list(o + rng.choice(size, size=min(3, size), replace=False))
    h = 1e-6
    for i in idx:
        tp, tm = th.copy(), th.copy()
        tp[i] += h
        tm[i] -= h
        fd = (_loss(tp, sysd, e)[0][0] - _loss(tm, sysd, e)[0][0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-5 * np.abs(g).max(), (i, fd, g[i])


def test_gamma_matches_finite_differences():
    from oracle.learned import twogrid_loss

    A, n_u, nbr, sd, so, agg, n_agg = small_system()
    E = len(sd)
    rng = np.random.default_rng(4)
    Wd = 0.65 / sd * (1 + 0.2 * rng.standard_normal(E))
    Wo = np.zeros((E, 4))
    has = nbr >= 0
    for p in range(E):
        for j in range(4):
            q = nbr[p, j]
            if q > p:
                v = 0.1 * rng.standard_normal() / np.sqrt(sd[p] * sd[q])
                Wo[p, j] = v
                Wo[q, list(nbr[q]).index(p)] = v
    e = rng.standard_normal((E, 5))
    L, gd, go = twogrid_loss(Wd, Wo, nbr, sd, so, agg, n_agg, e)
    h = 1e-7 / sd.mean()
    for p in rng.choice(E, 6, replace=False):
        Wp, Wm = Wd.copy(), Wd.copy()
        Wp[p] += h
        Wm[p] -= h
        fd = (twogrid_loss(Wp, Wo, nbr, sd, so, agg, n_agg, e)[0] -
              twogrid_loss(Wm, Wo, nbr, sd, so, agg, n_agg, e)[0]) / (2 * h)
        assert abs(fd - gd[p]) <= 1e-5 * abs(fd) + 1e-9 * np.abs(gd).max()
        for j in range(4):
            q = nbr[p, j]
            if q < 0:
                continue
            k = list(nbr[q]).index(p)
            Op, Om = Wo.copy(), Wo.copy()
            Op[p, j] += h
            Op[q, k] += h
            Om[p, j] -= h
            Om[q, k] -= h
            fd = (twogrid_loss(Wd, Op, nbr, sd, so, agg, n_agg, e)[0] -
                  twogrid_loss(Wd, Om, nbr, sd, so, agg, n_agg, e)[0]) / (2 * h)
            assert abs(fd - go[p, j]) <= 1e-5 * abs(fd) + 1e-9 * np.abs(go).max()
    assert np.allclose(go[~has], 0)


def test_twogrid_apply_is_the_symmetric_two_grid():
    from oracle.learned import coarse_matrix, ell_to_csr, forward, twogrid_apply

    A, n_u, nbr, sd, so, agg, n_agg = small_system()
    E = len(sd)
    th = _theta(5)
    Wd, Wo, _ = forward(th, nbr, sd, so, emulate_bf16=False)
    S = ell_to_csr(nbr, sd, so).toarray()
    W = ell_to_csr(nbr, Wd, Wo).toarray()
    assert np.allclose(W, W.T, rtol=0, atol=1e-14 * abs(W).max())
    P = np.zeros((E, n_agg))
    P[np.arange(E), agg] = 1
    Ac = coarse_matrix(nbr, sd, so, agg, n_agg)
    assert np.allclose(Ac, P.T @ S @ P)
    M = np.stack([twogrid_apply(Wd, Wo, nbr, sd, so, agg, n_agg, col) for col in np.eye(E)], axis=1)
    I = np.eye(E)
    C = P @ np.linalg.solve(Ac, P.T)
    assert np.allclose(I - M @ S, (I - W @ S) @ (I - C @ S) @ (I - W @ S), atol=1e-10)
    assert np.allclose(M, M.T, atol=1e-10 * abs(M).max())


def test_planted_message_mistake_is_caught():
    """Dropping the coupling weights from the message sum (forward only) breaks the
    agreement of backward with finite differences."""
    import oracle.learned as ol

    sysd = small_system()
    E = len(sysd[3])
    e = np.random.default_rng(2).standard_normal((E, 4))
    th = _theta()

    def bad_forward(theta, nbr, sd, so, emulate_bf16=False):
        a_true = ol.coupling

        def flat(nbr_, sd_, so_):
            return np.where(nbr_ >= 0, 1.0, 0.0)

        ol.coupling = flat
        try:
            return ol.forward(theta, nbr, sd, so, emulate_bf16)
        finally:
            ol.coupling = a_true

    (L, gd, go), cache = _loss(th, sysd, e)
    g = ol.backward(th, cache, gd, go)  # true backward, applied to the bad forward below
    lay, _ = ol.param_layout()
    o = lay["Wm1"][0]
    worst = 0.0
    h = 1e-6
    for i in range(o, o + 64, 8):
        tp, tm = th.copy(), th.copy()
        tp[i] += h
        tm[i] -= h
        fd = (_loss(tp, sysd, e, bad_forward)[0][0] - _loss(tm, sysd, e, bad_forward)[0][0]) / (2 * h)
        worst = max(worst, abs(fd - g[i]) / max(abs(fd), 1e-12))
    assert worst > 1e-3


def test_host_layout_and_aggregation_match_oracle():
    import nfempy_amd.learned as hl
    import oracle.learned as ol

    assert hl.param_layout() == ol.param_layout()
    assert (hl.F_IN, hl.WIDTH, hl.LAYERS, hl.OMEGA0) == (ol.F_IN, ol.WIDTH, ol.LAYERS, ol.OMEGA0)
    rng = np.random.default_rng(0)
    c = rng.uniform(0, 1, (500, 3))
    a1, n1 = hl.aggregate_cells(c, 0.25)
    a2, n2 = ol.aggregate_cells(c, 0.25)
    assert n1 == n2 and np.array_equal(a1, a2)


def test_aggregates_of_kuhn_cube_are_hex_blocks():
    from nfempy_amd.learned import aggregate_cells
    from nfempy_amd.mesh import unit_cube

    n = 8
    pts, blocks = unit_cube(n)
    tets = [b for b in blocks if b.dim == 3][0].data
    cent = pts[tets].mean(axis=1)
    agg, na = aggregate_cells(cent, 2.0 / n)
    assert na == (n // 2) ** 3
    hexid = np.arange(len(tets)) // 6
    i, j, k = hexid % n, (hexid // n) % n, hexid // (n * n)
    expect = (i // 2) + (n // 2) * ((j // 2) + (n // 2) * (k // 2))
    assert np.array_equal(agg, expect)
    assert np.all(np.bincount(agg) == 48)


def test_flux_sweeps_pins():
    """Two sweeps on a diagonal M are omega (2 - omega) diag(M)^-1; many sweeps converge
    to M^-1 r on the assembled flux mass; one sweep is the diagonal scaling. A planted
    sign slip in the correction (z - omega D^-1 (r - M z)) fails the limit check."""
    import scipy.sparse as sps
    import scipy.sparse.linalg as spla

    from oracle.learned import flux_sweeps

    A, n_u, *_ = small_system(2, seed=1)
    rng = np.random.default_rng(4)
    r = rng.standard_normal(n_u)
    d = A.diagonal()[:n_u]
    assert np.allclose(flux_sweeps(A.indptr, A.indices, A.data, n_u, r, 0.8, sweeps=1), 0.8 * r / d, rtol=1e-14)
    D = sps.csr_matrix(sps.diags(np.r_[d, np.ones(A.shape[0] - n_u)]))
    got = flux_sweeps(D.indptr, D.indices, D.data, n_u, r, 0.8)
    assert np.allclose(got, 0.8 * 1.2 * r / d, rtol=1e-14)
    M = A[:n_u, :n_u].tocsc()
    exact = spla.spsolve(M, r)
    lim = flux_sweeps(A.indptr, A.indices, A.data, n_u, r, 0.8, sweeps=200)
    assert np.linalg.norm(lim - exact) <= 1e-10 * np.linalg.norm(exact)
    z = np.zeros(n_u)
    for _ in range(200):
        z = z - 0.8 * (r - M @ z) / d
    assert not np.linalg.norm(z - exact) <= 1e-2 * np.linalg.norm(exact)


def test_flux_mass_spectrum_of_kuhn_cube():
    """diag(M)^-1 M of the RT1 flux mass on the Kuhn cube has its spectrum in [0.5, 2]
    (for any element-wise constant permeability): the two flux sweeps' weight 0.8 is
    2 / (0.5 + 2), with error factor (1 - 0.8 l)^2 <= 0.36 over the spectrum."""
    from nfempy_amd.learned import FLUX_OMEGA

    for seed in (0, 1):
        A, n_u, *_ = small_system(2, seed=seed)
        M = A[:n_u, :n_u].toarray()
        s = 1 / np.sqrt(np.diag(M))
        ev = np.linalg.eigvalsh(s[:, None] * M * s[None, :])
        assert ev.min() >= 0.5 - 1e-10 and ev.max() <= 2 + 1e-10, (ev.min(), ev.max())
    assert FLUX_OMEGA == 2 / (0.5 + 2)
