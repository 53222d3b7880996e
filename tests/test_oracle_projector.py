"""The L2-projection oracle pinned by closed forms (CPU only).

Fields inside the space are reproduced exactly (linear scalar in P1, cell values of a
linear scalar in P0 at the centroid, constant vector in RT1, linear vector in RT2 and in
BDM1 rows); the norm of a projected constant is its closed form; a field outside the
space is checked against a least-squares fit of the discrete field to the function's
values at the quadrature points (the same minimisation, done without the mass matrix);
the reference's value filter on a mesh scaled so that exactly the off-diagonal P1
contributions fall below it leaves the closed-form diagonal sum_K |K| / 10."""

import os

import numpy as np
import pytest

from tests._cases import MESHES


def oracle_space(dim, fam, k, nc, n=2, msh=None, scale=1.0, disc=False):
    import oracle.mesh as om
    import oracle.spaces as osp

    if msh:
        pts, bl = om.read_msh41(os.path.join(MESHES, msh))
    else:
        pts, bl = om.structured_square(n) if dim == 2 else om.structured_cube(n)
    m = om.ConformalMesh(dim, np.asarray(pts, float) * scale, bl)
    ps = osp.ProductSpace({"q": (dim, nc, fam, k, m)}, {"q": disc})
    ps.build_structures({"q": [1]}, {"q": []})
    return ps


def max_field_error(ps, alpha, f, vector):
    from oracle.projector import field_at_points

    err = 0.0
    for i in range(len(ps.spaces["q"].element_ids)):
        x, fh = field_at_points(ps, "q", alpha, i)
        fe = np.asarray(f(x[:, 0], x[:, 1], x[:, 2]), dtype=float)
        if vector:
            fe = np.moveaxis(fe.reshape(fe.shape[0], fe.shape[1], -1), -1, 0)[:, :, :fh.shape[2]]
        else:
            fe = fe.reshape(fe.shape[0], -1).T
        err = max(err, float(np.abs(fe - fh).max()))
    return err


def lin3(x, y, z):
    return np.array([1 + 2 * x - y + 3 * z])


def vec_const(x, y, z):
    o = np.ones_like(np.asarray(x, float))
    return np.array([[1.0 * o, -2.0 * o, 0.5 * o]])


@pytest.mark.parametrize("case", [
    ("P1 linear", 3, "Lagrange", 1, 1, lin3, False),
    ("RT1 constant", 3, "RT", 1, 1, vec_const, True),
    ("RT2 linear 2d", 2, "RT", 2, 1, lambda x, y, z: np.array([[1 + x - 2 * y, 3 * x + y]]), True),
    ("BDM1 linear rows", 3, "BDM", 1, 3, lambda x, y, z: np.array([
        [1 + x, y - z, 2 * z], [x * 0 + 1, 2 * x + y, z], [x - y, 3 + 0 * x, y + z]]), True),
])
def test_projection_reproduces_fields_in_the_space(case):
    from oracle.projector import l2_projector

    name, dim, fam, k, nc, f, vector = case
    ps = oracle_space(dim, fam, k, nc, n=2)
    alpha = l2_projector(ps, {"q": f})
    assert max_field_error(ps, alpha, f, vector) < 1e-10, name


def test_p0_projection_is_the_cell_mean():
    from oracle.projector import field_at_points, l2_projector

    ps = oracle_space(3, "Lagrange", 0, 1, n=2, disc=True)
    alpha = l2_projector(ps, {"q": lin3})
    sp = ps.spaces["q"]
    for i in range(len(sp.element_ids)):
        X = sp.mesh.points[sp.mesh.cells[sp.element_ids[i]].node_tags]
        c = X.mean(axis=0)
        _, fh = field_at_points(ps, "q", alpha, i)
        assert np.allclose(fh, lin3(*c)[0], rtol=0, atol=1e-12)


def test_projected_constant_norm_closed_form():
    from oracle.projector import l2_error_projected, l2_projector

    ps = oracle_space(3, "RT", 1, 1, n=2)
    alpha = l2_projector(ps, {"q": vec_const})
    (nrm,) = l2_error_projected(ps, alpha, 3)
    assert abs(nrm - np.sqrt(1 + 4 + 0.25)) < 1e-12


@pytest.mark.parametrize("fam,k,dim", [("RT", 1, 3), ("Lagrange", 1, 3), ("RT", 2, 2)])
def test_projection_is_the_least_squares_fit(fam, k, dim):
    """For f outside the space the projection minimises sum_e sum_q w detJ |f - f_h|^2:
    the same problem solved as weighted least squares on the point values."""
    from oracle.forms import _cell, _tabulate
    from oracle.geometry import evaluate_mapping
    from oracle.projector import l2_projector

    vector = fam != "Lagrange"
    if vector:
        f = lambda x, y, z: np.array([[np.sin(3 * x) * y, np.cos(2 * y) + z * x, np.exp(x * z)]])
    else:
        f = lambda x, y, z: np.array([np.sin(3 * x) * np.cos(2 * y) + z * z])
    ps = oracle_space(dim, fam, k, 1, n=2)
    alpha = l2_projector(ps, {"q": f})
    sp = ps.spaces["q"]
    rows, rhs = [], []
    for i in range(len(sp.element_ids)):
        cell = _cell(sp, i)
        points, weights = ps.quadrature[dim]
        x, jac, det, _ = evaluate_mapping(dim, points, sp.mesh.points[cell.node_tags])
        tab = _tabulate(sp, cell, points, jac, det)
        dest = ps.destination_indexes(i, "q")
        fv = np.asarray(f(x[:, 0], x[:, 1], x[:, 2]), float)
        for q in range(len(weights)):
            sw = np.sqrt(weights[q] * det[q])
            for d in range(dim if vector else 1):
                row = np.zeros(ps.n_dof)
                np.add.at(row, dest, sw * (tab[0, q, :, d] if vector else tab[0, q, :, 0]))
                rows.append(row)
                rhs.append(sw * (fv[0, d, q] if vector else fv[0, q]))
    fit = np.linalg.lstsq(np.array(rows), np.array(rhs), rcond=None)[0]
    assert np.abs(fit - alpha).max() < 1e-9 * max(1.0, np.abs(fit).max())


def test_value_filter_drops_small_contributions():
    """Scaled so that |K| ~ 1.5e-7: P1 off-diagonal contributions |K|/20 fall under the
    reference's filter, the diagonal ones |K|/10 do not: the matrix is diag(sum |K|/10)."""
    from oracle.projector import assemble_projector

    s = 0.0193
    ps = oracle_space(3, "Lagrange", 1, 1, n=2, scale=s)
    A, rg = assemble_projector(ps, {"q": lin3})
    vol = (s / 2) ** 3 / 6
    assert 1e-8 < vol / 10 and vol / 20 < 1e-8
    off = A.tocoo()
    assert np.all(off.row[off.data != 0] == off.col[off.data != 0])
    sp = ps.spaces["q"]
    expect = np.zeros(ps.n_dof)
    for i in range(len(sp.element_ids)):
        expect[ps.destination_indexes(i, "q")] += vol / 10
    assert np.allclose(A.diagonal(), expect, rtol=1e-12)
