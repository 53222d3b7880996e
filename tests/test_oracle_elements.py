"""Pin the oracle's reference elements and orientation handling.

- RT1 on the reference triangle / tetrahedron against the textbook closed forms
  with basix's (non-normalised) facet normals;
- the Ciarlet dual property l_i(phi_j) = delta_ij re-evaluated by an independent
  numeric quadrature on the facets;
- normal continuity of every global facet DoF across every interior facet of the
  reference's own gmsh meshes (RT1, RT2, BDM1 in 2-D and 3-D) after the
  reference's permute_and_transform: this only holds with the right facet
  transformation matrices and (rotation, reflection) table, so it pins them;
- the product's numeric elements agree with the oracle's exact ones.
"""

import os

import numpy as np
import pytest

from oracle.reference_elements import create_element, facet_normal, ref_facets, ref_vertices
from tests._cases import MESHES

X = np.random.default_rng(1).random((9, 3)) * 0.33


def test_rt1_triangle_closed_form():
    tab = create_element("RT", 2, 1).tabulate(X[:, :2])[0]
    x, y = X[:, 0], X[:, 1]
    expect = [np.stack([-x, -y], 1), np.stack([x - 1, y], 1), np.stack([-x, 1 - y], 1)]
    for j in range(3):
        assert np.allclose(tab[:, j, :], expect[j], atol=1e-14)


def test_rt1_tet_closed_form():
    tab = create_element("RT", 3, 1).tabulate(X)[0]
    x, y, z = X.T
    expect = [2 * np.stack([x, y, z], 1), np.stack([2 - 2 * x, -2 * y, -2 * z], 1),
              np.stack([2 * x, 2 * y - 2, 2 * z], 1), np.stack([-2 * x, -2 * y, 2 - 2 * z], 1)]
    for j in range(4):
        assert np.allclose(tab[:, j, :], expect[j], atol=1e-14)


def _facet_moment(el, j, f, q_index, moment_degree):
    """Independent numeric evaluation of the facet moment functional."""
    d = el.dim
    verts = np.array(ref_vertices(d), dtype=float)[list(ref_facets(d)[f])]
    n = np.array(facet_normal(d, ref_facets(d)[f]), dtype=float).ravel()
    # a dense tensor rule on the facet
    g, w = np.polynomial.legendre.leggauss(8)
    g = 0.5 * (g + 1)
    w = 0.5 * w
    if d == 2:
        s = g[:, None]
        ws = w
    else:
        S0, S1 = np.meshgrid(g, g, indexing="ij")
        W0, W1 = np.meshgrid(w, w, indexing="ij")
        s = np.stack([S0.ravel(), (S1 * (1 - S0)).ravel()], 1)
        ws = (W0 * W1 * (1 - S0)).ravel()
    pts = verts[0] + s @ (verts[1:] - verts[0])
    phi = el.tabulate(pts)[0][:, j, :]
    if moment_degree == 0:
        q = np.ones(len(ws))
    else:
        bary = np.concatenate([1 - s.sum(1, keepdims=True), s], 1)
        q = bary[:, q_index]
    return np.sum(ws * (phi @ n) * q)


@pytest.mark.parametrize("fam,d,k", [("RT", 2, 1), ("RT", 3, 1), ("RT", 2, 2), ("RT", 3, 2), ("BDM", 2, 1), ("BDM", 3, 1)])
def test_dual_property_facets(fam, d, k):
    el = create_element(fam, d, k)
    mdeg = k - 1 if fam == "RT" else k
    for f in range(d + 1):
        for qi, dof in enumerate(el.entity_dofs[d - 1][f]):
            for j in range(el.ndofs):
                v = _facet_moment(el, j, f, qi, mdeg)
                assert abs(v - (1.0 if j == dof else 0.0)) < 1e-12, (f, qi, j, v)


def test_rt2_interior_moments():
    for d in (2, 3):
        el = create_element("RT", d, 2)
        from oracle.quadrature import gauss_jacobi

        pts, w = gauss_jacobi(d, 7)
        tab = el.tabulate(pts)[0]
        for c, dof in enumerate(el.entity_dofs[d][0]):
            for j in range(el.ndofs):
                v = np.sum(w * tab[:, j, c])
                assert abs(v - (1.0 if j == dof else 0.0)) < 1e-12


def test_product_elements_match_oracle():
    from nfempy_amd.basis import create_element as pcreate

    for fam, d, k in [("RT", 2, 1), ("RT", 3, 1), ("RT", 2, 2), ("RT", 3, 2), ("BDM", 2, 1), ("BDM", 3, 1),
                      ("P", 2, 0), ("P", 3, 0), ("P", 2, 1), ("P", 3, 1)]:
        a = pcreate(fam, d, k).tabulate(X[:, :d])
        b = create_element(fam, d, k).tabulate(X[:, :d])
        assert np.allclose(a, b, atol=1e-12), (fam, d, k)
        assert pcreate(fam, d, k).entity_dofs == create_element(fam, d, k).entity_dofs or fam == "P"


def _normal_jumps(msh, d, fam, k, nmax=400):
    import oracle.mesh as om
    import oracle.spaces as osp
    from oracle.forms import _tabulate
    from oracle.geometry import evaluate_mapping

    pts, bl = om.read_msh41(os.path.join(MESHES, msh))
    m = om.ConformalMesh(d, pts, bl)
    ps = osp.ProductSpace({"s": (d, 1, fam, k, m)})
    ps.build_structures({"s": [1]}, {})
    sp = ps.spaces["s"]
    topo = sp.topology
    # interior facets: listed by exactly two cells
    owners = {}
    for i, cid in enumerate(sp.element_ids):
        for f in m.cells[cid].sub_cells_ids[d - 1].tolist():
            owners.setdefault(f, []).append(i)
    interior = [f for f, o in owners.items() if len(o) == 2][:nmax]
    worst = 0.0
    for f in interior:
        fn = m.cells[f].node_tags
        Xf = m.points[fn]
        bary = np.array([[0.6, 0.3, 0.1][:d], [0.2, 0.5, 0.3][:d], [1 / d] * d])
        bary = bary / bary.sum(1, keepdims=True)
        xp = bary @ Xf
        if d == 2:
            t = Xf[1] - Xf[0]
            n = np.array([-t[1], t[0], 0.0])
        else:
            n = np.cross(Xf[1] - Xf[0], Xf[2] - Xf[0])
        vals = {}
        for i in owners[f]:
            cell = m.cells[sp.element_ids[i]]
            Xc = m.points[cell.node_tags]
            J = (Xc[1:] - Xc[0]).T  # 3 x d
            ref = np.linalg.lstsq(J, (xp - Xc[0]).T, rcond=None)[0].T
            x, jac, det, _ = evaluate_mapping(d, ref, Xc)
            tab = _tabulate(sp, cell, ref, jac, det)
            dest = sp.destination_indices(cell.id)
            for j, g in enumerate(dest.tolist()):
                vals.setdefault(g, []).append(tab[0, :, j, :3] @ n)
        for g, v in vals.items():
            if len(v) == 2:
                worst = max(worst, np.abs(v[0] - v[1]).max())
            else:
                # a DoF of one side only must have zero normal trace on the facet
                worst = max(worst, np.abs(v[0]).max())
    return worst


@pytest.mark.parametrize("msh,d", [("example_3_2d_l_0.msh", 2), ("example_3_3d_l_0.msh", 3)])
@pytest.mark.parametrize("fam,k", [("RT", 1), ("RT", 2), ("BDM", 1)])
def test_normal_continuity_on_reference_meshes(msh, d, fam, k):
    assert _normal_jumps(msh, d, fam, k) < 1e-9
