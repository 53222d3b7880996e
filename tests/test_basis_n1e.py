"""The Nedelec (first kind, H(curl)) element family N1E (``src/basis/element_family.py:4-12``)
and its edge orientation (``src/basis/permute_and_transform.py:8-35, 37-55, 129-139``),
pinned by properties the mathematics fixes rather than by restating the construction:

- the space: the curl of every degree-1 function is constant, the span holds the constants
  and the rotations b x x (3-D) / b (-y, x) (2-D), and its dimension is 3 / 6 / 8;
- conformity of the global space: after the covariant Piola map and the orientation signs,
  the tangential component of every global basis function agrees from both sides of every
  interior edge (2-D) / face (3-D) of a mesh (a wrong reflection or sign breaks it);
- exactness: the L2 projection onto the global space reproduces a field of the space
  (a + b x x) to round-off on every element.
"""

import numpy as np
import pytest

from nfempy_amd.basis import covariant_piola, create_element, curl_2d
from nfempy_amd.mesh import Mesh, unit_cube, unit_square
from nfempy_amd.quadrature import make_quadrature
from nfempy_amd.spaces import ProductSpace


@pytest.mark.parametrize("d,k,n", [(2, 1, 3), (3, 1, 6), (2, 2, 8)])
def test_n1e_space(d, k, n):
    el = create_element("N1E", d, k)
    assert el.ndofs == n
    pts, _ = make_quadrature(d, 4)
    tab = el.tabulate(pts)  # (1 + d, nq, nb, d)
    if k == 1:
        if d == 2:
            curl = tab[1, :, :, 1] - tab[2, :, :, 0]
        else:
            curl = np.stack([tab[2, :, :, 2] - tab[3, :, :, 1], tab[3, :, :, 0] - tab[1, :, :, 2],
                             tab[1, :, :, 1] - tab[2, :, :, 0]])
        assert np.allclose(curl, curl[..., :1, :] if d == 2 else curl[:, :1, :])
    # constants and rotations are combinations of the basis: least squares residual ~ 0
    vals = tab[0].transpose(0, 2, 1).reshape(-1, el.ndofs)  # (nq * d, nb)
    fields = [np.tile(np.eye(d)[c], (len(pts), 1)) for c in range(d)]
    if d == 2:
        fields.append(np.stack([-pts[:, 1], pts[:, 0]], axis=1))
    else:
        fields += [np.cross(np.eye(3)[c], pts) for c in range(3)]
    for f in fields:
        coef, *_ = np.linalg.lstsq(vals, f.reshape(-1), rcond=None)
        assert np.abs(vals @ coef - f.reshape(-1)).max() < 1e-12


def _space(d, n):
    mesh = Mesh.from_blocks(d, *(unit_square(n) if d == 2 else unit_cube(n)))
    ps = ProductSpace({"e": (d, 1, "N1E", 1, mesh)})
    ps.build_structures({"e": [1]}, {})
    return ps, mesh


def _global_values(sp, mesh, d, e, pts):
    """(nq, n_dof, d) physical values of the global basis restricted to element e."""
    el = sp.element
    cells = mesh.entities[d].nodes[sp.topology.cell_index]
    X = mesh.points[cells[e]][:, :d]
    J = (X[1:] - X[0]).T
    phi = covariant_piola(el.tabulate(pts)[0], J)  # (nq, nb, d)
    out = np.zeros((len(pts), sp.n_dof, d))
    for b in range(el.ndofs):
        out[:, sp.ref_dest[e, b], :] += sp.ref_sign[e, b] * phi[:, b, :]
    return out, X


@pytest.mark.parametrize("d", [2, 3])
def test_n1e_tangential_continuity(d):
    ps, mesh = _space(d, 2)
    sp = ps.discrete_spaces["e"]
    cells = mesh.entities[d].nodes[sp.topology.cell_index]
    # interior facets: pairs of elements sharing d vertices
    E = len(cells)
    checked = 0
    for e0 in range(E):
        for e1 in range(e0 + 1, E):
            common = np.intersect1d(cells[e0], cells[e1])
            if len(common) != d:
                continue
            P = mesh.points[common][:, :d]
            s, _ = make_quadrature(d - 1, 3)
            xs = P[0] + s @ (P[1:] - P[0])  # physical facet points
            if d == 2:
                tans = [P[1] - P[0]]
            else:
                tans = [P[1] - P[0], P[2] - P[0]]
            vals = []
            for e in (e0, e1):
                X = mesh.points[cells[e]][:, :d]
                J = (X[1:] - X[0]).T
                ref = np.linalg.solve(J, (xs - X[0]).T).T
                g, _ = _global_values(sp, mesh, d, e, ref)
                vals.append(np.stack([g @ t for t in tans]))
            assert np.abs(vals[0] - vals[1]).max() < 1e-12, (e0, e1)
            checked += 1
    assert checked > 0


@pytest.mark.parametrize("d", [2, 3])
def test_n1e_projection_reproduces_the_space(d):
    ps, mesh = _space(d, 2)
    sp = ps.discrete_spaces["e"]
    pts, w = make_quadrature(d, 4)
    a = np.array([0.3, -0.7, 0.2])[:d]
    b = np.array([0.5, 0.25, -0.4])

    def field(x):
        if d == 2:
            return a + b[2] * np.stack([-x[:, 1], x[:, 0]], axis=1)
        return a + np.cross(b, x)

    n = sp.n_dof
    M = np.zeros((n, n))
    rhs = np.zeros(n)
    per = []
    for e in range(len(sp.element_ids)):
        g, X = _global_values(sp, mesh, d, e, pts)
        J = (X[1:] - X[0]).T
        det = abs(np.linalg.det(J))
        x = X[0] + pts @ J.T
        f = field(x)
        M += np.einsum("q,qid,qjd->ij", w * det, g, g)
        rhs += np.einsum("q,qid,qd->i", w * det, g, f)
        per.append((g, f))
    c = np.linalg.solve(M, rhs)
    err = max(np.abs(np.einsum("qid,i->qd", g, c) - f).max() for g, f in per)
    assert err < 1e-12, err
