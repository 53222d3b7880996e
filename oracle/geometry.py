"""Affine simplex mapping, facet normals and basis push-forward (FP64 oracle).

- ``evaluate_mapping``: restates ``src/geometry/mapping.py:9-55``: x = P1(xi) . X,
  J = (3 x d) since mesh points are 3-D, per-point QR of J with the signs of R's
  diagonal normalised, ``detJ = det(R) > 0`` (|det J|, or the length/area measure
  of a facet), ``J+ = R^-1 Q^T``.
- ``normal``: restates ``src/geometry/compute_normal.py:7-24`` (unit, outward by
  the centroid difference).
- ``push_forward``: basix's contravariant Piola map ``J phi / detJ`` for RT/BDM
  and the identity for Lagrange, as used by ``FiniteElement.evaluate_basis``
  (``src/basis/finite_element.py:156-186``), including the zero third component
  inserted for 2-D vector tables.
- ``permute_and_transform``: restates ``src/basis/permute_and_transform.py:8-149``
  including the position -> (rotations, reflections) table and the forced
  reflection of tagged faces at local index 1 or 3.
"""

from __future__ import annotations

from itertools import permutations

import numpy as np

from .mesh import TET_EDGES, TET_FACES, TRI_EDGES


def p1_shapes(dim, points):
    pts = np.asarray(points, dtype=np.float64)
    nq = pts.shape[0]
    phi = np.zeros((1 + dim, nq, dim + 1))
    phi[0, :, 0] = 1.0 - pts.sum(axis=1)
    for d in range(dim):
        phi[0, :, 1 + d] = pts[:, d]
        phi[1 + d, :, 0] = -1.0
        phi[1 + d, :, 1 + d] = 1.0
    return phi


def _det_and_pinv(J):
    q, r = np.linalg.qr(J)
    s = np.diag(np.sign(np.diag(r)))
    q = q @ s
    r = s @ r
    return np.linalg.det(r), np.linalg.inv(r) @ q.T


def evaluate_mapping(dim, points, cell_points):
    phi = p1_shapes(dim, points)
    x = phi[0] @ cell_points
    jac = np.rollaxis(phi[1:dim + 1], 1) @ cell_points
    jac = np.transpose(jac, (0, 2, 1))
    res = [_det_and_pinv(j) for j in jac]
    det = np.array([r[0] for r in res])
    inv = np.array([r[1] for r in res])
    return x, jac, det, inv


def normal(points, cell_nodes, facet_nodes):
    xc1 = points[facet_nodes].mean(axis=0)
    xc0 = points[cell_nodes].mean(axis=0)
    outward = (xc1 - xc0) / np.linalg.norm(xc1 - xc0)
    if len(facet_nodes) == 2:
        v = points[facet_nodes[1]] - points[facet_nodes[0]]
        v = v / np.linalg.norm(v)
        w = np.cross(v, outward) / np.linalg.norm(np.cross(v, outward))
        n = np.cross(v, w)
    else:
        v = points[facet_nodes[0]] - xc1
        w = points[facet_nodes[1]] - xc1
        n = np.cross(v, w) / np.linalg.norm(np.cross(v, w))
    if n @ outward < 0.0:
        n = -n
    return n


def push_forward(tab, family, jac, det):
    """tab (1+dim, nq, ndofs, vs) -> pushed table; vector tables gain 3 components."""
    if family == "P":
        return tab.copy()
    dim = tab.shape[0] - 1
    mapped = np.einsum("qij,qnj->qni", jac, tab[0]) / det[:, None, None]
    if dim == 2:
        tab = np.insert(tab, 2, 0.0, axis=3)
    else:
        tab = tab.copy()
    tab[0] = mapped
    return tab


_POSITION_TO_ROT_REF = {0: (0, 0), 1: (0, 1), 2: (2, 1), 3: (1, 0), 4: (2, 0), 5: (1, 1)}


def face_orientation_3d(cell, bc_entities):
    """Per local face: (oriented, rotations, reflections), with the tagged-face rule."""
    out = []
    for i, f in enumerate(TET_FACES):
        face = tuple(cell.node_tags[list(f)].tolist())
        face_ref = tuple(sorted(face))
        oriented = face == face_ref
        rot, ref = 0, 0
        if not oriented:
            pos = [k for k, perm in enumerate(permutations(face_ref)) if perm == face]
            assert len(pos) == 1 and pos[0] != 0
            rot, ref = _POSITION_TO_ROT_REF[pos[0]]
        if bc_entities[i] and i in (1, 3):
            oriented = False
            ref = 1
        out.append((oriented, rot, ref))
    return out


def permute_and_transform(tab, element, cell, mesh, discontinuous=False):
    if element.transformations_are_identity or discontinuous and element.family == "P":
        return tab
    dim = cell.dimension
    T = element.entity_transformations()
    if dim == 2:
        for i, (a, b) in enumerate(TRI_EDGES):
            edge = cell.node_tags[[a, b]]
            stored = mesh.cells[cell.sub_cells_ids[1][i]].node_tags
            if np.all(edge == stored):
                continue
            dofs = element.entity_dofs[1][i]
            tab[:, :, dofs, :] = np.einsum("ab,dqbc->dqac", T["interval"][0], tab[:, :, dofs, :])
        return tab
    if dim == 3:
        bc_entities = [mesh.cells[fid].material_id is not None for fid in cell.sub_cells_ids[2].tolist()]
        rot_t, ref_t = T["triangle"]
        for i, (oriented, rot, ref) in enumerate(face_orientation_3d(cell, bc_entities)):
            if oriented:
                continue
            t = np.identity(rot_t.shape[0])
            for _ in range(rot):
                t = rot_t @ t
            for _ in range(ref):
                t = ref_t @ t
            dofs = element.entity_dofs[2][i]
            tab[:, :, dofs, :] = np.einsum("ab,dqbc->dqac", t, tab[:, :, dofs, :])
        for i, (a, b) in enumerate(TET_EDGES):
            edge = cell.node_tags[[a, b]]
            stored = np.sort(mesh.cells[cell.sub_cells_ids[1][i]].node_tags)
            if np.all(edge == stored):
                continue
            dofs = element.entity_dofs[1][i]
            if len(dofs):
                tab[:, :, dofs, :] = np.einsum("ab,dqbc->dqac", T["interval"][0], tab[:, :, dofs, :])
        return tab
    return tab
