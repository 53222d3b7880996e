"""L2 projection (FP64 oracle; test infrastructure only).

Restates ``l2_projector`` (``src/postprocess/projectors.py:8-91``) with
``L2ProjectorWeakForm.evaluate_form`` (``src/weak_forms/l2_projector_weak_form.py:8-93``),
and ``l2_error_projected`` (``src/postprocess/l2_error_post_processor.py:66-117``):

- per element and field (in the product space's order), with the pushed-forward basis
  of the weak forms: Lagrange ``j = sum_q detJ w psi psi^T`` and
  ``r[c::nc] -= (detJ w psi^T) @ f_c``; vector families ``j = sum_q sum_{d<dim} detJ w
  phi_d phi_d^T`` and ``r[c::nc] -= sum_q detJ w phi[:, :dim] @ f_c[:dim]``; fields are
  stacked block-diagonally;
- scatter: ``rg[dest] += r_el``; each matrix entry enters only if
  ``not isclose(v, 1e-16)`` and, for the symmetric solver, ``row <= col`` (SBAIJ upper
  part, mirrored here into the full symmetric matrix);
- ``A alpha = -rg`` solved directly (scipy ``spsolve`` in place of MUMPS Cholesky / LU).

Pinned by: exact reproduction of fields inside the space (a linear scalar in P1, cell
means of a linear scalar in P0, a constant vector in RT1, a linear vector in BDM1 and
RT2), the closed-form norm of a projected constant, and, for a field outside the space,
a least-squares fit of the discrete field to the function's point values at the
quadrature points (the same minimisation, computed without the mass matrix).
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sps
import scipy.sparse.linalg as spla

from .assemble import csr_from_triplets
from .forms import _cell, _tabulate
from .geometry import evaluate_mapping


def projector_element(ps, i, functions):
    """(r_el, j_el) of element i over all fields, as the reference's evaluate_form."""
    rs, js = [], []
    for name, sp in ps.spaces.items():
        f = functions[name]
        nc = sp.n_comp
        cell = _cell(sp, i)
        dim = cell.dimension
        points, weights = ps.quadrature[dim]
        x, jac, det, _ = evaluate_mapping(dim, points, sp.mesh.points[cell.node_tags])
        tab = _tabulate(sp, cell, points, jac, det)
        nphi = tab.shape[2]
        n = nphi * nc
        r = np.zeros(n)
        j = np.zeros((n, n))
        fv = np.asarray(f(x[:, 0], x[:, 1], x[:, 2]), dtype=np.float64)
        if sp.family == "P":
            phis = det * weights * tab[0, :, :, 0].T
            fv = fv.reshape(nc, -1)
            for c in range(nc):
                r[c:n:nc] -= phis @ fv[c]
            for c in range(nc):
                for q, w in enumerate(weights):
                    phi = tab[0, q, :, 0]
                    j[c:n:nc, c:n:nc] += det[q] * w * np.outer(phi, phi)
        else:
            fv = fv.reshape(nc, fv.shape[1], -1)
            for c in range(nc):
                for q, w in enumerate(weights):
                    phis = det[q] * w * tab[0, q, :, 0:dim].T
                    r[c:n:nc] -= phis.T @ fv[c, 0:dim, q]
            for c in range(nc):
                for q, w in enumerate(weights):
                    for d in range(dim):
                        j[c:n:nc, c:n:nc] += det[q] * w * np.outer(tab[0, q, :, d], tab[0, q, :, d])
        rs.append(r)
        js.append(j)
    n_tot = sum(len(r) for r in rs)
    r_el = np.zeros(n_tot)
    j_el = np.zeros((n_tot, n_tot))
    o = 0
    for r, j in zip(rs, js):
        r_el[o:o + len(r)] += r
        j_el[o:o + len(r), o:o + len(r)] += j
        o += len(r)
    return r_el, j_el


def assemble_projector(ps, functions, symmetric=True):
    """(A full symmetric CSR, rg) of the projection."""
    n = ps.n_dof
    rg = np.zeros(n)
    rows, cols, vals = [], [], []
    first = next(iter(ps.spaces.values()))
    for i in range(len(first.element_ids)):
        dest = np.asarray(ps.destination_indexes(i), dtype=np.int64)
        r_el, j_el = projector_element(ps, i, functions)
        rg[dest] += r_el
        data = j_el.ravel()
        row = np.repeat(dest, len(dest))
        col = np.tile(dest, len(dest))
        keep = np.logical_not(np.isclose(data, 1.0e-16))
        if symmetric:
            keep &= row <= col
        rows.append(row[keep])
        cols.append(col[keep])
        vals.append(data[keep])
    indptr, indices, data = csr_from_triplets(n, np.concatenate(rows), np.concatenate(cols), np.concatenate(vals))
    A = sps.csr_matrix((data, indices, indptr), shape=(n, n))
    if symmetric:
        A = (A + sps.triu(A, k=1).T).tocsr()
    return A, rg


def l2_projector(ps, functions, symmetric=True):
    A, rg = assemble_projector(ps, functions, symmetric)
    return spla.spsolve(A.tocsc(), -rg)


def l2_error_projected(ps, alpha, dim, skip_fields=()):
    out = []
    for name, sp in ps.spaces.items():
        if name in skip_fields:
            continue
        points, weights = ps.quadrature[dim]
        err = 0.0
        if sp.dimension == dim:
            for i in range(len(sp.element_ids)):
                cell = _cell(sp, i)
                x, jac, det, _ = evaluate_mapping(dim, points, sp.mesh.points[cell.node_tags])
                tab = _tabulate(sp, cell, points, jac, det)
                a = np.array(np.split(alpha[ps.destination_indexes(i, name)], tab.shape[2]))
                if sp.family == "P":
                    fh = (tab[0, :, :, 0] @ a[:, 0:dim]).T
                    err += np.sum(det * weights * fh * fh)
                else:
                    fh = np.array([np.vstack([tab[0, k, :, 0:dim].T @ a[:, c] for c in range(sp.n_comp)])
                                   for k in range(len(points))])
                    err += np.sum(det * weights * np.array([np.trace(e.T @ e) for e in fh]))
        out.append(np.sqrt(err))
    return out


def field_at_points(ps, name, alpha, i):
    """(x (nq, 3), f_h (nq, n_comp, dim) or (nq, n_comp)) of field name on element i."""
    sp = ps.spaces[name]
    cell = _cell(sp, i)
    dim = cell.dimension
    points, _ = ps.quadrature[dim]
    x, jac, det, _ = evaluate_mapping(dim, points, sp.mesh.points[cell.node_tags])
    tab = _tabulate(sp, cell, points, jac, det)
    a = np.array(np.split(alpha[ps.destination_indexes(i, name)], tab.shape[2]))
    if sp.family == "P":
        return x, tab[0, :, :, 0] @ a
    return x, np.einsum("qad,ac->qcd", tab[0, :, :, 0:dim], a)
