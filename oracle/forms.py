"""Element kernels of the hot path, per element, FP64 (oracle).

Each function restates one reference weak form for one element and returns
``(r_el, j_el)`` exactly as the reference's ``evaluate_form`` would:

- ``darcy_element``: ``LaplaceDualWeakForm.evaluate_form``
  (``src/weak_forms/laplace_dual_weak_form.py:11-87``). alpha is reset to zero
  (``:46``), so the AutoDiff Jacobian is the bilinear form:
  ``J_uu = sum w detJ kappa^-1 psi_i.psi_j``, ``J_up = -sum w detJ w_k div psi_i``,
  ``J_pu = +sum w detJ div psi_j w_k``, ``J_pp = 0``, ``r_p = -sum w detJ f w_k``.
- ``darcy_bc_facet``: ``LaplaceDualWeakFormBCDirichlet.evaluate_form`` (``:91-154``).
- ``elasticity_element``: ``LEDualWeakForm.evaluate_form_vectorized``
  (``src/weak_forms/le_dual_weak_form.py:354-510``).
- ``elasticity_bc_facet``: ``LEDualWeakFormBCDirichlet.evaluate_form`` (``:514-578``).
- ``riesz_element``: ``LERieszMapWeakForm.evaluate_form_vectorized``
  (``src/weak_forms/le_riesz_map_weak_form.py:258-360``).

The basis evaluation follows ``FiniteElement.evaluate_basis``: tabulate the
reference element, push the values forward (derivatives stay reference
derivatives), then ``permute_and_transform``.

Pinned by: hand-derived one-element matrices (closed forms in the tests), the
symmetry / sign structure of each block, the invariants ``sum_j J_pu[k, j]`` =
flux balance, and manufactured-solution convergence of the assembled systems.
"""

from __future__ import annotations

import numpy as np

from .geometry import evaluate_mapping, normal, permute_and_transform, push_forward
from .mesh import TET_FACES, TRI_EDGES


def coefficient(functions, key, i, x):
    """Per-quadrature-point values of a coefficient.

    ``functions[key]`` is either a callable of (x, y, z) as in the reference
    (vectorised over points, scalars broadcast) or a per-element array, in which
    case element ``i`` takes the constant ``value[i]``. Vector-valued callables
    return (n_comp, nq).
    """
    v = functions[key]
    nq = x.shape[0]
    if isinstance(v, np.ndarray):
        e = np.asarray(v[i], dtype=np.float64)
        if e.ndim == 0:
            return np.full(nq, float(e))
        return np.repeat(e.reshape(-1, 1), nq, axis=1)
    out = np.asarray(v(x[:, 0], x[:, 1], x[:, 2]), dtype=np.float64)
    if out.ndim == 0:
        return np.full(nq, float(out))
    if out.ndim == 1 and out.shape[0] == nq:
        return out
    out = out.reshape(out.shape[0], -1)
    if out.shape[1] == 1 and nq > 1:
        out = np.repeat(out, nq, axis=1)
    return out


def _tabulate(space, cell, points, jac, det):
    el = space.element
    tab = el.tabulate(points)
    tab = push_forward(tab, el.family, jac, det)
    return permute_and_transform(tab, el, cell, space.mesh, space.discontinuous)


def _cell(space, i):
    return space.mesh.cells[space.element_ids[i]]


def darcy_element(ps, i, functions):
    u_sp, p_sp = ps.spaces["u"], ps.spaces["p"]
    cell = _cell(u_sp, i)
    dim = cell.dimension
    points, weights = ps.quadrature[dim]
    x, jac, det, _ = evaluate_mapping(dim, points, u_sp.mesh.points[cell.node_tags])
    u_tab = _tabulate(u_sp, cell, points, jac, det)
    p_tab = _tabulate(p_sp, cell, points, jac, det)
    nu = u_tab.shape[2] * u_sp.n_comp
    npp = p_tab.shape[2] * p_sp.n_comp
    n = nu + npp
    r = np.zeros(n)
    J = np.zeros((n, n))
    f = coefficient(functions, "rhs", i, x).reshape(-1, len(weights))
    kinv_q = 1.0 / coefficient(functions, "kappa", i, x).reshape(-1)
    phi_s = det * weights * p_tab[0, :, :, 0].T
    for c in range(p_sp.n_comp):
        r[nu + c:n:p_sp.n_comp] -= phi_s @ f[c]
    for q, w in enumerate(weights):
        psi = u_tab[0, q, :, 0:dim]
        wp = p_tab[0, q, :, 0]
        div = np.trace(u_tab[1:, q, :, 0:dim], axis1=0, axis2=2) / det[q]
        kinv = kinv_q[q]
        dw = det[q] * w
        J[:nu, :nu] += dw * kinv * (psi @ psi.T)
        J[:nu, nu:] -= dw * np.outer(div, wp)
        J[nu:, :nu] += dw * np.outer(wp, div)
    return r, J


def _facet_trace_points(facet, neigh, points):
    """``transform_lower_to_higher`` (``src/basis/parametric_transformation.py:30-98``)."""
    dim = neigh.dimension
    local = TRI_EDGES if dim == 2 else TET_FACES
    fidx = neigh.sub_cells_ids[dim - 1].tolist().index(facet.id)
    loc = list(local[fidx])
    sub = neigh.node_tags[loc]
    if tuple(facet.node_tags.tolist()) != tuple(sub.tolist()):
        index_map = dict(zip(sub.tolist(), loc))
        loc = [index_map[k] for k in facet.node_tags.tolist()]
    verts = np.eye(dim + 1, dim, -1)  # reference vertices: 0, e_0, e_1, ...
    if dim == 2:
        return np.array([verts[loc[0]] * (1 - xi) + verts[loc[1]] * xi for (xi,) in points]), fidx
    return np.array([verts[loc[0]] * (1 - a - b) + verts[loc[1]] * a + verts[loc[2]] * b for a, b in points]), fidx


def _bc_trace(ps, field, j):
    sp = ps.spaces[field]
    fid = sp.bc_element_ids[j]
    facet = sp.mesh.cells[fid]
    D = sp.dimension
    points, weights = ps.bc_quadrature[D - 1]
    x, _, det, _ = evaluate_mapping(D - 1, points, sp.mesh.points[facet.node_tags])
    neigh = sp.mesh.cells[sp.bc_neighbour(fid)]
    mpts, fidx = _facet_trace_points(facet, neigh, points)
    _, jac0, det0, _ = evaluate_mapping(D, mpts, sp.mesh.points[neigh.node_tags])
    tab = _tabulate(sp, neigh, mpts, jac0, det0)
    dofs = sp.element.entity_dofs[D - 1][fidx]
    n = normal(sp.mesh.points, neigh.node_tags, facet.node_tags)
    return x, det, weights, tab, dofs, n, sp


def darcy_bc_facet(ps, j, functions):
    x, det, weights, tab, dofs, n, sp = _bc_trace(ps, "u", j)
    D = sp.dimension
    nphi = len(dofs)
    r = np.zeros(nphi * sp.n_comp)
    for c in range(sp.n_comp):
        blk = np.zeros(nphi)
        for q, w in enumerate(weights):
            pD = np.asarray(functions["p"](x[q, 0], x[q, 1], x[q, 2]), dtype=np.float64).reshape(-1)
            phi = tab[0, q, dofs, 0:D] @ n[0:D]
            blk += det[q] * w * pD[c] * phi
        r[c::sp.n_comp] += blk
    return r, np.zeros((len(r), len(r)))


def elasticity_bc_facet(ps, j, functions):
    x, det, weights, tab, dofs, n, sp = _bc_trace(ps, "s", j)
    D = sp.dimension
    nphi = len(dofs)
    r = np.zeros(nphi * sp.n_comp)
    for c in range(sp.n_comp):
        blk = np.zeros(nphi)
        for q, w in enumerate(weights):
            uD = np.asarray(functions["u"](x[q, 0], x[q, 1], x[q, 2])).reshape(-1)
            phi = tab[0, q, dofs, 0:D] @ n[0:D]
            blk -= det[q] * w * uD[c] * phi
        r[c::sp.n_comp] += blk
    return r, np.zeros((len(r), len(r)))


def _elasticity_common(ps, i, functions):
    s_sp, u_sp, t_sp = ps.spaces["s"], ps.spaces["u"], ps.spaces["t"]
    cell = _cell(s_sp, i)
    dim = cell.dimension
    points, weights = ps.quadrature[dim]
    x, jac, det, _ = evaluate_mapping(dim, points, s_sp.mesh.points[cell.node_tags])
    s_tab = _tabulate(s_sp, cell, points, jac, det)
    u_tab = _tabulate(u_sp, cell, points, jac, det)
    t_tab = _tabulate(t_sp, cell, points, jac, det)
    lam = coefficient(functions, "lambda", i, x).reshape(-1)
    mu = coefficient(functions, "mu", i, x).reshape(-1)
    return s_sp, u_sp, t_sp, dim, points, weights, x, det, s_tab, u_tab, t_tab, lam, mu


def volumetric_factor(lam, mu, dim):
    """``(1 / 2 mu) (lambda / (2 mu + dim lambda))`` (``le_dual_weak_form.py:434``): the
    trace part of the compliance ``A sigma = (sigma - lambda / (2 mu + dim lambda) tr(sigma) I) / 2 mu``."""
    return (1.0 / (2.0 * mu)) * (lam / (2.0 * mu + dim * lam))


# rotation component c -> [(stress row, direction d, sign)]: the (t, s) block entry is
# -sign * int t_k phi_{j,d} on stress row cs (``le_dual_weak_form.py:473-491``); rotation c
# pairs with tau_{r1 r2} - tau_{r2 r1} for (r1, r2) = (y, z), (z, x), (x, y) in 3-D and
# (x, y) in 2-D
SKEW_MAP = {
    3: {0: [(2, 1, 1.0), (1, 2, -1.0)], 1: [(0, 2, 1.0), (2, 0, -1.0)], 2: [(1, 0, 1.0), (0, 1, -1.0)]},
    2: {0: [(1, 0, 1.0), (0, 1, -1.0)]},
}


def _ss_blocks(J, s_tab, dim, sc, weights, det, lam, mu):
    ns = s_tab.shape[2] * sc
    phi = s_tab[0, :, :, 0:dim]
    ss = np.einsum("qic,qjc,q->ij", phi, phi, (1.0 / (2.0 * mu)) * det * weights)
    for c in range(sc):
        J[c:ns:sc, c:ns:sc] += ss
    vol = volumetric_factor(lam, mu, dim)
    flat = phi.reshape(phi.shape[0], -1)
    J[:ns, :ns] -= np.einsum("qa,qb,q->ab", flat, flat, vol * det * weights)
    return ns


def elasticity_element(ps, i, functions):
    s_sp, u_sp, t_sp, dim, points, weights, x, det, s_tab, u_tab, t_tab, lam, mu = _elasticity_common(ps, i, functions)
    sc, uc, tc = s_sp.n_comp, u_sp.n_comp, t_sp.n_comp
    ns = s_tab.shape[2] * sc
    nu = u_tab.shape[2] * uc
    nt = t_tab.shape[2] * tc
    n = ns + nu + nt
    J = np.zeros((n, n))
    r = np.zeros(n)
    f = coefficient(functions, "rhs", i, x).reshape(-1, len(weights))
    u_s = det * weights * u_tab[0, :, :, 0].T
    t_s = det * weights * t_tab[0, :, :, 0].T
    for c in range(uc):
        r[ns + c:ns + nu:uc] += -(u_s @ f[c]).ravel()
    for c in range(tc):
        r[ns + nu + c:n:tc] += -(t_s @ f[c + uc]).ravel()
    _ss_blocks(J, s_tab, dim, sc, weights, det, lam, mu)
    phi = s_tab[0, :, :, 0:dim]
    # A_d[k, j] = sum_q detJ w t_k(q) phi_{j,d}(q)
    A = [np.einsum("qk,qj,q->kj", t_tab[0, :, :, 0], phi[:, :, d], det * weights) for d in range(dim)]
    for c in range(tc):
        b = ns + nu + c
        for cs, d, sgn in SKEW_MAP[dim][c]:
            J[b:n:tc, cs:ns:sc] += -sgn * A[d]
            J[cs:ns:sc, b:n:tc] += -sgn * A[d].T
    div = np.trace(s_tab[1:, :, :, 0:dim], axis1=0, axis2=3)  # (nq, nphi)
    ub = (weights * u_tab[0, :, :, 0].T) @ div
    for c in range(min(uc, sc)):
        J[ns + c:ns + nu:uc, c:ns:sc] += ub
        J[c:ns:sc, ns + c:ns + nu:uc] += ub.T
    return r, J


def riesz_element(ps, i, functions):
    s_sp, u_sp, t_sp, dim, points, weights, x, det, s_tab, u_tab, t_tab, lam, mu = _elasticity_common(ps, i, functions)
    sc, uc, tc = s_sp.n_comp, u_sp.n_comp, t_sp.n_comp
    ns = s_tab.shape[2] * sc
    nu = u_tab.shape[2] * uc
    nt = t_tab.shape[2] * tc
    n = ns + nu + nt
    J = np.zeros((n, n))
    _ss_blocks(J, s_tab, dim, sc, weights, det, lam, mu)
    div = np.trace(s_tab[1:, :, :, 0:dim], axis1=0, axis2=3) / det[:, None]
    dd = np.einsum("qi,qj,q->ij", div, div, det * weights)
    for c in range(sc):
        J[c:ns:sc, c:ns:sc] += dd
    uu = np.einsum("qi,qj,q->ij", u_tab[0, :, :, 0], u_tab[0, :, :, 0], det * weights)
    for c in range(uc):
        J[ns + c:ns + nu:uc, ns + c:ns + nu:uc] += uu
    tt = np.einsum("qi,qj,q->ij", t_tab[0, :, :, 0], t_tab[0, :, :, 0], det * weights)
    for c in range(tc):
        J[ns + nu + c:n:tc, ns + nu + c:n:tc] += tt
    return np.zeros(n), J
