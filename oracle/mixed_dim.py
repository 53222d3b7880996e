"""FP64 oracle of the mixed-dimensional transient Darcy system (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU baseline may import
this module; the product path never does.

Restates, per element and per fracture triangle, with the oracle's own spaces:
- the matrix and fracture mixed Darcy forms and the matrix Dirichlet facet term
  (``LaplaceDualWeakForm`` / ``LaplaceDualWeakFormBCDirichlet`` via ``forms.py``) and their
  scatter into one system (``examples/md_elliptic_scalar/md_elliptic_scalar.py:250-343``);
- the cut of the matrix along the fractures (``src/mesh/mesh_operations.py:274-411``) as
  a new facet DoF for the element on the negative side of each fracture face;
- the interface coupling form (``examples/md_elliptic_scalar/CouplingWeakForm.py:11-151``):
  for side s with facet DoF trace s_q = phi . n_s (n_s out of that side's element) and the
  fracture pressure basis psi = 1 (P0),
    (u_s, u_s) += sum_q w |det J_F| (1/alpha) s_q^2,   (u_s, p) += sum_q w |det J_F| s_q,
    (p, u_s) += sum_q w |det J_F| s_q,                  alpha = kappa_n / (delta / 2);
  (``exchange_sign = -1``: the (p, u_s) entries negated, the conservative exchange);
- the backward-Euler storage term of ``examples/parabolic_diffusive_problem/
  SundusWeakForm.py:148-150`` on both pressures (fracture: times the aperture);
- the elimination of the fracture-tip flux DoFs (no flow through the tips).

Global DoF order: [matrix flux (original, then the split DoFs in fracture-face order) |
fracture flux | matrix pressure | fracture pressure].
"""

from __future__ import annotations

import numpy as np

from .assemble import TripletCollector
from .forms import _facet_trace_points, _tabulate, darcy_bc_facet, darcy_element
from .geometry import evaluate_mapping, normal
from .mesh import Block, ConformalMesh, structured_cube
from .spaces import ProductSpace


def matrix_space(n):
    pts, blocks = structured_cube(n)
    mesh = ConformalMesh(3, pts, blocks)
    ps = ProductSpace({"u": (3, 1, "RT", 1, mesh), "p": (3, 1, "Lagrange", 0, mesh)}, {"p": True})
    ps.build_structures({"u": [1], "p": [1]}, {"u": [2, 3, 4, 5, 6, 7]})
    return ps, mesh


def _plane_frame(n):
    a = np.array([0.0, 0.0, 1.0]) if abs(n[2]) < 0.9 else np.array([1.0, 0.0, 0.0])
    t1 = a - n * (a @ n)
    t1 /= np.linalg.norm(t1)
    return t1, np.cross(n, t1)


def fracture_mesh(mesh, face_ids, normals=None):
    """The fracture triangles (matrix faces, in this order) as a 2-D mesh: vertices in
    ascending matrix id, tips (edges of one triangle) tagged 2, triangles tagged 1.

    The reference's 2-D forms use the x, y components of the mapped basis
    (``src/weak_forms/laplace_dual_weak_form.py:62``: ``tab[..., 0:dim]``), i.e. they hold for
    meshes in the plane z = 0. With ``normals`` (unit normal of each face's plane) every
    vertex is placed at its in-plane coordinates (t1 . x, t2 . x, 0) in its fracture's
    frame: an isometric copy on which the reference form gives the element's intrinsic
    matrices (fractures share no vertex, so each vertex has one frame)."""
    tri = np.array([mesh.cells[f].node_tags for f in face_ids], dtype=np.int64)
    vid, tri_local = np.unique(tri.reshape(-1), return_inverse=True)
    tri_local = tri_local.reshape(-1, 3)
    edges = np.sort(np.stack([tri_local[:, [0, 1]], tri_local[:, [1, 2]], tri_local[:, [0, 2]]], 1).reshape(-1, 2),
                    axis=1)
    ue, cnt = np.unique(edges, axis=0, return_counts=True)
    tips = ue[cnt == 1]
    blocks = [Block(1, tips, np.full(len(tips), 2)), Block(2, tri_local, np.full(len(tri_local), 1))]
    pts = mesh.points[vid].copy()
    if normals is not None:
        frame = np.zeros((len(vid), 3, 3))
        for t, nv in zip(tri_local, normals):
            t1, t2 = _plane_frame(np.asarray(nv, dtype=np.float64))
            frame[t] = np.stack([t1, t2, np.zeros(3)])
        pts = np.einsum("vij,vj->vi", frame, mesh.points[vid])
    fmesh = ConformalMesh(2, pts, blocks)
    return fmesh, vid


def _face_cells(mesh):
    out = {}
    for c in mesh.cells:
        if c is not None and c.dimension == 3:
            for f in c.sub_cells_ids[2].tolist():
                out.setdefault(f, []).append(c.id)
    return out


def _side_cells(mesh, face_id, n_unit, face_cells):
    """(positive-side cell id, negative-side cell id) of an interior face."""
    cells = face_cells[face_id]
    assert len(cells) == 2
    fc = mesh.points[mesh.cells[face_id].node_tags].mean(axis=0)
    c0 = mesh.points[mesh.cells[cells[0]].node_tags].mean(axis=0)
    return (cells[0], cells[1]) if (c0 - fc) @ n_unit > 0 else (cells[1], cells[0])


def coupling_traces(ps, mesh, face_id, cell_id):
    """(facet DoF index within the cell's flux DoFs, w |det J_F| per point, s_q per point)
    of the cell's facet DoF on face `face_id` (CouplingWeakForm.compute_trace_space)."""
    sp = ps.spaces["u"]
    facet = mesh.cells[face_id]
    neigh = mesh.cells[cell_id]
    points, weights = ps.bc_quadrature[2]
    _, _, det, _ = evaluate_mapping(2, points, mesh.points[facet.node_tags])
    mpts, fidx = _facet_trace_points(facet, neigh, points)
    _, jac0, det0, _ = evaluate_mapping(3, mpts, mesh.points[neigh.node_tags])
    tab = _tabulate(sp, neigh, mpts, jac0, det0)
    dofs = sp.element.entity_dofs[2][fidx]
    assert len(dofs) == 1
    nrm = normal(mesh.points, neigh.node_tags, facet.node_tags)
    s = np.array([tab[0, q, dofs[0], 0:3] @ nrm[0:3] for q in range(len(weights))])
    return int(dofs[0]), det * weights, s


def assemble_md(n, face_ids, normals, apertures, kappa_normal=1.0, storage=1.0, dt=1e-3, kappa_matrix=1.0,
                p_boundary=None, exchange_sign=1):
    """(A (CSR), b0, mass, n_u, sizes) of the coupled system; b0 the static right-hand side
    (-rg), mass the storage coefficients of the pressure rows (global pressure order)."""
    import scipy.sparse as sps

    ps, mesh = matrix_space(n)
    sp = ps.spaces["u"]
    m_n, m_nu = ps.n_dof, sp.n_dof
    face_ids = list(face_ids)
    fc = _face_cells(mesh)
    sides = [_side_cells(mesh, f, nv, fc) for f, nv in zip(face_ids, normals)]
    # split DoFs: the negative-side cell's facet DoF of every fracture face
    n_split = len(face_ids)
    split = {}
    for i, (f, (cp, cn)) in enumerate(zip(face_ids, sides)):
        e = sp.id_to_element[cn]
        loc, _, _ = coupling_traces(ps, mesh, f, cn)
        split[(e, loc)] = m_nu + i
    m_nu2 = m_nu + n_split
    m_n2 = m_n + n_split
    fmesh, _ = fracture_mesh(mesh, face_ids, normals)
    fps = ProductSpace({"u": (2, 1, "RT", 1, fmesh), "p": (2, 1, "Lagrange", 0, fmesh)}, {"p": True})
    fps.build_structures({"u": [1], "p": [1]}, {"u": [2]})
    f_n, f_nu = fps.n_dof, fps.spaces["u"].n_dof
    N = m_n2 + f_n
    n_u = m_nu2 + f_nu
    m_np = m_n2 - m_nu2

    def gm(i):  # matrix DoF after the split -> global
        i = np.asarray(i, np.int64)
        return np.where(i < m_nu2, i, i + f_nu)

    def gf(i):
        i = np.asarray(i, np.int64)
        return np.where(i < f_nu, m_nu2 + i, m_nu2 + m_np + i)

    def mdest(e):
        d = ps.destination_indexes(e).copy()
        d[d >= m_nu] += n_split
        for a in range(sp.element.ndofs):
            if (e, a) in split:
                d[a] = split[(e, a)]
        return gm(d)

    coll = TripletCollector(N)
    rg = np.zeros(N)
    pb = p_boundary if p_boundary is not None else (lambda x, y, z: 1.0 - x)
    fun_m = {"kappa": np.full(len(sp.element_ids), kappa_matrix), "rhs": np.zeros(len(sp.element_ids))}
    for e in range(len(sp.element_ids)):
        dest = mdest(e)
        r, J = darcy_element(ps, e, fun_m)
        rg[dest] += r
        coll.add_element(dest, J)
    for j in range(len(sp.bc_element_ids)):
        d = ps.bc_destination_indexes(j)
        d[d >= m_nu] += n_split
        dest = gm(d)
        r, J = darcy_bc_facet(ps, j, {"p": pb})
        rg[dest] += r
        coll.add_element(dest, J)
    fsp = fps.spaces["u"]
    delta_e = np.asarray(apertures, dtype=np.float64)
    fun_f = {"kappa": delta_e ** 3 / 12.0, "rhs": np.zeros(len(fsp.element_ids))}
    for i in range(len(fsp.element_ids)):
        dest = gf(fps.destination_indexes(i))
        r, J = darcy_element(fps, i, fun_f)
        rg[dest] += r
        coll.add_element(dest, J)
    # interface coupling, triangle by triangle (element i of the fracture space is triangle i)
    for i, (f, (cp, cn)) in enumerate(zip(face_ids, sides)):
        pdof = gf(fps.destination_indexes(i, "p"))[0]
        inv_alpha = delta_e[i] / (2.0 * kappa_normal)
        for cell in (cp, cn):
            e = sp.id_to_element[cell]
            loc, wd, s = coupling_traces(ps, mesh, f, cell)
            u = mdest(e)[loc]
            uu = float(np.sum(wd * inv_alpha * s * s))
            up = float(np.sum(wd * s))
            coll.add_element(np.array([u, pdof]), np.array([[uu, up], [exchange_sign * up, 0.0]]))
    A = coll.csr().tocsr()
    pattern = (A.indptr.copy(), A.indices.copy())
    # storage terms
    mass = np.zeros(N - n_u)
    for e in range(len(sp.element_ids)):
        X = mesh.points[mesh.cells[sp.element_ids[e]].node_tags]
        vol = abs(np.linalg.det(X[1:] - X[0])) / 6.0
        pd = mdest(e)[sp.element.ndofs:]
        mass[pd[0] - n_u] = storage * vol / dt
    for i in range(len(fsp.element_ids)):
        X = fmesh.points[fmesh.cells[fsp.element_ids[i]].node_tags]
        area = 0.5 * np.linalg.norm(np.cross(X[1] - X[0], X[2] - X[0]))
        pd = gf(fps.destination_indexes(i, "p"))[0]
        mass[pd - n_u] = storage * delta_e[i] * area / dt
    A = A + sps.diags(np.concatenate([np.zeros(n_u), mass]), format="csr")
    # tips: u = 0 (row -> identity, column -> 0)
    tips = set()
    for j in range(len(fsp.bc_element_ids)):
        for d in gf(fps.bc_destination_indexes(j, "u")):
            tips.add(int(d))
    tips = np.array(sorted(tips), dtype=np.int64)
    A = A.tolil()
    for t in tips:
        A[t, :] = 0.0
        A[:, t] = 0.0
        A[t, t] = 1.0
    A = A.tocsr()
    b0 = -rg
    b0[tips] = 0.0
    sizes = {"N": N, "n_u": n_u, "matrix_flux": m_nu2, "fracture_flux": f_nu, "matrix_pressure": m_np,
             "fracture_pressure": f_n - f_nu, "split": n_split, "tips": tips, "pattern": pattern}
    return A, b0, mass, n_u, sizes


def step(A, b0, mass, n_u, x_prev):
    """One backward-Euler step by a direct solve: A x = b0 + mass x_prev (pressure rows)."""
    import scipy.sparse.linalg as spla

    b = b0.copy()
    b[n_u:] += mass * x_prev[n_u:]
    return spla.spsolve(A.tocsc(), b)
