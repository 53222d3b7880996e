"""Mixed-dimensional transient Darcy flow: a 3-D RT1-P0 matrix cut by planar fractures.

The coupled problem of ``examples/md_elliptic_scalar/md_elliptic_scalar.py:136-343`` (matrix
and fracture mixed Darcy forms, the interface coupling form of
``examples/md_elliptic_scalar/CouplingWeakForm.py:11-151``, Dirichlet facet terms) with the
backward-Euler storage term of ``examples/parabolic_diffusive_problem/SundusWeakForm.py:11-197``
and the time loop of ``examples/parabolic_diffusive_problem/mixed_transport_model.py:254-350``,
in 3-D:

- fractures are planar patches made of faces of the tetrahedral mesh (the Kuhn cube's faces
  lie on the planes x_i = c and x_i - x_j = c, so six plane families are available); they are
  drawn at random (family, position, extent along both in-plane directions, log-normal
  aperture) and rejected when they touch the domain boundary or another fracture;
- the cut that the reference makes in the mesh (``cut_conformity_along_c1_lines``,
  ``src/mesh/mesh_operations.py:274-411``) is made on the DoF map: the face DoF of a fracture
  face gets a second DoF for the element on the negative side of the fracture plane
  (``DarcySystem(split=...)``), so the normal flux may jump across the fracture;
- the fractures carry their own RT1-P0 system on their triangles (2-D elements in 3-D),
  permeability delta^2/12 times the aperture delta (``md_elliptic_scalar.py:206-207``);
  the fracture tips carry no flow (their flux DoFs are eliminated);
- each fracture triangle couples the two matrix face fluxes to its pressure through the
  Robin terms of the coupling form, alpha = kappa_n / (delta / 2); as written there, the
  fracture mass row adds the matrix outflow with the sign of the fracture divergence
  (``conservative=False``); the default flips it, making the outflow a source of the
  fracture, because with the reference's sign the backward-Euler loop is unstable (the
  state grows without bound although every step's solve converges);
- storage: the pressure rows carry (phi/dt) |K| p (matrix) and (phi/dt) delta |T| p
  (fracture); the matrix, assembled once on the device, stays fixed while dt does, and a
  step re-forms only the right-hand side from the previous pressure.

Global DoF order: [matrix flux (with the split DoFs) | fracture flux | matrix pressure |
fracture pressure], so that the flux block comes first as the block preconditioner expects.
"""

from __future__ import annotations

import numpy as np

from . import _native as nat
from .device import CoefficientArg, DarcySystem, torch, to_dev
from .mesh import Block, Mesh, unit_cube
from .spaces import ProductSpace

FAMILIES = np.array([(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, -1, 0), (0, 1, -1), (1, 0, -1)], dtype=np.float64)


def _tangents(nv):
    n = nv / np.linalg.norm(nv)
    a = np.array([0.0, 0.0, 1.0]) if abs(n[2]) < 0.9 else np.array([1.0, 0.0, 0.0])
    t1 = a - n * (a @ n)
    t1 /= np.linalg.norm(t1)
    t2 = np.cross(n, t1)
    return t1, t2


def face_incidence(mesh):
    """(F, 2, 2) (tet index, local face) of every face, -1 where absent."""
    t3 = mesh.entities[3]
    foff = mesh.entities[2].offset
    tf = t3.subs[2] - foff  # (T, 4) face positions
    F = mesh.entities[2].n
    inc = np.full((F, 2, 2), -1, np.int64)
    flat_f = tf.reshape(-1)
    order = np.argsort(flat_f, kind="stable")
    fs = flat_f[order]
    tets = order // 4
    loc = order % 4
    first = np.ones(len(fs), bool)
    first[1:] = fs[1:] != fs[:-1]
    slot = np.where(first, 0, 1)
    inc[fs, slot, 0] = tets
    inc[fs, slot, 1] = loc
    return inc


def sample_fractures(mesh, count=8, seed=0, extent=(0.3, 0.6), mean_log_aperture=-4.0, sigma=0.5,
                     max_tries=5000):
    """Planar fracture patches made of mesh faces, pairwise disjoint and interior.

    Returns a list of dicts: family, normal (unit), offset, faces (face positions), aperture."""
    rng = np.random.default_rng(seed)
    pts = mesh.points
    fnodes = mesh.entities[2].nodes
    inc = face_incidence(mesh)
    interior = inc[:, 1, 0] >= 0
    lo, hi = pts.min(axis=0), pts.max(axis=0)
    h = np.min(np.diff(np.unique(pts[:, 0])))
    cent = pts[fnodes].mean(axis=1)
    used = np.zeros(len(pts), bool)
    out = []
    tries = 0
    while len(out) < count and tries < max_tries:
        tries += 1
        fam = int(rng.integers(len(FAMILIES)))
        nv = FAMILIES[fam]
        proj = pts @ nv
        levels = np.unique(np.round(proj / h).astype(np.int64))
        k = int(rng.choice(levels[1:-1]))
        c = k * h
        on = np.abs(pts[fnodes] @ nv - c).max(axis=1) < 1e-9 * max(1.0, abs(c)) + 1e-12
        cand = np.flatnonzero(on & interior)
        if len(cand) == 0:
            continue
        n_unit = nv / np.linalg.norm(nv)
        t1, t2 = _tangents(nv)
        L1, L2 = rng.uniform(extent[0], extent[1], 2)
        centre = cent[cand[int(rng.integers(len(cand)))]]
        d1 = (cent[cand] - centre) @ t1
        d2 = (cent[cand] - centre) @ t2
        sel = cand[(np.abs(d1) <= L1 / 2) & (np.abs(d2) <= L2 / 2)]
        if len(sel) < 4:
            continue
        v = np.unique(fnodes[sel])
        pv = pts[v]
        if np.any(pv <= lo + 0.5 * h) or np.any(pv >= hi - 0.5 * h) or np.any(used[v]):
            continue
        used[v] = True
        out.append({"family": fam, "normal": n_unit, "offset": c / np.linalg.norm(nv), "faces": sel,
                    "aperture": float(np.exp(rng.normal(mean_log_aperture, sigma)))})
    if len(out) < count:
        raise RuntimeError(f"placed {len(out)} of {count} fractures after {max_tries} tries")
    return out


class FractureNetwork:
    """Host-side geometry and DoF data of the cut matrix and its fractures."""

    def __init__(self, n, count=8, seed=0, extent=(0.3, 0.6), mean_log_aperture=-4.0, sigma=0.5,
                 kappa_normal=None):
        self.n = n
        self.mesh = Mesh.from_blocks(3, *unit_cube(n))
        mesh = self.mesh
        self.fractures = sample_fractures(mesh, count, seed, extent, mean_log_aperture, sigma)
        inc = face_incidence(mesh)
        pts = mesh.points
        fnodes = mesh.entities[2].nodes
        T = mesh.entities[3]
        tcent = pts[T.subs[0]].mean(axis=1)
        faces, frac_of, pos_side, neg_side = [], [], [], []
        for i, fr in enumerate(self.fractures):
            for f in fr["faces"]:
                (ta, la), (tb, lb) = inc[f]
                sa = (tcent[ta] - pts[fnodes[f]].mean(axis=0)) @ fr["normal"]
                if sa > 0:
                    pos_side.append((ta, la))
                    neg_side.append((tb, lb))
                else:
                    pos_side.append((tb, lb))
                    neg_side.append((ta, la))
                faces.append(f)
                frac_of.append(i)
        self.faces = np.asarray(faces, np.int64)  # face positions, one per fracture triangle
        self.frac_of = np.asarray(frac_of, np.int64)
        self.pos_side = np.asarray(pos_side, np.int64)  # (m, 2) tet index, local face
        self.neg_side = np.asarray(neg_side, np.int64)
        self.aperture = np.array([fr["aperture"] for fr in self.fractures])[self.frac_of]
        # fracture 2-D mesh: its triangles are the fracture faces (in this order)
        tri = fnodes[self.faces]
        vid, tri_local = np.unique(tri.reshape(-1), return_inverse=True)
        tri_local = tri_local.reshape(-1, 3)
        self.frac_vertices = vid
        edges = np.sort(np.stack([tri_local[:, [0, 1]], tri_local[:, [1, 2]], tri_local[:, [0, 2]]], 1).reshape(-1, 2),
                        axis=1)
        ue, cnt = np.unique(edges, axis=0, return_counts=True)
        tips = ue[cnt == 1]
        self.fmesh = Mesh.from_blocks(2, pts[vid], [Block.make(1, tips, 2), Block.make(2, tri_local, 1)])
        self.tip_edges = tips


def matrix_space(mesh):
    ps = ProductSpace({"u": (3, 1, "RT", 1, mesh), "p": (3, 1, "Lagrange", 0, mesh)})
    ps.make_subspaces_discontinuous({"p": True})
    ps.build_structures({"u": [1], "p": [1]}, {"u": [2, 3, 4, 5, 6, 7]})
    return ps


def fracture_space(fmesh):
    ps = ProductSpace({"u": (2, 1, "RT", 1, fmesh), "p": (2, 1, "Lagrange", 0, fmesh)})
    ps.make_subspaces_discontinuous({"p": True})
    ps.build_structures({"u": [1], "p": [1]}, {"u": [2]})
    return ps


def _cell_measures(points, cells):
    X = points[cells]
    if cells.shape[1] == 4:
        J = X[:, 1:] - X[:, :1]
        return np.abs(np.linalg.det(J)) / 6.0
    return 0.5 * np.linalg.norm(np.cross(X[:, 1] - X[:, 0], X[:, 2] - X[:, 0]), axis=1)


class MixedDimDarcy:
    """The coupled matrix + fracture system on the device and its backward-Euler loop.

    Assembled once (FP64 and FP32 values of one CSR system): the matrix and fracture
    block systems by the row kernels, placed into the coupled CSR, the interface
    coupling (``nfx_md_coupling``), the storage terms on the pressure diagonal and the
    elimination of the fracture-tip flux DoFs. A step forms b = b0 + (phi/dt) M p_prev
    and solves with FGMRES + FP64 refinement, warm-started from the previous state, with
    the diagonal block preconditioner (``nfx_md_block_diag``), set up once."""

    def __init__(self, net, dt=1e-3, storage=1.0, kappa_matrix=1.0, kappa_normal=1.0, p_boundary=None,
                 device="cuda", conservative=True, flux_sweeps=True):
        from .device import p1_bary  # noqa: F401
        from .pattern import row_codes

        t = torch()
        self.net, self.dt, self.storage, self.device = net, float(dt), float(storage), device
        # the two-grid preconditioner's flux rows: two damped Jacobi sweeps on the flux mass
        # block (True) or its diagonal
        self.flux_sweeps = bool(flux_sweeps)
        # +1: the reference's coupling form; -1: the matrix outflow is a fracture source
        self.exchange_sign = -1 if conservative else 1
        mesh = net.mesh
        ps_m = matrix_space(mesh)
        u_m = ps_m.discrete_spaces["u"]
        e_of_tet = np.full(mesh.entities[3].n, -1, np.int64)
        e_of_tet[u_m.topology.cell_index] = np.arange(len(u_m.topology.cell_index))
        pos_e, neg_e = e_of_tet[net.pos_side[:, 0]], e_of_tet[net.neg_side[:, 0]]
        if np.any(pos_e < 0) or np.any(neg_e < 0):
            raise ValueError("fracture face next to an untagged cell")
        el = u_m.element
        face_dof = np.array([el.entity_dofs[2][a][0] for a in range(4)], np.int64)
        dof_p, dof_n = face_dof[net.pos_side[:, 1]], face_dof[net.neg_side[:, 1]]
        self.m = DarcySystem(ps_m, device=device, split=np.stack([neg_e, dof_n], 1))
        self.f = DarcySystem(fracture_space(net.fmesh), device=device)
        m, f = self.m, self.f
        # the fracture system's element i is fracture triangle i
        tri_nodes = np.sort(net.frac_vertices[f.cells_np], axis=1)
        if not np.array_equal(tri_nodes, np.sort(mesh.entities[2].nodes[net.faces], axis=1)):
            raise RuntimeError("fracture elements are not in triangle order")
        n_um, n_uf = m.n_u, f.n_u
        n_pm = m.n - n_um
        N = m.n + f.n
        self.N, self.n_u = N, n_um + n_uf

        def gm(i):
            i = np.asarray(i, np.int64)
            return np.where(i < n_um, i, i + n_uf)

        def gf(i):
            i = np.asarray(i, np.int64)
            return np.where(i < n_uf, n_um + i, n_um + n_pm + i)

        self._gm, self._gf = gm, gf
        rm = np.repeat(np.arange(m.n), np.diff(m.pattern.rowptr))
        rf = np.repeat(np.arange(f.n), np.diff(f.pattern.rowptr))
        keys_m = gm(rm) * N + gm(m.pattern.col)
        keys_f = gf(rf) * N + gf(f.pattern.col)
        u0 = gm(m.ref_dest[pos_e, dof_p])
        u1 = gm(m.ref_dest[neg_e, dof_n])
        pf = gf(f.ref_dest[:, f.nb])
        cr = np.stack([u0, u0, pf, u1, u1, pf], 1)
        cc = np.stack([u0, pf, u0, u1, pf, u1], 1)
        keys_c = (cr * N + cc).reshape(-1)
        keys = np.unique(np.concatenate([keys_m, keys_f, keys_c]))
        rows, cols = keys // N, keys % N
        rowptr = np.zeros(N + 1, np.int64)
        np.add.at(rowptr, rows + 1, 1)
        rowptr = np.cumsum(rowptr)
        self.rowptr_np, self.col_np = rowptr, cols
        self.nnz = len(keys)
        pos_m = np.searchsorted(keys, keys_m)
        pos_f = np.searchsorted(keys, keys_f)
        pos_c = np.searchsorted(keys, keys_c).reshape(-1, 6)
        self.coupling_rows = (u0, u1, pf)
        # pressure diagonal and storage masses (global pressure order)
        vol_m = _cell_measures(mesh.points, m.cells_np)
        area_f = _cell_measures(net.fmesh.points, f.cells_np)
        delta_f = net.aperture
        mass = np.zeros(N - self.n_u)
        mass[gm(m.ref_dest[:, m.nb]) - self.n_u] = self.storage * vol_m / self.dt
        mass[gf(f.ref_dest[:, f.nb]) - self.n_u] = self.storage * delta_f * area_f / self.dt
        self.mass_np = mass
        pr = np.arange(self.n_u, N)
        pos_d = np.searchsorted(keys, pr * N + pr)
        # signed Schur diagonal of the pressure rows, a_kk - sum_j a_kj a_jk / a_jj over flux
        # columns j: positions of a_kj, a_jk and a_jj
        kj = np.flatnonzero((rows >= self.n_u) & (cols < self.n_u))
        jk = np.searchsorted(keys, cols[kj] * N + rows[kj])
        jj = np.searchsorted(keys, cols[kj] * N + cols[kj])
        if np.any(jk >= len(keys)) or np.any(keys[np.minimum(jk, len(keys) - 1)] != cols[kj] * N + rows[kj]):
            raise RuntimeError("coupled pattern is not structurally symmetric")
        # fracture-tip flux DoFs: u = 0 (row -> identity, column -> 0)
        tips = np.unique(gf(f.bc_rows.cpu().numpy().reshape(-1) & nat.SLOT_INDEX_MASK)) if f.F else np.zeros(0, np.int64)
        self.tips = tips
        is_tip = np.zeros(N, bool)
        is_tip[tips] = True
        hit = np.flatnonzero(is_tip[rows] | is_tip[cols])
        tip_v = (rows[hit] == cols[hit]).astype(np.float64)
        # device data
        dv = lambda a, dt=np.int32: to_dev(np.ascontiguousarray(np.asarray(a).astype(dt)), device=device)
        self.rowptr, self.col = dv(rowptr), dv(cols)
        self.vals32 = t.zeros(self.nnz, dtype=t.float32, device=device)
        self.vals64 = t.zeros(self.nnz, dtype=t.float64, device=device)
        self._pos_m, self._pos_f, self._pos_c = dv(pos_m), dv(pos_f), dv(pos_c)
        self._pos_d, self._mass_d = dv(pos_d), dv(mass, np.float64)
        # block-LU elimination of the fracture pressures in the preconditioner: for face i,
        # the matrix cells on its two sides (pressure rows a, b), their facet DoFs u0, u1
        # and the fracture pressure pf; positions of a_{a,u0}, a_{u0,pf}, a_{u0,u0},
        # a_{pf,u0}, a_{u0,a} (and the same on the negative side)
        def at(r, c):
            k = np.asarray(r, np.int64) * N + np.asarray(c, np.int64)
            i = np.searchsorted(keys, k)
            if np.any(i >= len(keys)) or np.any(keys[np.minimum(i, len(keys) - 1)] != k):
                raise RuntimeError("coupling entry missing from the coupled pattern")
            return i

        ra, rb = gm(m.ref_dest[pos_e, m.nb]), gm(m.ref_dest[neg_e, m.nb])
        lk = [at(ra, u0), at(u0, pf), at(u0, u0), at(pf, u0), at(u0, ra),
              at(rb, u1), at(u1, pf), at(u1, u1), at(pf, u1), at(u1, rb)]
        self._lk = [dv(a, np.int64) for a in lk]
        self._link_dofs = (m.ref_dest[pos_e, dof_p], m.ref_dest[neg_e, dof_n])
        self._link_rows = (ra - self.n_u, rb - self.n_u, pf - self.n_u)
        self._pos_m_np = pos_m
        self._sd_row = to_dev(rows[kj] - self.n_u, device=device)
        self._sd_kj, self._sd_jk, self._sd_jj = (to_dev(a, device=device) for a in (kj, jk, jj))
        self._pos_tip, self._tip_v = dv(hit), dv(tip_v, np.float64)
        self._map_m, self._map_f = dv(gm(np.arange(m.n))), dv(gf(np.arange(f.n)))
        self._side_elem = dv(np.stack([pos_e, neg_e], 1))
        from .device import facet_combo_index

        fid = net.faces + mesh.entities[2].offset
        c0 = facet_combo_index(mesh, 3, fid, net.pos_side[:, 0], net.pos_side[:, 1])
        c1 = facet_combo_index(mesh, 3, fid, net.neg_side[:, 0], net.neg_side[:, 1])
        self._side_combo = dv(np.stack([c0, c1], 1))
        code0 = row_codes(m.ref_dest[pos_e, dof_p], m.ref_sign[pos_e, dof_p])
        code1 = row_codes(m.ref_dest[neg_e, dof_n], m.ref_sign[neg_e, dof_n])
        self._side_code = dv(np.stack([code0, code1], 1))
        self._facet_nodes = dv(mesh.entities[2].nodes[net.faces])
        self.inv_alpha_np = delta_f / (2.0 * kappa_normal)
        self._inv_alpha64 = dv(self.inv_alpha_np, np.float64)
        self._inv_alpha32 = dv(self.inv_alpha_np, np.float32)
        if m.nqf != 4:
            raise ValueError("coupling kernel expects the 4-point facet quadrature of RT1 in 3-D")
        # coefficients
        self.kinv_m = CoefficientArg(t.full((1,), 1.0 / kappa_matrix, dtype=t.float64, device=device), 0, 0)
        self.kinv_f = CoefficientArg(to_dev(12.0 / delta_f ** 3, device=device), 1, 0)
        self.zero_m = CoefficientArg(t.zeros(1, dtype=t.float64, device=device), 0, 0)
        pb = p_boundary if p_boundary is not None else (lambda x, y, z: 1.0 - x)
        self.pD = m.dirichlet_values(pb)
        self.b0 = t.zeros(N, dtype=t.float64, device=device)
        self.pinv = t.empty(N, dtype=t.float32, device=device)
        self._solver = None

    def assemble(self, fp64, stream=None):
        """Values of the coupled system (and, with fp64, the static right-hand side)."""
        t = torch()
        m, f = self.m, self.f
        s = nat.stream_handle(stream)
        m.assemble(self.kinv_m, self.zero_m, self.pD, fp64=fp64, stream=stream)
        f.assemble_elements(self.kinv_f, self.zero_m, fp64=fp64, stream=stream)
        vals = self.vals64 if fp64 else self.vals32
        vals.zero_()
        fl = int(fp64)
        nat.call("nfx_md_scatter", m.nnz, nat.ptr(m.vals64 if fp64 else m.vals32), nat.ptr(self._pos_m), fl,
                 nat.ptr(vals), s)
        nat.call("nfx_md_scatter", f.nnz, nat.ptr(f.vals64 if fp64 else f.vals32), nat.ptr(self._pos_f), fl,
                 nat.ptr(vals), s)
        nat.call("nfx_md_coupling", len(self.net.faces), nat.ptr(self._side_elem), nat.ptr(self._side_combo),
                 nat.ptr(self._side_code), nat.ptr(self._facet_nodes), nat.ptr(m.cells),
                 nat.ptr(m.coords64 if fp64 else m.coords), nat.ptr(m.bc_trace64 if fp64 else m.bc_trace32),
                 nat.ptr(m.bc_fw64 if fp64 else m.bc_fw32), nat.ptr(self._inv_alpha64 if fp64 else self._inv_alpha32),
                 nat.ptr(self._pos_c), self.exchange_sign, fl, nat.ptr(vals), s)
        nat.call("nfx_md_update", len(self.mass_np), nat.ptr(self._pos_d), nat.ptr(self._mass_d), 1, fl,
                 nat.ptr(vals), s)
        nat.call("nfx_md_update", int(self._pos_tip.numel()), nat.ptr(self._pos_tip), nat.ptr(self._tip_v), 0, fl,
                 nat.ptr(vals), s)
        if fp64:
            rg = t.zeros(self.N, dtype=t.float64, device=self.device)
            nat.call("nfx_md_scatter", m.n, nat.ptr(m.rhs64), nat.ptr(self._map_m), 1, nat.ptr(rg), s)
            nat.call("nfx_md_scatter", f.n, nat.ptr(f.rhs64), nat.ptr(self._map_f), 1, nat.ptr(rg), s)
            self.b0.copy_(-rg)
            if len(self.tips):
                self.b0[to_dev(self.tips, device=self.device)] = 0.0
        return vals

    def setup(self, restart=60, preconditioner="learned"):
        """Assemble both precisions and set up the preconditioner and the solver, once.

        preconditioner: "diag" (the diagonal block preconditioner), "twogrid" (the two-grid
        Schur preconditioner on the matrix pressures, storage term included in S, Jacobi
        smoother) or "learned" (the same with the trained network smoother; it falls back to
        Jacobi when the network output is not finite); see _coupled_pc."""
        from .learned import Fgmres, default_weights_path, load_params

        self.assemble(True)
        self.assemble(False)
        nat.call("nfx_md_block_diag", self.N, self.n_u, nat.ptr(self.rowptr), nat.ptr(self.col),
                 nat.ptr(self.vals64), nat.ptr(self.pinv), nat.stream_handle())
        self.pc = None
        if preconditioner in ("twogrid", "learned"):
            theta = load_params(default_weights_path()) if preconditioner == "learned" else None
            self.pc = self._coupled_pc(theta)
        elif preconditioner != "diag":
            raise ValueError(f"unknown preconditioner {preconditioner!r}")
        if self._solver is None:
            self._solver = Fgmres(self.N, restart)
        return self

    def schur_diag(self):
        """a_kk - sum_j a_kj a_jk / a_jj over the pressure rows (FP64, global pressure order),
        from the FP64 values: the pressure Schur diagonal with the flux block replaced by its
        diagonal."""
        v = self.vals64
        s = v[self._pos_d].clone()
        s.index_add_(0, self._sd_row, -(v[self._sd_kj] * v[self._sd_jk] / v[self._sd_jj]))
        return s

    def schur_diag_inverse(self):
        """1 / schur_diag() in FP32 (1 where it vanishes)."""
        t = torch()
        s = self.schur_diag()
        return t.where(s != 0, 1.0 / s, t.ones_like(s)).float()

    def _coupled_pc(self, theta):
        """Two-grid preconditioner of the coupled system. Flux rows: 1/|a_kk|. Pressures: with
        the flux block replaced by its diagonal D, the pressure Schur complement
        [[S_mm, S_mf], [S_fm, S_ff]] over matrix cells m and fracture cells f. S_ff is taken
        as its diagonal (the tangential fracture conduction, ~delta^3, is dropped: a
        fracture cell then couples only to its two side cells), and the fracture pressures
        are eliminated by one block-LU step: the cells see S_mm - S_mf S_ff^-1 S_fm, an ELL
        matrix in which
        the two sides of a fracture face are neighbours through the face's slot, solved by
        the two-grid cycle; the fracture pressures follow by block back-substitution.
        S_mm is the Schur ELL of the matrix block read from the coupled values (so D
        includes the interface terms on the facet DoFs) plus the storage. The flux blocks
        of the matrix system are antisymmetric (a_up = -a_pu^T), so S_mm = C + B D^-1 B^T is
        positive and the two-grid sign psign is +1. With the reference exchange sign the
        coupling is not antisymmetric (S_mf != S_fm^T, S_ff of either sign): there the
        fracture pressures keep the diagonal 1 / S_ff and the cells S_mm + C (no
        elimination, no links)."""
        from .learned import LearnedSchurPC

        t = self.coupled_pc_terms()
        # (the flux rows are the coupled system's, matrix and fracture: its diagonal in
        # flux_dinv, and the two mass sweeps built from its own CSR, not the matrix system's)
        pc = LearnedSchurPC(self.m, theta, psign=1.0, graph=self._schur_graph, extra_diag=t["extra_diag"],
                            extra_off=t["extra_off"], flux_sweeps=1)
        pc.setup(vals64=self.vals64, network=theta is not None)
        pc.flux_dinv = self.pinv[: self.n_u].contiguous()
        pc.tail_dinv = t["tail_dinv"]
        if "slot_tail" in t:
            pc.tail_lu = (t["slot_tail"], t["slot_coef"], t["tail_cells"], t["tail_coef"])
        if self.flux_sweeps:
            pc.use_flux_rows(self.rowptr_np, self.col_np, self.n_u, self.vals64)
        return pc

    def coupled_pc_terms(self):
        """The host-built parts of the coupled preconditioner from the FP64 values (no
        library call): extra_diag / extra_off (added to the Schur ELL of the cells),
        tail_dinv, and with the conservative exchange the block-LU coupling slot_tail,
        slot_coef (n_cells * 4), tail_cells, tail_coef (n_tail * 2)."""
        from .learned import SchurGraph

        t = torch()
        m = self.m
        E = m.n - m.n_u
        lu = self.exchange_sign < 0
        if getattr(self, "_schur_graph", None) is None:
            self._schur_graph = SchurGraph(m, device=self.device, links=self._link_dofs if lu else None,
                                           pos_map=self._pos_m_np)
        g = self._schur_graph
        v = self.vals64
        sd = self.schur_diag()
        out = {"tail_dinv": t.where(sd[E:] != 0, 1.0 / sd[E:], t.ones_like(sd[E:])).float().contiguous(),
               "extra_diag": to_dev(self.mass_np[:E], device=self.device), "extra_off": None}
        if not lu:
            return out
        lk = [v[i] for i in self._lk]
        s_af = -lk[0] * lk[1] / lk[2]
        s_fa = -lk[3] * lk[4] / lk[2]
        s_bf = -lk[5] * lk[6] / lk[7]
        s_fb = -lk[8] * lk[9] / lk[7]
        ca, cb, fi = (to_dev(a, device=self.device) for a in self._link_rows)
        fi = fi - E
        s_ff = sd[fi + E]
        extra_diag = out["extra_diag"]
        extra_diag.index_add_(0, ca, -s_af * s_fa / s_ff)
        extra_diag.index_add_(0, cb, -s_bf * s_fb / s_ff)
        la, ja, lb, jb = (to_dev(a, device=self.device) for a in g.link_slots)
        sa, sb = la * 4 + ja, lb * 4 + jb
        extra_off = t.zeros(4 * E, dtype=t.float64, device=self.device)
        extra_off.index_add_(0, sa, -s_af * s_fb / s_ff)
        extra_off.index_add_(0, sb, -s_bf * s_fa / s_ff)
        n_tail = self.N - self.n_u - E
        slot_tail = t.full((4 * E,), -1, dtype=t.int32, device=self.device)
        slot_coef = t.zeros(4 * E, dtype=t.float32, device=self.device)
        slot_tail[sa] = fi.int()
        slot_tail[sb] = fi.int()
        slot_coef[sa] = (s_af / s_ff).float()
        slot_coef[sb] = (s_bf / s_ff).float()
        tail_cells = t.full((n_tail, 2), -1, dtype=t.int32, device=self.device)
        tail_coef = t.zeros((n_tail, 2), dtype=t.float32, device=self.device)
        tail_cells[fi, 0] = ca.int()
        tail_cells[fi, 1] = cb.int()
        tail_coef[fi, 0] = s_fa.float()
        tail_coef[fi, 1] = s_fb.float()
        # indices the kernels follow: cells in [0, E), tail rows in [0, n_tail)
        if int(fi.min()) < 0 or int(fi.max()) >= n_tail or int(t.maximum(ca, cb).max()) >= E:
            raise RuntimeError("fracture link indices out of range")
        out.update(extra_off=extra_off, slot_tail=slot_tail, slot_coef=slot_coef,
                   tail_cells=tail_cells.reshape(-1), tail_coef=tail_coef.reshape(-1))
        return out

    def rhs(self, x_prev, out=None):
        t = torch()
        b = out if out is not None else t.empty(self.N, dtype=t.float64, device=self.device)
        nat.call("nfx_md_rhs", self.N, self.n_u, nat.ptr(self.b0), nat.ptr(self._mass_d), nat.ptr(x_prev), nat.ptr(b),
                 nat.stream_handle())
        return b

    def step(self, x_prev, rtol=1e-8, inner_rtol=1e-5, max_inner=20000, pc_max_inner=2000):
        """One backward-Euler step from state x_prev (FP64, global order); returns (x, info).

        With a two-grid preconditioner set up, a solve that has not converged after
        pc_max_inner iterations continues from its iterate with the diagonal block
        preconditioner (info["fallback"] = True)."""
        b = self.rhs(x_prev)
        x = x_prev.clone()
        pc = getattr(self, "pc", None)
        if pc is not None:
            x, info = self._solver.solve(self.rowptr, self.col, self.vals32, self.vals64, b, x=x, pc=pc,
                                         rtol=rtol, inner_rtol=inner_rtol, max_inner=pc_max_inner, reorth=True,
                                         check=4)
            info["fallback"] = False
            if info["converged"]:
                return x, info
            its = info["iterations"]
            if not bool(torch().isfinite(x).all()):
                x = x_prev.clone()
        else:
            its = 0
        x, info = self._solver.solve(self.rowptr, self.col, self.vals32, self.vals64, b, x=x, pinv=self.pinv,
                                     rtol=rtol, inner_rtol=inner_rtol, max_inner=max_inner, reorth=True, check=4)
        info["fallback"] = pc is not None
        info["iterations"] += its
        return x, info

    def run(self, steps, x0=None, rtol=1e-8, callback=None):
        t = torch()
        x = x0 if x0 is not None else t.zeros(self.N, dtype=t.float64, device=self.device)
        infos = []
        for k in range(steps):
            x, info = self.step(x, rtol=rtol)
            if not info["converged"]:
                raise RuntimeError(f"step {k}: FGMRES did not converge ({info})")
            infos.append(info)
            if callback is not None:
                callback(k, x, info)
        return x, infos

    def csr_numpy(self, fp64=True):
        import scipy.sparse as sps

        v = (self.vals64 if fp64 else self.vals32).double().cpu().numpy()
        return sps.csr_matrix((v, self.col_np, self.rowptr_np), shape=(self.N, self.N))
