"""Dual-mixed elasticity (BDM1 - P0 - P0 on tetrahedra) resident on the GPU.

``ElasticitySystem(space, functions)`` mirrors the elasticity driver
``examples/cauchy_elasticity/mfe_elasticity_example_1.py:68-294``:

- operator A from ``LEDualWeakForm.evaluate_form_vectorized`` and preconditioning
  operator P from ``LERieszMapWeakForm.evaluate_form_vectorized``, both stored
  like PETSc SBAIJ (upper triangle, entries of an element contribution kept only
  when ``not isclose(v, 1e-16)``, explicit zero diagonals for A only,
  ``:128-192``). The value filter is decided once, in FP64 on the device (cold);
  warm assemblies reuse the pattern;
- residual ``rg`` including the stress Dirichlet term
  (``LEDualWeakFormBCDirichlet``, ``le_dual_weak_form.py:514-578``);
- the solve: the reference's MINRES + PCFIELDSPLIT additive {sigma: Cholesky of
  P_ss, (u, t): ICC} becomes FP32 MINRES with the block-diagonal inverse of P
  (one 9 x 9 block per face for sigma - the three BDM1 DoFs of the face times the
  three stress rows - and the diagonal mass of u, t) inside FP64 refinement.
"""

from __future__ import annotations

import numpy as np

from . import _native as nat
from .device import (CoefficientArg, check_coefficient, coefficient, facet_combo_index, facet_combo_tables,
                     p1_bary, to_dev, torch)
from .pattern import build_pattern, row_codes

NL = 42


def elasticity_tables(s_el, points, weights):
    """w (nq) | phi (nq, 12, 3) | reference divergence (nq, 12) | Uref (12) | Phi (12, 3) |
    DDref (12, 12): Uref_b = sum_q w divref_b, Phi_be = sum_q w phi_be,
    DDref_ab = sum_q w divref_a divref_b (the element-independent sums of the face kernel)."""
    tab = s_el.tabulate(points)
    div = np.einsum("kqjk->qj", tab[1:])
    phi = tab[0]
    uref = weights @ div
    Phi = np.einsum("q,qbe->be", weights, phi)
    dd = np.einsum("q,qa,qb->ab", weights, div, div)
    return np.concatenate([weights, phi.ravel(), div.ravel(), uref, Phi.ravel(), dd.ravel()])


def _vector_coefficient(value, xq, nq, ncomp, device):
    """(E, nq, ncomp) FP64 device array from a callable returning (ncomp, npts) or an array."""
    t = torch()
    if callable(value):
        x = xq.reshape(-1, 3)
        v = np.asarray(value(x[:, 0], x[:, 1], x[:, 2]), dtype=np.float64)
        v = np.broadcast_to(v.reshape(v.shape[0], -1), (v.shape[0], len(x)))[:ncomp]
        arr = np.ascontiguousarray(v.T).reshape(xq.shape[0], nq, ncomp)
        return to_dev(arr, device=device)
    arr = value if isinstance(value, t.Tensor) else t.from_numpy(np.asarray(value, dtype=np.float64))
    arr = arr.to(device=device, dtype=t.float64)
    if arr.dim() == 2:  # (E, ncomp) per element
        arr = arr[:, None, :].expand(arr.shape[0], nq, ncomp)
    return arr.contiguous()


def _block3_structure(n, rowptr, col, device):
    """3 x 3 block CSR structure of a full CSR pattern (n a multiple of 3): block row
    pointer, block columns, and for every CSR entry its value slot (3a + b) * n_blocks + j
    in the block copy (entry (a, b) of block j)."""
    if n % 3:
        return None
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rowptr))
    cols = col.astype(np.int64)
    nbc = n // 3
    key = (rows // 3) * nbc + cols // 3
    ukey, inv = np.unique(key, return_inverse=True)
    nblk = len(ukey)
    brp = np.zeros(nbc + 1, np.int64)
    np.add.at(brp, ukey // nbc + 1, 1)
    brp = np.cumsum(brp)
    dest = ((rows % 3) * 3 + cols % 3) * nblk + inv
    if 9 * nblk > np.iinfo(np.int32).max:
        return None
    t = torch()
    return {"n_blocks": nblk, "browptr": to_dev(brp.astype(np.int32), device=device),
            "bcol": to_dev((ukey % nbc).astype(np.int32), device=device),
            "dest": to_dev(dest.astype(np.int32), device=device),
            "bval": t.zeros(9 * nblk, dtype=t.float32, device=device)}


def _pair_slots(slot):
    """(E, NL (NL + 1) / 2) codes, one per local upper-triangle entry (row-major), from an
    entry-major (NL * NL, E) map of an upper (SBAIJ) operator: the code of whichever of
    (I, J), (J, I) the operator keeps (-1 when neither)."""
    iu, ju = np.triu_indices(NL)
    a = slot[iu * NL + ju]
    b = slot[ju * NL + iu]
    if np.any((a >= 0) & (b >= 0) & (iu != ju)[:, None]):
        raise RuntimeError("slot map keeps both orientations of an entry: not an upper operator")
    return np.ascontiguousarray(np.where(a >= 0, a, b).T).astype(np.int32)


def _sbaij_mask(vals, ref_dest, ref_sign, upper=True):
    """Entries the reference inserts: not isclose(v, 1e-16) (numpy defaults), row <= col."""
    v = vals * (ref_sign[:, :, None] * ref_sign[:, None, :])
    keep = ~np.isclose(v, 1.0e-16)
    if upper:
        keep &= ref_dest[:, :, None] <= ref_dest[:, None, :]
    return keep


FACE_NO_SRC = 63


def el_sizes(nl):
    """BDM1 - P0 - P0 sizes from the local DoF count (42: tetrahedra, 15: triangles): basis
    functions NB, stress rows / displacement components NC, rotation components NT, BDM1
    DoFs per facet NBF, facets per element NFC, rows of one facet block ROWS = NBF NC."""
    if nl == 42:
        d = dict(dim=3, NB=12, NC=3, NT=3, NBF=3, NFC=4)
    elif nl == 15:
        d = dict(dim=2, NB=6, NC=2, NT=1, NBF=2, NFC=3)
    else:
        raise NotImplementedError(f"BDM1 - P0 - P0: unexpected local size {nl}")
    d["ROWS"] = d["NBF"] * d["NC"]
    d["NSS"] = d["NB"] * d["NC"]
    d["NL"] = nl
    return d


def face_owner_layout(n, ref_dest, ref_sign):
    """Per-facet and per-element data of the facet-owner kernel (nfx_assemble_elasticity_faces),
    tetrahedra (faces) or triangles (edges).

    The stress DoFs come in blocks of ROWS per facet (NBF BDM1 DoFs x NC rows, ``j * NC + c``),
    the displacement and rotation DoFs per cell (``n_s + NC e + c``, ``n_s + n_u + NT e + c``).
    Returns a dict with face_inc (n_f, 2) (incident elements in element order, -1 none),
    face_lf (n_f,) (local facet index of each incidence, 2 bits each), el_face (E, NFC) (the
    DoF block of each local facet) and el_orient (E,) uint64 (bits 2 b, 2 b + 1: facet slot of
    reference basis b; bit 32 + b: its orientation sign), or raises NotImplementedError
    when the DoF layout is not this one."""
    E = ref_dest.shape[0]
    z = el_sizes(ref_dest.shape[1])
    NB, NC, NT, NBF, NFC, ROWS, NSS = z["NB"], z["NC"], z["NT"], z["NBF"], z["NFC"], z["ROWS"], z["NSS"]
    sd = ref_dest[:, :NSS].reshape(E, NB, NC).astype(np.int64)
    if np.any(sd != sd[:, :, :1] + np.arange(NC)):
        raise NotImplementedError("face-owner layout: stress components must be consecutive")
    base = sd[:, :, 0]
    if np.any(base % NC):
        raise NotImplementedError("face-owner layout: stress DoF blocks")
    block, slot = base // ROWS, (base % ROWS) // NC
    el_face = block[:, ::NBF]
    if np.any(block != np.repeat(el_face, NBF, axis=1)):
        raise NotImplementedError("face-owner layout: a facet's BDM1 DoFs must share one DoF block")
    if np.any(np.sort(slot.reshape(E, NFC, NBF), axis=2) != np.arange(NBF)):
        raise NotImplementedError("face-owner layout: facet slots")
    sg = ref_sign[:, :NSS].reshape(E, NB, NC)
    if np.any(sg != sg[:, :, :1]):
        raise NotImplementedError("face-owner layout: signs must not depend on the stress row")
    neg = (sg[:, :, 0] < 0).astype(np.uint64)
    b = np.arange(NB, dtype=np.uint64)[None, :]
    orient = ((slot.astype(np.uint64) << (np.uint64(2) * b)) | (neg << (np.uint64(32) + b))).sum(axis=1,
                                                                                                dtype=np.uint64)
    n_f = int(el_face.max()) + 1 if E else 0
    n_s = ROWS * n_f
    n_u = NC * E
    if n != n_s + n_u + NT * E:
        raise NotImplementedError("face-owner layout: n != ROWS n_facets + (NC + NT) n_cells")
    e_idx = np.arange(E)[:, None]
    if np.any(ref_dest[:, NSS:NSS + NC] != n_s + NC * e_idx + np.arange(NC)) or \
            np.any(ref_dest[:, NSS + NC:] != n_s + n_u + NT * e_idx + np.arange(NT)):
        raise NotImplementedError("face-owner layout: cell DoFs must follow the element order")
    flat = el_face.reshape(-1)
    order = np.argsort(flat, kind="stable")  # element order inside each facet
    counts = np.bincount(flat, minlength=n_f)
    if counts.min(initial=1) < 1 or counts.max(initial=1) > 2:
        raise NotImplementedError("face-owner layout: every facet needs 1 or 2 incident elements")
    first = np.concatenate([[0], np.cumsum(counts)[:-1]])
    face_inc = np.full((n_f, 2), -1, np.int64)
    lf = np.zeros((n_f, 2), np.int64)
    face_inc[:, 0] = order[first] // NFC
    lf[:, 0] = order[first] % NFC
    two = counts == 2
    face_inc[two, 1] = order[first[two] + 1] // NFC
    lf[two, 1] = order[first[two] + 1] % NFC
    return {"n_faces": n_f, "n_s": n_s, "n_u": n_u, "face_inc": face_inc, "face_lf": lf[:, 0] | (lf[:, 1] << 2),
            "el_face": el_face, "el_orient": orient, "rows": ROWS}


def face_owner_codes(n, lay, ref_dest, keep, pattern, zero_diag):
    """16-bit source code of every CSR entry of the stress rows (bits 0..3: row inside the
    facet block, j NC + c; bits 4..9 / 10..15: element-local reference column of the entry in
    the facet's first / second incident element, 63 = none), after checking that every cell
    row holds exactly its diagonal (its only entry in both SBAIJ operators)."""
    E = ref_dest.shape[0]
    z = el_sizes(ref_dest.shape[1])
    nl, ROWS, NSS = z["NL"], z["ROWS"], z["NSS"]
    n_s = lay["n_s"]
    rowptr = pattern.rowptr.astype(np.int64)
    col = pattern.col.astype(np.int64)
    if np.any(np.diff(rowptr[n_s:]) != 1) or np.any(col[rowptr[n_s:n]] != np.arange(n_s, n)):
        raise NotImplementedError("face-owner layout: cell rows must hold only their diagonal")
    rows = np.repeat(ref_dest[:, :, None], nl, axis=2)
    cols = np.repeat(ref_dest[:, None, :], nl, axis=1)
    src = keep & (rows <= cols) & (rows < n_s)
    e, I, Jl = np.nonzero(src)
    R = ref_dest[e, I].astype(np.int64)
    C = ref_dest[e, Jl].astype(np.int64)
    keys = np.repeat(np.arange(n, dtype=np.int64), np.diff(rowptr)) * n + col
    k = np.searchsorted(keys, R * n + C)
    if np.any(k >= len(keys)) or np.any(keys[np.minimum(k, len(keys) - 1)] != R * n + C):
        raise ValueError("face-owner codes: an element entry is missing from the pattern")
    p = R // ROWS
    inc = (lay["face_inc"][p, 1] == e).astype(np.int64)
    if np.any((inc == 0) & (lay["face_inc"][p, 0] != e)):
        raise ValueError("face-owner codes: an element entry of a face row from a non-incident element")
    ns_entries = int(rowptr[n_s])
    srcs = np.full((ns_entries, 2), FACE_NO_SRC, np.int64)
    flat = k * 2 + inc
    if len(np.unique(flat)) != len(flat):
        raise ValueError("face-owner codes: two contributions of one element to one entry")
    srcs.reshape(-1)[flat] = Jl
    row_of = np.repeat(np.arange(n_s, dtype=np.int64), np.diff(rowptr[: n_s + 1]))
    none = (srcs[:, 0] == FACE_NO_SRC) & (srcs[:, 1] == FACE_NO_SRC)
    if np.any(none & (row_of != col[:ns_entries])) or (not zero_diag and np.any(none)):
        raise ValueError("face-owner codes: an entry without a contribution that is not an explicit zero diagonal")
    # the kernel evaluates the stress columns of facets whose block is >= the row's facet
    sc = srcs < NSS
    blk_col = np.where(sc, col[:ns_entries, None] // ROWS, 1 << 40)
    if np.any(blk_col < (row_of // ROWS)[:, None]):
        raise ValueError("face-owner codes: a stress column below the row's face block")
    code = (row_of % ROWS) | (srcs[:, 0] << 4) | (srcs[:, 1] << 10)
    return code.astype(np.uint16)


def full_from_upper(n, rowptr, col):
    """Full symmetric CSR pattern from an upper-triangle pattern + gather map into it."""
    r = np.repeat(np.arange(n, dtype=np.int64), np.diff(rowptr))
    c = col.astype(np.int64)
    idx = np.arange(len(c), dtype=np.int64)
    off = r != c
    rows = np.concatenate([r, c[off]])
    cols = np.concatenate([c, r[off]])
    src = np.concatenate([idx, idx[off]])
    order = np.lexsort((cols, rows))
    rows, cols, src = rows[order], cols[order], src[order]
    fr = np.zeros(n + 1, np.int64)
    np.add.at(fr, rows + 1, 1)
    return np.cumsum(fr).astype(np.int32), cols.astype(np.int32), src.astype(np.int32)


class ElasticitySystem:
    def __init__(self, space, functions, device="cuda"):
        t = torch()
        self.space = space
        self.device = device
        # inner MINRES on the 3 x 3 block copy of the operator (False: the scalar CSR)
        self.use_block3 = True
        names = space.names
        self.s_sp, self.u_sp, self.t_sp = (space.discrete_spaces[n] for n in names)
        s_sp = self.s_sp
        mesh = s_sp.mesh
        D = s_sp.dimension
        if D != 3 or s_sp.element.family != "BDM" or s_sp.k_order != 1 or s_sp.n_comp != 3:
            raise NotImplementedError("GPU elasticity path: BDM1 - P0 - P0 on tetrahedra")
        self.dim = 3
        self.E = len(s_sp.element_ids)
        self.n = space.n_dof
        self.n_s = s_sp.n_dof
        # DoF ranges of the fields (stress | displacement | rotation), in the space's order
        self.field_bounds = [0, self.n_s, self.n_s + self.u_sp.n_dof, self.n]
        points, weights = space.quadrature[3]
        if len(weights) != 8:
            raise ValueError("expected the degree-3 tetrahedral rule (8 points)")
        self.nq = 8
        cells = mesh.entities[3].nodes[s_sp.topology.cell_index].astype(np.int32)
        self.cells_np = cells
        if cells.min() < 0 or cells.max() >= len(mesh.points):
            raise ValueError("cells out of range")
        self.cells = to_dev(cells, device=device)
        self.coords = to_dev(mesh.points.astype(np.float32), device=device)
        self.coords64 = to_dev(mesh.points, device=device)
        tab = elasticity_tables(s_sp.element, points, weights)
        self.tables64 = to_dev(tab, device=device)
        self.tables32 = to_dev(tab.astype(np.float32), device=device)
        X = mesh.points[cells]
        self.xq = np.einsum("qa,eac->eqc", p1_bary(points), X)
        # per-element geometry record of the face kernel: J / |det J| row-major, then |det J|,
        # with J[c][d] = X_{d+1,c} - X_{0,c}
        J = np.transpose(X[:, 1:] - X[:, :1], (0, 2, 1))
        det = np.abs(np.linalg.det(J))
        if self.E and det.min() <= 0:
            raise ValueError("degenerate element")
        geom = np.concatenate([(J / det[:, None, None]).reshape(self.E, 9), det[:, None]], axis=1)
        self.geom64 = to_dev(np.ascontiguousarray(geom), device=device)
        self.geom32 = to_dev(np.ascontiguousarray(geom.astype(np.float32)), device=device)
        if self.E * NL * NL > nat.SLOT_INDEX_MASK:
            raise ValueError("mesh too large for the dense cold pass (28-bit local codes)")
        ref_dest, ref_sign = space.reference_destinations()
        if ref_dest.shape[1] != NL:
            raise ValueError("unexpected local size")
        self.ref_dest, self.ref_sign = ref_dest, ref_sign
        self.functions = functions
        self._coef = self._coefficients(functions)
        # cold: FP64 element matrices -> value filter -> SBAIJ patterns
        # element-major slot map (the kernel's layout): code e * NL * NL + i * NL + j
        loc_slot = np.arange(self.E * NL * NL, dtype=np.int32)
        d_loc = to_dev(loc_slot, device=device)
        KA = self._local(0, d_loc)
        KP = self._local(1, d_loc)
        keepA = _sbaij_mask(KA, ref_dest, ref_sign, upper=False)
        keepP = _sbaij_mask(KP, ref_dest, ref_sign, upper=False)
        # signed (u, t) x stress coupling of every element, coefficient-independent (the
        # two-grid preconditioner's B blocks)
        self._b_ref = KA[:, 36:42, :36] * (ref_sign[:, 36:42, None] * ref_sign[:, None, :36])
        del KA, KP
        self.pattern_A = build_pattern(self.n, ref_dest, ref_sign, keep=keepA, zero_diag_dest=ref_dest, upper=True)
        self.pattern_P = build_pattern(self.n, ref_dest, ref_sign, keep=keepP, upper=True)
        self._upload_patterns()
        # face-owner kernel (the default warm path): one wave per face block of 9 CSR rows
        self.face_layout = None
        try:
            lay = face_owner_layout(self.n, ref_dest, ref_sign)
            codes = {"A": face_owner_codes(self.n, lay, ref_dest, keepA, self.pattern_A, zero_diag=True),
                     "P": face_owner_codes(self.n, lay, ref_dest, keepP, self.pattern_P, zero_diag=False)}
        except NotImplementedError:
            lay = None
        del keepA, keepP
        if lay is not None:
            self.face_layout = lay
            self.face_inc = to_dev(lay["face_inc"].astype(np.int32), device=device)
            self.face_lf = to_dev(lay["face_lf"].astype(np.int32), device=device)
            self.el_face = to_dev(lay["el_face"].astype(np.int32), device=device)
            self.el_orient = to_dev(lay["el_orient"].view(np.int64), device=device)
            for key in ("A", "P"):
                self.ops[key]["codes_np"] = codes[key]
                self.ops[key]["codes"] = to_dev(codes[key].view(np.int16), device=device)
        self.use_faces = lay is not None
        self.rhs_row = to_dev(np.ascontiguousarray(row_codes(ref_dest[:, 36:]).T), device=device)
        self.rhs32 = t.zeros(self.n, dtype=t.float32, device=device)
        self.rhs64 = t.zeros(self.n, dtype=t.float64, device=device)
        self._setup_bc(mesh, device)
        self._validate(mesh)

    # -- setup ---------------------------------------------------------------------
    def _coefficients(self, functions):
        lam = coefficient(functions["lambda"], self.xq, self.nq, device=self.device)
        mu = coefficient(functions["mu"], self.xq, self.nq, device=self.device)
        if (lam.se, lam.sq) != (mu.se, mu.sq):
            t = torch()
            E, nq = self.E, self.nq

            def full(c):
                a = c.get(True)
                if c.se == 0:
                    return a.expand(E * nq).contiguous()
                if c.sq == 0:
                    return a.reshape(-1, 1).expand(E, nq).reshape(-1).contiguous()
                return a

            lam = CoefficientArg(full(lam), nq, 1)
            mu = CoefficientArg(full(mu), nq, 1)
        check_coefficient(lam, self.E, self.nq, "lambda")
        check_coefficient(mu, self.E, self.nq, "mu")
        f = functions.get("rhs")
        fv = _vector_coefficient(f, self.xq, self.nq, 6, self.device) if f is not None else \
            torch().zeros((self.E, self.nq, 6), dtype=torch().float64, device=self.device)
        if fv.shape != (self.E, self.nq, 6):
            raise ValueError("rhs must give 6 components (3 force + 3 rotation) per point")
        return {"lam": lam, "mu": mu, "f": CoefficientArg(fv.reshape(-1), self.nq * 6, 6)}

    def _local(self, mode, d_loc):
        t = torch()
        out = t.zeros(self.E * NL * NL, dtype=t.float64, device=self.device)
        rhs = t.zeros(max(self.E * NL, 1), dtype=t.float64, device=self.device)
        rows = to_dev(np.ascontiguousarray((np.arange(self.E)[:, None] * NL + np.arange(36, 42)[None, :]).T)
                      .astype(np.int32), device=self.device)
        self._launch(mode, True, d_loc, rows, out, out, rhs, pairs=0)
        self._held = rows
        self._local_rhs = rhs
        return out.cpu().numpy().reshape(self.E, NL, NL)

    def _launch(self, mode, fp64, slot, rhs_row, vals, stage, rhs, stream=None, pairs=1):
        c = self._coef
        lam, mu, f = c["lam"], c["mu"], c["f"]
        nat.call("nfx_assemble_elasticity", mode, self.E, nat.ptr(self.cells),
                 nat.ptr(self.coords64 if fp64 else self.coords), nat.ptr(self.tables64 if fp64 else self.tables32),
                 nat.ptr(lam.get(fp64)), nat.ptr(mu.get(fp64)), lam.se, lam.sq, nat.ptr(f.get(fp64)),
                 self.nq * 6, 6, 1, nat.ptr(slot), int(pairs), nat.ptr(rhs_row), int(fp64), nat.ptr(vals),
                 nat.ptr(stage),
                 nat.ptr(rhs), nat.stream_handle(stream))

    def _upload_patterns(self):
        t = torch()
        dev = self.device
        self.ops = {}
        for key, pat in (("A", self.pattern_A), ("P", self.pattern_P)):
            frp, fcol, fsrc = full_from_upper(self.n, pat.rowptr, pat.col)
            self.ops[key] = {
                "pattern": pat,
                "rowptr": to_dev(pat.rowptr, device=dev), "col": to_dev(pat.col, device=dev),
                # the host pattern is entry-major (NL * NL, E); the kernel reads one code per
                # upper-triangle entry, element-major
                "slot": to_dev(_pair_slots(pat.slot), device=dev), "segptr": to_dev(pat.segptr, device=dev),
                "target": to_dev(pat.target, device=dev),
                "vals32": t.zeros(pat.nnz, dtype=t.float32, device=dev),
                "vals64": t.zeros(pat.nnz, dtype=t.float64, device=dev),
                "stage32": t.zeros(max(pat.n_stage, 1), dtype=t.float32, device=dev),
                "stage64": t.zeros(max(pat.n_stage, 1), dtype=t.float64, device=dev),
                "full_rowptr": to_dev(frp, device=dev), "full_col": to_dev(fcol, device=dev),
                "full_src": to_dev(fsrc, device=dev), "full_nnz": int(len(fcol)),
            }
            if key == "A":
                self.ops[key]["block3"] = _block3_structure(self.n, frp, fcol, dev)

    def _setup_bc(self, mesh, device):
        s_sp = self.s_sp
        self.F = len(s_sp.bc_element_ids)
        if not self.F:
            return
        fpts, fw = self.space.bc_quadrature[2]
        self.nqf = len(fw)
        neigh_cells = s_sp.topology.cell_index[s_sp.bc_neighbour]
        combo = facet_combo_index(mesh, 3, s_sp.bc_element_ids, neigh_cells, s_sp.bc_local)
        trace = facet_combo_tables(s_sp.element, 3, fpts)
        fnodes = mesh.entities[2].nodes[s_sp.bc_element_ids - mesh.entities[2].offset]
        self.bc_facet = to_dev(fnodes.astype(np.int32), device=device)
        self.bc_neigh = to_dev(s_sp.bc_neighbour.astype(np.int32), device=device)
        self.bc_combo = to_dev(combo.astype(np.int32), device=device)
        self.bc_trace64 = to_dev(trace, device=device)
        self.bc_trace32 = to_dev(trace.astype(np.float32), device=device)
        self.bc_fw64 = to_dev(fw, device=device)
        self.bc_fw32 = to_dev(fw.astype(np.float32), device=device)
        shift = self.space.strides[self.space.names[0]] + self.space.dof_shift
        self.bc_rows_np = row_codes(s_sp.bc_ref_dest + shift, s_sp.bc_ref_sign)
        self.bc_rows = to_dev(self.bc_rows_np, device=device)
        self.bc_xq = np.einsum("qa,fac->fqc", p1_bary(fpts), mesh.points[fnodes])

    def _validate(self, mesh):
        for key in ("A", "P"):
            op = self.ops[key]
            pat = op["pattern"]
            if pat.slot.shape != (NL * NL, self.E):
                raise ValueError("slot map shape")
            c = pat.slot.astype(np.int64).ravel()
            live = c[c >= 0]
            idx = live & nat.SLOT_INDEX_MASK
            st = (live & nat.SLOT_STAGE) != 0
            if np.any(idx[~st] >= pat.nnz) or np.any(idx[st] >= max(pat.n_stage, 1)):
                raise ValueError(f"slot codes of {key} out of range")
        rr = self.ref_dest[:, 36:]
        if rr.min() < 0 or rr.max() >= self.n:
            raise ValueError("residual rows out of range")
        if self.F:
            if self.nqf != 4 or self.bc_trace64.shape[1:] != (4, 3, 3):
                raise ValueError("facet tables do not match the compiled kernel")
            if self.bc_rows_np.shape != (self.F, 9) or (self.bc_rows_np & nat.SLOT_INDEX_MASK).max() >= self.n:
                raise ValueError("facet rows out of range")
        if self.use_faces:
            # every index the face-owner kernel forms stays inside its buffer
            lay = self.face_layout
            nf, E = lay["n_faces"], self.E
            fi = lay["face_inc"]
            if fi.shape != (nf, 2) or fi[:, 0].min(initial=0) < 0 or fi.max(initial=-1) >= E or \
                    np.any(fi[:, 1] < -1) or np.any((fi[:, 1] >= 0) & (fi[:, 1] <= fi[:, 0])):
                raise ValueError("face incidences out of range or not in element order")
            if lay["el_face"].shape != (E, 4) or lay["el_face"].min(initial=0) < 0 or \
                    lay["el_face"].max(initial=-1) >= nf or np.any(lay["face_lf"] >> 4):
                raise ValueError("element face blocks out of range")
            if lay["n_s"] != self.n_s or lay["n_u"] != self.u_sp.n_dof or self.n != lay["n_s"] + 2 * lay["n_u"]:
                raise ValueError("face-owner layout sizes")
            if self.tables32.numel() != 8 + 8 * 36 + 8 * 12 + 12 + 36 + 144:
                raise ValueError("elasticity table size")
            if tuple(self.geom32.shape) != (E, 10) or tuple(self.geom64.shape) != (E, 10):
                raise ValueError("element geometry record size")
            for key in ("A", "P"):
                op = self.ops[key]
                code = op["codes_np"].astype(np.int64)
                rp = op["pattern"].rowptr
                if len(code) != int(rp[self.n_s]) or len(rp) != self.n + 1:
                    raise ValueError(f"face codes of {key}: size")
                s0, s1 = (code >> 4) & 63, code >> 10
                if np.any((code & 15) >= 9) or np.any((s0 >= NL) & (s0 != FACE_NO_SRC)) or \
                        np.any((s1 >= NL) & (s1 != FACE_NO_SRC)):
                    raise ValueError(f"face codes of {key} out of range")
                p_of = np.repeat(np.arange(self.n_s) // 9, np.diff(rp[: self.n_s + 1]))
                if np.any((s1 != FACE_NO_SRC) & (fi[p_of, 1] < 0)):
                    raise ValueError(f"face codes of {key}: a second source on a one-element face")

    # -- warm assembly ---------------------------------------------------------------
    def assemble(self, which="A", fp64=False, stream=None, u_dirichlet=None, method=None):
        """Values of operator ``which`` ("A" or "P") and, for A, the residual. method "faces"
        (the default when the DoF layout allows it): the face-owner kernel; "element": one
        wave per element through the slot map and the staged segment sum."""
        op = self.ops[which]
        s = nat.stream_handle(stream)
        vals = op["vals64" if fp64 else "vals32"]
        stage = op["stage64" if fp64 else "stage32"]
        rg = self.rhs64 if fp64 else self.rhs32
        mode = 0 if which == "A" else 1
        method = method or ("faces" if self.use_faces else "element")
        if method == "faces":
            if not self.use_faces:
                raise NotImplementedError("face-owner assembly: DoF layout not supported")
            c = self._coef
            lam, mu, f = c["lam"], c["mu"], c["f"]
            lay = self.face_layout
            nat.call("nfx_assemble_elasticity_faces", self.dim, mode, lay["n_faces"], self.E,
                     nat.ptr(self.face_inc), nat.ptr(self.face_lf), nat.ptr(self.el_face), nat.ptr(self.el_orient),
                     nat.ptr(self.geom64 if fp64 else self.geom32),
                     nat.ptr(self.tables64 if fp64 else self.tables32), nat.ptr(lam.get(fp64)),
                     nat.ptr(mu.get(fp64)), lam.se, lam.sq, nat.ptr(f.get(fp64)), self.nq * 6, 6, 1,
                     nat.ptr(op["rowptr"]), nat.ptr(op["codes"]), int(fp64), nat.ptr(vals), nat.ptr(rg), s)
            if mode == 0 and u_dirichlet is not None and self.F:
                self.assemble_dirichlet(u_dirichlet, fp64=fp64, stream=stream)
            return vals, rg
        if mode == 0:
            nat.call("nfx_fill", int(fp64), self.n, 0.0, nat.ptr(rg), s)
        self._launch(mode, fp64, op["slot"], self.rhs_row, vals, stage, rg, stream)
        pat = op["pattern"]
        nat.call("nfx_stage_reduce", len(pat.target), nat.ptr(op["segptr"]), nat.ptr(op["target"]), int(fp64),
                 nat.ptr(stage), nat.ptr(vals), s)
        if mode == 0 and u_dirichlet is not None and self.F:
            self.assemble_dirichlet(u_dirichlet, fp64=fp64, stream=stream)
        return vals, rg

    def assemble_dirichlet(self, u_dirichlet, fp64=False, stream=None, rows=None, out=None):
        uD = _vector_coefficient(u_dirichlet, self.bc_xq, self.nqf, 3, self.device)
        if uD.shape != (self.F, self.nqf, 3):
            raise ValueError("u_D must give 3 components per facet point")
        rg = out if out is not None else (self.rhs64 if fp64 else self.rhs32)
        g = (uD if fp64 else uD.to(torch().float32)).contiguous()
        nat.call("nfx_assemble_elasticity_bc", 3, self.F, nat.ptr(self.bc_facet), nat.ptr(self.bc_neigh),
                 nat.ptr(self.bc_combo), nat.ptr(self.cells), nat.ptr(self.coords64 if fp64 else self.coords),
                 nat.ptr(self.bc_trace64 if fp64 else self.bc_trace32), nat.ptr(self.bc_fw64 if fp64 else self.bc_fw32),
                 self.nqf, nat.ptr(g), nat.ptr(self.bc_rows if rows is None else rows), int(fp64),
                 nat.ptr(rg), nat.stream_handle(stream))

    def _full_buffers(self, which):
        t = torch()
        op = self.ops[which]
        if "full32" not in op:
            op["full32"] = t.empty(op["full_nnz"], dtype=t.float32, device=self.device)
            op["full64"] = t.empty(op["full_nnz"], dtype=t.float64, device=self.device)

    def full(self, which="A"):
        """Full symmetric (rowptr, col, vals32, vals64) for the Krylov solver."""
        t = torch()
        op = self.ops[which]
        n = op["full_nnz"]
        self._full_buffers(which)
        s = nat.stream_handle()
        nat.call("nfx_gather", 0, n, nat.ptr(op["full_src"]), nat.ptr(op["vals32"]), nat.ptr(op["full32"]), s)
        nat.call("nfx_gather", 1, n, nat.ptr(op["full_src"]), nat.ptr(op["vals64"]), nat.ptr(op["full64"]), s)
        return op["full_rowptr"], op["full_col"], op["full32"], op["full64"]

    def csr_numpy(self, which="A", fp64=True):
        import scipy.sparse as sps

        op = self.ops[which]
        pat = op["pattern"]
        vals = op["vals64" if fp64 else "vals32"].cpu().numpy()
        return sps.csr_matrix((vals, pat.col, pat.rowptr), shape=(self.n, self.n))

    # -- preconditioner -----------------------------------------------------------
    def preconditioner(self, P=None, kind="mass_schur", device=True):
        """(sign, pc) for MinresIR: an SPD block-diagonal preconditioner as a CSR matrix.

        device=True builds "mass_schur" on the device from the FP64 values
        (nfx_elasticity_mass_schur_pc: block gathers and FP64 block inverses, FP32 output);
        device=False (and kind "riesz_block") build it on the host with numpy/scipy.

        kind "mass_schur" (default): sigma rows get the inverse of the 9 x 9 face blocks
        of the compliance block A_ss (each face's 3 BDM1 DoFs x 3 stress rows are
        consecutive in the DoF numbering); the (u, t) rows get the inverse of the 6 x 6
        cell blocks of the Schur complement S = B D^-1 B^T, D the 9 x 9 face blocks,
        B the (u, t) x sigma block of A.
        kind "riesz_block": 9 x 9 face blocks of the Riesz operator P and its diagonal
        on u, t (the reference factorises P's sigma block exactly,
        ``mfe_elasticity_example_1.py:246-262``; a block-Jacobi of that H(div) operator
        is a much weaker stand-in)."""
        import scipy.sparse as sps

        if kind == "mass_schur" and device:
            return None, self._mass_schur_pc_device()
        which = "A" if kind == "mass_schur" else "P"
        if which == "P":
            self.assemble("P", fp64=True)
        # A: its FP64 values from the caller's assembly (re-assembling here would replace
        # the residual, which carries the caller's Dirichlet data)
        rp, col, _, v64 = self.full(which)
        Mh = sps.csr_matrix((v64.cpu().numpy(), col.cpu().numpy(), rp.cpu().numpy()), shape=(self.n, self.n))
        ns = self.n_s
        nb = ns // 9
        blocks = np.zeros((nb, 9, 9))
        Ms = Mh[:ns, :ns].tocoo()
        same = (Ms.row // 9) == (Ms.col // 9)
        blocks[Ms.row[same] // 9, Ms.row[same] % 9, Ms.col[same] % 9] = Ms.data[same]
        Dinv = np.linalg.inv(blocks)
        Dinv = 0.5 * (Dinv + Dinv.transpose(0, 2, 1))
        rows_b = (np.arange(nb)[:, None, None] * 9 + np.arange(9)[None, :, None]).repeat(9, axis=2)
        cols_b = (np.arange(nb)[:, None, None] * 9 + np.arange(9)[None, None, :]).repeat(9, axis=1)
        r_list, c_list, v_list = [rows_b.ravel()], [cols_b.ravel()], [Dinv.ravel()]
        if kind == "mass_schur":
            nu = self.u_sp.n_dof
            E = nu // 3
            B = Mh[ns:, :ns].tocoo()
            loc = B.row
            cell = (loc % nu) // 3
            comp = (loc // nu) * 3 + loc % 3
            face = B.col // 9
            key = cell.astype(np.int64) * nb + face
            uk, inv = np.unique(key, return_inverse=True)
            Bblk = np.zeros((len(uk), 6, 9))
            np.add.at(Bblk, (inv, comp, B.col % 9), B.data)
            pc_cell = uk // nb
            pc_face = uk % nb
            contrib = np.einsum("pai,pij,pbj->pab", Bblk, Dinv[pc_face], Bblk)
            S = np.zeros((E, 6, 6))
            np.add.at(S, pc_cell, contrib)
            Sinv = np.linalg.inv(S)
            Sinv = 0.5 * (Sinv + Sinv.transpose(0, 2, 1))
            dof = ns + np.concatenate([np.arange(E)[:, None] * 3 + np.arange(3),
                                       nu + np.arange(E)[:, None] * 3 + np.arange(3)], axis=1)  # (E, 6)
            r_list.append(np.repeat(dof, 6, axis=1).ravel())
            c_list.append(np.tile(dof, (1, 6)).ravel())
            v_list.append(Sinv.ravel())
        else:
            dg = Mh.diagonal()[ns:]
            r_list.append(np.arange(ns, self.n))
            c_list.append(np.arange(ns, self.n))
            v_list.append(1.0 / dg)
        M = sps.csr_matrix((np.concatenate(v_list), (np.concatenate(r_list), np.concatenate(c_list))),
                           shape=(self.n, self.n))
        M.sort_indices()
        pc = (to_dev(M.indptr.astype(np.int32), device=self.device), to_dev(M.indices.astype(np.int32), device=self.device),
              to_dev(M.data.astype(np.float32), device=self.device))
        return None, pc

    def _mass_schur_pc_device(self):
        """The "mass_schur" preconditioner built on the device: its CSR pattern (9 entries
        per stress row, the row's face block; 6 per u / t row, the cell's u and t DoFs) is
        set up once, its values come from the current FP64 values of A."""
        self._pc_buffers()
        d = self._pc_dev
        ns, nu, n = self.n_s, self.u_sp.n_dof, self.n
        rpA, colA, _, v64 = self.full("A")
        d["status"].zero_()
        nat.call("nfx_elasticity_mass_schur_pc", n, ns, nu, nat.ptr(rpA), nat.ptr(colA), nat.ptr(v64),
                 nat.ptr(d["dinv"]), nat.ptr(d["val"]), nat.ptr(d["status"]), nat.stream_handle())
        st = int(d["status"].item())
        if st:
            raise RuntimeError(f"elasticity preconditioner: non-SPD block (status {st})")
        return d["rowptr"], d["col"], d["val"]

    def _pc_buffers(self):
        t = torch()
        ns, nu, n = self.n_s, self.u_sp.n_dof, self.n
        if getattr(self, "_pc_dev", None) is None:
            if ns % 9 or nu % 3 or n != ns + 2 * nu:
                raise ValueError("unexpected elasticity DoF layout")
            rp = np.concatenate([9 * np.arange(ns + 1), 9 * ns + 6 * np.arange(1, n - ns + 1)]).astype(np.int32)
            cols_s = (np.arange(ns)[:, None] // 9) * 9 + np.arange(9)[None, :]
            cell = (np.arange(n - ns) % nu) // 3
            cols_ut = np.concatenate([ns + cell[:, None] * 3 + np.arange(3), ns + nu + cell[:, None] * 3 + np.arange(3)],
                                     axis=1)
            col = np.concatenate([cols_s.ravel(), cols_ut.ravel()]).astype(np.int32)
            self._pc_dev = {"rowptr": to_dev(rp, device=self.device), "col": to_dev(col, device=self.device),
                            "val": t.empty(len(col), dtype=t.float32, device=self.device),
                            "dinv": t.empty(max(ns // 9, 1) * 81, dtype=t.float64, device=self.device),
                            "status": t.zeros(1, dtype=t.int32, device=self.device)}

    def twogrid(self, agg_h=None, omega=None):
        """The two-grid block preconditioner of this system (built once, set up per solve);
        agg_h / omega rebuild it with other aggregates / smoother damping (None: keep the
        current one, or the defaults of ``ElasticityTwoGrid`` for the first build)."""
        tg = getattr(self, "_tg", None)
        if tg is None or (agg_h is not None and tg.h != agg_h):
            kw = {} if omega is None else {"omega": omega}
            tg = self._tg = ElasticityTwoGrid(self, agg_h=agg_h, **kw)
        elif omega is not None:
            tg.omega = float(omega)
        return tg

    def solve(self, rtol=1e-8, inner_rtol=None, max_inner=50000, max_outer=30, pc_kind="twogrid"):
        """MINRES + FP64 refinement on A x = -rg to ||b - A x|| / ||b|| <= rtol (FP64). pc_kind
        "twogrid" (default when the DoF layout allows it): ElasticityTwoGrid; "mass_schur" /
        "riesz_block": block-Jacobi preconditioners (``preconditioner``). inner_rtol: stopping
        tolerance of each FP32 MINRES on its preconditioned residual estimate (default 3e-4
        with the two-grid preconditioner: fewer iterations in total over the FP64 refinement
        steps, measured at 16^3 x 6; 1e-5 otherwise)."""
        import time

        from .solvers import MinresIR

        t = torch()
        t0 = time.perf_counter()
        if pc_kind == "twogrid" and not self.use_faces:
            pc_kind = "mass_schur"
        if inner_rtol is None:
            inner_rtol = 3e-4 if pc_kind == "twogrid" else 1e-5
        tg = None
        if pc_kind == "twogrid":
            tg = self.twogrid()
            tg_desc = tg.setup()
            pc = None
        else:
            sign, pc = self.preconditioner(kind=pc_kind)
        rp, col, v32, v64 = self.full("A")
        b3 = self.ops["A"].get("block3")
        block3 = None
        if b3 is not None and self.use_block3:
            # the block copy of the FP32 values (explicit zeros inside blocks stay zero)
            nat.call("nfx_md_scatter", int(v32.numel()), nat.ptr(v32), nat.ptr(b3["dest"]), 0, nat.ptr(b3["bval"]),
                     nat.stream_handle())
            block3 = (b3["n_blocks"], b3["browptr"], b3["bcol"], b3["bval"])
        solver = MinresIR(rp, col, v32, v64, sign=None, pc=pc, block3=block3, tg=tg_desc if tg else None)
        t.cuda.synchronize()
        t1 = time.perf_counter()
        x, info = solver.solve(-self.rhs64, rtol=rtol, inner_rtol=inner_rtol, max_inner=max_inner,
                               max_outer=max_outer)
        solver.close()
        info["pc"] = pc_kind
        # wall time of the preconditioner set-up (values, blocks, coarse inverse) and of the
        # iterations (the solve call synchronises)
        info["setup_s"] = t1 - t0
        info["iterate_s"] = time.perf_counter() - t1
        return x, info


class ElasticityTwoGrid:
    """Two-grid block preconditioner of an ``ElasticitySystem`` (``nfx_elasticity_tg``).

    P^-1 = diag(D^-1, TG(S)): D the 9 x 9 face blocks of the compliance block of A, S =
    B D^-1 B^T on the cell unknowns (u, t: 6 per cell), TG one symmetric two-grid cycle
    with the block-Jacobi smoother omega S_cc^-1 (omega 0.8: 0.7 needs 10 % more iterations
    at 16^3 x 6, 1.0 twice as many, 1.2 loses positive definiteness) and piecewise-constant
    aggregation of cells into boxes of side ``agg_h`` (default: twice the edge of the cube
    holding six tetrahedra of the mean volume; three times needs 40 % more iterations). The reference preconditions with the Riesz map
    instead (exact Cholesky of its stress block, ICC on the rest,
    ``examples/cauchy_elasticity/mfe_elasticity_example_1.py:238-269``); both are
    block-diagonal SPD preconditioners of MINRES, and the residual the solve is judged by
    is the FP64 residual of A either way."""

    def __init__(self, sysd, agg_h=None, omega=0.8):
        from .learned import aggregate_cells, coupling_plan

        t = torch()
        if not sysd.use_faces:
            raise NotImplementedError("two-grid preconditioner: face-owner DoF layout required")
        self.sys = sysd
        self.omega = float(omega)
        dev = sysd.device
        lay = sysd.face_layout
        E, ns, nu = sysd.E, sysd.n_s, sysd.u_sp.n_dof
        efac = lay["el_face"].astype(np.int64)
        fi = lay["face_inc"].astype(np.int64)
        e_idx = np.arange(E)[:, None]
        other = np.where(fi[efac, 0] == e_idx, fi[efac, 1], fi[efac, 0])
        nbrk = np.zeros((E, 4), np.int64)
        has = other >= 0
        on = np.where(has, other, 0)
        match = efac[on] == efac[:, :, None]  # (E, 4, 4): the neighbour's local face of face k
        nbrk[has] = np.argmax(match, axis=2)[has]
        if np.any(has & ~match.any(axis=2)):
            raise RuntimeError("two-grid preconditioner: inconsistent face incidences")
        # B blocks (E, 4, 6, 9) from the signed element couplings: local stress DoF l of
        # element e sits at face block dest // 9, position dest % 9
        dest = sysd.ref_dest[:, :36].astype(np.int64)
        fblk, fpos = dest // 9, dest % 9
        kof = np.argmax(fblk[:, :, None] == efac[:, None, :], axis=2)  # (E, 36) local face of l
        bblk = np.zeros((E, 4, 6, 9))
        ee = np.repeat(np.arange(E), 36)
        for i in range(6):
            np.add.at(bblk, (ee, kof.ravel(), i, fpos.ravel()), sysd._b_ref[:, i, :].ravel())
        # aggregates
        vol = sysd.geom64[:, -1].double().cpu().numpy() / 6.0
        cent = sysd.xq.mean(axis=1)
        h = float(agg_h) if agg_h is not None else 2.0 * (6.0 * vol.mean()) ** (1.0 / 3.0)
        agg_of, n_agg = aggregate_cells(cent, h)
        order = np.argsort(agg_of, kind="stable")
        agg_ptr = np.concatenate([[0], np.cumsum(np.bincount(agg_of, minlength=n_agg))])
        self.n_agg, self.h = int(n_agg), h
        nc = 6 * self.n_agg
        # coarse coupling pattern (aggregate pairs joined by a face) for the blocked inversion
        a0 = np.repeat(agg_of, 4)[other.ravel() >= 0]
        a1 = agg_of[other.ravel()[other.ravel() >= 0]]
        pa = np.unique(np.stack([a0, a1], 1), axis=0)
        rr = (6 * pa[:, 0, None] + np.arange(6)[None, :])[:, :, None] + np.zeros((1, 1, 6), np.int64)
        cc = (6 * pa[:, 1, None] + np.arange(6)[None, :])[:, None, :] + np.zeros((1, 6, 1), np.int64)
        self.plan = coupling_plan(nc, rr.ravel(), cc.ravel())
        i32 = lambda a: to_dev(np.ascontiguousarray(a).astype(np.int32), device=dev)
        # cell order (set-up kernels) and aggregate order (the application: position m holds
        # cell order[m], the positions of one aggregate are contiguous)
        pos = np.empty(E, np.int64)
        pos[order] = np.arange(E)
        other_o = other[order]
        nbr_m = np.where(other_o >= 0, pos[np.where(other_o >= 0, other_o, 0)], -1)
        self.efac, self.nbr, self.nbrk = i32(efac), i32(other), i32(nbrk)
        self.bblk = to_dev(bblk.reshape(-1), device=dev)
        self.agg_of, self.agg_ptr, self.agg_mem = i32(agg_of), i32(agg_ptr), i32(order)
        self.pos, self.nbr_m, self.agg_of_m = i32(pos), i32(nbr_m), i32(agg_of[order])
        f64 = lambda m: t.empty(m, dtype=t.float64, device=dev)
        f32 = lambda m: t.empty(m, dtype=t.float32, device=dev)
        self.part, self.S64 = f64(E * 4 * 36), f64(E * 180)
        self.S32, self.W32 = f32(E * 180), f32(E * 36)
        self.Ac64 = t.empty((nc, nc), dtype=t.float64, device=dev)
        self.Ainv = t.empty((nc, nc), dtype=t.float32, device=dev)
        self.dinv32 = f32(max(ns // 9, 1) * 81)
        self.work = f32(12 * E + 12 * self.n_agg)
        self.status = t.zeros(1, dtype=t.int32, device=dev)
        self.E, self.ns, self.nu = E, ns, nu
        self._desc = None
        # buffers of the solve (full CSR copies, face-block inverses) and one inversion of
        # the identity: allocations and first launches belong to the set-up of the mesh
        from .learned import spd_inverse_

        sysd._full_buffers("A")
        sysd._pc_buffers()
        self.Ainv.copy_(t.eye(nc, dtype=t.float32, device=dev))
        spd_inverse_(self.Ainv, self.plan)

    def setup(self):
        """Blocks, smoother and coarse inverse from the current FP64 values of A."""
        from .learned import spd_inverse_

        s = nat.stream_handle()
        self.sys._mass_schur_pc_device()  # FP64 face-block inverses (and the block-Jacobi values)
        dinv64 = self.sys._pc_dev["dinv"]
        self.dinv32.copy_(dinv64)
        self.status.zero_()
        nat.call("nfx_elasticity_tg_blocks", self.E, nat.ptr(self.efac), nat.ptr(self.nbr), nat.ptr(self.nbrk),
                 nat.ptr(self.bblk), nat.ptr(dinv64), nat.ptr(self.pos), nat.ptr(self.part), nat.ptr(self.S64),
                 nat.ptr(self.S32), nat.ptr(self.W32), self.omega, nat.ptr(self.status), s)
        self.Ac64.zero_()
        nat.call("nfx_elasticity_tg_coarse", self.n_agg, nat.ptr(self.agg_ptr), nat.ptr(self.agg_mem),
                 nat.ptr(self.agg_of), nat.ptr(self.nbr), nat.ptr(self.S64), nat.ptr(self.Ac64), s)
        if int(self.status.item()):
            raise RuntimeError("elasticity two-grid: a cell block of S is not SPD")
        self.Ainv.copy_(self.Ac64)
        spd_inverse_(self.Ainv, self.plan)
        self._desc = nat.ElasticityTG(self.ns, self.nu, self.E, self.n_agg, nat.ptr(self.dinv32),
                                      nat.ptr(self.nbr_m), nat.ptr(self.S32), nat.ptr(self.W32),
                                      nat.ptr(self.agg_ptr), nat.ptr(self.agg_mem), nat.ptr(self.agg_of_m),
                                      nat.ptr(self.Ainv), 1.0, nat.ptr(self.work))
        return self._desc

    def __call__(self, r, out=None):
        import ctypes

        z = out if out is not None else torch().empty_like(r)
        nat.call("nfx_elasticity_tg_apply", ctypes.byref(self._desc), nat.ptr(r), nat.ptr(z), nat.stream_handle())
        return z


def _local_matrices_host(el, points, weights, X, lam_q, mu_q, mode):
    """(E, NL, NL) FP64 element matrices in the reference basis order with the facet-owner
    kernel's formulas (the cold pass of ``ElasticitySystem2D``: the value filter and the
    patterns are decided on them). X (E, dim + 1, dim) vertices, lam_q / mu_q (E, nq)."""
    dim = X.shape[2]
    tab = elasticity_tables(el, points, weights)
    nq, NB = len(weights), el.tabulate(points).shape[2]
    z = el_sizes(NB * dim + dim + (3 if dim == 3 else 1))
    NC, NT, NSS, NL_ = z["NC"], z["NT"], z["NSS"], z["NL"]
    o = nq
    phi = tab[o:o + nq * NB * dim].reshape(nq, NB, dim)
    o += nq * NB * dim + nq * NB
    uref = tab[o:o + NB]
    Phi = tab[o + NB:o + NB + NB * dim].reshape(NB, dim)
    ddref = tab[o + NB + NB * dim:o + NB + NB * dim + NB * NB].reshape(NB, NB)
    J = np.transpose(X[:, 1:] - X[:, :1], (0, 2, 1))  # J[c, d] = X[d + 1, c] - X[0, c]
    det = np.abs(np.linalg.det(J))
    psi = np.einsum("ecd,qbd->ebcq", J, phi) / det[:, None, None, None]
    k1 = weights[None, :] * det[:, None] / (2 * mu_q)
    k2 = k1 * lam_q / (2 * mu_q + dim * lam_q)
    E = len(det)
    K = np.zeros((E, NL_, NL_))
    D = np.einsum("eacq,ebcq,eq->eab", psi, psi, k1)
    V = np.einsum("eacq,ebdq,eq->eacbd", psi, psi, k2)
    ss = -V
    for c in range(NC):
        ss[:, :, c, :, c] += D
        if mode == 1:
            ss[:, :, c, :, c] += ddref[None] / det[:, None, None]
    K[:, :NSS, :NSS] = ss.reshape(E, NSS, NSS)
    if mode == 0:
        A = np.einsum("ekx,bx->ebk", J, Phi)  # A[e, b, k] = sum_x J_kx Phi_bx
        for b in range(NB):
            for d in range(NC):
                K[:, NC * b + d, NSS + d] = uref[b]
                for r in range(NT):
                    ct = r if dim == 3 else 2  # the rotation about axis ct (2-D: z)
                    if ct == d:
                        continue
                    k = 3 - d - ct
                    even = (ct, d) in ((0, 1), (1, 2), (2, 0))
                    K[:, NC * b + d, NSS + NC + r] = A[:, b, k] if even else -A[:, b, k]
        K[:, NSS:, :NSS] = np.transpose(K[:, :NSS, NSS:], (0, 2, 1))
    else:
        for r in range(NSS, NL_):
            K[:, r, r] = det * weights.sum()
    return K, det


def _host_block_preconditioner(Mh, n_s, rows, cell_dofs):
    """Block-diagonal SPD preconditioner as a scipy CSR: the inverted ROWS x ROWS facet blocks
    of the compliance block on the stress rows, the inverted cell blocks of S = B D^-1 B^T
    (B the cell rows' stress columns, D the facet blocks) on the cell rows ``cell_dofs``
    (E, k), both symmetrised."""
    import scipy.sparse as sps

    nb = n_s // rows
    blocks = np.zeros((nb, rows, rows))
    Ms = Mh[:n_s, :n_s].tocoo()
    same = (Ms.row // rows) == (Ms.col // rows)
    blocks[Ms.row[same] // rows, Ms.row[same] % rows, Ms.col[same] % rows] = Ms.data[same]
    Dinv = np.linalg.inv(blocks)
    Dinv = 0.5 * (Dinv + Dinv.transpose(0, 2, 1))
    Dm = sps.block_diag(list(Dinv), format="csr") if nb else sps.csr_matrix((0, 0))
    E, k = cell_dofs.shape
    B = Mh[cell_dofs.ravel()][:, :n_s].tocsr()
    S = (B @ Dm @ B.T).tocsr()
    Sb = np.stack([S[k * c:k * c + k, k * c:k * c + k].toarray() for c in range(E)]) if E else np.zeros((0, k, k))
    Sinv = np.linalg.inv(Sb)
    Sinv = 0.5 * (Sinv + Sinv.transpose(0, 2, 1))
    rows_b = np.repeat(np.arange(n_s), rows)
    cols_b = ((np.arange(n_s) // rows) * rows)[:, None] + np.arange(rows)[None, :]
    r_list = [rows_b, np.repeat(cell_dofs, k, axis=1).ravel()]
    c_list = [cols_b.ravel(), np.tile(cell_dofs, (1, k)).ravel()]
    v_list = [Dinv.reshape(nb, rows, rows).reshape(-1), Sinv.ravel()]
    n = Mh.shape[0]
    M = sps.csr_matrix((np.concatenate(v_list), (np.concatenate(r_list), np.concatenate(c_list))), shape=(n, n))
    M.sort_indices()
    return M


class ElasticitySystem2D:
    """Dual-mixed elasticity on triangles (BDM1 stress, P0 displacement, P0 rotation about z)
    resident on the GPU: the 2-D case of the reference's elasticity driver
    (``examples/cauchy_elasticity/mfe_elasticity_example_1.py:68-294`` with ``dimension = 2``,
    ``:729``). Operators A (``LEDualWeakForm``) and P (``LERieszMapWeakForm``) stored like
    PETSc SBAIJ (upper triangle, the reference's value filter, explicit zero diagonals for
    A), residual with the stress Dirichlet term; warm assemblies by the facet-owner kernel
    (``nfx_assemble_elasticity_faces`` with dim 2: one wavefront per edge owns its 4 CSR
    rows); the cold pass (element matrices for the filter and the patterns) on the host in
    FP64. Solve: FP32 MINRES with the block-diagonal preconditioner (inverted edge blocks of
    the compliance block, inverted cell blocks of B D^-1 B^T) inside FP64 refinement."""

    def __init__(self, space, functions, device="cuda"):
        t = torch()
        self.space = space
        self.device = device
        names = space.names
        self.s_sp, self.u_sp, self.t_sp = (space.discrete_spaces[n] for n in names)
        s_sp = self.s_sp
        mesh = s_sp.mesh
        if s_sp.dimension != 2 or s_sp.element.family != "BDM" or s_sp.k_order != 1 or s_sp.n_comp != 2:
            raise NotImplementedError("2-D GPU elasticity path: BDM1 - P0 - P0 on triangles")
        self.dim = 2
        self.E = len(s_sp.element_ids)
        self.n = space.n_dof
        self.n_s = s_sp.n_dof
        self.field_bounds = [0, self.n_s, self.n_s + self.u_sp.n_dof, self.n]
        points, weights = space.quadrature[2]
        if len(weights) != 4:
            raise ValueError("expected the degree-3 triangle rule (4 points)")
        self.nq = 4
        cells = mesh.entities[2].nodes[s_sp.topology.cell_index].astype(np.int64)
        if cells.min() < 0 or cells.max() >= len(mesh.points):
            raise ValueError("cells out of range")
        X = mesh.points[cells]
        if np.abs(X[:, :, 2]).max() > 0:
            raise NotImplementedError("2-D elasticity: the mesh must lie in the plane z = 0")
        self.xq = np.einsum("qa,eac->eqc", p1_bary(points), X)
        X2 = X[:, :, :2]
        J = np.transpose(X2[:, 1:] - X2[:, :1], (0, 2, 1))
        det = np.abs(np.linalg.det(J))
        if self.E and det.min() <= 0:
            raise ValueError("degenerate element")
        geom = np.concatenate([(J / det[:, None, None]).reshape(self.E, 4), det[:, None]], axis=1)
        self.geom64 = to_dev(np.ascontiguousarray(geom), device=device)
        self.geom32 = to_dev(np.ascontiguousarray(geom.astype(np.float32)), device=device)
        tab = elasticity_tables(s_sp.element, points, weights)
        self.tables64 = to_dev(tab, device=device)
        self.tables32 = to_dev(tab.astype(np.float32), device=device)
        ref_dest, ref_sign = space.reference_destinations()
        if ref_dest.shape[1] != 15:
            raise ValueError("unexpected local size")
        self.ref_dest, self.ref_sign = ref_dest, ref_sign
        self.functions = functions
        self._coef = self._coefficients(functions)
        lq = _expand_coef(self._coef["lam"], self.E, self.nq)
        mq = _expand_coef(self._coef["mu"], self.E, self.nq)
        # cold pass on the host: FP64 element matrices -> value filter -> SBAIJ patterns
        self._K = {}
        KA, _ = _local_matrices_host(s_sp.element, points, weights, X2, lq, mq, 0)
        KP, _ = _local_matrices_host(s_sp.element, points, weights, X2, lq, mq, 1)
        self._K = {0: KA, 1: KP}
        keepA = _sbaij_mask(KA, ref_dest, ref_sign, upper=False)
        keepP = _sbaij_mask(KP, ref_dest, ref_sign, upper=False)
        self.pattern_A = build_pattern(self.n, ref_dest, ref_sign, keep=keepA, zero_diag_dest=ref_dest, upper=True)
        self.pattern_P = build_pattern(self.n, ref_dest, ref_sign, keep=keepP, upper=True)
        lay = face_owner_layout(self.n, ref_dest, ref_sign)
        self.face_layout = lay
        codes = {"A": face_owner_codes(self.n, lay, ref_dest, keepA, self.pattern_A, zero_diag=True),
                 "P": face_owner_codes(self.n, lay, ref_dest, keepP, self.pattern_P, zero_diag=False)}
        self.ops = {}
        for key, pat in (("A", self.pattern_A), ("P", self.pattern_P)):
            frp, fcol, fsrc = full_from_upper(self.n, pat.rowptr, pat.col)
            self.ops[key] = {"pattern": pat, "rowptr": to_dev(pat.rowptr, device=device),
                             "col": to_dev(pat.col, device=device), "codes_np": codes[key],
                             "codes": to_dev(codes[key].view(np.int16), device=device),
                             "vals32": t.zeros(pat.nnz, dtype=t.float32, device=device),
                             "vals64": t.zeros(pat.nnz, dtype=t.float64, device=device),
                             "full_rowptr": to_dev(frp, device=device), "full_col": to_dev(fcol, device=device),
                             "full_src": to_dev(fsrc, device=device), "full_nnz": int(len(fcol))}
        self.face_inc = to_dev(lay["face_inc"].astype(np.int32), device=device)
        self.face_lf = to_dev(lay["face_lf"].astype(np.int32), device=device)
        self.el_face = to_dev(lay["el_face"].astype(np.int32), device=device)
        self.el_orient = to_dev(lay["el_orient"].view(np.int64), device=device)
        self.rhs32 = t.zeros(self.n, dtype=t.float32, device=device)
        self.rhs64 = t.zeros(self.n, dtype=t.float64, device=device)
        self.cells = to_dev(cells.astype(np.int32), device=device)
        self.coords = to_dev(mesh.points.astype(np.float32), device=device)
        self.coords64 = to_dev(mesh.points, device=device)
        self._setup_bc(mesh, device)
        self._validate()

    def _coefficients(self, functions):
        lam = coefficient(functions["lambda"], self.xq, self.nq, device=self.device)
        mu = coefficient(functions["mu"], self.xq, self.nq, device=self.device)
        if (lam.se, lam.sq) != (mu.se, mu.sq):
            lam = CoefficientArg(_expand_coef_dev(lam, self.E, self.nq), self.nq, 1)
            mu = CoefficientArg(_expand_coef_dev(mu, self.E, self.nq), self.nq, 1)
        check_coefficient(lam, self.E, self.nq, "lambda")
        check_coefficient(mu, self.E, self.nq, "mu")
        f = functions.get("rhs")
        fv = _vector_coefficient(f, self.xq, self.nq, 3, self.device) if f is not None else \
            torch().zeros((self.E, self.nq, 3), dtype=torch().float64, device=self.device)
        if fv.shape != (self.E, self.nq, 3):
            raise ValueError("rhs must give 3 components (2 force + 1 rotation) per point")
        return {"lam": lam, "mu": mu, "f": CoefficientArg(fv.reshape(-1), self.nq * 3, 3)}

    def _setup_bc(self, mesh, device):
        s_sp = self.s_sp
        self.F = len(s_sp.bc_element_ids)
        if not self.F:
            return
        fpts, fw = self.space.bc_quadrature[1]
        self.nqf = len(fw)
        if self.nqf != 2:
            raise ValueError("expected the 2-point edge rule")
        neigh_cells = s_sp.topology.cell_index[s_sp.bc_neighbour]
        combo = facet_combo_index(mesh, 2, s_sp.bc_element_ids, neigh_cells, s_sp.bc_local)
        trace = facet_combo_tables(s_sp.element, 2, fpts)
        fnodes = mesh.entities[1].nodes[s_sp.bc_element_ids - mesh.entities[1].offset]
        self.bc_facet = to_dev(fnodes.astype(np.int32), device=device)
        self.bc_neigh = to_dev(s_sp.bc_neighbour.astype(np.int32), device=device)
        self.bc_combo = to_dev(combo.astype(np.int32), device=device)
        self.bc_trace64 = to_dev(trace, device=device)
        self.bc_trace32 = to_dev(trace.astype(np.float32), device=device)
        self.bc_fw64 = to_dev(fw, device=device)
        self.bc_fw32 = to_dev(fw.astype(np.float32), device=device)
        shift = self.space.strides[self.space.names[0]] + self.space.dof_shift
        self.bc_rows_np = row_codes(s_sp.bc_ref_dest + shift, s_sp.bc_ref_sign)
        self.bc_rows = to_dev(self.bc_rows_np, device=device)
        self.bc_xq = np.einsum("qa,fac->fqc", p1_bary(fpts), mesh.points[fnodes])

    def _validate(self):
        lay, E = self.face_layout, self.E
        nf = lay["n_faces"]
        fi = lay["face_inc"]
        if fi.shape != (nf, 2) or fi[:, 0].min(initial=0) < 0 or fi.max(initial=-1) >= E or \
                np.any(fi[:, 1] < -1) or np.any((fi[:, 1] >= 0) & (fi[:, 1] <= fi[:, 0])):
            raise ValueError("facet incidences out of range or not in element order")
        if lay["el_face"].shape != (E, 3) or lay["el_face"].min(initial=0) < 0 or \
                lay["el_face"].max(initial=-1) >= nf or np.any(lay["face_lf"] >> 4):
            raise ValueError("element facet blocks out of range")
        if self.tables32.numel() != 4 + 4 * 12 + 4 * 6 + 6 + 12 + 36:
            raise ValueError("elasticity table size")
        for key in ("A", "P"):
            op = self.ops[key]
            code = op["codes_np"].astype(np.int64)
            rp = op["pattern"].rowptr
            if len(code) != int(rp[self.n_s]) or len(rp) != self.n + 1:
                raise ValueError(f"facet codes of {key}: size")
            s0, s1 = (code >> 4) & 63, code >> 10
            if np.any((code & 15) >= 4) or np.any((s0 >= 15) & (s0 != FACE_NO_SRC)) or \
                    np.any((s1 >= 15) & (s1 != FACE_NO_SRC)):
                raise ValueError(f"facet codes of {key} out of range")
            p_of = np.repeat(np.arange(self.n_s) // 4, np.diff(rp[: self.n_s + 1]))
            if np.any((s1 != FACE_NO_SRC) & (fi[p_of, 1] < 0)):
                raise ValueError(f"facet codes of {key}: a second source on a one-element facet")
        if self.F:
            rr = self.bc_rows_np.astype(np.int64)
            if rr.shape != (self.F, 4) or np.any((rr >= 0) & ((rr & nat.SLOT_INDEX_MASK) >= self.n)):
                raise ValueError("facet rows out of range")

    def local_matrices(self, mode):
        """The cold pass's FP64 element matrices (reference order) and element residuals."""
        K = self._K[mode]
        r = np.zeros((self.E, 15))
        if mode == 0:
            f = self._coef["f"].get(True).reshape(self.E, self.nq, 3).cpu().numpy()
            pts, w = self.space.quadrature[2]
            det = self.geom64[:, -1].cpu().numpy()
            r[:, 12:] = -np.einsum("q,e,eqc->ec", w, det, f)
        return K, r

    def assemble(self, which="A", fp64=False, stream=None, u_dirichlet=None):
        op = self.ops[which]
        s = nat.stream_handle(stream)
        vals = op["vals64" if fp64 else "vals32"]
        rg = self.rhs64 if fp64 else self.rhs32
        mode = 0 if which == "A" else 1
        c = self._coef
        lam, mu, f = c["lam"], c["mu"], c["f"]
        lay = self.face_layout
        nat.call("nfx_assemble_elasticity_faces", 2, mode, lay["n_faces"], self.E, nat.ptr(self.face_inc),
                 nat.ptr(self.face_lf), nat.ptr(self.el_face), nat.ptr(self.el_orient),
                 nat.ptr(self.geom64 if fp64 else self.geom32), nat.ptr(self.tables64 if fp64 else self.tables32),
                 nat.ptr(lam.get(fp64)), nat.ptr(mu.get(fp64)), lam.se, lam.sq, nat.ptr(f.get(fp64)), self.nq * 3, 3,
                 1, nat.ptr(op["rowptr"]), nat.ptr(op["codes"]), int(fp64), nat.ptr(vals), nat.ptr(rg), s)
        if mode == 0 and u_dirichlet is not None and self.F:
            self.assemble_dirichlet(u_dirichlet, fp64=fp64, stream=stream)
        return vals, rg

    def assemble_dirichlet(self, u_dirichlet, fp64=False, stream=None, rows=None, out=None):
        uD = _vector_coefficient(u_dirichlet, self.bc_xq, self.nqf, 2, self.device)
        if uD.shape != (self.F, self.nqf, 2):
            raise ValueError("u_D must give 2 components per facet point")
        rg = out if out is not None else (self.rhs64 if fp64 else self.rhs32)
        g = (uD if fp64 else uD.to(torch().float32)).contiguous()
        self._held = g
        nat.call("nfx_assemble_elasticity_bc", 2, self.F, nat.ptr(self.bc_facet), nat.ptr(self.bc_neigh),
                 nat.ptr(self.bc_combo), nat.ptr(self.cells), nat.ptr(self.coords64 if fp64 else self.coords),
                 nat.ptr(self.bc_trace64 if fp64 else self.bc_trace32),
                 nat.ptr(self.bc_fw64 if fp64 else self.bc_fw32), self.nqf, nat.ptr(g),
                 nat.ptr(self.bc_rows if rows is None else rows), int(fp64), nat.ptr(rg), nat.stream_handle(stream))

    def full(self, which="A"):
        """Full symmetric (rowptr, col, vals32, vals64) for the Krylov solver."""
        t = torch()
        op = self.ops[which]
        n = op["full_nnz"]
        if "full32" not in op:
            op["full32"] = t.empty(n, dtype=t.float32, device=self.device)
            op["full64"] = t.empty(n, dtype=t.float64, device=self.device)
        s = nat.stream_handle()
        nat.call("nfx_gather", 0, n, nat.ptr(op["full_src"]), nat.ptr(op["vals32"]), nat.ptr(op["full32"]), s)
        nat.call("nfx_gather", 1, n, nat.ptr(op["full_src"]), nat.ptr(op["vals64"]), nat.ptr(op["full64"]), s)
        return op["full_rowptr"], op["full_col"], op["full32"], op["full64"]

    def csr_numpy(self, which="A", fp64=True):
        import scipy.sparse as sps

        op = self.ops[which]
        pat = op["pattern"]
        vals = op["vals64" if fp64 else "vals32"].cpu().numpy()
        return sps.csr_matrix((vals, pat.col, pat.rowptr), shape=(self.n, self.n))

    def preconditioner(self):
        """(rowptr, col, val) FP32 CSR of the block-diagonal preconditioner (built on the host
        from the current FP64 values of A)."""
        import scipy.sparse as sps

        rp, col, _, v64 = self.full("A")
        Mh = sps.csr_matrix((v64.cpu().numpy(), col.cpu().numpy(), rp.cpu().numpy()), shape=(self.n, self.n))
        nu = self.u_sp.n_dof
        E = self.E
        cell_dofs = np.concatenate([self.n_s + 2 * np.arange(E)[:, None] + np.arange(2),
                                    (self.n_s + nu + np.arange(E))[:, None]], axis=1)
        M = _host_block_preconditioner(Mh, self.n_s, 4, cell_dofs)
        return (to_dev(M.indptr.astype(np.int32), device=self.device),
                to_dev(M.indices.astype(np.int32), device=self.device),
                to_dev(M.data.astype(np.float32), device=self.device))

    def solve(self, rtol=1e-8, inner_rtol=1e-5, max_inner=100000, max_outer=30):
        from .solvers import MinresIR

        pc = self.preconditioner()
        rp, col, v32, v64 = self.full("A")
        solver = MinresIR(rp, col, v32, v64, sign=None, pc=pc)
        x, info = solver.solve(-self.rhs64, rtol=rtol, inner_rtol=inner_rtol, max_inner=max_inner,
                               max_outer=max_outer)
        solver.close()
        info["pc"] = "mass_schur"
        return x, info


def _expand_coef_dev(c, E, nq):
    a = c.get(True)
    if c.se == 0:
        return a.expand(E * nq).contiguous()
    if c.sq == 0:
        return a.reshape(-1, 1).expand(E, nq).reshape(-1).contiguous()
    return a


def _expand_coef(c, E, nq):
    """(E, nq) FP64 host array of a coefficient argument."""
    return _expand_coef_dev(c, E, nq).reshape(E, nq).cpu().numpy()
