"""Device-resident systems: mesh/space arrays in HBM and the warm assembly calls.

``DarcySystem`` holds, for one ``ProductSpace`` with an RT_k flux field and a
discontinuous P_{k-1} pressure field (the spaces built by
``examples/md_elliptic_scalar/md_elliptic_scalar.py:44-86``), everything the
HIP kernels read:

  cells (E, D+1) int32, coords (V, 3) f32/f64, reference tables (f32 and f64),
  slot map (nl*nl, E) int32, pressure rows (np, E), CSR rowptr/col int32,
  stage segments, and the boundary-facet arrays of the Dirichlet term.

``assemble`` is the warm path: coefficient arrays in, CSR values + residual out
(FP32 by default, FP64 on request), pattern reused.
"""

from __future__ import annotations

from itertools import permutations

import numpy as np

from . import _native as nat
from .basis import create_element
from .mesh import TET_FACES, TRI_EDGES
from .pattern import build_pattern, row_codes

_TORCH = None


def torch():
    global _TORCH
    if _TORCH is None:
        import torch as _t

        _TORCH = _t
    return _TORCH


def to_dev(a, dtype=None, device="cuda"):
    t = torch()
    if isinstance(a, t.Tensor):
        return a.to(device=device, dtype=dtype) if dtype is not None else a.to(device)
    arr = np.ascontiguousarray(a)
    out = t.from_numpy(arr)
    if dtype is not None:
        out = out.to(dtype)
    return out.to(device)


def p1_bary(points):
    pts = np.asarray(points, dtype=np.float64)
    return np.concatenate([1.0 - pts.sum(axis=1, keepdims=True), pts], axis=1)


def darcy_tables(u_el, p_el, points, weights):
    """w | phi (nq, nb, dim) | div (nq, nb) | psi (nq, np) | Rs (nb, nb, ns) | Bc (nb, np) |
    ck (np) as one FP64 vector.

    The last three are reference tensors of the same quadrature sums, used when the
    coefficients are constant per element:
    Rs[a, b, s] = sum_q w phi_a . (E_s phi_b) over the ns = dim (dim + 1) / 2 symmetric
    pairs s = (d <= d') (E_s the symmetric unit pair), so that the mass row is
    kinv / detJ sum_s G_s Rs[a, b, s] with G = J^T J; Bc[a, k] = sum_q w div_a psi_k;
    ck[k] = sum_q w psi_k."""
    tu = u_el.tabulate(points)
    tp = p_el.tabulate(points)
    phi = tu[0]
    div = np.einsum("kqjk->qj", tu[1:])
    psi = tp[0, :, :, 0]
    dim = phi.shape[2]
    R = np.einsum("q,qax,qby->abxy", weights, phi, phi)
    pairs = [(d, d2) for d in range(dim) for d2 in range(d, dim)]
    Rs = np.stack([R[:, :, d2, d] + (R[:, :, d, d2] if d != d2 else 0.0) for d, d2 in pairs], axis=-1)
    Bc = np.einsum("q,qa,qk->ak", weights, div, psi)
    ck = np.einsum("q,qk->k", weights, psi)
    return np.concatenate([weights, phi.ravel(), div.ravel(), psi.ravel(), Rs.ravel(), Bc.ravel(), ck])


def facet_combo_tables(el, D, fpoints):
    """Reference flux basis of each facet's DoFs at facet points mapped into the cell.

    Combo index = local_facet * D! + index of the facet-vertex order in
    permutations(range(D)), following ``transform_lower_to_higher``
    (``src/basis/parametric_transformation.py:30-98``). Returns (ncombo, nqf, fd, D).
    """
    local = TRI_EDGES if D == 2 else TET_FACES
    verts = np.eye(D + 1, D, -1)
    perms = list(permutations(range(D)))
    s = np.asarray(fpoints, dtype=np.float64)
    bary = p1_bary(s)  # (nqf, D)
    out = []
    for lf in range(D + 1):
        dofs = el.entity_dofs[D - 1][lf]
        for p in perms:
            loc = local[lf][list(p)]
            mapped = bary @ verts[loc]
            tab = el.tabulate(mapped)[0]  # (nqf, nb, D)
            out.append(tab[:, dofs, :])
    return np.array(out)


def facet_combo_index(mesh, D, facet_ids, neigh_cells, local_facet):
    """Combo index of each boundary facet (see ``facet_combo_tables``)."""
    local = TRI_EDGES if D == 2 else TET_FACES
    perms = {p: i for i, p in enumerate(permutations(range(D)))}
    fnodes = mesh.entities[D - 1].nodes[facet_ids - mesh.entities[D - 1].offset]  # (F, D)
    cnodes = mesh.entities[D].nodes[neigh_cells]  # (F, D+1)
    # local vertex index (in the cell) of each facet node
    locv = np.argmax(cnodes[:, None, :] == fnodes[:, :, None], axis=2)  # (F, D)
    base = local[local_facet]  # (F, D) basix order of that facet's local vertices
    pos = np.argmax(base[:, None, :] == locv[:, :, None], axis=2)  # (F, D) position in base
    pidx = np.array([perms[tuple(r)] for r in pos.tolist()], dtype=np.int64) if len(pos) else np.zeros(0, np.int64)
    return local_facet * len(perms) + pidx


class CoefficientArg:
    """A coefficient at quadrature points: device array + (stride_e, stride_q).

    Holds the values in FP64 and hands out an FP32 or FP64 copy matching the
    kernel precision (both cached).
    """

    def __init__(self, arr, se, sq):
        t = torch()
        self.se, self.sq = se, sq
        self._by = {arr.dtype: arr.contiguous()}
        self.arr = arr

    def get(self, fp64):
        t = torch()
        dt = t.float64 if fp64 else t.float32
        if dt not in self._by:
            self._by[dt] = self.arr.to(dt).contiguous()
        return self._by[dt]


def coefficient(value, xq, nq, comp=0, invert=False, device="cuda"):
    """Turn a reference-style coefficient into a kernel argument.

    value: callable f(x, y, z) (vectorised over points; returns a scalar, an
    (n_comp, npts) array or an (npts,) array), a per-element array (E,) or
    (E, nq), a torch tensor of the same shapes, or a number.
    """
    t = torch()
    if callable(value):
        x = xq.reshape(-1, 3)
        v = np.asarray(value(x[:, 0], x[:, 1], x[:, 2]), dtype=np.float64)
        if v.ndim == 0:
            arr = np.array([float(v)])
            se = sq = 0
        else:
            if v.ndim >= 2:
                v = v.reshape(v.shape[0], -1)[comp]
            arr = v.reshape(-1)
            if arr.size == 1:
                se = sq = 0
            else:
                se, sq = nq, 1
        if invert:
            arr = 1.0 / arr
        return CoefficientArg(to_dev(arr.astype(np.float64), device=device), se, sq)
    if isinstance(value, (int, float, np.floating)):
        v = 1.0 / float(value) if invert else float(value)
        return CoefficientArg(t.tensor([v], dtype=t.float64, device=device), 0, 0)
    arr = value if isinstance(value, t.Tensor) else t.from_numpy(np.asarray(value, dtype=np.float64))
    arr = arr.to(device=device, dtype=t.float64)
    if invert:
        arr = 1.0 / arr
    if arr.dim() == 1:
        return CoefficientArg(arr.contiguous(), 1, 0)
    return CoefficientArg(arr.reshape(arr.shape[0], -1).contiguous(), arr.shape[1], 1)


# (dim, k) -> (NB, NP, NQ, FD, NQF) compiled into nfx_assemble_darcy / _bc
DARCY_KERNELS = {(2, 1): (3, 1, 4, 1, 2), (2, 2): (8, 3, 9, 2, 3), (3, 1): (4, 1, 8, 1, 4), (3, 2): (15, 4, 27, 3, 9)}


def _check_codes(codes, n_direct, n_stage, what):
    c = np.asarray(codes).astype(np.int64).reshape(-1)
    live = c[c >= 0]
    idx = live & nat.SLOT_INDEX_MASK
    staged = (live & nat.SLOT_STAGE) != 0
    if np.any(idx[~staged] >= n_direct) or np.any(idx[staged] >= max(n_stage, 0)) or (np.any(staged) and n_stage == 0):
        raise ValueError(f"{what}: slot/row code out of range")


def check_coefficient(arg, E, nq, what):
    n = int(arg.arr.numel())
    last = (E - 1) * arg.se + (nq - 1) * arg.sq
    if E > 0 and last >= n:
        raise ValueError(f"{what}: coefficient array too short ({n} <= {last})")


def row_pos_device(pos, nl, sign_mask, rowptr=None, tile=128):
    """The row kernel's column-position array from the (n, 2, nlp) byte positions (255 =
    not present) and the (n, 2) entry-sign masks (bit b: entry (a, b) of the incident
    element row is negated). nl <= 6 packs each incidence into one uint32: bits 4b..4b+3
    the position of b (15 = not present; rows of at most 12 entries), bits 4 nl + b the
    sign mask, the rest zero. For nl <= 5 with ``rowptr`` given, bits 25..31 of the two
    words of row r hold its CSR offset inside its ``tile``-row tile, rowptr[r] -
    rowptr[tile * (r // tile)] (low 7 bits in word 0, high 7 in word 1). Larger elements keep
    the bytes as they are (the kernel's row_pos_stride)."""
    if nl > 6:
        return pos.reshape(-1)
    p = pos[:, :, :nl].astype(np.uint64)
    if np.any((p != 255) & (p >= 15)):
        raise ValueError("row too long for 4-bit column positions")
    nib = np.where(p == 255, 15, p)
    words = (nib << (4 * np.arange(nl, dtype=np.uint64))[None, None, :]).sum(axis=2, dtype=np.uint64)
    words |= (sign_mask.astype(np.uint64) & np.uint64((1 << nl) - 1)) << np.uint64(4 * nl)
    if nl <= 5:
        # the row kernel reads these offsets unconditionally for such elements
        if rowptr is None:
            raise ValueError("rowptr is required for nl <= 5: the row kernel reads each row's "
                             "in-tile CSR offset from the position words")
        rp = np.asarray(rowptr, dtype=np.int64)
        r = np.arange(len(rp) - 1)
        off = rp[:-1] - rp[(r // tile) * tile]
        if off.size and off.max() >= 1 << 14:
            raise ValueError("row offset inside its tile does not fit 14 bits")
        words[:, 0] |= (off.astype(np.uint64) & np.uint64(0x7F)) << np.uint64(25)
        words[:, 1] |= (off.astype(np.uint64) >> np.uint64(7)) << np.uint64(25)
    return np.ascontiguousarray(words.astype(np.uint32)).view(np.uint8).reshape(-1)


ROW_TILE = 128  # rows per workgroup of the row kernel (kRowsPerBlock)
N_XCD = 8  # accelerator complex dies; workgroups are dispatched to them round-robin


def xcd_tile_map(first_elem, n_rows, tile=ROW_TILE, n_xcd=N_XCD):
    """Launch slots of the row kernel: tiles (runs of ``tile`` consecutive CSR rows)
    ordered by the median element of their rows' first incidences and cut into
    ``n_xcd`` equal runs; run x fills slots x, x + n_xcd, ... (ascending tile order,
    -1 = empty slot). Round-robin dispatch then places the face rows and the cell rows of
    one element range on the same XCD, so each element's metric and coefficient are
    fetched into one L2 instead of up to five."""
    nt = -(-n_rows // tile)
    if nt == 0:
        return np.zeros(0, np.int32)
    idx = np.minimum(np.arange(nt * tile), n_rows - 1).reshape(nt, tile)
    key = np.median(first_elem[idx], axis=1)
    order = np.argsort(key, kind="stable")
    runs = np.array_split(order, n_xcd)
    slots = np.full((len(runs[0]), n_xcd), -1, np.int32)
    for x, run in enumerate(runs):
        slots[: len(run), x] = np.sort(run)
    return slots.reshape(-1)


def slot_positions(pattern, code):
    """CSR value positions of slot codes (any shape): the value index of a direct code, the
    target of its stage segment for a staged one. A dropped entry (code < 0) is an error."""
    code = np.asarray(code, dtype=np.int64)
    if (code < 0).any():
        raise ValueError("an element entry is missing from the pattern")
    idx = code & nat.SLOT_INDEX_MASK
    staged = (code & nat.SLOT_STAGE) != 0
    if staged.any():
        # stage index -> value index of its segment, built once per pattern
        st = getattr(pattern, "_stage_target", None)
        if st is None:
            st = np.repeat(pattern.target.astype(np.int64), np.diff(pattern.segptr.astype(np.int64)))
            pattern._stage_target = st
        idx[staged] = st[idx[staged]]
    return idx


class DarcySystem:
    """Mixed Darcy RT_k - P_{k-1} on one product space, resident on the GPU."""

    def __init__(self, space, flux="u", potential="p", device="cuda", split=None, cold_device=None,
                 pattern=None):
        """cold_device: run the pattern's key sort there ("cuda"; same pattern as the host).
        pattern: the space's pattern built beforehand (``build_pattern`` of its reference
        destinations and signs, e.g. in a worker process; not with split).
        split: optional (m, 2) (element index, local flux DoF) pairs whose flux DoF is
        replaced by a new one of their own (appended after the flux DoFs, pressure DoFs
        shifted): the facet DoF of a face cut by a fracture, on one side of it, so that
        the normal flux may jump across the fracture."""
        t = torch()
        self.space = space
        self.device = device
        u_sp = space.discrete_spaces[flux]
        p_sp = space.discrete_spaces[potential]
        self.u_sp, self.p_sp = u_sp, p_sp
        mesh = u_sp.mesh
        D = u_sp.dimension
        self.dim = D
        self.k = u_sp.k_order
        self.u_el = u_sp.element
        self.p_el = p_sp.element
        self.nb = self.u_el.ndofs
        self.np_ = self.p_el.ndofs
        self.nl = self.nb + self.np_
        self.E = len(u_sp.element_ids)
        self.n = space.n_dof
        self.n_u = u_sp.n_dof
        points, weights = space.quadrature[D]
        self.qpoints, self.qweights = points, weights
        self.nq = len(weights)
        cell_tab = mesh.entities[D]
        cells = cell_tab.nodes[u_sp.topology.cell_index].astype(np.int32)
        self.cells_np = cells
        self.cells = to_dev(cells, device=device)
        self.coords = to_dev(mesh.points.astype(np.float32), device=device)
        self.coords64 = to_dev(mesh.points, device=device)
        # per-element metric of the row kernels (G / detJ, detJ), once per mesh: the warm
        # assembly then reads it instead of the vertex coordinates (use_geometry = False
        # derives it in the kernel; same values)
        self.geo32 = self.geo64 = None
        self.use_geometry = True
        self._launch_cache = {}
        if str(device).startswith("cuda"):
            t = torch()
            gs = 8 if D == 3 else 4
            self.geo32 = t.empty(self.E * gs, dtype=t.float32, device=device)
            self.geo64 = t.empty(self.E * gs, dtype=t.float64, device=device)
            for fl, crd, geo in ((0, self.coords, self.geo32), (1, self.coords64, self.geo64)):
                nat.call("nfx_element_geometry", D, self.E, nat.ptr(self.cells), nat.ptr(crd), fl, nat.ptr(geo),
                         nat.stream_handle())
        tab = darcy_tables(self.u_el, self.p_el, points, weights)
        self.tables64 = to_dev(tab, device=device)
        self.tables32 = to_dev(tab.astype(np.float32), device=device)
        # x_q on the host (FP64): coefficient callables are evaluated there
        X = mesh.points[cells]  # (E, D+1, 3)
        self.xq = np.einsum("qa,eac->eqc", p1_bary(points), X)
        ref_dest, ref_sign = space.reference_destinations()
        self.n_split = 0
        if split is not None and len(split):
            split = np.asarray(split, dtype=np.int64).reshape(-1, 2)
            if np.any(split[:, 0] < 0) or np.any(split[:, 0] >= self.E) or np.any(split[:, 1] < 0) \
                    or np.any(split[:, 1] >= self.nb):
                raise ValueError("split: (element, flux DoF) out of range")
            if len(np.unique(split[:, 0] * self.nb + split[:, 1])) != len(split):
                raise ValueError("split: repeated (element, flux DoF)")
            m, n_u0 = len(split), self.n_u
            ref_dest = ref_dest.copy()
            ref_dest[ref_dest >= n_u0] += m
            ref_dest[split[:, 0], split[:, 1]] = n_u0 + np.arange(m)
            self.n_u += m
            self.n += m
            self.n_split = m
        self.split = split
        self.ref_dest, self.ref_sign = ref_dest, ref_sign
        if pattern is not None:
            if split is not None and len(split):
                raise ValueError("a prebuilt pattern cannot be combined with split DoFs")
            if pattern.n != self.n or pattern.slot.shape != (self.nl * self.nl, self.E):
                raise ValueError("the prebuilt pattern does not belong to this space")
            self.pattern = pattern
        else:
            self.pattern = build_pattern(self.n, ref_dest, ref_sign, device=cold_device)
        pat = self.pattern
        self.rowptr = to_dev(pat.rowptr, device=device)
        self.col = to_dev(pat.col, device=device)
        self.slot = to_dev(pat.slot, device=device)
        self.segptr = to_dev(pat.segptr, device=device)
        self.target = to_dev(pat.target, device=device)
        self.nnz = pat.nnz
        self.rhs_row = to_dev(np.ascontiguousarray(row_codes(ref_dest[:, self.nb:]).T), device=device)
        self.vals32 = t.zeros(self.nnz, dtype=t.float32, device=device)
        self.vals64 = t.zeros(self.nnz, dtype=t.float64, device=device)
        self.stage32 = t.zeros(max(pat.n_stage, 1), dtype=t.float32, device=device)
        self.stage64 = t.zeros(max(pat.n_stage, 1), dtype=t.float64, device=device)
        self.rhs32 = t.zeros(self.n, dtype=t.float32, device=device)
        self.rhs64 = t.zeros(self.n, dtype=t.float64, device=device)
        self._setup_bc(mesh, D, device)
        self._setup_rows(device)
        self._validate(mesh)

    def _setup_rows(self, device):
        """Row-owner gather data: incidences, column positions, sign masks."""
        n, nl, E = self.n, self.nl, self.E
        flat = self.ref_dest.reshape(-1)
        order = np.argsort(flat, kind="stable")  # element order inside each row
        rows_sorted = flat[order]
        counts = np.bincount(rows_sorted, minlength=n)
        if counts.min() < 1 or counts.max() > 2:
            raise ValueError("row-owner assembly needs 1 or 2 incidences per row")
        first = np.concatenate([[0], np.cumsum(counts)[:-1]])
        inc = np.full((2, n), -1, np.int64)
        inc[0] = order[first]
        two = counts == 2
        inc[1, two] = order[first[two] + 1]
        e_of = np.where(inc >= 0, inc // nl, 0)
        a_of = np.where(inc >= 0, inc % nl, 0)
        packed = np.where(inc >= 0, e_of * 32 + a_of, -1)
        if self.F:
            # rows owned by Dirichlet facet f (facet DoF a): second incidence -2 - (f * fd + a)
            bc_r = (self.bc_rows_np & nat.SLOT_INDEX_MASK).reshape(-1)
            if len(np.unique(bc_r)) != bc_r.size or np.any(packed[1, bc_r] != -1):
                raise ValueError("a Dirichlet facet row must have one element incidence and one facet")
            if np.any(e_of[0, bc_r] != np.repeat(self.bc_neigh_np, self.bc_rows_np.shape[1])):
                raise ValueError("a Dirichlet facet row's element must be the facet's neighbour")
            packed[1, bc_r] = -2 - np.arange(bc_r.size)
        # column positions of every reference DoF b of each incident element, inside row r:
        # the CSR index of entry (a, b) of element e is its slot code's value index, or for
        # a staged entry the target of its stage segment
        pat = self.pattern
        rowptr = pat.rowptr.astype(np.int64)
        # bytes per incidence: nl, padded to whole words for the larger elements; the
        # device copy packs nl <= 7 into 4-bit fields (row_pos_device)
        nlp = nl if nl <= 8 else (nl + 3) // 4 * 4
        pos = np.full((n, 2, nlp), 255, np.uint8)
        if (np.diff(rowptr).max(initial=0)) >= 255:
            raise ValueError("row too long for uint8 column positions")
        b_idx = np.arange(nl)
        for j in range(2):
            live = inc[j] >= 0
            r = np.flatnonzero(live)
            e, a = e_of[j, live], a_of[j, live]
            idx = slot_positions(pat, pat.slot[a[:, None] * nl + b_idx[None, :], e[:, None]])  # (m, nl)
            cols = self.ref_dest[e]
            if not np.array_equal(pat.col[idx], cols):
                raise ValueError("incident column missing from the pattern")
            pos[r, j, :nl] = (idx - rowptr[r][:, None]).astype(np.uint8)
        signs = ((self.ref_sign < 0).astype(np.uint64) << np.arange(nl, dtype=np.uint64)[None, :]).sum(axis=1)
        # entry (a, b) of element e's row a enters with (-1)^(sign bit a xor sign bit b)
        sign_mask = np.zeros((n, 2), np.uint64)
        for j in range(2):
            live = inc[j] >= 0
            s_e = signs[e_of[j, live]]
            own = (s_e >> a_of[j, live].astype(np.uint64)) & np.uint64(1)
            sign_mask[live, j] = s_e ^ (own * np.uint64((1 << nl) - 1))
        self.row_inc_np, self.row_pos_np = packed.astype(np.int32), pos
        self.row_sign_mask_np = sign_mask
        self.row_tile_map_np = xcd_tile_map(e_of[0], n)
        self.row_tile_map = to_dev(self.row_tile_map_np, device=device)
        self.use_tile_map = True
        self.row_inc = to_dev(np.ascontiguousarray(packed.astype(np.int32)), device=device)
        self.row_pos = to_dev(row_pos_device(pos, nl, sign_mask, self.pattern.rowptr, ROW_TILE), device=device)
        self.row_signs = to_dev(signs.astype(np.uint32).view(np.int32), device=device)

    def entry_positions(self, a, b):
        """CSR value positions of the element entries (a, b) (local reference DoFs, integer
        arrays of one shape S) of every element: (E,) + S, from the pattern's slot codes;
        checked against the pattern's columns."""
        a, b = np.broadcast_arrays(np.asarray(a, dtype=np.int64), np.asarray(b, dtype=np.int64))
        flat = (a * self.nl + b).ravel()
        pos = slot_positions(self.pattern, self.pattern.slot[flat].T).reshape((self.E,) + a.shape)
        want = self.ref_dest[:, b.ravel()].reshape((self.E,) + a.shape)
        if not np.array_equal(self.pattern.col[pos], want):
            raise ValueError("slot codes do not point at the entries' columns")
        return pos

    def _validate(self, mesh):
        """Host-side checks that every kernel index stays inside its buffer."""
        key = (self.dim, self.k)
        if key not in DARCY_KERNELS:
            raise ValueError(f"no Darcy kernel for dim={self.dim}, k={self.k}")
        NB, NP, NQ, FD, NQF = DARCY_KERNELS[key]
        if (self.nb, self.np_, self.nq) != (NB, NP, NQ):
            raise ValueError("element sizes do not match the compiled kernel")
        V = len(mesh.points)
        if self.cells_np.shape != (self.E, self.dim + 1) or (self.E and (self.cells_np.min() < 0 or self.cells_np.max() >= V)):
            raise ValueError("cells out of range")
        NS = self.dim * (self.dim + 1) // 2
        if self.tables64.numel() != NQ * (1 + NB * self.dim + NB + NP) + NB * NB * NS + NB * NP + NP:
            raise ValueError("table size mismatch")
        pat = self.pattern
        if pat.slot.shape != (self.nl * self.nl, self.E):
            raise ValueError("slot map shape")
        _check_codes(pat.slot, pat.nnz, pat.n_stage, "slot map")
        _check_codes(self.ref_dest[:, self.nb:], self.n, 0, "pressure rows")
        if len(pat.target) and (pat.target.max() >= pat.nnz or pat.segptr[-1] != pat.n_stage):
            raise ValueError("stage segments out of range")
        inc = self.row_inc_np
        live = inc >= 0
        if np.any((inc[live] >> 5) >= self.E) or np.any((inc[live] & 31) >= self.nl) or not np.all(live[0]):
            raise ValueError("row incidences out of range")
        if np.any(-inc[1][inc[1] <= -2] - 2 >= self.F * FD):
            raise ValueError("row facet codes out of range")
        lens = np.diff(pat.rowptr)
        if lens.max(initial=0) > 2 * self.nl:
            raise ValueError("row longer than the row kernel's LDS slice (2 * n_loc)")
        pp = self.row_pos_np
        if pp.shape != (self.n, 2, self.nl if self.nl <= 8 else (self.nl + 3) // 4 * 4):
            raise ValueError("row column positions shape")
        if np.any((pp != 255) & (pp >= lens[:, None, None])):
            raise ValueError("row column positions out of range")
        tm = self.row_tile_map_np
        nt = -(-self.n // ROW_TILE)
        if not np.array_equal(np.sort(tm[tm >= 0]), np.arange(nt)) or np.any(tm < -1):
            raise ValueError("row tile map is not a permutation of the tiles")
        stride = 4 if self.nl <= 6 else pp.shape[2]
        if self.row_pos.numel() != self.n * 2 * stride:
            raise ValueError("device row column positions size")
        if self.nl <= 5:
            # the in-tile CSR offsets the kernel decodes from bits 25..31 of the two words
            w = self.row_pos.cpu().numpy().view(np.uint32).reshape(self.n, 2).astype(np.int64)
            off = (w[:, 0] >> 25) | ((w[:, 1] >> 25) << 7)
            rp = pat.rowptr.astype(np.int64)
            r = np.arange(self.n)
            if not np.array_equal(off, rp[:-1] - rp[(r // ROW_TILE) * ROW_TILE]):
                raise ValueError("row position words carry wrong in-tile CSR offsets")
        if self.F:
            if self.nqf != NQF:
                raise ValueError("facet quadrature does not match the compiled kernel")
            fn = self.bc_facet.cpu().numpy()
            if fn.min() < 0 or fn.max() >= V:
                raise ValueError("facet nodes out of range")
            nb_ = self.bc_neigh.cpu().numpy()
            if nb_.min() < 0 or nb_.max() >= self.E:
                raise ValueError("facet neighbours out of range")
            cb = self.bc_combo.cpu().numpy()
            if cb.min() < 0 or cb.max() >= self.bc_trace64.shape[0] or self.bc_trace64.shape[1:] != (NQF, FD, self.dim):
                raise ValueError("facet trace tables out of range")
            rows = self.bc_rows.cpu().numpy()
            if rows.shape != (self.F, FD):
                raise ValueError("facet rows shape")
            _check_codes(rows, self.n, 0, "facet rows")

    def _setup_bc(self, mesh, D, device):
        u_sp = self.u_sp
        self.F = len(u_sp.bc_element_ids)
        if not self.F:
            return
        fpts, fw = self.space.bc_quadrature[D - 1]
        self.nqf = len(fw)
        neigh_cells = u_sp.topology.cell_index[u_sp.bc_neighbour]
        combo = facet_combo_index(mesh, D, u_sp.bc_element_ids, neigh_cells, u_sp.bc_local)
        trace = facet_combo_tables(self.u_el, D, fpts)
        fnodes = mesh.entities[D - 1].nodes[u_sp.bc_element_ids - mesh.entities[D - 1].offset]
        self.bc_facet = to_dev(fnodes.astype(np.int32), device=device)
        self.bc_neigh = to_dev(u_sp.bc_neighbour.astype(np.int32), device=device)
        self.bc_combo = to_dev(combo.astype(np.int32), device=device)
        self.bc_trace64 = to_dev(trace, device=device)
        self.bc_trace32 = to_dev(trace.astype(np.float32), device=device)
        self.bc_fw64 = to_dev(fw, device=device)
        self.bc_fw32 = to_dev(fw.astype(np.float32), device=device)
        shift = self.space.strides[self.space.names[0]] + self.space.dof_shift
        self.bc_rows_np = row_codes(u_sp.bc_ref_dest + shift, u_sp.bc_ref_sign)
        self.bc_neigh_np = u_sp.bc_neighbour.astype(np.int32)
        self.bc_rows = to_dev(self.bc_rows_np, device=device)
        Xf = mesh.points[fnodes]  # (F, D, 3)
        self.bc_xq = np.einsum("qa,fac->fqc", p1_bary(fpts), Xf)

    # -- warm assembly ------------------------------------------------------------
    def assemble(self, kappa, rhs, p_dirichlet=None, fp64=False, stream=None):
        """Values + residual rg (the reference's ``rg``, so b = -rg): one launch of the
        row kernel with the Dirichlet facet term folded in."""
        if p_dirichlet is not None and self.F:
            return self._assemble_rows(kappa, rhs, fp64, stream, p_dirichlet=p_dirichlet)
        return self.assemble_elements(kappa, rhs, fp64=fp64, stream=stream)

    def assemble_elements(self, kappa, rhs, fp64=False, stream=None, method="rows"):
        """Matrix values + element residual (rg of every row rewritten).

        method "rows": row-owner gather kernel (one pass, coalesced stores);
        method "element": element kernel + slot map + staged-segment reduction.
        """
        if method == "rows":
            return self._assemble_rows(kappa, rhs, fp64, stream)
        s = nat.stream_handle(stream)
        vals = self.vals64 if fp64 else self.vals32
        stage = self.stage64 if fp64 else self.stage32
        rg = self.rhs64 if fp64 else self.rhs32
        tables = self.tables64 if fp64 else self.tables32
        ka = kappa if isinstance(kappa, CoefficientArg) else coefficient(kappa, self.xq, self.nq, invert=True, device=self.device)
        fa = rhs if isinstance(rhs, CoefficientArg) else coefficient(rhs, self.xq, self.nq, device=self.device)
        check_coefficient(ka, self.E, self.nq, "kappa")
        check_coefficient(fa, self.E, self.nq, "rhs")
        coords = self.coords64 if fp64 else self.coords
        nat.call("nfx_fill", int(fp64), self.n, 0.0, nat.ptr(rg), s)
        nat.call("nfx_assemble_darcy", self.dim, self.k, self.E, nat.ptr(self.cells), nat.ptr(coords),
                 nat.ptr(tables), nat.ptr(ka.get(fp64)), ka.se, ka.sq, nat.ptr(fa.get(fp64)), fa.se, fa.sq,
                 nat.ptr(self.slot), nat.ptr(self.rhs_row), int(fp64), nat.ptr(vals), nat.ptr(stage),
                 nat.ptr(rg), s)
        nat.call("nfx_stage_reduce", len(self.pattern.target), nat.ptr(self.segptr), nat.ptr(self.target),
                 int(fp64), nat.ptr(stage), nat.ptr(vals), s)
        # the coefficient arrays stay referenced until the next launch of this system
        self._held = (ka, fa)
        return vals, rg

    def _assemble_rows(self, kappa, rhs, fp64, stream, p_dirichlet=None):
        s = nat.stream_handle(stream)
        vals = self.vals64 if fp64 else self.vals32
        rg = self.rhs64 if fp64 else self.rhs32
        # repeated launches with the same device coefficients (CoefficientArg objects) reuse
        # the prepared argument list; the entry holds the objects, so an identity match is
        # the same arrays (callables and host arrays are re-evaluated every call)
        key = (fp64, self.use_geometry, self.use_tile_map, p_dirichlet is None)
        ent = self._launch_cache.get(key)
        if ent is not None and ent[0] is kappa and ent[1] is rhs and ent[2] is p_dirichlet:
            nat.check(ent[3](*ent[4], s), ent[5])
            return vals, rg
        ka = kappa if isinstance(kappa, CoefficientArg) else coefficient(kappa, self.xq, self.nq, invert=True, device=self.device)
        fa = rhs if isinstance(rhs, CoefficientArg) else coefficient(rhs, self.xq, self.nq, device=self.device)
        check_coefficient(ka, self.E, self.nq, "kappa")
        check_coefficient(fa, self.E, self.nq, "rhs")
        coords = nat.ptr(self.coords64 if fp64 else self.coords)
        tables = nat.ptr(self.tables64 if fp64 else self.tables32)
        geo = nat.ptr(self.geo64 if fp64 else self.geo32) if self.use_geometry else None
        tmap, n_map = (nat.ptr(self.row_tile_map), self.row_tile_map.numel()) if self.use_tile_map else (None, 0)
        if p_dirichlet is None:
            name = "nfx_assemble_darcy_rows"
            args = (self.dim, self.k, self.n, nat.ptr(self.row_inc), nat.ptr(self.row_pos), nat.ptr(self.rowptr),
                    nat.ptr(self.row_signs), nat.ptr(self.cells), coords, tables, nat.ptr(ka.get(fp64)), ka.se,
                    ka.sq, nat.ptr(fa.get(fp64)), fa.se, fa.sq, geo, tmap, n_map, int(fp64), nat.ptr(vals),
                    nat.ptr(rg))
            held = (ka, fa)
        else:
            pd = self.dirichlet_values(p_dirichlet)
            name = "nfx_assemble_darcy_rows_bc"
            args = (self.dim, self.k, self.n, nat.ptr(self.row_inc), nat.ptr(self.row_pos), nat.ptr(self.rowptr),
                    nat.ptr(self.row_signs), nat.ptr(self.cells), coords, tables, nat.ptr(ka.get(fp64)), ka.se,
                    ka.sq, nat.ptr(fa.get(fp64)), fa.se, fa.sq, self.F, nat.ptr(self.bc_facet),
                    nat.ptr(self.bc_combo), nat.ptr(self.bc_trace64 if fp64 else self.bc_trace32),
                    nat.ptr(self.bc_fw64 if fp64 else self.bc_fw32), self.nqf, nat.ptr(pd.get(fp64)),
                    nat.ptr(self.bc_rows), geo, tmap, n_map, int(fp64), nat.ptr(vals), nat.ptr(rg))
            held = (ka, fa, pd)
        fn = getattr(nat.lib(), name)
        nat.check(fn(*args, s), name)
        self._held = held
        if all(isinstance(c, CoefficientArg) for c in (kappa, rhs, p_dirichlet) if c is not None):
            cargs = tuple(t(v) for t, v in zip(fn.argtypes, args))
            self._launch_cache[key] = (kappa, rhs, p_dirichlet, fn, cargs, name, held)
        return vals, rg

    def dirichlet_values(self, p_dirichlet):
        """Dirichlet data at facet points as a CoefficientArg with one value per point."""
        pd = p_dirichlet if isinstance(p_dirichlet, CoefficientArg) else coefficient(
            p_dirichlet, self.bc_xq, self.nqf, device=self.device)
        if pd.se == 0 or pd.sq == 0:
            t = torch()
            base = pd.get(True)
            full = base.expand(self.F * self.nqf) if pd.se == 0 else base.reshape(-1, 1).expand(self.F, self.nqf)
            pd = CoefficientArg(full.reshape(-1).contiguous(), self.nqf, 1)
        if pd.arr.numel() != self.F * self.nqf:
            raise ValueError("Dirichlet data must have one value per facet point")
        return pd

    def assemble_dirichlet(self, p_dirichlet, fp64=False, stream=None, rows=None, out=None):
        """Facet term into rg (each boundary facet owns its rows)."""
        if not self.F:
            return
        s = nat.stream_handle(stream)
        pd = self.dirichlet_values(p_dirichlet)
        rg = out if out is not None else (self.rhs64 if fp64 else self.rhs32)
        coords = self.coords64 if fp64 else self.coords
        self._held = (pd, rows)
        nat.call("nfx_assemble_darcy_bc", self.dim, self.k, self.F, nat.ptr(self.bc_facet),
                 nat.ptr(self.bc_neigh), nat.ptr(self.bc_combo), nat.ptr(self.cells), nat.ptr(coords),
                 nat.ptr(self.bc_trace64 if fp64 else self.bc_trace32),
                 nat.ptr(self.bc_fw64 if fp64 else self.bc_fw32), self.nqf, nat.ptr(pd.get(fp64)),
                 nat.ptr(self.bc_rows if rows is None else rows), int(fp64), nat.ptr(rg), s)

    # -- element-local matrices (the reference's evaluate_form API) ----------------
    def local_matrices(self, kappa, rhs, fp64=True):
        """All element matrices / residuals in the reference's element-local order.

        Returns numpy (E, nl, nl) and (E, nl): what ``LaplaceDualWeakForm.evaluate_form(i, alpha)``
        returns for each element i (``src/weak_forms/laplace_dual_weak_form.py:11-87``).
        The same kernel runs with slot codes that address a dense per-element array.
        """
        from .pattern import local_slot_map

        t = torch()
        if not hasattr(self, "_local_slot"):
            loc = self.local_index_of_reference()
            self._local_slot = to_dev(local_slot_map(loc, self.ref_sign, self.nl), device=self.device)
            rows = (np.arange(self.E)[:, None] * self.nl + loc[:, self.nb:])
            self._local_rows = to_dev(np.ascontiguousarray(rows.T).astype(np.int32), device=self.device)
        dt = t.float64 if fp64 else t.float32
        out = t.zeros(self.E * self.nl * self.nl, dtype=dt, device=self.device)
        res = t.zeros(self.E * self.nl, dtype=dt, device=self.device)
        ka = kappa if isinstance(kappa, CoefficientArg) else coefficient(kappa, self.xq, self.nq, invert=True, device=self.device)
        fa = rhs if isinstance(rhs, CoefficientArg) else coefficient(rhs, self.xq, self.nq, device=self.device)
        check_coefficient(ka, self.E, self.nq, "kappa")
        check_coefficient(fa, self.E, self.nq, "rhs")
        s = nat.stream_handle()
        nat.call("nfx_assemble_darcy", self.dim, self.k, self.E, nat.ptr(self.cells),
                 nat.ptr(self.coords64 if fp64 else self.coords), nat.ptr(self.tables64 if fp64 else self.tables32),
                 nat.ptr(ka.get(fp64)), ka.se, ka.sq, nat.ptr(fa.get(fp64)), fa.se, fa.sq,
                 nat.ptr(self._local_slot), nat.ptr(self._local_rows), int(fp64), nat.ptr(out), nat.ptr(out),
                 nat.ptr(res), s)
        self._held = (ka, fa)
        return (out.cpu().numpy().reshape(self.E, self.nl, self.nl), res.cpu().numpy().reshape(self.E, self.nl))

    def local_index_of_reference(self):
        """(E, nl): element-local index of every reference basis function (fields stacked)."""
        parts = []
        off = 0
        for name in self.space.names:
            sp = self.space.discrete_spaces[name]
            nb = sp.element.ndofs
            nc = sp.n_comp
            inv = np.argsort(sp.local_src, axis=1)  # reference a -> local i
            loc = (inv[:, :, None] * nc + np.arange(nc)[None, None, :]).reshape(len(inv), nb * nc) + off
            parts.append(loc)
            off += nb * nc
        return np.concatenate(parts, axis=1)

    def csr_numpy(self, fp64=True):
        import scipy.sparse as sps

        vals = (self.vals64 if fp64 else self.vals32).cpu().numpy()
        return sps.csr_matrix((vals, self.pattern.col, self.pattern.rowptr), shape=(self.n, self.n))
