"""Learned correction of the Darcy block preconditioner (RT1-P0), on the device.

The plain block preconditioner of ``solvers.darcy_preconditioner`` uses
``diag(M)^-1`` on the flux rows and ``diag(S)^-1`` on the pressure rows, with
``S = B diag(M)^-1 B^T``. The learned preconditioner replaces ``diag(S)^-1`` by one
symmetric two-grid cycle on ``S``:

    x1 = W r;  x2 = x1 + P Ac^-1 P^T (r - S x1);  z = x2 + W (r - S x2),

with ``P`` the piecewise-constant prolongation from aggregates of cells (boxes of
side ``agg_h``, two hexes of a Kuhn mesh by default), ``Ac = P^T S P`` inverted
exactly (dense FP32 recursive block inverse, replayed from a captured graph), and
``W`` a symmetric sparse smoother
with the pattern of ``S`` whose entries come from a message-passing network:
4 layers of width 64 on the cell graph of ``S`` (bf16 weights and activations,
FP32 accumulation), run once per system. With the output heads at zero ``W`` is
the damped Jacobi smoother ``0.65 diag(S)^-1``: that is the method without the
network, and the trained heads are what the network adds.

The network is trained on the device with the library's forward and backward
kernels. The loss for a system is the mean over sample errors ``e`` of
``||(I - Pi)(I - W S) e||_S^2 / ||e||_S^2`` (``Pi = P Ac^-1 P^T S``), the squared
two-grid error reduction; samples are Gaussian vectors and the pressure errors of
plain-preconditioned Krylov iterates of the system (``x* - x_k``).

The outer solver is FGMRES with FP64 iterative refinement (``nfx_solve_fgmres_ir``). A network
whose smoother has non-finite entries, or a learned solve that does not converge,
falls back to the plain block preconditioner with MINRES.

Reference boundary: the KSP solve of the Darcy driver
(``examples/md_elliptic_scalar/md_elliptic_scalar.py:351-364``); the network has no
reference counterpart.
"""

from __future__ import annotations

import numpy as np

from . import _native as nat
from .device import to_dev, torch

F_IN = 8
WIDTH = 64
LAYERS = 4
OMEGA0 = 0.65
# weight of the two Jacobi sweeps on the flux mass block: 2 / (lmin + lmax) for the
# spectrum [0.5, 2] of diag(M)^-1 M of the RT1 flux mass on the Kuhn meshes (the
# sweeps' error factor (1 - w l)^2 is then at most 0.36 over the spectrum)
FLUX_OMEGA = 0.8


def param_layout():
    """Flat parameter layout (name -> (offset, shape)); the library uses the same."""
    out, off = {}, 0
    names = [("W0", (F_IN, WIDTH)), ("b0", (WIDTH,))]
    for l in range(LAYERS):
        names += [(f"Wh{l}", (WIDTH, WIDTH)), (f"Wm{l}", (WIDTH, WIDTH)), (f"b{l + 1}", (WIDTH,))]
    names += [("ud", (WIDTH,)), ("cd", (1,)), ("ue", (WIDTH,)), ("ce", (1,))]
    for name, shape in names:
        out[name] = (off, shape)
        off += int(np.prod(shape))
    return out, off


def init_params(seed=0, head_scale=0.0):
    """He-initialised encoder and layers; output heads at ``head_scale`` (0 = Jacobi)."""
    lay, n = param_layout()
    rng = np.random.default_rng(seed)
    th = np.zeros(n, dtype=np.float32)
    for name, (o, shape) in lay.items():
        size = int(np.prod(shape))
        if name == "W0":
            th[o:o + size] = rng.standard_normal(size) * np.sqrt(2.0 / F_IN)
        elif name.startswith("Wh") or name.startswith("Wm"):
            th[o:o + size] = rng.standard_normal(size) * (0.5 / np.sqrt(2 * WIDTH))
        elif name in ("ud", "ue"):
            th[o:o + size] = rng.standard_normal(size) * head_scale / np.sqrt(WIDTH)
    return th


def load_params(path):
    with np.load(path, allow_pickle=False) as z:
        th = np.asarray(z["theta"], dtype=np.float32)
    if th.size != param_layout()[1]:
        raise ValueError(f"{path}: {th.size} parameters, expected {param_layout()[1]}")
    return th


def default_weights_path():
    from pathlib import Path

    return Path(__file__).resolve().parent / "data" / "learned_schur_rt1p0.npz"


class SchurGraph:
    """Host-side structure of S and the aggregates, once per Darcy system pattern."""

    def __init__(self, sysd, agg_h=None, device="cuda", links=None, pos_map=None):
        """links: optional (dofs_a, dofs_b), pairs of flux DoFs of sysd that each belong to
        one cell (split faces); the two cells become ELL neighbours in the slots of those
        DoFs (their S entries are the caller's: the Schur ELL gives 0 there), recorded in
        link_slots = (cell_a, slot_a, cell_b, slot_b). pos_map: maps positions in sysd's
        pattern to positions in a larger value array (a coupled system) that the Schur ELL
        then reads."""
        if sysd.k != 1 or sysd.dim != 3:
            raise NotImplementedError("the learned Schur preconditioner covers 3-D RT1-P0 systems")
        pat = sysd.pattern
        rowptr, col = pat.rowptr.astype(np.int64), pat.col.astype(np.int64)
        n, n_u = sysd.n, sysd.n_u
        E = n - n_u
        self.E, self.n, self.n_u = E, n, n_u
        # diagonal positions of the flux block
        pos_diag = np.full(n, -1, dtype=np.int64)
        rows = np.repeat(np.arange(n), np.diff(rowptr))
        d = np.flatnonzero(rows == col)
        pos_diag[rows[d]] = d
        # B rows: pressure row n_u + p; its flux columns (ascending) precede its
        # pressure columns, so the j-th flux entry sits at rowptr[n_u + p] + j
        ent = np.arange(rowptr[n_u], rowptr[n])
        fl = col[ent] < n_u
        ent = ent[fl]
        r = rows[ent] - n_u
        j = ent - rowptr[n_u + r]
        if len(ent) and j.max() >= 4:
            raise ValueError("pressure row with more than 4 flux couplings")
        pos_b = -np.ones((E, 4), dtype=np.int64)
        face = -np.ones((E, 4), dtype=np.int64)
        pos_b[r, j] = ent
        face[r, j] = col[ent]
        has = face >= 0
        # faces touched by two cells: pair the two (cell, slot) incidences
        order = np.argsort(col[ent], kind="stable")
        fs = col[ent][order]
        if len(fs) > 2 and (fs[2:] == fs[:-2]).any():
            raise ValueError("face shared by more than two cells")
        dup = np.flatnonzero(fs[1:] == fs[:-1])
        a1, a2 = order[dup], order[dup + 1]
        nbr = -np.ones((E, 4), dtype=np.int64)
        pos_qb = -np.ones((E, 4), dtype=np.int64)
        nbr[r[a1], j[a1]] = r[a2]
        nbr[r[a2], j[a2]] = r[a1]
        pos_qb[r[a1], j[a1]] = ent[a2]
        pos_qb[r[a2], j[a2]] = ent[a1]
        pos_m = np.where(has, pos_diag[np.where(has, face, 0)], -1)
        if (pos_m[has] < 0).any():
            raise ValueError("flux diagonal missing from the pattern")
        self.link_slots = None
        if links is not None:
            da, db = (np.asarray(d, dtype=np.int64) for d in links)
            cnt = np.bincount(face[has], minlength=n_u)
            owner = -np.ones((n_u, 2), dtype=np.int64)
            fr, fj = np.nonzero(has)
            owner[face[fr, fj]] = np.stack([fr, fj], 1)
            if len(da) != len(db) or (cnt[da] != 1).any() or (cnt[db] != 1).any():
                raise ValueError("linked flux DoFs must each belong to exactly one cell")
            ca, ja = owner[da, 0], owner[da, 1]
            cb, jb = owner[db, 0], owner[db, 1]
            if (nbr[ca, ja] >= 0).any() or (nbr[cb, jb] >= 0).any() or (ca == cb).any():
                raise ValueError("linked slot already has a neighbour")
            nbr[ca, ja] = cb
            nbr[cb, jb] = ca
            self.link_slots = (ca, ja, cb, jb)
        if pos_map is not None:
            pos_map = np.asarray(pos_map, dtype=np.int64)
            remap = lambda a: np.where(a >= 0, pos_map[np.maximum(a, 0)], -1)
            pos_b, pos_qb, pos_m, pos_diag = remap(pos_b), remap(pos_qb), remap(pos_m), remap(pos_diag)
        self.nbr_np = nbr
        # aggregates: boxes of side agg_h over the cell centroids
        pts = sysd.u_sp.mesh.points
        X = pts[sysd.cells_np]
        cent = X.mean(axis=1)
        if agg_h is None:
            vol = np.abs(np.einsum("ij,ij->i", X[:, 1] - X[:, 0],
                                   np.cross(X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]))) / 6.0
            agg_h = 2.0 * float(np.cbrt(6.0 * vol.mean()))
        self.agg_h = agg_h
        agg, n_agg = aggregate_cells(cent, agg_h)
        # cells of the pressure block follow the element order (P0: one dof per cell)
        p_rows = sysd.ref_dest[:, sysd.nb:].ravel() - n_u
        if not np.array_equal(np.sort(p_rows), np.arange(E)):
            raise ValueError("pressure dofs are not one per cell")
        agg_of = np.empty(E, dtype=np.int64)
        agg_of[p_rows] = agg
        self.agg_of_np, self.n_agg = agg_of, n_agg
        hasn = nbr >= 0
        ca = np.repeat(agg_of, 4)[hasn.ravel()]
        cb = agg_of[nbr[hasn]]
        off = ca != cb
        self.coarse_plan = coupling_plan(n_agg, ca[off], cb[off])
        order = np.argsort(agg_of, kind="stable")
        ptr = np.zeros(n_agg + 1, dtype=np.int64)
        np.add.at(ptr, agg_of + 1, 1)
        ptr = np.cumsum(ptr)
        i32 = lambda a: to_dev(np.ascontiguousarray(a, dtype=np.int32), device=device)
        self.nbr, self.pos_b, self.pos_qb, self.pos_m = i32(nbr), i32(pos_b), i32(pos_qb), i32(pos_m)
        self.agg_ptr, self.agg_mem, self.agg_of = i32(ptr), i32(order), i32(agg_of)
        if n_u and (pos_diag[:n_u] < 0).any():
            raise ValueError("a flux row without its diagonal entry in the pattern")
        self.diag_pos_u = to_dev(pos_diag[:n_u], device=device).long()
        self.device = device


def flux_ell(rowptr, col, n_u):
    """(pos, col), each (width, n_u) int32 slot-major: the CSR positions and columns of the
    entries of flux row i with column < n_u, in CSR order; padding slots pos -1, column i."""
    rowptr = np.asarray(rowptr, dtype=np.int64)
    col = np.asarray(col, dtype=np.int64)
    end = int(rowptr[n_u])
    rows = np.repeat(np.arange(n_u), np.diff(rowptr[: n_u + 1]))
    k = np.flatnonzero(col[:end] < n_u)
    r = rows[k]
    cnt = np.bincount(r, minlength=n_u)
    width = int(cnt.max()) if n_u else 0
    first = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    slot = np.arange(len(k)) - first[r]
    pos = np.full((width, n_u), -1, dtype=np.int32)
    fc = np.broadcast_to(np.arange(n_u, dtype=np.int32), (width, n_u)).copy()
    pos[slot, r] = k
    fc[slot, r] = col[k]
    return pos, fc


def aggregate_cells(centroids, h):
    """Aggregate per cell: boxes of side h anchored at the smallest centroid
    coordinate (shifted by h/16 so that box faces fall between cell centroids of a
    structured mesh), numbered by ascending (z, y, x) box among the non-empty boxes."""
    lo = centroids.min(axis=0)
    key = np.floor((centroids - lo) / h + 1.0 / 16.0).astype(np.int64)
    nb = key.max(axis=0) + 1
    lin = key[:, 0] + nb[0] * (key[:, 1] + nb[1] * key[:, 2])
    u, inv = np.unique(lin, return_inverse=True)
    return inv.astype(np.int64), len(u)


def _split(n):
    h = n // 2
    return (h // 128) * 128 if h > 128 else h


def coupling_plan(n, rows, cols):
    """Coupling envelopes of the recursive 2 x 2 inversion (``spd_inverse_``) of a
    symmetric matrix with off-diagonal nonzeros at (rows, cols): for the node
    (offset o, size m) split at h, (lo, hi) with every nonzero of the block
    A[o:o+h, o+h:o+m] inside rows [lo, h) and columns [h, hi) (node-relative;
    lo = hi = h when the halves do not couple). The fill of each Schur complement
    (a dense block on F x F of the second half) is propagated symbolically."""
    P = np.zeros((n, n), dtype=bool)
    P[np.asarray(rows, dtype=np.int64), np.asarray(cols, dtype=np.int64)] = True
    P |= P.T
    plan = {}

    def rec(o, m):
        if m <= 128:
            return
        h = _split(m)
        blk = P[o:o + h, o + h:o + m]
        rr = np.flatnonzero(blk.any(axis=1))
        cc = np.flatnonzero(blk.any(axis=0))
        rec(o, h)
        if len(rr):
            lo, hi = int(rr[0]), int(h + cc[-1] + 1)
            plan[(o, m)] = (lo, hi)
            P[o + h:o + hi, o + h:o + hi] = True
        else:
            plan[(o, m)] = (h, h)
        rec(o + h, m - h)

    rec(0, n)
    return plan


def _sgemm(M, N, K, alpha, A, transA, B, transB, beta, C, stream=None, C2=None):
    """C = beta C + alpha op(A) op(B) with ``nfx_sgemm`` on row-major views (A and B
    are the stored matrices; op transposes them when the flag is set); C2, when given,
    receives the result transposed."""
    nat.call("nfx_sgemm", M, N, K, alpha, A.data_ptr(), A.stride(0), int(transA), B.data_ptr(), B.stride(0),
             int(transB), beta, C.data_ptr(), C.stride(0), None if C2 is None else C2.data_ptr(),
             0 if C2 is None else C2.stride(0), nat.stream_handle(stream))


def spd_inverse_(A, plan=None, stream=None, o=0, gemm="nfx"):
    """In-place inverse of an SPD matrix (FP32, row-major view with unit column stride)
    by recursive 2 x 2 blocking. With B = A12[L, F] the only nonzero part of the
    coupling block (rows L = [lo, h), columns F = [h, hi), from ``coupling_plan``):
    X11 = A11^-1, T = X11[:, L] B, S = A22 - B^T T[L, :] (on F x F only), X22 = S^-1,
    X12 = -T X22[F, :], X11 <- X11 - X12[:, F] T^T, X21 = X12^T. Without a plan the
    envelope is the whole block. Diagonal blocks of size <= 128 are inverted by
    ``nfx_spd_inverse_small`` (block Gauss-Jordan on one CU); the products run on
    ``nfx_sgemm`` (gemm="nfx", which also writes T^T and X21 = X12^T as transposed
    outputs) or the library GEMM (gemm="torch"), in place in the strided blocks."""
    t = torch()
    n = A.shape[0]
    if n <= 128:
        nat.call("nfx_spd_inverse_small", n, A.data_ptr(), A.stride(0), nat.stream_handle(stream))
        return
    h = _split(n)
    lo, hi = plan[(o, n)] if plan is not None else (0, n)
    A11, A12, A21, A22 = A[:h, :h], A[:h, h:], A[h:, :h], A[h:, h:]
    spd_inverse_(A11, plan, stream, o, gemm)
    if hi <= h:
        spd_inverse_(A22, plan, stream, o + h, gemm)
        return
    f = hi - h
    l = h - lo
    B = A12[lo:, :f]
    if gemm == "torch":
        T = t.mm(A11[:, lo:], B)
        A22[:f, :f].addmm_(B.T, T[lo:, :], alpha=-1.0)
        spd_inverse_(A22, plan, stream, o + h, gemm)
        A12.addmm_(T, A22[:f, :], beta=0.0, alpha=-1.0)
        A11.addmm_(A12[:, :f], T.T, alpha=-1.0)
        A21.copy_(A12.T)
    else:
        T = t.empty((h, f), dtype=A.dtype, device=A.device)
        Tt = t.empty((f, h), dtype=A.dtype, device=A.device)
        _sgemm(h, f, l, 1.0, A11[:, lo:], False, B, False, 0.0, T, stream, C2=Tt)
        _sgemm(f, f, l, -1.0, B, True, T[lo:, :], False, 1.0, A22[:f, :f], stream)
        spd_inverse_(A22, plan, stream, o + h, gemm)
        _sgemm(h, n - h, f, -1.0, T, False, A22[:f, :], False, 0.0, A12, stream, C2=A21)
        _sgemm(h, h, f, -1.0, A12[:, :f], False, Tt, False, 1.0, A11, stream)


class LearnedSchurPC:
    """Two-grid Schur block preconditioner with the network smoother (callable z = M^-1 r)."""

    def __init__(self, sysd, theta=None, agg_h=None, psign=-1.0, graph=None, coarse_fp16=True, extra_diag=None,
                 extra_off=None, flux_sweeps=2, flux_omega=None):
        """extra_diag / extra_off: optional FP64 (n_cells) / (n_cells * 4) added to S (diagonal,
        ELL off-diagonals) before the network inputs and the coarse operator are formed (a
        storage term; the Schur complement of an eliminated block).
        flux_sweeps: 2 applies two damped Jacobi sweeps (weight flux_omega) on the flux mass
        block, as the operator 2 w D^-1 - w^2 D^-1 M D^-1 built at set-up from the FP64
        values; 1 the diagonal scaling alone. flux_omega None: w = 2 / (0.5 + b) with b the
        row-sum bound of the spectral radius of D^-1 M, computed at set-up (0.8 on the Kuhn
        meshes, where b = 2 is the spectral radius). The sweeps apply to the system's own flux rows
        only (not with the descriptor overrides of a larger system) and need a positive mass
        diagonal."""
        t = torch()
        self.sys = sysd
        self.g = graph if graph is not None else SchurGraph(sysd, agg_h, device=sysd.device)
        g = self.g
        dev = sysd.device
        E, n_agg = g.E, g.n_agg
        self.E = E
        f32 = lambda *s: t.empty(*s, dtype=t.float32, device=dev)
        self.sdiag = t.empty(E, dtype=t.float64, device=dev)
        self.soff = t.empty(E * 4, dtype=t.float64, device=dev)
        self.a, self.x = f32(E * 4), f32(E * F_IN)
        self.sd32, self.so32 = f32(E), f32(E * 4)
        self.Wd, self.Wo = f32(E), f32(E * 4)
        self.work = f32(4 * E + 2 * n_agg)
        self.Ainv = t.empty((n_agg, n_agg), dtype=t.float32, device=dev)
        self._Ac64 = t.empty((n_agg, n_agg), dtype=t.float64, device=dev)
        self.coarse_fp16 = bool(coarse_fp16)
        self.Ainv16 = t.empty((n_agg, n_agg), dtype=t.int16, device=dev) if self.coarse_fp16 else None
        self.Ascale = t.empty(n_agg, dtype=t.float32, device=dev) if self.coarse_fp16 else None
        self._coarse_graph = None
        self.dMinv = f32(g.n_u)
        self.psign = float(psign)
        self.extra_diag = extra_diag
        self.extra_off = extra_off
        # descriptor overrides for a larger system: flux diagonal of all flux rows, a
        # diagonal block of rows after the cells, and its block-LU coupling to the cells
        # (slot_tail, slot_coef, tail_cells, tail_coef; see nfx_twogrid_pc)
        self.flux_dinv = None
        if flux_sweeps not in (1, 2):
            raise ValueError("flux_sweeps must be 1 or 2")
        self.flux_sweeps = int(flux_sweeps)
        self.flux_omega = None if flux_omega is None else float(flux_omega)
        self._mass_positive = False
        self.flux_width = 0
        self.flux_pos = self.flux_col = self.flux_coef = None
        self._flux_rows = 0
        self._flux_external = False
        if self.flux_sweeps == 2 and g.n_u > 0:
            pos, col = flux_ell(sysd.pattern.rowptr, sysd.pattern.col, g.n_u)
            self._flux_alloc(pos, col, self.g.diag_pos_u.to(t.int32).contiguous())
        self.tail_dinv = None
        self.tail_lu = None
        self.h = self.m = self.z = None
        self.n_params = nat.call("nfx_gnn_num_params")
        if self.n_params != param_layout()[1]:
            raise RuntimeError("network parameter layout differs between host and library")
        self.set_params(theta)
        self.network_used = False
        self.fallback_reason = None

    def set_params(self, theta):
        t = torch()
        if theta is None:
            self.theta32 = None
            self.theta16 = None
            return
        th = t.as_tensor(np.asarray(theta, dtype=np.float32) if not hasattr(theta, "device") else theta,
                         dtype=t.float32, device=self.sys.device).contiguous()
        self.theta32 = th
        self.theta16 = th.to(t.bfloat16).view(t.int16).contiguous()

    def _activations(self):
        t = torch()
        if self.h is None:
            EW = self.E * WIDTH
            dev = self.sys.device
            self.h = t.empty(5 * EW, dtype=t.int16, device=dev)
            self.m = t.empty(4 * EW, dtype=t.int16, device=dev)
            self.z = t.empty(5 * EW, dtype=t.float32, device=dev)

    def schur(self, vals64, stream=None):
        """S (FP64 ELL) and the network inputs from the assembled FP64 values."""
        g, s = self.g, nat.stream_handle(stream)
        nat.call("nfx_schur_ell", g.E, nat.ptr(g.pos_b), nat.ptr(g.pos_qb), nat.ptr(g.pos_m), nat.ptr(vals64),
                 nat.ptr(self.sdiag), nat.ptr(self.soff), s)
        if self.extra_diag is not None:
            self.sdiag.add_(self.extra_diag)
        if self.extra_off is not None:
            self.soff.add_(self.extra_off)
        nat.call("nfx_gnn_features", g.E, nat.ptr(g.nbr), nat.ptr(self.sdiag), nat.ptr(self.soff), nat.ptr(self.a),
                 nat.ptr(self.x), nat.ptr(self.sd32), nat.ptr(self.so32), s)
        dm = vals64[self.g.diag_pos_u]
        self.dMinv.copy_(1.0 / dm.abs())
        if self.flux_coef is not None and not self._flux_external:
            self._flux_build(vals64, stream)

    def _flux_alloc(self, pos, col, diag_pos32):
        t = torch()
        dev = self.sys.device
        self.flux_width = pos.shape[0]
        self._flux_rows = pos.shape[1]
        self.flux_pos, self.flux_col = to_dev(pos, device=dev), to_dev(col, device=dev)
        self._diag_pos32 = diag_pos32
        self.flux_coef = t.empty(pos.size, dtype=t.float32, device=dev)
        self._flux_bound = t.zeros(1, dtype=t.float32, device=dev)

    def _flux_build(self, vals64, stream=None):
        dm = vals64[self._diag_pos32.long()]
        self._mass_positive = bool((dm > 0).all())
        if self._mass_positive:
            nat.call("nfx_flux_sweep_coef", self._flux_rows, self.flux_width, nat.ptr(self.flux_pos),
                     nat.ptr(self.flux_col), nat.ptr(self._diag_pos32), nat.ptr(vals64), self.flux_omega or 0.0,
                     nat.ptr(self.flux_coef), nat.ptr(self._flux_bound), nat.stream_handle(stream))

    def use_flux_rows(self, rowptr, col, n_u, vals64, stream=None):
        """Flux sweeps on the flux rows [0, n_u) of a larger system that owns the descriptor's
        flux rows (with ``flux_dinv``): host CSR pattern (rowptr, col) and its FP64 device
        values. Builds the flux ELL once and its coefficients from vals64 (call again after
        the values change). Inactive when the mass diagonal is not positive."""
        t = torch()
        rp = np.asarray(rowptr, dtype=np.int64)
        cc = np.asarray(col, dtype=np.int64)
        rows = np.repeat(np.arange(n_u), np.diff(rp[: n_u + 1]))
        d = np.flatnonzero(rows == cc[: rp[n_u]])
        diag = np.full(n_u, -1, dtype=np.int64)
        diag[rows[d]] = d
        if n_u == 0 or (diag < 0).any():
            raise ValueError("every flux row needs a diagonal entry")
        pos, fcol = flux_ell(rp, cc, n_u)
        self._flux_alloc(pos, fcol, to_dev(diag.astype(np.int32), device=self.sys.device))
        self._flux_external = True
        self.flux_sweeps = 2
        self._flux_build(vals64, stream)
        return self.sweeps_active

    def forward(self, stream=None):
        g = self.g
        self._activations()
        nat.call("nfx_gnn_forward", g.E, nat.ptr(g.nbr), nat.ptr(self.a), nat.ptr(self.x), nat.ptr(self.theta32),
                 nat.ptr(self.theta16), nat.ptr(self.h), nat.ptr(self.m), nat.ptr(self.z), nat.ptr(self.sdiag),
                 nat.ptr(self.Wd), nat.ptr(self.Wo), nat.stream_handle(stream))

    def _coarse_eager(self, stream=None):
        g = self.g
        self._Ac64.zero_()
        nat.call("nfx_coarse_dense", g.n_agg, nat.ptr(g.agg_ptr), nat.ptr(g.agg_mem), nat.ptr(g.agg_of),
                 nat.ptr(g.nbr), nat.ptr(self.sdiag), nat.ptr(self.soff), nat.ptr(self._Ac64),
                 nat.stream_handle(stream))
        self.Ainv.copy_(self._Ac64)
        spd_inverse_(self.Ainv, self.g.coarse_plan, stream)
        if self.coarse_fp16:
            nat.call("nfx_half_rows", g.n_agg, nat.ptr(self.Ainv), nat.ptr(self.Ainv16), nat.ptr(self.Ascale),
                     nat.stream_handle(stream))

    def coarse(self, stream=None, graph=True):
        """Ainv = (P^T S P)^-1: deterministic coarse kernel (FP64), FP32 recursive block
        inverse. With graph=True the sequence is captured once and replayed."""
        t = torch()
        if not graph or stream is not None:
            self._coarse_eager(stream)
            return
        if self._coarse_graph is None:
            side = t.cuda.Stream()
            side.wait_stream(t.cuda.current_stream())
            with t.cuda.stream(side):
                self._coarse_eager()
            t.cuda.current_stream().wait_stream(side)
            t.cuda.synchronize()
            gr = t.cuda.CUDAGraph()
            with t.cuda.graph(gr):
                self._coarse_eager()
            self._coarse_graph = gr
        self._coarse_graph.replay()

    def setup(self, vals64=None, network=True, stream=None):
        """Per-system set-up after FP64 assembly. Returns True if the network smoother is used."""
        t = torch()
        vals64 = self.sys.vals64 if vals64 is None else vals64
        self.schur(vals64, stream)
        self.network_used = False
        self.fallback_reason = None
        if network and self.theta32 is not None:
            if not bool(t.isfinite(self.theta32).all()):
                # a NaN weight can vanish inside relu (max(NaN, 0) = 0): check the weights themselves
                self.fallback_reason = "non-finite network parameters"
            else:
                self.forward(stream)
                if bool(t.isfinite(self.Wd).all()) and bool(t.isfinite(self.Wo).all()):
                    self.network_used = True
                else:
                    self.fallback_reason = "non-finite smoother entries"
        if not self.network_used:
            self.Wd.copy_(OMEGA0 / self.sd32)
            self.Wo.zero_()
        self.coarse(stream)
        return self.network_used

    def __call__(self, r, out=None, stream=None):
        t = torch()
        g = self.g
        import ctypes

        z = out if out is not None else t.empty_like(r)
        self._desc = self.desc()
        nat.call("nfx_twogrid_apply", ctypes.byref(self._desc), nat.ptr(r), nat.ptr(z), nat.stream_handle(stream))
        return z

    def desc(self):
        """The ``nfx_twogrid_pc`` descriptor of this preconditioner (for the C++ FGMRES)."""
        g = self.g
        dminv = self.dMinv if self.flux_dinv is None else self.flux_dinv
        tail = self.tail_dinv
        lu = self.tail_lu if tail is not None else None
        if lu is not None:
            st, sc, tc, tk = lu
            if st.numel() != 4 * g.E or sc.numel() != 4 * g.E or tc.numel() != 2 * tail.numel() \
                    or tk.numel() != 2 * tail.numel():
                raise ValueError("tail coupling arrays do not match the cell and tail counts")
        sweeps = self.sweeps_active
        return nat.TwoGridPC(int(dminv.numel()) if self.flux_dinv is not None else g.n_u, g.E, g.n_agg,
                             nat.ptr(g.nbr), nat.ptr(self.Wd), nat.ptr(self.Wo),
                             nat.ptr(self.sd32), nat.ptr(self.so32), nat.ptr(g.agg_ptr), nat.ptr(g.agg_mem),
                             nat.ptr(g.agg_of), nat.ptr(self.Ainv), nat.ptr(self.Ainv16), nat.ptr(self.Ascale),
                             nat.ptr(dminv), self.psign, nat.ptr(self.work), 0 if tail is None else int(tail.numel()),
                             None if tail is None else nat.ptr(tail),
                             *(nat.ptr(a) for a in lu) if lu is not None else (None,) * 4,
                             self.flux_width if sweeps else 0, nat.ptr(self.flux_col) if sweeps else None,
                             nat.ptr(self.flux_coef) if sweeps else None)

    @property
    def flux_weight(self):
        """The sweeps' weight in use (reads the device bound when it is automatic)."""
        if self.flux_omega is not None:
            return self.flux_omega
        if self.flux_coef is None:
            return None
        return 2.0 / (0.5 + float(self._flux_bound.item()))

    @property
    def flux_bound(self):
        """Row-sum bound of the spectral radius of D^-1 M from the last set-up."""
        return None if self.flux_coef is None else float(self._flux_bound.item())

    @property
    def sweeps_active(self):
        """Whether the flux rows get the two mass sweeps (else the diagonal scaling)."""
        n_u = int(self.flux_dinv.numel()) if self.flux_dinv is not None else self.g.n_u
        return self.flux_coef is not None and self._mass_positive and self._flux_rows == n_u

    def smoother_numpy(self):
        """(Wd, Wo (E,4)) on the host (tests)."""
        return self.Wd.cpu().numpy(), self.Wo.cpu().numpy().reshape(-1, 4)


class Adam:
    """Adam on the FP32 master parameters (device tensors)."""

    def __init__(self, theta, lr=1e-3, betas=(0.9, 0.999), eps=1e-8, clip=1.0):
        t = torch()
        self.theta = theta
        self.m = t.zeros_like(theta)
        self.v = t.zeros_like(theta)
        self.lr, self.b1, self.b2, self.eps, self.clip = lr, betas[0], betas[1], eps, clip
        self.k = 0

    def step(self, grad):
        t = torch()
        gn = t.linalg.vector_norm(grad)
        if self.clip and float(gn) > self.clip:
            grad = grad * (self.clip / gn)
        self.k += 1
        self.m.mul_(self.b1).add_(grad, alpha=1 - self.b1)
        self.v.mul_(self.b2).addcmul_(grad, grad, value=1 - self.b2)
        mh = self.m / (1 - self.b1 ** self.k)
        vh = self.v / (1 - self.b2 ** self.k)
        self.theta.sub_(self.lr * mh / (vh.sqrt() + self.eps))
        return float(gn)


class SmootherTrainer:
    """Loss, gradient and Adam steps of the network smoother on device-resident systems."""

    def __init__(self, pc: LearnedSchurPC, theta0, lr=1e-3, n_random=12, seed=0):
        t = torch()
        self.pc = pc
        pc.set_params(theta0)
        self.opt = Adam(pc.theta32, lr=lr)
        self.n_random = n_random
        self.gen = t.Generator(device=pc.sys.device)
        self.gen.manual_seed(seed)
        E = pc.E
        dev = pc.sys.device
        self.dh = t.empty(E * WIDTH, dtype=t.float32, device=dev)
        self.dm = t.empty(E * WIDTH, dtype=t.float32, device=dev)
        self.partial = t.empty(nat.call("nfx_gnn_backward_workspace", E), dtype=t.float32, device=dev)
        self.grad = t.zeros(pc.n_params, dtype=t.float32, device=dev)
        self.gd = t.empty(E, dtype=t.float32, device=dev)
        self.go = t.empty(E * 4, dtype=t.float32, device=dev)

    def samples(self, extra=None):
        t = torch()
        e = t.randn((self.n_random, self.pc.E), generator=self.gen, device=self.pc.sys.device, dtype=t.float32)
        if extra is not None and len(extra):
            ex = extra / extra.norm(dim=1, keepdim=True).clamp_min(1e-30) * np.sqrt(self.pc.E)
            e = t.cat([e, ex.to(t.float32)], dim=0)
        return e.contiguous()

    def loss_and_gamma(self, e, stream=None):
        """Two-grid loss of the current smoother on samples e (ns, E); fills gamma."""
        t = torch()
        pc, g = self.pc, self.pc.g
        s = nat.stream_handle(stream)
        ns, E = e.shape
        Se = t.empty_like(e)
        v = t.empty_like(e)
        Sv = t.empty_like(e)
        Su = t.empty_like(e)
        ell = lambda d, o, X, Y, alpha=1.0, beta=0.0, Y0=None: nat.call(
            "nfx_ell_spmm", E, ns, nat.ptr(g.nbr), nat.ptr(d), nat.ptr(o), alpha, nat.ptr(X), beta, nat.ptr(Y0),
            nat.ptr(Y), s)
        ell(pc.sd32, pc.so32, e, Se)
        en = (e.double() * Se.double()).sum(dim=1)
        ell(pc.Wd, pc.Wo, Se, v, alpha=-1.0, beta=1.0, Y0=e)
        ell(pc.sd32, pc.so32, v, Sv)
        c = t.empty((ns, g.n_agg), dtype=t.float32, device=e.device)
        nat.call("nfx_agg_sum", g.n_agg, ns, E, nat.ptr(g.agg_ptr), nat.ptr(g.agg_mem), nat.ptr(Sv), nat.ptr(c), s)
        y = t.mm(c, pc.Ainv)
        u = (v - y[:, g.agg_of.long()]).contiguous()
        ell(pc.sd32, pc.so32, u, Su)
        per = (u.double() * Su.double()).sum(dim=1) / en
        w = (-2.0 / (ns * en)).to(t.float32).contiguous()
        nat.call("nfx_smoother_grad", E, ns, nat.ptr(g.nbr), nat.ptr(Su), nat.ptr(Se), nat.ptr(w), nat.ptr(self.gd),
                 nat.ptr(self.go), s)
        return per.mean()

    def backward(self, stream=None):
        pc, g = self.pc, self.pc.g
        nat.call("nfx_gnn_backward", g.E, nat.ptr(g.nbr), nat.ptr(pc.a), nat.ptr(pc.x), nat.ptr(pc.theta16),
                 nat.ptr(pc.h), nat.ptr(pc.m), nat.ptr(pc.z), nat.ptr(pc.sdiag), nat.ptr(self.gd), nat.ptr(self.go),
                 nat.ptr(self.dh), nat.ptr(self.dm), nat.ptr(self.partial), nat.ptr(self.grad),
                 nat.stream_handle(stream))
        return self.grad

    def step(self, vals64, extra=None):
        """One Adam step on one assembled system (FP64 values). Returns the loss before the step."""
        t = torch()
        pc = self.pc
        pc.schur(vals64)
        pc.forward()
        pc.coarse()
        e = self.samples(extra)
        loss = self.loss_and_gamma(e)
        self.backward()
        self.opt.step(self.grad)
        pc.theta16 = pc.theta32.to(t.bfloat16).view(t.int16).contiguous()
        return float(loss)

    def evaluate(self, vals64, n=32, network=True):
        """Loss on fresh Gaussian samples without a step."""
        pc = self.pc
        pc.setup(vals64, network=network)
        e = self.samples()
        return float(self.loss_and_gamma(e))


class Fgmres:
    """FGMRES(restart) + FP64 refinement on the device (``nfx_solve_fgmres_ir``)."""

    def __init__(self, n, restart=40, stream=None):
        import ctypes

        self.n, self.restart = int(n), int(restart)
        h = ctypes.c_void_p()
        nat.call("nfx_fgmres_create", ctypes.byref(h), self.n, self.restart, nat.stream_handle(stream))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            nat.lib().nfx_fgmres_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, rowptr, col, vals32, vals64, b, x=None, pc=None, pinv=None, rtol=1e-8, inner_rtol=1e-5,
              max_inner=4000, max_outer=20, reorth=True, check=4):
        """pc: a LearnedSchurPC (two-grid block preconditioner) or None; pinv: diagonal."""
        import ctypes

        t = torch()
        b = b.to(dtype=t.float64).contiguous()
        if x is None:
            x = t.zeros(self.n, dtype=t.float64, device=b.device)
        info = (ctypes.c_int * 2)()
        resid = (ctypes.c_double * 1)()
        hist = (ctypes.c_double * (max_outer + 1))()
        desc = pc.desc() if pc is not None else None
        rc = nat.lib().nfx_solve_fgmres_ir(
            self._h, nat.ptr(rowptr), nat.ptr(col), nat.ptr(vals32), nat.ptr(vals64),
            ctypes.byref(desc) if desc is not None else None, nat.ptr(pinv), nat.ptr(b), nat.ptr(x), rtol,
            inner_rtol, max_inner, max_outer, int(bool(reorth)), check, info, resid, hist, max_outer + 1)
        if rc not in (nat.NFX_OK, nat.NFX_E_NOT_CONVERGED):
            nat.check(rc, "nfx_solve_fgmres_ir")
        return x, {"converged": rc == nat.NFX_OK, "outer": info[0], "iterations": info[1],
                   "residual": resid[0], "history": [hist[i] for i in range(min(info[0] + 1, max_outer + 1))]}


def solve_learned(sysd, pc: LearnedSchurPC, b64, rtol=1e-8, restart=40, maxit=4000, inner_tol=1e-5,
                  fallback=True, reorth=True, check=4, solver=None):
    """FGMRES + FP64 refinement with the learned preconditioner; plain MINRES fallback.

    ``pc.setup`` must have run on the current FP64 values. Returns (x64, info).
    restart 40: with the flux sweeps the 32^3 x 6 systems take 104 iterations (120 at 60,
    110 at 50, 118 at 30; FP32 classical Gram-Schmidt without re-orthogonalisation loses
    orthogonality over longer cycles)."""
    from .solvers import MinresIR, darcy_preconditioner

    t = torch()
    info = {"fallback": None}
    x = None
    if pc.network_used or pc.fallback_reason is None:
        fg = solver if solver is not None else Fgmres(sysd.n, restart)
        x, inf = fg.solve(sysd.rowptr, sysd.col, sysd.vals32, sysd.vals64, b64, pc=pc, rtol=rtol,
                          inner_rtol=inner_tol, max_inner=maxit, reorth=reorth, check=check)
        if solver is None:
            fg.close()
        info.update(inf)
        info["method"] = "fgmres+twogrid" + ("+network" if pc.network_used else "+jacobi")
        if inf["converged"] and np.isfinite(inf["residual"]):
            return x, info
        info["fallback"] = "learned solve did not converge"
    else:
        info["fallback"] = pc.fallback_reason
    if not fallback:
        return x, info
    sign, pinv = darcy_preconditioner(sysd.rowptr, sysd.col, sysd.vals64, sysd.n, sysd.n_u)
    solver = MinresIR(sysd.rowptr, sysd.col, sysd.vals32, sysd.vals64, sign, pinv)
    x0 = x if (x is not None and bool(t.isfinite(x).all())) else None
    xs, inf = solver.solve(b64, x=x0, rtol=rtol, max_inner=100000)
    solver.close()
    info.update(inf)
    info["method"] = "minres+block-diagonal (fallback)"
    return xs, info
