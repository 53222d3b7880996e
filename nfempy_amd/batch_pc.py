"""Two-grid Schur preconditioner of the pressure rows of a batch of Darcy instances, for the
batched MINRES (``nfx_batch_tg``, ``nfx_minres_batch_set_pc``).

The plain block preconditioner of each instance is ``diag(M)^-1`` on the flux rows and
``|diag(S)|^-1`` on the pressure rows, ``S = B diag(M)^-1 B^T``. Here the pressure rows of
every instance of a group (one group per degree: RT1-P0 with one pressure unknown per cell,
RT2-P1 with four) get one symmetric two-grid cycle on ``S`` instead, all instances of the
group in the same launches:

- ``S`` in cell blocks (the cell's own and one per face neighbour), formed on the device from
  each instance's FP64 values;
- RT2-P1: the flux rows get the inverse 3 x 3 blocks of the flux mass (a face's three DoFs,
  a cell's three interior DoFs) instead of its diagonal, and ``S = B D^-1 B^T`` uses the same
  blocks: on the 32-instance batch the RT2 instances then need 240-368 iterations instead of
  576-640 (with the diagonal) and 2,000-2,600 (plain);
- smoother: block Jacobi ``omega S_cc^-1``, or for RT1-P0 the learned smoother of
  ``learned.py`` (the network runs on the union of the instances' cell graphs; it only looks
  at a cell's 4-hop neighbourhood, so an instance's smoother does not depend on the others);
- aggregates of 2 x 2 x 2 hexes of each instance's mesh, piecewise-constant prolongation of
  every unknown; coarse operator ``P^T S P`` per instance (FP64), inverted per instance
  (FP32, ``learned.spd_inverse_``): the coarse solve is block diagonal over the batch.

The per-(n, k) structure (CSR positions of B and of the flux diagonals, neighbours,
aggregates, coupling plan of the coarse inverse) is built once per mesh and shared by every
instance on it. No reference counterpart: the preconditioned solve replaces the KSP solve of
``examples/md_elliptic_scalar/md_elliptic_scalar.py:351-364`` for each instance.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _native as nat
from .device import to_dev, torch
from .learned import OMEGA0, WIDTH, F_IN, aggregate_cells, coupling_plan, spd_inverse_

# block-Jacobi weights: the single-system Jacobi weight for RT1-P0; for RT2-P1 (with the
# flux blocks) 0.65 / 0.8 / 0.95 gave at most 416 / 368 / 368 iterations on the 32-instance
# batch (n 4..16), and weights above 1 lose the definiteness MINRES needs
OMEGA = {1: OMEGA0, 4: 0.85}


def cell_structure(sysd):
    """Host structure of one Darcy system for the cycle (cached on the system)."""
    st = getattr(sysd, "_btg_struct", None)
    if st is not None:
        return st
    if sysd.dim != 3 or sysd.k not in (1, 2):
        raise NotImplementedError("the batched two-grid covers 3-D RT1-P0 and RT2-P1")
    nbf, nb, E, n, n_u = sysd.nb, sysd.np_, sysd.E, sysd.n, sysd.n_u
    rd = np.asarray(sysd.ref_dest, dtype=np.int64)
    fl, pr = rd[:, :nbf], rd[:, nbf:]
    # CSR positions of element entries (from the slot codes): B (pressure row i, flux
    # column j) as (E, nbf, nb), and the flux diagonals (E, nbf)
    jf, ip = np.arange(nbf)[:, None], nbf + np.arange(nb)[None, :]
    posb = sysd.entry_positions(ip, jf)
    posd = sysd.entry_positions(np.arange(nbf), np.arange(nbf))
    fdofs = np.asarray(sysd.u_el.entity_dofs[2], dtype=np.int64)  # (4, fd)
    fd = fdofs.shape[1]
    # incidences of every flux DoF: (element, local index), at most two
    flat = fl.ravel()
    order = np.argsort(flat, kind="stable")
    fs = flat[order]
    if len(fs) > 2 and (fs[2:] == fs[:-2]).any():
        raise ValueError("flux DoF shared by more than two cells")
    inc = -np.ones((n_u, 2), dtype=np.int64)
    first = np.r_[True, fs[1:] != fs[:-1]]
    inc[fs[first], 0] = order[first]
    inc[fs[~first], 1] = order[~first]
    ge = fl[:, fdofs]  # (E, 4, fd) global DoFs of each local face
    e_idx = np.arange(E)[:, None, None]
    o0, o1 = inc[ge, 0], inc[ge, 1]
    other = np.where(o0 // nbf == e_idx, o1, o0)  # the neighbour's incidence, or -1
    has = (other >= 0) & (o1 >= 0)
    nbr_e = np.where(has, other // nbf, -1)
    if not (nbr_e == nbr_e[:, :, :1]).all():
        raise ValueError("face DoFs of one face lead to different neighbours")
    nbr = nbr_e[:, :, 0]
    nloc = np.where(has, other % nbf, 0)
    # aggregates: boxes of side 2 hexes over the cell centroids
    X = sysd.u_sp.mesh.points[sysd.cells_np]
    cent = X.mean(axis=1)
    vol = np.abs(np.einsum("ij,ij->i", X[:, 1] - X[:, 0], np.cross(X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]))) / 6.0
    agg, n_agg = aggregate_cells(cent, 2.0 * float(np.cbrt(6.0 * vol.mean())))
    mem = np.argsort(agg, kind="stable")
    ptr = np.concatenate([[0], np.cumsum(np.bincount(agg, minlength=n_agg))])
    blk = {}
    if sysd.k == 2:
        # flux mass in 3 x 3 blocks: a face's DoFs, the cell's interior DoFs
        inter = np.asarray(sysd.u_el.entity_dofs[3][0], dtype=np.int64)
        if fd != 3 or list(inter) != [12, 13, 14]:
            raise ValueError("unexpected RT2 DoF layout")
        gf = fl[:, fdofs]  # (E, 4, 3)
        gi = fl[:, inter]  # (E, 3)
        posmf = sysd.entry_positions(fdofs[:, :, None], fdofs[:, None, :])  # (E, 4, 3, 3)
        posmi = sysd.entry_positions(inter[:, None], inter[None, :])  # (E, 3, 3)
        # first occurrence of each face in (element, local face) order; a face DoF belongs to
        # one face only, so its smallest DoF names it
        _, first = np.unique(gf.reshape(-1, 3).min(axis=1), return_index=True)
        first = np.sort(first)
        blk = {"posmf": posmf.astype(np.int32), "posmi": posmi.astype(np.int32),
               "fb_rows": np.concatenate([gf.reshape(-1, 3)[first], gi]),
               "fb_pos": np.concatenate([posmf.reshape(-1, 9)[first], posmi.reshape(-1, 9)]).astype(np.int32)}
    hn = nbr >= 0
    ca = np.repeat(agg, 4)[hn.ravel()]
    cb = agg[nbr[hn]]
    off = ca != cb
    st = {
        "nbf": nbf, "nb": nb, "fd": fd, "E": E, "n_agg": n_agg,
        "posb": posb.astype(np.int32), "posd": posd.astype(np.int32), "fdofs": fdofs.astype(np.int32),
        "nloc": nloc.astype(np.int32), "nbr": nbr.astype(np.int64), "prow": pr,
        "agg": agg, "agg_ptr": ptr, "agg_mem": mem, "plan": coupling_plan(n_agg, ca[off], cb[off]),
        **blk,
    }
    sysd._btg_struct = st
    return st


class _Group:
    """One degree's instances, concatenated."""

    def __init__(self, nb, members, device, flux_blocks=True):
        """members: list of (batch index, structure, row offset of the instance);
        flux_blocks: RT2-P1 flux mass in 3 x 3 blocks (else its diagonal)."""
        t = torch()
        self.nb = nb
        self.members = members
        s0 = members[0][1]
        self.nbf, self.fd = s0["nbf"], s0["fd"]
        cells = [m[1]["E"] for m in members]
        aggs = [m[1]["n_agg"] for m in members]
        self.cell_off = np.concatenate([[0], np.cumsum(cells)])
        self.agg_off = np.concatenate([[0], np.cumsum(aggs)])
        self.C, self.A = int(self.cell_off[-1]), int(self.agg_off[-1])
        self.coff = np.concatenate([[0], np.cumsum(np.asarray(aggs, dtype=np.int64) ** 2)])
        cat = lambda f: np.concatenate([f(j, m) for j, m in enumerate(members)])
        prow = cat(lambda j, m: m[1]["prow"] + m[2])
        if prow.max() >= 1 << 31:
            raise ValueError("batch rows exceed int32")
        nbr = cat(lambda j, m: np.where(m[1]["nbr"] >= 0, m[1]["nbr"] + self.cell_off[j], -1))
        agg_of = cat(lambda j, m: m[1]["agg"] + self.agg_off[j])
        agg_mem = cat(lambda j, m: m[1]["agg_mem"] + self.cell_off[j])
        agg_ptr = np.concatenate([[0]] + [m[1]["agg_ptr"][1:] + self.cell_off[j] for j, m in enumerate(members)])
        inst = cat(lambda j, m: np.full(m[1]["E"], m[0], dtype=np.int64))
        arow = cat(lambda j, m: self.coff[j] + np.arange(m[1]["n_agg"], dtype=np.int64) * m[1]["n_agg"])
        a0 = cat(lambda j, m: np.full(m[1]["n_agg"], self.agg_off[j], dtype=np.int64))
        alen = cat(lambda j, m: np.full(m[1]["n_agg"], m[1]["n_agg"], dtype=np.int64))
        i32 = lambda a: to_dev(np.ascontiguousarray(a, dtype=np.int32), device=device)
        self.d = {
            "prow": i32(prow), "nbr": i32(nbr), "agg_of": i32(agg_of), "agg_mem": i32(agg_mem),
            "agg_ptr": i32(agg_ptr), "cell_inst": i32(inst), "a0": i32(a0), "alen": i32(alen),
            "arow": to_dev(np.ascontiguousarray(arow, dtype=np.int64), device=device),
            "posb": i32(cat(lambda j, m: m[1]["posb"].reshape(m[1]["E"], -1))),
            "posd": i32(cat(lambda j, m: m[1]["posd"])),
            "nloc": i32(cat(lambda j, m: m[1]["nloc"].reshape(m[1]["E"], -1))),
            "fdofs": i32(s0["fdofs"]),
        }
        self.blocks = bool(flux_blocks) and "posmf" in s0
        self.n_fb = 0
        if self.blocks:
            self.d["posmf"] = i32(cat(lambda j, m: m[1]["posmf"].reshape(m[1]["E"], -1)))
            self.d["posmi"] = i32(cat(lambda j, m: m[1]["posmi"].reshape(m[1]["E"], -1)))
            fb_rows = cat(lambda j, m: m[1]["fb_rows"] + m[2])
            self.n_fb = len(fb_rows)
            self.d["fb_rows"] = i32(fb_rows)
            self.d["fb_pos"] = i32(cat(lambda j, m: m[1]["fb_pos"]))
            self.d["fb_inst"] = i32(cat(lambda j, m: np.full(len(m[1]["fb_rows"]), m[0], dtype=np.int64)))
            self.fb_inv = t.empty(self.n_fb * 9, dtype=t.float32, device=device)
        C, A, NB = self.C, self.A, nb
        f64 = lambda k: t.empty(k, dtype=t.float64, device=device)
        f32 = lambda k: t.empty(k, dtype=t.float32, device=device)
        self.sd64, self.so64 = f64(C * NB * NB), f64(C * 4 * NB * NB)
        self.sd, self.so = f32(C * NB * NB), f32(C * 4 * NB * NB)
        self.rs = f32(C * 5 * NB)
        self.wd, self.wo = f32(C * NB * NB), None
        self.ac64 = f64(int(self.coff[-1]))
        self.ainv = f32(int(self.coff[-1]))
        self.work = f32(4 * C * NB + 2 * A + 4)
        self.status = t.zeros(1, dtype=t.int32, device=device)
        self.network_used = False

    def desc(self):
        d = self.d
        return nat.BatchTG(self.nb, self.C, self.A, nat.ptr(d["cell_inst"]), nat.ptr(d["prow"]), nat.ptr(d["nbr"]),
                           nat.ptr(self.sd), nat.ptr(self.so), nat.ptr(self.rs), nat.ptr(self.wd),
                           None if self.wo is None else nat.ptr(self.wo), nat.ptr(d["agg_ptr"]),
                           nat.ptr(d["agg_mem"]), nat.ptr(d["agg_of"]), nat.ptr(d["arow"]), nat.ptr(d["a0"]),
                           nat.ptr(d["alen"]), nat.ptr(self.ainv), nat.ptr(self.work), self.n_fb,
                           nat.ptr(d.get("fb_inst")), nat.ptr(d.get("fb_rows")),
                           nat.ptr(self.fb_inv) if self.n_fb else None)


class BatchTwoGrid:
    """The two-grid groups of a batched solve. ``systems``: the DarcySystem of each instance
    in batch order (shared objects allowed), ``offsets``: row offsets of the instances in the
    stacked vector; ``theta``: learned smoother parameters for the RT1-P0 group (None: block
    Jacobi for every group)."""

    def __init__(self, systems, offsets, theta=None, device="cuda", flux_blocks=True):
        t = torch()
        self.device = device
        by_nb = {}
        for i, sysd in enumerate(systems):
            st = cell_structure(sysd)
            by_nb.setdefault(st["nb"], []).append((i, st, int(offsets[i])))
        self.m = len(systems)
        self.groups = [_Group(nb, mem, device, flux_blocks) for nb, mem in sorted(by_nb.items())]
        self.theta = theta
        self._gnn = None
        g1 = [g for g in self.groups if g.nb == 1]
        if theta is not None and g1:
            g = g1[0]
            E = g.C
            th = t.as_tensor(np.asarray(theta, dtype=np.float32), device=device)
            self._gnn = {
                "theta32": th, "theta16": th.to(t.bfloat16).view(t.int16).contiguous(),
                "a": t.empty(E * 4, dtype=t.float32, device=device),
                "x": t.empty(E * F_IN, dtype=t.float32, device=device),
                "h": t.empty(5 * E * WIDTH, dtype=t.int16, device=device),
                "mm": t.empty(4 * E * WIDTH, dtype=t.int16, device=device),
                "z": t.empty(5 * E * WIDTH, dtype=t.float32, device=device),
            }
            g.wo = t.empty(E * 4, dtype=t.float32, device=device)
        self.vptr = t.zeros(self.m, dtype=t.int64, device=device)
        self._descs = None

    def setup(self, vals64, network=True, stream=None):
        """vals64: FP64 CSR values of every instance (batch order; kept referenced until the
        next set-up). Returns the ctypes array of group descriptors."""
        t = torch()
        s = nat.stream_handle(stream)
        self._vals = list(vals64)
        self.vptr.copy_(t.tensor([v.data_ptr() for v in self._vals], dtype=t.int64))
        for g in self.groups:
            d = g.d
            g.status.zero_()
            if g.blocks:
                nat.call("nfx_batch_tg_fluxinv", g.n_fb, nat.ptr(d["fb_inst"]), nat.ptr(self.vptr),
                         nat.ptr(d["fb_pos"]), nat.ptr(g.fb_inv), nat.ptr(g.status), s)
                nat.call("nfx_batch_tg_schur_blk", g.C, nat.ptr(d["cell_inst"]), nat.ptr(self.vptr),
                         nat.ptr(d["posb"]), nat.ptr(d["posmf"]), nat.ptr(d["posmi"]), nat.ptr(d["nbr"]),
                         nat.ptr(d["fdofs"]), nat.ptr(d["nloc"]), nat.ptr(g.sd64), nat.ptr(g.so64), nat.ptr(g.sd),
                         nat.ptr(g.so), nat.ptr(g.status), s)
            else:
                nat.call("nfx_batch_tg_schur", g.nb, g.C, g.nbf, g.fd, nat.ptr(d["cell_inst"]), nat.ptr(self.vptr),
                         nat.ptr(d["posb"]), nat.ptr(d["posd"]), nat.ptr(d["nbr"]), nat.ptr(d["fdofs"]),
                         nat.ptr(d["nloc"]), nat.ptr(g.sd64), nat.ptr(g.so64), nat.ptr(g.sd), nat.ptr(g.so), s)
            nat.call("nfx_batch_tg_rowsums", g.nb, g.C, nat.ptr(g.sd64), nat.ptr(g.so64), nat.ptr(g.rs), s)
            g.network_used = False
            if g.nb == 1 and self._gnn is not None and network:
                q = self._gnn
                nat.call("nfx_gnn_features", g.C, nat.ptr(d["nbr"]), nat.ptr(g.sd64), nat.ptr(g.so64), nat.ptr(q["a"]),
                         nat.ptr(q["x"]), nat.ptr(g.sd), nat.ptr(g.so), s)
                nat.call("nfx_gnn_forward", g.C, nat.ptr(d["nbr"]), nat.ptr(q["a"]), nat.ptr(q["x"]),
                         nat.ptr(q["theta32"]), nat.ptr(q["theta16"]), nat.ptr(q["h"]), nat.ptr(q["mm"]),
                         nat.ptr(q["z"]), nat.ptr(g.sd64), nat.ptr(g.wd), nat.ptr(g.wo), s)
                g.network_used = bool(t.isfinite(g.wd).all()) and bool(t.isfinite(g.wo).all())
            if not g.network_used:
                nat.call("nfx_batch_tg_smoother", g.nb, g.C, nat.ptr(g.sd64), OMEGA[g.nb], nat.ptr(g.wd),
                         nat.ptr(g.status), s)
                if g.wo is not None:
                    g.wo.zero_()
            if int(g.status.item()):
                raise RuntimeError("a flux block or a cell block of the Schur complement is not SPD")
            g.ac64.zero_()
            nat.call("nfx_batch_tg_coarse", g.nb, g.A, nat.ptr(d["agg_ptr"]), nat.ptr(d["agg_mem"]),
                     nat.ptr(d["agg_of"]), nat.ptr(d["nbr"]), nat.ptr(g.sd64), nat.ptr(g.so64), nat.ptr(d["arow"]),
                     nat.ptr(d["a0"]), nat.ptr(g.ac64), s)
            g.ainv.copy_(g.ac64)
            for j, (_, st, _) in enumerate(g.members):
                na = st["n_agg"]
                blk = g.ainv[int(g.coff[j]):int(g.coff[j + 1])].view(na, na)
                spd_inverse_(blk, st["plan"], stream)
        arr = (nat.BatchTG * len(self.groups))(*[g.desc() for g in self.groups])
        self._descs = arr
        return arr

    def apply(self, r, z=None, stream=None):
        """z = the preconditioner's two-grid part on the groups' rows of r (tests)."""
        t = torch()
        z = t.zeros_like(r) if z is None else z
        nat.call("nfx_batch_tg_apply", len(self._descs), self._descs, None, nat.ptr(r), nat.ptr(z),
                 nat.stream_handle(stream))
        return z
