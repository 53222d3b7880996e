"""L2 projection on the device.

``l2_projector(fe_space, functions, dof_shift=0, symmetric_solver_q=True)`` keeps the
signature of ``src/postprocess/projectors.py:8-91``: for every field of the product space,
the mass matrix and right-hand side of ``L2ProjectorWeakForm``
(``src/weak_forms/l2_projector_weak_form.py:8-93``) are assembled over the elements, with
the reference's value filter on every matrix contribution, and ``M alpha = b`` is solved.
The reference factorises M (Cholesky or LU by MUMPS); here M is assembled on the device in
both precisions (``nfx_assemble_projector``, a staged deterministic sum into CSR), and the
solve is FP32 MINRES with the Jacobi preconditioner inside FP64 iterative refinement on the
FP64 matrix, to ``||b - M alpha|| <= rtol ||b||`` (1e-12 by default; M is SPD and
Jacobi-preconditioned mass matrices are well conditioned). The functions are evaluated on
the host at the element quadrature points, as the reference evaluates its callables.

The pattern holds the (row, col) pairs of every element's same-field, same-component
entries; where the reference's filter drops all contributions to an entry (|value| within
1e-8 of 1e-16) the reference's PETSc pattern has no entry and this one holds an explicit
zero: the matrices are equal as operators.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _native as nat
from .device import p1_bary, to_dev, torch
from .pattern import build_pattern, row_codes, vector_slots
from .postprocess import _eval_exact


def _field_tables(sp, points, weights, vector, dim):
    tab = sp.element.tabulate(points)
    if vector:
        return np.concatenate([weights, tab[0, :, :, :dim].ravel()])
    return np.concatenate([weights, tab[0, :, :, 0].ravel()])


def l2_projector(fe_space, functions, dof_shift=0, symmetric_solver_q=True, rtol=1e-12, as_numpy=True,
                 device="cuda", info=None):
    """alpha with M alpha = b over all fields of fe_space (numpy FP64 by default; a device
    FP64 tensor with as_numpy=False). info: optional dict that receives the solve report.
    symmetric_solver_q selects the reference's SBAIJ + Cholesky or AIJ + LU; both solve the
    same symmetric system, so it changes nothing here."""
    from .solvers import MinresIR

    t = torch()
    n = int(fe_space.n_dof)
    shift_saved = fe_space.dof_shift
    fe_space.dof_shift = 0
    try:
        dest, sign = fe_space.reference_destinations()
    finally:
        fe_space.dof_shift = shift_saved
    dest = dest + int(dof_shift)
    if dest.size and (dest.min() < 0 or dest.max() >= n):
        raise IndexError("destination index outside the projected system (dof_shift too large)")
    E, nl = dest.shape
    # element-local field blocks and components
    blocks = []
    fid = np.zeros(nl, np.int64)
    cid = np.zeros(nl, np.int64)
    off = 0
    for k, (name, sp) in enumerate(fe_space.discrete_spaces.items()):
        nb, nc = sp.element.ndofs, sp.n_comp
        if len(sp.element_ids) != E:
            raise ValueError("the fields of the projected space must share their elements")
        fid[off:off + nb * nc] = k
        cid[off:off + nb * nc] = np.arange(nb * nc) % nc
        blocks.append((name, sp, off, nb, nc))
        off += nb * nc
    if off != nl:
        raise RuntimeError("element-local DoF count mismatch")
    same = (fid[:, None] == fid[None, :]) & (cid[:, None] == cid[None, :])
    pat = build_pattern(n, dest, sign, keep=np.broadcast_to(same, (E, nl, nl)))
    rslot, rsegptr, rtarget = vector_slots(n, dest, sign)
    dv = lambda a, dt=np.int32: to_dev(np.ascontiguousarray(np.asarray(a).astype(dt)), device=device)
    rowptr, col = dv(pat.rowptr), dv(pat.col)
    d_slot, d_rslot = dv(pat.slot), dv(rslot)
    segptr, target = dv(pat.segptr), dv(pat.target)
    d_rsegptr, d_rtarget = dv(rsegptr), dv(rtarget)
    vals64 = t.zeros(pat.nnz, dtype=t.float64, device=device)
    vals32 = t.zeros(pat.nnz, dtype=t.float32, device=device)
    st64 = t.zeros(max(pat.n_stage, 1), dtype=t.float64, device=device)
    st32 = t.zeros(max(pat.n_stage, 1), dtype=t.float32, device=device)
    b = t.zeros(n, dtype=t.float64, device=device)
    b32 = t.zeros(n, dtype=t.float32, device=device)
    rst64 = t.zeros(max(int(rsegptr[-1]), 1), dtype=t.float64, device=device)
    rst32 = t.zeros(max(int(rsegptr[-1]), 1), dtype=t.float32, device=device)
    s = nat.stream_handle()
    for name, sp, off, nb, nc in blocks:
        dim = sp.dimension
        points, weights = fe_space.quadrature[dim]
        vector = sp.family != "P"
        tables = _field_tables(sp, points, weights, vector, dim)
        mesh = sp.mesh
        cells = mesh.entities[dim].nodes[sp.topology.cell_index]
        xq = np.einsum("qa,eac->eqc", p1_bary(points), mesh.points[cells])
        fv = _eval_exact(functions[name], xq, nc, vector, dim)
        want = (E, len(weights), nc, dim) if vector else (E, len(weights), nc)
        if fv.shape != want:
            raise ValueError(f"function for field {name!r} has shape {fv.shape[2:]} per point, expected {want[2:]}")
        d_cells, d_coords, d_tab, d_f = dv(cells), dv(mesh.points, np.float64), dv(tables, np.float64), dv(fv, np.float64)
        kind = 0 if vector else 1
        for fp64, vals, st, rhs, rst in ((1, vals64, st64, b, rst64), (0, vals32, st32, b32, rst32)):
            nat.call("nfx_assemble_projector", kind, dim, nb, len(weights), nc, nl, off, E, nat.ptr(d_cells),
                     nat.ptr(d_coords), nat.ptr(d_tab), nat.ptr(d_f), nat.ptr(d_slot), nat.ptr(d_rslot), fp64,
                     nat.ptr(vals), nat.ptr(st), nat.ptr(rhs), nat.ptr(rst), s)
    for fp64, vals, st in ((1, vals64, st64), (0, vals32, st32)):
        nat.call("nfx_stage_reduce", len(pat.target), nat.ptr(segptr), nat.ptr(target), fp64, nat.ptr(st),
                 nat.ptr(vals), s)
    nat.call("nfx_stage_reduce", len(rtarget), nat.ptr(d_rsegptr), nat.ptr(d_rtarget), 1, nat.ptr(rst64), nat.ptr(b), s)
    diag = t.empty(n, dtype=t.float64, device=device)
    nat.call("nfx_csr_diag64", n, nat.ptr(rowptr), nat.ptr(col), nat.ptr(vals64), nat.ptr(diag), s)
    pinv = t.where(diag != 0, 1.0 / diag, t.ones_like(diag)).float()
    solver = MinresIR(rowptr, col, vals32, vals64, pinv=pinv)
    try:
        x, rep = solver.solve(b, rtol=rtol, inner_rtol=1e-6, max_inner=20000, max_outer=30)
    finally:
        solver.close()
    if not rep["converged"]:
        raise RuntimeError(f"l2_projector: solve did not reach {rtol:.1e} ({rep['residual']:.2e})")
    if info is not None:
        info.update(rep)
        info.update(nnz=pat.nnz, rowptr=pat.rowptr, col=pat.col, vals64=vals64, vals32=vals32, b=b)
    return x.cpu().numpy() if as_numpy else x


def l2_error_projected(dim, fe_space, alpha, skip_fields=(), dof_shift=0):
    """The L2 norm of each discrete field (``src/postprocess/l2_error_post_processor.py:66-117``):
    sqrt(sum_e sum_q detJ w |f_h(x_q)|^2), vector fields in the Frobenius norm over
    (row, direction), Lagrange fields over their first dim components (as the reference's
    ``alpha_star[:, 0:dim]``). FP64 on the device (``nfx_field_error`` with a zero exact
    field)."""
    t = torch()
    out = []
    a = alpha if isinstance(alpha, t.Tensor) else t.from_numpy(np.asarray(alpha, dtype=np.float64))
    a = a.to(device="cuda", dtype=t.float64).contiguous()
    for name, sp in fe_space.discrete_spaces.items():
        if name in skip_fields:
            continue
        if sp.dimension != dim:
            out.append(0.0)  # no element of this dimension: the reference's sum is empty
            continue
        points, weights = fe_space.quadrature[dim]
        el = sp.element
        tab = el.tabulate(points)
        vector = sp.family != "P"
        if vector:
            div = np.einsum("kqjk->qj", tab[1:])
            tables = np.concatenate([weights, tab[0].ravel(), div.ravel()])
            nc = sp.n_comp
        else:
            tables = np.concatenate([weights, tab[0, :, :, 0].ravel()])
            nc = min(sp.n_comp, dim)
        mesh = sp.mesh
        cells = mesh.entities[dim].nodes[sp.topology.cell_index]
        nb = el.ndofs
        shift = fe_space.strides[name] + fe_space.dof_shift + int(dof_shift)
        codes = row_codes(sp.ref_dest + shift, sp.ref_sign).reshape(len(cells), nb, sp.n_comp)
        codes = np.ascontiguousarray(codes[:, :, :nc])
        if int((codes & nat.SLOT_INDEX_MASK).max(initial=0)) >= a.numel():
            raise ValueError("alpha is shorter than the space's DoF range")
        zero = t.zeros(len(cells) * len(weights) * nc * (dim if vector else 1), dtype=t.float64, device="cuda")
        res = (ctypes.c_double * 1)()
        dv = lambda x, dt: to_dev(np.ascontiguousarray(np.asarray(x).astype(dt)), device="cuda")
        d_cells, d_coords = dv(cells, np.int32), dv(mesh.points, np.float64)
        d_tab, d_rows = dv(tables, np.float64), dv(codes, np.int32)
        nat.call("nfx_field_error", 0 if vector else 1, dim, nb, len(weights), nc, len(cells), nat.ptr(d_cells),
                 nat.ptr(d_coords), nat.ptr(d_tab), nat.ptr(d_rows), nat.ptr(a), nat.ptr(zero), res,
                 nat.stream_handle())
        out.append(float(np.sqrt(res[0])))
    return out
