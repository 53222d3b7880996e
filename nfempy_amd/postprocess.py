"""FP64 error norms on the device.

``l2_error(dim, fe_space, functions, alpha, skip_fields=[])`` and
``div_error(...)`` keep the signatures of
``src/postprocess/l2_error_post_processor.py:7-63,185-237``: per field,
``sqrt(sum_e sum_q detJ w |f_exact(x_q) - f_h(x_q)|^2)`` with vector fields
measured in the Frobenius norm over (row, direction) and ``div f_h =
tr(grad_ref phi) / detJ``. Exact fields are evaluated on the host at the
element quadrature points (as the reference does), the discrete field and the
per-element sums run in FP64 through ``nfx_field_error``.
"""

from __future__ import annotations

import numpy as np

from . import _native as nat
from .device import p1_bary, to_dev, torch
from .pattern import row_codes


def _eval_exact(exact, xq, n_out, vector, dim):
    """exact at points xq (E, nq, 3) -> (E, nq, n_out[, dim]) FP64."""
    E, nq, _ = xq.shape
    x = xq.reshape(-1, 3)
    npts = len(x)
    try:
        v = np.asarray(exact(x[:, 0], x[:, 1], x[:, 2]), dtype=np.float64)
        if vector:
            # (n_comp, dim(+), npts) -> (npts, n_comp, dim)
            v = np.moveaxis(v.reshape(v.shape[0], v.shape[1], -1), -1, 0)[:, :n_out, :dim]
        else:
            v = np.moveaxis(v.reshape(v.shape[0], -1), -1, 0)[:, :n_out]
        if v.shape[0] != npts:
            raise ValueError
    except Exception:
        v = np.array([np.asarray(exact(p[0], p[1], p[2]), dtype=np.float64) for p in x])
        v = v[:, :n_out, :dim] if vector else v.reshape(npts, -1)[:, :n_out]
    return v.reshape((E, nq) + v.shape[1:])


def _field_error(kind, dim, fe_space, name, exact, alpha):
    t = torch()
    sp = fe_space.discrete_spaces[name]
    el = sp.element
    mesh = sp.mesh
    points, weights = fe_space.quadrature[dim]
    tab = el.tabulate(points)
    if kind == 1:
        tables = np.concatenate([weights, tab[0, :, :, 0].ravel()])
    else:
        div = np.einsum("kqjk->qj", tab[1:])
        tables = np.concatenate([weights, tab[0].ravel(), div.ravel()])
    cells = mesh.entities[dim].nodes[sp.topology.cell_index]
    X = mesh.points[cells]
    xq = np.einsum("qa,eac->eqc", p1_bary(points), X)
    nb = el.ndofs
    nc_all = sp.n_comp
    shift = fe_space.strides[name] + fe_space.dof_shift
    codes = row_codes(sp.ref_dest + shift, sp.ref_sign).reshape(len(cells), nb, nc_all)
    if kind == 1:
        nc = min(nc_all, dim)
        ex = _eval_exact(exact, xq, nc, False, dim)
    elif kind == 0:
        nc = nc_all
        ex = _eval_exact(exact, xq, nc, True, dim)
    else:
        nc = nc_all
        ex = _eval_exact(exact, xq, nc, False, dim)
    codes = np.ascontiguousarray(codes[:, :, :nc])
    dev = "cuda"
    a = alpha if isinstance(alpha, t.Tensor) else t.from_numpy(np.asarray(alpha, dtype=np.float64))
    a = a.to(device=dev, dtype=t.float64).contiguous()
    if int((codes & nat.SLOT_INDEX_MASK).max(initial=0)) >= a.numel():
        raise ValueError("alpha is shorter than the space's DoF range")
    out = (nat.ctypes.c_double * 1)()
    d_cells = to_dev(cells.astype(np.int32), device=dev)
    d_coords = to_dev(mesh.points, device=dev)
    d_tab = to_dev(tables, device=dev)
    d_rows = to_dev(codes.astype(np.int32), device=dev)
    d_ex = to_dev(np.ascontiguousarray(ex), device=dev)
    nat.call("nfx_field_error", kind, dim, nb, len(weights), nc, len(cells), nat.ptr(d_cells), nat.ptr(d_coords),
             nat.ptr(d_tab), nat.ptr(d_rows), nat.ptr(a), nat.ptr(d_ex), out, nat.stream_handle())
    return float(np.sqrt(out[0]))


def l2_error(dim, fe_space, functions, alpha, skip_fields=()):
    errs = []
    for name, sp in fe_space.discrete_spaces.items():
        if name in skip_fields:
            continue
        kind = 1 if sp.family == "P" else 0
        errs.append(_field_error(kind, dim, fe_space, name, functions[name], alpha))
    return errs


def div_error(dim, fe_space, functions, alpha, skip_fields=()):
    errs = []
    for name, sp in fe_space.discrete_spaces.items():
        if sp.family not in ("RT", "BDM") or name in skip_fields:
            continue
        errs.append(_field_error(2, dim, fe_space, name, functions["div_" + name], alpha))
    return errs
