"""Krylov solves on the device: MINRES in FP32 inside FP64 iterative refinement.

Replaces the reference's PETSc KSP blocks:
- Darcy: ``preonly`` + LU(MUMPS) on ``[[M, -B^T], [B, 0]] x = -rg``
  (``examples/md_elliptic_scalar/md_elliptic_scalar.py:351-364``);
- elasticity: MINRES + PCFIELDSPLIT additive
  (``examples/cauchy_elasticity/mfe_elasticity_example_1.py:238-274``).

Every answer is judged by the FP64 residual ``||b - A x||_2 / ||b||_2`` of the
as-assembled FP64 matrix.

Block preconditioner for the Darcy saddle point (``darcy_preconditioner``): the
pressure rows are negated so that ``D A`` is symmetric indefinite, and MINRES runs
with the SPD block-diagonal preconditioner ``diag(M)^-1`` on the flux rows and
``|S|^-1`` on the pressure rows, ``S = diag(B diag(M)^-1 B^T)`` evaluated in FP64.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _native as nat
from .device import torch


class MinresIR:
    """FP32 MINRES + FP64 refinement on one CSR operator resident on the device."""

    def __init__(self, rowptr, col, vals32, vals64, sign=None, pinv=None, stream=None, pc=None, block3=None,
                 tg=None):
        """pinv: diagonal preconditioner (FP32 inverse diagonal) or pc: (rowptr, col, val)
        CSR of a sparse SPD preconditioner P^-1 (FP32 values), or tg: an ``nfx_elasticity_tg``
        descriptor (the elasticity two-grid preconditioner, applied as a function). block3:
        optional (n_blocks, browptr, bcol, bval), a 3 x 3 block CSR copy of the FP32 operator
        that the inner MINRES multiplies with (nfx_solver_set_block3)."""
        t = torch()
        self.n = int(rowptr.numel() - 1)
        self.nnz = int(col.numel())
        self.rowptr, self.col, self.vals32, self.vals64 = rowptr, col, vals32, vals64
        self.sign = sign
        if pc is None:
            dev = rowptr.device
            diag = pinv if pinv is not None else t.ones(self.n, dtype=t.float32, device=dev)
            ar = t.arange(self.n + 1, dtype=t.int32, device=dev)
            pc = (ar, ar[:-1].contiguous(), diag.to(t.float32).contiguous())
        self.pc = pc
        self.tg = tg
        if tg is not None and sign is not None:
            raise ValueError("the two-grid preconditioner solves the symmetric operator (sign must be None)")
        self.stream = stream
        h = ctypes.c_void_p()
        nat.call("nfx_solver_create", ctypes.byref(h), self.n, self.nnz, nat.stream_handle(stream))
        self._h = h
        self.block3 = block3
        if block3 is not None:
            nb, brp, bcol, bval = block3
            nat.call("nfx_solver_set_block3", h, int(nb), nat.ptr(brp), nat.ptr(bcol), nat.ptr(bval))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            nat.lib().nfx_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, b, x=None, rtol=1e-8, inner_rtol=1e-5, max_inner=20000, max_outer=30, raise_on_fail=False):
        t = torch()
        b = b.to(dtype=t.float64).contiguous()
        if x is None:
            x = t.zeros(self.n, dtype=t.float64, device=b.device)
        info = (ctypes.c_int * 2)()
        resid = (ctypes.c_double * 1)()
        hist = (ctypes.c_double * (max_outer + 1))()
        if self.tg is not None:
            name = "nfx_solve_minres_ir_tg"
            rc = nat.lib().nfx_solve_minres_ir_tg(
                self._h, nat.ptr(self.rowptr), nat.ptr(self.col), nat.ptr(self.vals32), nat.ptr(self.vals64),
                ctypes.byref(self.tg), nat.ptr(b), nat.ptr(x), rtol, inner_rtol, max_inner, max_outer, info, resid,
                hist, max_outer + 1)
        else:
            name = "nfx_solve_minres_ir"
            rc = nat.lib().nfx_solve_minres_ir(
                self._h, nat.ptr(self.rowptr), nat.ptr(self.col), nat.ptr(self.vals32), nat.ptr(self.vals64),
                nat.ptr(self.sign), nat.ptr(self.pc[0]), nat.ptr(self.pc[1]), nat.ptr(self.pc[2]), nat.ptr(b),
                nat.ptr(x), rtol, inner_rtol, max_inner,
                max_outer, info, resid, hist, max_outer + 1)
        if rc not in (nat.NFX_OK, nat.NFX_E_NOT_CONVERGED) or (raise_on_fail and rc != nat.NFX_OK):
            nat.check(rc, name)
        out = {
            "converged": rc == nat.NFX_OK,
            "outer": info[0],
            "iterations": info[1],
            "residual": resid[0],
            "history": [hist[i] for i in range(min(info[0] + 1, max_outer + 1))],
        }
        return x, out


class MinresBatch:
    """Batched FP32 MINRES + FP64 refinement over independent systems (``nfx_minres_batch``):
    one set of launches per iteration for all instances, each with its own recurrence,
    stopping test and refinement. ``systems``: dicts with rowptr, col, vals32, vals64 and
    optional sign / pinv (the symmetric form D A and a diagonal preconditioner)."""

    def __init__(self, systems, stream=None):
        self._systems = list(systems)
        m = len(self._systems)
        arr = (nat.MinresBatchItem * m)()
        self.sizes = []
        for i, sy in enumerate(self._systems):
            n = int(sy["rowptr"].numel() - 1)
            for key in ("vals32", "vals64"):
                if int(sy[key].numel()) != int(sy["col"].numel()):
                    raise ValueError(f"instance {i}: {key} size differs from the pattern")
            for key in ("sign", "pinv"):
                if sy.get(key) is not None and int(sy[key].numel()) != n:
                    raise ValueError(f"instance {i}: {key} size differs from the rows")
            arr[i] = nat.MinresBatchItem(n, 0, 0, nat.ptr(sy["rowptr"]), nat.ptr(sy["col"]), nat.ptr(sy["vals32"]),
                                         nat.ptr(sy["vals64"]), nat.ptr(sy.get("sign")), nat.ptr(sy.get("pinv")))
            self.sizes.append(n)
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        h = ctypes.c_void_p()
        nat.call("nfx_minres_batch_create", ctypes.byref(h), m, arr, nat.stream_handle(stream))
        self._h = h
        if nat.lib().nfx_minres_batch_rows(h) != self.offsets[-1]:
            raise RuntimeError("batch row count differs between host and library")

    def set_pc(self, groups):
        """Attach the two-grid groups (a ctypes array of ``nat.BatchTG``, or None to detach)."""
        self._pc = groups
        n = 0 if groups is None else len(groups)
        nat.call("nfx_minres_batch_set_pc", self._h, n, groups if n else None)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            nat.lib().nfx_minres_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, b, x=None, rtol=1e-8, inner_rtol=1e-5, max_inner=100000, max_outer=30):
        """b: concatenated FP64 right-hand sides (instance order). Returns (x, infos)."""
        t = torch()
        b = b.to(dtype=t.float64).contiguous()
        if int(b.numel()) != int(self.offsets[-1]):
            raise ValueError("right-hand side length differs from the batch")
        if x is None:
            x = t.zeros_like(b)
        m = len(self.sizes)
        iters = (ctypes.c_int * m)()
        outer = (ctypes.c_int * m)()
        resid = (ctypes.c_double * m)()
        rc = nat.lib().nfx_solve_minres_batch_ir(self._h, nat.ptr(b), nat.ptr(x), rtol, inner_rtol, max_inner,
                                                 max_outer, iters, outer, resid)
        if rc not in (nat.NFX_OK, nat.NFX_E_NOT_CONVERGED):
            nat.check(rc, "nfx_solve_minres_batch_ir")
        infos = [{"iterations": iters[i], "outer": outer[i], "residual": resid[i], "converged": resid[i] <= rtol}
                 for i in range(m)]
        return x, infos


def darcy_preconditioner(rowptr, col, vals64, n, n_u, stream=None):
    """(sign, pinv) device FP32 vectors of the Darcy block preconditioner."""
    t = torch()
    s = nat.stream_handle(stream)
    dev = rowptr.device
    diag = t.empty(n, dtype=t.float64, device=dev)
    nat.call("nfx_csr_diag64", n, nat.ptr(rowptr), nat.ptr(col), nat.ptr(vals64), nat.ptr(diag), s)
    schur = t.empty(max(n - n_u, 1), dtype=t.float64, device=dev)
    nat.call("nfx_schur_diag64", n, n_u, nat.ptr(rowptr), nat.ptr(col), nat.ptr(vals64), nat.ptr(diag),
             nat.ptr(schur), s)
    pinv = t.empty(n, dtype=t.float64, device=dev)
    pinv[:n_u] = 1.0 / diag[:n_u].abs()
    sa = schur[: n - n_u].abs()
    pinv[n_u:] = t.where(sa > 0, 1.0 / sa, t.ones_like(sa))
    sign = t.ones(n, dtype=t.float32, device=dev)
    sign[n_u:] = -1.0
    return sign, pinv.to(t.float32)
