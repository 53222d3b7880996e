"""Flexible GMRES on the device, inside FP64 iterative refinement.

Right-preconditioned FGMRES(m) (Saad 1993): the preconditioner may change from
one iteration to the next (the learned correction is a nonlinear map), so the
preconditioned directions Z are kept next to the Arnoldi basis V. Vectors are
FP32 in HBM; the operator is the FP32 CSR through ``nfx_spmv``; Gram-Schmidt
coefficients are computed as dense products over the basis (classical
Gram-Schmidt applied twice), the small Hessenberg least-squares problem on the
host in FP64. The outer FP64 refinement evaluates ``b - A x`` with the FP64
matrix (``nfx_residual64``) and stops at ``rtol``.
"""

from __future__ import annotations

import numpy as np

from . import _native as nat
from .device import torch


class CsrOperator:
    def __init__(self, rowptr, col, vals32, vals64):
        self.rowptr, self.col, self.vals32, self.vals64 = rowptr, col, vals32, vals64
        self.n = int(rowptr.numel() - 1)

    def matvec32(self, x, y):
        nat.call("nfx_spmv", self.n, nat.ptr(self.rowptr), nat.ptr(self.col), nat.ptr(self.vals32), 1.0,
                 nat.ptr(x), 0.0, nat.ptr(y), nat.stream_handle())
        return y

    def residual64(self, x, b, r):
        nat.call("nfx_residual64", self.n, nat.ptr(self.rowptr), nat.ptr(self.col), nat.ptr(self.vals64),
                 nat.ptr(x), nat.ptr(b), nat.ptr(r), nat.stream_handle())
        return r


def fgmres32(op, precond, b, tol=1e-5, restart=60, maxit=20000):
    """FP32 FGMRES(restart) for op x = b from x = 0. Returns (x, iterations, rel. residual)."""
    t = torch()
    n = op.n
    dev = b.device
    m = restart
    V = t.zeros((m + 1, n), dtype=t.float32, device=dev)
    Z = t.zeros((m, n), dtype=t.float32, device=dev)
    w = t.empty(n, dtype=t.float32, device=dev)
    x = t.zeros(n, dtype=t.float32, device=dev)
    bnorm = float(t.linalg.vector_norm(b.double()))
    if bnorm == 0:
        return x, 0, 0.0
    r = b.clone()
    beta = bnorm
    it = 0
    rel = 1.0
    while it < maxit:
        V[0].copy_(r / beta)
        H = np.zeros((m + 1, m))
        g = np.zeros(m + 1)
        g[0] = beta
        cs, sn = np.zeros(m), np.zeros(m)
        j_end = 0
        for j in range(m):
            precond(V[j], out=Z[j])
            op.matvec32(Z[j], w)
            # classical Gram-Schmidt, twice
            h = t.mv(V[: j + 1], w)
            w -= t.mv(V[: j + 1].T, h)
            h2 = t.mv(V[: j + 1], w)
            w -= t.mv(V[: j + 1].T, h2)
            hh = (h + h2).double().cpu().numpy()
            hn = float(t.linalg.vector_norm(w.double()))
            H[: j + 1, j] = hh
            H[j + 1, j] = hn
            if hn > 0:
                V[j + 1].copy_(w / hn)
            for i in range(j):
                tmp = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                H[i + 1, j] = -sn[i] * H[i, j] + cs[i] * H[i + 1, j]
                H[i, j] = tmp
            den = np.hypot(H[j, j], H[j + 1, j])
            cs[j], sn[j] = (1.0, 0.0) if den == 0 else (H[j, j] / den, H[j + 1, j] / den)
            H[j, j] = cs[j] * H[j, j] + sn[j] * H[j + 1, j]
            H[j + 1, j] = 0.0
            g[j + 1] = -sn[j] * g[j]
            g[j] = cs[j] * g[j]
            it += 1
            j_end = j + 1
            rel = abs(g[j + 1]) / bnorm
            if rel <= tol or it >= maxit or hn == 0:
                break
        y = np.linalg.solve(np.triu(H[:j_end, :j_end]), g[:j_end])
        x += t.mv(Z[:j_end].T, t.from_numpy(y.astype(np.float32)).to(dev))
        if rel <= tol or it >= maxit:
            break
        op.matvec32(x, w)
        r = b - w
        beta = float(t.linalg.vector_norm(r.double()))
        rel = beta / bnorm
        if rel <= tol:
            break
    return x, it, rel


def fgmres_ir(op, precond, b64, rtol=1e-8, inner_tol=1e-5, restart=60, maxit=20000, max_outer=20):
    """FP64-refined FGMRES: returns (x64, info)."""
    t = torch()
    n = op.n
    x = t.zeros(n, dtype=t.float64, device=b64.device)
    r = t.empty_like(x)
    bn = float(t.linalg.vector_norm(b64))
    hist, total = [], 0
    outer = 0
    rel = 1.0
    for outer in range(max_outer + 1):
        op.residual64(x, b64, r)
        rn = float(t.linalg.vector_norm(r))
        rel = rn / bn if bn > 0 else rn
        hist.append(rel)
        if rel <= rtol or outer == max_outer:
            break
        d, it, _ = fgmres32(op, precond, (r / rn).to(t.float32), tol=inner_tol, restart=restart, maxit=maxit)
        total += it
        x += rn * d.double()
    return x, {"converged": rel <= rtol, "outer": outer, "iterations": total, "residual": rel, "history": hist}
