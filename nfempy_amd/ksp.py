"""KSP-like solver facade over the device Krylov solver.

Mirrors the petsc4py calls of the reference drivers:
``KSP().create()``, ``setOperators(A[, P])``, ``setType``, ``getPC().setType`` /
``setFactorSolverType`` / ``setFieldSplitIS`` / ``setFieldSplitType`` /
``getFieldSplitSubKSP``, ``setTolerances``, ``setConvergenceHistory``,
``setFromOptions``, ``solve(b, x)``, ``getConvergenceHistory``
(``examples/md_elliptic_scalar/md_elliptic_scalar.py:351-364``,
``examples/cauchy_elasticity/mfe_elasticity_example_1.py:238-276``).

Whatever PC the caller selects, the solve is FP32 MINRES inside FP64 iterative
refinement on the device (``nfx_solve_minres_ir``), stopped at the FP64 relative
residual ``rtol`` (default 1e-8): "preonly + LU" becomes an iterative solve
with the Darcy block preconditioner, "minres + fieldsplit" keeps MINRES with a
block-diagonal preconditioner built from the preconditioning operator.
"""

from __future__ import annotations

import numpy as np


class _PC:
    def __init__(self):
        self.type = None
        self.factor = None
        self.splits = []
        self.split_type = None

    def setType(self, t):
        self.type = t

    def setFactorSolverType(self, t):
        self.factor = t

    def setFieldSplitIS(self, *pairs):
        self.splits = list(pairs)

    def setFieldSplitType(self, t):
        self.split_type = t

    def getFieldSplitSubKSP(self):
        return [KSP() for _ in self.splits]


class KSP:
    def __init__(self):
        self.A = None
        self.P = None
        self.type = "gmres"
        self.pc = _PC()
        self.rtol = 1e-8
        self.max_it = 20000
        self.history = []
        self.info = None

    def create(self, comm=None):
        return self

    def setOperators(self, A, P=None):
        self.A, self.P = A, P

    def setType(self, t):
        self.type = t

    def getPC(self):
        return self.pc

    def setTolerances(self, rtol=None, atol=None, divtol=None, max_it=None):
        if rtol:
            self.rtol = rtol
        if max_it:
            self.max_it = max_it

    def setConvergenceHistory(self):
        self.history = []

    def setFromOptions(self):
        pass

    def solve(self, b, x):
        import torch

        from .solvers import MinresIR, darcy_preconditioner

        rowptr, col, v32, v64 = self.A.device_csr
        sysd = self.A._sys
        n = sysd.n
        if hasattr(sysd, "preconditioner"):
            sign, pc = sysd.preconditioner(self.P)
            solver = MinresIR(rowptr, col, v32, v64, sign, pc=pc)
        else:
            sign, pinv = darcy_preconditioner(rowptr, col, v64, n, sysd.n_u)
            solver = MinresIR(rowptr, col, v32, v64, sign, pinv)
        bd = torch.as_tensor(np.asarray(b.array), dtype=torch.float64, device=rowptr.device)
        xd, info = solver.solve(bd, rtol=self.rtol, max_inner=self.max_it)
        solver.close()
        x.array[:] = xd.cpu().numpy()
        self.info = info
        self.history = info["history"]

    def getConvergenceHistory(self):
        return np.asarray(self.history)

    def getIterationNumber(self):
        return 0 if self.info is None else self.info["iterations"]

    def destroy(self):
        pass
