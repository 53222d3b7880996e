"""KSP-like solver facade over the device Krylov solver.

Mirrors the petsc4py calls of the reference drivers:
``KSP().create()``, ``setOperators(A[, P])``, ``setType``, ``getPC().setType`` /
``setFactorSolverType`` / ``setFieldSplitIS`` / ``setFieldSplitType`` /
``getFieldSplitSubKSP``, ``setTolerances``, ``setConvergenceHistory``,
``setFromOptions``, ``solve(b, x)``, ``getConvergenceHistory``,
``getConvergedReason`` (``examples/md_elliptic_scalar/md_elliptic_scalar.py:351-364``,
``examples/cauchy_elasticity/mfe_elasticity_example_1.py:238-276``).

Whatever PC the caller selects, the solve is FP32 MINRES inside FP64 iterative
refinement on the device (``nfx_solve_minres_ir``):

- "preonly" + LU / Cholesky (a direct solve in the reference) becomes an iterative solve
  to an FP64 relative residual of 1e-8 unless ``setTolerances`` asks for another target,
  with the Darcy block preconditioner or the elasticity block preconditioner; a
  ``RuntimeWarning`` says so once per solver;
- "minres" + "fieldsplit" keeps MINRES with the elasticity block-diagonal preconditioner.
  The split must be the reference's (stress DoFs; displacement and rotation DoFs, two
  contiguous ranges at the field boundaries), additive; anything else raises
  ``NotImplementedError`` at ``solve``.

Tolerances follow PETSc's convergence test on the unpreconditioned FP64 residual 2-norm:
converged when ``||b - A x|| <= max(rtol ||b||, atol)``; diverged when
``||b - A x|| > divtol ||b - A x_0||``; ``max_it`` caps the total number of FP32 MINRES
iterations over all refinement steps. ``rtol = 0`` is honoured (pure ``atol`` test);
``None`` keeps the current value, as ``PETSC_DEFAULT`` does.
"""

from __future__ import annotations

import warnings

import numpy as np

# KSPConvergedReason values of PETSc
CONVERGED_RTOL = 2
CONVERGED_ATOL = 3
DIVERGED_ITS = -3
DIVERGED_DTOL = -4


class _PC:
    def __init__(self):
        self.type = None
        self.factor = None
        self.splits = []
        self.split_type = None
        self.sub = []

    def setType(self, t):
        self.type = str(t).lower()

    def getType(self):
        return self.type

    def setFactorSolverType(self, t):
        self.factor = str(t).lower()

    def setFieldSplitIS(self, *pairs):
        self.splits = [(name, np.asarray(getattr(iset, "indices", iset))) for name, iset in pairs]

    def setFieldSplitType(self, t):
        self.split_type = t

    def getFieldSplitSubKSP(self):
        self.sub = [KSP() for _ in self.splits]
        return self.sub


class IS:
    """``PETSc.IS`` stand-in: ``createGeneral(indices)``."""

    def __init__(self):
        self.indices = np.zeros(0, np.int64)

    def createGeneral(self, indices, comm=None):
        self.indices = np.asarray(indices, dtype=np.int64)
        return self


class KSP:
    def __init__(self):
        self.A = None
        self.P = None
        self.type = "gmres"
        self.pc = _PC()
        self.rtol = 1e-8
        self.atol = 0.0
        self.divtol = 1e4
        self.max_it = 20000
        self.history = []
        self.info = None
        self.reason = 0
        self._warned = False

    def create(self, comm=None):
        return self

    def setOperators(self, A, P=None):
        self.A, self.P = A, P

    def setType(self, t):
        self.type = str(t).lower()

    def getType(self):
        return self.type

    def getPC(self):
        return self.pc

    def setTolerances(self, rtol=None, atol=None, divtol=None, max_it=None):
        for name, v in (("rtol", rtol), ("atol", atol), ("divtol", divtol)):
            if v is not None:
                if not np.isfinite(v) or v < 0:
                    raise ValueError(f"{name} must be a finite non-negative number")
                setattr(self, name, float(v))
        if max_it is not None:
            if int(max_it) <= 0:
                raise ValueError("max_it must be positive")
            self.max_it = int(max_it)

    def getTolerances(self):
        return self.rtol, self.atol, self.divtol, self.max_it

    def setConvergenceHistory(self):
        self.history = []

    def setFromOptions(self):
        pass

    # -- checks of what the device solver can honour ---------------------------------
    def _check_split(self, sysd):
        pc = self.pc
        if pc.type != "fieldsplit":
            return
        bounds = getattr(sysd, "field_bounds", None)
        if bounds is None or len(pc.splits) != 2:
            raise NotImplementedError("field split: only the stress | (displacement, rotation) split of the "
                                      "elasticity system is supported")
        (_, i0), (_, i1) = pc.splits
        s_end = bounds[1]
        ok = (np.array_equal(i0, np.arange(0, s_end)) and np.array_equal(i1, np.arange(s_end, bounds[-1])))
        if not ok:
            raise NotImplementedError("field split: the index sets must be the stress DoFs and the "
                                      "displacement + rotation DoFs (contiguous field ranges)")
        st = pc.split_type
        if st is not None and str(st).lower() not in ("additive", "0"):
            raise NotImplementedError(f"field split type {st!r}: only ADDITIVE is supported")

    def _note_substitution(self):
        direct = self.pc.type in ("lu", "cholesky") or any(
            k.pc.type in ("lu", "cholesky", "icc", "ilu") for k in self.pc.sub)
        if direct and not self._warned:
            warnings.warn("direct / incomplete factorisations run as FP32 MINRES with a block preconditioner "
                          "inside FP64 iterative refinement to the requested tolerance", RuntimeWarning,
                          stacklevel=3)
            self._warned = True

    def solve(self, b, x):
        import torch

        from .solvers import MinresIR, darcy_preconditioner

        rowptr, col, v32, v64 = self.A.device_csr
        sysd = self.A._sys
        n = sysd.n
        self._check_split(sysd)
        self._note_substitution()
        if hasattr(sysd, "preconditioner"):
            sign, pc = sysd.preconditioner(self.P)
            solver = MinresIR(rowptr, col, v32, v64, sign, pc=pc)
        else:
            sign, pinv = darcy_preconditioner(rowptr, col, v64, n, sysd.n_u)
            solver = MinresIR(rowptr, col, v32, v64, sign, pinv)
        dev = rowptr.device
        bd = torch.as_tensor(np.asarray(b.array), dtype=torch.float64, device=dev)
        xd = torch.as_tensor(np.asarray(x.array), dtype=torch.float64, device=dev).clone()
        bn = float(torch.linalg.norm(bd))
        if self.rtol == 0.0 and self.atol == 0.0:
            raise ValueError("rtol and atol are both zero: no convergence test")
        target = max(self.rtol * bn, self.atol)
        rel_target = target / bn if bn > 0 else 0.0
        # refinement one step per call so that max_it caps the total inner iterations
        hist, total, reason, r0 = [], 0, 0, None
        try:
            for _ in range(64):
                budget = self.max_it - total
                xd, info = solver.solve(bd, xd, rtol=rel_target, max_inner=max(budget, 1), max_outer=1)
                h = [v * bn for v in info["history"]]
                if r0 is None:
                    r0 = h[0]
                    hist.append(h[0])
                hist.extend(h[1:])
                total += info["iterations"]
                res = h[-1]
                if res <= target:
                    reason = CONVERGED_RTOL if res <= self.rtol * bn else CONVERGED_ATOL
                    break
                if r0 > 0 and res > self.divtol * r0:
                    reason = DIVERGED_DTOL
                    break
                if total >= self.max_it or info["iterations"] == 0:
                    reason = DIVERGED_ITS
                    break
            else:
                reason = DIVERGED_ITS
        finally:
            solver.close()
        x.array[:] = xd.cpu().numpy()
        self.reason = reason
        self.info = {"iterations": total, "residual": hist[-1] / bn if bn > 0 else hist[-1],
                     "converged": reason > 0, "history": hist}
        self.history = hist

    def getConvergedReason(self):
        return self.reason

    def getConvergenceHistory(self):
        return np.asarray(self.history)

    def getIterationNumber(self):
        return 0 if self.info is None else self.info["iterations"]

    def getResidualNorm(self):
        return 0.0 if not self.history else float(self.history[-1])

    def destroy(self):
        pass
