"""Weak forms: the drop-in surface of ``src/weak_forms`` for the hot path.

``WeakForm(space)`` with ``.space``, ``.functions`` and
``evaluate_form(element_index, alpha) -> (r_el, j_el)`` as in
``src/weak_forms/weak_from.py:6-29``; ``alpha`` is accepted and ignored exactly
where the reference ignores it (the mixed Darcy form resets it to zero,
``laplace_dual_weak_form.py:46``; the forms are linear).

Unlike the reference, the element kernels run for all elements at once on the
GPU: ``evaluate_form`` serves element ``i`` from one cached batched evaluation,
and ``scatter(jacobian, residual)`` (used by ``SequentialAssembler``) writes the
whole CSR system in HBM through the sorted-segment scatter.

Forms:
- ``LaplaceDualWeakForm`` / ``LaplaceDualWeakFormBCDirichlet``
  (``src/weak_forms/laplace_dual_weak_form.py:10-154``), RT_k - P_{k-1}, k = 1, 2.
- ``LEDualWeakForm`` / ``LEDualWeakFormBCDirichlet`` / ``LERieszMapWeakForm``
  (``src/weak_forms/le_dual_weak_form.py``, ``le_riesz_map_weak_form.py``),
  BDM1 - P0 - P0 in 3-D (see ``nfempy_amd.elasticity``).
"""

from __future__ import annotations

from abc import ABC, abstractmethod

import numpy as np


class WeakForm(ABC):
    def __init__(self, space):
        self.space = space
        self._functions = None
        self._cache = None

    @property
    def space(self):
        return self._space

    @space.setter
    def space(self, space):
        self._space = space
        self._cache = None

    @property
    def functions(self):
        return self._functions

    @functions.setter
    def functions(self, functions):
        self._functions = functions
        self._cache = None

    @abstractmethod
    def evaluate_form(self, element_index, alpha):
        pass


def _darcy_system(space):
    from .device import DarcySystem

    sysd = getattr(space, "_nfx_darcy_system", None)
    if sysd is None:
        sysd = DarcySystem(space)
        space._nfx_darcy_system = sysd
    return sysd


class LaplaceDualWeakForm(WeakForm):
    """Mixed Darcy element: ``functions = {"rhs": f, "kappa": K}``."""

    def _check(self):
        if self.space is None or self.functions is None:
            raise ValueError

    def evaluate_form(self, element_index, alpha=None):
        self._check()
        if self._cache is None:
            sysd = _darcy_system(self.space)
            J, r = sysd.local_matrices(self.functions["kappa"], self.functions["rhs"], fp64=True)
            self._cache = (J, r)
        J, r = self._cache
        return r[element_index].copy(), J[element_index].copy()

    def scatter(self, jacobian, residual):
        """Assemble every element into ``jacobian`` (a ``Mat``) and ``residual`` (numpy rg).

        FP64 and FP32 values are both produced (the FP64 copy judges the solve)."""
        self._check()
        sysd = _darcy_system(self.space)
        sysd.assemble_elements(self.functions["kappa"], self.functions["rhs"], fp64=True)
        sysd.assemble_elements(self.functions["kappa"], self.functions["rhs"], fp64=False)
        if jacobian is not None:
            jacobian._attach(sysd)
        if residual is not None:
            residual += sysd.rhs64.cpu().numpy()
        return sysd


class LaplaceDualWeakFormBCDirichlet(WeakForm):
    """Dirichlet pressure facet term: ``functions = {"p": p_D, ...}``."""

    def evaluate_form(self, element_index, alpha=None):
        r = self.evaluate_all()[element_index]
        return r.copy(), np.zeros((len(r), len(r)))

    def scatter(self, jacobian, residual):
        """Facet residuals into the device rg (both precisions) and the host ``residual``."""
        sysd = _darcy_system(self.space)
        if not sysd.F:
            return
        sysd.assemble_dirichlet(self.functions["p"], fp64=True)
        sysd.assemble_dirichlet(self.functions["p"], fp64=False)
        if residual is not None:
            vals = self.evaluate_all()
            u = self.space.discrete_spaces[self.space.names[0]]
            shift = self.space.strides[self.space.names[0]] + self.space.dof_shift
            for j in range(vals.shape[1]):
                np.add.at(residual, u.bc_dest[:, j] + shift, vals[:, j])

    def evaluate_all(self):
        if self._cache is None:
            self._cache = _bc_facet_values(self.space, self.functions["p"])
        return self._cache


def _bc_facet_values(space, p_dirichlet):
    """(F, fd) facet residuals in the reference's local (transformed) facet order."""
    import torch

    from . import _native as nat
    from .device import to_dev

    sysd = _darcy_system(space)
    if not sysd.F:
        return np.zeros((0, 0))
    fd = sysd.bc_rows.shape[1]
    u = space.discrete_spaces[space.names[0]]
    # local codes: facet f, reference facet dof a -> f*fd + j(a), with the orientation sign
    bcd = u.bc_dest
    ref = u.bc_ref_dest
    j_of_a = np.argmax(ref[:, :, None] == bcd[:, None, :], axis=2)
    codes = (np.arange(sysd.F)[:, None] * fd + j_of_a).astype(np.int64)
    codes |= np.where(u.bc_ref_sign < 0, nat.SLOT_NEG, 0)
    d_codes = to_dev(codes.astype(np.int32))
    out = torch.zeros(sysd.F * fd, dtype=torch.float64, device="cuda")
    sysd.assemble_dirichlet(p_dirichlet, fp64=True, rows=d_codes, out=out)
    return out.cpu().numpy().reshape(sysd.F, fd)


def _elasticity_system(space, functions):
    from .elasticity import ElasticitySystem, ElasticitySystem2D

    sysd = getattr(space, "_nfx_elasticity_system", None)
    if sysd is None or sysd.functions is not functions:
        two_d = space.discrete_spaces[space.names[0]].dimension == 2
        sysd = (ElasticitySystem2D if two_d else ElasticitySystem)(space, functions)
        space._nfx_elasticity_system = sysd
    return sysd


class LEDualWeakForm(WeakForm):
    """Dual-mixed elasticity element: ``functions = {"rhs": f (6 comps), "lambda", "mu"}``.

    ``evaluate_form_vectorized`` is the reference's live path
    (``le_dual_weak_form.py:354-510``); ``evaluate_form`` is the same operator here
    (the reference's AD variant is not usable in 3-D)."""

    _mode = 0

    def _local(self):
        if self.space is None or self.functions is None:
            raise ValueError
        if self._cache is None:
            import numpy as _np

            from .elasticity import NL

            sysd = _elasticity_system(self.space, self.functions)
            if sysd.dim == 2:
                # the 2-D system keeps its FP64 element matrices from the host cold pass
                self._cache = sysd.local_matrices(self._mode)
                return self._cache
            d_loc = _dense_local_slots(sysd)
            K = sysd._local(self._mode, d_loc)
            r = sysd._local_rhs.cpu().numpy().reshape(sysd.E, NL)
            self._cache = (K, r)
        return self._cache

    def evaluate_form_vectorized(self, element_index, alpha=None):
        K, r = self._local()
        return r[element_index].copy(), K[element_index].copy()

    evaluate_form = evaluate_form_vectorized

    def scatter(self, jacobian, residual):
        sysd = _elasticity_system(self.space, self.functions)
        which = "A" if self._mode == 0 else "P"
        sysd.assemble(which, fp64=True)
        sysd.assemble(which, fp64=False)
        if jacobian is not None:
            jacobian._attach_elasticity(sysd, which)
        if residual is not None and self._mode == 0:
            residual += sysd.rhs64.cpu().numpy()
        return sysd


class LERieszMapWeakForm(LEDualWeakForm):
    """Riesz-map preconditioning operator (``le_riesz_map_weak_form.py:258-360``)."""

    _mode = 1


class LEDualWeakFormBCDirichlet(WeakForm):
    """Stress Dirichlet term: ``functions = {"u": u_D, ...}`` (``le_dual_weak_form.py:514-578``).

    Needs the operator weak form's coefficients: pass ``operator_functions`` (the
    ``functions`` of the ``LEDualWeakForm`` on the same space)."""

    def __init__(self, space, operator_functions=None):
        super().__init__(space)
        self.operator_functions = operator_functions

    def _system(self):
        sysd = getattr(self.space, "_nfx_elasticity_system", None)
        if sysd is None:
            if self.operator_functions is None:
                raise ValueError("scatter the LEDualWeakForm first or pass operator_functions")
            sysd = _elasticity_system(self.space, self.operator_functions)
        return sysd

    def evaluate_all(self):
        if self._cache is None:
            import torch

            from . import _native as nat
            from .device import to_dev

            sysd = self._system()
            s = self.space.discrete_spaces[self.space.names[0]]
            fd = s.bc_dest.shape[1]
            ref = s.bc_ref_dest
            j_of_a = np.argmax(ref[:, :, None] == s.bc_dest[:, None, :], axis=2)
            codes = (np.arange(sysd.F)[:, None] * fd + j_of_a).astype(np.int64)
            codes |= np.where(s.bc_ref_sign < 0, nat.SLOT_NEG, 0)
            out = torch.zeros(sysd.F * fd, dtype=torch.float64, device="cuda")
            sysd.assemble_dirichlet(self.functions["u"], fp64=True, rows=to_dev(codes.astype(np.int32)), out=out)
            self._cache = out.cpu().numpy().reshape(sysd.F, fd)
        return self._cache

    def evaluate_form(self, element_index, alpha=None):
        r = self.evaluate_all()[element_index]
        return r.copy(), np.zeros((len(r), len(r)))

    def scatter(self, jacobian, residual):
        sysd = self._system()
        if not sysd.F:
            return
        sysd.assemble_dirichlet(self.functions["u"], fp64=True)
        sysd.assemble_dirichlet(self.functions["u"], fp64=False)
        if residual is not None:
            vals = self.evaluate_all()
            s = self.space.discrete_spaces[self.space.names[0]]
            shift = self.space.strides[self.space.names[0]] + self.space.dof_shift
            for j in range(vals.shape[1]):
                np.add.at(residual, s.bc_dest[:, j] + shift, vals[:, j])


def _dense_local_slots(sysd):
    """Slot codes writing each element matrix densely in the reference's local order."""
    from .device import to_dev
    from .elasticity import NL

    if not hasattr(sysd, "_dense_local"):
        loc = np.argsort(sysd.s_sp.local_src, axis=1)  # reference a -> local i (stress)
        nb = loc.shape[1]
        li = np.concatenate([(loc[:, :, None] * 3 + np.arange(3)).reshape(sysd.E, nb * 3),
                             np.tile(np.arange(36, 42), (sysd.E, 1))], axis=1)
        from .pattern import local_slot_map

        # the elasticity kernel reads its slot map element-major
        sysd._dense_local = to_dev(np.ascontiguousarray(local_slot_map(li, sysd.ref_sign, NL).T))
    return sysd._dense_local
