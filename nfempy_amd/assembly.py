"""Assembler and PETSc-like operator facades (drop-in for the hot path).

- ``SequentialAssembler(space, jacobian, residual)`` with ``form_to_input_list``
  and ``scatter_forms(measure_time_q)`` mirrors
  ``src/assembly/SequentialAssembler.py:13-86``. Entries are
  ``{name: [form_type, sequence, weak_form, alpha(, ...)]}``; form types "form"
  and "bc_form" run through the batched GPU scatter (the sequence must cover
  every element / boundary facet, which is how the reference drivers use it).
- ``Mat`` mimics the PETSc ``Mat`` calls the reference makes on its Jacobian
  (``createAIJ``, ``setType("sbaij")``, ``setValue(addv=True)``, ``assemble``,
  ``getValuesCSR``, ``getInfo``, ``createVecLeft/Right``, ``scale``): the CSR lives
  in HBM (int32 rowptr/col, FP32 + FP64 values). ``setValue`` on a matrix that a
  weak form has not filled collects host triplets (slow path, for small
  hand-built systems).
- ``Vec`` holds a host array as ``.array`` (what the drivers read/write).
"""

from __future__ import annotations

import time

import numpy as np


class Vec:
    def __init__(self, n):
        self.array = np.zeros(n)

    def getSize(self):
        return len(self.array)


class Mat:
    def __init__(self):
        self.n = 0
        self.type = "aij"
        self._sys = None
        self._trip = ([], [], [])
        self._host = None

    # -- creation ---------------------------------------------------------------
    def createAIJ(self, size, comm=None):
        n = size[0] if isinstance(size, (list, tuple)) else int(size)
        self.n = int(n)
        return self

    def setType(self, t):
        self.type = t

    def getSize(self):
        return (self.n, self.n)

    # -- slow host path ---------------------------------------------------------
    def setValue(self, row, col, value, addv=True):
        self._trip[0].append(int(row))
        self._trip[1].append(int(col))
        self._trip[2].append(float(value))

    def assemble(self):
        if self._sys is None and self._trip[0]:
            from .pattern import build_pattern  # noqa: F401  (same sort rule)
            import scipy.sparse as sps

            r, c, v = (np.asarray(a) for a in self._trip)
            self._host = sps.coo_matrix((v, (r, c)), shape=(self.n, self.n)).tocsr()
        return self

    # -- device path -------------------------------------------------------------
    def _attach(self, system):
        if self.n and system.n != self.n:
            raise ValueError(f"matrix size {self.n} != system size {system.n}")
        self.n = system.n
        self._sys = system

    def _attach_elasticity(self, system, which):
        self._attach(system)
        self._which = which

    @property
    def device_csr(self):
        """(rowptr, col, vals32, vals64) torch tensors in HBM."""
        s = self._sys
        if s is None:
            raise RuntimeError("matrix has no device values (no weak form scattered into it)")
        if getattr(self, "_which", None):
            return s.full(self._which)
        return s.rowptr, s.col, s.vals32, s.vals64

    def getValuesCSR(self):
        if self._sys is not None and getattr(self, "_which", None):
            op = self._sys.ops[self._which]
            pat = op["pattern"]
            return pat.rowptr.astype(np.int64), pat.col.astype(np.int64), op["vals64"].cpu().numpy()
        if self._sys is not None:
            s = self._sys
            return (s.pattern.rowptr.astype(np.int64), s.pattern.col.astype(np.int64), s.vals64.cpu().numpy())
        if self._host is not None:
            return self._host.indptr, self._host.indices, self._host.data
        return np.zeros(self.n + 1, np.int64), np.zeros(0, np.int64), np.zeros(0)

    def getInfo(self):
        if self._sys is not None and getattr(self, "_which", None):
            nnz = self._sys.ops[self._which]["pattern"].nnz
        else:
            nnz = self._sys.nnz if self._sys is not None else (self._host.nnz if self._host is not None else 0)
        return {"nz_allocated": nnz, "nz_used": nnz, "nz_unneeded": 0}

    def createVecLeft(self):
        return Vec(self.n)

    def createVecRight(self):
        return Vec(self.n)

    def scale(self, a):
        if self._sys is not None and getattr(self, "_which", None):
            op = self._sys.ops[self._which]
            op["vals32"].mul_(a)
            op["vals64"].mul_(a)
        elif self._sys is not None:
            self._sys.vals32.mul_(a)
            self._sys.vals64.mul_(a)
        elif self._host is not None:
            self._host = self._host * a

    def zeroEntries(self):
        self.scale(0.0)


class SequentialAssembler:
    """``SequentialAssembler(space, jacobian, residual)`` (``SequentialAssembler.py:13``)."""

    def __init__(self, space, jacobian, residual):
        self.space = space
        self.jacobian = jacobian
        self.residual = residual
        self._form_to_input_list = {}

    @property
    def form_to_input_list(self):
        return self._form_to_input_list

    @form_to_input_list.setter
    def form_to_input_list(self, value):
        self._form_to_input_list = value

    def scatter_forms(self, measure_time_q=False):
        for name, entry in self._form_to_input_list.items():
            form_type, sequence, weak_form = entry[0], entry[1], entry[2]
            st = time.time()
            if form_type not in ("form", "bc_form"):
                raise NotImplementedError(f"form type {form_type!r} is not on the accelerated path")
            expected = (len(weak_form.space.discrete_spaces[weak_form.space.names[0]].elements) if form_type == "form"
                        else len(weak_form.space.discrete_spaces[weak_form.space.names[0]].bc_elements))
            if len(sequence) != expected:
                raise NotImplementedError("the batched scatter needs the full element / facet sequence")
            weak_form.scatter(self.jacobian, self.residual)
            if measure_time_q:
                print(f"Assembly time for {name}: {time.time() - st} seconds")
