"""Assembler and PETSc-like operator facades (drop-in for the hot path).

- ``SequentialAssembler(space, jacobian, residual)`` with ``form_to_input_list``
  and ``scatter_forms(measure_time_q)`` mirrors
  ``src/assembly/SequentialAssembler.py:13-86``. Entries are
  ``{name: [form_type, sequence, weak_form, alpha(, ...)]}`` (time-dependent types:
  ``[form_type, sequence, weak_form, (alpha_n, alpha), t]``). Weak forms of this package
  with a batched GPU path ("form" / "bc_form" over every element / boundary facet, how the
  reference drivers use them) run it; every other form and every other type - "form",
  "bc_form", "interface_form", "bc_interface_form", "time_dependent_form",
  "time_dependent_bc_form" with the argument conventions of
  ``src/assembly/ScatterFormData.py:28-76`` - is evaluated element by element (the user's
  ``evaluate_form``), and its contributions are summed into CSR on the device by a sorted
  segmented reduction in insertion order (the reference's ``setValue(addv=True)`` order).
- ``Mat`` mimics the PETSc ``Mat`` calls the reference makes on its Jacobian
  (``createAIJ``, ``setType("sbaij")``, ``setValue(addv=True)``, ``assemble``,
  ``getValuesCSR``, ``getInfo``, ``createVecLeft/Right``, ``scale``): the CSR lives
  in HBM (int32 rowptr/col, FP32 + FP64 values). ``setValue`` on a matrix that a
  weak form has not filled collects host triplets (slow path, for small
  hand-built systems).
- ``Vec`` holds a host array as ``.array`` (what the drivers read/write).
"""

from __future__ import annotations

import hashlib
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
        self._gen = []  # triplet blocks of generically scattered forms
        self._merged = None  # device CSR of every contribution (when there are generic ones)
        self._merge_cache = None

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
        if self._gen:
            self._assemble_generic()
            return self
        if self._sys is None and self._trip[0]:
            from .pattern import build_pattern  # noqa: F401  (same sort rule)
            import scipy.sparse as sps

            r, c, v = (np.asarray(a) for a in self._trip)
            self._host = sps.coo_matrix((v, (r, c)), shape=(self.n, self.n)).tocsr()
        return self

    def _assemble_generic(self):
        """Every contribution - the device system's CSR values first (when a weak form with a
        batched path filled this matrix), then the generic blocks in scatter order - summed
        into one CSR on the device (``DeviceCooSum``)."""
        from .device import torch

        t = torch()
        if self.type == "sbaij" and self._sys is not None:
            raise NotImplementedError("generic forms into an SBAIJ matrix that holds a device elasticity operator")
        blocks = []
        dev_vals = None
        if self._sys is not None:
            rp, col, v32, v64 = self.device_csr_system()
            r = np.repeat(np.arange(len(rp) - 1, dtype=np.int64), np.diff(rp.cpu().numpy()))
            blocks.append((r, col.cpu().numpy().astype(np.int64)))
            dev_vals = v64
        gr = np.concatenate([b[0] for b in self._gen])
        gc = np.concatenate([b[1] for b in self._gen])
        gv = np.concatenate([b[2] for b in self._gen])
        if self.type == "sbaij":
            # SBAIJ stores the upper triangle; lower entries are ignored
            up = gr <= gc
            gr, gc, gv = gr[up], gc[up], gv[up]
        blocks.append((gr, gc))
        rows = np.concatenate([b[0] for b in blocks])
        cols = np.concatenate([b[1] for b in blocks])
        h = hashlib.blake2b(rows.tobytes(), digest_size=16)
        h.update(cols.tobytes())
        key = (self.n, len(rows), h.digest())
        if self._merge_cache is None or self._merge_cache[0] != key:
            self._merge_cache = (key, DeviceCooSum(self.n, rows, cols))
        summer = self._merge_cache[1]
        vals = t.from_numpy(gv).to("cuda")
        if dev_vals is not None:
            vals = t.cat([dev_vals.to(t.float64), vals])
        v64 = summer.sum(vals)
        self._merged = (summer.rowptr, summer.col, v64.to(t.float32), v64)

    def device_csr_system(self):
        s = self._sys
        if getattr(self, "_which", None):
            return s.full(self._which)
        return s.rowptr, s.col, s.vals32, s.vals64

    # -- device path -------------------------------------------------------------
    def _attach(self, system):
        if self.n and system.n != self.n:
            raise ValueError(f"matrix size {self.n} != system size {system.n}")
        self.n = system.n
        self._sys = system

    def _add_triplets(self, rows, cols, vals):
        """Element contributions of a generically scattered form (host FP64, in order). They
        stay until ``zeroEntries`` (PETSc's add semantics: every ``assemble`` sums all of
        them with the current values of an attached device system)."""
        self._gen.append((np.asarray(rows, np.int64), np.asarray(cols, np.int64), np.asarray(vals, np.float64)))

    def _attach_elasticity(self, system, which):
        self._attach(system)
        self._which = which

    @property
    def device_csr(self):
        """(rowptr, col, vals32, vals64) torch tensors in HBM."""
        if self._merged is not None:
            return self._merged
        s = self._sys
        if s is None:
            raise RuntimeError("matrix has no device values (no weak form scattered into it)")
        return self.device_csr_system()

    def getValuesCSR(self):
        if self._merged is not None:
            rp, col, _, v64 = self._merged
            return rp.cpu().numpy().astype(np.int64), col.cpu().numpy().astype(np.int64), v64.cpu().numpy()
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
        if self._merged is not None:
            nnz = int(self._merged[1].numel())
        elif self._sys is not None and getattr(self, "_which", None):
            nnz = self._sys.ops[self._which]["pattern"].nnz
        else:
            nnz = self._sys.nnz if self._sys is not None else (self._host.nnz if self._host is not None else 0)
        return {"nz_allocated": nnz, "nz_used": nnz, "nz_unneeded": 0}

    def createVecLeft(self):
        return Vec(self.n)

    def createVecRight(self):
        return Vec(self.n)

    def scale(self, a):
        self._gen = [(r, c, v * a) for r, c, v in self._gen]
        if self._merged is not None:
            self._merged[2].mul_(a)
            self._merged[3].mul_(a)
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
        """Values to zero (the pattern stays); contributions added afterwards start over."""
        self.scale(0.0)
        self._gen = []
        if self._merged is not None:
            self._merged = None


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

    FORM_TYPES = ("form", "bc_form", "interface_form", "bc_interface_form", "time_dependent_form",
                  "time_dependent_bc_form")

    def scatter_forms(self, measure_time_q=False):
        for name, entry in self._form_to_input_list.items():
            form_type, sequence, weak_form = entry[0], entry[1], entry[2]
            st = time.time()
            if form_type not in self.FORM_TYPES:
                raise ValueError(f"unknown form type {form_type!r}")
            batched = form_type in ("form", "bc_form") and hasattr(weak_form, "scatter")
            if batched:
                sp = weak_form.space.discrete_spaces[weak_form.space.names[0]]
                expected = len(sp.elements) if form_type == "form" else len(sp.bc_elements)
                batched = len(sequence) == expected
            if batched:
                weak_form.scatter(self.jacobian, self.residual)
            else:
                self._scatter_generic(entry)
            if measure_time_q:
                print(f"Assembly time for {name}: {time.time() - st} seconds")

    def _scatter_generic(self, entry):
        """The element loop of ``src/assembly/ScatterFormData.py:28-76`` for one entry: local
        destinations and values per sequence item with the reference's argument conventions;
        the residual is summed in order on the host (``res_g[d] += r``, repeated indices
        included), the Jacobian contributions go to the matrix's device sum."""
        form_type, sequence, form = entry[0], entry[1], entry[2]
        space = form.space
        timed = form_type.startswith("time_dependent")
        if timed:
            (alpha_n, alpha), t_val = entry[3], entry[4]
        else:
            alpha = entry[3] if len(entry) > 3 else None
        rows, cols, vals = [], [], []
        for item in sequence:
            if form_type == "form":
                d = space.destination_indexes(item)
                r, j = form.evaluate_form(item, None if alpha is None else alpha[d])
            elif form_type == "bc_form":
                d = space.bc_destination_indexes(item)
                r, j = form.evaluate_form(item, None if alpha is None else alpha[d])
            elif form_type == "time_dependent_form":
                d = space.destination_indexes(item)
                r, j = form.evaluate_form(item, alpha_n[d], alpha[d], t_val)
            elif form_type == "time_dependent_bc_form":
                d = space.bc_destination_indexes(item)
                r, j = form.evaluate_form(item, alpha_n[d], t_val)
            elif form_type == "interface_form":
                idx, (idx_p, idx_n) = item
                dp, dn = space.destination_indexes(idx_p), space.destination_indexes(idx_n)
                d = np.concatenate((dp, dn))
                r, j = form.evaluate_form(idx, (idx_p, idx_n), (alpha[dp], alpha[dn]))
            else:  # bc_interface_form
                idx, idx_p = item
                d = space.destination_indexes(idx_p)
                r, j = form.evaluate_form(idx, idx_p, alpha[d])
            d = np.asarray(d, np.int64)
            if self.residual is not None:
                np.add.at(self.residual, d, np.asarray(r, np.float64).ravel())
            if self.jacobian is not None:
                rows.append(np.repeat(d, len(d)))
                cols.append(np.tile(d, len(d)))
                vals.append(np.asarray(j, np.float64).ravel())
        if self.jacobian is not None and rows:
            self.jacobian._add_triplets(np.concatenate(rows), np.concatenate(cols), np.concatenate(vals))


def coo_pattern(n, rows, cols):
    """CSR pattern of COO indices and the segmented-sum plan: ``perm`` orders the
    contributions by (row, column), ties in insertion order (``np.lexsort`` is stable);
    CSR entry m sums ``perm[segptr[m]:segptr[m + 1]]`` left to right. Returns int64
    (rowptr (n+1,), col (nnz,), perm (n_in,), segptr (nnz+1,))."""
    rows = np.asarray(rows, np.int64).ravel()
    cols = np.asarray(cols, np.int64).ravel()
    if rows.shape != cols.shape:
        raise ValueError("rows and cols differ in length")
    if len(rows) and (rows.min() < 0 or cols.min() < 0 or rows.max() >= n or cols.max() >= n):
        raise ValueError("COO index out of range")
    order = np.lexsort((cols, rows))
    key = rows[order] * n + cols[order]
    uniq, start = np.unique(key, return_index=True)
    counts = np.bincount(uniq // n, minlength=n) if n else np.zeros(0, np.int64)
    rowptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    return rowptr, (uniq % n).astype(np.int64), order.astype(np.int64), np.append(start, len(key)).astype(np.int64)


class DeviceCooSum:
    """Sum of COO contributions into CSR on the device, reproducible: entries are sorted by
    (row, column) with ties kept in insertion order (``np.lexsort`` is stable), and every
    CSR value is the sum of its contributions in that order (``nfx_stage_reduce``), the
    order of the reference's ``setValue(addv=True)`` loop. The pattern is built once per
    sequence of indices; ``sum`` runs on the device for every new set of values."""

    def __init__(self, n, rows, cols):
        from . import _native as nat
        from .device import to_dev

        rowptr, col, perm, segptr = coo_pattern(n, rows, cols)
        self.n, self.nnz, self.n_in = n, len(col), len(perm)
        if self.n_in >= 1 << 31 or self.nnz >= 1 << 31:
            raise ValueError("too many contributions for int32 indices")
        self.rowptr = to_dev(rowptr.astype(np.int32))
        self.col = to_dev(col.astype(np.int32))
        self.perm = to_dev(perm.astype(np.int32))
        self.segptr = to_dev(segptr.astype(np.int32))
        self.target = to_dev(np.arange(self.nnz, dtype=np.int32))
        self._nat = nat

    def sum(self, vals):
        """FP64 CSR values from the (n_in,) FP64 device contributions in insertion order."""
        from .device import torch

        t = torch()
        nat = self._nat
        vals = vals.to(dtype=t.float64).contiguous()
        if int(vals.numel()) != self.n_in:
            raise ValueError("contribution count differs from the pattern")
        stage = t.empty(max(self.n_in, 1), dtype=t.float64, device=vals.device)
        out = t.empty(max(self.nnz, 1), dtype=t.float64, device=vals.device)
        s = nat.stream_handle()
        if self.n_in:
            nat.call("nfx_gather", 1, self.n_in, nat.ptr(self.perm), nat.ptr(vals), nat.ptr(stage), s)
            nat.call("nfx_stage_reduce", self.nnz, nat.ptr(self.segptr), nat.ptr(self.target), 1, nat.ptr(stage),
                     nat.ptr(out), s)
        self._held = (vals, stage)
        return out[: self.nnz]
