"""Scatter of element contributions into CSR (FP64 oracle).

Restates the drivers' scatter closures:

- Darcy (``examples/md_elliptic_scalar/md_elliptic_scalar.py:250-282``): every one
  of the ``n_loc^2`` entries is inserted with ``addv=True`` (zeros included), so the
  pattern is the union of ``dest x dest``; ``rg[dest] += r_el``.
- Elasticity (``examples/cauchy_elasticity/mfe_elasticity_example_1.py:128-222``):
  SBAIJ keeps ``row <= col``; an entry of an element contribution is inserted only if
  ``not isclose(v, 1e-16)`` (numpy defaults ``rtol=1e-5, atol=1e-8``); every
  ``dest`` diagonal receives an explicit ``0.0`` (operator ``A`` only, not ``P``).
- ``ScatterFormData.__scatter_generic_rhs_data`` (``src/assembly/ScatterFormData.py:14-18``)
  accumulates repeated destination indices one by one.

The sum over contributions to one CSR slot runs in element order (stable sort on
(row, col)), which is the order of the reference's serial ``setValue(addv=True)``.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sps


def csr_from_triplets(n, rows, cols, vals):
    """Stable (row, col) sort + sequential sum -> (indptr, indices, data)."""
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    vals = np.asarray(vals, dtype=np.float64)
    key = rows * n + cols
    order = np.argsort(key, kind="stable")
    key = key[order]
    v = vals[order]
    if len(key) == 0:
        return np.zeros(n + 1, dtype=np.int64), np.zeros(0, dtype=np.int64), np.zeros(0)
    start = np.flatnonzero(np.r_[True, key[1:] != key[:-1]])
    data = np.add.reduceat(v, start)
    ukey = key[start]
    r = ukey // n
    indices = ukey % n
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(indptr, r + 1, 1)
    indptr = np.cumsum(indptr)
    return indptr, indices, data


class TripletCollector:
    """Collects (row, col, value) in scatter order, as ``Mat.setValue(addv=True)``."""

    def __init__(self, n, symmetric_filter=False, explicit_zero_diagonal=False):
        self.n = n
        self.symmetric_filter = symmetric_filter
        self.explicit_zero_diagonal = explicit_zero_diagonal
        self.rows, self.cols, self.vals = [], [], []

    def add_element(self, dest, j_el):
        dest = np.asarray(dest, dtype=np.int64)
        data = np.asarray(j_el, dtype=np.float64).ravel()
        row = np.repeat(dest, len(dest))
        col = np.tile(dest, len(dest))
        if self.symmetric_filter:
            keep = np.logical_not(np.isclose(data, 1.0e-16)) & (row <= col)
            row, col, data = row[keep], col[keep], data[keep]
        self.rows.append(row)
        self.cols.append(col)
        self.vals.append(data)
        if self.explicit_zero_diagonal:
            self.rows.append(dest)
            self.cols.append(dest)
            self.vals.append(np.zeros(len(dest)))

    def csr(self):
        if not self.rows:
            return sps.csr_matrix((self.n, self.n))
        indptr, indices, data = csr_from_triplets(
            self.n, np.concatenate(self.rows), np.concatenate(self.cols), np.concatenate(self.vals)
        )
        return sps.csr_matrix((data, indices, indptr), shape=(self.n, self.n))


def scatter_rhs(rg, dest, r_el):
    for k, d in enumerate(np.asarray(dest).tolist()):
        rg[d] += r_el[k]


def coo_plan(n, rows, cols):
    """The summation plan of ``csr_from_triplets`` without values: (indptr, indices, perm,
    segptr) such that ``data[m] = sum(vals[perm[segptr[m]:segptr[m + 1]]])`` left to right
    (same stable (row, col) order)."""
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    key = rows * n + cols
    order = np.argsort(key, kind="stable")
    key = key[order]
    if len(key) == 0:
        return np.zeros(n + 1, dtype=np.int64), np.zeros(0, dtype=np.int64), order, np.zeros(1, dtype=np.int64)
    start = np.flatnonzero(np.r_[True, key[1:] != key[:-1]])
    ukey = key[start]
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(indptr, ukey // n + 1, 1)
    return np.cumsum(indptr), ukey % n, order.astype(np.int64), np.r_[start, len(key)].astype(np.int64)


def assemble_darcy(ps, functions, bc_functions=None, element_range=None):
    """Darcy system ``(A, rg)`` exactly as ``md_elliptic_scalar.py:250-343``."""
    from .forms import darcy_bc_facet, darcy_element

    n = ps.n_dof
    rg = np.zeros(n)
    coll = TripletCollector(n)
    ne = len(ps.spaces["u"].element_ids)
    rng = range(ne) if element_range is None else element_range
    for i in rng:
        dest = ps.destination_indexes(i)
        r, J = darcy_element(ps, i, functions)
        rg[dest] += r
        coll.add_element(dest, J)
    if bc_functions is not None and element_range is None:
        for j in range(len(ps.spaces["u"].bc_element_ids)):
            dest = ps.bc_destination_indexes(j)
            r, J = darcy_bc_facet(ps, j, bc_functions)
            rg[dest] += r
            coll.add_element(dest, J)
    return coll.csr(), rg


def assemble_elasticity(ps, functions, bc_functions=None):
    """``(A, P, rg)`` as ``mfe_elasticity_example_1.py:128-222`` (SBAIJ upper parts)."""
    from .forms import elasticity_bc_facet, elasticity_element, riesz_element

    n = ps.n_dof
    rg = np.zeros(n)
    ca = TripletCollector(n, symmetric_filter=True, explicit_zero_diagonal=True)
    cp = TripletCollector(n, symmetric_filter=True)
    ne = len(ps.spaces["s"].element_ids)
    for i in range(ne):
        dest = ps.destination_indexes(i)
        r, J = elasticity_element(ps, i, functions)
        rg[dest] += r
        ca.add_element(dest, J)
    if bc_functions is not None:
        for j in range(len(ps.spaces["s"].bc_element_ids)):
            dest = ps.bc_destination_indexes(j)
            r, _ = elasticity_bc_facet(ps, j, bc_functions)
            rg[dest] += r
    for i in range(ne):
        dest = ps.destination_indexes(i)
        _, J = riesz_element(ps, i, functions)
        cp.add_element(dest, J)
    return ca.csr(), cp.csr(), rg


def symmetric_from_upper(U):
    """Full symmetric matrix from an SBAIJ-style upper triangle."""
    U = sps.csr_matrix(U)
    L = sps.triu(U, k=1).T
    return sps.csr_matrix(U + L)
