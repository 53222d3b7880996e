"""CSR pattern and element -> slot maps (the cold half of assembly).

The reference inserts every element contribution with ``Mat.setValue(row, col,
v, addv=True)`` into an unpreallocated PETSc matrix
(``src/assembly/ScatterFormData.py:5-11``,
``examples/md_elliptic_scalar/md_elliptic_scalar.py:250-265``), which makes the
pattern the union of ``dest x dest`` and sums repeated entries in element order.
Here the pattern is built once by a stable sort of the (row, col) keys of all
element entries; each entry gets a slot code:

- a CSR value index when it is the only contribution to its (row, col);
- a stage index (bit 29) when several elements contribute; the stage array is
  then summed per slot in element order by ``nfx_stage_reduce`` (deterministic,
  no atomics);
- the orientation sign of the reference's ``permute_and_transform`` as bit 30.

The SBAIJ variant of the elasticity driver
(``examples/cauchy_elasticity/mfe_elasticity_example_1.py:128-154``) keeps
``row <= col`` entries passing a value filter and adds explicit zero diagonals.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ._native import SLOT_INDEX_MASK, SLOT_NEG, SLOT_STAGE, SLOT_ZERO


@dataclass
class Pattern:
    n: int
    rowptr: np.ndarray  # int32 (n+1,)
    col: np.ndarray  # int32 (nnz,)
    slot: np.ndarray  # int32 (nl*nl, E) entry-major codes
    segptr: np.ndarray  # int32 (n_seg+1,)
    target: np.ndarray  # int32 (n_seg,)
    n_stage: int

    @property
    def nnz(self):
        return int(self.rowptr[-1])


def build_pattern(n, ref_dest, ref_sign, keep=None, zero_diag_dest=None, upper=False, device=None):
    """Pattern of the union of element (row, col) entries.

    ref_dest, ref_sign: (E, nl) destination rows and signs in reference-basis order.
    keep: optional (E, nl, nl) bool mask of entries inserted (value filter).
    upper: keep only row <= col (SBAIJ).
    zero_diag_dest: if given, (E, nl) rows that receive an explicit 0.0 diagonal
    (these contribute ZERO codes at the (a, a) entries; an entry kept by the mask
    keeps its value).
    device: run the stable sort of the keys there (``numbering.stable_argsort``).
    """
    E, nl = ref_dest.shape
    rd = np.asarray(ref_dest, dtype=np.int64)
    # keys of every element entry (a, b): row ref_dest[e, a], column ref_dest[e, b]
    key = (rd[:, :, None] * n + rd[:, None, :]).reshape(-1)
    plain = keep is None and not upper and zero_diag_dest is None
    if plain:
        flat_sel = None
        zero_flat = None
    else:
        sel = np.ones((E, nl, nl), dtype=bool) if keep is None else keep.copy()
        if upper:
            sel &= rd[:, :, None] <= rd[:, None, :]
        zero = np.zeros((E, nl, nl), dtype=bool)
        if zero_diag_dest is not None:
            diag = np.zeros((E, nl, nl), dtype=bool)
            idx = np.arange(nl)
            diag[:, idx, idx] = True
            zero = diag & ~sel
            sel = sel | diag
        flat_sel = sel.reshape(-1)
        key = key[flat_sel]
        zero_flat = zero.reshape(-1)[flat_sel]
    if device is not None:
        from .numbering import stable_argsort

        order, ks = stable_argsort(key, device)
    else:
        order = np.argsort(key, kind="stable")
        ks = key[order]
    del key
    start = np.r_[True, ks[1:] != ks[:-1]] if len(ks) else np.zeros(0, bool)
    uid_sorted = np.cumsum(start) - 1
    ukeys = ks[start]
    del ks
    nnz = len(ukeys)
    counts = np.bincount(uid_sorted, minlength=nnz)
    seg_first = np.flatnonzero(start)
    del start
    multi = counts > 1
    seg_ids = np.flatnonzero(multi)
    seg_len = counts[seg_ids]
    segptr = np.concatenate([[0], np.cumsum(seg_len)]).astype(np.int64)
    seg_of_uid = np.full(nnz, -1, np.int64)
    seg_of_uid[seg_ids] = np.arange(len(seg_ids))
    n_stage = int(segptr[-1])
    if max(nnz, n_stage) > SLOT_INDEX_MASK:
        raise ValueError("pattern too large for 28-bit slot codes")
    # codes in sorted order (rank = position inside the entry's run, i.e. element order),
    # then one scatter back to entry order
    m_sorted = multi[uid_sorted]
    rank_sorted = np.arange(len(uid_sorted)) - seg_first[uid_sorted]
    code_sorted = np.where(m_sorted, SLOT_STAGE | (segptr[np.maximum(seg_of_uid[uid_sorted], 0)] + rank_sorted),
                           uid_sorted)
    del m_sorted, rank_sorted, uid_sorted
    code = np.empty(len(code_sorted), np.int64)
    code[order] = code_sorted
    del code_sorted, order
    sgn = (ref_sign[:, :, None] * ref_sign[:, None, :]).reshape(-1)
    if flat_sel is not None:
        sgn = sgn[flat_sel]
    code |= np.where(sgn < 0, SLOT_NEG, 0)
    if zero_flat is not None:
        code |= np.where(zero_flat, SLOT_ZERO, 0)
    if flat_sel is None:
        full = code
    else:
        full = np.full(E * nl * nl, -1, np.int64)
        full[flat_sel] = code
    slot = np.ascontiguousarray(full.reshape(E, nl * nl).T).astype(np.int32)
    urow = ukeys // n
    ucol = ukeys % n
    rowptr = np.zeros(n + 1, np.int64)
    np.add.at(rowptr, urow + 1, 1)
    rowptr = np.cumsum(rowptr)
    return Pattern(
        n=n,
        rowptr=rowptr.astype(np.int32),
        col=ucol.astype(np.int32),
        slot=slot,
        segptr=segptr.astype(np.int32),
        target=seg_ids.astype(np.int32),
        n_stage=n_stage,
    )


def vector_slots(n, ref_dest, ref_sign, keep=None):
    """Slot codes of a right-hand side assembled by ``rg[dest] += r_el`` over elements:
    (nl, E) codes (row index when one element contributes to the row, else a stage
    index (bit 29), orientation sign as bit 30; -1 where keep is False), and the
    segments (segptr, target) that ``nfx_stage_reduce`` sums in element order."""
    E, nl = ref_dest.shape
    sel = np.ones((E, nl), dtype=bool) if keep is None else np.asarray(keep, bool)
    r = ref_dest.reshape(-1)[sel.reshape(-1)].astype(np.int64)
    if len(r) and (r.min() < 0 or r.max() >= n):
        raise ValueError("destination outside the vector")
    order = np.argsort(r, kind="stable")
    rs = r[order]
    start = np.r_[True, rs[1:] != rs[:-1]] if len(rs) else np.zeros(0, bool)
    uid_sorted = np.cumsum(start) - 1
    counts = np.bincount(uid_sorted, minlength=int(start.sum()))
    seg_first = np.flatnonzero(start)
    rank_sorted = np.arange(len(rs)) - seg_first[uid_sorted]
    uid = np.empty(len(rs), np.int64)
    uid[order] = uid_sorted
    rank = np.empty(len(rs), np.int64)
    rank[order] = rank_sorted
    multi = counts > 1
    seg_ids = np.flatnonzero(multi)
    segptr = np.concatenate([[0], np.cumsum(counts[seg_ids])]).astype(np.int64)
    seg_of_uid = np.full(len(counts), -1, np.int64)
    seg_of_uid[seg_ids] = np.arange(len(seg_ids))
    urow = rs[start]
    code = np.where(multi[uid], SLOT_STAGE | (segptr[np.maximum(seg_of_uid[uid], 0)] + rank), urow[uid])
    code = code | np.where(ref_sign.reshape(-1)[sel.reshape(-1)] < 0, SLOT_NEG, 0)
    if max(n, int(segptr[-1])) > SLOT_INDEX_MASK:
        raise ValueError("vector too large for 28-bit slot codes")
    full = np.full(E * nl, -1, np.int64)
    full[sel.reshape(-1)] = code
    return (np.ascontiguousarray(full.reshape(E, nl).T).astype(np.int32), segptr.astype(np.int32),
            urow[seg_ids].astype(np.int32))


def row_codes(dest, sign=None):
    """int32 row codes (bit 30 = negate) from destinations and signs."""
    code = np.asarray(dest, dtype=np.int64).copy()
    if sign is not None:
        code |= np.where(np.asarray(sign) < 0, SLOT_NEG, 0)
    return code.astype(np.int32)


def local_slot_map(ref_src_local, ref_sign, nl):
    """Slot codes that write each element's local matrix densely, in the
    reference's element-local (transformed) order: element e entry (i, j) at
    e*nl*nl + i*nl + j. ref_src_local (E, nl): local index of each reference dof.
    """
    E = ref_src_local.shape[0]
    i = ref_src_local[:, :, None]
    j = ref_src_local[:, None, :]
    idx = np.arange(E)[:, None, None] * nl * nl + i * nl + j
    sgn = ref_sign[:, :, None] * ref_sign[:, None, :]
    code = idx | np.where(sgn < 0, SLOT_NEG, 0)
    return np.ascontiguousarray(code.reshape(E, nl * nl).T).astype(np.int32)
