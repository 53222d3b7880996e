"""Generic form types of the drop-in ``SequentialAssembler``: the six scatter rules of
``src/assembly/ScatterFormData.py:28-76`` ("form", "bc_form", "interface_form",
"bc_interface_form", "time_dependent_form", "time_dependent_bc_form") with forms that have
no batched path of their own.

Expected values come from the reference's loop restated on dictionaries: for every item,
``res_g[d_k] += r_k`` one index at a time and ``jac_g[row_k, col_k] += v_k`` for
``row = repeat(d, len(d))``, ``col = tile(d, len(d))`` - the scatter the reference performs
with PETSc's ``setValue(addv=True)``. Both sums run in insertion order, so the device
segmented sum must equal them bit for bit.

The CPU tests pin the segmented-sum plan (``coo_pattern``) and the argument conventions
(which alpha each type receives: ``alpha_n`` for the time-dependent boundary form, the pair
for interfaces); the GPU tests run the device sum, the merge with a batched Darcy scatter
(against the oracle) and ``zeroEntries`` / ``scale``.
"""

import zlib

import numpy as np
import pytest

from nfempy_amd.assembly import SequentialAssembler, coo_pattern
from tests._cases import product_darcy_space


class _Rec:
    """A form of the reference's shape: deterministic local values from the item and the
    alpha it receives (so a wrong alpha changes the result), recording every call."""

    def __init__(self, space, kind):
        self.space, self.kind, self.calls = space, kind, []

    def _vals(self, key, n, *alphas):
        rng = np.random.default_rng(zlib.crc32(repr(key).encode()))
        r = rng.standard_normal(n)
        j = rng.standard_normal((n, n))
        for k, a in enumerate(alphas):
            r = r + (k + 1) * np.asarray(a, float).ravel()[:n]
        return r, j

    def evaluate_form(self, *args):
        self.calls.append(args)
        k = self.kind
        if k in ("form", "bc_form"):
            idx, a = args
            return self._vals((k, idx), len(a), a)
        if k == "time_dependent_form":
            idx, an, a, t = args
            r, j = self._vals((k, idx), len(a), an, a)
            return r * t, j
        if k == "time_dependent_bc_form":
            idx, an, t = args
            r, j = self._vals((k, idx), len(an), an)
            return r * t, j
        if k == "interface_form":
            idx, (p, q), (ap, aq) = args
            return self._vals((k, idx, p, q), len(ap) + len(aq), np.concatenate((ap, aq)))
        idx, p, a = args  # bc_interface_form
        return self._vals((k, idx, p), len(a), a)


def _items(kind, ne, nbc, rng):
    if kind in ("form", "time_dependent_form"):
        return [int(i) for i in rng.permutation(ne)[: max(2, ne // 2)]]
    if kind in ("bc_form", "time_dependent_bc_form"):
        return [int(i) for i in rng.permutation(nbc)[: max(2, nbc // 2)]]
    if kind == "interface_form":
        return [(k, (int(a), int(b))) for k, (a, b) in enumerate(rng.integers(0, ne, (6, 2)))]
    return [(k, int(a)) for k, a in enumerate(rng.integers(0, ne, 5))]


def _expected(space, entries, n):
    """The reference's scatter restated on dicts (``ScatterFormData.py:5-76``)."""
    jac, res = {}, np.zeros(n)
    for kind, seq, form, *rest in entries:
        for item in seq:
            if kind == "form":
                d = space.destination_indexes(item)
                r, j = form._vals((kind, item), len(d), rest[0][d])
            elif kind == "bc_form":
                d = space.bc_destination_indexes(item)
                r, j = form._vals((kind, item), len(d), rest[0][d])
            elif kind == "time_dependent_form":
                (an, a), t = rest
                d = space.destination_indexes(item)
                r, j = form._vals((kind, item), len(d), an[d], a[d])
                r = r * t
            elif kind == "time_dependent_bc_form":
                (an, a), t = rest
                d = space.bc_destination_indexes(item)
                r, j = form._vals((kind, item), len(d), an[d])
                r = r * t
            elif kind == "interface_form":
                idx, (p, q) = item
                d = np.concatenate((space.destination_indexes(p), space.destination_indexes(q)))
                r, j = form._vals((kind, idx, p, q), len(d), rest[0][d])
            else:
                idx, p = item
                d = space.destination_indexes(p)
                r, j = form._vals((kind, idx, p), len(d), rest[0][d])
            for k, dk in enumerate(d):
                res[dk] += r[k]
            row, col, data = np.repeat(d, len(d)), np.tile(d, len(d)), j.ravel()
            for k in range(len(data)):
                jac[(int(row[k]), int(col[k]))] = jac.get((int(row[k]), int(col[k])), 0.0) + data[k]
    return jac, res


def _entries(space, rng):
    n = space.n_dof
    ne = len(space.discrete_spaces["u"].elements)
    nbc = len(space.discrete_spaces["u"].bc_elements)
    a = rng.standard_normal(n)
    an = rng.standard_normal(n)
    out = []
    for kind in SequentialAssembler.FORM_TYPES:
        f = _Rec(space, kind)
        seq = _items(kind, ne, nbc, rng)
        if kind.startswith("time_dependent"):
            out.append([kind, seq, f, (an, a), 0.75])
        else:
            out.append([kind, seq, f, a])
    return out


class _HostJac:
    """A Jacobian that records the generic triplets (the CPU side of ``Mat``)."""

    def __init__(self):
        self.blocks = []

    def _add_triplets(self, r, c, v):
        self.blocks.append((r, c, v))


def test_generic_scatter_plan_matches_reference_loop():
    space = product_darcy_space(2, 2, 1, None)
    n = space.n_dof
    entries = _entries(space, np.random.default_rng(5))
    jac = _HostJac()
    res = np.zeros(n)
    asm = SequentialAssembler(space, jac, res)
    asm.form_to_input_list = {e[0]: e for e in entries}
    asm.scatter_forms()
    jref, rref = _expected(space, entries, n)
    assert np.array_equal(res, rref)
    rows = np.concatenate([b[0] for b in jac.blocks])
    cols = np.concatenate([b[1] for b in jac.blocks])
    vals = np.concatenate([b[2] for b in jac.blocks])
    rowptr, col, perm, segptr = coo_pattern(n, rows, cols)
    got = {}
    for i in range(n):
        for m in range(rowptr[i], rowptr[i + 1]):
            s = vals[perm[segptr[m]]]
            for q in perm[segptr[m] + 1: segptr[m + 1]]:
                s += vals[q]
            got[(i, int(col[m]))] = s
    assert got == jref
    # the argument conventions: the time-dependent boundary form receives alpha_n
    tbc = next(e for e in entries if e[0] == "time_dependent_bc_form")
    item = tbc[1][0]
    call = tbc[2].calls[0]
    assert np.array_equal(call[1], tbc[3][0][space.bc_destination_indexes(item)]) and call[2] == 0.75


def test_coo_pattern_edges():
    rp, col, perm, seg = coo_pattern(4, [], [])
    assert np.array_equal(rp, np.zeros(5)) and len(col) == 0 and len(perm) == 0 and np.array_equal(seg, [0])
    rp, col, perm, seg = coo_pattern(3, [2, 0, 2, 2], [1, 0, 1, 0])
    assert np.array_equal(rp, [0, 1, 1, 3]) and np.array_equal(col, [0, 0, 1])
    assert np.array_equal(perm, [1, 3, 0, 2]) and np.array_equal(seg, [0, 1, 2, 4])
    with pytest.raises(ValueError):
        coo_pattern(3, [3], [0])
    asm = SequentialAssembler(None, None, None)
    asm.form_to_input_list = {"x": ["cell_form", [0], None, None]}
    with pytest.raises(ValueError, match="unknown form type"):
        asm.scatter_forms()


def _csr_dict(indptr, indices, data):
    return {(i, int(indices[m])): data[m] for i in range(len(indptr) - 1) for m in range(indptr[i], indptr[i + 1])}


@pytest.mark.gpu
def test_generic_forms_on_device():
    import torch
    from nfempy_amd.assembly import Mat

    print("[generic forms] six types, device sum", flush=True)
    space = product_darcy_space(2, 3, 1, None)
    n = space.n_dof
    entries = _entries(space, np.random.default_rng(11))
    A = Mat().createAIJ([n, n])
    rg = np.zeros(n)
    asm = SequentialAssembler(space, A, rg)
    asm.form_to_input_list = {e[0]: e for e in entries}
    asm.scatter_forms()
    A.assemble()
    torch.cuda.synchronize()
    jref, rref = _expected(space, entries, n)
    assert np.array_equal(rg, rref)
    got = _csr_dict(*A.getValuesCSR())
    assert got == jref
    assert A.getInfo()["nz_used"] == len(jref)
    # assemble again: the same values (contributions persist until zeroEntries)
    A.assemble()
    assert _csr_dict(*A.getValuesCSR()) == jref
    A.scale(2.0)
    assert _csr_dict(*A.getValuesCSR()) == {k: 2.0 * v for k, v in jref.items()}
    A.zeroEntries()
    asm.scatter_forms()
    A.assemble()
    assert _csr_dict(*A.getValuesCSR()) == jref


@pytest.mark.gpu
def test_generic_form_merged_with_batched_darcy():
    """A batched Darcy scatter and a generic interface form in one Jacobian: the merged CSR
    equals the oracle's Darcy matrix plus the reference loop of the interface form."""
    import torch
    from nfempy_amd.assembly import Mat
    from nfempy_amd.weak_forms import LaplaceDualWeakForm
    from oracle.assemble import assemble_darcy
    from tests._cases import darcy_functions, oracle_darcy_space

    print("[generic forms] merged with the batched Darcy form", flush=True)
    d, nn, k = 2, 3, 1
    fun, ex = darcy_functions(d)
    space = product_darcy_space(d, nn, k, None)
    n = space.n_dof
    ne = len(space.discrete_spaces["u"].elements)
    wf = LaplaceDualWeakForm(space)
    wf.functions = fun
    rng = np.random.default_rng(3)
    a = rng.standard_normal(n)
    itf = ["interface_form", _items("interface_form", ne, 0, rng), _Rec(space, "interface_form"), a]
    A = Mat().createAIJ([n, n])
    rg = np.zeros(n)
    asm = SequentialAssembler(space, A, rg)
    asm.form_to_input_list = {"mixed": ["form", range(ne), wf, np.zeros(n)], "itf": itf}
    asm.scatter_forms()
    A.assemble()
    torch.cuda.synchronize()
    Aref, rgref = assemble_darcy(oracle_darcy_space(d, nn, k, None), fun)
    jitf, ritf = _expected(space, [itf], n)
    exp = _csr_dict(Aref.indptr, Aref.indices, Aref.data)
    for key, v in jitf.items():
        exp[key] = exp.get(key, 0.0) + v
    got = _csr_dict(*A.getValuesCSR())
    assert set(got) == set(exp)
    scale = max(abs(v) for v in exp.values())
    assert max(abs(got[key] - exp[key]) for key in exp) <= 1e-12 * scale
    assert np.abs(rg - (rgref + ritf)).max() <= 1e-12 * np.abs(rgref + ritf).max()
    # the device CSR the solvers read is the merged one
    rp, col, v32, v64 = A.device_csr
    assert int(col.numel()) == len(exp) and v32.dtype == torch.float32
