"""Batched persistent assembly of uneven Darcy instances (RT1-P0 and RT2-P1 mixed).

Every instance's CSR values and residual from the single batched launch are
bitwise identical to the single-system row kernel with the same coefficients,
before and after autotuning, and on every replay of the captured graph; one
instance is checked against the FP64 oracle (pattern bit-exact, values 1e-5);
per-instance solves reach an FP64 relative residual of 1e-8.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SPECS = [(2, 1), (3, 2), (2, 2), (3, 1), (2, 1), (4, 1)]


def _single(inst):
    s = inst["system"]
    v, r = s.assemble_elements(inst["kinv"], inst["f"], fp64=False)
    return v.clone(), r.clone()


def test_batched_matches_single_and_graph():
    import torch
    from nfempy_amd.batch import DarcyBatch

    print("[batch] build", flush=True)
    b = DarcyBatch(SPECS, seed=3)
    b.assemble()
    torch.cuda.synchronize()
    ref = [_single(inst) for inst in b.instances]
    for inst, (v, r) in zip(b.instances, ref):
        assert torch.equal(inst["vals"], v) and torch.equal(inst["rhs"], r)
    trials = b.autotune(reps=2)
    # RT2 with per-element coefficients runs the compact launch (degree code 3)
    assert len(trials) == 12 * len(b.groups) and sorted(b.groups) == [1, 3]
    for inst in b.instances:
        inst["vals"].zero_()
    b.assemble()
    torch.cuda.synchronize()
    for inst, (v, r) in zip(b.instances, ref):
        assert torch.equal(inst["vals"], v) and torch.equal(inst["rhs"], r)
    print("[batch] graph", flush=True)
    b.capture()
    for rep in range(3):
        for inst in b.instances:
            inst["vals"].fill_(123.0)
            inst["rhs"].fill_(-7.0)
        torch.cuda.synchronize()
        b.assemble()  # graph replay
        torch.cuda.synchronize()
        for inst, (v, r) in zip(b.instances, ref):
            assert torch.equal(inst["vals"], v) and torch.equal(inst["rhs"], r), rep
        _single(b.instances[0])  # eager work between replays
        torch.cuda.synchronize()


def test_batched_one_launch_over_both_degrees():
    """degree = 0 (one launch, generic kernel over RT1 and RT2 tiles) gives the same bits
    as the degree-specialised launches."""
    import torch
    from nfempy_amd.batch import DarcyBatch

    b1 = DarcyBatch(SPECS, seed=4)
    b0 = DarcyBatch(SPECS, seed=4, split=False)
    assert sorted(b0.groups) == [0]
    b1.assemble()
    b0.assemble()
    torch.cuda.synchronize()
    for i1, i0 in zip(b1.instances, b0.instances):
        assert torch.equal(i1["vals"], i0["vals"]) and torch.equal(i1["rhs"], i0["rhs"])


def test_batched_instance_vs_oracle_and_solve():
    import torch
    from nfempy_amd.batch import DarcyBatch
    from oracle.assemble import assemble_darcy
    from tests._cases import oracle_darcy_space

    b = DarcyBatch([(3, 1), (2, 2)], seed=9)
    b.assemble()
    torch.cuda.synchronize()
    for i, (n, k) in enumerate(b.specs):
        inst = b.instances[i]
        ops = oracle_darcy_space(3, n, k)
        kappa = np.exp(inst["logk"].cpu().numpy())
        E = len(kappa)
        A, rg = assemble_darcy(ops, {"rhs": np.ones(E), "kappa": kappa})
        s = inst["system"]
        assert np.array_equal(s.pattern.rowptr, A.indptr) and np.array_equal(s.pattern.col, A.indices)
        got = inst["vals"].double().cpu().numpy()
        rowmax = np.repeat(abs(A).max(axis=1).toarray().ravel(), np.diff(A.indptr))
        assert np.max(np.abs(got - A.data) / rowmax) <= 1e-5
        x, info = b.solve(i)
        assert info["converged"]
        xs = x.cpu().numpy()
        assert np.linalg.norm(-rg - A @ xs) / np.linalg.norm(rg) <= 1.2e-8
