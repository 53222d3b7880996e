"""Batched MINRES over uneven Darcy instances (nfx_minres_batch).

Every instance of a mixed RT1-P0 / RT2-P1 batch reaches an FP64 relative residual of
1e-8 against the oracle-assembled system (scipy's sparse LU of that system agrees to
1e-6); an instance's result does not depend on the other instances of the batch
(bitwise equal when it is solved alone); the per-instance iteration counts stay within a
check interval of the single-system MINRES's or below (same recurrence, other reduction
order, refinement steps that stop at the reduction still missing).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SPECS = [(3, 1), (2, 2), (4, 1), (2, 1), (3, 2)]


def test_batched_minres_against_oracle_and_alone():
    import scipy.sparse.linalg as spla
    import torch
    from nfempy_amd.batch import DarcyBatch
    from oracle.assemble import assemble_darcy
    from tests._cases import oracle_darcy_space

    print("[batch solve] build", flush=True)
    b = DarcyBatch(SPECS, seed=21)
    b.assemble()
    torch.cuda.synchronize()
    xs, infos, dt = b.solve_batched()
    print(f"[batch solve] batched: {[i['iterations'] for i in infos]} it, {dt * 1e3:.1f} ms", flush=True)
    for i, (n, k) in enumerate(b.specs):
        info = infos[i]
        assert info["converged"] and info["residual"] <= 1e-8, (i, info)
        inst = b.instances[i]
        ops = oracle_darcy_space(3, n, k)
        kappa = np.exp(inst["logk"].cpu().numpy())
        A, rg = assemble_darcy(ops, {"rhs": np.ones(len(kappa)), "kappa": kappa})
        x = xs[i].cpu().numpy()
        res = np.linalg.norm(-rg - A @ x) / np.linalg.norm(rg)
        xr = spla.spsolve(A.tocsc(), -rg)
        err = np.linalg.norm(x - xr) / np.linalg.norm(xr)
        print(f"[batch solve] instance {i} (n={n}, k={k}): residual vs oracle {res:.1e}, vs LU {err:.1e}", flush=True)
        assert res <= 1.2e-8 and err <= 1e-6, (i, res, err)
        # the single-system MINRES: same recurrence; the batched refinement steps stop at the
        # reduction still missing, so never more iterations than it (within a check interval)
        _, single = b.solve(i)
        assert info["iterations"] <= single["iterations"] + 32, (i, single["iterations"], info["iterations"])
    # instance 2 alone: bitwise the same result
    xa, ia, _ = b.solve_batched([2])
    assert ia[0]["iterations"] == infos[2]["iterations"]
    assert torch.equal(xa[0], xs[2])
