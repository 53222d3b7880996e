"""GPU parity of the timed warm-assembly launch against the FP64 oracle.

The benchmark's step is ``DarcySystem.assemble(kinv, f, p_D)`` with every coefficient a
device array: a per-element inverse permeability from a log-normal Gaussian random field
(squared-exponential, correlation length 0.05, sigma 2), a constant source and Dirichlet
pressure data given per facet quadrature point. That call takes the fused row-kernel
launch (``nfx_assemble_darcy_rows_bc``): per-element reference tensors (not the
per-quadrature-point branch), the per-element metric, the XCD-grouped tile map, the CSR
offsets packed into the row words for the lowest-order elements, and the Dirichlet facet
term folded in. Its second call with the same arrays replays the prepared argument list.

Checked against ``oracle.assemble.assemble_darcy`` with the same per-element permeability
(``kappa = 1 / kinv`` as an array, one value per element), for FP32 and FP64, 2-D and
3-D, k = 1 and 2, including the benchmark's own element (3-D RT1-P0) at 16^3 x 6:
indices bit-exact, FP64 values and residual within 1e-12 (per row, relative), FP32 within
1e-5.
"""

import numpy as np
import pytest

from tests._cases import oracle_darcy_space, product_darcy_space

pytestmark = pytest.mark.gpu

CASES = [(2, 8, 1), (2, 4, 2), (3, 4, 1), (3, 2, 2), (3, 16, 1)]


def _row_rel_err(A_ref, A_got):
    D = (A_got - A_ref).tocsr()
    rowmax = np.maximum(abs(A_ref).max(axis=1).toarray().ravel(), 1e-300)
    return (abs(D).max(axis=1).toarray().ravel() / rowmax).max()


@pytest.mark.parametrize("case", CASES, ids=[f"d{d}-n{n}-k{k}" for d, n, k in CASES])
def test_benched_launch_matches_oracle(case):
    import torch

    from nfempy_amd.device import CoefficientArg, DarcySystem
    from nfempy_amd.fields import grf_host
    from oracle.assemble import assemble_darcy

    d, n, k = case
    print(f"[benched launch] d={d} n={n} k={k}", flush=True)
    ps = product_darcy_space(d, n, k)
    sysd = DarcySystem(ps)
    dev = sysd.vals32.device
    cells = sysd.cells.long().cpu().numpy()
    X = sysd.coords64.cpu().numpy()
    cent = X[cells].mean(axis=1)
    logk = grf_host(cent, seed=10000 + n, ell=0.05, sigma=2.0)
    kappa = np.exp(logk)
    f0 = 1.25
    pd_fun = lambda x, y, z: np.array([1.0 - x + 0.5 * y * y])
    xq = sysd.bc_xq.reshape(-1, 3)
    pd_vals = pd_fun(xq[:, 0], xq[:, 1], xq[:, 2]).reshape(-1)
    kinv = CoefficientArg(torch.tensor(1.0 / kappa, dtype=torch.float64, device=dev), 1, 0)
    fsrc = CoefficientArg(torch.full((1,), f0, dtype=torch.float64, device=dev), 0, 0)
    pd = CoefficientArg(torch.tensor(pd_vals, dtype=torch.float64, device=dev), sysd.nqf, 1)

    ops = oracle_darcy_space(d, n, k)
    A, rg = assemble_darcy(ops, {"kappa": kappa, "rhs": np.full(len(kappa), f0)}, {"p": pd_fun})
    assert np.array_equal(sysd.pattern.rowptr, A.indptr) and np.array_equal(sysd.pattern.col, A.indices)
    assert sysd.use_geometry and sysd.use_tile_map
    for fp64, tol in ((False, 1e-5), (True, 1e-12)):
        for call in range(2):  # the second call replays the prepared launch
            vals = sysd.vals64 if fp64 else sysd.vals32
            rhs = sysd.rhs64 if fp64 else sysd.rhs32
            vals.fill_(float("nan"))
            rhs.fill_(float("nan"))
            sysd.assemble(kinv, fsrc, pd, fp64=fp64)
            torch.cuda.synchronize()
            e_val = _row_rel_err(A, sysd.csr_numpy(fp64=fp64))
            r = rhs.double().cpu().numpy()
            e_rhs = np.abs(r - rg).max() / np.abs(rg).max()
            print(f"[benched launch] fp64={fp64} call {call}: values {e_val:.1e}, residual {e_rhs:.1e}", flush=True)
            assert e_val <= tol and e_rhs <= tol, (fp64, call, e_val, e_rhs)
