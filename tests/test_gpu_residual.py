"""GPU parity of the FP64 refinement residual r = b - A x (``nfx_residual64`` and the
refinement loops of both Krylov solvers, which share its kernel) against scipy in FP64 on
matrices with rows of 0 to 70 entries: every entry of r within 1e-13 of the row's scale
sum_j |a_ij x_j| + |b_i|, and two calls bitwise equal (fixed-order row sums)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _random_csr(n, max_len, seed):
    import scipy.sparse as sps

    rng = np.random.default_rng(seed)
    lens = rng.integers(0, max_len + 1, n)
    lens[: min(n - 1, 5)] = 0  # empty rows
    lens[-1] = max(lens[-1], 1)  # at least one entry (col / val pointers not null)
    rows = np.repeat(np.arange(n), lens)
    cols = np.concatenate([np.sort(rng.choice(n, size=l, replace=False)) for l in lens]) if lens.sum() else \
        np.zeros(0, np.int64)
    vals = rng.standard_normal(len(rows)) * np.exp(3 * rng.standard_normal(len(rows)))
    return sps.csr_matrix((vals, (rows, cols)), shape=(n, n))


@pytest.mark.parametrize("n,max_len", [(1, 1), (777, 12), (5000, 70), (300000, 9)])
def test_residual64_matches_scipy(n, max_len):
    import torch

    from nfempy_amd import _native as nat

    print(f"[residual64] n={n} max_len={max_len}", flush=True)
    A = _random_csr(n, max_len, seed=n)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(n)
    b = rng.standard_normal(n)
    dev = torch.device("cuda")
    rp = torch.tensor(A.indptr.astype(np.int32), device=dev)
    ci = torch.tensor(A.indices.astype(np.int32), device=dev)
    va = torch.tensor(A.data, dtype=torch.float64, device=dev)
    xd = torch.tensor(x, device=dev)
    bd = torch.tensor(b, device=dev)
    r1 = torch.full((n,), float("nan"), dtype=torch.float64, device=dev)
    r2 = torch.full((n,), float("nan"), dtype=torch.float64, device=dev)
    for r in (r1, r2):
        nat.call("nfx_residual64", n, nat.ptr(rp), nat.ptr(ci), nat.ptr(va), nat.ptr(xd), nat.ptr(bd), nat.ptr(r),
                 nat.stream_handle())
    torch.cuda.synchronize()
    ref = b - A @ x
    scale = abs(A) @ np.abs(x) + np.abs(b)
    got = r1.cpu().numpy()
    assert np.all(np.abs(got - ref) <= 1e-13 * scale), np.max(np.abs(got - ref) / scale)
    assert torch.equal(r1, r2)
