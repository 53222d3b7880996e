"""The allocation-free residual entry point (nfx_residual64_noalloc) against nfx_residual64
through the C ABI: bitwise equal outputs on two calls in a row, the second with a smaller
input (the static scratch array keeps the first call's partial sums; nothing may leak from
them into the second result)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _system(n, max_len, seed):
    import scipy.sparse as sps
    import torch

    rng = np.random.default_rng(seed)
    lens = rng.integers(0, max_len + 1, n)
    rows = np.repeat(np.arange(n), lens)
    cols = np.concatenate([np.sort(rng.choice(n, size=l, replace=False)) for l in lens])
    A = sps.csr_matrix((rng.standard_normal(len(rows)), (rows, cols)), shape=(n, n))
    dev = torch.device("cuda")
    return (torch.tensor(A.indptr.astype(np.int32), device=dev), torch.tensor(A.indices.astype(np.int32), device=dev),
            torch.tensor(A.data, device=dev), torch.tensor(rng.standard_normal(n), device=dev),
            torch.tensor(rng.standard_normal(n), device=dev))


def test_noalloc_residual_matches_allocating_entry_point():
    import torch

    from nfempy_amd import _native as nat

    for n, max_len in ((200000, 20), (1500, 9)):
        print(f"[no-alloc residual] n={n}", flush=True)
        rp, ci, va, x, b = _system(n, max_len, seed=n)
        outs = []
        for name in ("nfx_residual64", "nfx_residual64_noalloc"):
            r = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
            nat.call(name, n, nat.ptr(rp), nat.ptr(ci), nat.ptr(va), nat.ptr(x), nat.ptr(b), nat.ptr(r),
                     nat.stream_handle())
            outs.append(r)
        torch.cuda.synchronize()
        assert torch.equal(outs[0], outs[1]), n
        assert bool(torch.isfinite(outs[1]).all())
