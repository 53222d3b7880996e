"""The 3 x 3 block CSR copy used by the inner MINRES (host structure, CPU only): every CSR
entry lands in its own slot, and the block product equals the CSR product."""

import numpy as np
import scipy.sparse as sps


def test_block3_structure_reproduces_the_product():
    from nfempy_amd.elasticity import _block3_structure

    rng = np.random.default_rng(3)
    n = 3 * 40
    A = sps.random(n, n, density=0.05, random_state=4, format="csr")
    A = (A + A.T + sps.eye(n)).tocsr()
    A.sort_indices()
    b3 = _block3_structure(n, A.indptr.astype(np.int32), A.indices.astype(np.int32), "cpu")
    nblk = b3["n_blocks"]
    dest = b3["dest"].numpy()
    assert len(np.unique(dest)) == A.nnz and dest.max() < 9 * nblk
    bval = np.zeros(9 * nblk)
    bval[dest] = A.data
    brp, bcol = b3["browptr"].numpy(), b3["bcol"].numpy()
    x = rng.standard_normal(n)
    y = np.zeros(n)
    for R in range(n // 3):
        for j in range(brp[R], brp[R + 1]):
            blk = bval[np.arange(9) * nblk + j].reshape(3, 3)
            y[3 * R: 3 * R + 3] += blk @ x[3 * bcol[j]: 3 * bcol[j] + 3]
    assert np.allclose(y, A @ x, rtol=1e-12, atol=1e-12)
    # block columns ascend within each block row
    assert all(np.all(np.diff(bcol[brp[R]:brp[R + 1]]) > 0) for R in range(n // 3))
