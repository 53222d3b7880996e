"""Write ``elasticity_c2_n8.npz``: the FP64 solution of the 3-D dual-mixed elasticity
system (BDM1 stress, P0 displacement, P0 rotation) on the 8^3 x 6 Kuhn cube with the
elasticity workload's coefficient model (per-element log-normal Lame parameters, sigma 0.5,
seed 2024; body force (1, 1/2, 1/4); homogeneous Dirichlet displacement), assembled by the
oracle and solved by scipy's sparse LU (about two minutes on one core).

Run from the repository root: ``python -m tests.golden.make_elasticity_c2``.
"""

import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

N = 8
SEED = 2024


def main():
    import scipy.sparse.linalg as spla

    from nfempy_amd.fields import elasticity_body_force, lame_lognormal
    from oracle.assemble import assemble_elasticity, symmetric_from_upper
    from tests._cases import oracle_elasticity_space

    ops = oracle_elasticity_space(N)
    E = len(ops.spaces["s"].element_ids)
    lam, mu = lame_lognormal(E, SEED)
    fun = {"lambda": lam, "mu": mu, "rhs": elasticity_body_force}
    uD = {"u": lambda x, y, z: np.zeros((3,) + np.shape(x))}
    t0 = time.time()
    A, P, rg = assemble_elasticity(ops, fun, uD)
    Af = symmetric_from_upper(A).tocsc()
    x = spla.spsolve(Af, -rg)
    res = float(np.linalg.norm(-rg - Af @ x) / np.linalg.norm(rg))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), f"elasticity_c2_n{N}.npz")
    np.savez_compressed(out, x=x, n=N, seed=SEED, residual=res,
                        n_dofs=np.array([ops.n_dofs["s"], ops.n_dofs["u"], ops.n_dofs["t"]]))
    print(f"wrote {out}: N={len(x)} residual {res:.2e} ({time.time() - t0:.0f} s)")


if __name__ == "__main__":
    main()
