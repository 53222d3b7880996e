"""GPU parity of the 2-D dual-mixed elasticity path (BDM1 - P0 - P0 on triangles, the
reference driver's ``dimension = 2`` case) against the oracle.

SBAIJ patterns of A and P bit-exact; FP64 values to 1e-12 and FP32 to 1e-5 per row (the
edge-owner kernel, dim 2); the residual with the stress Dirichlet term of the reference's
exact displacement; the solve to an FP64 relative residual of 1e-8 agreeing with scipy's LU
of the oracle-assembled system, on structured squares and the reference's own gmsh
triangle mesh; the discrete solution converging to the reference's exact fields
(``examples/cauchy_elasticity/strong_solution_elasticity_example_1.py``) at O(h) in the
stress and displacement (the oracle's errors at the same meshes, which its own tests pin).
"""

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from tests._cases import elasticity_functions, oracle_elasticity_space, product_elasticity_space

pytestmark = pytest.mark.gpu


def _row_err(A, G):
    D = (G - A).tocsr()
    rowmax = np.maximum(abs(A).max(axis=1).toarray().ravel(), 1e-300)
    return (abs(D).max(axis=1).toarray().ravel() / rowmax).max()


def _fun(kind, E):
    fun, ex = elasticity_functions(dim=2)
    if kind == "lognormal":
        rng = np.random.default_rng(13)
        fun = dict(fun, **{"lambda": np.exp(0.5 * rng.standard_normal(E)), "mu": np.exp(0.5 * rng.standard_normal(E))})
    return fun, ex


@pytest.mark.parametrize("case", [(2, None), (5, None), (None, "example_3_2d_l_1.msh")],
                         ids=lambda c: str(c[1] or c[0]))
@pytest.mark.parametrize("coef", ["const", "lognormal"])
def test_elasticity_2d_parity(case, coef):
    import torch
    from nfempy_amd.elasticity import ElasticitySystem2D
    from oracle.assemble import assemble_elasticity, symmetric_from_upper

    n, msh = case
    print(f"[elasticity 2d] {case} {coef}", flush=True)
    ops = oracle_elasticity_space(n, msh, dim=2)
    ps = product_elasticity_space(n, msh, dim=2)
    E = len(ps.discrete_spaces["s"].element_ids)
    fun, ex = _fun(coef, E)
    A, P, rg = assemble_elasticity(ops, fun, {"u": ex["u"]})
    sysd = ElasticitySystem2D(ps, fun)
    for which, R in (("A", A), ("P", P)):
        pat = sysd.ops[which]["pattern"]
        assert np.array_equal(pat.rowptr, R.indptr) and np.array_equal(pat.col, R.indices), which
        for fp64, tol in ((True, 1e-12), (False, 1e-5)):
            sysd.assemble(which, fp64=fp64, u_dirichlet=ex["u"])
            torch.cuda.synchronize()
            err = _row_err(R, sysd.csr_numpy(which, fp64))
            assert err <= tol, (which, fp64, err)
            if which == "A":
                r = (sysd.rhs64 if fp64 else sysd.rhs32).double().cpu().numpy()
                assert np.abs(r - rg).max() <= tol * np.abs(rg).max(), fp64
    sysd.assemble("A", fp64=True, u_dirichlet=ex["u"])
    sysd.assemble("A", fp64=False, u_dirichlet=ex["u"])
    x, info = sysd.solve(rtol=1e-8)
    assert info["converged"], info
    Af = symmetric_from_upper(A).tocsc()
    xs = x.cpu().numpy()
    assert np.linalg.norm(-rg - Af @ xs) / np.linalg.norm(rg) <= 1.2e-8
    xr = spla.spsolve(Af, -rg)
    assert np.linalg.norm(xs - xr) / np.linalg.norm(xr) <= 1e-6


def test_elasticity_2d_converges_to_reference_solution():
    import torch
    from nfempy_amd.elasticity import ElasticitySystem2D
    from oracle.assemble import assemble_elasticity, symmetric_from_upper
    from oracle.norms import l2_error

    errs = []
    for n in (4, 8, 16):
        ops = oracle_elasticity_space(n, dim=2)
        ps = product_elasticity_space(n, dim=2)
        fun, ex = elasticity_functions(dim=2)
        sysd = ElasticitySystem2D(ps, fun)
        sysd.assemble("A", fp64=True, u_dirichlet=ex["u"])
        sysd.assemble("A", fp64=False, u_dirichlet=ex["u"])
        torch.cuda.synchronize()
        x, info = sysd.solve(rtol=1e-10)
        assert info["converged"], info
        # the oracle's error functional on the GPU solution (same DoF numbering)
        e = l2_error(ops, ex, x.cpu().numpy())
        print(f"[elasticity 2d] n={n}: L2 errors s, u, t = {e}", flush=True)
        errs.append(e)
    errs = np.asarray(errs)
    rates = np.log2(errs[:-1] / errs[1:])
    assert np.all(rates[-1][:2] > 0.85), (rates, errs)
