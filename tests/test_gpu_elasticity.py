"""GPU parity of the dual-mixed elasticity path (BDM1 - P0 - P0, 3-D) against the oracle.

SBAIJ patterns of A (value filter + explicit zero diagonals) and of the Riesz map P
bit-exact; FP64 values to 1e-12 and FP32 to 1e-5 (per row); residual with the stress
Dirichlet term; element-local evaluate_form; MINRES with the block-Jacobi field split
to an FP64 relative residual of 1e-8, agreeing with scipy's solution; the block
preconditioner built on the device equal (pattern bit-exact, FP32 values 1e-6 of the
largest) to the host construction.
"""

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from tests._cases import elasticity_functions, oracle_elasticity_space, product_elasticity_space

pytestmark = pytest.mark.gpu

CASES = [(1, None), (2, None), (None, "example_3_3d_l_0.msh")]


def _coefs(kind, E):
    if kind == "const":
        return elasticity_functions(lam=1.0, mu=1.0)
    fun, ex = elasticity_functions()
    rng = np.random.default_rng(11)
    fun = dict(fun)
    fun["lambda"] = np.exp(0.5 * rng.standard_normal(E))
    fun["mu"] = np.exp(0.5 * rng.standard_normal(E))
    return fun, ex


def _row_err(A, G):
    D = (G - A).tocsr()
    rowmax = np.maximum(abs(A).max(axis=1).toarray().ravel(), 1e-300)
    return (abs(D).max(axis=1).toarray().ravel() / rowmax).max()


@pytest.mark.parametrize("case", CASES, ids=lambda c: str(c[1] or c[0]))
@pytest.mark.parametrize("coef", ["const", "lognormal"])
def test_elasticity_parity(case, coef):
    import torch
    from nfempy_amd.elasticity import ElasticitySystem
    from nfempy_amd.weak_forms import LEDualWeakForm
    from oracle.assemble import assemble_elasticity, symmetric_from_upper
    from oracle.forms import elasticity_element, riesz_element

    n, msh = case
    print(f"[elasticity] {case} {coef}", flush=True)
    ops = oracle_elasticity_space(n, msh)
    E = len(ops.spaces["s"].element_ids)
    fun, ex = _coefs(coef, E)
    A, P, rg = assemble_elasticity(ops, fun, ex)
    ps = product_elasticity_space(n, msh)
    sysd = ElasticitySystem(ps, fun)
    for which, R in (("A", A), ("P", P)):
        pat = sysd.ops[which]["pattern"]
        assert np.array_equal(pat.rowptr, R.indptr) and np.array_equal(pat.col, R.indices), which
        for fp64, tol in ((True, 1e-12), (False, 1e-5)):
            sysd.assemble(which, fp64=fp64, u_dirichlet=ex["u"])
            torch.cuda.synchronize()
            assert _row_err(R, sysd.csr_numpy(which, fp64)) <= tol, (which, fp64)
            if which == "A":
                r = (sysd.rhs64 if fp64 else sysd.rhs32).double().cpu().numpy()
                assert np.abs(r - rg).max() <= tol * np.abs(rg).max(), fp64
    # element-local operator in the reference's local order
    wf = LEDualWeakForm(ps)
    wf.functions = fun
    for i in sorted({0, E // 2, E - 1}):
        r, J = wf.evaluate_form_vectorized(i, None)
        r0, J0 = elasticity_element(ops, i, fun)
        assert np.abs(J - J0).max() <= 1e-12 * np.abs(J0).max()
        assert np.abs(r - r0).max() <= 1e-12 * max(np.abs(r0).max(), 1e-300)
    # solve
    sysd.assemble("A", fp64=True, u_dirichlet=ex["u"])
    sysd.assemble("A", fp64=False, u_dirichlet=ex["u"])
    # the block preconditioner built on the device equals the host construction
    _, (prp, pcol, pval) = sysd.preconditioner(device=True)
    _, (hrp, hcol, hval) = sysd.preconditioner(device=False)
    assert torch.equal(prp, hrp) and torch.equal(pcol, hcol)
    hv = hval.double()
    assert float((pval.double() - hv).abs().max()) <= 1e-6 * float(hv.abs().max())
    x, info = sysd.solve(rtol=1e-8)
    assert info["converged"], info
    Af = symmetric_from_upper(A)
    xs = x.cpu().numpy()
    res = np.linalg.norm(-rg - Af @ xs) / np.linalg.norm(rg)
    assert res <= 1.2e-8, res
    xref = spla.spsolve(Af.tocsc(), -rg)
    assert np.linalg.norm(xs - xref) <= 1e-5 * np.linalg.norm(xref)


def test_elasticity_workload_model_at_8():
    """The elasticity workload's coefficient model (per-element log-normal Lame parameters,
    sigma 0.5, seed 2024; constant body force; zero Dirichlet displacement) at 8^3 x 6
    (3,072 tets, 77,184 DoFs): A and P against the oracle's assembly (patterns bit-exact,
    FP64 1e-12, FP32 1e-5 per row), and the solve to 1e-8 against the committed FP64
    oracle solution (scipy LU of the oracle-assembled system), per field."""
    import os

    import torch
    from nfempy_amd.elasticity import ElasticitySystem
    from nfempy_amd.fields import elasticity_body_force, lame_lognormal
    from oracle.assemble import assemble_elasticity

    print("[elasticity] workload model 8^3", flush=True)
    ops = oracle_elasticity_space(8)
    E = len(ops.spaces["s"].element_ids)
    lam, mu = lame_lognormal(E, 2024)
    fun = {"lambda": lam, "mu": mu, "rhs": elasticity_body_force}
    uD = lambda x, y, z: np.zeros((3,) + np.shape(x))
    A, P, rg = assemble_elasticity(ops, fun, {"u": uD})
    ps = product_elasticity_space(8)
    sysd = ElasticitySystem(ps, fun)
    for which, R in (("A", A), ("P", P)):
        pat = sysd.ops[which]["pattern"]
        assert np.array_equal(pat.rowptr, R.indptr) and np.array_equal(pat.col, R.indices), which
        for fp64, tol in ((True, 1e-12), (False, 1e-5)):
            sysd.assemble(which, fp64=fp64, u_dirichlet=uD)
            torch.cuda.synchronize()
            err = _row_err(R, sysd.csr_numpy(which, fp64))
            print(f"[elasticity] 8^3 {which} fp64={fp64}: {err:.1e}", flush=True)
            assert err <= tol, (which, fp64, err)
            if which == "A":
                r = (sysd.rhs64 if fp64 else sysd.rhs32).double().cpu().numpy()
                assert np.abs(r - rg).max() <= tol * np.abs(rg).max(), fp64
    ref = np.load(os.path.join(os.path.dirname(__file__), "golden", "elasticity_c2_n8.npz"))
    sysd.assemble("A", fp64=True, u_dirichlet=uD)
    sysd.assemble("A", fp64=False, u_dirichlet=uD)
    x, info = sysd.solve(rtol=1e-8)
    assert info["converged"], info
    xs = x.cpu().numpy()
    xr = ref["x"]
    off = 0
    for name, nd in zip(("s", "u", "t"), ref["n_dofs"].tolist()):
        e = np.linalg.norm(xs[off:off + nd] - xr[off:off + nd]) / np.linalg.norm(xr[off:off + nd])
        print(f"[elasticity] 8^3 field {name}: rel error vs oracle solution {e:.1e}", flush=True)
        assert e <= 1e-6, (name, e)
        off += nd


@pytest.mark.parametrize("case", [(1, None), (3, None), (None, "example_3_3d_l_1.msh")], ids=lambda c: str(c[1] or c[0]))
def test_face_kernel_matches_element_kernel(case):
    """The face-owner launch (the default) against the per-element launch with its slot map
    and staged sum: A and P, FP32 and FP64, with per-element and per-point (callable) Lame
    parameters; values within a few ulps of each other (different summation order), the
    residual rows likewise, and two face launches bitwise equal."""
    import torch
    from nfempy_amd.elasticity import ElasticitySystem

    n, msh = case
    print(f"[elasticity faces] {case}", flush=True)
    ps = product_elasticity_space(n, msh)
    E = len(ps.discrete_spaces["s"].element_ids)
    for coef in ("lognormal", "callable"):
        fun, ex = _coefs("lognormal", E)
        if coef == "callable":
            fun = dict(fun)
            fun["lambda"] = lambda x, y, z: 1.0 + 0.5 * x * y
            fun["mu"] = lambda x, y, z: 0.7 + 0.2 * z
        sysd = ElasticitySystem(ps, fun)
        assert sysd.use_faces
        for which in ("A", "P"):
            for fp64, tol in ((True, 1e-13), (False, 2e-6)):
                vf, rf = (t.clone() for t in sysd.assemble(which, fp64=fp64, u_dirichlet=ex["u"], method="faces"))
                vf2, rf2 = (t.clone() for t in sysd.assemble(which, fp64=fp64, u_dirichlet=ex["u"], method="faces"))
                ve, re = (t.clone() for t in sysd.assemble(which, fp64=fp64, u_dirichlet=ex["u"], method="element"))
                torch.cuda.synchronize()
                assert torch.equal(vf, vf2) and torch.equal(rf, rf2)
                scale = float(ve.abs().max())
                err = float((vf.double() - ve.double()).abs().max()) / scale
                print(f"[elasticity faces] {coef} {which} fp64={fp64}: {err:.1e}", flush=True)
                assert err <= tol, (coef, which, fp64, err)
                if which == "A":
                    er = float((rf.double() - re.double()).abs().max()) / max(float(re.abs().max()), 1e-300)
                    assert er <= tol, (coef, fp64, er)
