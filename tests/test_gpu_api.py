"""The drop-in API driven like the reference's Darcy driver
(``examples/md_elliptic_scalar/md_elliptic_scalar.py:206-376``), checked against
the oracle: per-element ``evaluate_form`` (local matrices in the reference's
transformed local order), the assembled system through ``SequentialAssembler``
+ ``Mat``, the KSP solve, and the FP64 error norms of ``l2_error`` / ``div_error``.
"""

import numpy as np
import pytest

from tests._cases import darcy_functions, oracle_darcy_space, product_darcy_space

pytestmark = pytest.mark.gpu

CASES = [(2, None, 1, "example_3_2d_l_1.msh"), (3, None, 1, "example_3_3d_l_0.msh"), (3, 2, 2, None)]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"d{c[0]}k{c[2]}{c[3] or c[1]}")
def test_driver_flow(case):
    from nfempy_amd.assembly import Mat, SequentialAssembler
    from nfempy_amd.ksp import KSP
    from nfempy_amd.postprocess import div_error, l2_error
    from nfempy_amd.weak_forms import LaplaceDualWeakForm, LaplaceDualWeakFormBCDirichlet
    from oracle.assemble import assemble_darcy
    from oracle.forms import darcy_bc_facet, darcy_element
    import oracle.norms as onorms

    d, n, k, msh = case
    print(f"[api] {case}", flush=True)
    fun, ex = darcy_functions(d)
    space = product_darcy_space(d, n, k, msh)
    wf = LaplaceDualWeakForm(space)
    wf.functions = fun
    bc = LaplaceDualWeakFormBCDirichlet(space)
    bc.functions = ex
    ndof = space.n_dof
    A = Mat()
    A.createAIJ([ndof, ndof])
    rg = np.zeros(ndof)
    ne = len(space.discrete_spaces["u"].elements)
    nbc = len(space.discrete_spaces["u"].bc_elements)
    asm = SequentialAssembler(space, A, rg)
    asm.form_to_input_list = {"mixed": ["form", range(ne), wf, np.zeros(ndof)],
                              "dirichlet": ["bc_form", range(nbc), bc, np.zeros(ndof)]}
    asm.scatter_forms()
    A.assemble()

    ops = oracle_darcy_space(d, n, k, msh)
    Aref, rgref = assemble_darcy(ops, fun, ex)
    indptr, indices, data = A.getValuesCSR()
    assert np.array_equal(indptr, Aref.indptr) and np.array_equal(indices, Aref.indices)
    assert np.abs(data - Aref.data).max() <= 1e-12 * np.abs(Aref.data).max()
    assert np.abs(rg - rgref).max() <= 1e-12 * np.abs(rgref).max()

    # element-local evaluate_form in the reference's local order
    for i in sorted({0, ne // 2, ne - 1}):
        r, J = wf.evaluate_form(i, None)
        r0, J0 = darcy_element(ops, i, fun)
        assert np.abs(J - J0).max() <= 1e-12 * max(1.0, np.abs(J0).max())
        assert np.abs(r - r0).max() <= 1e-12 * max(1.0, np.abs(r0).max())
        assert np.array_equal(space.destination_indexes(i), ops.destination_indexes(i))
    for j in sorted({0, nbc - 1}):
        r, _ = bc.evaluate_form(j, None)
        r0, _ = darcy_bc_facet(ops, j, ex)
        assert np.abs(r - r0).max() <= 1e-12 * max(1.0, np.abs(r0).max())

    ksp = KSP().create()
    ksp.setOperators(A)
    b = A.createVecLeft()
    b.array[:] = -rg
    x = A.createVecRight()
    ksp.setType("preonly")
    ksp.getPC().setType("lu")
    ksp.getPC().setFactorSolverType("mumps")
    ksp.setConvergenceHistory()
    ksp.solve(b, x)
    alpha = x.array
    res = np.linalg.norm(-rgref - Aref @ alpha) / np.linalg.norm(rgref)
    assert res <= 1.2e-8

    e_gpu = l2_error(d, space, ex, alpha)
    e_ref = onorms.l2_error(ops, ex, alpha)
    assert np.allclose(e_gpu, e_ref, rtol=1e-10, atol=1e-14), (e_gpu, e_ref)
    dfun = dict(ex)
    c = d * np.pi * np.pi
    dfun["div_u"] = (lambda x, y, z: np.array([c * np.sin(np.pi * x) * np.sin(np.pi * y)])) if d == 2 else \
        (lambda x, y, z: np.array([c * np.sin(np.pi * x) * np.sin(np.pi * y) * np.sin(np.pi * z)]))
    dv_gpu = div_error(d, space, dfun, alpha)
    dv_ref = onorms.div_error(ops, dfun, alpha)
    assert np.allclose(dv_gpu, dv_ref, rtol=1e-10, atol=1e-14), (dv_gpu, dv_ref)


def test_elasticity_driver_ksp():
    """The reference elasticity driver's solver set-up (``mfe_elasticity_example_1.py:238-276``):
    MINRES + additive field split (stress | displacement and rotation), ``rtol = 0``,
    ``atol = 1e-12``, ``divtol = 5000``, ``max_it = 20000``. The facade must stop on the
    absolute test (``||b - A x|| <= 1e-12``, reason CONVERGED_ATOL), a relative test gives
    CONVERGED_RTOL, and a total iteration cap gives DIVERGED_ITS within the cap."""
    import warnings

    from nfempy_amd.assembly import Mat, SequentialAssembler
    from nfempy_amd.ksp import CONVERGED_ATOL, CONVERGED_RTOL, DIVERGED_ITS, IS, KSP
    from nfempy_amd.weak_forms import LEDualWeakForm, LEDualWeakFormBCDirichlet, LERieszMapWeakForm
    from oracle.assemble import assemble_elasticity, symmetric_from_upper
    from tests._cases import elasticity_functions, oracle_elasticity_space, product_elasticity_space

    print("[api] elasticity driver ksp", flush=True)
    fun, ex = elasticity_functions()
    space = product_elasticity_space(2)
    n = space.n_dof
    A, P = Mat().createAIJ([n, n]), Mat().createAIJ([n, n])
    A.setType("sbaij")
    P.setType("sbaij")
    rg = np.zeros(n)
    wf, pf = LEDualWeakForm(space), LERieszMapWeakForm(space)
    wf.functions = pf.functions = fun
    bc = LEDualWeakFormBCDirichlet(space, fun)
    bc.functions = ex
    ne = len(space.discrete_spaces["s"].elements)
    nbc = len(space.discrete_spaces["s"].bc_elements)
    asm = SequentialAssembler(space, A, rg)
    asm.form_to_input_list = {"le": ["form", range(ne), wf, np.zeros(n)],
                              "bc": ["bc_form", range(nbc), bc, np.zeros(n)]}
    asm.scatter_forms()
    pf.scatter(P, None)
    Aref, _, rgref = assemble_elasticity(oracle_elasticity_space(2), fun, ex)
    assert np.abs(rg - rgref).max() <= 1e-12 * np.abs(rgref).max()
    Af = symmetric_from_upper(Aref)
    bounds = np.add.accumulate([0] + list(space.discrete_spaces_dofs.values()))

    def run(rtol, atol, max_it):
        ksp = KSP().create()
        ksp.setOperators(A, P)
        b = A.createVecLeft()
        b.array[:] = -rg
        x = A.createVecRight()
        ksp.setType("minres")
        ksp.getPC().setType("fieldsplit")
        ksp.getPC().setFieldSplitIS(("gen_sigma", IS().createGeneral(np.arange(bounds[0], bounds[1]))),
                                    ("gen_u", IS().createGeneral(np.arange(bounds[1], bounds[3]))))
        ksp.getPC().setFieldSplitType("additive")
        ks, ku = ksp.getPC().getFieldSplitSubKSP()
        ks.setType("preonly")
        ks.getPC().setType("cholesky")
        ks.getPC().setFactorSolverType("mumps")
        ku.setType("preonly")
        ku.getPC().setType("icc")
        ksp.setTolerances(rtol=rtol, atol=atol, divtol=5000, max_it=max_it)
        ksp.setConvergenceHistory()
        ksp.setFromOptions()
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", RuntimeWarning)
            ksp.solve(b, x)
        res = np.linalg.norm(-rgref - Af @ x.array)
        return ksp, res

    ksp, res = run(0.0, 1e-12, 20000)
    assert ksp.getConvergedReason() == CONVERGED_ATOL and res <= 1.5e-12, (ksp.getConvergedReason(), res)
    assert ksp.getConvergenceHistory()[-1] <= 1e-12
    ksp, res = run(1e-8, 0.0, 20000)
    assert ksp.getConvergedReason() == CONVERGED_RTOL and res <= 1.2e-8 * np.linalg.norm(rgref)
    ksp, res = run(0.0, 1e-12, 5)
    assert ksp.getConvergedReason() == DIVERGED_ITS and ksp.getIterationNumber() <= 5
