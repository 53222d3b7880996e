"""The all-core C/OpenMP restatement of the RT1-P0 Darcy path (the "cpu-opt" baseline,
``oracle/cpu_opt.py``) against the per-element oracle (``oracle.assemble.assemble_darcy``)
and scipy: pattern bit-exact, values and residual to round-off, the vectorised face-sign
rule equal to ``oracle.geometry.face_orientation_3d`` element by element (a mesh with tagged
boundary faces, where dropping the tagged-face rule changes signs), the MINRES solve to an
FP64 relative residual of 1e-8 agreeing with scipy's sparse LU."""

import numpy as np
import pytest
import scipy.sparse as sps
import scipy.sparse.linalg as spla

from tests._cases import oracle_darcy_space


@pytest.fixture(scope="module")
def co():
    import oracle.cpu_opt as co

    co.build()
    return co


@pytest.mark.parametrize("case", [(3, 2, 1, None), (3, 3, 1, None), (3, None, 1, "example_3_3d_l_0.msh")],
                         ids=lambda c: str(c[3] or c[1]))
def test_all_core_assembly_matches_oracle(co, case):
    from oracle.assemble import assemble_darcy

    ops = oracle_darcy_space(*case)
    E = len(ops.spaces["u"].element_ids)
    rng = np.random.default_rng(7)
    kappa, f = np.exp(rng.standard_normal(E)), rng.standard_normal(E)
    A, rg = assemble_darcy(ops, {"rhs": f, "kappa": kappa})
    s = co.DarcyAllCore(co.darcy_inputs(ops))
    ip, ix, data, r = s.assemble(kappa, f)
    assert np.array_equal(ip, A.indptr) and np.array_equal(ix, A.indices)
    assert np.abs(data - A.data).max() <= 1e-13 * np.abs(A.data).max()
    assert np.abs(r - rg).max() <= 1e-13 * np.abs(rg).max()


def test_face_signs_follow_the_orientation_rule(co):
    from oracle.geometry import face_orientation_3d

    ops = oracle_darcy_space(3, None, 1, "example_3_3d_l_0.msh")
    u = ops.spaces["u"]
    m = u.mesh
    cells = [m.cells[c] for c in u.element_ids]
    want = []
    for c in cells:
        bc = [m.cells[f].material_id is not None for f in c.sub_cells_ids[2].tolist()]
        want.append([-1.0 if (not o and ref % 2) else 1.0 for o, _, ref in face_orientation_3d(c, bc)])
    want = np.array(want)
    tags = np.array([c.node_tags for c in cells])
    tagged = np.array([[m.cells[f].material_id is not None for f in c.sub_cells_ids[2].tolist()] for c in cells])
    assert np.array_equal(co.face_signs(tags, tagged), want)
    assert tagged[:, [1, 3]].any()
    # the tagged-face rule matters on this mesh
    assert not np.array_equal(co.face_signs(tags, np.zeros_like(tagged)), want)


def test_all_core_solve(co):
    ops = oracle_darcy_space(3, 4, 1, None)
    E = len(ops.spaces["u"].element_ids)
    kappa = np.exp(np.random.default_rng(3).standard_normal(E))
    s = co.DarcyAllCore(co.darcy_inputs(ops))
    ip, ix, data, rg = s.assemble(kappa, np.ones(E))
    x, its, res = s.solve(data, -rg, ops.n_dofs["u"], rtol=1e-8)
    A = sps.csr_matrix((data, ix, ip), shape=(len(rg), len(rg)))
    assert res <= 1e-8 and its > 0
    assert np.linalg.norm(-rg - A @ x) / np.linalg.norm(rg) <= 1e-8
    xr = spla.spsolve(A.tocsc(), -rg)
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) <= 1e-6
