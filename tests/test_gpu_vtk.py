"""GPU parity of the VTK point values (nfx_field_values) against the oracle, and the files
written from device solutions read back (legacy binary VTK).

For RT1-P0 on a cube, the elasticity space (BDM1 rows - P0 - P0) and RT2-P1 on the
reference's 2-D gmsh mesh, with seeded random coefficients: vertex means per mesh point
and centroid values per cell to 1e-12 of the largest value."""

import numpy as np
import pytest

from tests.test_gpu_projector import spaces

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["rt1p0-cube3", "bdm1-elasticity-cube2", "rt2p1-gmsh2d"])
def test_point_and_cell_values(case, tmp_path):
    import oracle.solution as osol
    from nfempy_amd.vtk import (cell_centered_values, node_average_values, read_legacy_vtk,
                                write_vtk_file_with_exact_solution)

    print(f"[vtk] {case}", flush=True)
    ops, ps, fn, dim = spaces(case)
    alpha = np.random.default_rng(7).standard_normal(ps.n_dof)
    names = list(ps.discrete_spaces)
    mesh = next(iter(ps.discrete_spaces.values())).mesh
    n_pts = len(mesh.points)
    for name in names:
        nd = node_average_values(ps, name, alpha, n_pts)
        no = osol.node_average(ops, name, alpha, n_pts)
        sc = max(np.abs(no).max(), 1e-300)
        assert nd.shape == no.shape and np.abs(nd - no).max() <= 1e-12 * sc, name
        cd = cell_centered_values(ps, name, alpha)
        co = osol.cell_centered(ops, name, alpha)
        assert np.abs(cd - co).max() <= 1e-12 * max(np.abs(co).max(), 1e-300), name
    f = tmp_path / "sol.vtk"
    cc = names[1:]
    write_vtk_file_with_exact_solution(str(f), mesh, ps, fn, alpha, cc)
    P, C, PD, CD = read_legacy_vtk(str(f))
    assert np.array_equal(P, mesh.points) and np.array_equal(C, mesh.entities[dim].nodes)
    assert np.allclose(PD[names[0] + "_h"], node_average_values(ps, names[0], alpha, n_pts), rtol=0, atol=0)
    for name in cc:
        assert np.array_equal(CD[name + "_h"], cell_centered_values(ps, name, alpha))
        assert CD[name + "_e"].shape[0] == len(C)
