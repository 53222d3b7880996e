"""VTK output: the point-value oracle pinned by fields inside the space, and the legacy /
XML writers round-tripped (CPU only).

A linear P1 field gives its nodal values at the mesh points and its centroid values at the
cells; a constant RT1 vector gives the constant everywhere; BDM1 rows of a linear matrix
field give the field at every vertex (the mean of equal values). The writer is read back
by its own reader (binary and ASCII legacy VTK) and by an XML parser (.vtu)."""

import xml.etree.ElementTree as ET

import numpy as np
import pytest

from tests.test_oracle_projector import lin3, oracle_space, vec_const


def mat_lin(x, y, z):
    return np.array([[1 + x, y - z, 2 * z], [x * 0 + 1, 2 * x + y, z], [x - y, 3 + 0 * x, y + z]])


@pytest.mark.parametrize("fam,k,nc,f,vector", [("Lagrange", 1, 1, lin3, False), ("RT", 1, 1, vec_const, True),
                                              ("BDM", 1, 3, mat_lin, True)])
def test_point_values_reproduce_fields_in_the_space(fam, k, nc, f, vector):
    from oracle.projector import l2_projector
    from oracle.solution import cell_centered, node_average

    ps = oracle_space(3, fam, k, nc, n=2)
    alpha = l2_projector(ps, {"q": f})
    sp = ps.spaces["q"]
    pts = sp.mesh.points
    nodal = node_average(ps, "q", alpha, len(pts))
    exact = np.array([np.asarray(f(*p), float).ravel() for p in pts])
    assert np.abs(nodal - exact).max() < 1e-10
    cc = cell_centered(ps, "q", alpha)
    cent = np.array([pts[sp.mesh.cells[c].node_tags].mean(axis=0) for c in sp.element_ids])
    exact_c = np.array([np.asarray(f(*p), float).ravel() for p in cent])
    assert np.abs(cc - exact_c).max() < 1e-10


@pytest.mark.parametrize("binary", [True, False])
def test_legacy_vtk_round_trip(tmp_path, binary):
    from nfempy_amd.vtk import read_legacy_vtk, write_unstructured

    rng = np.random.default_rng(0)
    pts = rng.random((9, 3))
    cells = np.array([[0, 1, 2, 3], [4, 5, 6, 7], [1, 8, 3, 5]])
    pd = {"p_h": rng.random((9, 1)), "u_h": rng.random((9, 3))}
    cd = {"s_h": rng.random((3, 9))}
    f = tmp_path / "out.vtk"
    write_unstructured(str(f), pts, cells, 3, pd, cd, binary=binary)
    P, C, PD, CD = read_legacy_vtk(str(f))
    assert np.array_equal(P, pts) and np.array_equal(C, cells)
    for k in pd:
        assert np.array_equal(PD[k], pd[k])
    assert np.array_equal(CD["s_h"], cd["s_h"])
    text = f.read_bytes()
    assert text.startswith(b"# vtk DataFile Version") and b"CELL_TYPES 3" in text


def test_vtu_round_trip(tmp_path):
    from nfempy_amd.vtk import write_unstructured

    rng = np.random.default_rng(1)
    pts = rng.random((4, 2))
    cells = np.array([[0, 1, 2], [1, 3, 2]])
    f = tmp_path / "out.vtu"
    write_unstructured(str(f), pts, cells, 2, {"p_h": rng.random((4, 1))}, {"u_h": rng.random((2, 2))})
    root = ET.parse(str(f)).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    assert piece.get("NumberOfPoints") == "4" and piece.get("NumberOfCells") == "2"
    arrays = {a.get("Name"): np.array(a.text.split(), float) for a in root.iter("DataArray") if a.get("Name")}
    assert np.array_equal(arrays["connectivity"], cells.ravel())
    assert np.array_equal(arrays["types"], [5, 5])
    P = np.array(piece.find("Points/DataArray").text.split(), float).reshape(-1, 3)
    assert np.array_equal(P[:, :2], pts) and np.all(P[:, 2] == 0)
