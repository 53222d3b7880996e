"""The mixed-dimensional oracle pinned by closed forms, and the host-side structure of the
coupled system (pattern, storage coefficients, tip DoFs) against it (CPU only).

Pins: the RT1 facet DoF's normal trace is constant on its face and integrates to +-1, so the
coupling entries are (u, p) = +-1 and (u, u) = (1/alpha) / |F|; the two sides of a fracture
face see opposite traces (normals out of each side); the storage coefficients are
phi |K| / dt and phi delta |T| / dt; the coupled matrix is the block-diagonal sum of the
matrix and fracture systems plus the coupling, storage and tip terms (a planted wrong sign
in the coupling fails the closed forms)."""

import numpy as np
import pytest


@pytest.fixture(scope="module")
def case():
    import oracle.mixed_dim as omd
    from nfempy_amd.mixed_dim import FractureNetwork

    net = FractureNetwork(6, count=2, seed=3)
    fo = net.mesh.entities[2].offset
    normals = [net.fractures[i]["normal"] for i in net.frac_of]
    A, b0, mass, n_u, sz = omd.assemble_md(6, net.faces + fo, normals, net.aperture, kappa_normal=2.0)
    return net, fo, normals, A, b0, mass, n_u, sz


def test_coupling_traces_closed_form(case):
    import oracle.mixed_dim as omd

    net, fo, normals = case[:3]
    ps, mesh = omd.matrix_space(6)
    fc = omd._face_cells(mesh)
    for f, nv in zip(net.faces + fo, normals):
        sides = omd._side_cells(mesh, f, nv, fc)
        tot = []
        for c in sides:
            _, wd, s = omd.coupling_traces(ps, mesh, f, c)
            assert np.ptp(s) < 1e-12
            tot.append(float(np.sum(wd * s)))
            # the positive side's element lies on the +normal side
        assert abs(abs(tot[0]) - 1) < 1e-12 and abs(abs(tot[1]) - 1) < 1e-12
        # orientation: each side's trace sign is its element's facet sign, consistent with
        # the outward normals of the two sides being opposite
        Xf = mesh.points[mesh.cells[f].node_tags].mean(axis=0)
        c0 = mesh.points[mesh.cells[sides[0]].node_tags].mean(axis=0)
        assert (c0 - Xf) @ nv > 0


def test_coupling_entries_closed_form(case):
    net, fo, normals, A, b0, mass, n_u, sz = case
    import oracle.mixed_dim as omd

    ps, mesh = omd.matrix_space(6)
    # the fracture pressure rows hold exactly +-1 in the columns of the two sides' facet DoFs
    n_pf = sz["fracture_pressure"]
    N = sz["N"]
    Ac = A.tocsr()
    for i in range(n_pf):
        r = N - n_pf + i
        row = Ac.getrow(r)
        cols = row.indices
        vals = row.data
        coup = [(c, v) for c, v in zip(cols, vals) if c < sz["matrix_flux"]]
        assert len(coup) == 2, coup
        assert all(abs(abs(v) - 1.0) < 1e-12 for _, v in coup)
        # symmetric placement: (u, p) equals (p, u)
        for c, v in coup:
            assert abs(Ac[c, r] - v) < 1e-12
    # (u, u): the coupling adds (1/alpha) / |F| to the diagonal of both sides' facet DoFs
    # (the difference to an assembly with 1/alpha -> 0)
    A0, *_ = omd.assemble_md(6, net.faces + fo, normals, net.aperture, kappa_normal=1e300)
    d = (Ac.diagonal() - A0.tocsr().diagonal())
    area = np.array([0.5 * np.linalg.norm(np.cross(*(mesh.points[mesh.cells[f].node_tags][1:]
                                                      - mesh.points[mesh.cells[f].node_tags][0])))
                     for f in net.faces + fo])
    expect = np.zeros(N)
    for i in range(len(net.faces)):
        rows = [c for c in Ac.getrow(N - n_pf + i).indices if c < sz["matrix_flux"]]
        for r in rows:
            expect[r] = (net.aperture[i] / (2 * 2.0)) / area[i]
    assert np.allclose(d, expect, rtol=1e-12, atol=1e-14)


def test_storage_closed_form(case):
    net, fo, normals, A, b0, mass, n_u, sz = case
    n_pm = sz["matrix_pressure"]
    h = 1.0 / 6
    assert np.allclose(mass[:n_pm], (h ** 3 / 6) / 1e-3, rtol=1e-12)
    assert np.all(mass[n_pm:] > 0)


def test_host_structure_matches_oracle(case):
    from nfempy_amd.mixed_dim import MixedDimDarcy

    net, fo, normals, A, b0, mass, n_u, sz = case
    md = MixedDimDarcy(net, device="cpu", kappa_normal=2.0)
    assert md.N == sz["N"] and md.n_u == n_u
    assert np.array_equal(md.rowptr_np, sz["pattern"][0]) and np.array_equal(md.col_np, sz["pattern"][1])
    assert np.allclose(md.mass_np, mass, rtol=1e-13, atol=0)
    assert np.array_equal(md.tips, sz["tips"])
    assert np.allclose(md.inv_alpha_np, net.aperture / 4.0, rtol=1e-15)


def test_planted_sign_error_fails_closed_form(case):
    """A coupling assembled with the (p, u) sign flipped (the Darcy convention instead of the
    reference's symmetric coupling) breaks the symmetric-placement check above."""
    net, fo, normals, A, b0, mass, n_u, sz = case
    N, n_pf = sz["N"], sz["fracture_pressure"]
    B = A.tolil()
    r = N - n_pf
    row = A.getrow(r)
    c = [c for c in row.indices if c < sz["matrix_flux"]][0]
    B[r, c] = -A[r, c]
    assert abs(B.tocsr()[c, r] - B.tocsr()[r, c]) > 1.0


def test_fracture_element_is_intrinsic():
    """A fracture triangle's local matrix does not depend on the plane it lies in: the
    oracle evaluates it in its plane's frame, where the reference's 2-D form (which takes
    the x, y components of the mapped basis) is exact; in the plane z = 0 the frame is the
    identity up to a rotation about z, so both give the same matrix."""
    import oracle.mesh as om
    from oracle.forms import darcy_element
    from oracle.mixed_dim import _plane_frame
    from oracle.spaces import ProductSpace as OPS

    def local(pts):
        tri = np.array([[0, 1, 2]])
        ed = np.array([[0, 1], [1, 2], [0, 2]])
        blocks = [om.Block(1, ed, np.full(3, 2)), om.Block(2, tri, np.full(1, 1))]
        m = om.ConformalMesh(2, np.asarray(pts, float), blocks)
        ps = OPS({"u": (2, 1, "RT", 1, m), "p": (2, 1, "Lagrange", 0, m)}, {"p": True})
        ps.build_structures({"u": [1], "p": [1]}, {"u": [2]})
        return darcy_element(ps, 0, {"kappa": np.ones(1), "rhs": np.zeros(1)})[1]

    rng = np.random.default_rng(0)
    P = rng.random((3, 3))
    n = np.cross(P[1] - P[0], P[2] - P[0])
    n /= np.linalg.norm(n)
    t1, t2 = _plane_frame(n)
    flat = np.stack([P @ t1, P @ t2, np.zeros(3)], 1)
    J_flat = local(flat)
    # the same triangle rotated about z stays z = 0: same matrix
    c, s_ = np.cos(0.7), np.sin(0.7)
    Rz = np.array([[c, -s_, 0], [s_, c, 0], [0, 0, 1]])
    assert np.allclose(local(flat @ Rz.T), J_flat, rtol=1e-12, atol=1e-12)
    # the mass block is symmetric positive definite
    M = J_flat[:3, :3]
    assert np.allclose(M, M.T) and np.linalg.eigvalsh(M).min() > 0
