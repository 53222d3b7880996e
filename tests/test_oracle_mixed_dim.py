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


def test_signed_schur_diagonal_positions(case):
    """The pressure rows' signed Schur diagonal a_kk - sum_j a_kj a_jk / a_jj (the fracture
    pressure block of the coupled preconditioner) from the host-side positions, on the
    oracle's values in the coupled pattern, against the same quantity from scipy's blocks;
    conservative exchange: positive everywhere (storage plus B M^-1 B^T)."""
    import torch
    import oracle.mixed_dim as omd
    from nfempy_amd.mixed_dim import MixedDimDarcy

    net, fo, normals = case[:3]
    for sign, conservative in ((1, False), (-1, True)):
        A, b0, mass, n_u, sz = omd.assemble_md(6, net.faces + fo, normals, net.aperture, kappa_normal=2.0,
                                               exchange_sign=sign)
        A = A.tocsr()
        md = MixedDimDarcy(net, device="cpu", kappa_normal=2.0, conservative=conservative)
        rows = np.repeat(np.arange(md.N), np.diff(md.rowptr_np))
        md.vals64 = torch.as_tensor(np.asarray(A[rows, md.col_np]).ravel())
        got = md.schur_diag_inverse().double().numpy()
        Auu, Aup, Apu, App = A[:n_u, :n_u], A[:n_u, n_u:], A[n_u:, :n_u], A[n_u:, n_u:]
        S = App.diagonal() - np.asarray((Apu.multiply((Aup / Auu.diagonal()[:, None]).T)).sum(axis=1)).ravel()
        assert np.allclose(got, 1.0 / S, rtol=1e-6, atol=0)
        if conservative:
            assert S.min() > 0


def test_schur_graph_of_split_matrix(case):
    """The two-grid graph of the matrix system with split fracture faces: one pressure row
    per cell, symmetric neighbours, and no neighbour link across a fracture face (the two
    sides' flux DoFs are separate, flow passes through the fracture)."""
    from nfempy_amd.learned import SchurGraph
    from nfempy_amd.mixed_dim import MixedDimDarcy

    net = case[0]
    md = MixedDimDarcy(net, device="cpu", kappa_normal=2.0)
    g = SchurGraph(md.m, device="cpu")
    nbr = g.nbr_np
    E = md.m.n - md.m.n_u
    assert g.E == E and nbr.shape == (E, 4)
    has = nbr >= 0
    p, j = np.nonzero(has)
    assert all(p_ in nbr[q] for p_, q in zip(p, nbr[p, j]))
    cell_row = md.m.ref_dest[:, md.m.nb] - md.m.n_u
    e_of_tet = {c: e for e, c in enumerate(md.m.u_sp.topology.cell_index)}
    for a, b in zip(net.pos_side[:, 0], net.neg_side[:, 0]):
        ra, rb = cell_row[e_of_tet[a]], cell_row[e_of_tet[b]]
        assert rb not in nbr[ra] and ra not in nbr[rb]
    # an interior face away from the fractures still links its two cells
    assert has.sum() > 0.9 * 4 * E - 2 * 6 * 6 * 6 * 2


def test_coupled_preconditioner_elimination(case):
    """The cells' operator of the coupled preconditioner (Schur ELL read from the coupled
    values through the remapped positions, plus storage and the eliminated fracture
    pressures) against the same reduction done densely with scipy on the oracle's matrix:
    P = A_pp - A_pu diag(A_uu)^-1 A_up split into cells m and fracture cells f,
    S' = P_mm - P_mf diag(P_ff)^-1 P_fm; and the block-LU coefficients P_mf / P_ff, P_fm."""
    import scipy.sparse as sps
    import torch
    import oracle.mixed_dim as omd
    from nfempy_amd.mixed_dim import MixedDimDarcy

    net, fo, normals = case[:3]
    A, b0, mass, n_u, sz = omd.assemble_md(6, net.faces + fo, normals, net.aperture, kappa_normal=2.0,
                                           exchange_sign=-1)
    A = A.tocsr()
    md = MixedDimDarcy(net, device="cpu", kappa_normal=2.0, conservative=True)
    rows = np.repeat(np.arange(md.N), np.diff(md.rowptr_np))
    vals = np.asarray(A[rows, md.col_np]).ravel()
    md.vals64 = torch.as_tensor(vals)
    terms = md.coupled_pc_terms()
    g = md._schur_graph
    E = md.m.n - md.m.n_u
    # the Schur ELL (the library's nfx_schur_ell) restated on the host
    pb, pq, pm = (a.numpy().reshape(E, 4) for a in (g.pos_b, g.pos_qb, g.pos_m))
    has = (pb >= 0) & (pm >= 0)
    bv = np.where(has, vals[np.maximum(pb, 0)], 0.0)
    mv = np.where(has, vals[np.maximum(pm, 0)], 1.0)
    qv = np.where(pq >= 0, vals[np.maximum(pq, 0)], 0.0)
    sdiag = (bv * bv / mv).sum(1) + terms["extra_diag"].numpy()
    soff = bv * qv / mv + terms["extra_off"].numpy().reshape(E, 4)
    nbr = g.nbr_np
    r_ = np.repeat(np.arange(E), 4)[nbr.ravel() >= 0]
    S_ell = sps.csr_matrix((soff.ravel()[nbr.ravel() >= 0], (r_, nbr.ravel()[nbr.ravel() >= 0])), shape=(E, E))
    S_ell = S_ell + sps.diags(sdiag)
    # dense reduction
    Dinv = sps.diags(1.0 / A[:n_u, :n_u].diagonal())
    P = (A[n_u:, n_u:] - A[n_u:, :n_u] @ Dinv @ A[:n_u, n_u:]).tocsr()
    Pmm, Pmf, Pfm, Pff = P[:E, :E], P[:E, E:], P[E:, :E], P[E:, E:].diagonal()
    S_red = (Pmm - Pmf @ sps.diags(1.0 / Pff) @ Pfm).toarray()
    err = np.abs(S_ell.toarray() - S_red).max() / np.abs(S_red).max()
    assert err < 1e-12, err
    # symmetric, positive diagonal, links across every fracture face
    assert np.abs(S_red - S_red.T).max() < 1e-12 * np.abs(S_red).max() and sdiag.min() > 0
    assert np.allclose(terms["tail_dinv"].numpy(), 1.0 / Pff, rtol=1e-6)
    st = terms["slot_tail"].numpy().reshape(E, 4)
    sc = terms["slot_coef"].numpy().reshape(E, 4)
    ref = (Pmf @ sps.diags(1.0 / Pff)).tocsr()
    pi, pj = np.nonzero(st >= 0)
    assert len(pi) == 2 * len(net.faces)
    assert np.allclose(sc[pi, pj], np.asarray(ref[pi, st[pi, pj]]).ravel(), rtol=1e-6)
    tc = terms["tail_cells"].numpy().reshape(-1, 2)
    tk = terms["tail_coef"].numpy().reshape(-1, 2)
    fi = np.repeat(np.arange(len(tc)), 2)
    assert np.allclose(tk.ravel(), np.asarray(Pfm[fi, tc.ravel()]).ravel(), rtol=1e-6)
    assert Pfm.nnz == 2 * len(net.faces)
