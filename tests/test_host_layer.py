"""Host layer of the product against the oracle (CPU only).

Entity numbering, DoF destinations (element and boundary facet), the CSR
pattern of the Darcy scatter and the reference-order destination maps are
integer work: they must agree bit for bit. Also checks that the launch-data
validation of ``DarcySystem`` accepts every test case and rejects corrupted
slot codes (the host checks that keep every kernel index inside its buffer).
"""

import numpy as np
import pytest

import oracle.assemble as oa
from tests._cases import DARCY_CASES, case_id, oracle_darcy_space, product_darcy_space


@pytest.mark.parametrize("case", DARCY_CASES, ids=[case_id(c) for c in DARCY_CASES])
def test_destinations_and_pattern(case):
    from nfempy_amd.pattern import build_pattern

    ops = oracle_darcy_space(*case)
    pps = product_darcy_space(*case)
    assert ops.n_dof == pps.n_dof
    E = len(ops.spaces["u"].element_ids)
    assert np.array_equal(np.array(ops.spaces["u"].element_ids), pps.discrete_spaces["u"].element_ids)
    dest_o = np.array([ops.destination_indexes(i) for i in range(E)])
    assert np.array_equal(dest_o, pps.all_destination_indexes())
    nb = len(ops.spaces["u"].bc_element_ids)
    assert np.array_equal(np.array(ops.spaces["u"].bc_element_ids), pps.discrete_spaces["u"].bc_element_ids)
    for j in range(nb):
        assert np.array_equal(ops.bc_destination_indexes(j), pps.bc_destination_indexes(j))
    ref_dest, ref_sign = pps.reference_destinations()
    pat = build_pattern(pps.n_dof, ref_dest, ref_sign)
    rows = np.repeat(dest_o, dest_o.shape[1], axis=1).ravel()
    cols = np.tile(dest_o, (1, dest_o.shape[1])).ravel()
    ip, ix, _ = oa.csr_from_triplets(ops.n_dof, rows, cols, np.ones(len(rows)))
    assert np.array_equal(ip, pat.rowptr) and np.array_equal(ix, pat.col)
    # every reference-order destination is a permutation of the local destinations
    assert np.array_equal(np.sort(ref_dest, axis=1), np.sort(dest_o, axis=1))


def test_structured_generators_match_oracle():
    import oracle.mesh as om
    from nfempy_amd.mesh import unit_cube, unit_square

    for n in (1, 3):
        a = om.structured_square(n)
        b = unit_square(n)
        assert np.allclose(a[0], b[0])
        for x, y in zip(a[1], b[1]):
            assert x.dim == y.dim and np.array_equal(x.data, y.data) and np.array_equal(x.physical, y.physical)
        a = om.structured_cube(n)
        b = unit_cube(n)
        assert np.allclose(a[0], b[0])
        for x, y in zip(a[1], b[1]):
            assert x.dim == y.dim and np.array_equal(x.data, y.data) and np.array_equal(x.physical, y.physical)


def test_entity_counts_of_benchmark_meshes():
    from nfempy_amd.mesh import Mesh, unit_cube, unit_square

    m = Mesh.from_blocks(2, *unit_square(32))
    assert (m.n_vertices, m.n_edges, m.n_faces) == (1089, 3136, 2048)
    m = Mesh.from_blocks(3, *unit_cube(16))
    assert (m.n_vertices, m.n_faces, m.n_volumes) == (4913, 50688, 24576)


def test_msh_reader_matches_oracle(fixtures_dir):
    import os

    import oracle.mesh as om
    from nfempy_amd.mesh import read_msh41

    for f in ("example_3_2d_l_1.msh", "example_3_3d_l_1.msh"):
        a = om.read_msh41(os.path.join(fixtures_dir, "meshes", f))
        b = read_msh41(os.path.join(fixtures_dir, "meshes", f))
        assert np.allclose(a[0], b[0])
        assert len(a[1]) == len(b[1])
        for x, y in zip(a[1], b[1]):
            assert x.dim == y.dim and np.array_equal(x.data, y.data) and np.array_equal(x.physical, y.physical)


def test_oracle_msh_roundtrip(tmp_path):
    import oracle.mesh as om

    pts, bl = om.structured_cube(2)
    p = tmp_path / "cube.msh"
    om.write_msh41(str(p), pts, bl)
    pts2, bl2 = om.read_msh41(str(p))
    assert np.allclose(pts, pts2)
    for x, y in zip(bl, bl2):
        assert np.array_equal(x.data, y.data) and np.array_equal(x.physical, y.physical)


@pytest.mark.parametrize("case", DARCY_CASES[:5], ids=[case_id(c) for c in DARCY_CASES[:5]])
def test_launch_validation(case):
    from nfempy_amd.device import DarcySystem

    sysd = DarcySystem(product_darcy_space(*case), device="cpu")
    bad = sysd.pattern.slot.copy()
    bad[0, 0] = sysd.nnz + 5  # direct slot past the end of the CSR values
    sysd.pattern.slot = bad
    with pytest.raises(ValueError):
        sysd._validate(sysd.u_sp.mesh)


@pytest.mark.parametrize("case", DARCY_CASES[:5], ids=[case_id(c) for c in DARCY_CASES[:5]])
def test_row_positions_device_layout(case):
    """The device copy of the row kernel's column positions decodes back to the byte
    positions and the entry signs: for nl <= 6 one little-endian word per incidence,
    4-bit position fields (15 = absent), then one bit per column b equal to
    ref_sign[e, a] * ref_sign[e, b] < 0, and for nl <= 5 the row's offset inside its
    128-row tile in bits 25..31 of the two words; the bytes themselves otherwise."""
    from nfempy_amd.device import DarcySystem

    sysd = DarcySystem(product_darcy_space(*case), device="cpu")
    pp, nl = sysd.row_pos_np, sysd.nl
    dev = sysd.row_pos.numpy()
    if nl > 6:
        assert np.array_equal(dev, pp.reshape(-1))
        return
    assert dev.size == sysd.n * 2 * 4
    words = dev.reshape(sysd.n, 2, 4).astype(np.int64)
    words = words[..., 0] | words[..., 1] << 8 | words[..., 2] << 16 | words[..., 3] << 24
    for b in range(nl):
        field = (words >> (4 * b)) & 15
        assert np.array_equal(np.where(field == 15, 255, field), pp[:, :, b])
    hi = words >> 25
    if nl <= 5:
        # bits 25..31: the row's CSR offset inside its 128-row tile, split over the two words
        rp = np.asarray(sysd.pattern.rowptr, np.int64)
        r = np.arange(sysd.n)
        assert np.array_equal(hi[:, 0] | hi[:, 1] << 7, rp[:-1] - rp[(r // 128) * 128])
        assert np.all((words >> (5 * nl)) & ((1 << (25 - 5 * nl)) - 1) == 0)
    else:
        assert np.all(words >> (5 * nl) == 0)
    inc = sysd.row_inc_np
    for j in range(2):
        live = np.flatnonzero(inc[j] >= 0)
        e, a = inc[j][live] >> 5, inc[j][live] & 31
        for b in range(nl):
            bit = (words[live, j] >> (4 * nl + b)) & 1
            want = sysd.ref_sign[e, a] * sysd.ref_sign[e, b] < 0
            assert np.array_equal(bit.astype(bool), want)
        dead = np.flatnonzero(inc[j] < 0)
        assert np.all((words[dead, j] >> (4 * nl)) & ((1 << nl) - 1) == 0)


@pytest.mark.parametrize("dim,k", [(2, 1), (2, 2), (3, 1), (3, 2)])
def test_reference_tensors_match_quadrature_sums(dim, k):
    """The per-element fast path of the row assembly kernel (reference tensors Rs, Bc,
    ck appended to the Darcy tables) equals the quadrature sums of the general path
    for a random affine map."""
    from nfempy_amd.device import darcy_tables
    from tests._cases import product_darcy_space

    ps = product_darcy_space(dim, n=2, k=k)
    u_el, p_el = ps.discrete_spaces["u"].element, ps.discrete_spaces["p"].element
    pts, wts = ps.quadrature[dim]
    tab = darcy_tables(u_el, p_el, pts, wts)
    nq, nb, npp = len(wts), u_el.ndofs, p_el.ndofs
    ns = dim * (dim + 1) // 2
    o = 0
    w = tab[o:o + nq]; o += nq
    phi = tab[o:o + nq * nb * dim].reshape(nq, nb, dim); o += nq * nb * dim
    div = tab[o:o + nq * nb].reshape(nq, nb); o += nq * nb
    psi = tab[o:o + nq * npp].reshape(nq, npp); o += nq * npp
    Rs = tab[o:o + nb * nb * ns].reshape(nb, nb, ns); o += nb * nb * ns
    Bc = tab[o:o + nb * npp].reshape(nb, npp); o += nb * npp
    ck = tab[o:o + npp]; o += npp
    assert o == len(tab)
    rng = np.random.default_rng(dim * 10 + k)
    J = rng.standard_normal((3, dim))
    G = J.T @ J
    kinv = 0.37
    M_q = kinv * np.einsum("q,qax,xy,qby->ab", w, phi, G, phi)
    gs = np.array([G[d, d2] for d in range(dim) for d2 in range(d, dim)])
    M_t = kinv * np.einsum("s,abs->ab", gs, Rs)
    assert np.allclose(M_t, M_q, rtol=1e-12, atol=1e-12 * np.abs(M_q).max())
    assert np.allclose(Bc, np.einsum("q,qa,qk->ak", w, div, psi), rtol=1e-12, atol=1e-14)
    assert np.allclose(ck, np.einsum("q,qk->k", w, psi), rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("tile_rows", [64, 128, 256])
def test_span_tiles_cover_rows_within_staging(tile_rows):
    """The compact RT2 tiling: consecutive tiles cover every row once, each holds at most
    tile_rows rows and a CSR span within tile_rows * 33 entries, and a tile stops short of
    tile_rows only where one more row would pass the span limit."""
    from nfempy_amd.batch import RT2_SPAN_PER_ROW, span_tiles
    from nfempy_amd.device import DarcySystem

    sysd = DarcySystem(product_darcy_space(3, n=3, k=2), device="cpu")
    rp = np.asarray(sysd.pattern.rowptr, np.int64)
    starts, counts = span_tiles(rp, tile_rows)
    assert starts[0] == 0 and np.array_equal(starts[1:], (starts + counts)[:-1]) and starts[-1] + counts[-1] == sysd.n
    cap = tile_rows * RT2_SPAN_PER_ROW
    ends = starts + counts
    assert np.all(counts <= tile_rows) and np.all(rp[ends] - rp[starts] <= cap)
    short = (counts < tile_rows) & (ends < sysd.n)
    assert np.all(rp[ends[short] + 1] - rp[starts[short]] > cap)
    assert np.diff(rp).max() > RT2_SPAN_PER_ROW  # the RT2 face rows are longer than the mean


def test_packed_offsets_required_and_validated():
    """For nl <= 5 the row kernel reads each row's in-tile CSR offset from its position
    words: building the words without rowptr is refused, and words whose offsets disagree
    with rowptr fail the launch validation."""
    import torch

    from nfempy_amd.device import DarcySystem, row_pos_device

    sysd = DarcySystem(product_darcy_space(3, n=3, k=1), device="cpu")
    assert sysd.nl <= 5
    with pytest.raises(ValueError):
        row_pos_device(sysd.row_pos_np, sysd.nl, sysd.row_sign_mask_np)
    sysd._validate(sysd.u_sp.mesh)
    good = sysd.row_pos
    w = good.numpy().view(np.uint32).copy()
    w[2 * 5] ^= np.uint32(1 << 25)  # flip one offset bit of row 5
    sysd.row_pos = torch.from_numpy(w.view(np.uint8).copy())
    with pytest.raises(ValueError):
        sysd._validate(sysd.u_sp.mesh)
    sysd.row_pos = good


def test_flux_ell_layout():
    """The slot-major flux-row ELL of the two-grid preconditioner holds exactly the entries
    of rows < n_u with columns < n_u, in CSR order; padding slots point at the row itself."""
    import scipy.sparse as sps

    from nfempy_amd.learned import flux_ell

    A = (sps.random(40, 40, density=0.15, random_state=3, format="csr") + sps.eye(40)).tocsr()
    A.sort_indices()
    n_u = 29
    pos, col = flux_ell(A.indptr, A.indices, n_u)
    blk = A[:n_u, :n_u].tocsr()
    assert pos.shape == col.shape and pos.shape[0] == int(np.diff(blk.indptr).max())
    got = np.zeros((n_u, n_u))
    for s in range(pos.shape[0]):
        for i in range(n_u):
            if pos[s, i] >= 0:
                assert A.indices[pos[s, i]] == col[s, i] and col[s, i] < n_u
                got[i, col[s, i]] += A.data[pos[s, i]]
            else:
                assert col[s, i] == i
    assert np.array_equal(got, blk.toarray())
    # slots in CSR order within each row
    for i in range(n_u):
        p = pos[:, i][pos[:, i] >= 0]
        assert np.all(np.diff(p) > 0)


def test_batch_cold_systems_in_worker_processes():
    """The host half of a batch's systems built in spawned worker processes (space and
    pattern) equals the one built in this process; the small-batch rule keeps tiny batches
    in-process."""
    import concurrent.futures
    import multiprocessing

    from nfempy_amd.batch import _cold_system, cold_workers, darcy_space
    from nfempy_amd.pattern import build_pattern

    keys = [(3, 1), (4, 2), (5, 1)]
    assert cold_workers(keys) == 1
    assert cold_workers([(32, 2), (30, 2), (28, 1)]) > 1 or multiprocessing.cpu_count() == 1
    with concurrent.futures.ProcessPoolExecutor(max_workers=2, mp_context=multiprocessing.get_context("spawn")) as ex:
        got = {k: (ps, pat) for k, ps, pat in ex.map(_cold_system, keys)}
    for key in keys:
        ps = darcy_space(*key)
        rd, rs = ps.reference_destinations()
        ref = build_pattern(ps.n_dof, rd, rs)
        ps2, pat = got[key]
        rd2, rs2 = ps2.reference_destinations()
        assert ps2.n_dof == ps.n_dof and np.array_equal(rd2, rd) and np.array_equal(rs2, rs)
        for f in ("rowptr", "col", "slot", "segptr", "target"):
            assert np.array_equal(getattr(pat, f), getattr(ref, f)), (key, f)
        assert pat.n_stage == ref.n_stage


def test_darcy_system_with_prebuilt_pattern():
    """DarcySystem given the pattern built beforehand has the same pattern and row tables
    as the one that builds it; a pattern of another space is refused."""
    from nfempy_amd.batch import darcy_space
    from nfempy_amd.device import DarcySystem
    from nfempy_amd.pattern import build_pattern

    ps = darcy_space(3, 2)
    rd, rs = ps.reference_destinations()
    pat = build_pattern(ps.n_dof, rd, rs)
    a = DarcySystem(ps, device="cpu")
    b = DarcySystem(ps, device="cpu", pattern=pat)
    assert b.pattern is pat
    for f in ("rowptr", "col", "slot", "segptr", "target"):
        assert np.array_equal(getattr(a.pattern, f), getattr(b.pattern, f))
    for f in ("row_inc_np", "row_pos_np", "row_sign_mask_np", "row_tile_map_np"):
        assert np.array_equal(getattr(a, f), getattr(b, f))
    other = darcy_space(2, 2)
    rdo, rso = other.reference_destinations()
    with pytest.raises(ValueError):
        DarcySystem(ps, device="cpu", pattern=build_pattern(other.n_dof, rdo, rso))
