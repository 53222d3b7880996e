"""The batched two-grid Schur preconditioner (``nfempy_amd.batch_pc``, ``nfx_batch_tg``).

CPU: the per-mesh structure - the S blocks gathered through the stored CSR positions
(``S_cc = sum_a B_a B_a^T / M_a`` over the cell's flux DoFs, ``S_ck`` over the face's DoFs;
for RT2-P1 the same with the 3 x 3 face / interior blocks of M) equal the blocks of the dense
``B D^-1 B^T`` of the same values, every other block of that product is zero (so the
neighbour lists are complete), neighbours are mutual, and the flux blocks cover every flux
row once.

GPU: one application on a mixed RT1-P0 / RT2-P1 batch equals the dense cycle
``x1 = W r; x2 = x1 + P Ac^-1 P^T (r - S x1); z = x2 + W (r - S x2)`` built with scipy from the
instances' FP64 matrices (W = omega S_cc^-1 per cell, P the aggregates' indicator), to 1e-4
relative; the batched MINRES with the cycle reaches 1e-8 against the oracle-assembled
systems in far fewer iterations than with the diagonal, and an instance's result is bitwise
the same when it is solved alone.
"""

import numpy as np
import pytest
import scipy.sparse as sps

SPECS = [(3, 1), (2, 2), (4, 1), (3, 2)]


def _blocks_numpy(st, vals, blocks=False):
    """The S blocks from the stored positions (what the set-up kernels compute): with the
    flux diagonal, or (blocks, RT2-P1) with the 3 x 3 face / interior flux blocks."""
    E, nbf, nb, fd = st["E"], st["nbf"], st["nb"], st["fd"]
    posb = st["posb"].astype(np.int64)
    posd = st["posd"].astype(np.int64)
    sd = np.zeros((E, nb, nb))
    so = np.zeros((E, 4, nb, nb))
    for c in range(E):
        if blocks:
            groups = [(st["fdofs"][k], vals[st["posmf"][c, k]]) for k in range(4)]
            groups.append((np.arange(12, 15), vals[st["posmi"][c]]))
            dinv = [np.linalg.inv(Mg) for _, Mg in groups]
            for (la, _), Di in zip(groups, dinv):
                Bg = vals[posb[c, la]]  # (3, nb)
                sd[c] += Bg.T @ Di @ Bg
        else:
            for a in range(nbf):
                b = vals[posb[c, a]]
                sd[c] += np.outer(b, b) / vals[posd[c, a]]
        for k in range(4):
            d = st["nbr"][c, k]
            if d < 0:
                continue
            if blocks:
                la, la2 = st["fdofs"][k], st["nloc"][c, k]
                so[c, k] = vals[posb[c, la]].T @ dinv[k] @ vals[posb[d, la2]]
                continue
            for m in range(fd):
                a, a2 = st["fdofs"][k, m], st["nloc"][c, k, m]
                so[c, k] += np.outer(vals[posb[c, a]], vals[posb[d, a2]]) / vals[posd[c, a]]
    return sd, so


def _flux_dinv(A, n_u, st):
    """D^-1 of the flux block: diagonal, or the 3 x 3 face / interior blocks (RT2-P1)."""
    M = A[:n_u, :n_u].tocsr()
    if "fb_rows" not in st:
        return sps.diags(1.0 / M.diagonal())
    rows, cols, vals = [], [], []
    for g in st["fb_rows"]:
        inv = np.linalg.inv(M[g][:, g].toarray())
        rows.append(np.repeat(g, 3))
        cols.append(np.tile(g, 3))
        vals.append(inv.ravel())
    Dinv = sps.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=M.shape)
    assert np.array_equal(np.sort(np.concatenate([g for g in st["fb_rows"]])), np.arange(n_u))
    return Dinv


@pytest.mark.parametrize("nk", [(3, 1), (2, 2), (3, 2)])
def test_cell_structure_blocks_match_dense_schur(nk):
    """RT1-P0 with the flux diagonal, RT2-P1 with the flux blocks (random values on the
    pattern, an SPD flux block for RT2)."""
    from nfempy_amd.batch import darcy_space
    from nfempy_amd.batch_pc import cell_structure
    from nfempy_amd.device import DarcySystem

    sysd = DarcySystem(darcy_space(*nk), device="cpu")
    st = cell_structure(sysd)
    n, n_u = sysd.n, sysd.n_u
    rng = np.random.default_rng(1)
    vals = rng.standard_normal(sysd.nnz)
    rp, col = sysd.pattern.rowptr, sysd.pattern.col
    A = sps.csr_matrix((vals, col, rp), shape=(n, n))
    dg = A.diagonal()[:n_u]
    vals[np.flatnonzero(np.repeat(np.arange(n), np.diff(rp)) == col)[:n_u]] = np.abs(dg) + 1.0
    if "posmf" in st:
        # an SPD flux block: M = G G^T + I on the pattern's flux block
        M = A[:n_u, :n_u].toarray()
        M = np.where(M != 0, 0.1 * rng.standard_normal(M.shape), 0.0)
        M = 0.5 * (M + M.T) + np.diag(np.abs(M).sum(axis=1) + 1.0)
        Ad = A.toarray()
        Ad[:n_u, :n_u] = np.where(A[:n_u, :n_u].toarray() != 0, M, 0.0)
        A = sps.csr_matrix(Ad)
        A = sps.csr_matrix((np.asarray(A[np.repeat(np.arange(n), np.diff(rp)), col]).ravel(), col, rp),
                           shape=(n, n))
        vals = A.data
    A = sps.csr_matrix((vals, col, rp), shape=(n, n))
    B = A[n_u:, :n_u]
    blocks = "posmf" in st
    S = (B @ (_flux_dinv(A, n_u, st) if blocks else sps.diags(1.0 / A.diagonal()[:n_u])) @ B.T).toarray()
    sd, so = _blocks_numpy(st, vals, blocks)
    pr = st["prow"] - n_u
    covered = np.zeros_like(S, dtype=bool)
    scale = np.abs(S).max()
    for c in range(st["E"]):
        assert np.abs(sd[c] - S[np.ix_(pr[c], pr[c])]).max() <= 1e-12 * scale
        covered[np.ix_(pr[c], pr[c])] = True
        for k in range(4):
            d = st["nbr"][c, k]
            if d < 0:
                continue
            assert c in st["nbr"][d]
            assert np.abs(so[c, k] - S[np.ix_(pr[c], pr[d])]).max() <= 1e-12 * scale
            covered[np.ix_(pr[c], pr[d])] = True
    assert np.abs(S[~covered]).max(initial=0.0) == 0.0


def _dense_cycle(sysd_vals, st, n_u, r_p, omega, wd_wo=None):
    """The cycle on one instance, dense, from its FP64 matrix (S with the flux blocks of
    RT2-P1); also returns the flux block inverse."""
    n = len(sysd_vals[0]) - 1
    rp, col, vals = sysd_vals
    A = sps.csr_matrix((vals, col, rp), shape=(n, n))
    B = A[n_u:, :n_u]
    Dinv = _flux_dinv(A, n_u, st)
    S = (B @ Dinv @ B.T).toarray()
    pr = st["prow"] - n_u
    E, nb = st["E"], st["nb"]
    W = np.zeros_like(S)
    if wd_wo is None:
        for c in range(E):
            W[np.ix_(pr[c], pr[c])] = omega * np.linalg.inv(S[np.ix_(pr[c], pr[c])])
    else:
        wd, wo = wd_wo
        for c in range(E):
            W[pr[c, 0], pr[c, 0]] = wd[c]
            for k in range(4):
                d = st["nbr"][c, k]
                if d >= 0:
                    W[pr[c, 0], pr[d, 0]] = wo[c, k]
    P = np.zeros((S.shape[0], st["n_agg"]))
    for c in range(E):
        P[pr[c], st["agg"][c]] = 1.0
    Ac = P.T @ S @ P
    x1 = W @ r_p
    x2 = x1 + P @ np.linalg.solve(Ac, P.T @ (r_p - S @ x1))
    return x2 + W @ (r_p - S @ x2), Dinv


@pytest.mark.gpu
@pytest.mark.parametrize("learned", [False, True])
def test_batch_twogrid_apply_matches_dense_cycle(learned):
    import torch
    from nfempy_amd.batch import DarcyBatch
    from nfempy_amd.batch_pc import OMEGA, BatchTwoGrid, cell_structure
    from nfempy_amd.learned import default_weights_path, load_params

    print(f"[batch tg] apply, learned={learned}", flush=True)
    b = DarcyBatch(SPECS, seed=5)
    b.assemble()
    systems = [b.fp64_system(i) for i in range(len(SPECS))]
    sizes = [int(s["rowptr"].numel() - 1) for s in systems]
    offsets = np.concatenate([[0], np.cumsum(sizes)])
    theta = load_params(default_weights_path()) if learned else None
    tg = BatchTwoGrid([inst["system"] for inst in b.instances], offsets, theta=theta)
    tg.setup([s["vals64"] for s in systems])
    torch.cuda.synchronize()
    rng = np.random.default_rng(2)
    r = rng.standard_normal(int(offsets[-1]))
    z = tg.apply(torch.tensor(r, dtype=torch.float32, device="cuda")).double().cpu().numpy()
    g1 = [g for g in tg.groups if g.nb == 1][0]
    if learned:
        assert g1.network_used
        wd = g1.wd.cpu().numpy().astype(np.float64)
        wo = g1.wo.cpu().numpy().astype(np.float64).reshape(-1, 4)
    for i, inst in enumerate(b.instances):
        sysd = inst["system"]
        st = cell_structure(sysd)
        n_u = sysd.n_u
        rows = np.arange(offsets[i] + n_u, offsets[i + 1])
        ww = None
        if learned and st["nb"] == 1:
            j = [m[0] for m in g1.members].index(i)
            c0, c1 = g1.cell_off[j], g1.cell_off[j + 1]
            ww = (wd[c0:c1], wo[c0:c1])
        ref, Dinv = _dense_cycle((sysd.pattern.rowptr, sysd.pattern.col, systems[i]["vals64"].cpu().numpy()), st,
                                 n_u, r[rows], OMEGA[st["nb"]], ww)
        err = np.abs(z[rows] - ref).max() / np.abs(ref).max()
        print(f"[batch tg] instance {i} {SPECS[i]}: rel err {err:.2e}", flush=True)
        assert err <= 1e-4, (i, err)
        zu = z[offsets[i]:offsets[i] + n_u]
        if "fb_rows" in st:  # RT2-P1: the flux blocks
            fu = Dinv @ r[offsets[i]:offsets[i] + n_u]
            assert np.abs(zu - fu).max() <= 1e-5 * np.abs(fu).max(), i
        else:  # RT1-P0: flux rows are the diagonal preconditioner's, not the cycle's
            assert not zu.any()


@pytest.mark.gpu
@pytest.mark.parametrize("pc", ["twogrid", "learned"])
def test_batched_solve_with_twogrid(pc):
    import scipy.sparse.linalg as spla
    import torch
    from nfempy_amd.batch import DarcyBatch
    from oracle.assemble import assemble_darcy
    from tests._cases import oracle_darcy_space

    print(f"[batch tg] solve pc={pc}", flush=True)
    b = DarcyBatch(SPECS, seed=21)
    b.assemble()
    torch.cuda.synchronize()
    _, diag_infos, _ = b.solve_batched()
    xs, infos, dt = b.solve_batched(pc=pc)
    print(f"[batch tg] {pc}: {[i['iterations'] for i in infos]} it vs diag {[i['iterations'] for i in diag_infos]}, "
          f"{dt * 1e3:.1f} ms", flush=True)
    assert 1.6 * sum(q["iterations"] for q in infos) < sum(q["iterations"] for q in diag_infos)
    for i, (n, k) in enumerate(b.specs):
        info = infos[i]
        assert info["converged"] and info["residual"] <= 1e-8, (i, info)
        assert info["iterations"] <= diag_infos[i]["iterations"], (i, info, diag_infos[i])
        inst = b.instances[i]
        ops = oracle_darcy_space(3, n, k)
        kappa = np.exp(inst["logk"].cpu().numpy())
        A, rg = assemble_darcy(ops, {"rhs": np.ones(len(kappa)), "kappa": kappa})
        x = xs[i].cpu().numpy()
        res = np.linalg.norm(-rg - A @ x) / np.linalg.norm(rg)
        xr = spla.spsolve(A.tocsc(), -rg)
        assert res <= 1.2e-8 and np.linalg.norm(x - xr) / np.linalg.norm(xr) <= 1e-6, (i, res)
    # instance 3 alone: bitwise the same
    xa, ia, _ = b.solve_batched([3], pc=pc)
    assert ia[0]["iterations"] == infos[3]["iterations"]
    assert torch.equal(xa[0], xs[3])
