"""GPU checks of the elasticity two-grid block preconditioner (nfx_elasticity_tg).

Every piece against a dense FP64 numpy restatement built from the assembled FP64
operator itself (no shared code with the kernels): the coupling blocks B (the (u, t) x
stress block of A), the Schur blocks S = B D^-1 B^T (D the 9 x 9 face blocks of the
compliance block), the coarse operator P^T S P, one application of the preconditioner
(stress rows D^-1 r, cell rows the symmetric two-grid cycle), its symmetry and
positivity (MINRES needs an SPD preconditioner); then the solve of the 8^3 x 6 workload
model to 1e-8 against the committed FP64 oracle solution, in fewer MINRES iterations than
the block-Jacobi preconditioner.
"""

import os

import numpy as np
import pytest

from tests._cases import product_elasticity_space

pytestmark = pytest.mark.gpu


def _system(n, seed=2024):
    import torch
    from nfempy_amd.elasticity import ElasticitySystem
    from nfempy_amd.fields import elasticity_body_force, lame_lognormal

    ps = product_elasticity_space(n)
    E = len(ps.discrete_spaces["s"].element_ids)
    lam, mu = lame_lognormal(E, seed)
    fun = {"lambda": lam, "mu": mu, "rhs": elasticity_body_force}
    sysd = ElasticitySystem(ps, fun)
    uD = lambda x, y, z: np.zeros((3,) + np.shape(x))
    sysd.assemble("A", fp64=True, u_dirichlet=uD)
    sysd.assemble("A", fp64=False, u_dirichlet=uD)
    torch.cuda.synchronize()
    return sysd


def _reference(sysd, tg):
    """Dense FP64 restatement: B, D^-1, S, P, Ac and the cell ordering (u0 u1 u2 t0 t1 t2)."""
    import scipy.sparse as sps

    rp, col, _, v64 = sysd.full("A")
    A = sps.csr_matrix((v64.cpu().numpy(), col.cpu().numpy(), rp.cpu().numpy()), shape=(sysd.n, sysd.n)).toarray()
    ns, nu, E = sysd.n_s, sysd.u_sp.n_dof, sysd.E
    cell_rows = np.concatenate([ns + 3 * np.arange(E)[:, None] + np.arange(3),
                                ns + nu + 3 * np.arange(E)[:, None] + np.arange(3)], axis=1)  # (E, 6)
    order = cell_rows.ravel()
    B = A[order][:, :ns]
    Dinv = np.zeros((ns, ns))
    for f in range(ns // 9):
        sl = slice(9 * f, 9 * f + 9)
        Dinv[sl, sl] = np.linalg.inv(A[sl, sl])
    S = B @ Dinv @ B.T
    agg_of = tg.agg_of.cpu().numpy()
    P = np.zeros((6 * E, 6 * tg.n_agg))
    for c in range(E):
        for i in range(6):
            P[6 * c + i, 6 * agg_of[c] + i] = 1.0
    return A, B, Dinv, S, P, order


def test_twogrid_pieces_against_dense_restatement():
    import torch

    print("[elasticity tg] pieces at 3^3 x 6", flush=True)
    sysd = _system(3)
    tg = sysd.twogrid(agg_h=2.0 / 3.0)
    tg.setup()
    torch.cuda.synchronize()
    A, B, Dinv, S, P, order = _reference(sysd, tg)
    E = sysd.E
    # B blocks: the element couplings placed at the cell's faces reproduce the assembled block
    bblk = tg.bblk.cpu().numpy().reshape(E, 4, 6, 9)
    efac = tg.efac.cpu().numpy()
    Bg = np.zeros_like(B)
    for c in range(E):
        for k in range(4):
            Bg[6 * c:6 * c + 6, 9 * efac[c, k]:9 * efac[c, k] + 9] += bblk[c, k]
    assert np.abs(Bg - B).max() <= 1e-12 * np.abs(B).max()
    # S blocks
    S64 = tg.S64.cpu().numpy().reshape(E, 5, 6, 6)
    nbr = tg.nbr.cpu().numpy()
    Sg = np.zeros_like(S)
    for c in range(E):
        Sg[6 * c:6 * c + 6, 6 * c:6 * c + 6] = S64[c, 0]
        for k in range(4):
            if nbr[c, k] >= 0:
                n = nbr[c, k]
                Sg[6 * c:6 * c + 6, 6 * n:6 * n + 6] += S64[c, 1 + k]
    err = np.abs(Sg - S).max() / np.abs(S).max()
    print(f"[elasticity tg] S blocks rel err {err:.1e}", flush=True)
    assert err <= 1e-12
    # coarse operator
    Ac = P.T @ S @ P
    errc = np.abs(tg.Ac64.cpu().numpy() - Ac).max() / np.abs(Ac).max()
    print(f"[elasticity tg] coarse rel err {errc:.1e}", flush=True)
    assert errc <= 1e-12
    # one application against the FP64 cycle
    omega = tg.omega
    W = np.zeros_like(S)
    for c in range(E):
        sl = slice(6 * c, 6 * c + 6)
        W[sl, sl] = omega * np.linalg.inv(S[sl, sl])
    Ainv = np.linalg.inv(Ac)
    rng = np.random.default_rng(3)
    zs = []
    rs = []
    for _ in range(2):
        r = rng.standard_normal(sysd.n).astype(np.float32)
        rc = r[order].astype(np.float64)
        x1 = W @ rc
        x2 = x1 + P @ (Ainv @ (P.T @ (rc - S @ x1)))
        zc = x2 + W @ (rc - S @ x2)
        z = np.zeros(sysd.n)
        z[:sysd.n_s] = Dinv @ r[:sysd.n_s]
        z[order] = zc
        zg = tg(torch.tensor(r, device="cuda")).double().cpu().numpy()
        e = np.linalg.norm(zg - z) / np.linalg.norm(z)
        print(f"[elasticity tg] apply rel err {e:.1e}", flush=True)
        assert e <= 1e-3, e
        assert float(zg @ r) > 0
        zs.append(zg)
        rs.append(r.astype(np.float64))
    # symmetric: r2 . P r1 = r1 . P r2 (FP32 application)
    a, b = rs[1] @ zs[0], rs[0] @ zs[1]
    assert abs(a - b) <= 1e-5 * np.linalg.norm(rs[1]) * np.linalg.norm(zs[0]), (a, b)


def test_twogrid_solve_workload_model_at_8():
    """8^3 x 6 workload model: two-grid MINRES to 1e-8 against the committed FP64 oracle
    solution, fewer iterations than the block-Jacobi (mass_schur) preconditioner."""
    print("[elasticity tg] solve at 8^3 x 6", flush=True)
    sysd = _system(8)
    x, info = sysd.solve(rtol=1e-8, pc_kind="twogrid")
    print(f"[elasticity tg] twogrid: {info['iterations']} it, {info['outer']} refinements, "
          f"residual {info['residual']:.1e}", flush=True)
    assert info["converged"] and info["residual"] <= 1e-8, info
    ref = np.load(os.path.join(os.path.dirname(__file__), "golden", "elasticity_c2_n8.npz"))
    xs, xr = x.cpu().numpy(), ref["x"]
    off = 0
    for name, nd in zip(("s", "u", "t"), ref["n_dofs"].tolist()):
        e = np.linalg.norm(xs[off:off + nd] - xr[off:off + nd]) / np.linalg.norm(xr[off:off + nd])
        print(f"[elasticity tg] field {name}: rel error vs oracle solution {e:.1e}", flush=True)
        assert e <= 1e-6, (name, e)
        off += nd
    _, info_j = sysd.solve(rtol=1e-8, pc_kind="mass_schur")
    print(f"[elasticity tg] mass_schur: {info_j['iterations']} it", flush=True)
    assert info_j["converged"]
    assert info["iterations"] < 0.5 * info_j["iterations"], (info["iterations"], info_j["iterations"])
