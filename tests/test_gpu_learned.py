"""GPU parity and behaviour of the learned Schur preconditioner (RT1-P0 Darcy).

Against the FP64 oracle (``oracle/learned.py``) on the same assembled system:
- S in ELL form (pattern exact, values 1e-12), node features (1e-6);
- network forward with bf16 weights/activations: stored activations agree with
  the bf16-emulating oracle except for rare last-bit rounding flips, smoother
  entries within 2e-2 of the Jacobi scale (max) and 2e-3 (mean);
- network backward on the device activations: gradient per parameter group
  within 1e-4 (relative L2) of the FP64 oracle backward;
- loss and its gradient w.r.t. the smoother entries (1e-4 / 1e-3);
- one two-grid preconditioner application (1e-4).
Then: the learned preconditioner in FGMRES reaches an FP64 relative residual of
1e-8; a short training run lowers the two-grid loss on held-out systems and does
not need more iterations than the method without the network; non-finite network
weights fall back to the plain block preconditioner and still converge.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _system(n, seed, sigma=2.0, ell=0.1):
    import torch
    from nfempy_amd.batch import darcy_space
    from nfempy_amd.device import CoefficientArg, DarcySystem
    from nfempy_amd.fields import grf_device

    sysd = DarcySystem(darcy_space(n, 1))
    cent = sysd.coords64[sysd.cells.long()].mean(dim=1)
    logk = grf_device(cent, seed=seed, ell=ell, sigma=sigma)
    kinv = CoefficientArg(torch.exp(-logk).contiguous(), 1, 0)
    f = CoefficientArg(torch.ones(1, dtype=torch.float64, device="cuda"), 0, 0)
    sysd.assemble_elements(kinv, f, fp64=True)
    sysd.assemble_elements(kinv, f, fp64=False)
    torch.cuda.synchronize()
    return sysd, kinv, f


def _theta(seed=11, scale=0.15):
    from nfempy_amd.learned import param_layout

    rng = np.random.default_rng(seed)
    return (rng.standard_normal(param_layout()[1]) * scale).astype(np.float32)


def _bf(t):
    import torch

    return t.view(torch.bfloat16).float().cpu().numpy().astype(np.float64)


def test_learned_parity_with_oracle():
    import scipy.sparse as sps
    import torch
    import oracle.learned as ol
    from nfempy_amd.learned import LearnedSchurPC, SmootherTrainer

    print("[learned] parity n=4", flush=True)
    sysd, _, _ = _system(4, seed=3)
    th = _theta()
    pc = LearnedSchurPC(sysd, th, coarse_fp16=False)
    assert pc.setup()
    torch.cuda.synchronize()
    g = pc.g
    E = g.E
    rp, ci, v64 = sysd.pattern.rowptr, sysd.pattern.col, sysd.vals64.cpu().numpy()
    nbr, sd, so = ol.schur_ell(rp, ci, v64, sysd.n_u)
    assert np.array_equal(g.nbr_np, nbr)
    assert np.allclose(pc.sdiag.cpu().numpy(), sd, rtol=1e-12, atol=0)
    assert np.allclose(pc.soff.cpu().numpy().reshape(E, 4), so, rtol=1e-12, atol=1e-300)
    x_or = ol.features(nbr, sd, so)
    assert np.allclose(pc.x.cpu().numpy().reshape(E, 8), x_or, rtol=1e-6, atol=1e-7)

    # forward
    Wd_or, Wo_or, cache = ol.forward(th.astype(np.float64), nbr, sd, so, emulate_bf16=True)
    EW = E * 64
    h_dev = _bf(pc.h).reshape(5, E, 64)
    for l in range(5):
        ref = cache["h"][l]
        diff = np.abs(h_dev[l] - ref)
        flips = np.mean(diff > 0)
        assert flips < 0.02, (l, flips)
        assert np.all(diff <= 2 ** -6 * np.abs(ref) + 1e-6), l
    Wd, Wo = pc.smoother_numpy()
    errd = np.abs(Wd - Wd_or) * sd
    q = np.where(nbr >= 0, nbr, 0)
    erro = np.abs(Wo - Wo_or) * np.sqrt(sd[:, None] * sd[q])
    print(f"[learned] forward max/mean err diag {errd.max():.2e}/{errd.mean():.2e} off {erro.max():.2e}", flush=True)
    assert errd.max() < 2e-2 and errd.mean() < 2e-3
    assert erro.max() < 2e-2 and erro.mean() < 2e-3

    # loss + gamma on device samples, against the oracle with the device smoother
    tr = SmootherTrainer(pc, th, n_random=6, seed=5)
    pc.setup()
    e = tr.samples()
    loss = float(tr.loss_and_gamma(e))
    L_or, gd_or, go_or = ol.twogrid_loss(Wd.astype(np.float64), Wo.astype(np.float64), nbr, sd, so,
                                         g.agg_of_np, g.n_agg, e.double().cpu().numpy().T)
    assert abs(loss - L_or) <= 1e-4 * abs(L_or)
    gd, go = tr.gd.cpu().numpy(), tr.go.cpu().numpy().reshape(E, 4)
    assert np.abs(gd - gd_or).max() <= 1e-3 * np.abs(gd_or).max()
    assert np.abs(go - go_or).max() <= 1e-3 * np.abs(go_or).max()

    # backward on the device activations
    grad = tr.backward().cpu().numpy().astype(np.float64)
    dev_cache = dict(cache)
    dev_cache["h"] = [h_dev[l] for l in range(5)]
    dev_cache["m"] = [m for m in _bf(pc.m).reshape(4, E, 64)]
    zz = pc.z.cpu().numpy().astype(np.float64).reshape(5, E, 64)
    dev_cache["z0"] = zz[0]
    dev_cache["z"] = [zz[l + 1] for l in range(4)]
    dev_cache["a"] = pc.a.cpu().numpy().astype(np.float64).reshape(E, 4)
    dev_cache["x"] = pc.x.cpu().numpy().astype(np.float64).reshape(E, 8)
    dev_cache["s"] = dev_cache["h"][4][:, None, :] + dev_cache["h"][4][q]
    g_or = ol.backward(th.astype(np.float64), dev_cache, tr.gd.cpu().numpy().astype(np.float64),
                       tr.go.cpu().numpy().astype(np.float64).reshape(E, 4))
    lay, _ = ol.param_layout()
    for name, (o, shape) in lay.items():
        size = int(np.prod(shape))
        a, b = grad[o:o + size], g_or[o:o + size]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        assert rel < 1e-4, (name, rel)

    # one preconditioner application
    r = torch.randn(sysd.n, dtype=torch.float32, device="cuda")
    z = pc(r)
    torch.cuda.synchronize()
    rn = r.double().cpu().numpy()
    zp_or = ol.twogrid_apply(Wd.astype(np.float64), Wo.astype(np.float64), nbr, sd, so, g.agg_of_np, g.n_agg,
                             rn[sysd.n_u:])
    zn = z.double().cpu().numpy()
    assert np.abs(zn[sysd.n_u:] - pc.psign * zp_or).max() <= 1e-4 * np.abs(zp_or).max()
    # flux rows: two damped Jacobi sweeps on the flux mass block (FP32 values on the device)
    assert pc.sweeps_active
    A64 = sps.csr_matrix((v64, ci, rp))
    Mh = A64[:sysd.n_u, :sysd.n_u]
    bound = float((abs(Mh).sum(axis=1).A1 / Mh.diagonal()).max())
    assert abs(pc.flux_bound - bound) <= 1e-6 * bound
    w = pc.flux_weight
    assert abs(w - 2 / (0.5 + bound)) <= 1e-6 and w * bound < 2
    zu_or = ol.flux_sweeps(rp, ci, v64, sysd.n_u, rn[:sysd.n_u], w)
    assert np.abs(zn[:sysd.n_u] - zu_or).max() <= 1e-5 * np.abs(zu_or).max()
    # a fixed weight gives the same operator as the oracle with that weight
    pcw = LearnedSchurPC(sysd, th, graph=g, coarse_fp16=False, flux_omega=0.7)
    assert pcw.setup() and pcw.flux_weight == 0.7
    zw = pcw(r)
    torch.cuda.synchronize()
    zu_w = ol.flux_sweeps(rp, ci, v64, sysd.n_u, rn[:sysd.n_u], 0.7)
    assert np.abs(zw.double().cpu().numpy()[:sysd.n_u] - zu_w).max() <= 1e-5 * np.abs(zu_w).max()
    # with flux_sweeps=1 the flux rows are the diagonal scaling and the pressure rows unchanged
    pc1 = LearnedSchurPC(sysd, th, graph=g, coarse_fp16=False, flux_sweeps=1)
    assert pc1.setup() and not pc1.sweeps_active
    z1 = pc1(r)
    torch.cuda.synchronize()
    z1n = z1.double().cpu().numpy()
    dM = np.array([v64[rp[i]:rp[i + 1]][ci[rp[i]:rp[i + 1]] == i][0] for i in range(sysd.n_u)])
    assert np.allclose(z1n[:sysd.n_u], rn[:sysd.n_u] / np.abs(dM), rtol=1e-5)
    assert np.array_equal(z1n[sysd.n_u:], zn[sysd.n_u:])
    # row-scaled FP16 coarse inverse: same application within 3e-3
    pc16 = LearnedSchurPC(sysd, th, graph=g, coarse_fp16=True)
    assert pc16.setup()
    z16 = pc16(r)
    torch.cuda.synchronize()
    d = (z16 - z).abs().max() / z.abs().max()
    assert float(d) < 3e-3, float(d)


def test_coarse_inverse_and_graph_replays():
    """Recursive block inverse of the 512 x 512 coarse operator against an FP64 inverse;
    graph replays (other systems, eager work in between) against the eager sequence."""
    import torch
    from nfempy_amd.device import CoefficientArg
    from nfempy_amd.fields import grf_device
    from nfempy_amd.learned import LearnedSchurPC

    print("[learned] coarse inverse n=16", flush=True)
    sysd, _, f = _system(16, seed=4)
    pc = LearnedSchurPC(sysd, None)
    assert pc.g.n_agg == 512
    cent = sysd.coords64[sysd.cells.long()].mean(dim=1)
    for rep, seed in enumerate((4, 5, 4, 6)):
        logk = grf_device(cent, seed=seed, ell=0.1, sigma=2.0)
        sysd.assemble_elements(CoefficientArg(torch.exp(-logk).contiguous(), 1, 0), f, fp64=True)
        pc.schur(sysd.vals64)
        pc.coarse(graph=False)
        eager = pc.Ainv.clone()
        Ac = pc._Ac64.clone()
        pc.Ainv.fill_(7.0)
        torch.cuda.synchronize()
        pc.coarse(graph=True)
        torch.cuda.synchronize()
        assert torch.allclose(pc.Ainv, eager, rtol=1e-6, atol=1e-6 * float(eager.abs().max())), rep
        ref = torch.linalg.inv(Ac)
        err = float((pc.Ainv.double() - ref).abs().max() / ref.abs().max())
        print(f"[learned] seed {seed}: inverse rel err {err:.2e}", flush=True)
        assert err < 1e-4
        assert torch.allclose(pc.Ainv, pc.Ainv.T, rtol=0, atol=1e-6 * float(eager.abs().max()))
        _ = (sysd.vals32 * 2).sum()  # eager work between replays
        torch.cuda.synchronize()
    # the 32^3 x 6 size: 4096 aggregates, banded recursion with fill (eager and graph)
    print("[learned] coarse inverse n=32", flush=True)
    sysd, _, f = _system(32, seed=10000, ell=0.05)
    pc = LearnedSchurPC(sysd, None)
    assert pc.g.n_agg == 4096
    pc.schur(sysd.vals64)
    for graph in (False, True):
        pc.coarse(graph=graph)
        torch.cuda.synchronize()
        ref = torch.linalg.inv(pc._Ac64)
        err = float((pc.Ainv.double() - ref).abs().max() / ref.abs().max())
        print(f"[learned] n_agg=4096 graph={graph}: inverse rel err {err:.2e}", flush=True)
        assert err < 1e-4


@pytest.mark.parametrize("n", [1, 5, 16, 17, 100, 127, 128])
def test_spd_inverse_small_block_gauss_jordan(n):
    """nfx_spd_inverse_small (block Gauss-Jordan over 16 x 16 pivots, MFMA updates) on a
    strided SPD block of every size class: under one pivot block, exact multiples,
    ragged tails; entries outside the n x n block are left untouched."""
    import torch
    from nfempy_amd import _native as nat

    g = torch.Generator(device="cuda").manual_seed(n)
    X = torch.randn((n, n), generator=g, device="cuda", dtype=torch.float64)
    A64 = X @ X.T / n + torch.diag(torch.rand(n, generator=g, device="cuda", dtype=torch.float64) + 0.5)
    buf = torch.full((n + 2, n + 9), 3.0, device="cuda")
    A = buf[1:n + 1, 4:n + 4]
    A.copy_(A64.float())
    nat.call("nfx_spd_inverse_small", n, A.data_ptr(), A.stride(0), nat.stream_handle())
    torch.cuda.synchronize()
    ref = torch.linalg.inv(A64)
    err = float((A.double() - ref).abs().max() / ref.abs().max())
    cond = float(torch.linalg.cond(A64))
    print(f"[learned] inverse n={n}: rel err {err:.2e} (cond {cond:.1e})", flush=True)
    assert err < 2e-6 * max(cond, 10.0)
    assert (buf[0] == 3.0).all() and (buf[:, :4] == 3.0).all() and (buf[:, n + 4:] == 3.0).all()


@pytest.mark.parametrize("M,N,K,ta,tb", [(64, 64, 16, 0, 0), (200, 130, 37, 0, 0), (130, 200, 256, 1, 0),
                                         (256, 96, 70, 0, 1), (77, 301, 5, 1, 1), (2048, 2048, 256, 0, 1),
                                         (1500, 1400, 300, 1, 0), (100, 90, 600, 0, 0)])
def test_sgemm_against_library(M, N, K, ta, tb):
    """nfx_sgemm (MFMA FP32, 64 x 64 and 32 x 32 tiles, K in panels of 256) on strided
    row-major views, both transposes, sizes off the tile grid, K past one panel, beta 0
    and 1, and the transposed second output."""
    import torch
    from nfempy_amd.learned import _sgemm

    g = torch.Generator(device="cuda").manual_seed(M * 7 + N)
    Abuf = torch.randn((K + 3, M + 5) if ta else (M + 3, K + 5), generator=g, device="cuda")
    Bbuf = torch.randn((N + 2, K + 4) if tb else (K + 2, N + 4), generator=g, device="cuda")
    A = Abuf[1:K + 1, 2:M + 2] if ta else Abuf[1:M + 1, 2:K + 2]
    B = Bbuf[1:N + 1, 2:K + 2] if tb else Bbuf[1:K + 1, 2:N + 2]
    opA = A.T if ta else A
    opB = B.T if tb else B
    ref = opA.double() @ opB.double()
    Cbuf = torch.randn((M + 1, N + 6), generator=g, device="cuda")
    C = Cbuf[1:, 3:N + 3]
    C0 = C.clone()
    _sgemm(M, N, K, -0.5, A, ta, B, tb, 1.0, C)
    torch.cuda.synchronize()
    exp = C0.double() - 0.5 * ref
    tol = 1e-5 * (ref.abs().max().item() + 1) * max(K, 1) ** 0.5
    assert (C.double() - exp).abs().max().item() <= tol
    C.fill_(float("nan"))
    C2buf = torch.full((N + 2, M + 3), 7.0, device="cuda")
    C2 = C2buf[1:N + 1, 1:M + 1]
    _sgemm(M, N, K, 2.0, A, ta, B, tb, 0.0, C, C2=C2)
    torch.cuda.synchronize()
    assert (C.double() - 2.0 * ref).abs().max().item() <= 2 * tol
    assert torch.isfinite(Cbuf[0]).all() and torch.isfinite(Cbuf[:, :3]).all()
    assert torch.equal(C2, C.T)
    assert (C2buf[0] == 7.0).all() and (C2buf[:, 0] == 7.0).all() and (C2buf[:, M + 1:] == 7.0).all()


def test_fgmres_diagonal_and_none():
    """The C++ FGMRES with a diagonal preconditioner (no two-grid) on a small system."""
    import torch
    from nfempy_amd.learned import Fgmres
    from nfempy_amd.solvers import darcy_preconditioner

    sysd, _, _ = _system(3, seed=1, sigma=1.0)
    sign, pinv = darcy_preconditioner(sysd.rowptr, sysd.col, sysd.vals64, sysd.n, sysd.n_u)
    fg = Fgmres(sysd.n, restart=40)
    x, info = fg.solve(sysd.rowptr, sysd.col, sysd.vals32, sysd.vals64, -sysd.rhs64, pinv=pinv * sign,
                       rtol=1e-8, max_inner=5000)
    fg.close()
    assert info["converged"] and info["residual"] <= 1e-8
    import scipy.sparse as sps
    import scipy.sparse.linalg as spla

    A = sps.csr_matrix((sysd.vals64.cpu().numpy(), sysd.pattern.col, sysd.pattern.rowptr))
    xr = spla.spsolve(A.tocsc(), -sysd.rhs64.cpu().numpy())
    assert np.abs(x.cpu().numpy() - xr).max() <= 1e-6 * np.abs(xr).max()


def test_learned_pc_solves_and_beats_plain_iterations():
    import torch
    from nfempy_amd.learned import LearnedSchurPC, solve_learned
    from nfempy_amd.solvers import MinresIR, darcy_preconditioner

    print("[learned] solve n=8", flush=True)
    sysd, _, _ = _system(8, seed=10000)
    b = -sysd.rhs64
    pc = LearnedSchurPC(sysd, None)
    pc.setup(network=False)
    x, info = solve_learned(sysd, pc, b, rtol=1e-8)
    print(f"[learned] two-grid(jacobi) {info['method']} it={info['iterations']} res={info['residual']:.2e}",
          flush=True)
    assert info["fallback"] is None and info["converged"] and info["residual"] <= 1e-8
    sign, pinv = darcy_preconditioner(sysd.rowptr, sysd.col, sysd.vals64, sysd.n, sysd.n_u)
    solver = MinresIR(sysd.rowptr, sysd.col, sysd.vals32, sysd.vals64, sign, pinv)
    x2, info2 = solver.solve(b, rtol=1e-8, max_inner=100000)
    solver.close()
    print(f"[learned] plain minres it={info2['iterations']}", flush=True)
    assert info2["converged"]
    assert info["iterations"] < info2["iterations"]
    assert torch.allclose(x, x2, rtol=0, atol=1e-6 * float(x2.abs().max()))


def test_flux_sweeps_on_unstructured_mesh():
    """On the reference's gmsh tetrahedral mesh (spectrum of D^-1 M past [0.5, 2]) the
    automatic weight follows the row-sum bound, the device operator C = 2 w D^-1 -
    w^2 D^-1 M D^-1 is symmetric positive definite, and the learned solve with the sweeps
    converges in no more FGMRES iterations than with diag(M)^-1."""
    import scipy.sparse as sps
    import torch
    from nfempy_amd.device import CoefficientArg, DarcySystem
    from nfempy_amd.fields import grf_device
    from nfempy_amd.learned import LearnedSchurPC, default_weights_path, load_params, solve_learned
    from tests._cases import product_darcy_space

    print("[learned] flux sweeps on example_3_3d_l_0.msh", flush=True)
    sysd = DarcySystem(product_darcy_space(3, None, 1, "example_3_3d_l_0.msh"))
    cent = sysd.coords64[sysd.cells.long()].mean(dim=1)
    kinv = CoefficientArg(torch.exp(-grf_device(cent, seed=5, ell=0.1, sigma=1.5)).contiguous(), 1, 0)
    f = CoefficientArg(torch.ones(1, dtype=torch.float64, device="cuda"), 0, 0)
    sysd.assemble_elements(kinv, f, fp64=True)
    sysd.assemble_elements(kinv, f, fp64=False)
    torch.cuda.synchronize()
    theta = load_params(default_weights_path())
    pc = LearnedSchurPC(sysd, theta)
    assert pc.setup() and pc.sweeps_active
    n_u = sysd.n_u
    A = sps.csr_matrix((sysd.vals64.cpu().numpy(), sysd.pattern.col, sysd.pattern.rowptr))
    M = A[:n_u, :n_u].toarray()
    d = np.diag(M)
    bound = float((np.abs(M).sum(1) / d).max())
    lmax = float(np.linalg.eigvalsh(M / np.sqrt(np.outer(d, d))).max())
    w = pc.flux_weight
    assert abs(pc.flux_bound - bound) <= 1e-6 * bound and bound > 2.0 and lmax > 2.0
    assert abs(w - 2 / (0.5 + bound)) <= 1e-6 and w * lmax < 2
    # C from the device coefficients
    coef = pc.flux_coef.cpu().numpy().astype(np.float64).reshape(pc.flux_width, n_u)
    col = pc.flux_col.cpu().numpy().reshape(pc.flux_width, n_u)
    C = np.zeros((n_u, n_u))
    np.add.at(C, (np.broadcast_to(np.arange(n_u), col.shape), col), coef)
    assert np.abs(C - C.T).max() <= 1e-5 * np.abs(C).max()
    Cref = 2 * w * np.diag(1 / d) - w * w * (M / np.outer(d, d))
    assert np.abs(C - Cref).max() <= 1e-5 * np.abs(Cref).max()
    assert np.linalg.eigvalsh(0.5 * (C + C.T)).min() > 0
    x, info = solve_learned(sysd, pc, -sysd.rhs64, rtol=1e-8)
    pc1 = LearnedSchurPC(sysd, theta, graph=pc.g, flux_sweeps=1)
    pc1.setup()
    x1, info1 = solve_learned(sysd, pc1, -sysd.rhs64, rtol=1e-8)
    print(f"[learned] gmsh: sweeps {info['iterations']} it, diag {info1['iterations']} it", flush=True)
    assert info["fallback"] is None and info["converged"] and info1["converged"]
    assert info["iterations"] <= info1["iterations"]
    assert float((x - x1).norm() / x1.norm()) <= 1e-6


def test_training_gain_on_held_out_systems():
    import torch
    from nfempy_amd.learned import LearnedSchurPC, SmootherTrainer, init_params, solve_learned

    print("[learned] train n=8", flush=True)
    sysd, _, _ = _system(8, seed=0)
    pc = LearnedSchurPC(sysd, None)
    tr = SmootherTrainer(pc, init_params(0), lr=3e-3, n_random=12, seed=1)
    from nfempy_amd.device import CoefficientArg
    from nfempy_amd.fields import grf_device

    cent = sysd.coords64[sysd.cells.long()].mean(dim=1)
    f = CoefficientArg(torch.ones(1, dtype=torch.float64, device="cuda"), 0, 0)

    def assemble(seed):
        logk = grf_device(cent, seed=seed, ell=0.1, sigma=2.0)
        kinv = CoefficientArg(torch.exp(-logk).contiguous(), 1, 0)
        sysd.assemble_elements(kinv, f, fp64=True)
        sysd.assemble_elements(kinv, f, fp64=False)

    losses = []
    for step in range(80):
        assemble(step % 40)
        losses.append(tr.step(sysd.vals64))
    print(f"[learned] train loss first {np.mean(losses[:10]):.4f} last {np.mean(losses[-10:]):.4f}", flush=True)
    theta = pc.theta32.clone()
    res = []
    for seed in (10000, 10001):
        assemble(seed)
        pc.set_params(theta)
        l_net = tr.evaluate(sysd.vals64, network=True)
        x, inf_net = solve_learned(sysd, pc, -sysd.rhs64, rtol=1e-8)
        l_jac = tr.evaluate(sysd.vals64, network=False)
        x, inf_jac = solve_learned(sysd, pc, -sysd.rhs64, rtol=1e-8)
        print(f"[learned] seed {seed}: loss net {l_net:.4f} jacobi {l_jac:.4f}; it net {inf_net['iterations']} "
              f"jacobi {inf_jac['iterations']}", flush=True)
        assert inf_net["fallback"] is None and inf_net["converged"]
        res.append((l_net, l_jac, inf_net["iterations"], inf_jac["iterations"]))
    assert all(l_net < 0.8 * l_jac for l_net, l_jac, _, _ in res)
    assert sum(r[2] for r in res) <= sum(r[3] for r in res)


def test_nonfinite_network_falls_back_to_plain():
    import torch
    from nfempy_amd.learned import LearnedSchurPC, solve_learned

    from nfempy_amd.learned import param_layout

    sysd, _, _ = _system(4, seed=2)
    th = _theta()
    th[100] = np.nan
    pc = LearnedSchurPC(sysd, th)
    assert not pc.setup()
    assert pc.fallback_reason == "non-finite network parameters"
    x, info = solve_learned(sysd, pc, -sysd.rhs64, rtol=1e-8)
    assert info["fallback"] == "non-finite network parameters"
    assert info["converged"] and info["residual"] <= 1e-8
    assert bool(torch.isfinite(x).all())
    # finite weights whose smoother overflows: the output check catches it
    th = _theta()
    o = param_layout()[0]["ud"][0]
    th[o:o + 64] = 3e38
    pc = LearnedSchurPC(sysd, th)
    assert not pc.setup()
    assert pc.fallback_reason == "non-finite smoother entries"
    x, info = solve_learned(sysd, pc, -sysd.rhs64, rtol=1e-8)
    assert info["fallback"] == "non-finite smoother entries" and info["converged"]
