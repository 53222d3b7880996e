"""GPU parity of the mixed-dimensional transient Darcy path against the FP64 oracle.

6^3 x 6 tets cut by two seeded fractures: the coupled system assembled on the device (row
kernels for the matrix and fracture blocks, placement into the coupled CSR, interface
coupling, storage terms, tip elimination) against ``oracle.mixed_dim.assemble_md``: pattern
bit-exact, FP64 values and right-hand side to 1e-12 of the row maximum, FP32 values to
1e-5; then backward-Euler steps (FGMRES + FP64 refinement to 1e-10, warm-started, with the
diagonal block, two-grid and learned preconditioners) against the oracle's direct solves, to
1e-8 relative, the coupled preconditioner's rows, and the 24^3 x 6 configuration with eight
fractures for a few steps (convergence and a positive, bounded pressure)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small():
    import oracle.mixed_dim as omd
    from nfempy_amd.mixed_dim import FractureNetwork

    net = FractureNetwork(6, count=2, seed=3)
    fo = net.mesh.entities[2].offset
    normals = [net.fractures[i]["normal"] for i in net.frac_of]
    A, b0, mass, n_u, sz = omd.assemble_md(6, net.faces + fo, normals, net.aperture, kappa_normal=2.0)
    return net, A, b0, mass, n_u, sz


@pytest.mark.parametrize("conservative", [False, True])
def test_md_assembly_parity(small, conservative):
    import torch
    import oracle.mixed_dim as omd
    from nfempy_amd.mixed_dim import MixedDimDarcy

    net, A, b0, mass, n_u, sz = small
    if conservative:
        fo = net.mesh.entities[2].offset
        normals = [net.fractures[i]["normal"] for i in net.frac_of]
        A, b0, mass, n_u, sz = omd.assemble_md(6, net.faces + fo, normals, net.aperture, kappa_normal=2.0,
                                               exchange_sign=-1)
    print(f"[md] assembly parity 6^3, 2 fractures, conservative={conservative}", flush=True)
    md = MixedDimDarcy(net, kappa_normal=2.0, conservative=conservative).setup()
    torch.cuda.synchronize()
    assert np.array_equal(md.rowptr_np, sz["pattern"][0]) and np.array_equal(md.col_np, sz["pattern"][1])
    rowmax = np.maximum(abs(A).max(axis=1).toarray().ravel(), 1e-300)
    for fp64, tol in ((True, 1e-12), (False, 1e-5)):
        G = md.csr_numpy(fp64)
        D = (G - A).tocsr()
        err = (abs(D).max(axis=1).toarray().ravel() / rowmax).max()
        print(f"[md] fp64={fp64}: max row-relative error {err:.2e}", flush=True)
        assert err <= tol
    bb = md.b0.cpu().numpy()
    assert np.abs(bb - b0).max() <= 1e-12 * max(np.abs(b0).max(), 1e-300)


@pytest.mark.parametrize("conservative,pc", [(False, "diag"), (True, "diag"), (True, "twogrid"),
                                             (True, "learned"), (False, "learned")])
def test_md_transient_steps_vs_oracle(small, conservative, pc):
    import torch
    import oracle.mixed_dim as omd
    from nfempy_amd.mixed_dim import MixedDimDarcy

    net, A, b0, mass, n_u, sz = small
    if conservative:
        fo = net.mesh.entities[2].offset
        normals = [net.fractures[i]["normal"] for i in net.frac_of]
        A, b0, mass, n_u, sz = omd.assemble_md(6, net.faces + fo, normals, net.aperture, kappa_normal=2.0,
                                               exchange_sign=-1)
    print(f"[md] transient 6^3, conservative={conservative}, preconditioner {pc}", flush=True)
    md = MixedDimDarcy(net, kappa_normal=2.0, conservative=conservative).setup(preconditioner=pc)
    x = torch.zeros(md.N, dtype=torch.float64, device="cuda")
    xo = np.zeros(md.N)
    for k in range(3):
        x, info = md.step(x, rtol=1e-10)
        xo = omd.step(A, b0, mass, n_u, xo)
        rel = np.linalg.norm(x.cpu().numpy() - xo) / np.linalg.norm(xo)
        print(f"[md] step {k}: {info['iterations']} FGMRES iterations, rel diff {rel:.2e}", flush=True)
        assert info["converged"] and rel <= 1e-8


@pytest.mark.parametrize("conservative", [False, True])
def test_md_coupled_preconditioner_rows(small, conservative):
    """The two-grid descriptor over the coupled vector: flux rows two damped Jacobi sweeps on
    the coupled flux mass block (against the oracle's sweeps on the coupled CSR), or
    1/|a_kk| r without the sweeps; matrix cells
    the two-grid cycle on r_m - sum slot_coef r_f (against the same cycle applied to a
    vector laid out as the matrix system alone); fracture pressures by back-substitution,
    z_f = tail_dinv (r_f - sum tail_coef z_m) (conservative exchange), or tail_dinv r_f."""
    import torch
    import oracle.learned as ol
    from nfempy_amd.mixed_dim import MixedDimDarcy

    net = small[0]
    print(f"[md] coupled preconditioner rows, conservative={conservative}", flush=True)
    md = MixedDimDarcy(net, kappa_normal=2.0, conservative=conservative).setup(preconditioner="twogrid")
    pc = md.pc
    E = md.m.n - md.m.n_u
    n_u = md.n_u
    g = torch.Generator(device="cuda").manual_seed(5)
    r = torch.randn(md.N, generator=g, device="cuda", dtype=torch.float32)
    z = pc(r)
    torch.cuda.synchronize()
    zn, rn = z.cpu().numpy(), r.cpu().numpy()
    assert pc.sweeps_active
    A = md.csr_numpy(fp64=True)
    zu = ol.flux_sweeps(A.indptr, A.indices, A.data, n_u, rn[:n_u].astype(np.float64), pc.flux_weight)
    assert np.abs(zn[:n_u] - zu).max() <= 1e-5 * np.abs(zu).max()
    md0 = MixedDimDarcy(net, kappa_normal=2.0, conservative=conservative, flux_sweeps=False).setup(
        preconditioner="twogrid")
    assert not md0.pc.sweeps_active
    z0 = md0.pc(r).cpu().numpy()
    fl = md0.pinv[:n_u].cpu().numpy()
    assert np.allclose(z0[:n_u], fl * rn[:n_u], rtol=1e-6, atol=0)
    assert np.array_equal(z0[n_u:], zn[n_u:])
    tail = pc.tail_dinv.cpu().numpy()
    rf = rn[n_u + E:]
    rq = rn[n_u: n_u + E].astype(np.float64)
    lu = pc.tail_lu
    assert (lu is not None) == conservative
    if conservative:
        st, sc, tc, tk = (a.cpu().numpy() for a in lu)
        st, sc, tc, tk = st.reshape(E, 4), sc.reshape(E, 4), tc.reshape(-1, 2), tk.reshape(-1, 2)
        rq -= np.where(st >= 0, sc * rf[np.maximum(st, 0)], 0).sum(1)
        zm = zn[n_u: n_u + E]
        expect_f = tail * (rf - np.where(tc >= 0, tk * zm[np.maximum(tc, 0)], 0).sum(1))
        assert np.allclose(zn[n_u + E:], expect_f, rtol=1e-4, atol=1e-5 * np.abs(expect_f).max())
    else:
        assert np.allclose(zn[n_u + E:], tail * rf, rtol=1e-6, atol=0)
    # the cell block: the same cycle on the matrix system's own layout
    saved = pc.flux_dinv, pc.tail_dinv
    pc.flux_dinv = pc.tail_dinv = None
    rm = torch.cat([torch.zeros(md.m.n_u, device="cuda"), torch.as_tensor(rq, dtype=torch.float32, device="cuda")])
    zc = pc(rm).cpu().numpy()
    pc.flux_dinv, pc.tail_dinv = saved
    assert np.allclose(zn[n_u: n_u + E], zc[md.m.n_u:], rtol=1e-4, atol=1e-5 * np.abs(zc).max())


@pytest.mark.parametrize("conservative", [False, True])
def test_md_full_configuration_few_steps(conservative):
    import torch
    from nfempy_amd.mixed_dim import FractureNetwork, MixedDimDarcy

    print(f"[md] 24^3 x 6, 8 fractures, conservative={conservative}: build", flush=True)
    net = FractureNetwork(24, count=8, seed=2026)
    md = MixedDimDarcy(net, conservative=conservative).setup()
    print(f"[md] network smoother used: {md.pc.network_used}", flush=True)
    x = torch.zeros(md.N, dtype=torch.float64, device="cuda")
    for k in range(3):
        x, info = md.step(x, rtol=1e-8)
        print(f"[md] step {k}: {info['iterations']} iterations, residual {info['residual']:.2e}, "
              f"fallback {info['fallback']}", flush=True)
        assert info["converged"] and info["residual"] <= 1e-8
    p = x[md.n_u:].cpu().numpy()
    assert np.all(np.isfinite(p)) and p.max() > 1e-3
    if conservative:
        # boundary data 1 - x, zero initial state, no sources, the exchange conserving mass:
        # 0 <= p <= 1 up to the small over/undershoot RT1-P0 allows
        assert p.min() > -1e-3 and p.max() < 1 + 1e-3
