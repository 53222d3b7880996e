"""GPU parity of the device L2 projector against the FP64 oracle.

For each space (RT1-P0 on a structured cube, RT2-P1 on the reference's 2-D gmsh mesh, the
elasticity space BDM1 rows - P0 - P0, continuous vector P1 on the reference's 3-D gmsh
mesh, and a P1 space on a cube scaled so that the reference's value filter drops every
off-diagonal contribution): the assembled FP64 matrix equals the oracle's as an operator
(1e-13 of its largest entry; the device pattern holds the oracle's), the right-hand side to
1e-13, the FP32 values to 1e-6, alpha to 1e-10 relative (FP64-refined solve to 1e-12
against the oracle's direct solve), and l2_error_projected of the same alpha to 1e-12
relative."""

import os

import numpy as np
import pytest
import scipy.sparse as sps

from tests._cases import MESHES

pytestmark = pytest.mark.gpu

PI = np.pi


def sfun(x, y, z):
    return np.array([np.sin(PI * x) * np.cos(2 * y) + z])


def vfun(x, y, z):
    return np.array([[np.sin(2 * x) * y, np.cos(3 * y) + x * z, np.exp(0.5 * z) - x]])


def mfun(x, y, z):
    return np.array([[np.sin(2 * x) * y, np.cos(3 * y) + x * z, np.exp(0.5 * z)],
                     [x * y, np.cos(z), y - z], [np.sin(x + y), 2 + 0 * x, x * z]])


def s3fun(x, y, z):
    return np.array([np.sin(x) + y, np.cos(y) * z, x * y * z])


def spaces(case):
    """(oracle ProductSpace, device ProductSpace, functions, dim)."""
    import oracle.mesh as om
    import oracle.spaces as osp
    from nfempy_amd.mesh import Mesh, unit_cube
    from nfempy_amd.spaces import ProductSpace

    if case == "rt1p0-cube3":
        fields = {"u": (3, 1, "RT", 1), "p": (3, 1, "Lagrange", 0)}
        disc = {"p": True}
        src, dim, scale = ("cube", 3), 3, 1.0
        fn = {"u": vfun, "p": sfun}
    elif case == "rt2p1-gmsh2d":
        fields = {"u": (2, 1, "RT", 2), "p": (2, 1, "Lagrange", 1)}
        disc = {"p": True}
        src, dim, scale = ("example_3_2d_l_0.msh", None), 2, 1.0
        fn = {"u": lambda x, y, z: vfun(x, y, z)[:, :2], "p": sfun}
    elif case == "bdm1-elasticity-cube2":
        fields = {"s": (3, 3, "BDM", 1), "u": (3, 3, "Lagrange", 0), "t": (3, 3, "Lagrange", 0)}
        disc = {"u": True, "t": True}
        src, dim, scale = ("cube", 2), 3, 1.0
        fn = {"s": mfun, "u": s3fun, "t": s3fun}
    elif case == "p1vec-gmsh3d":
        fields = {"q": (3, 3, "Lagrange", 1)}
        disc = {"q": False}
        src, dim, scale = ("example_3_3d_l_1.msh", None), 3, 1.0
        fn = {"q": s3fun}
    elif case == "p1-filtered-cube2":
        fields = {"q": (3, 1, "Lagrange", 1)}
        disc = {"q": False}
        src, dim, scale = ("cube", 2), 3, 0.0193
        fn = {"q": sfun}
    else:
        raise ValueError(case)
    if src[0] == "cube":
        pts, bl = om.structured_cube(src[1])
        dpts, dbl = unit_cube(src[1])
        mesh = Mesh.from_blocks(3, np.asarray(dpts, float) * scale, dbl)
    else:
        pts, bl = om.read_msh41(os.path.join(MESHES, src[0]))
        mesh = Mesh(dim, os.path.join(MESHES, src[0]))
        mesh.build_conformal_mesh()
    m = om.ConformalMesh(dim, np.asarray(pts, float) * scale, bl)
    ops = osp.ProductSpace({k: v + (m,) for k, v in fields.items()}, disc)
    ops.build_structures({k: [1] for k in fields}, {k: [] for k in fields})
    ps = ProductSpace({k: v + (mesh,) for k, v in fields.items()})
    ps.make_subspaces_discontinuous(disc)
    ps.build_structures({k: [1] for k in fields}, {})
    return ops, ps, fn, dim


CASES = ["rt1p0-cube3", "rt2p1-gmsh2d", "bdm1-elasticity-cube2", "p1vec-gmsh3d", "p1-filtered-cube2"]


@pytest.mark.parametrize("case", CASES)
def test_projector_parity(case):
    import torch
    import oracle.projector as op
    from nfempy_amd.projector import l2_error_projected, l2_projector

    print(f"[projector] {case}", flush=True)
    ops, ps, fn, dim = spaces(case)
    assert ops.n_dof == ps.n_dof
    A, rg = op.assemble_projector(ops, fn)
    alpha_o = op.l2_projector(ops, fn)
    info = {}
    alpha = l2_projector(ps, fn, info=info)
    torch.cuda.synchronize()
    n = ps.n_dof
    G = sps.csr_matrix((info["vals64"].cpu().numpy(), info["col"], info["rowptr"]), shape=(n, n))
    scale = abs(A).max()
    assert abs(G - A).max() <= 1e-13 * scale
    # the device pattern holds every entry of the oracle's
    Ao = A.tocoo()
    Gb = (G != 0) + sps.csr_matrix((np.ones(len(info["col"])), info["col"], info["rowptr"]), shape=(n, n))
    assert np.all(np.asarray(Gb[Ao.row, Ao.col]).ravel() != 0)
    v32 = info["vals32"].double().cpu().numpy()
    assert np.abs(v32 - info["vals64"].cpu().numpy()).max() <= 1e-6 * scale
    b = info["b"].cpu().numpy()
    assert np.abs(b + rg).max() <= 1e-13 * np.abs(rg).max()
    rel = np.linalg.norm(alpha - alpha_o) / np.linalg.norm(alpha_o)
    print(f"[projector] {case}: n={n} nnz={info['nnz']} MINRES {info['iterations']} it, alpha rel {rel:.2e}",
          flush=True)
    assert rel <= 1e-10
    e_dev = l2_error_projected(dim, ps, alpha)
    e_or = op.l2_error_projected(ops, alpha, dim)
    assert np.allclose(e_dev, e_or, rtol=1e-12, atol=0)
    if case == "p1-filtered-cube2":
        # the filter left a diagonal matrix
        Gc = G.tocoo()
        assert np.all(Gc.row[Gc.data != 0] == Gc.col[Gc.data != 0])


def test_projector_outputs_and_dof_shift_error():
    import torch
    from nfempy_amd.projector import l2_projector

    ops, ps, fn, dim = spaces("rt1p0-cube3")
    with pytest.raises(IndexError):
        l2_projector(ps, fn, dof_shift=5)
    a_np = l2_projector(ps, fn)
    a_dev = l2_projector(ps, fn, as_numpy=False)
    assert isinstance(a_dev, torch.Tensor) and a_dev.is_cuda
    assert np.array_equal(a_np, a_dev.cpu().numpy())
