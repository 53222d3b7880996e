"""Shared problem definitions for the tests (oracle side and product side)."""

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MESHES = os.path.join(HERE, "golden", "meshes")
PI = np.pi


def darcy_functions(dim):
    """Manufactured p = prod sin(pi x_i), K = 1: f = dim pi^2 p, p_D = p, u = -grad p."""
    if dim == 2:
        p = lambda x, y, z: np.array([np.sin(PI * x) * np.sin(PI * y)])
        f = lambda x, y, z: np.array([2 * PI * PI * np.sin(PI * x) * np.sin(PI * y)])
        u = lambda x, y, z: np.array([[-PI * np.cos(PI * x) * np.sin(PI * y), -PI * np.sin(PI * x) * np.cos(PI * y)]])
    else:
        p = lambda x, y, z: np.array([np.sin(PI * x) * np.sin(PI * y) * np.sin(PI * z)])
        f = lambda x, y, z: np.array([3 * PI * PI * np.sin(PI * x) * np.sin(PI * y) * np.sin(PI * z)])
        u = lambda x, y, z: np.array([[
            -PI * np.cos(PI * x) * np.sin(PI * y) * np.sin(PI * z),
            -PI * np.sin(PI * x) * np.cos(PI * y) * np.sin(PI * z),
            -PI * np.sin(PI * x) * np.sin(PI * y) * np.cos(PI * z)]])
    return {"rhs": f, "kappa": lambda x, y, z: 1.0}, {"p": p, "u": u}


def oracle_darcy_space(dim, n=None, k=1, msh=None):
    import oracle.mesh as om
    import oracle.spaces as osp

    if msh:
        pts, bl = om.read_msh41(os.path.join(MESHES, msh))
    else:
        pts, bl = om.structured_square(n) if dim == 2 else om.structured_cube(n)
    m = om.ConformalMesh(dim, pts, bl)
    ps = osp.ProductSpace({"u": (dim, 1, "RT", k, m), "p": (dim, 1, "Lagrange", k - 1, m)}, {"p": True})
    btags = [2, 3, 4, 5] if dim == 2 else [2, 3, 4, 5, 6, 7]
    ps.build_structures({"u": [1], "p": [1]}, {"u": btags})
    return ps


def product_darcy_space(dim, n=None, k=1, msh=None):
    from nfempy_amd.mesh import Mesh, unit_cube, unit_square
    from nfempy_amd.spaces import ProductSpace

    if msh:
        mesh = Mesh(dim, os.path.join(MESHES, msh))
        mesh.build_conformal_mesh()
    else:
        pts, bl = unit_square(n) if dim == 2 else unit_cube(n)
        mesh = Mesh.from_blocks(dim, pts, bl)
    ps = ProductSpace({"u": (dim, 1, "RT", k, mesh), "p": (dim, 1, "Lagrange", k - 1, mesh)})
    ps.make_subspaces_discontinuous({"p": True})
    btags = [2, 3, 4, 5] if dim == 2 else [2, 3, 4, 5, 6, 7]
    ps.build_structures({"u": [1], "p": [1]}, {"u": btags})
    return ps


DARCY_CASES = [
    (2, 8, 1, None),
    (2, 32, 1, None),
    (2, 6, 2, None),
    (3, 3, 1, None),
    (3, 2, 2, None),
    (2, None, 1, "example_3_2d_l_1.msh"),
    (2, None, 2, "example_3_2d_l_0.msh"),
    (3, None, 1, "example_3_3d_l_1.msh"),
    (3, None, 2, "example_3_3d_l_0.msh"),
]


def case_id(c):
    d, n, k, msh = c
    return f"d{d}-k{k}-" + (msh.replace(".msh", "") if msh else f"n{n}")


def elasticity_functions(lam=1.0, mu=1.0):
    """Manufactured 3-D solution: sigma = 2 mu eps(u) + lam tr(eps) I, f_u = div sigma, f_t = 0."""
    import sympy as sp

    X = sp.symbols("x y z")
    u = sp.Matrix([sp.sin(sp.pi * X[0]) * sp.sin(sp.pi * X[1]) * sp.sin(sp.pi * X[2]),
                   X[0] * X[1] * X[2] * (1 - X[0]), sp.cos(X[1]) * X[2] ** 2])
    G = u.jacobian(X)
    eps = (G + G.T) / 2
    sig = 2 * mu * eps + lam * eps.trace() * sp.eye(3)
    f = sp.Matrix([sum(sp.diff(sig[i, j], X[j]) for j in range(3)) for i in range(3)])
    fu = sp.lambdify(X, list(f), "numpy")
    us = sp.lambdify(X, list(u), "numpy")
    ss = sp.lambdify(X, sig.tolist(), "numpy")

    def vec(fun, n):
        def g(x, y, z):
            out = fun(x, y, z)
            return np.array([np.broadcast_to(np.asarray(v, dtype=float), np.shape(x)) for v in out][:n])
        return g

    fu6 = vec(fu, 3)

    def rhs(x, y, z):
        a = fu6(x, y, z)
        return np.concatenate([a, np.zeros((3,) + np.shape(x))], axis=0)

    def s_exact(x, y, z):
        return np.array([[np.broadcast_to(np.asarray(v, dtype=float), np.shape(x)) for v in row] for row in ss(x, y, z)])

    fun = {"rhs": rhs, "lambda": lambda x, y, z: lam, "mu": lambda x, y, z: mu}
    exact = {"u": vec(us, 3), "s": s_exact, "t": lambda x, y, z: np.zeros((3,) + np.shape(x))}
    return fun, exact


def oracle_elasticity_space(n=None, msh=None):
    import oracle.mesh as om
    import oracle.spaces as osp

    if msh:
        pts, bl = om.read_msh41(os.path.join(MESHES, msh))
    else:
        pts, bl = om.structured_cube(n)
    m = om.ConformalMesh(3, pts, bl)
    ps = osp.ProductSpace({"s": (3, 3, "BDM", 1, m), "u": (3, 3, "Lagrange", 0, m), "t": (3, 3, "Lagrange", 0, m)},
                          {"u": True, "t": True})
    ps.build_structures({"s": [1], "u": [1], "t": [1]}, {"s": [2, 3, 4, 5, 6, 7]})
    return ps


def product_elasticity_space(n=None, msh=None):
    from nfempy_amd.mesh import Mesh, unit_cube
    from nfempy_amd.spaces import ProductSpace

    if msh:
        mesh = Mesh(3, os.path.join(MESHES, msh))
        mesh.build_conformal_mesh()
    else:
        mesh = Mesh.from_blocks(3, *unit_cube(n))
    ps = ProductSpace({"s": (3, 3, "BDM", 1, mesh), "u": (3, 3, "Lagrange", 0, mesh), "t": (3, 3, "Lagrange", 0, mesh)})
    ps.make_subspaces_discontinuous({"u": True, "t": True})
    ps.build_structures({"s": [1], "u": [1], "t": [1]}, {"s": [2, 3, 4, 5, 6, 7]})
    return ps
