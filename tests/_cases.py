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


def elasticity_functions(lam=1.0, mu=1.0, dim=3, u_expr=None):
    """Manufactured solution of the dual-mixed elasticity forms from a displacement u:
    sigma = 2 mu eps(u) + lam tr(eps) I, f_u = div sigma (the forms' sign: the u rows read
    (div sigma, v) = -r_u = (f_u, v)), f_t = 0, and the rotation the (t, s) coupling of
    ``le_dual_weak_form.py:473-491`` enforces, t = skew(grad u) as
    (T_yz, T_zx, T_xy) with T = (grad u - grad u^T) / 2 in 3-D and T_xy in 2-D.

    3-D default u = (sin(pi x) sin(pi y) sin(pi z), x y z (1 - x), cos(y) z^2); 2-D
    default u is the displacement of ``examples/cauchy_elasticity/
    strong_solution_elasticity_example_1.py:4-21`` (zero on the unit square's boundary).
    Returns (functions, exact) with exact fields "s", "u", "t", "div_s"."""
    import sympy as sp

    X = sp.symbols("x y z")
    if u_expr is None:
        if dim == 3:
            u_expr = [sp.sin(sp.pi * X[0]) * sp.sin(sp.pi * X[1]) * sp.sin(sp.pi * X[2]),
                      X[0] * X[1] * X[2] * (1 - X[0]), sp.cos(X[1]) * X[2] ** 2]
        else:
            s2x, s2y = sp.sin(2 * sp.pi * X[0]), sp.sin(2 * sp.pi * X[1])
            c2x, c2y = sp.cos(2 * sp.pi * X[0]), sp.cos(2 * sp.pi * X[1])
            u_expr = [4 * sp.pi * c2y * s2x ** 2 * s2y, -4 * sp.pi * c2x * s2x * s2y ** 2]
    Xd = X[:dim]
    u = sp.Matrix(u_expr)
    G = u.jacobian(Xd)  # G[c, j] = d u_c / d x_j
    eps = (G + G.T) / 2
    sig = 2 * mu * eps + lam * eps.trace() * sp.eye(dim)
    f = [sum(sp.diff(sig[i, j], Xd[j]) for j in range(dim)) for i in range(dim)]
    T = (G - G.T) / 2
    t = [T[1, 2], T[2, 0], T[0, 1]] if dim == 3 else [T[0, 1]]

    def lamb(exprs):
        g = sp.lambdify(X, exprs, "numpy")

        def out(x, y, z):
            shape = np.shape(x)
            v = g(x, y, z)
            return np.array([np.broadcast_to(np.asarray(e, dtype=float), shape) for e in v])
        return out

    fu = lamb(f)
    us = lamb(list(u))
    ts = lamb(t)
    sflat = lamb([sig[i, j] for i in range(dim) for j in range(dim)])
    zeros_t = len(t)

    def rhs(x, y, z):
        a = fu(x, y, z)
        return np.concatenate([a, np.zeros((zeros_t,) + np.shape(x))], axis=0)

    def s_exact(x, y, z):
        v = sflat(x, y, z)
        return v.reshape((dim, dim) + v.shape[1:])

    fun = {"rhs": rhs, "lambda": lambda x, y, z: lam, "mu": lambda x, y, z: mu}
    exact = {"u": us, "s": s_exact, "t": ts, "div_s": fu}
    return fun, exact


def oracle_elasticity_space(n=None, msh=None, dim=3):
    import oracle.mesh as om
    import oracle.spaces as osp

    if msh:
        pts, bl = om.read_msh41(os.path.join(MESHES, msh))
    else:
        pts, bl = om.structured_cube(n) if dim == 3 else om.structured_square(n)
    m = om.ConformalMesh(dim, pts, bl)
    tc = 3 if dim == 3 else 1
    ps = osp.ProductSpace({"s": (dim, dim, "BDM", 1, m), "u": (dim, dim, "Lagrange", 0, m),
                           "t": (dim, tc, "Lagrange", 0, m)}, {"u": True, "t": True})
    ps.build_structures({"s": [1], "u": [1], "t": [1]}, {"s": [2, 3, 4, 5, 6, 7] if dim == 3 else [2, 3, 4, 5]})
    return ps


def product_elasticity_space(n=None, msh=None, dim=3):
    from nfempy_amd.mesh import Mesh, unit_cube, unit_square
    from nfempy_amd.spaces import ProductSpace

    if msh:
        mesh = Mesh(dim, os.path.join(MESHES, msh))
        mesh.build_conformal_mesh()
    else:
        mesh = Mesh.from_blocks(dim, *(unit_cube(n) if dim == 3 else unit_square(n)))
    tc = 3 if dim == 3 else 1
    ps = ProductSpace({"s": (dim, dim, "BDM", 1, mesh), "u": (dim, dim, "Lagrange", 0, mesh),
                       "t": (dim, tc, "Lagrange", 0, mesh)})
    ps.make_subspaces_discontinuous({"u": True, "t": True})
    ps.build_structures({"s": [1], "u": [1], "t": [1]}, {"s": [2, 3, 4, 5, 6, 7] if dim == 3 else [2, 3, 4, 5]})
    return ps
