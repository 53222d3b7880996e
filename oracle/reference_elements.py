"""Reference finite elements (P0, P1, RT1, RT2, BDM1 on triangle/tetrahedron), exact.

The reference creates every element through ``basix.create_element`` with the
``gll_centroid`` Lagrange variant (``src/basis/finite_element.py:85-91``,
``src/basis/element_family.py:4-16``). basix is not vendored, so this module
restates basix's published element definitions (the Ciarlet triple: polynomial
span + degree-of-freedom functionals) and derives the dual basis exactly with
sympy rationals:

- Lagrange P0 / P1 (``gll_centroid``): P0 is the constant 1, P1 is nodal at the
  vertices in vertex order.
- Raviart-Thomas RT_k (basix degree k, k = 1, 2): span (P_{k-1})^d + x P~_{k-1};
  DOFs are facet normal moments against Lagrange P_{k-1} on each facet, then
  (k = 2) interior moments against vector P0 (one per direction).
- Brezzi-Douglas-Marini BDM_1: span (P_1)^d; DOFs are facet normal moments
  against Lagrange P_1 on each facet.

Facet conventions (basix sub-entity tables): triangle edges (1,2), (0,2), (0,1);
tetrahedron faces (1,2,3), (0,2,3), (0,1,3), (0,1,2). The facet normal is NOT
normalised: ``(-t_y, t_x)`` with ``t = v1 - v0`` for an edge, ``(v1-v0)x(v2-v0)``
for a face; moments integrate over the reference facet (length 1 / area 1/2).

Base transformations (``entity_transformations()`` in the reference,
``finite_element.py:108``) are restated per facet-moment space: for P0 moments
a reflection flips the sign and a rotation is the identity; for P1 moments
``rotation = [[0,0,1],[1,0,0],[0,1,0]]`` and
``reflection = -[[1,0,0],[0,0,1],[0,1,0]]`` (triangle faces) and
``reflection = -[[0,1],[1,0]]`` (edges).

Pinned by: ``l_i(phi_j) = delta_ij`` re-evaluated with an independent numeric
quadrature, the textbook closed forms of RT1, and (in the tests) normal
continuity of the transformed bases across every interior facet of real
meshes, which only holds with the transformation matrices above combined with
the reference's (rotation, reflection) table (``permute_and_transform.py:83-100``).
The basix sign/scale per facet and the DOF order inside ``entity_dofs`` are
pinned only by the mathematics above: *parity unpinned* against a basix import.
"""

from __future__ import annotations

import functools
import itertools

import numpy as np
import sympy as sp

X = sp.symbols("x y z")

TRI_VERTS = [(0, 0), (1, 0), (0, 1)]
TET_VERTS = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)]
TRI_FACETS = [(1, 2), (0, 2), (0, 1)]
TET_FACETS = [(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)]


def ref_vertices(dim):
    return TRI_VERTS if dim == 2 else TET_VERTS


def ref_facets(dim):
    return TRI_FACETS if dim == 2 else TET_FACETS


def _monomials(dim, degree, homogeneous=False):
    xs = X[:dim]
    out = []
    for total in range(degree + 1):
        if homogeneous and total != degree:
            continue
        for powers in itertools.product(range(total + 1), repeat=dim):
            if sum(powers) == total:
                out.append(sp.Mul(*[v**p for v, p in zip(xs, powers)]))
    return out


def _cell_integral(expr, dim):
    x, y, z = X
    if dim == 2:
        return sp.integrate(sp.integrate(expr, (y, 0, 1 - x)), (x, 0, 1))
    return sp.integrate(
        sp.integrate(sp.integrate(expr, (z, 0, 1 - x - y)), (y, 0, 1 - x)), (x, 0, 1)
    )


def _facet_lagrange(fdim, degree, s):
    """Lagrange P_degree nodal basis on the reference facet (vertex order)."""
    if degree == 0:
        return [sp.Integer(1)]
    if degree == 1:
        if fdim == 1:
            return [1 - s[0], s[0]]
        return [1 - s[0] - s[1], s[0], s[1]]
    raise NotImplementedError(degree)


def _facet_integral(expr, fdim, s):
    if fdim == 1:
        return sp.integrate(expr, (s[0], 0, 1))
    return sp.integrate(sp.integrate(expr, (s[1], 0, 1 - s[0])), (s[0], 0, 1))


def facet_normal(dim, facet):
    v = [sp.Matrix(ref_vertices(dim)[i]) for i in facet]
    if dim == 2:
        t = v[1] - v[0]
        return sp.Matrix([-t[1], t[0]])
    return (v[1] - v[0]).cross(v[2] - v[0])


class ReferenceElement:
    """Exact reference element: basis functions as sympy expressions."""

    def __init__(self, family, dim, degree):
        self.family = family
        self.dim = dim
        self.degree = degree
        self.value_size = 1 if family == "P" else dim
        self._build()

    # -- span ---------------------------------------------------------------
    def _span(self):
        d = self.dim
        if self.family == "P":
            return [sp.Matrix([m]) for m in _monomials(d, self.degree)]
        if self.family == "BDM":
            out = []
            for c in range(d):
                for m in _monomials(d, self.degree):
                    v = [0] * d
                    v[c] = m
                    out.append(sp.Matrix(v))
            return out
        if self.family == "RT":
            out = []
            for c in range(d):
                for m in _monomials(d, self.degree - 1):
                    v = [0] * d
                    v[c] = m
                    out.append(sp.Matrix(v))
            for m in _monomials(d, self.degree - 1, homogeneous=True):
                out.append(sp.Matrix([X[c] * m for c in range(d)]))
            return out
        raise ValueError(self.family)

    # -- functionals ----------------------------------------------------------
    def _functionals(self):
        d = self.dim
        fdim = d - 1
        s = sp.symbols("s0 s1")[:fdim]
        funcs = []
        entity_dofs = [[[] for _ in range(n)] for n in ((3, 3, 1) if d == 2 else (4, 6, 4, 1))]
        if self.family == "P":
            if self.degree == 0:
                funcs.append(lambda v: _cell_integral(v[0], d) / _cell_integral(1, d))
                entity_dofs[d][0] = [0]
            else:
                for vi, vert in enumerate(ref_vertices(d)):
                    sub = dict(zip(X[:d], vert))
                    funcs.append(lambda v, sub=sub: v[0].subs(sub))
                    entity_dofs[0][vi] = [vi]
            return funcs, entity_dofs
        moment_degree = self.degree - 1 if self.family == "RT" else self.degree
        for fi, facet in enumerate(ref_facets(d)):
            verts = [sp.Matrix(ref_vertices(d)[i]) for i in facet]
            pt = verts[0] + sum(((verts[j + 1] - verts[0]) * s[j] for j in range(fdim)), sp.zeros(d, 1))
            n = facet_normal(d, facet)
            sub = dict(zip(X[:d], list(pt)))
            for q in _facet_lagrange(fdim, moment_degree, s):
                def ell(v, sub=sub, n=n, q=q):
                    vn = sum(v[c].subs(sub, simultaneous=True) * n[c] for c in range(d))
                    return _facet_integral(sp.expand(vn * q), fdim, s)
                entity_dofs[d - 1][fi].append(len(funcs))
                funcs.append(ell)
        if self.family == "RT" and self.degree == 2:
            for c in range(d):
                entity_dofs[d][0].append(len(funcs))
                funcs.append(lambda v, c=c: _cell_integral(v[c], d))
        return funcs, entity_dofs

    def _build(self):
        span = self._span()
        funcs, self.entity_dofs = self._functionals()
        n = len(span)
        assert len(funcs) == n, (self.family, self.dim, self.degree, len(funcs), n)
        D = sp.Matrix(n, n, lambda i, j: funcs[i](span[j]))
        C = D.inv()  # phi_j = sum_k C[k, j] span_k
        self.basis = []
        for j in range(n):
            v = sp.zeros(self.value_size, 1)
            for k in range(n):
                if C[k, j] != 0:
                    v += C[k, j] * span[k]
            self.basis.append(sp.simplify(v))
        self.functionals = funcs
        self.ndofs = n
        xs = X[: self.dim]
        self._f = sp.lambdify(xs, [list(b) for b in self.basis], "numpy")
        grads = [[[sp.diff(b[c], xs[k]) for c in range(self.value_size)] for b in self.basis] for k in range(self.dim)]
        self._g = sp.lambdify(xs, grads, "numpy")

    # -- tabulation -----------------------------------------------------------
    def tabulate(self, points):
        """Return (1+dim, nq, ndofs, value_size) like basix ``tabulate(1, pts)``."""
        pts = np.asarray(points, dtype=np.float64)
        nq = pts.shape[0]
        cols = [pts[:, k] for k in range(self.dim)]
        out = np.zeros((1 + self.dim, nq, self.ndofs, self.value_size))
        vals = self._f(*cols)
        grads = self._g(*cols)
        for j in range(self.ndofs):
            for c in range(self.value_size):
                out[0, :, j, c] = np.broadcast_to(np.asarray(vals[j][c], dtype=np.float64), (nq,))
                for k in range(self.dim):
                    out[1 + k, :, j, c] = np.broadcast_to(np.asarray(grads[k][j][c], dtype=np.float64), (nq,))
        return out

    @property
    def num_entity_dofs(self):
        return [[len(e) for e in ed] for ed in self.entity_dofs]

    @property
    def transformations_are_identity(self):
        # basix: vector H(div) elements carry facet orientation; Lagrange P0/P1 do not.
        return self.family == "P"

    def entity_transformations(self):
        """Base transformations per sub-entity type (restated, see module doc)."""
        if self.family == "P":
            return {}
        moment_degree = self.degree - 1 if self.family == "RT" else self.degree
        out = {}
        if self.dim == 2:
            if moment_degree == 0:
                out["interval"] = [np.array([[-1.0]])]
            else:
                out["interval"] = [-np.array([[0.0, 1.0], [1.0, 0.0]])]
        else:
            if moment_degree == 0:
                out["triangle"] = [np.array([[1.0]]), np.array([[-1.0]])]
            else:
                rot = np.array([[0.0, 0.0, 1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
                ref = -np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 1.0], [0.0, 1.0, 0.0]])
                out["triangle"] = [rot, ref]
            out["interval"] = [np.zeros((0, 0))]
        return out


@functools.lru_cache(maxsize=None)
def create_element(family, dim, degree):
    """Cached constructor. ``family`` in {"P", "RT", "BDM"}."""
    return ReferenceElement(family, dim, degree)
