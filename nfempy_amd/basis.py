"""Reference finite elements and their tables (host side, FP64).

Replaces the basix calls of the reference (``basix.create_element(family, cell,
k, lagrange_variant=gll_centroid, discontinuous)`` in
``src/basis/finite_element.py:85-91`` and ``src/spaces/dof_map.py:16-22``) for
the families of ``src/basis/element_family.py:4-12``: Lagrange P0/P1, Raviart-Thomas
RT1/RT2 and Brezzi-Douglas-Marini BDM1 on triangles and tetrahedra, and Nedelec of the
first kind N1E (H(curl); the reference's "N2E" name maps to it too) of degree 1 on
triangles and tetrahedra and degree 2 on triangles.

Each element is built numerically from its Ciarlet definition: a monomial span
and degree-of-freedom functionals (facet normal moments against Lagrange
P_{k-1} (RT) / P_k (BDM) with the non-normalised facet normal of basix's
sub-entity vertex order; RT2 interior moments against vector P0); the dual basis
is ``span @ inv(D)`` with ``D[i, j] = l_i(span_j)`` evaluated by quadrature. N1E:
span (P_{k-1})^d plus the homogeneous degree-k vector polynomials p with p . x = 0;
functionals the tangential edge moments against Lagrange P_{k-1} with the
non-normalised edge tangent v1 - v0 of basix's sub-entity vertex order (edges of the
tetrahedron (2,3), (1,3), (1,2), (0,3), (0,2), (0,1)), and for degree 2 on triangles
the interior moments against vector P0. Its push-forward is the covariant Piola map
phi = J^-T phihat (``covariant_piola``).

``entity_transformations`` restates basix's base transformations for these
moment spaces (P0 moments: reflection = -1, rotation = 1; P1 moments: a cyclic
vertex permutation for the rotation and a vertex swap with a sign flip for the
reflection).
"""

from __future__ import annotations

import functools
import itertools

import numpy as np

from .quadrature import make_quadrature

REF_VERTICES = {
    1: np.array([[0.0], [1.0]]),
    2: np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]]),
    3: np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]]),
}
FACETS = {
    2: [(1, 2), (0, 2), (0, 1)],
    3: [(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)],
}
EDGES = {
    2: [(1, 2), (0, 2), (0, 1)],
    3: [(2, 3), (1, 3), (1, 2), (0, 3), (0, 2), (0, 1)],
}
N_SUB = {2: (3, 3, 1), 3: (4, 6, 4, 1)}


def _exponents(dim, degree, homogeneous=False):
    out = []
    for total in range(degree + 1):
        if homogeneous and total != degree:
            continue
        for p in itertools.product(range(total + 1), repeat=dim):
            if sum(p) == total:
                out.append(p)
    return out


def _eval_mono(exps, pts):
    pts = np.asarray(pts, dtype=np.float64)
    out = np.ones((pts.shape[0], len(exps)))
    for j, e in enumerate(exps):
        for d, p in enumerate(e):
            if p:
                out[:, j] *= pts[:, d] ** p
    return out


def _eval_mono_grad(exps, pts):
    pts = np.asarray(pts, dtype=np.float64)
    dim = pts.shape[1]
    out = np.zeros((dim, pts.shape[0], len(exps)))
    for k in range(dim):
        for j, e in enumerate(exps):
            if e[k] == 0:
                continue
            v = np.full(pts.shape[0], float(e[k]))
            for d, p in enumerate(e):
                q = p - 1 if d == k else p
                if q:
                    v = v * pts[:, d] ** q
            out[k, :, j] = v
    return out


def _lagrange_facet(fdim, degree, s):
    """Nodal Lagrange P_degree (vertex nodes) on the reference facet at points s."""
    if degree == 0:
        return np.ones((s.shape[0], 1))
    if degree == 1:
        return np.concatenate([1.0 - s.sum(axis=1, keepdims=True), s], axis=1)
    raise NotImplementedError(degree)


def facet_normal(dim, facet):
    v = REF_VERTICES[dim][list(facet)]
    if dim == 2:
        t = v[1] - v[0]
        return np.array([-t[1], t[0]])
    return np.cross(v[1] - v[0], v[2] - v[0])


class Element:
    """A reference element: ``tabulate`` returns (1+dim, nq, ndofs, value_size)."""

    def __init__(self, family: str, dim: int, degree: int, discontinuous: bool = False):
        self.family = family
        self.dim = dim
        self.degree = degree
        self.discontinuous = discontinuous
        self.value_size = 1 if family == "P" else dim
        self._build()

    def _span(self):
        d = self.dim
        if self.family == "P":
            exps = _exponents(d, self.degree)
            coeffs = [np.eye(len(exps))[j][None, :] for j in range(len(exps))]
            return exps, coeffs
        if self.family == "N1E":
            return self._span_nedelec()
        top = self.degree
        exps = _exponents(d, top)
        idx = {e: i for i, e in enumerate(exps)}
        coeffs = []
        base_deg = self.degree if self.family == "BDM" else self.degree - 1
        for c in range(d):
            for e in _exponents(d, base_deg):
                m = np.zeros((d, len(exps)))
                m[c, idx[e]] = 1.0
                coeffs.append(m)
        if self.family == "RT":
            for e in _exponents(d, self.degree - 1, homogeneous=True):
                m = np.zeros((d, len(exps)))
                for c in range(d):
                    ee = list(e)
                    ee[c] += 1
                    m[c, idx[tuple(ee)]] = 1.0
                coeffs.append(m)
        return exps, coeffs

    def _span_nedelec(self):
        """(P_{k-1})^d + {p in (homogeneous P_k)^d : p . x = 0} (a basis of the second part
        from the null space of p -> p . x)."""
        d, k = self.dim, self.degree
        exps = _exponents(d, k)
        idx = {e: i for i, e in enumerate(exps)}
        coeffs = []
        for c in range(d):
            for e in _exponents(d, k - 1):
                m = np.zeros((d, len(exps)))
                m[c, idx[e]] = 1.0
                coeffs.append(m)
        hk = _exponents(d, k, homogeneous=True)
        hk1 = {e: i for i, e in enumerate(_exponents(d, k + 1, homogeneous=True))}
        M = np.zeros((len(hk1), d * len(hk)))
        for c in range(d):
            for j, e in enumerate(hk):
                ee = list(e)
                ee[c] += 1
                M[hk1[tuple(ee)], c * len(hk) + j] = 1.0
        _, sv, vt = np.linalg.svd(M)
        rank = int(np.sum(sv > 1e-12))
        for v in vt[rank:]:
            m = np.zeros((d, len(exps)))
            for c in range(d):
                for j, e in enumerate(hk):
                    m[c, idx[e]] = v[c * len(hk) + j]
            coeffs.append(m)
        return exps, coeffs

    def _build(self):
        d = self.dim
        exps, span = self._span()
        self._exps = exps
        n = len(span)
        rows = []  # each functional as a linear map on (vs, nmono) coefficients
        ent = [[[] for _ in range(k)] for k in N_SUB[d]]
        if self.family == "P":
            if self.degree == 0:
                pts, wts = make_quadrature(d, 2)
                vals = _eval_mono(exps, pts)
                rows.append(((wts @ vals) / wts.sum())[None, :])
                ent[d][0] = [0]
            else:
                vals = _eval_mono(exps, REF_VERTICES[d])
                for vi in range(d + 1):
                    rows.append(vals[vi][None, :])
                    ent[0][vi] = [vi]
        elif self.family == "N1E":
            if d == 3 and self.degree > 1:
                raise NotImplementedError("N1E of degree > 1 on tetrahedra (face moments)")
            if self.degree > 2:
                raise NotImplementedError("N1E of degree > 2")
            s, ws = make_quadrature(1, 2 * self.degree + 2)
            q = _lagrange_facet(1, self.degree - 1, s)
            for ei, edge in enumerate(EDGES[d]):
                V = REF_VERTICES[d][list(edge)]
                x = V[0] + s @ (V[1:] - V[0])
                vals = _eval_mono(exps, x)
                tan = V[1] - V[0]
                for j in range(q.shape[1]):
                    wv = ws * q[:, j]
                    r = np.zeros((d, len(exps)))
                    for c in range(d):
                        r[c] = tan[c] * (wv @ vals)
                    ent[1][ei].append(len(rows))
                    rows.append(r)
            if self.degree == 2:
                pts, wts = make_quadrature(d, 2 * self.degree + 2)
                vals = _eval_mono(exps, pts)
                for c in range(d):
                    r = np.zeros((d, len(exps)))
                    r[c] = wts @ vals
                    ent[d][0].append(len(rows))
                    rows.append(r)
        else:
            mdeg = self.degree - 1 if self.family == "RT" else self.degree
            s, ws = make_quadrature(d - 1, 2 * self.degree + 2)
            q = _lagrange_facet(d - 1, mdeg, s)
            for fi, facet in enumerate(FACETS[d]):
                V = REF_VERTICES[d][list(facet)]
                x = V[0] + s @ (V[1:] - V[0])
                vals = _eval_mono(exps, x)
                nrm = facet_normal(d, facet)
                for j in range(q.shape[1]):
                    wv = ws * q[:, j]
                    r = np.zeros((d, len(exps)))
                    for c in range(d):
                        r[c] = nrm[c] * (wv @ vals)
                    ent[d - 1][fi].append(len(rows))
                    rows.append(r)
            if self.family == "RT" and self.degree == 2:
                pts, wts = make_quadrature(d, 2 * self.degree + 2)
                vals = _eval_mono(exps, pts)
                for c in range(d):
                    r = np.zeros((d, len(exps)))
                    r[c] = wts @ vals
                    ent[d][0].append(len(rows))
                    rows.append(r)
        assert len(rows) == n, (self.family, d, self.degree, len(rows), n)
        D = np.array([[np.sum(rows[i] * span[j]) for j in range(n)] for i in range(n)])
        C = np.linalg.inv(D)
        self.coeffs = np.einsum("kj,kcm->jcm", C, np.array(span))  # (ndofs, vs, nmono)
        self.ndofs = n
        if self.discontinuous:
            ent = [[[] for _ in e] for e in ent]
            ent[d][0] = list(range(n))
        self.entity_dofs = ent

    @property
    def num_entity_dofs(self):
        return [[len(e) for e in ed] for ed in self.entity_dofs]

    @property
    def dof_transformations_are_identity(self):
        return self.family == "P"

    def tabulate(self, points):
        pts = np.asarray(points, dtype=np.float64)
        vals = _eval_mono(self._exps, pts)
        grads = _eval_mono_grad(self._exps, pts)
        out = np.zeros((1 + self.dim, pts.shape[0], self.ndofs, self.value_size))
        out[0] = np.einsum("qm,jcm->qjc", vals, self.coeffs)
        out[1:] = np.einsum("kqm,jcm->kqjc", grads, self.coeffs)
        return out

    def facet_moment_degree(self):
        return self.degree - 1 if self.family in ("RT", "N1E") else self.degree

    def entity_transformations(self):
        if self.family == "P":
            return {}
        mdeg = self.facet_moment_degree()
        if self.family == "N1E":
            # tangential edge moments: a reversed edge flips the tangent (and, for P1 moments,
            # swaps the two vertex functions)
            iv = np.array([[-1.0]]) if mdeg == 0 else np.array([[0.0, -1.0], [-1.0, 0.0]])
            if self.dim == 2:
                return {"interval": [iv]}
            return {"interval": [iv], "triangle": [np.zeros((0, 0)), np.zeros((0, 0))]}
        if self.dim == 2:
            if mdeg == 0:
                return {"interval": [np.array([[-1.0]])]}
            return {"interval": [np.array([[0.0, -1.0], [-1.0, 0.0]])]}
        if mdeg == 0:
            tri = [np.array([[1.0]]), np.array([[-1.0]])]
        else:
            rot = np.zeros((3, 3))
            rot[0, 2] = rot[1, 0] = rot[2, 1] = 1.0
            ref = np.zeros((3, 3))
            ref[0, 0] = ref[1, 2] = ref[2, 1] = -1.0
            tri = [rot, ref]
        return {"triangle": tri, "interval": [np.zeros((0, 0))]}


@functools.lru_cache(maxsize=None)
def create_element(family: str, dim: int, degree: int, discontinuous: bool = False) -> Element:
    fam = {"Lagrange": "P", "P": "P", "RT": "RT", "BDM": "BDM", "N1E": "N1E", "N2E": "N1E"}[family]
    return Element(fam, dim, degree, discontinuous)


def covariant_piola(phi_ref, jac):
    """H(curl) push-forward phi = J^-T phihat: phi_ref (..., nb, dim) reference values at
    points of one element, jac (dim, dim) its affine Jacobian."""
    return np.einsum("dc,...bd->...bc", np.linalg.inv(jac), phi_ref)


def curl_2d(grad_ref, jac):
    """Scalar curl d phi_y / dx - d phi_x / dy of covariantly mapped 2-D fields from the
    reference gradients grad_ref (2, nq, nb, 2): curl phi = curl phihat / det J."""
    return (grad_ref[0, :, :, 1] - grad_ref[1, :, :, 0]) / np.linalg.det(jac)


_POS_TO_ROT_REF = np.array([[0, 0], [0, 1], [2, 1], [1, 0], [2, 0], [1, 1]])


def facet_code_transform(element: Element, code: int) -> np.ndarray:
    """Signed-permutation matrix T of one facet for an orientation code.

    code bit 3 set: transformation applies, rot = bits 0-1, ref = bit 2
    (T = reflect^ref @ rotate^rot, ``src/basis/permute_and_transform.py:117-127``);
    bit 3 clear: identity.
    """
    tr = element.entity_transformations()
    if element.dim == 2:
        n = len(element.entity_dofs[1][0])
        if not code & 8:
            return np.eye(n)
        return tr["interval"][0] if code & 4 else np.eye(n)
    rot_t, ref_t = tr["triangle"]
    n = rot_t.shape[0]
    if not code & 8:
        return np.eye(n)
    t = np.eye(n)
    for _ in range(code & 3):
        t = rot_t @ t
    for _ in range((code >> 2) & 1):
        t = ref_t @ t
    return t
