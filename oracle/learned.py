"""FP64 oracle of the learned Schur smoother and its two-grid preconditioner.

TEST INFRASTRUCTURE ONLY (see the package header). There is no reference
counterpart: the reference solves its Darcy systems with a direct LU
(``examples/md_elliptic_scalar/md_elliptic_scalar.py:351-364``). This module pins
the device implementation of the learned correction of the block preconditioner:

- ``schur_ell``: the P0 pressure Schur approximation ``S = B diag(M)^-1 B^T`` of an
  assembled RT1-P0 matrix ``[[M, -B^T], [B, 0]]`` in a 4-wide ELL layout (row p lists
  the cells across its faces in ascending face order);
- ``features`` / ``forward`` / ``backward``: the 4-layer message-passing network
  (width 64) that maps the graph of ``S`` to the entries of a symmetric sparse
  smoother ``W`` with the pattern of ``S``; ``forward(emulate_bf16=True)`` rounds
  weights and stored activations to bfloat16 exactly where the device kernels do;
- ``twogrid_loss``: mean over samples of ``||(I - Pi)(I - W S) e||_S^2 / ||e||_S^2``
  with ``Pi`` the S-orthogonal projector on piecewise constants over aggregates,
  and its gradient with respect to the entries of ``W``;
- ``twogrid_apply``: one symmetric two-grid cycle (pre-smooth W, exact coarse
  solve, post-smooth W);
- ``flux_sweeps``: the flux rows of the block preconditioner, damped Jacobi sweeps
  on the flux mass block ``M`` from a zero start (one sweep is the diagonal scaling
  ``omega diag(M)^-1``).

Pinned by: central finite differences of ``forward``+``twogrid_loss`` against
``backward`` (FP64, no rounding), ``schur_ell`` against a dense ``B diag(M)^-1 B^T``,
the two-grid cycle against its closed form on a small dense problem, and
``flux_sweeps`` against ``omega (2 - omega) diag(M)^-1`` for a diagonal ``M`` and
against ``M^-1`` in the limit of many sweeps (``tests/test_oracle_learned.py``).
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sps

F_IN = 8
WIDTH = 64
LAYERS = 4
OMEGA0 = 0.65


def param_layout():
    """Flat parameter vector layout: name -> (offset, shape)."""
    out, off = {}, 0
    for name, shape in [("W0", (F_IN, WIDTH)), ("b0", (WIDTH,))] + [
            item for l in range(LAYERS) for item in
            ((f"Wh{l}", (WIDTH, WIDTH)), (f"Wm{l}", (WIDTH, WIDTH)), (f"b{l + 1}", (WIDTH,)))] + [
            ("ud", (WIDTH,)), ("cd", (1,)), ("ue", (WIDTH,)), ("ce", (1,))]:
        out[name] = (off, shape)
        off += int(np.prod(shape))
    return out, off


def unpack(theta):
    lay, _ = param_layout()
    return {k: np.asarray(theta[o:o + int(np.prod(s))], dtype=np.float64).reshape(s) for k, (o, s) in lay.items()}


def bf16(x):
    """Round to bfloat16 (round to nearest, ties to even), returned as float64."""
    a = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((a + 0x7FFF + ((a >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def schur_ell(indptr, indices, data, n_u):
    """(nbr (E,4) int, sdiag (E,), soff (E,4)) of S = B diag(M)^-1 B^T."""
    A = sps.csr_matrix((data, indices, indptr))
    n = A.shape[0]
    B = A[n_u:, :n_u].tocsr()
    B.sort_indices()
    dM = A[:n_u, :n_u].diagonal()
    S = (B @ sps.diags(1.0 / dM) @ B.T).tocsr()
    E = n - n_u
    Bt = B.T.tocsr()
    nbr = -np.ones((E, 4), dtype=np.int64)
    soff = np.zeros((E, 4))
    for p in range(E):
        faces = B.indices[B.indptr[p]:B.indptr[p + 1]]
        assert len(faces) <= 4
        for j, f in enumerate(faces):
            cells = Bt.indices[Bt.indptr[f]:Bt.indptr[f + 1]]
            other = [c for c in cells if c != p]
            if other:
                nbr[p, j] = other[0]
                soff[p, j] = S[p, other[0]]
    return nbr, S.diagonal().copy(), soff


def coupling(nbr, sdiag, soff):
    """a_pj = -S_pq / sqrt(S_pp S_qq) (0 where there is no neighbour)."""
    has = nbr >= 0
    q = np.where(has, nbr, 0)
    return np.where(has, -soff / np.sqrt(sdiag[:, None] * sdiag[q]), 0.0)


def features(nbr, sdiag, soff):
    """(E, 8) node features, computed in FP64 and rounded to FP32."""
    has = nbr >= 0
    q = np.where(has, nbr, 0)
    a = coupling(nbr, sdiag, soff)
    cnt = has.sum(1)
    lr = np.where(has, np.abs(np.log(sdiag[q] / sdiag[:, None])), 0.0)
    x = np.zeros((len(sdiag), F_IN))
    x[:, 0] = a.sum(1)
    x[:, 1] = a.max(1)
    x[:, 2] = np.where(cnt > 0, np.where(has, a, np.inf).min(1), 0.0)
    x[:, 3] = (sdiag + soff.sum(1)) / sdiag
    x[:, 4] = np.where(cnt > 0, lr.sum(1) / np.maximum(cnt, 1), 0.0)
    x[:, 5] = lr.max(1)
    x[:, 6] = cnt / 4.0
    x[:, 7] = 1.0
    return x.astype(np.float32).astype(np.float64)


def forward(theta, nbr, sdiag, soff, emulate_bf16=True):
    """Network forward. Returns (Wd (E,), Wo (E,4), cache)."""
    P = unpack(theta)
    rnd = bf16 if emulate_bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    W = {k: (rnd(v) if k.startswith("W") or k in ("ud", "ue") else v) for k, v in P.items()}
    has = nbr >= 0
    q = np.where(has, nbr, 0)
    a = coupling(nbr, sdiag, soff)
    if emulate_bf16:
        a = a.astype(np.float32).astype(np.float64)
    x = features(nbr, sdiag, soff)
    z0 = x @ W["W0"] + W["b0"]
    h = rnd(np.maximum(z0, 0.0))
    cache = {"x": x, "z0": z0, "h": [h], "m": [], "z": [], "a": a, "q": q, "has": has, "W": W}
    for l in range(LAYERS):
        m = rnd(np.einsum("pj,pjc->pc", a, h[q]))
        z = h @ W[f"Wh{l}"] + m @ W[f"Wm{l}"] + W[f"b{l + 1}"]
        h = rnd(h + np.maximum(z, 0.0))
        cache["m"].append(m)
        cache["z"].append(z)
        cache["h"].append(h)
    od = h @ W["ud"] + W["cd"][0]
    s = h[:, None, :] + h[q]
    oe = a * (s @ W["ue"] + W["ce"][0])
    Wd = (OMEGA0 + od) / sdiag
    Wo = np.where(has, oe / np.sqrt(sdiag[:, None] * sdiag[q]), 0.0)
    cache.update(s=s, sdiag=sdiag)
    return Wd, Wo, cache


def backward(theta, cache, gamma_d, gamma_o):
    """Gradient of the loss w.r.t. theta, given the loss gradient w.r.t. the smoother
    entries: gamma_d[p] = dL/dW_pp, gamma_o[p, j] = dL/dW_pq + dL/dW_qp (q = nbr[p, j]).
    FP64 throughout on the cached (possibly rounded) activations."""
    lay, n_par = param_layout()
    W = cache["W"]
    a, q, has, sdiag = cache["a"], cache["q"], cache["has"], cache["sdiag"]
    h4 = cache["h"][-1]
    grad = {k: np.zeros(s) for k, (_, s) in lay.items()}
    gN = gamma_d / sdiag
    gE = np.where(has, gamma_o / np.sqrt(sdiag[:, None] * sdiag[q]), 0.0)
    grad["ud"] = gN @ h4
    grad["cd"][0] = gN.sum()
    ag = a * gE
    grad["ue"] = 0.5 * np.einsum("pj,pjc->c", ag, cache["s"])
    grad["ce"][0] = 0.5 * ag.sum()
    dh = np.outer(gN, W["ud"]) + np.outer(ag.sum(1), W["ue"])
    for l in reversed(range(LAYERS)):
        h, m, z = cache["h"][l], cache["m"][l], cache["z"][l]
        dz = dh * (z > 0)
        grad[f"Wh{l}"] = h.T @ dz
        grad[f"Wm{l}"] = m.T @ dz
        grad[f"b{l + 1}"] = dz.sum(0)
        dm = dz @ W[f"Wm{l}"].T
        dh_prev = dh + dz @ W[f"Wh{l}"].T
        # m_p = sum_j a_pj h_q  ->  dh_q += a_pj dm_p; with a symmetric this is a gather
        np.add.at(dh_prev, q[has], (a[:, :, None] * dm[:, None, :])[has])
        dh = dh_prev
    dz0 = dh * (cache["z0"] > 0)
    grad["W0"] = cache["x"].T @ dz0
    grad["b0"] = dz0.sum(0)
    out = np.zeros(n_par)
    for k, (o, s) in lay.items():
        out[o:o + int(np.prod(s))] = grad[k].ravel()
    return out


def ell_mul(nbr, d, off, v):
    """(diag + ELL off-diagonal) times v; v may be (E,) or (E, ns)."""
    has = nbr >= 0
    q = np.where(has, nbr, 0)
    vv = v if v.ndim == 2 else v[:, None]
    out = d[:, None] * vv + np.einsum("pj,pjs->ps", np.where(has, off, 0.0), vv[q])
    return out if v.ndim == 2 else out[:, 0]


def aggregate_cells(centroids, h):
    """Aggregate id per cell: boxes of side h anchored h/16 below the smallest centroid
    coordinate, numbered by ascending (z, y, x) box index among the non-empty boxes."""
    lo = centroids.min(axis=0)
    key = np.floor((centroids - lo) / h + 1.0 / 16.0).astype(np.int64)
    nb = key.max(axis=0) + 1
    lin = key[:, 0] + nb[0] * (key[:, 1] + nb[1] * key[:, 2])
    u, inv = np.unique(lin, return_inverse=True)
    return inv.astype(np.int64), len(u)


def coarse_matrix(nbr, sdiag, soff, agg, n_agg):
    P = sps.csr_matrix((np.ones(len(agg)), (np.arange(len(agg)), agg)), shape=(len(agg), n_agg))
    S = ell_to_csr(nbr, sdiag, soff)
    return (P.T @ S @ P).toarray()


def ell_to_csr(nbr, d, off):
    E = len(d)
    has = nbr >= 0
    r = np.r_[np.arange(E), np.repeat(np.arange(E), 4)[has.ravel()]]
    c = np.r_[np.arange(E), nbr.ravel()[has.ravel()]]
    v = np.r_[d, off.ravel()[has.ravel()]]
    return sps.csr_matrix((v, (r, c)), shape=(E, E))


def twogrid_loss(Wd, Wo, nbr, sdiag, soff, agg, n_agg, e):
    """(loss, gamma_d, gamma_o) for samples e (E, ns)."""
    ns = e.shape[1]
    has = nbr >= 0
    q = np.where(has, nbr, 0)
    Ac = coarse_matrix(nbr, sdiag, soff, agg, n_agg)
    Se = ell_mul(nbr, sdiag, soff, e)
    en = (e * Se).sum(0)
    v = e - ell_mul(nbr, Wd, Wo, Se)
    Sv = ell_mul(nbr, sdiag, soff, v)
    c = np.zeros((n_agg, ns))
    np.add.at(c, agg, Sv)
    y = np.linalg.solve(Ac, c)
    u = v - y[agg]
    Su = ell_mul(nbr, sdiag, soff, u)
    loss = float(np.mean((u * Su).sum(0) / en))
    w = -2.0 / (ns * en)
    gamma_d = (Su * Se) @ w
    gamma_o = np.where(has, np.einsum("ps,pjs->pj", Su * w, Se[q]) + np.einsum("ps,pjs->pj", Se * w, Su[q]), 0.0)
    return loss, gamma_d, gamma_o


def twogrid_apply(Wd, Wo, nbr, sdiag, soff, agg, n_agg, r):
    """z = TG(r): x1 = W r; x2 = x1 + P Ac^-1 P^T (r - S x1); z = x2 + W (r - S x2)."""
    Ac = coarse_matrix(nbr, sdiag, soff, agg, n_agg)
    x1 = ell_mul(nbr, Wd, Wo, r)
    t = r - ell_mul(nbr, sdiag, soff, x1)
    c = np.zeros(n_agg)
    np.add.at(c, agg, t)
    x2 = x1 + np.linalg.solve(Ac, c)[agg]
    t2 = r - ell_mul(nbr, sdiag, soff, x2)
    return x2 + ell_mul(nbr, Wd, Wo, t2)


def flux_sweeps(indptr, indices, data, n_u, r, omega, sweeps=2):
    """``sweeps`` damped Jacobi sweeps on ``M z = r`` from z = 0, ``M`` the rows and
    columns < n_u of the CSR matrix: z <- z + omega diag(M)^-1 (r - M z)."""
    A = sps.csr_matrix((data, indices, indptr))
    M = A[:n_u, :n_u].tocsr()
    d = M.diagonal()
    r = np.asarray(r, dtype=np.float64)
    z = np.zeros_like(r)
    for _ in range(sweeps):
        z = z + omega * (r - M @ z) / d
    return z
