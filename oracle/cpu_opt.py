"""All-core FP64 CPU baseline of the Darcy RT1-P0 path ("cpu-opt"): test and benchmark
infrastructure, never a product path. The element loop, the COO-to-CSR summation and a
diagonally preconditioned MINRES are restated in C with OpenMP (``oracle/cpu/darcy_omp.c``)
and timed on every host core beside the GPU.

Inputs come from the oracle's own space (``oracle.spaces.ProductSpace``): the DoF map of
``destination_indexes`` (``src/spaces/product_space.py`` via
``src/spaces/discrete_space.py:destination_indices``) and the RT1 face-DoF signs of the
reference's orientation rule (``src/basis/permute_and_transform.py:57-127``, restated in
``oracle.geometry.face_orientation_3d``), vectorised here: a face whose vertex tuple is the
k-th lexicographic permutation of its sorted tuple has (rotations, reflections) =
``_POSITION_TO_ROT_REF[k]``, tagged boundary faces 1 and 3 take one reflection, and for
the degree-0 face moments of RT1 the rotation is the identity and the reflection -1.
``tests/test_oracle_cpu_opt.py`` pins the whole chain against ``oracle.assemble.assemble_darcy``
(the per-element restatement) on structured and gmsh meshes, and the solve against scipy.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
import time

import numpy as np

from .geometry import _POSITION_TO_ROT_REF
from .mesh import TET_FACES

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpu", "darcy_omp.c")
LIB = os.path.join(HERE, "cpu", "liboracle_omp.so")
# x86-64-v3 (AVX2 + FMA) runs on every EPYC / Xeon host these GPUs sit in
CFLAGS = ["-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared"]

_lib = None


def build(verbose=False):
    cmd = ["gcc", *CFLAGS, SRC, "-o", LIB, "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(SRC):
            build()
        L = ctypes.CDLL(LIB)
        p, i64, d = ctypes.c_void_p, ctypes.c_int64, ctypes.c_double
        L.ocpu_threads.restype = ctypes.c_int
        L.ocpu_darcy_rt1p0.argtypes = [i64, ctypes.c_int] + [p] * 9
        L.ocpu_darcy_rt1p0.restype = None
        L.ocpu_segsum.argtypes = [i64, p, p, p, p]
        L.ocpu_segsum.restype = None
        L.ocpu_minres.argtypes = [i64, p, p, p, p, p, p, p, d, i64, p]
        L.ocpu_minres.restype = i64
        L.ocpu_residual.argtypes = [i64, p, p, p, p, p, p]
        L.ocpu_residual.restype = d
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def face_signs(tags, face_tagged):
    """(E, 4) sign of each RT1 face DoF: tags (E, 4) cell vertex tags in the cell's order,
    face_tagged (E, 4) whether the local face carries a boundary tag."""
    out = np.ones(tags.shape, np.float64)
    lut = np.array([_POSITION_TO_ROT_REF[k][1] for k in range(6)])
    for i, f in enumerate(TET_FACES):
        a, b, c = (tags[:, f[0]], tags[:, f[1]], tags[:, f[2]])
        # lexicographic index among the permutations of the sorted tuple: rank of the first
        # entry, then whether the remaining two are in descending order
        pos = 2 * ((a > b).astype(np.int64) + (a > c)) + (b > c)
        ref = lut[pos]
        if i in (1, 3):
            ref = np.where(face_tagged[:, i], 1, ref)
        out[:, i] = np.where(ref % 2 == 1, -1.0, 1.0)
    return out


def darcy_inputs(ps):
    """Geometry, DoF map and signs of an RT1-P0 oracle product space (FP64, untimed set-up)."""
    from .reference_elements import create_element

    u = ps.spaces["u"]
    m = u.mesh
    if u.dimension != 3 or u.element.family != "RT" or u.element.degree != 1:
        raise ValueError("the all-core baseline covers RT1-P0 on tetrahedra")
    cells = [m.cells[c] for c in u.element_ids]
    tags = np.array([c.node_tags for c in cells], dtype=np.int64)
    fids = np.array([c.sub_cells_ids[2] for c in cells], dtype=np.int64)
    tagged_face = np.zeros(int(fids.max()) + 1, bool)
    for fid in np.unique(fids):
        tagged_face[fid] = m.cells[fid].material_id is not None
    sign = face_signs(tags, tagged_face[fids])
    dest = np.stack([ps.destination_indexes(i) for i in range(len(cells))]).astype(np.int64)
    points, weights = ps.quadrature[3]
    tab = create_element("RT", 3, 1).tabulate(points)  # (4, nq, 4, 3)
    div = np.einsum("kqjk->qj", tab[1:])[0]  # constant over the element
    return {
        "X": np.ascontiguousarray(m.points[tags][:, :, :3], dtype=np.float64),
        "sign": sign,
        "dest": dest,
        "phi": np.ascontiguousarray(tab[0], dtype=np.float64),
        "div": np.ascontiguousarray(div, dtype=np.float64),
        "w": np.ascontiguousarray(weights, dtype=np.float64),
        "n": ps.n_dof,
    }


class DarcyAllCore:
    """Cold set-up (the COO -> CSR plan, once per mesh) and warm assembly (element loop +
    slot sums, all cores) of the RT1-P0 Darcy system, and the all-core MINRES solve."""

    def __init__(self, inp):
        from .assemble import coo_plan

        self.inp = inp
        E = len(inp["dest"])
        d = inp["dest"]
        t0 = time.perf_counter()
        rows = np.repeat(d, 5, axis=1).ravel()
        cols = np.tile(d, (1, 5)).ravel()
        self.indptr, self.indices, self.perm, self.segptr = coo_plan(inp["n"], rows, cols)
        self.plan_s = time.perf_counter() - t0
        self.E = E
        self.stage = np.empty(E * 25)
        self.rstage = np.empty(E * 5)
        self.data = np.empty(len(self.indices))
        self.rd = d.ravel()

    def assemble(self, kappa, f):
        L = lib()
        inp = self.inp
        kinv = np.ascontiguousarray(1.0 / np.asarray(kappa, np.float64))
        f = np.ascontiguousarray(np.asarray(f, np.float64))
        L.ocpu_darcy_rt1p0(self.E, len(inp["w"]), _p(inp["X"]), _p(inp["sign"]), _p(kinv), _p(f), _p(inp["phi"]),
                           _p(inp["div"]), _p(inp["w"]), _p(self.stage), _p(self.rstage))
        L.ocpu_segsum(len(self.indices), _p(self.segptr), _p(self.perm), _p(self.stage), _p(self.data))
        rg = np.bincount(self.rd, weights=self.rstage, minlength=inp["n"])
        return self.indptr, self.indices, self.data, rg

    def solve(self, data, b, n_u, rtol=1e-8, maxit=200000, refine=30):
        """MINRES on diag(1 | -1) A with the block-diagonal preconditioner
        |diag(M)|^-1 | |diag(B diag(M)^-1 B^T)|^-1 inside FP64 refinement on the true
        residual. Returns (x, iterations, relative residual)."""
        L = lib()
        n = len(b)
        rp, col = self.indptr, self.indices
        sgn = np.ones(n)
        sgn[n_u:] = -1.0
        rows = np.repeat(np.arange(n), np.diff(rp))
        diag = np.zeros(n)
        on = rows == col
        diag[rows[on]] = data[on]
        d = diag[:n_u]
        sel = (rows >= n_u) & (col < n_u)
        sd = np.abs(np.bincount(rows[sel] - n_u, weights=data[sel] ** 2 / d[col[sel]], minlength=n - n_u))
        pinv = np.concatenate([1.0 / np.abs(d), np.where(sd > 0, 1.0 / np.where(sd > 0, sd, 1.0), 1.0)])
        x = np.zeros(n)
        r = np.empty(n)
        dx = np.empty(n)
        work = np.empty(7 * n)
        b = np.ascontiguousarray(b, np.float64)
        bn = float(np.linalg.norm(b))
        its = 0
        res = 1.0
        for _ in range(refine):
            rn = L.ocpu_residual(n, _p(rp), _p(col), _p(data), _p(b), _p(x), _p(r))
            res = rn / bn
            if res <= rtol:
                break
            # the correction needs only the reduction still missing (the recurrence's
            # estimate is in the preconditioner's norm: half of it as a margin)
            inner = min(0.5, 0.5 * rtol / res)
            its += L.ocpu_minres(n, _p(rp), _p(col), _p(data), _p(sgn), _p(pinv), _p(r), _p(dx), inner, maxit,
                                 _p(work))
            x += dx
        return x, its, res


def threads():
    return int(lib().ocpu_threads())
