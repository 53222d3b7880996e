"""Batched assembly and solve of many independent mixed Darcy systems.

The benchmark's batched configuration: hundreds of unit-cube tetrahedral Darcy
instances of uneven size (n^3 x 6 tets, n drawn per instance) and degree
(RT1-P0 or RT2-P1), log-normal permeability per instance, assembled by ONE
persistent launch (``nfx_assemble_darcy_batched``):

- instances with the same (n, k) share their mesh arrays, pattern and row data
  (one ``DarcySystem`` per distinct (n, k)); coefficients, values and residuals
  are per instance;
- work is cut into tiles of CSR rows, sorted by cost (rows x local DoFs x
  quadrature points) and walked in a static order by a fixed grid, so the
  heaviest tiles start first and no atomics are needed;
- the tile size and grid are chosen by ``autotune`` (timed on the batch itself);
- ``capture`` records the whole batched assembly in a CUDA/HIP graph (kernels
  only, no memsets) for replay.

Solves: ``solve_batched`` runs the batched MINRES (one set of launches per iteration for
all instances, per-instance recurrences and FP64 refinement to 1e-8); ``solve`` solves one
instance on its own.
"""

from __future__ import annotations

import ctypes
import time

import numpy as np

from . import _native as nat
from .device import CoefficientArg, DarcySystem, check_coefficient, torch
from .fields import grf_device
from .mesh import Mesh, unit_cube
from .spaces import ProductSpace


class _Item(ctypes.Structure):
    _fields_ = [("k", ctypes.c_int), ("n_rows", ctypes.c_int), ("ks_e", ctypes.c_int), ("ks_q", ctypes.c_int),
                ("fs_e", ctypes.c_int), ("fs_q", ctypes.c_int),
                ("inc", ctypes.c_void_p), ("pos", ctypes.c_void_p), ("rowptr", ctypes.c_void_p),
                ("signs", ctypes.c_void_p), ("cells", ctypes.c_void_p), ("coords", ctypes.c_void_p),
                ("tab", ctypes.c_void_p), ("kinv", ctypes.c_void_p), ("f", ctypes.c_void_p),
                ("vals", ctypes.c_void_p), ("rhs", ctypes.c_void_p), ("geo", ctypes.c_void_p)]


def darcy_space(n, k):
    mesh = Mesh.from_blocks(3, *unit_cube(n))
    ps = ProductSpace({"u": (3, 1, "RT", k, mesh), "p": (3, 1, "Lagrange", k - 1, mesh)})
    ps.make_subspaces_discontinuous({"p": True})
    ps.build_structures({"u": [1], "p": [1]}, {"u": [2, 3, 4, 5, 6, 7]})
    return ps


def _cold_system(key):
    """Host half of one batch system (runs in a worker process, no device work): the
    product space of (n, k) and its CSR pattern."""
    from .pattern import build_pattern

    ps = darcy_space(*key)
    rd, rs = ps.reference_destinations()
    return key, ps, build_pattern(ps.n_dof, rd, rs)


def _cold_work(key):
    """Relative host cost of building system (n, k): cells times the RT2 factor."""
    return key[0] ** 3 * (6 if key[1] == 2 else 1)


# below this total (four 16^3 RT2-P1 systems) the systems are built in this process: the
# workers' start-up (a fresh interpreter and its imports each) would cost more
COLD_POOL_MIN_WORK = 4 * _cold_work((16, 2))


def cold_workers(keys):
    """Worker processes for the host half of a batch's distinct systems: one per system, at
    most 16 and at most the host's cores; 1 (build them in this process) for a small batch."""
    import os

    if sum(_cold_work(k) for k in keys) < COLD_POOL_MIN_WORK:
        return 1
    return max(1, min(16, os.cpu_count() or 1, len(keys)))


RT2_SPAN_PER_ROW = 33  # staging entries per tile row of the compact RT2 launch (kRt2SpanPerRow)


def span_tiles(rowptr, tile_rows, span_per_row=RT2_SPAN_PER_ROW):
    """Greedy cut of rows [0, n) into tiles of at most ``tile_rows`` rows whose CSR span
    (rowptr[end] - rowptr[start]) is at most ``tile_rows * span_per_row``: (starts, counts)."""
    rp = np.asarray(rowptr, dtype=np.int64)
    n, cap = len(rp) - 1, tile_rows * span_per_row
    starts, counts = [], []
    r0 = 0
    while r0 < n:
        end = min(int(np.searchsorted(rp, rp[r0] + cap, side="right")) - 1, r0 + tile_rows, n)
        if end <= r0:
            raise ValueError("a CSR row is longer than the compact staging slice")
        starts.append(r0)
        counts.append(end - r0)
        r0 = end
    return np.asarray(starts, np.int64), np.asarray(counts, np.int64)


class DarcyBatch:
    def __init__(self, specs, seed=0, ell=0.1, sigma=1.0, device="cuda", log=None, split=True, workers=None):
        """specs: list of (n, k) per instance; log: optional progress callable; split: one
        launch per degree with a degree-specialised kernel (else one launch over both);
        workers: processes that build the distinct systems' spaces and patterns (host work,
        largest first) while this process uploads each one as it arrives (default
        ``cold_workers``; 1 builds them here)."""
        t = torch()
        if nat.call("nfx_darcy_batch_item_size") != ctypes.sizeof(_Item):
            raise RuntimeError("batch descriptor layout differs between host and library")
        self.device = device
        self.specs = [(int(n), int(k)) for n, k in specs]
        keys = sorted(set(self.specs))
        workers = cold_workers(keys) if workers is None else max(1, int(workers))
        built = {}

        def add(key, ps, pattern=None):
            built[key] = DarcySystem(ps, device=device, pattern=pattern)
            if log is not None:
                log(f"batch system {len(built)}/{len(keys)}: n={key[0]} k={key[1]} E={built[key].E} "
                    f"nnz={built[key].nnz}")

        if workers > 1 and len(keys) > 1:
            import concurrent.futures
            import multiprocessing

            # spawn: fresh interpreters that never touch the device; largest systems first
            order = sorted(keys, key=lambda kk: -_cold_work(kk))
            with concurrent.futures.ProcessPoolExecutor(max_workers=workers,
                                                        mp_context=multiprocessing.get_context("spawn")) as ex:
                futs = [ex.submit(_cold_system, key) for key in order]
                for fut in concurrent.futures.as_completed(futs):
                    key, ps, pattern = fut.result()
                    add(key, ps, pattern)
        else:
            for key in keys:
                add(key, darcy_space(*key))
        self.systems = {key: built[key] for key in keys}
        self.instances = []
        for i, (n, k) in enumerate(self.specs):
            sysd = self.systems[(n, k)]
            cent = sysd.coords64[sysd.cells.long()].mean(dim=1)
            logk = grf_device(cent, seed=seed + i, ell=ell, sigma=sigma)
            kinv = CoefficientArg(t.exp(-logk).contiguous(), 1, 0)
            f = CoefficientArg(t.ones(1, dtype=t.float64, device=device), 0, 0)
            check_coefficient(kinv, sysd.E, sysd.nq, "kappa")
            check_coefficient(f, sysd.E, sysd.nq, "rhs")
            self.instances.append({
                "system": sysd, "kinv": kinv, "f": f, "logk": logk,
                "vals": t.zeros(sysd.nnz, dtype=t.float32, device=device),
                "rhs": t.zeros(sysd.n, dtype=t.float32, device=device),
            })
        self.n_elements = sum(inst["system"].E for inst in self.instances)
        # launch groups: degree -> tiles, tile rows, grid (degree 0: both degrees in one launch)
        self.split = bool(split)
        degrees = sorted({k for _, k in self.specs}, reverse=True) if self.split else [0]
        # RT2 with per-element coefficients only: the compact launch (degree code 3: LDS holds
        # the reference tensors alone and a staging slice sized for the average row)
        rt2 = [inst for inst in self.instances if inst["system"].k == 2]
        if self.split and rt2 and all(inst["kinv"].sq == 0 and inst["f"].sq == 0 for inst in rt2):
            degrees = [3 if d == 2 else d for d in degrees]
        self.groups = {d: {"tile_rows": 128, "grid": None} for d in degrees}
        self._build_items()
        for d in degrees:
            self._build_tiles(d, 128)
        self.graph = None

    def _build_items(self):
        # one table per degree (the kernel keeps a table in LDS while the pointer repeats)
        tables = {}
        for s in self.systems.values():
            ref = tables.setdefault(s.k, s.tables32)
            if not torch().equal(ref, s.tables32):
                raise RuntimeError("instances of one degree must share their reference tables")
        if any(inst["system"].geo32 is None for inst in self.instances):
            raise RuntimeError("batched assembly needs the per-element metric of every instance (device systems)")
        arr = (_Item * len(self.instances))()
        for i, inst in enumerate(self.instances):
            s = inst["system"]
            kinv32, f32 = inst["kinv"].get(False), inst["f"].get(False)
            arr[i] = _Item(s.k, s.n, inst["kinv"].se, inst["kinv"].sq, inst["f"].se, inst["f"].sq,
                           nat.ptr(s.row_inc), nat.ptr(s.row_pos), nat.ptr(s.rowptr), nat.ptr(s.row_signs),
                           nat.ptr(s.cells), nat.ptr(s.coords), nat.ptr(tables[s.k]), nat.ptr(kinv32),
                           nat.ptr(f32), nat.ptr(inst["vals"]), nat.ptr(inst["rhs"]), nat.ptr(s.geo32))
        self._tables = tables
        raw = np.frombuffer(bytes(arr), dtype=np.uint8).copy()
        self._items = torch().from_numpy(raw).to(self.device)

    def _build_tiles(self, degree, tile_rows):
        """Tiles of the instances of `degree` (0: all; 3: RT2 compact), heaviest first.
        Degree 3 cuts each tile where its CSR span would pass tile_rows * RT2_SPAN_PER_ROW
        staging entries and carries the row count in bits 16.. of the item field."""
        tiles, cost = [], []
        if len(self.instances) >= 1 << 16:
            raise ValueError("at most 65535 instances per batch")
        for i, inst in enumerate(self.instances):
            s = inst["system"]
            if degree and s.k != (2 if degree == 3 else degree):
                continue
            c = s.nl * s.nq
            if degree != 3:
                for r0 in range(0, s.n, tile_rows):
                    tiles.append((i, r0))
                    cost.append(min(tile_rows, s.n - r0) * c)
                continue
            for r0, cnt in zip(*span_tiles(s.pattern.rowptr, tile_rows)):
                tiles.append((i | int(cnt) << 16, int(r0)))
                cost.append(int(cnt) * c)
        order = np.argsort(-np.asarray(cost), kind="stable")
        tl = np.asarray(tiles, dtype=np.int32).reshape(-1, 2)[order]
        g = self.groups[degree]
        g["n_tiles"] = len(tl)
        g["tiles"] = torch().from_numpy(np.ascontiguousarray(tl)).to(self.device)
        g["tile_rows"] = tile_rows

    @property
    def n_tiles(self):
        return sum(g["n_tiles"] for g in self.groups.values())

    def _launch(self, stream=None, only=None):
        for d, g in self.groups.items():
            if only is not None and d != only:
                continue
            grid = g["grid"] or min(g["n_tiles"], 2048)
            nat.call("nfx_assemble_darcy_batched", len(self.instances), nat.ptr(self._items), g["n_tiles"],
                     nat.ptr(g["tiles"]), g["tile_rows"], max(grid, 1), d, nat.stream_handle(stream))

    def assemble(self, stream=None):
        """All instances' FP32 CSR values and residuals (homogeneous Dirichlet data)."""
        if self.graph is not None:
            self.graph.replay()
        else:
            self._launch(stream)

    def autotune(self, reps=5):
        """Pick tile rows and grid of every launch group by timing the batch itself;
        returns the table of trials."""
        t = torch()
        cus = t.cuda.get_device_properties(t.cuda.current_device()).multi_processor_count
        trials = []
        for d, g in self.groups.items():
            best = None
            for tile in (64, 128, 256):
                self._build_tiles(d, tile)
                for mult in (1, 2, 4, 8):
                    g["grid"] = max(1, min(g["n_tiles"], cus * mult))
                    self._launch(only=d)
                    t.cuda.synchronize()
                    e0, e1 = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
                    e0.record()
                    for _ in range(reps):
                        self._launch(only=d)
                    e1.record()
                    t.cuda.synchronize()
                    ms = e0.elapsed_time(e1) / reps
                    trials.append({"degree": d, "tile_rows": tile, "grid": g["grid"], "ms": ms})
                    if best is None or ms < best[0]:
                        best = (ms, tile, g["grid"])
            self._build_tiles(d, best[1])
            g["grid"] = best[2]
            g["ms"] = best[0]
        self.graph = None
        return trials

    def capture(self):
        """Record the batched assembly in a graph (replayed by ``assemble``)."""
        t = torch()
        s = t.cuda.Stream()
        s.wait_stream(t.cuda.current_stream())
        with t.cuda.stream(s):
            self._launch()
        t.cuda.current_stream().wait_stream(s)
        t.cuda.synchronize()
        g = t.cuda.CUDAGraph()
        with t.cuda.graph(g):
            self._launch()
        self.graph = g
        return g

    def fp64_system(self, i):
        """Instance i's own FP64 values, residual and (sign, pinv) preconditioner (the shared
        per-(n, k) system's buffers are overwritten by the next instance)."""
        from .solvers import darcy_preconditioner

        inst = self.instances[i]
        s = inst["system"]
        s.assemble_elements(inst["kinv"], inst["f"], fp64=True)
        sign, pinv = darcy_preconditioner(s.rowptr, s.col, s.vals64, s.n, s.n_u)
        return {"rowptr": s.rowptr, "col": s.col, "vals32": inst["vals"], "vals64": s.vals64.clone(),
                "sign": sign, "pinv": pinv, "b": -s.rhs64.clone()}

    def solve_batched(self, indices=None, rtol=1e-8, inner_rtol=1e-5, max_inner=100000, max_outer=30,
                      pc="diag", theta=None):
        """Instances ``indices`` (default all) solved together by the batched MINRES: one set
        of launches per iteration, per-instance recurrences, stopping tests and FP64
        refinement. The FP32 values are those of the last batched assembly. ``pc``: "diag"
        (the block-diagonal preconditioner of every instance), "twogrid" (its pressure rows
        replaced by the batched two-grid Schur cycle, block-Jacobi smoothers) or "learned"
        (the same with the learned smoother on the RT1-P0 instances: ``theta``, default the
        committed weights). Returns (list of x, list of infos, seconds of the preconditioner
        set-up and the batched solve); the cold structure of the cycle is built once per
        set of instances and kept."""
        from .solvers import MinresBatch

        t = torch()
        idx = list(range(len(self.instances))) if indices is None else list(indices)
        systems = [self.fp64_system(i) for i in idx]
        b = t.cat([sy["b"] for sy in systems])
        mb = MinresBatch(systems)
        tg = None
        if pc != "diag":
            if pc not in ("twogrid", "learned"):
                raise ValueError(f"unknown batched preconditioner {pc!r}")
            from .batch_pc import BatchTwoGrid

            if pc == "learned" and theta is None:
                from .learned import default_weights_path, load_params

                theta = load_params(default_weights_path())
            key = (tuple(idx), pc)
            cache = self.__dict__.setdefault("_tg_cache", {})
            if key not in cache:
                cache.clear()
                cache[key] = BatchTwoGrid([self.instances[i]["system"] for i in idx], mb.offsets,
                                          theta=theta if pc == "learned" else None, device=self.device)
            tg = cache[key]
        t.cuda.synchronize()
        t0 = time.perf_counter()
        if tg is not None:
            mb.set_pc(tg.setup([sy["vals64"] for sy in systems]))
        x, infos = mb.solve(b, rtol=rtol, inner_rtol=inner_rtol, max_inner=max_inner, max_outer=max_outer)
        t.cuda.synchronize()
        dt = time.perf_counter() - t0
        mb.close()
        if tg is not None:
            for info in infos:
                info["preconditioner"] = pc
            self.last_tg = tg
        xs = [x[mb.offsets[j]:mb.offsets[j + 1]] for j in range(len(idx))]
        return xs, infos, dt

    def solve(self, i, rtol=1e-8):
        """FP64-refined solve of instance i (FP64 values assembled on demand)."""
        from .solvers import MinresIR, darcy_preconditioner

        inst = self.instances[i]
        s = inst["system"]
        s.assemble_elements(inst["kinv"], inst["f"], fp64=True)
        sign, pinv = darcy_preconditioner(s.rowptr, s.col, s.vals64, s.n, s.n_u)
        solver = MinresIR(s.rowptr, s.col, inst["vals"], s.vals64, sign, pinv)
        x, info = solver.solve(-s.rhs64, rtol=rtol, max_inner=100000)
        solver.close()
        return x, info
