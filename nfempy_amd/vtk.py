"""VTK output of device solutions.

``write_vtk_file(file_name, gmesh, fe_space, alpha, cell_centered=[])`` and
``write_vtk_file_with_exact_solution(file_name, gmesh, fe_space, functions, alpha,
cell_centered=[])`` keep the signatures of ``src/postprocess/solution_post_processor.py:106-242``.
Each field is written as ``<name>_h`` (and ``<name>_e`` for the exact solution):

- fields named in ``cell_centered``: one value per cell of the mesh dimension, the
  discrete field at the reference cell's centroid (``cell_centered_quatity``, ``:69-103``);
- the others: one value per mesh point, the mean over the field's elements around the
  point of the discrete field at that vertex (``node_average_quatity``, ``:11-66``).

Vector families (RT, BDM rows) give ``n_comp * dim`` values (the first ``dim`` components
of the Piola-mapped field), Lagrange fields ``n_comp`` values (the reference's
``alpha_star[:, 0:dim]``). The values are computed on the device in FP64
(``nfx_field_values``); the per-point mean is taken on the host in element order. The
reference writes through meshio (absent here); the writer below emits the legacy VTK
format (binary big-endian by default, as meshio does for ``.vtk``) or VTK XML (``.vtu``,
ASCII), chosen by the file extension.
"""

from __future__ import annotations

import os

import numpy as np

from . import _native as nat
from .device import to_dev, torch
from .pattern import row_codes

_VTK_TYPE = {1: 3, 2: 5, 3: 10}  # line, triangle, tetra
_REF_VERTICES = {1: np.array([[0.0], [1.0]]), 2: np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]]),
                 3: np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])}


def _device_alpha(alpha):
    t = torch()
    a = alpha if isinstance(alpha, t.Tensor) else t.from_numpy(np.asarray(alpha, dtype=np.float64))
    return a.to(device="cuda", dtype=t.float64).contiguous()


def field_values_at(fe_space, name, alpha, ref_points):
    """(E, n_points, n_data) values of field ``name`` at reference points of every element."""
    t = torch()
    sp = fe_space.discrete_spaces[name]
    dim = sp.dimension
    el = sp.element
    tab = el.tabulate(np.asarray(ref_points, dtype=np.float64))
    vector = sp.family != "P"
    nb = el.ndofs
    if vector:
        tables, nc, kind = tab[0, :, :, :dim], sp.n_comp, 0
    else:
        tables, nc, kind = tab[0, :, :, 0], min(sp.n_comp, dim), 1
    mesh = sp.mesh
    cells = mesh.entities[dim].nodes[sp.topology.cell_index]
    E, npt = len(cells), len(ref_points)
    shift = fe_space.strides[name] + fe_space.dof_shift
    codes = row_codes(sp.ref_dest + shift, sp.ref_sign).reshape(E, nb, sp.n_comp)[:, :, :nc]
    a = _device_alpha(alpha)
    if E and int((codes & nat.SLOT_INDEX_MASK).max()) >= a.numel():
        raise ValueError("alpha is shorter than the space's DoF range")
    if E and int(cells.max()) >= len(mesh.points):
        raise ValueError("element vertex outside the mesh points")
    dv = lambda x, dt: to_dev(np.ascontiguousarray(np.asarray(x).astype(dt)), device="cuda")
    # every device array stays referenced until the values are copied back: a temporary
    # would return its block to the caching allocator before the kernel reads it
    d_cells, d_coords = dv(cells, np.int32), dv(mesh.points, np.float64)
    d_tab, d_rows = dv(tables, np.float64), dv(codes, np.int32)
    if d_coords.numel() != 3 * len(mesh.points):
        raise ValueError("mesh points must have three coordinates")
    out = t.empty(E * npt * nc * (dim if vector else 1), dtype=t.float64, device="cuda")
    nat.call("nfx_field_values", kind, dim, nb, npt, nc, E, nat.ptr(d_cells), nat.ptr(d_coords), nat.ptr(d_tab),
             nat.ptr(d_rows), nat.ptr(a), nat.ptr(out), nat.stream_handle())
    vals = out.cpu().numpy().reshape(E, npt, -1)
    del d_cells, d_coords, d_tab, d_rows
    return vals


def _n_data(sp, dim):
    n = sp.n_comp * (dim if sp.family != "P" else 1)
    if sp.family == "P" and sp.n_comp > dim:
        raise ValueError("a Lagrange field with more components than the dimension has no VTK layout "
                         "(the reference keeps the first dim components)")
    return n


def node_average_values(fe_space, name, alpha, n_points):
    """(n_points, n_data): per mesh point, the mean over the field's elements that have the
    point as a vertex of the field at that vertex (zero at points of no element)."""
    sp = fe_space.discrete_spaces[name]
    dim = sp.dimension
    _n_data(sp, dim)
    vals = field_values_at(fe_space, name, alpha, _REF_VERTICES[dim])  # (E, dim+1, n_data)
    cells = sp.mesh.entities[dim].nodes[sp.topology.cell_index]
    acc = np.zeros((n_points, vals.shape[2]))
    cnt = np.zeros(n_points)
    flat_nodes = cells.reshape(-1)
    np.add.at(acc, flat_nodes, vals.reshape(-1, vals.shape[2]))
    np.add.at(cnt, flat_nodes, 1.0)
    has = cnt > 0
    acc[has] /= cnt[has, None]
    return acc


def cell_centered_values(fe_space, name, alpha):
    """(n_cells, n_data): the field at the reference centroid of each cell of the mesh
    dimension (cells outside the field: zero)."""
    sp = fe_space.discrete_spaces[name]
    dim = sp.dimension
    _n_data(sp, dim)
    c = _REF_VERTICES[dim].mean(axis=0, keepdims=True)
    vals = field_values_at(fe_space, name, alpha, c)[:, 0, :]
    out = np.zeros((sp.mesh.entities[dim].n, vals.shape[1]))
    out[sp.topology.cell_index] = vals
    return out


def _exact_at(f, X):
    return np.array([np.asarray(f(x[0], x[1], x[2]), dtype=np.float64).ravel() for x in X])


def _collect(gmesh, fe_space, alpha, cell_centered, functions=None):
    dim = gmesh.dimension
    pdata, cdata = {}, {}
    for name, sp in fe_space.discrete_spaces.items():
        if name in cell_centered:
            cdata[name + "_h"] = cell_centered_values(fe_space, name, alpha)
            if functions is not None:
                X = gmesh.points[gmesh.entities[dim].nodes].mean(axis=1)
                cdata[name + "_e"] = _exact_at(functions[name], X)
        else:
            pdata[name + "_h"] = node_average_values(fe_space, name, alpha, len(gmesh.points))
            if functions is not None:
                pdata[name + "_e"] = _exact_at(functions[name], gmesh.points)
    return pdata, cdata


def write_vtk_file(file_name, gmesh, fe_space, alpha, cell_centered=()):
    pdata, cdata = _collect(gmesh, fe_space, alpha, cell_centered)
    dim = gmesh.dimension
    write_unstructured(file_name, gmesh.points, gmesh.entities[dim].nodes, dim, pdata, cdata)


def write_vtk_file_with_exact_solution(file_name, gmesh, fe_space, functions, alpha, cell_centered=()):
    pdata, cdata = _collect(gmesh, fe_space, alpha, cell_centered, functions)
    dim = fe_space.dimension()
    tab = gmesh.entities[dim]
    name = list(fe_space.discrete_spaces)[-1]
    tags = [t for t in fe_space.fields_physical_tags[name] if t is not None]
    sel = np.isin(tab.material, tags)
    if not sel.all() and cdata:
        cdata = {k: v[sel] for k, v in cdata.items()}
    write_unstructured(file_name, gmesh.points, tab.nodes[sel], dim, pdata, cdata)


def write_unstructured(file_name, points, cells, cell_dim, point_data=None, cell_data=None, binary=True):
    """One-cell-type unstructured grid: legacy VTK (``.vtk``; binary or ASCII) or VTK XML
    (``.vtu``, ASCII). Data arrays are (n, k) FP64."""
    points = np.asarray(points, dtype=np.float64)
    if points.shape[1] < 3:
        points = np.hstack([points, np.zeros((len(points), 3 - points.shape[1]))])
    cells = np.asarray(cells, dtype=np.int64)
    point_data = {k: np.asarray(v, dtype=np.float64).reshape(len(points), -1) for k, v in (point_data or {}).items()}
    cell_data = {k: np.asarray(v, dtype=np.float64).reshape(len(cells), -1) for k, v in (cell_data or {}).items()}
    ctype = _VTK_TYPE[cell_dim]
    ext = os.path.splitext(file_name)[1].lower()
    if ext == ".vtu":
        _write_vtu(file_name, points, cells, ctype, point_data, cell_data)
    else:
        _write_legacy(file_name, points, cells, ctype, point_data, cell_data, binary)


def _write_legacy(path, points, cells, ctype, pdata, cdata, binary):
    nv = cells.shape[1]
    conn = np.hstack([np.full((len(cells), 1), nv, np.int64), cells]).astype(">i4")
    with open(path, "wb") as fh:
        w = lambda s: fh.write(s.encode("ascii"))
        w("# vtk DataFile Version 4.2\nnfempy_amd\n" + ("BINARY\n" if binary else "ASCII\n"))
        w("DATASET UNSTRUCTURED_GRID\n")

        def arr(a, dt, per_line):
            if binary:
                fh.write(np.ascontiguousarray(a, dtype=dt).tobytes())
                w("\n")
            else:
                flat = np.asarray(a).reshape(-1, per_line)
                fmt = "%.17g" if np.dtype(dt).kind == "f" else "%d"
                for row in flat:
                    w(" ".join(fmt % v for v in row) + "\n")

        w(f"POINTS {len(points)} double\n")
        arr(points, ">f8", 3)
        w(f"CELLS {len(cells)} {conn.size}\n")
        arr(conn, ">i4", nv + 1)
        w(f"CELL_TYPES {len(cells)}\n")
        arr(np.full(len(cells), ctype), ">i4", 1)
        for kw, data, n in (("POINT_DATA", pdata, len(points)), ("CELL_DATA", cdata, len(cells))):
            if not data:
                continue
            w(f"{kw} {n}\nFIELD FieldData {len(data)}\n")
            for k, v in data.items():
                w(f"{k} {v.shape[1]} {n} double\n")
                arr(v, ">f8", v.shape[1])


def _write_vtu(path, points, cells, ctype, pdata, cdata):
    nv = cells.shape[1]
    fmt = lambda a: " ".join("%.17g" % v for v in np.asarray(a, dtype=np.float64).ravel())
    fmti = lambda a: " ".join("%d" % v for v in np.asarray(a).ravel())
    parts = ['<?xml version="1.0"?>\n<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian">\n',
             f'<UnstructuredGrid>\n<Piece NumberOfPoints="{len(points)}" NumberOfCells="{len(cells)}">\n',
             '<Points>\n<DataArray type="Float64" NumberOfComponents="3" format="ascii">\n', fmt(points),
             "\n</DataArray>\n</Points>\n<Cells>\n",
             '<DataArray type="Int64" Name="connectivity" format="ascii">\n', fmti(cells), "\n</DataArray>\n",
             '<DataArray type="Int64" Name="offsets" format="ascii">\n', fmti(nv * np.arange(1, len(cells) + 1)),
             "\n</DataArray>\n", '<DataArray type="UInt8" Name="types" format="ascii">\n',
             fmti(np.full(len(cells), ctype)), "\n</DataArray>\n</Cells>\n"]
    for tag, data in (("PointData", pdata), ("CellData", cdata)):
        if not data:
            continue
        parts.append(f"<{tag}>\n")
        for k, v in data.items():
            parts.append(f'<DataArray type="Float64" Name="{k}" NumberOfComponents="{v.shape[1]}" format="ascii">\n')
            parts.append(fmt(v) + "\n</DataArray>\n")
        parts.append(f"</{tag}>\n")
    parts.append("</Piece>\n</UnstructuredGrid>\n</VTKFile>\n")
    with open(path, "w") as fh:
        fh.write("".join(parts))


def read_legacy_vtk(path):
    """(points, cells, point_data, cell_data) of a file written by ``_write_legacy`` (tests)."""
    with open(path, "rb") as fh:
        raw = fh.read()
    pos = 0

    def line():
        nonlocal pos
        j = raw.index(b"\n", pos)
        s = raw[pos:j].decode("ascii")
        pos = j + 1
        return s

    line()
    line()
    binary = line().strip() == "BINARY"
    line()

    def arr(count, dt):
        nonlocal pos
        if binary:
            size = np.dtype(dt).itemsize * count
            a = np.frombuffer(raw[pos:pos + size], dtype=dt).copy()
            pos += size + 1
            return a
        vals = []
        while len(vals) < count:
            vals.extend(line().split())
        return np.array(vals, dtype=float if np.dtype(dt).kind == "f" else np.int64)

    n_pts = int(line().split()[1])
    points = arr(3 * n_pts, ">f8").reshape(-1, 3)
    _, n_cells, size = line().split()
    conn = arr(int(size), ">i4").reshape(int(n_cells), -1)
    line()
    arr(int(n_cells), ">i4")
    pdata, cdata = {}, {}
    while pos < len(raw):
        head = line().split()
        if not head:
            continue
        target = pdata if head[0] == "POINT_DATA" else cdata
        n = int(head[1])
        k = int(line().split()[2])
        for _ in range(k):
            name, nc, nt, _ = line().split()
            target[name] = arr(int(nc) * int(nt), ">f8").reshape(int(nt), int(nc))
    return points, conn[:, 1:], pdata, cdata
