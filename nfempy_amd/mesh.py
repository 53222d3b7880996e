"""Mesh ingest and conformal entity numbering (host side, vectorised).

Mirrors ``Mesh`` of the reference (``src/mesh/mesh.py``): ``Mesh(dimension,
file_name)`` reads a gmsh 4.1 ASCII file (the reference goes through meshio,
``mesh.py:19``), ``build_conformal_mesh()`` numbers every entity exactly as
``mesh.py:239-339`` does, and ``topology(physical_tags)`` gives the
tag-filtered entity order of ``build_graph_on_physical_tags``
(``mesh.py:668-681``) and ``MeshTopology`` (``src/mesh/mesh_topology.py:28-35``).

Numbering rules (all "first appearance" orders are computed with ``np.unique``):
- vertex id = node index;
- edges: over tets with local pairs (2,3),(1,3),(1,2),(0,3),(0,2),(0,1), then
  triangles (1,2),(0,2),(0,1), then lines; ids start at V;
- faces: over tets (1,2,3),(0,2,3),(0,1,3),(0,1,2), then triangles;
- volumes: over tets;
- the stored node order of an entity is fixed by the block element that creates
  it first (blocks in file order): edges from lines keep the file order, edges
  from triangles are sorted, edges and faces from tets keep the tet-local order;
- an entity's material is the physical tag of the last block element that IS the
  entity (None -> -1); sub-entities created on the way get none.

Also holds structured generators (``unit_square``, ``unit_cube``) for the
benchmark configurations; they produce the same (points, blocks) a gmsh file
would, positively oriented, with the physical tags of ``build_box_2D/3D``
(``src/topology/domain_market.py:58,94``).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from itertools import permutations

import numpy as np

TET_EDGES = np.array([(2, 3), (1, 3), (1, 2), (0, 3), (0, 2), (0, 1)])
TET_FACES = np.array([(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)])
TRI_EDGES = np.array([(1, 2), (0, 2), (0, 1)])


@dataclass
class Block:
    dim: int
    data: np.ndarray  # (n, dim+1) node ids
    physical: np.ndarray  # (n,) physical tag, -1 = none

    @classmethod
    def make(cls, dim, data, tag):
        data = np.asarray(data, dtype=np.int64).reshape(-1, dim + 1)
        return cls(dim, data, np.full(len(data), tag, dtype=np.int64))


# ----------------------------------------------------------------------------
# gmsh 4.1 ASCII reader
# ----------------------------------------------------------------------------

_GMSH = {15: (0, 1), 1: (1, 2), 2: (2, 3), 4: (3, 4)}


def read_msh41(path):
    """(points, blocks) of a gmsh 4.1 ASCII file, node order by tag, blocks in file order."""
    with open(path) as fh:
        toks = fh.read().split("\n")
    it = iter(range(len(toks)))
    phys = {0: {}, 1: {}, 2: {}, 3: {}}
    nodes_tag, nodes_xyz = [], []
    blocks = []
    i = 0
    n = len(toks)
    while i < n:
        line = toks[i].strip()
        if line == "$Entities":
            cnt = [int(v) for v in toks[i + 1].split()]
            i += 2
            for d in range(4):
                for _ in range(cnt[d]):
                    f = toks[i].split()
                    k = 4 if d == 0 else 7
                    npt = int(f[k])
                    phys[d][int(f[0])] = int(f[k + 1]) if npt > 0 else -1
                    i += 1
            continue
        if line == "$Nodes":
            nb = int(toks[i + 1].split()[0])
            i += 2
            for _ in range(nb):
                _, _, _, nn = (int(v) for v in toks[i].split())
                i += 1
                tags = np.array(toks[i:i + nn], dtype=np.int64) if nn else np.zeros(0, np.int64)
                i += nn
                xyz = np.array([r.split()[:3] for r in toks[i:i + nn]], dtype=np.float64).reshape(nn, 3)
                i += nn
                nodes_tag.append(tags)
                nodes_xyz.append(xyz)
            continue
        if line == "$Elements":
            nb = int(toks[i + 1].split()[0])
            i += 2
            for _ in range(nb):
                edim, etag, etype, ne = (int(v) for v in toks[i].split())
                i += 1
                dim, nv = _GMSH[etype]
                arr = np.array([r.split()[1:1 + nv] for r in toks[i:i + ne]], dtype=np.int64).reshape(ne, nv)
                i += ne
                blocks.append((dim, arr, phys[edim].get(etag, -1)))
            continue
        i += 1
    tags = np.concatenate(nodes_tag)
    xyz = np.concatenate(nodes_xyz)
    order = np.argsort(tags)
    points = xyz[order]
    lookup = np.full(tags.max() + 1, -1, dtype=np.int64)
    lookup[tags[order]] = np.arange(len(tags))
    return points, [Block.make(d, lookup[a], t) for d, a, t in blocks]


# ----------------------------------------------------------------------------
# structured generators
# ----------------------------------------------------------------------------

def unit_square(n, lx=1.0, ly=1.0):
    """n x n squares, 2 CCW triangles each: (v00,v10,v11), (v00,v11,v01).

    Vertex (i, j) = i + (n+1) j. Boundary lines run CCW: south (tag 2), east (3),
    north (4), west (5); then the triangle block (tag 1).
    """
    m = n + 1
    x = np.linspace(0.0, lx, m)
    y = np.linspace(0.0, ly, m)
    pts = np.zeros((m * m, 3))
    J, I = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    pts[:, 0] = x[I.ravel()]
    pts[:, 1] = y[J.ravel()]
    jj, ii = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    ii, jj = ii.ravel(), jj.ravel()
    v00 = ii + m * jj
    v10 = v00 + 1
    v11 = v00 + 1 + m
    v01 = v00 + m
    tris = np.stack([np.stack([v00, v10, v11], 1), np.stack([v00, v11, v01], 1)], axis=1).reshape(-1, 3)
    r = np.arange(n)
    south = np.stack([r, r + 1], 1)
    east = np.stack([n + m * r, n + m * (r + 1)], 1)
    rr = r[::-1]
    north = np.stack([(rr + 1) + m * n, rr + m * n], 1)
    west = np.stack([m * (rr + 1), m * rr], 1)
    blocks = [Block.make(1, south, 2), Block.make(1, east, 3), Block.make(1, north, 4), Block.make(1, west, 5),
              Block.make(2, tris, 1)]
    return pts, blocks


def _kuhn_local():
    """Six Kuhn tets of the unit cube as corner indices a + 2b + 4c, all positive."""
    out = []
    for perm in permutations(range(3)):
        p = [0, 0, 0]
        path = [0]
        for ax in perm:
            p[ax] = 1
            path.append(p[0] + 2 * p[1] + 4 * p[2])
        inv = sum(1 for a in range(3) for b in range(a + 1, 3) if perm[a] > perm[b])
        if inv % 2:
            path[2], path[3] = path[3], path[2]
        out.append(path)
    return np.array(out)


def unit_cube(n, lx=1.0, ly=1.0, lz=1.0):
    """n^3 hexes split into 6 positively oriented Kuhn tets (hex (i,j,k), i fastest).

    Vertex (i,j,k) = i + (n+1) j + (n+1)^2 k. Boundary triangle blocks with the tags
    of ``build_box_3D``: 2 y=0, 3 x=1, 4 y=1, 5 x=0, 6 z=1, 7 z=0, then the tet
    block (tag 1).
    """
    m = n + 1
    pts = np.zeros((m ** 3, 3))
    K, J, I = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    pts[:, 0] = np.linspace(0, lx, m)[I.ravel()]
    pts[:, 1] = np.linspace(0, ly, m)[J.ravel()]
    pts[:, 2] = np.linspace(0, lz, m)[K.ravel()]
    kk, jj, ii = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    base = (ii + m * jj + m * m * kk).ravel()
    corner = np.array([a + m * b + m * m * c for c in (0, 1) for b in (0, 1) for a in (0, 1)])
    # corner index a + 2b + 4c -> offset
    off = np.zeros(8, dtype=np.int64)
    for c in (0, 1):
        for b in (0, 1):
            for a in (0, 1):
                off[a + 2 * b + 4 * c] = a + m * b + m * m * c
    tets = (base[:, None, None] + off[_kuhn_local()][None, :, :]).reshape(-1, 4)

    def face(axis, value, tag):
        other = [a for a in range(3) if a != axis]
        qq, pp = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        pp, qq = pp.ravel(), qq.ravel()

        def v(dp, dq):
            idx = [None, None, None]
            idx[axis] = np.full_like(pp, value)
            idx[other[0]] = pp + dp
            idx[other[1]] = qq + dq
            return idx[0] + m * idx[1] + m * m * idx[2]

        t1 = np.stack([v(0, 0), v(1, 0), v(1, 1)], 1)
        t2 = np.stack([v(0, 0), v(1, 1), v(0, 1)], 1)
        return Block.make(2, np.stack([t1, t2], 1).reshape(-1, 3), tag)

    blocks = [face(1, 0, 2), face(0, n, 3), face(1, n, 4), face(0, 0, 5), face(2, n, 6), face(2, 0, 7),
              Block.make(3, tets, 1)]
    return pts, blocks


# ----------------------------------------------------------------------------
# numbering
# ----------------------------------------------------------------------------

def _first_appearance_ids(keys, device=None):
    """keys (K,) int64 -> (ids (K,) in first-appearance order, first index per id). With a
    device the sort runs there (``numbering.first_appearance_ids``, same integers)."""
    if device is not None:
        from .numbering import first_appearance_ids

        return first_appearance_ids(keys, device)
    uniq, first, inv = np.unique(keys, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")
    rank = np.empty(len(uniq), dtype=np.int64)
    rank[order] = np.arange(len(uniq))
    return rank[inv.reshape(-1)], first[order]


@dataclass
class EntityTable:
    """Entities of one dimension: ids are ``offset + arange(n)``."""

    offset: int
    nodes: np.ndarray  # (n, d+1) stored node order
    material: np.ndarray  # (n,) -1 = None
    subs: dict = field(default_factory=dict)  # d' -> (n, n_sub) global ids

    @property
    def n(self):
        return len(self.nodes)


class Mesh:
    """Reference ``Mesh`` (``src/mesh/mesh.py:13``) after ``build_conformal_mesh``."""

    def __init__(self, dimension, file_name=None, points=None, blocks=None, device=None):
        """device: run the numbering sorts there ("cuda"), else on the host."""
        self.dimension = dimension
        self.device = device
        if file_name is not None:
            points, blocks = read_msh41(file_name)
        self.points = np.asarray(points, dtype=np.float64)
        self.blocks = list(blocks)
        self.entities = None

    @classmethod
    def from_blocks(cls, dimension, points, blocks, device=None):
        m = cls(dimension, points=points, blocks=blocks, device=device)
        m.build_conformal_mesh()
        return m

    def build_conformal_mesh(self):
        dims = [b.dim for b in self.blocks]
        if dims != sorted(dims):
            raise NotImplementedError("element blocks must come in ascending dimension (gmsh order)")
        V = len(self.points)
        by = {d: [b for b in self.blocks if b.dim == d] for d in range(4)}
        cat = lambda d, k: (np.concatenate([b.data for b in by[d]]) if by[d] else np.zeros((0, k), np.int64))
        tets, tris, lines = cat(3, 4), cat(2, 3), cat(1, 2)

        def key2(a, b):
            lo, hi = np.minimum(a, b), np.maximum(a, b)
            return lo * V + hi

        # ---- edges: ids by first appearance (tets, tris, lines)
        ek = [key2(tets[:, TET_EDGES[:, 0]], tets[:, TET_EDGES[:, 1]]).reshape(-1),
              key2(tris[:, TRI_EDGES[:, 0]], tris[:, TRI_EDGES[:, 1]]).reshape(-1),
              key2(lines[:, 0], lines[:, 1]).reshape(-1)]
        all_ek = np.concatenate(ek)
        dev = self.device
        eid, efirst = _first_appearance_ids(all_ek, dev)
        n_edges = len(efirst)
        tet_edge = eid[: ek[0].size].reshape(-1, 6) + V
        tri_edge = eid[ek[0].size: ek[0].size + ek[1].size].reshape(-1, 3) + V
        line_edge = eid[ek[0].size + ek[1].size:] + V
        # stored node order: events in block (file) order
        ev_id, ev_nodes = [], []
        for b in self.blocks:
            if b.dim == 1:
                ev_id.append(key2(b.data[:, 0], b.data[:, 1]))
                ev_nodes.append(b.data)
            elif b.dim == 2:
                a, c = b.data[:, TRI_EDGES[:, 0]].reshape(-1), b.data[:, TRI_EDGES[:, 1]].reshape(-1)
                ev_id.append(key2(a, c))
                ev_nodes.append(np.stack([np.minimum(a, c), np.maximum(a, c)], 1))
            elif b.dim == 3:
                a, c = b.data[:, TET_EDGES[:, 0]].reshape(-1), b.data[:, TET_EDGES[:, 1]].reshape(-1)
                ev_id.append(key2(a, c))
                ev_nodes.append(np.stack([a, c], 1))
        edge_nodes = np.zeros((n_edges, 2), dtype=np.int64)
        if ev_id:
            kk = np.concatenate(ev_id)
            nn = np.concatenate(ev_nodes)
            uk = np.unique(all_ek)
            pos = np.searchsorted(uk, kk)
            # map key -> id via the first-appearance table
            key_to_id = np.empty(len(uk), dtype=np.int64)
            key_to_id[np.searchsorted(uk, all_ek)] = eid
            ids = key_to_id[pos]
            _, first_ev = np.unique(ids, return_index=True)
            edge_nodes[ids[first_ev]] = nn[first_ev]
        edge_mat = np.full(n_edges, -1, dtype=np.int64)
        off = 0
        for b in by[1]:
            ids = line_edge[off: off + len(b.data)] - V
            edge_mat[ids] = b.physical  # later blocks overwrite
            off += len(b.data)

        # ---- faces (tets, tris)
        def key3(a, b, c):
            s = np.sort(np.stack([a, b, c], -1), axis=-1)
            return (s[..., 0] * V + s[..., 1]) * V + s[..., 2]

        fk = [key3(tets[:, TET_FACES[:, 0]], tets[:, TET_FACES[:, 1]], tets[:, TET_FACES[:, 2]]).reshape(-1),
              key3(tris[:, 0], tris[:, 1], tris[:, 2]).reshape(-1)]
        all_fk = np.concatenate(fk)
        foff = V + n_edges
        fid, ffirst = _first_appearance_ids(all_fk, dev) if all_fk.size else (np.zeros(0, np.int64), np.zeros(0))
        n_faces = len(ffirst)
        tet_face = fid[: fk[0].size].reshape(-1, 4) + foff
        tri_face = fid[fk[0].size:] + foff
        face_nodes = np.zeros((n_faces, 3), dtype=np.int64)
        if n_faces:
            ev_k, ev_n = [], []
            for b in self.blocks:
                if b.dim == 2:
                    ev_k.append(key3(b.data[:, 0], b.data[:, 1], b.data[:, 2]))
                    ev_n.append(b.data)
                elif b.dim == 3:
                    f = b.data[:, TET_FACES]  # (n, 4, 3)
                    ev_k.append(key3(f[..., 0], f[..., 1], f[..., 2]).reshape(-1))
                    ev_n.append(f.reshape(-1, 3))
            kk = np.concatenate(ev_k)
            nn = np.concatenate(ev_n)
            uk = np.unique(all_fk)
            key_to_id = np.empty(len(uk), dtype=np.int64)
            key_to_id[np.searchsorted(uk, all_fk)] = fid
            ids = key_to_id[np.searchsorted(uk, kk)]
            _, first_ev = np.unique(ids, return_index=True)
            face_nodes[ids[first_ev]] = nn[first_ev]
        face_mat = np.full(n_faces, -1, dtype=np.int64)
        off = 0
        for b in by[2]:
            ids = tri_face[off: off + len(b.data)] - foff
            face_mat[ids] = b.physical
            off += len(b.data)

        # ---- volumes
        voff = foff + n_faces if n_faces else 0
        if len(tets):
            s = np.sort(tets, axis=1)
            vk = ((s[:, 0] * V + s[:, 1]) * V + s[:, 2]) * V + s[:, 3] if V < 40000 else None
            if vk is None:
                _, vinv = np.unique(s, axis=0, return_inverse=True)
                vk = vinv.reshape(-1)
            vid, vfirst = _first_appearance_ids(vk, dev)
            n_vol = len(vfirst)
            vol_nodes = np.zeros((n_vol, 4), dtype=np.int64)
            vol_nodes[vid[vfirst]] = tets[vfirst]
            vol_mat = np.full(n_vol, -1, dtype=np.int64)
            off = 0
            for b in by[3]:
                vol_mat[vid[off: off + len(b.data)]] = b.physical
                off += len(b.data)
            tet_vol = vid + voff
        else:
            n_vol = 0
            vol_nodes = np.zeros((0, 4), np.int64)
            vol_mat = np.zeros(0, np.int64)
            tet_vol = np.zeros(0, np.int64)

        ent = {
            0: EntityTable(0, np.arange(V)[:, None], np.full(V, -1, np.int64)),
            1: EntityTable(V, edge_nodes, edge_mat),
            2: EntityTable(foff, face_nodes, face_mat),
            3: EntityTable(voff, vol_nodes, vol_mat),
        }
        for b in by[0]:
            ent[0].material[b.data[:, 0]] = b.physical
        ent[1].subs[0] = edge_nodes
        # sub-entities of the top-dimensional cells
        if len(tets):
            order = np.argsort(tet_vol - voff)
            t3 = ent[3]
            t3.subs[0] = tets[order]
            t3.subs[1] = tet_edge[order]
            t3.subs[2] = tet_face[order]
            t3.subs[3] = (np.arange(n_vol) + voff)[:, None]
        if len(tris) and not len(tets):
            # 2-D cells: face ids follow triangle order
            order = np.argsort(tri_face - foff)
            t2 = ent[2]
            t2.subs[0] = tris[order]
            t2.subs[1] = tri_edge[order]
            t2.subs[2] = (np.arange(n_faces) + foff)[:, None]
        self.entities = ent
        self.n_vertices, self.n_edges, self.n_faces, self.n_volumes = V, n_edges, n_faces, n_vol
        return self

    # -- queries ----------------------------------------------------------------
    def cell_table(self):
        return self.entities[self.dimension]

    def topology(self, physical_tags, dimension=None):
        return Topology(self, self.dimension if dimension is None else dimension, physical_tags)


class Topology:
    """Tag-filtered entity order (``MeshTopology.build_data_on_pysical_tags``).

    cells: global ids of tagged D-cells in id order (``element_ids``);
    entity_ids[d]: global ids of d-entities by first appearance;
    position[d]: dict-free lookup global id -> position (array, -1 = absent);
    first_cell[d]: for each entity in entity_ids[d], the index (into cells) of the
    first tagged cell listing it (its DiGraph first predecessor).
    """

    def __init__(self, mesh: Mesh, D, physical_tags):
        tab = mesh.entities[D]
        tags = np.array([t for t in physical_tags if t is not None], dtype=np.int64)
        sel = np.flatnonzero(np.isin(tab.material, tags))
        self.mesh = mesh
        self.dimension = D
        self.cell_index = sel  # positions within the D table
        self.cells = sel + tab.offset
        self.entity_ids = {}
        self.position = {}
        self.first_cell = {}
        self.n_total = {d: mesh.entities[d].n for d in range(4)}
        for d in range(D + 1):
            subs = tab.subs[d][sel]  # (nc, nsub) global ids
            flat = subs.reshape(-1)
            if flat.size == 0:
                self.entity_ids[d] = np.zeros(0, np.int64)
                self.position[d] = np.full(self.n_total[d], -1, np.int64)
                self.first_cell[d] = np.zeros(0, np.int64)
                continue
            pos, first = _first_appearance_ids(flat, getattr(mesh, "device", None))
            ids = flat[first]
            self.entity_ids[d] = ids
            p = np.full(self.n_total[d], -1, np.int64)
            p[ids - mesh.entities[d].offset] = np.arange(len(ids))
            self.position[d] = p
            self.first_cell[d] = first // subs.shape[1]

    def sub_positions(self, d):
        """(n_cells, n_sub) positions in entity_ids[d] of each cell's d-sub-entities."""
        tab = self.mesh.entities[self.dimension]
        subs = tab.subs[d][self.cell_index]
        return self.position[d][subs - self.mesh.entities[d].offset]
