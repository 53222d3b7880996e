"""Meshes and conformal entity numbering (FP64 oracle).

Restates ``Mesh.build_conformal_mesh`` and the ``insert_*`` helpers of
``src/mesh/mesh.py:125-339``, the physical-tag graphs of
``Mesh.build_graph_on_physical_tags`` / ``gather_graph_edges``
(``src/mesh/mesh.py:639-681``) and ``MeshTopology`` (``src/mesh/mesh_topology.py:16-65``)
with plain Python dictionaries, in the same order of operations:

1. entity ids: vertices = node index; edges by first appearance over tets (local
   pairs (2,3),(1,3),(1,2),(0,3),(0,2),(0,1)), then triangles ((1,2),(0,2),(0,1)),
   then lines; faces over tets ((1,2,3),(0,2,3),(0,1,3),(0,1,2)) then triangles;
   volumes over tets (``mesh.py:239-324``);
2. cells are created block by block in file order; the first creation of an entity
   fixes its stored node order (``mesh.py:137-237``): edges made from line elements
   keep the file order, edges made from triangles are sorted, edges and faces made
   from tets keep the tet-local order;
3. ``entity_ids[d]`` is the networkx node-insertion order of the DiGraph
   cell -> sub-entity over tag-filtered cells in id order (first appearance).

Also: structured generators (2-triangle squares, 6-tet Kuhn cubes, positively
oriented, physical tags as ``build_box_2D/3D`` in ``src/topology/domain_market.py:58,94``)
and a gmsh ``.msh`` 4.1 ASCII reader/writer standing in for meshio.
"""

from __future__ import annotations

import itertools

import numpy as np

TET_EDGES = [(2, 3), (1, 3), (1, 2), (0, 3), (0, 2), (0, 1)]
TET_FACES = [(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)]
TRI_EDGES = [(1, 2), (0, 2), (0, 1)]


class Block:
    """One element block as meshio hands it to the reference: dim, node ids, tag."""

    def __init__(self, dim, data, physical):
        self.dim = dim
        self.data = np.asarray(data, dtype=np.int64)
        self.physical = np.asarray(physical, dtype=np.int64)

    @property
    def type(self):
        return {0: "vertex", 1: "line", 2: "triangle", 3: "tetra"}[self.dim]


# --------------------------------------------------------------------------
# structured generators
# --------------------------------------------------------------------------

def structured_square(n, lx=1.0, ly=1.0):
    """Unit square, n x n squares split into 2 CCW triangles each.

    Vertex (i, j) has index ``i + (n + 1) j``. Square (i, j) gives the triangles
    (v00, v10, v11) and (v00, v11, v01). Boundary lines follow the CCW loop of
    ``build_box_2D``: south (tag 2) left->right, east (3) bottom->top,
    north (4) right->left, west (5) top->bottom. Blocks: lines, then triangles.
    """
    xs = np.linspace(0.0, lx, n + 1)
    ys = np.linspace(0.0, ly, n + 1)
    X, Y = np.meshgrid(xs, ys, indexing="xy")
    pts = np.zeros(((n + 1) ** 2, 3))
    pts[:, 0] = X.ravel()
    pts[:, 1] = Y.ravel()

    def vid(i, j):
        return i + (n + 1) * j

    tris = []
    for j in range(n):
        for i in range(n):
            v00, v10, v11, v01 = vid(i, j), vid(i + 1, j), vid(i + 1, j + 1), vid(i, j + 1)
            tris.append((v00, v10, v11))
            tris.append((v00, v11, v01))
    south = [(vid(i, 0), vid(i + 1, 0)) for i in range(n)]
    east = [(vid(n, j), vid(n, j + 1)) for j in range(n)]
    north = [(vid(i + 1, n), vid(i, n)) for i in reversed(range(n))]
    west = [(vid(0, j + 1), vid(0, j)) for j in reversed(range(n))]
    blocks = [
        Block(1, south, [2] * n),
        Block(1, east, [3] * n),
        Block(1, north, [4] * n),
        Block(1, west, [5] * n),
        Block(2, tris, [1] * len(tris)),
    ]
    return pts, blocks


def kuhn_tets_of_cube(c):
    """The 6 Kuhn tetrahedra of a cube with corner ids ``c[(a,b,d)]``, all positive.

    Tet for the axis permutation sigma is the monotone path v000 -> v111; odd
    permutations have their last two vertices swapped to orient them positively.
    """
    out = []
    for perm in itertools.permutations(range(3)):
        p = [0, 0, 0]
        path = [c[tuple(p)]]
        for ax in perm:
            p[ax] = 1
            path.append(c[tuple(p)])
        inv = sum(1 for a in range(3) for b in range(a + 1, 3) if perm[a] > perm[b])
        if inv % 2 == 1:
            path[2], path[3] = path[3], path[2]
        out.append(tuple(path))
    return out


def structured_cube(n, lx=1.0, ly=1.0, lz=1.0):
    """Unit cube, n^3 hexes split into 6 Kuhn tets each (all positive).

    Vertex (i,j,k) has index ``i + (n+1) j + (n+1)^2 k``. Hexes are visited with
    i fastest. Boundary triangles (2 per boundary square, both diagonals through
    the Kuhn diagonal so they match the tets) carry the tags of ``build_box_3D``:
    2: y=0, 3: x=1, 4: y=1, 5: x=0, 6: z=1, 7: z=0. Blocks: six triangle blocks
    (one per face), then the tet block.
    """
    m = n + 1
    xs, ys, zs = np.linspace(0, lx, m), np.linspace(0, ly, m), np.linspace(0, lz, m)
    pts = np.zeros((m**3, 3))
    K, J, I = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    pts[:, 0] = xs[I.ravel()]
    pts[:, 1] = ys[J.ravel()]
    pts[:, 2] = zs[K.ravel()]

    def vid(i, j, k):
        return i + m * j + m * m * k

    tets = []
    for k in range(n):
        for j in range(n):
            for i in range(n):
                c = {(a, b, d): vid(i + a, j + b, k + d) for a in (0, 1) for b in (0, 1) for d in (0, 1)}
                tets.extend(kuhn_tets_of_cube(c))

    def face_tris(fixed_axis, value, tag):
        tris = []
        other = [a for a in range(3) if a != fixed_axis]
        for q in range(n):
            for p in range(n):
                def v(dp, dq):
                    idx = [0, 0, 0]
                    idx[fixed_axis] = value
                    idx[other[0]] = p + dp
                    idx[other[1]] = q + dq
                    return vid(*idx)
                # split along the (0,0)-(1,1) diagonal, which is a Kuhn edge
                tris.append((v(0, 0), v(1, 0), v(1, 1)))
                tris.append((v(0, 0), v(1, 1), v(0, 1)))
        return Block(2, tris, [tag] * len(tris))

    blocks = [
        face_tris(1, 0, 2),
        face_tris(0, n, 3),
        face_tris(1, n, 4),
        face_tris(0, 0, 5),
        face_tris(2, n, 6),
        face_tris(2, 0, 7),
        Block(3, tets, [1] * len(tets)),
    ]
    return pts, blocks


# --------------------------------------------------------------------------
# gmsh 4.1 ASCII
# --------------------------------------------------------------------------

_GMSH_TYPE = {1: (1, 2), 2: (2, 3), 4: (3, 4), 15: (0, 1)}
_GMSH_CODE = {0: 15, 1: 1, 2: 2, 3: 4}


def read_msh41(path):
    """Read a gmsh 4.1 ASCII file into (points, blocks) as meshio would.

    Points are ordered by node tag; each $Elements entity block becomes one block
    whose physical tag is the first physical tag of its entity (None -> -1).
    """
    with open(path) as fh:
        lines = [ln.strip() for ln in fh]
    i = 0
    phys = {0: {}, 1: {}, 2: {}, 3: {}}
    nodes = {}
    blocks = []
    while i < len(lines):
        ln = lines[i]
        if ln == "$Entities":
            i += 1
            counts = [int(v) for v in lines[i].split()]
            i += 1
            for dim in range(4):
                for _ in range(counts[dim]):
                    tok = lines[i].split()
                    tag = int(tok[0])
                    if dim == 0:
                        nphys = int(tok[4])
                        ptags = [int(v) for v in tok[5:5 + nphys]]
                    else:
                        nphys = int(tok[7])
                        ptags = [int(v) for v in tok[8:8 + nphys]]
                    phys[dim][tag] = ptags[0] if ptags else -1
                    i += 1
            continue
        if ln == "$Nodes":
            i += 1
            nblocks = int(lines[i].split()[0])
            i += 1
            for _ in range(nblocks):
                _, _, _, nn = [int(v) for v in lines[i].split()]
                i += 1
                tags = [int(lines[i + t]) for t in range(nn)]
                i += nn
                for t in range(nn):
                    nodes[tags[t]] = [float(v) for v in lines[i + t].split()[:3]]
                i += nn
            continue
        if ln == "$Elements":
            i += 1
            nblocks = int(lines[i].split()[0])
            i += 1
            for _ in range(nblocks):
                edim, etag, etype, ne = [int(v) for v in lines[i].split()]
                i += 1
                dim, nv = _GMSH_TYPE[etype]
                data = np.array([[int(v) for v in lines[i + t].split()[1:1 + nv]] for t in range(ne)], dtype=np.int64)
                i += ne
                blocks.append((dim, data, phys[edim].get(etag, -1)))
            continue
        i += 1
    node_tags = np.array(sorted(nodes))
    pts = np.array([nodes[t] for t in node_tags], dtype=np.float64)
    remap = {int(t): k for k, t in enumerate(node_tags)}
    out = []
    for dim, data, ptag in blocks:
        mapped = np.vectorize(remap.__getitem__)(data) if data.size else data
        out.append(Block(dim, mapped, [ptag] * len(data)))
    return pts, out


def write_msh41(path, points, blocks):
    """Write (points, blocks) as gmsh 4.1 ASCII, one entity per block."""
    lines = ["$MeshFormat", "4.1 0 8", "$EndMeshFormat"]
    ents = {0: [], 1: [], 2: [], 3: []}
    for bi, b in enumerate(blocks):
        ents[b.dim].append((bi + 1, int(b.physical[0]) if len(b.physical) else -1))
    lines.append("$Entities")
    lines.append(" ".join(str(len(ents[d])) for d in range(4)))
    lo, hi = points.min(axis=0), points.max(axis=0)
    for d in range(4):
        for tag, ptag in ents[d]:
            ph = f"1 {ptag}" if ptag >= 0 else "0"
            if d == 0:
                lines.append(f"{tag} {lo[0]} {lo[1]} {lo[2]} {ph}")
            else:
                lines.append(f"{tag} {lo[0]} {lo[1]} {lo[2]} {hi[0]} {hi[1]} {hi[2]} {ph} 0")
    lines.append("$EndEntities")
    nv = len(points)
    lines.append("$Nodes")
    lines.append(f"1 {nv} 1 {nv}")
    lines.append(f"3 1 0 {nv}")
    lines.extend(str(t + 1) for t in range(nv))
    lines.extend(f"{float(p[0])!r} {float(p[1])!r} {float(p[2])!r}" for p in points)
    lines.append("$EndNodes")
    ne = sum(len(b.data) for b in blocks)
    lines.append("$Elements")
    lines.append(f"{len(blocks)} {ne} 1 {ne}")
    eid = 1
    for bi, b in enumerate(blocks):
        lines.append(f"{b.dim} {bi + 1} {_GMSH_CODE[b.dim]} {len(b.data)}")
        for row in b.data:
            lines.append(f"{eid} " + " ".join(str(int(v) + 1) for v in row))
            eid += 1
    lines.append("$EndElements")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


# --------------------------------------------------------------------------
# conformal numbering
# --------------------------------------------------------------------------

class Cell:
    __slots__ = ("dimension", "id", "material_id", "node_tags", "sub_cells_ids")

    def __init__(self, dimension, node_tags, cid, material_id):
        self.dimension = dimension
        self.id = cid
        self.material_id = material_id
        self.node_tags = np.asarray(node_tags, dtype=np.int64)
        self.sub_cells_ids = [np.zeros(0, dtype=np.int64) for _ in range(dimension + 1)]


class ConformalMesh:
    """The reference ``Mesh`` after ``build_conformal_mesh`` (ids, cells, tags)."""

    def __init__(self, dimension, points, blocks):
        self.dimension = dimension
        self.points = np.asarray(points, dtype=np.float64)
        self.blocks = blocks
        self._build()

    def _build(self):
        e0, e1, e2, e3 = {}, {}, {}, {}
        for i in range(len(self.points)):
            e0[i] = i
        by_type = {d: [b.data for b in self.blocks if b.dim == d] for d in range(4)}
        tets = np.concatenate(by_type[3]) if by_type[3] else np.zeros((0, 4), np.int64)
        tris = np.concatenate(by_type[2]) if by_type[2] else np.zeros((0, 3), np.int64)
        lines = np.concatenate(by_type[1]) if by_type[1] else np.zeros((0, 2), np.int64)
        nid = len(e0)
        for nodes in tets.tolist():
            for a, b in TET_EDGES:
                key = tuple(sorted((nodes[a], nodes[b])))
                if key not in e1:
                    e1[key] = nid
                    nid += 1
        for nodes in tris.tolist():
            for a, b in TRI_EDGES:
                key = tuple(sorted((nodes[a], nodes[b])))
                if key not in e1:
                    e1[key] = nid
                    nid += 1
        for nodes in lines.tolist():
            key = tuple(sorted(nodes))
            if key not in e1:
                e1[key] = nid
                nid += 1
        nid = max(e1.values()) + 1 if e1 else nid
        for nodes in tets.tolist():
            for f in TET_FACES:
                key = tuple(sorted(nodes[v] for v in f))
                if key not in e2:
                    e2[key] = nid
                    nid += 1
        for nodes in tris.tolist():
            key = tuple(sorted(nodes))
            if key not in e2:
                e2[key] = nid
                nid += 1
        nid = max(e2.values()) + 1 if e2 else 0
        for nodes in tets.tolist():
            key = tuple(sorted(nodes))
            if key not in e3:
                e3[key] = nid
                nid += 1
        n_cells = len(e0) + len(e1) + len(e2) + len(e3)
        self.cells = [None] * n_cells
        self._e = (e0, e1, e2, e3)
        for b in self.blocks:
            for nodes, ptag in zip(b.data.tolist(), b.physical.tolist()):
                tag = None if ptag < 0 else int(ptag)
                if b.dim == 0:
                    self._insert_vertex(nodes[0], tag)
                elif b.dim == 1:
                    self._insert_edge(nodes, tag)
                elif b.dim == 2:
                    self._insert_polygon(nodes, tag)
                else:
                    self._insert_polyhedron(nodes, tag)
        self.n_vertices = len(e0)
        self.n_edges = len(e1)
        self.n_faces = len(e2)
        self.n_volumes = len(e3)
        del self._e

    def _insert_vertex(self, node, tag=None):
        cid = self._e[0][node]
        if self.cells[cid] is None:
            c = Cell(0, [node], cid, tag)
            c.sub_cells_ids[0] = np.array([node])
            self.cells[cid] = c
        if tag is not None:
            self.cells[cid].material_id = tag
        return cid

    def _insert_edge(self, nodes, tag=None):
        cid = self._e[1][tuple(sorted(nodes))]
        if self.cells[cid] is None:
            vids = [self._insert_vertex(v) for v in nodes]
            c = Cell(1, list(nodes), cid, tag)
            c.sub_cells_ids[0] = np.array(vids)
            c.sub_cells_ids[1] = np.array([cid])
            self.cells[cid] = c
        if tag is not None:
            self.cells[cid].material_id = tag
        return cid

    def _insert_polygon(self, nodes, tag=None):
        cid = self._e[2][tuple(sorted(nodes))]
        if self.cells[cid] is None:
            vids = [self._e[0][v] for v in nodes]
            eids = [self._insert_edge(sorted((nodes[a], nodes[b]))) for a, b in TRI_EDGES]
            c = Cell(2, list(nodes), cid, tag)
            c.sub_cells_ids[0] = np.array(vids)
            c.sub_cells_ids[1] = np.array(eids)
            c.sub_cells_ids[2] = np.array([cid])
            self.cells[cid] = c
        if tag is not None:
            self.cells[cid].material_id = tag
        return cid

    def _insert_polyhedron(self, nodes, tag=None):
        cid = self._e[3][tuple(sorted(nodes))]
        if self.cells[cid] is None:
            vids = [self._e[0][v] for v in nodes]
            eids = [self._insert_edge([nodes[a], nodes[b]]) for a, b in TET_EDGES]
            fids = [self._insert_polygon([nodes[v] for v in f]) for f in TET_FACES]
            c = Cell(3, list(nodes), cid, tag)
            c.sub_cells_ids[0] = np.array(vids)
            c.sub_cells_ids[1] = np.array(eids)
            c.sub_cells_ids[2] = np.array(fids)
            c.sub_cells_ids[3] = np.array([cid])
            self.cells[cid] = c
        if tag is not None:
            self.cells[cid].material_id = tag
        return cid


class Topology:
    """``MeshTopology.build_data_on_pysical_tags`` for one space (dimension D).

    ``entity_ids[d]``: d-entities in DiGraph node-insertion order.
    ``successors(cell_id, d)``: sub-entities of a tagged D-cell (local order).
    ``predecessor(entity_id, d)``: first tagged D-cell that lists the entity.
    """

    def __init__(self, mesh: ConformalMesh, dimension, physical_tags):
        self.mesh = mesh
        self.dimension = dimension
        D = dimension
        self.cells = [
            c for c in mesh.cells if c is not None and c.dimension == D and c.material_id in physical_tags
        ]
        self.entity_ids = {}
        self._pred = {}
        for d in range(D + 1):
            seen = {}
            pred = {}
            for c in self.cells:
                for sid in c.sub_cells_ids[d].tolist():
                    if sid not in seen:
                        seen[sid] = len(seen)
                    if sid not in pred:
                        pred[sid] = c.id
            self.entity_ids[d] = list(seen.keys())
            self._pred[d] = pred

    def successors(self, cell_id, d):
        return self.mesh.cells[cell_id].sub_cells_ids[d]

    def predecessor(self, entity_id, d):
        return self._pred[d].get(entity_id)
