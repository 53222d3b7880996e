"""Entity numbering, stored node orders, tag-filtered entity order, DoF layout,
destinations and facet orientation codes on four tiny meshes, against values
derived by hand from the reference's rules (not from either implementation):

- entity ids (``src/mesh/mesh.py:239-324``): vertices keep their node index; edge
  ids by first appearance over tets (local pairs (2,3),(1,3),(1,2),(0,3),(0,2),(0,1)),
  then triangles ((1,2),(0,2),(0,1)), then lines; face ids over tets (local faces
  (1,2,3),(0,2,3),(0,1,3),(0,1,2)) then triangles; volumes last;
- stored node order (``mesh.py:137-237``): the first insertion, blocks in file order;
  edges inserted by a line keep the file order, by a triangle are sorted, by a tet
  keep the tet-local order; faces keep the inserting block's order;
- ``entity_ids[d]`` (``src/mesh/mesh_topology.py:28-35``): first appearance over the
  tagged cells in id order (networkx node-insertion order of the cell -> sub-entity
  DiGraph, ``mesh.py:639-681``);
- DoFs (``src/spaces/dof_map.py:38-170``): per dimension, contiguous chunks of
  ``e_dofs * n_comp`` in ``entity_ids`` order, laid out ``[j * n_comp + c]``;
  fields stacked in dict order (``src/spaces/product_space.py:147-162``);
- boundary facets (``src/spaces/discrete_space.py:166-176``): ``entity_ids[D-1]``
  order filtered by tag; neighbour = first tagged cell listing the facet;
- orientation (``src/basis/permute_and_transform.py:8-149``): 2-D edges against the
  stored order; 3-D faces by the position in ``itertools.permutations(sorted)``
  mapped to (rotations, reflections), tagged faces at local index 1 or 3 forced to
  one reflection. Codes are written as 8 | rot | ref << 2 (0 = untouched).

Each mesh is built twice, by the oracle and by the product, and both must give
exactly the literal values below.
"""

import numpy as np
import pytest

import oracle.mesh as om
import oracle.spaces as osp


def _blocks(spec, cls):
    return [cls(d, data, tags) for d, data, tags in spec]


PTS2 = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], dtype=float)
PTS_SQ = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], dtype=float)
PTS_TET = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=float)
PTS_2TET = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1]], dtype=float)

MESHES = {
    # one triangle, its three sides given as lines in non-sorted order
    "triangle": dict(
        dim=2, pts=PTS2,
        blocks=[(1, [[1, 0], [1, 2], [2, 0]], [2, 3, 4]), (2, [[0, 1, 2]], [1])],
        edges={3: [1, 2], 4: [2, 0], 5: [1, 0]},
        edge_tags={3: 3, 4: 4, 5: 2},
        faces={6: [0, 1, 2]},
        cell_subs={6: {1: [3, 4, 5]}},
        entity_ids={1: [3, 4, 5], 2: [6]},
        codes=[[0, 12, 12]],
        btags=[2, 3, 4],
        bc_ids=[3, 4, 5],
        rt1_dest=[[0, 1, 2, 3]],
        rt1_bc=[[0], [1], [2]],
    ),
    # the 2-triangle unit square of the structured generator
    "square": dict(
        dim=2, pts=PTS_SQ,
        blocks=[(1, [[0, 1]], [2]), (1, [[1, 3]], [3]), (1, [[3, 2]], [4]), (1, [[2, 0]], [5]),
                (2, [[0, 1, 3], [0, 3, 2]], [1, 1])],
        edges={4: [1, 3], 5: [0, 3], 6: [0, 1], 7: [3, 2], 8: [2, 0]},
        edge_tags={4: 3, 5: None, 6: 2, 7: 4, 8: 5},
        faces={9: [0, 1, 3], 10: [0, 3, 2]},
        cell_subs={9: {1: [4, 5, 6]}, 10: {1: [7, 8, 5]}},
        entity_ids={1: [4, 5, 6, 7, 8], 2: [9, 10]},
        codes=[[0, 0, 0], [0, 12, 0]],
        btags=[2, 3, 4, 5],
        bc_ids=[4, 6, 7, 8],
        rt1_dest=[[0, 1, 2, 5], [3, 4, 1, 6]],
        rt1_bc=[[0], [2], [3], [4]],
        # RT2 (2 DoFs per edge, 2 interior) - discontinuous P1 (3 per cell)
        rt2_dest=[[0, 1, 2, 3, 4, 5, 10, 11, 14, 15, 16], [6, 7, 8, 9, 2, 3, 12, 13, 17, 18, 19]],
        rt2_bc=[[0, 1], [4, 5], [6, 7], [8, 9]],
    ),
    # one tet in non-sorted node order, three of its faces tagged (given in other orders)
    "tet": dict(
        dim=3, pts=PTS_TET,
        blocks=[(2, [[3, 2, 1], [0, 1, 3], [0, 2, 1]], [2, 3, 4]), (3, [[2, 0, 3, 1]], [1])],
        edges={4: [1, 3], 5: [0, 1], 6: [0, 3], 7: [1, 2], 8: [2, 3], 9: [0, 2]},
        faces={10: [0, 1, 3], 11: [3, 2, 1], 12: [0, 2, 1], 13: [2, 0, 3]},
        face_tags={10: 3, 11: 2, 12: 4, 13: None},
        cell_subs={14: {1: [4, 5, 6, 7, 8, 9], 2: [10, 11, 12, 13]}, 11: {1: [7, 4, 8]},
                   10: {1: [4, 6, 5]}, 12: {1: [7, 5, 9]}, 13: {1: [6, 8, 9]}},
        entity_ids={2: [10, 11, 12, 13], 3: [14]},
        # f0 [0,3,1]: position 1 -> (0, 1); f1 [2,3,1]: position 3 -> (1, 0), tagged at
        # index 1 -> one reflection forced; f2 [2,0,1]: position 4 -> (2, 0);
        # f3 [2,0,3]: position 2 -> (2, 1), untagged
        codes=[[12, 13, 10, 14]],
        btags=[2, 3, 4],
        bc_ids=[10, 11, 12],
        rt1_dest=[[0, 1, 2, 3, 4]],
        rt1_bc=[[0], [1], [2]],
    ),
    # two tets sharing the face {1, 2, 3}; the second in non-sorted node order
    "two_tets": dict(
        dim=3, pts=PTS_2TET,
        blocks=[(3, [[0, 1, 2, 3], [4, 2, 1, 3]], [1, 1])],
        edges={5: [2, 3], 6: [1, 3], 7: [1, 2], 8: [0, 3], 9: [0, 2], 10: [0, 1], 11: [4, 3], 12: [4, 1],
               13: [4, 2]},
        faces={14: [1, 2, 3], 15: [0, 2, 3], 16: [0, 1, 3], 17: [0, 1, 2], 18: [4, 1, 3], 19: [4, 2, 3],
               20: [4, 2, 1]},
        face_tags={f: None for f in range(14, 21)},
        cell_subs={21: {1: [5, 6, 7, 8, 9, 10], 2: [14, 15, 16, 17]},
                   22: {1: [6, 5, 7, 11, 12, 13], 2: [14, 18, 19, 20]}},
        entity_ids={2: [14, 15, 16, 17, 18, 19, 20], 3: [21, 22]},
        # second tet: f0 [2,1,3] position 2 -> (2, 1); f1 [4,1,3] position 4 -> (2, 0);
        # f2 [4,2,3] position 4 -> (2, 0); f3 [4,2,1] position 5 -> (1, 1)
        codes=[[0, 0, 0, 0], [14, 10, 10, 13]],
        btags=[],
        bc_ids=[],
        rt1_dest=[[0, 1, 2, 3, 7], [0, 4, 5, 6, 8]],
        rt1_bc=[],
        # BDM1 stress (3 DoFs per face x 3 components), P0 displacement and rotation (3 each)
        bdm_dest=[list(range(0, 36)) + [63, 64, 65, 69, 70, 71],
                  list(range(0, 9)) + list(range(36, 63)) + [66, 67, 68, 72, 73, 74]],
    ),
}


def _oracle(name):
    m = MESHES[name]
    return om.ConformalMesh(m["dim"], m["pts"], _blocks(m["blocks"], om.Block))


def _product(name):
    from nfempy_amd.mesh import Block, Mesh

    m = MESHES[name]
    return Mesh.from_blocks(m["dim"], m["pts"], [Block.make(d, data, tags) for d, data, tags in m["blocks"]])


def _product_nodes(mesh, cid):
    for d in range(4):
        t = mesh.entities[d]
        if t.offset <= cid < t.offset + t.n:
            return d, t.nodes[cid - t.offset].tolist(), int(t.material[cid - t.offset])
    raise KeyError(cid)


@pytest.mark.parametrize("name", list(MESHES))
def test_entities_by_hand(name):
    m = MESHES[name]
    o = _oracle(name)
    p = _product(name)
    for ids, dim in ((m["edges"], 1), (m["faces"], 2)):
        for cid, nodes in ids.items():
            assert o.cells[cid].dimension == dim and o.cells[cid].node_tags.tolist() == nodes, (name, cid)
            d, pn, _ = _product_nodes(p, cid)
            assert d == dim and pn == nodes, (name, cid, pn)
    for cid, tag in m.get("edge_tags", {}).items():
        assert o.cells[cid].material_id == tag
        assert _product_nodes(p, cid)[2] == (-1 if tag is None else tag)
    for cid, tag in m.get("face_tags", {}).items():
        assert o.cells[cid].material_id == tag
        assert _product_nodes(p, cid)[2] == (-1 if tag is None else tag)
    n_ent = len(m["pts"]) + len(m["edges"]) + len(m["faces"]) + len(m["entity_ids"].get(3, []))
    assert len(o.cells) == n_ent
    D = m["dim"]
    for cid, subs in m["cell_subs"].items():
        for d, expect in subs.items():
            assert o.cells[cid].sub_cells_ids[d].tolist() == expect, (name, cid, d)
            if o.cells[cid].dimension == D:
                tab = p.entities[D]
                assert tab.subs[d][cid - tab.offset].tolist() == expect, (name, cid, d)
    ot = om.Topology(o, D, [1])
    pt = p.topology([1], D)
    for d, expect in m["entity_ids"].items():
        assert list(ot.entity_ids[d]) == expect, (name, d)
        assert pt.entity_ids[d].tolist() == expect, (name, d)


def _spaces(name, fields, btag_field):
    from nfempy_amd.spaces import ProductSpace

    m = MESHES[name]
    D = m["dim"]
    o = _oracle(name)
    disc = {k: True for k, v in fields.items() if v[0] == "Lagrange"}
    ops = osp.ProductSpace({k: (D, nc, fam, k_, o) for k, (fam, k_, nc) in fields.items()}, disc)
    tags = {k: [1] for k in fields}
    btags = {btag_field: m["btags"]} if m["btags"] else {}
    ops.build_structures(tags, btags)
    pps = ProductSpace({k: (D, nc, fam, k_, _product(name)) for k, (fam, k_, nc) in fields.items()})
    pps.make_subspaces_discontinuous(disc)
    pps.build_structures(tags, btags)
    return ops, pps


@pytest.mark.parametrize("name", list(MESHES))
def test_rt1_p0_destinations_by_hand(name):
    m = MESHES[name]
    ops, pps = _spaces(name, {"u": ("RT", 1, 1), "p": ("Lagrange", 0, 1)}, "u")
    E = len(m["rt1_dest"])
    for i in range(E):
        assert ops.destination_indexes(i).tolist() == m["rt1_dest"][i], (name, i)
        assert pps.destination_indexes(i).tolist() == m["rt1_dest"][i], (name, i)
    assert list(ops.spaces["u"].bc_element_ids) == m["bc_ids"]
    assert list(pps.discrete_spaces["u"].bc_element_ids) == m["bc_ids"]
    for j, expect in enumerate(m["rt1_bc"]):
        assert ops.bc_destination_indexes(j).tolist() == expect, (name, j)
        assert pps.bc_destination_indexes(j).tolist() == expect, (name, j)
    assert pps.discrete_spaces["u"].codes.tolist() == m["codes"], name


def test_rt2_p1_destinations_by_hand():
    m = MESHES["square"]
    ops, pps = _spaces("square", {"u": ("RT", 2, 1), "p": ("Lagrange", 1, 1)}, "u")
    for i, expect in enumerate(m["rt2_dest"]):
        assert ops.destination_indexes(i).tolist() == expect
        assert pps.destination_indexes(i).tolist() == expect
    for j, expect in enumerate(m["rt2_bc"]):
        assert ops.bc_destination_indexes(j).tolist() == expect
        assert pps.bc_destination_indexes(j).tolist() == expect


def test_bdm1_three_field_destinations_by_hand():
    m = MESHES["two_tets"]
    ops, pps = _spaces("two_tets", {"s": ("BDM", 1, 3), "u": ("Lagrange", 0, 3), "t": ("Lagrange", 0, 3)}, "s")
    assert ops.n_dof == pps.n_dof == 75
    for i, expect in enumerate(m["bdm_dest"]):
        assert ops.destination_indexes(i).tolist() == expect
        assert pps.destination_indexes(i).tolist() == expect
    assert pps.discrete_spaces["s"].codes.tolist() == m["codes"]


def test_oracle_face_orientation_by_hand():
    """The oracle's (oriented, rotations, reflections) per face of the one-tet mesh."""
    from oracle.geometry import face_orientation_3d

    o = _oracle("tet")
    cell = o.cells[14]
    bc = [o.cells[f].material_id is not None for f in cell.sub_cells_ids[2].tolist()]
    assert bc == [True, True, True, False]
    assert face_orientation_3d(cell, bc) == [(False, 0, 1), (False, 1, 1), (False, 2, 0), (False, 2, 1)]
    o2 = _oracle("two_tets")
    assert face_orientation_3d(o2.cells[22], [False] * 4) == [(False, 2, 1), (False, 2, 0), (False, 2, 0),
                                                               (False, 1, 1)]
    assert face_orientation_3d(o2.cells[21], [False] * 4) == [(True, 0, 0)] * 4
