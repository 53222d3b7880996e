"""DoF maps and destination indices (FP64 oracle).

Restates ``DoFMap.build_entity_maps`` / ``destination_indices`` /
``bc_destination_indices`` (``src/spaces/dof_map.py:38-170``),
``DiscreteSpace`` (``src/spaces/discrete_space.py:14-221``) and
``ProductSpace`` (``src/spaces/product_space.py:9-201``):

- per dimension d the DoFs of the d-entities are contiguous chunks of
  ``e_dofs * n_comp`` ids, entities in ``entity_ids[d]`` order, dimension blocks
  vertex -> edge -> face -> volume, ``e_dofs = int(mean(num_entity_dofs[d]))``;
- element destination = concatenation over d of the chunks of the cell's
  d-sub-entities in local order;
- fields are stacked in dict insertion order (offset = sum of previous n_dof);
- boundary facets: the first tagged cell that lists the facet is its neighbour,
  destination = chunks of the facet's own sub-entities of dimension < D (with
  dimensions that carry no DoF skipped);
- quadrature degree ``2 * max(k) + 1`` shared by every field; the same degree
  on facets;
- BC facet families: RT -> discontinuous P_{k-1}, BDM -> discontinuous P_k.
"""

from __future__ import annotations

import numpy as np

from .mesh import ConformalMesh, Topology
from .quadrature import gauss_jacobi
from .reference_elements import create_element

FAMILY = {"Lagrange": "P", "RT": "RT", "BDM": "BDM"}


class DiscreteSpace:
    def __init__(self, dimension, n_comp, family, k, mesh: ConformalMesh, discontinuous=False):
        self.dimension = dimension
        self.n_comp = n_comp
        self.family = FAMILY[family]
        self.k = k
        self.mesh = mesh
        self.discontinuous = discontinuous
        self.element = create_element(self.family, dimension, k)

    def num_entity_dofs(self):
        e = self.element
        if self.discontinuous:
            counts = [[0] * len(ed) for ed in e.entity_dofs]
            counts[self.dimension][0] = e.ndofs
            return counts
        return e.num_entity_dofs

    def build_structures(self, physical_tags, bc_physical_tags):
        D = self.dimension
        self.topology = Topology(self.mesh, D, physical_tags)
        self.e_dofs = [int(np.mean(n)) for n in self.num_entity_dofs()]
        offsets = [0]
        for d in range(4):
            n_ent = len(self.topology.entity_ids[d]) if d <= D else 0
            e = self.e_dofs[d] if d < len(self.e_dofs) else 0
            offsets.append(offsets[-1] + n_ent * e * self.n_comp)
        self.n_dof = offsets[-1]
        self.chunk = [{} for _ in range(4)]
        for d in range(D + 1):
            size = self.e_dofs[d] * self.n_comp
            for p, eid in enumerate(self.topology.entity_ids[d]):
                self.chunk[d][eid] = np.arange(offsets[d] + p * size, offsets[d] + (p + 1) * size)
        self.element_ids = list(self.topology.entity_ids[D])
        self.id_to_element = {cid: i for i, cid in enumerate(self.element_ids)}
        self.bc_element_ids = []
        if bc_physical_tags:
            self.bc_element_ids = [
                fid for fid in self.topology.entity_ids[D - 1]
                if self.mesh.cells[fid].material_id in bc_physical_tags
            ]
        self.id_to_bc_element = {cid: i for i, cid in enumerate(self.bc_element_ids)}
        if self.family == "BDM":
            self.bc_family, self.bc_k = "P", self.k
        elif self.family == "RT":
            self.bc_family, self.bc_k = "P", self.k - 1
        else:
            self.bc_family, self.bc_k = "P", self.k
        # the facet element only fixes the number of traced DoFs per facet
        # (= DoFs of the cell element on that facet), see forms._bc_trace

    def destination_indices(self, cell_id):
        D = self.dimension
        parts = []
        for d in range(D + 1):
            for sid in self.topology.successors(cell_id, d).tolist():
                parts.append(self.chunk[d].get(sid, np.zeros(0, dtype=np.int64)))
        return np.concatenate(parts).astype(np.int64)

    def bc_neighbour(self, facet_id):
        n = self.topology.predecessor(facet_id, self.dimension - 1)
        assert n is not None
        return n

    def bc_destination_indices(self, facet_id):
        D = self.dimension
        parts = []
        neigh = self.bc_neighbour(facet_id)
        for d in range(D):
            if self.e_dofs[d] == 0:
                continue
            subs = set(self.topology.successors(neigh, d).tolist())
            for sid in self.mesh.cells[facet_id].sub_cells_ids[d].tolist():
                if sid in subs:
                    parts.append(self.chunk[d][sid])
        if not parts:
            return np.zeros(0, dtype=np.int64)
        return np.concatenate(parts).astype(np.int64)


class ProductSpace:
    """``discrete_spaces_data = {name: (dim, n_comp, family, k, mesh)}``."""

    def __init__(self, discrete_spaces_data, discontinuous=None):
        discontinuous = discontinuous or {}
        self.names = list(discrete_spaces_data)
        ks = [v[3] for v in discrete_spaces_data.values()]
        dims = [v[0] for v in discrete_spaces_data.values()]
        self.integration_order = 2 * max(ks) + 1
        self.quadrature = [None] * 4
        for d in range(max(dims) + 1):
            self.quadrature[d] = gauss_jacobi(d, self.integration_order)
        self.bc_quadrature = [None] * 4
        for d in dims:
            self.bc_quadrature[d - 1] = gauss_jacobi(d - 1, self.integration_order)
        self.spaces = {}
        for name, (dim, nc, fam, k, mesh) in discrete_spaces_data.items():
            self.spaces[name] = DiscreteSpace(dim, nc, fam, k, mesh, discontinuous.get(name, False))
        self.dof_shift = 0

    def build_structures(self, physical_tags, bc_physical_tags):
        for name, sp in self.spaces.items():
            sp.build_structures(physical_tags.get(name, [None]), bc_physical_tags.get(name, []))
        self.n_dofs = {n: s.n_dof for n, s in self.spaces.items()}
        self.n_dof = sum(self.n_dofs.values())
        self.strides = dict(zip(self.names, np.cumsum([0] + list(self.n_dofs.values()))[:-1].tolist()))

    def destination_indexes(self, i, field=None):
        out = {}
        for n, sp in self.spaces.items():
            out[n] = sp.destination_indices(sp.element_ids[i]) + self.strides[n] + self.dof_shift
        return out[field] if field else np.concatenate(list(out.values()))

    def bc_destination_indexes(self, i, field=None):
        out = {}
        for n, sp in self.spaces.items():
            if not sp.bc_element_ids:
                continue
            out[n] = sp.bc_destination_indices(sp.bc_element_ids[i]) + self.strides[n] + self.dof_shift
        return out[field] if field else np.concatenate(list(out.values()))
