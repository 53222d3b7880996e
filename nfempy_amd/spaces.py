"""Function spaces, DoF maps and destination indices (host side, vectorised).

Drop-in mirror of ``DiscreteSpace`` (``src/spaces/discrete_space.py:14-221``),
``DoFMap`` (``src/spaces/dof_map.py:5-170``) and ``ProductSpace``
(``src/spaces/product_space.py:9-201``):

- ``ProductSpace({name: (dim, n_comp, family, k, mesh)})``,
  ``make_subspaces_discontinuous({name: bool})``,
  ``build_structures(physical_tags, b_physical_tags)``,
  ``destination_indexes(i, field=None)``, ``bc_destination_indexes(i, field=None)``,
  ``n_dof``, ``discrete_spaces``, ``discrete_spaces_dofs``, ``quadrature``,
  ``bc_quadrature``, ``dof_shift``.

DoF rules: per dimension d the DoFs of d-entities are contiguous chunks of
``e_dofs * n_comp`` ids in first-appearance entity order, dimensions stacked
vertex -> edge -> face -> volume, ``e_dofs = int(mean(num_entity_dofs[d]))``
(``dof_map.py:38-129``); element destination = chunks of the cell's sub-entities
in local order (``dof_map.py:134-146``); fields stacked in dict order
(``product_space.py:147-162``); quadrature degree ``2*max(k)+1`` shared by all
fields (``product_space.py:23-50``).

Beyond the reference, every space also holds the whole-mesh arrays the device
kernels consume: ``dest`` (E, n_loc) destination indices in the element-local
(transformed) order, and per element the signed permutation of facet DoFs the
reference applies in ``permute_and_transform`` (``src/basis/permute_and_transform.py``),
as ``ref_dest`` / ``ref_sign``: the destination and sign of each REFERENCE basis
function.
"""

from __future__ import annotations

from itertools import permutations

import numpy as np

from .basis import create_element, facet_code_transform
from .mesh import TET_EDGES, TET_FACES, TRI_EDGES, Mesh
from .quadrature import make_quadrature

_POS_TO_ROT_REF = np.array([[0, 0], [0, 1], [2, 1], [1, 0], [2, 0], [1, 1]])


def orientation_codes(mesh: Mesh, D: int, cell_index: np.ndarray) -> np.ndarray:
    """(E, n_facets) uint8 facet codes of ``permute_and_transform``.

    2-D: bit 3 | bit 2 when the local edge (cell nodes at (1,2),(0,2),(0,1)) differs
    from the stored edge node order (``permute_and_transform.py:8-35``).
    3-D: compare the local face to its sorted version; the position in
    ``itertools.permutations(sorted)`` gives (rotations, reflections)
    (``:57-100``); tagged faces at local index 1 or 3 are forced to one
    reflection (``:102-105``). bit 3 = transformation applies, bits 0-1 rot,
    bit 2 ref.
    """
    tab = mesh.entities[D]
    nodes = tab.nodes[cell_index]
    if D == 2:
        edges = tab.subs[1][cell_index] - mesh.entities[1].offset
        stored = mesh.entities[1].nodes[edges]  # (E, 3, 2)
        local = nodes[:, TRI_EDGES]  # (E, 3, 2)
        rev = np.any(local != stored, axis=2)
        return np.where(rev, 8 | 4, 0).astype(np.uint8)
    faces = nodes[:, TET_FACES]  # (E, 4, 3)
    r0 = (faces[..., 0:1] > faces).sum(-1)  # rank of first vertex among the 3
    r1 = (faces[..., 1:2] > faces).sum(-1)
    r2 = (faces[..., 2:3] > faces).sum(-1)
    pos = 2 * r0 + (r1 > r2).astype(np.int64)
    rot = _POS_TO_ROT_REF[pos, 0]
    ref = _POS_TO_ROT_REF[pos, 1]
    apply = pos != 0
    fids = tab.subs[2][cell_index] - mesh.entities[2].offset
    tagged = mesh.entities[2].material[fids] >= 0
    hack = tagged & np.isin(np.arange(4), [1, 3])[None, :]
    ref = np.where(hack, 1, ref)
    apply = apply | hack
    code = np.where(apply, 8 | rot | (ref << 2), 0)
    return code.astype(np.uint8)


class _ElementList:
    """Sized stand-in for the reference's list of FiniteElement objects."""

    def __init__(self, ids):
        self._ids = ids

    def __len__(self):
        return len(self._ids)

    def __getitem__(self, i):
        return int(self._ids[i])


class DiscreteSpace:
    def __init__(self, dimension, n_components, family, k_order, mesh, integration_order=0):
        self.dimension = dimension
        self.n_comp = n_components
        self.family_name = family
        self.k_order = k_order
        self.mesh = mesh
        self.integration_order = integration_order
        self.discontinuous = False
        self.name = "Unnamed"
        self.elements = _ElementList(np.zeros(0, np.int64))
        self.bc_elements = _ElementList(np.zeros(0, np.int64))
        self.element_ids = []
        self.bc_element_ids = []

    @property
    def family(self):
        return {"Lagrange": "P", "RT": "RT", "BDM": "BDM", "N1E": "N1E", "N2E": "N1E"}[self.family_name]

    def make_discontinuous(self):
        self.discontinuous = True

    @property
    def element(self):
        return create_element(self.family, self.dimension, self.k_order, self.discontinuous)

    def build_structures(self, physical_tags=(None,)):
        mesh = self.mesh
        D = self.dimension
        if mesh.entities is None:
            mesh.build_conformal_mesh()
        self.topology = mesh.topology(physical_tags, D)
        el = self.element
        self.e_dofs = [int(np.mean(n)) for n in el.num_entity_dofs] + [0] * (4 - len(el.num_entity_dofs))
        n_ent = [len(self.topology.entity_ids[d]) if d <= D else 0 for d in range(4)]
        self.offsets = np.concatenate([[0], np.cumsum([n_ent[d] * self.e_dofs[d] * self.n_comp for d in range(4)])])
        self.n_dof = int(self.offsets[-1])
        self.element_ids = self.topology.cells
        self.elements = _ElementList(self.element_ids)
        self.id_to_element = {int(c): i for i, c in enumerate(self.element_ids)} if len(self.element_ids) < 200000 else None
        # local destinations (E, n_loc) in the reference's local order
        parts = []
        for d in range(D + 1):
            size = self.e_dofs[d] * self.n_comp
            if size == 0:
                continue
            pos = self.topology.sub_positions(d)  # (E, nsub)
            base = self.offsets[d] + pos * size
            parts.append((base[:, :, None] + np.arange(size)[None, None, :]).reshape(len(pos), -1))
        self.dest = np.concatenate(parts, axis=1).astype(np.int64) if parts else np.zeros((len(self.element_ids), 0), np.int64)
        self._build_reference_order()

    def _build_reference_order(self):
        """ref_dest/ref_sign: destination and sign of each reference basis function."""
        el = self.element
        D = self.dimension
        E = len(self.element_ids)
        nb = el.ndofs
        nc = self.n_comp
        if el.dof_transformations_are_identity:
            self.codes = np.zeros((E, D + 1), np.uint8)
            src = np.tile(np.arange(nb), (E, 1))
            sgn = np.ones((E, nb))
        else:
            self.codes = orientation_codes(self.mesh, D, self.topology.cell_index)
            src = np.tile(np.arange(nb), (E, 1))
            sgn = np.ones((E, nb))
            for f in range(D + 1):
                dofs = np.array(el.entity_dofs[D - 1][f])
                if len(dofs) == 0:
                    continue
                for code in np.unique(self.codes[:, f]):
                    T = facet_code_transform(el, int(code))
                    rows = np.flatnonzero(self.codes[:, f] == code)
                    # phi~_i = sum_a T[i, a] phi^_a with one nonzero per row
                    a = np.argmax(np.abs(T), axis=1)
                    s = T[np.arange(len(a)), a]
                    src[np.ix_(rows, dofs)] = dofs[a][None, :]
                    sgn[np.ix_(rows, dofs)] = s[None, :]
            if D == 3 and any(len(e) for e in el.entity_dofs[1]):
                # edge DoFs (N1E) after the face transformations: a local edge whose cell node
                # order is not ascending is reflected (``permute_and_transform.py:37-55``,
                # ``:129-139``)
                nodes = self.mesh.entities[3].nodes[self.topology.cell_index]
                refl = el.entity_transformations()["interval"][0]
                a = np.argmax(np.abs(refl), axis=1)
                sg = refl[np.arange(len(a)), a]
                for k, (i, j) in enumerate(TET_EDGES):
                    dofs = np.array(el.entity_dofs[1][k])
                    rows = np.flatnonzero(nodes[:, i] > nodes[:, j])
                    src[np.ix_(rows, dofs)] = dofs[a][None, :]
                    sgn[np.ix_(rows, dofs)] = sg[None, :]
        # local dof i (basis i, comp c) -> reference basis src[i]
        self.local_src = src
        self.local_sign = sgn
        ref_dest = np.empty((E, nb, nc), np.int64)
        ref_sign = np.empty((E, nb), np.float64)
        dest = self.dest.reshape(E, nb, nc)
        rows = np.arange(E)[:, None]
        ref_dest[rows, src] = dest
        ref_sign[rows, src] = sgn
        self.ref_dest = ref_dest.reshape(E, nb * nc)
        self.ref_sign = np.repeat(ref_sign, nc, axis=1)

    def build_boundary_structures(self, bc_physical_tags=()):
        if not bc_physical_tags:
            self.bc_element_ids = np.zeros(0, np.int64)
            self.bc_elements = _ElementList(self.bc_element_ids)
            return
        D = self.dimension
        mesh = self.mesh
        ids = self.topology.entity_ids[D - 1]
        mat = mesh.entities[D - 1].material[ids - mesh.entities[D - 1].offset]
        keep = np.isin(mat, np.array(list(bc_physical_tags), dtype=np.int64))
        self.bc_element_ids = ids[keep]
        self.bc_elements = _ElementList(self.bc_element_ids)
        self.bc_neighbour = self.topology.first_cell[D - 1][keep]  # element index
        # local facet index of each bc facet in its neighbour
        subs = mesh.entities[D].subs[D - 1][self.topology.cell_index[self.bc_neighbour]]
        self.bc_local = np.argmax(subs == self.bc_element_ids[:, None], axis=1)
        el = self.element
        nc = self.n_comp
        fd = len(el.entity_dofs[D - 1][0]) if D - 1 < len(el.entity_dofs) else 0
        # bc destination: chunk of the facet itself (dimensions below carry no DoF for RT/BDM)
        size = self.e_dofs[D - 1] * nc
        pos = self.topology.position[D - 1][self.bc_element_ids - mesh.entities[D - 1].offset]
        self.bc_dest = (self.offsets[D - 1] + pos[:, None] * size + np.arange(size)[None, :]).astype(np.int64)
        # reference order on the facet: reference facet dof a -> local j with src = a
        F = len(self.bc_element_ids)
        if F and fd:
            ent = np.array([el.entity_dofs[D - 1][f] for f in range(D + 1)])  # (nf, fd)
            loc_dofs = ent[self.bc_local]  # (F, fd) local dof indices of the facet
            src = self.local_src[self.bc_neighbour[:, None], loc_dofs]  # reference basis per local j
            sgn = self.local_sign[self.bc_neighbour[:, None], loc_dofs]
            a_of_j = np.argsort(np.argsort(src, axis=1), axis=1)  # rank of src within the facet block
            bc_ref_dest = np.empty((F, fd, nc), np.int64)
            bc_ref_sign = np.empty((F, fd))
            rows = np.arange(F)[:, None]
            bd = self.bc_dest.reshape(F, fd, nc)
            bc_ref_dest[rows, a_of_j] = bd
            bc_ref_sign[rows, a_of_j] = sgn
            self.bc_ref_dest = bc_ref_dest.reshape(F, fd * nc)
            self.bc_ref_sign = np.repeat(bc_ref_sign, nc, axis=1)
        else:
            self.bc_ref_dest = np.zeros((F, 0), np.int64)
            self.bc_ref_sign = np.zeros((F, 0))


class ProductSpace:
    """``ProductSpace(discrete_spaces_data)`` (``src/spaces/product_space.py:9``)."""

    def __init__(self, discrete_spaces_data):
        self.names = list(discrete_spaces_data.keys())
        ks = [v[3] for v in discrete_spaces_data.values()]
        dims = [v[0] for v in discrete_spaces_data.values()]
        self.integration_order = 2 * int(np.max(ks)) + 1
        self.quadrature = [None] * 4
        for d in range(int(np.max(dims)) + 1):
            self.quadrature[d] = (np.array([1.0]), np.array([1.0])) if d == 0 else make_quadrature(d, self.integration_order)
        self.bc_quadrature = [None] * 4
        for d in dims:
            dd = d - 1
            self.bc_quadrature[dd] = (np.array([1.0]), np.array([1.0])) if dd == 0 else make_quadrature(dd, self.integration_order)
        self.discrete_spaces = {}
        for name, (dim, nc, fam, k, mesh) in discrete_spaces_data.items():
            sp = DiscreteSpace(dim, nc, fam, k, mesh, integration_order=self.integration_order)
            sp.name = name
            self.discrete_spaces[name] = sp
        self._dof_shift = 0

    @property
    def dof_shift(self):
        return self._dof_shift

    @dof_shift.setter
    def dof_shift(self, value):
        if not isinstance(value, (int, np.integer)):
            raise TypeError("dof_shift must be an integer")
        if value < 0:
            raise ValueError("dof_shift must be non-negative")
        self._dof_shift = int(value)

    def dimension(self):
        return max(s.dimension for s in self.discrete_spaces.values())

    def make_subspaces_discontinuous(self, discrete_space_disc):
        for name in self.names:
            if discrete_space_disc.get(name, False):
                self.discrete_spaces[name].make_discontinuous()

    def build_structures(self, physical_tags, b_physical_tags):
        self.fields_physical_tags = physical_tags
        for name, sp in self.discrete_spaces.items():
            sp.build_structures(physical_tags.get(name, [None]))
            sp.build_boundary_structures(b_physical_tags.get(name, []))
        self.discrete_spaces_dofs = {n: s.n_dof for n, s in self.discrete_spaces.items()}
        self.n_dof = int(sum(self.discrete_spaces_dofs.values()))
        self.strides = dict(zip(self.names, np.cumsum([0] + list(self.discrete_spaces_dofs.values()))[:-1].tolist()))

    def _shift(self, name):
        return self.strides[name] + self._dof_shift

    def discrete_spaces_destination_indexes(self, cell_index):
        return {n: s.dest[cell_index] + self._shift(n) for n, s in self.discrete_spaces.items()}

    def destination_indexes(self, cell_index, field_name=None):
        d = self.discrete_spaces_destination_indexes(cell_index)
        return d[field_name] if field_name else np.concatenate(list(d.values()))

    def discrete_spaces_bc_destination_indexes(self, cell_index):
        return {n: s.bc_dest[cell_index] + self._shift(n) for n, s in self.discrete_spaces.items()
                if len(s.bc_element_ids)}

    def bc_destination_indexes(self, cell_index, field_name=None):
        d = self.discrete_spaces_bc_destination_indexes(cell_index)
        return d[field_name] if field_name else np.concatenate(list(d.values()))

    # -- whole-mesh arrays --------------------------------------------------------
    def all_destination_indexes(self):
        """(E, n_loc) destinations of every element, local order."""
        return np.concatenate([s.dest + self._shift(n) for n, s in self.discrete_spaces.items()], axis=1)

    def reference_destinations(self):
        """(E, n_loc) destinations / signs in reference-basis order, fields stacked."""
        dest = np.concatenate([s.ref_dest + self._shift(n) for n, s in self.discrete_spaces.items()], axis=1)
        sign = np.concatenate([s.ref_sign for s in self.discrete_spaces.values()], axis=1)
        return dest, sign
