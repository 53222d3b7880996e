"""Point values of discrete fields for the VTK output (FP64 oracle; test infrastructure only).

Restates ``node_average_quatity`` and ``cell_centered_quatity``
(``src/postprocess/solution_post_processor.py:11-103``): the field of one element at a
reference point (a vertex of the reference cell, or its centroid) with the pushed-forward,
permuted-and-transformed basis of the weak forms; Lagrange fields keep
``alpha_star[:, 0:dim]``, vector fields the first ``dim`` components of each row; a
point's value is the mean over the field's elements that have it as a vertex, in element
order.

Pinned by: a linear P1 field gives its nodal values at the points and its centroid value
at the cells; a constant vector in RT1 gives the constant at every point and cell; BDM1
rows of a linear field give the field's values.
"""

from __future__ import annotations

import numpy as np

from .forms import _cell, _tabulate
from .geometry import evaluate_mapping

REF_VERTICES = {2: np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]]),
                3: np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])}


def element_values(ps, name, alpha, i, ref_points):
    """(n_points, n_data) field values of element i at reference points."""
    sp = ps.spaces[name]
    cell = _cell(sp, i)
    dim = cell.dimension
    x, jac, det, _ = evaluate_mapping(dim, ref_points, sp.mesh.points[cell.node_tags])
    tab = _tabulate(sp, cell, ref_points, jac, det)
    a = np.array(np.split(alpha[ps.destination_indexes(i, name)], tab.shape[2]))
    out = []
    for q in range(len(ref_points)):
        if sp.family == "P":
            out.append((tab[0, q:q + 1, :, 0] @ a[:, 0:dim]).T.ravel())
        else:
            out.append(np.vstack([tab[0, q, :, 0:dim].T @ a[:, c] for c in range(sp.n_comp)]).ravel())
    return np.array(out)


def node_average(ps, name, alpha, n_points):
    sp = ps.spaces[name]
    dim = sp.dimension
    acc = {}
    for i in range(len(sp.element_ids)):
        cell = _cell(sp, i)
        vals = element_values(ps, name, alpha, i, REF_VERTICES[dim])
        for a, node in enumerate(cell.node_tags):
            acc.setdefault(int(node), []).append(vals[a])
    n_data = len(next(iter(acc.values()))[0]) if acc else 0
    out = np.zeros((n_points, n_data))
    for node, v in acc.items():
        out[node] = np.mean(np.array(v), axis=0)
    return out


def cell_centered(ps, name, alpha):
    sp = ps.spaces[name]
    dim = sp.dimension
    c = REF_VERTICES[dim].mean(axis=0, keepdims=True)
    return np.array([element_values(ps, name, alpha, i, c)[0] for i in range(len(sp.element_ids))])
