"""FP64 error norms (oracle).

Restates ``l2_error`` (``src/postprocess/l2_error_post_processor.py:7-63``) and
``div_error`` (``:185-237``): per field, per element,
``sum_q detJ w |f_exact(x_q) - f_h(x_q)|^2`` with the same basis evaluation as the
weak forms; vector fields use the Frobenius norm over (component, direction);
``div f_h = tr(grad_ref phi) / detJ``.

Pinned by: zero error when the exact field lies in the discrete space (lowest-order
interpolants), and the O(h) convergence of the RT1-P0 Darcy solution.
"""

from __future__ import annotations

import numpy as np

from .forms import _cell, _tabulate
from .geometry import evaluate_mapping


def l2_error(ps, functions, alpha, skip_fields=()):
    out = []
    for name, sp in ps.spaces.items():
        if name in skip_fields:
            continue
        dim = sp.dimension
        points, weights = ps.quadrature[dim]
        err = 0.0
        exact = functions[name]
        for i in range(len(sp.element_ids)):
            cell = _cell(sp, i)
            x, jac, det, _ = evaluate_mapping(dim, points, sp.mesh.points[cell.node_tags])
            tab = _tabulate(sp, cell, points, jac, det)
            dest = ps.destination_indexes(i, name)
            al = alpha[dest]
            nphi = tab.shape[2]
            a = np.array(np.split(al, nphi))
            if sp.family == "P":
                fe = np.asarray(exact(x[:, 0], x[:, 1], x[:, 2]), dtype=np.float64).reshape(-1, len(weights))[0:dim, :]
                fh = (tab[0, :, :, 0] @ a[:, 0:dim]).T
                err += np.sum(det * weights * (fe - fh) * (fe - fh))
            else:
                fe = np.array([np.asarray(exact(xv[0], xv[1], xv[2]), dtype=np.float64) for xv in x])[:, :, 0:dim]
                fh = np.array([
                    np.vstack([tab[0, k, :, 0:dim].T @ a[:, c] for c in range(sp.n_comp)])
                    for k in range(len(points))
                ])
                d = fe - fh
                err += np.sum(det * weights * np.array([np.trace(e.T @ e) for e in d]))
        out.append(np.sqrt(err))
    return out


def div_error(ps, functions, alpha, skip_fields=()):
    out = []
    for name, sp in ps.spaces.items():
        if sp.family not in ("RT", "BDM") or name in skip_fields:
            continue
        dim = sp.dimension
        points, weights = ps.quadrature[dim]
        exact = functions["div_" + name]
        err = 0.0
        for i in range(len(sp.element_ids)):
            cell = _cell(sp, i)
            x, jac, det, _ = evaluate_mapping(dim, points, sp.mesh.points[cell.node_tags])
            tab = _tabulate(sp, cell, points, jac, det)
            dest = ps.destination_indexes(i, name)
            a = np.array(np.split(alpha[dest], tab.shape[2]))
            divv = np.trace(tab[1:, :, :, 0:dim], axis1=0, axis2=3).T * (1.0 / det)
            fe = np.array([np.asarray(exact(xv[0], xv[1], xv[2]), dtype=np.float64) for xv in x])
            fh = divv.T @ a
            err += np.sum(((fe.reshape(fh.shape) - fh) ** 2).T @ (det * weights))
        out.append(np.sqrt(err))
    return out
