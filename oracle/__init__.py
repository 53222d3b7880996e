"""FP64 CPU oracle for the nfempy assembly-and-solve path.

TEST INFRASTRUCTURE ONLY. Nothing in the ``nfempy_amd`` package imports this
package. Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` may import it, and only as the checker (or, for
``cpu_baseline``, as the timed CPU restatement of the reference algorithm).

The oracle restates, in numpy/scipy/sympy at float64, the reference files of
OmarDuran/nfempy that define the hot path:

- ``src/mesh/mesh.py`` (entity numbering, physical-tag graphs),
- ``src/mesh/mesh_topology.py``, ``src/spaces/{dof_map,product_space,discrete_space}.py``,
- ``src/basis/{finite_element,permute_and_transform,parametric_transformation}.py``,
- ``src/geometry/{mapping,compute_normal}.py``,
- ``src/weak_forms/laplace_dual_weak_form.py``,
  ``src/weak_forms/le_dual_weak_form.py`` (``evaluate_form_vectorized``),
  ``src/weak_forms/le_riesz_map_weak_form.py`` (``evaluate_form_vectorized``),
- ``src/assembly/ScatterFormData.py`` and the driver scatter closures of
  ``examples/md_elliptic_scalar/md_elliptic_scalar.py`` and
  ``examples/cauchy_elasticity/mfe_elasticity_example_1.py``,
- ``src/postprocess/l2_error_post_processor.py``.

Third-party arithmetic that the reference delegates to un-vendored packages
(basix element tables and quadrature, PETSc/MUMPS solves, AutoDiff Jacobians)
is restated from its published definitions; each module says how it is pinned.
"""
