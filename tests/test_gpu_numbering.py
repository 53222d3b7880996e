"""Device sorts of the cold path (``nfx_first_appearance``, ``nfx_stable_sort64``) give the
host path's integers bit for bit: first-appearance ranks and first positions on random keys
with repeats (62-bit keys, one key, all equal, all distinct), the stable argsort, the mesh
builder's entities (stored node orders, sub-entity ids, tags) and the tag-filtered entity
orders on the structured cube and the reference's gmsh meshes, and the Darcy and SBAIJ
elasticity patterns with their slot codes."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _host_first(keys):
    from nfempy_amd.mesh import _first_appearance_ids

    return _first_appearance_ids(keys)


@pytest.mark.parametrize("case", ["random", "wide", "one", "equal", "distinct"])
def test_first_appearance_and_sort(case):
    from nfempy_amd.numbering import first_appearance_ids, stable_argsort

    rng = np.random.default_rng(3)
    keys = {"random": rng.integers(0, 5000, 200_000), "wide": rng.integers(0, 1 << 62, 50_000) // 7 * 7,
            "one": np.array([42]), "equal": np.full(1000, 9), "distinct": rng.permutation(300_000)}[case]
    keys = keys.astype(np.int64)
    if case == "wide":
        keys = np.concatenate([keys, keys[::3]])
    print(f"[numbering] {case} K={len(keys)}", flush=True)
    ids, first = first_appearance_ids(keys)
    hids, hfirst = _host_first(keys)
    assert np.array_equal(ids, hids) and np.array_equal(first, hfirst)
    order, sk = stable_argsort(keys)
    assert np.array_equal(order, np.argsort(keys, kind="stable")) and np.array_equal(sk, np.sort(keys))


def _mesh_pair(kind):
    from nfempy_amd.mesh import Mesh, unit_cube, unit_square

    if kind == "cube":
        args = (3, *unit_cube(6))
        return Mesh.from_blocks(*args), Mesh.from_blocks(*args, device="cuda"), [1]
    if kind == "square":
        args = (2, *unit_square(5))
        return Mesh.from_blocks(*args), Mesh.from_blocks(*args, device="cuda"), [1]
    from pathlib import Path

    f = str(Path(__file__).parent / "golden" / "meshes" / kind)
    d = 3 if "3d" in kind else 2
    h = Mesh(d, f)
    h.build_conformal_mesh()
    g = Mesh(d, f, device="cuda")
    g.build_conformal_mesh()
    return h, g, [1]


@pytest.mark.parametrize("kind", ["cube", "square", "example_3_2d_l_1.msh", "example_3_3d_l_0.msh"])
def test_mesh_and_topology_on_device(kind):
    print(f"[numbering] mesh {kind}", flush=True)
    h, g, tags = _mesh_pair(kind)
    for d in range(4):
        a, b = h.entities[d], g.entities[d]
        assert a.offset == b.offset and np.array_equal(a.nodes, b.nodes) and np.array_equal(a.material, b.material)
        assert sorted(a.subs) == sorted(b.subs)
        for k in a.subs:
            assert np.array_equal(a.subs[k], b.subs[k]), (d, k)
    th, tg = h.topology(tags), g.topology(tags)
    for d in th.entity_ids:
        assert np.array_equal(th.entity_ids[d], tg.entity_ids[d])
        assert np.array_equal(th.first_cell[d], tg.first_cell[d])


@pytest.mark.parametrize("nk", [(4, 1), (3, 2)])
def test_darcy_pattern_on_device(nk):
    from nfempy_amd.batch import darcy_space
    from nfempy_amd.pattern import build_pattern

    print(f"[numbering] darcy pattern {nk}", flush=True)
    ps = darcy_space(*nk)
    rd, rs = ps.reference_destinations()
    a = build_pattern(ps.n_dof, rd, rs)
    b = build_pattern(ps.n_dof, rd, rs, device="cuda")
    for f in ("rowptr", "col", "slot", "segptr", "target"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f


def test_sbaij_pattern_on_device():
    from nfempy_amd.pattern import build_pattern

    print("[numbering] sbaij pattern", flush=True)
    rng = np.random.default_rng(5)
    E, nl, n = 500, 12, 900
    rd = np.stack([rng.choice(n, nl, replace=False) for _ in range(E)])
    rs = rng.choice([-1, 1], (E, nl))
    keep = rng.random((E, nl, nl)) > 0.2
    a = build_pattern(n, rd, rs, keep=keep, zero_diag_dest=rd, upper=True)
    b = build_pattern(n, rd, rs, keep=keep, zero_diag_dest=rd, upper=True, device="cuda")
    for f in ("rowptr", "col", "slot", "segptr", "target"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
