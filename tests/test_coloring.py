"""Particle graph construction, coloring validity/balance and the per-world color merge.

The reference delegates the coloring itself to warp-lang's native code, whose
exact colors cannot be reproduced here; its own tests pin validity and balance
only (test_coloring.py:101-233), and so do these.  Graph construction and the
combine plan are restated exactly and checked against hand-derived answers.
"""

import numpy as np
import pytest

from newton_amd import ModelBuilder
from newton_amd.sim import graph_coloring as gc


def _colors_from_groups(n, groups):
    colors = np.full(n, -1, dtype=np.int64)
    for c, g in enumerate(groups):
        assert np.all(colors[g] == -1), "a vertex appears in two groups"
        colors[g] = c
    assert np.all(colors >= 0), "a vertex has no color"
    return colors


def test_simple_path_graph():
    """test_coloring.py:101-121."""
    edges = np.array([[0, 1], [1, 2], [2, 3]], dtype=np.int32)
    groups = gc.color_graph(4, edges, balance_colors=True, algorithm=gc.ColoringAlgorithm.MCS)
    colors = _colors_from_groups(4, groups)
    for a, b in edges:
        assert colors[a] != colors[b]


@pytest.mark.parametrize("alg", [gc.ColoringAlgorithm.MCS, gc.ColoringAlgorithm.GREEDY])
@pytest.mark.parametrize("seed", [0, 1])
def test_random_graph_valid_and_balanced(alg, seed):
    rng = np.random.default_rng(seed)
    n = 500
    edges = gc._pairs_to_graph_edges(rng.integers(0, n, size=(2000, 2)))
    edges = edges[edges[:, 0] != edges[:, 1]]
    groups = gc.color_graph(n, edges, balance_colors=True, target_max_min_color_ratio=1.1, algorithm=alg)
    colors = _colors_from_groups(n, groups)
    assert gc.validate_coloring(edges, colors)
    sizes = np.array([len(g) for g in groups])
    assert sizes.min() > 0
    for g in groups:
        assert np.all(np.diff(g) > 0), "groups are ascending vertex lists"


def test_trimesh_graph_with_bending_known_answer():
    """Two triangles (0,1,2),(0,2,3) sharing edge 0-2 with hinge o0=1, o1=3: bending adds the 1-3 pair."""
    tris = np.array([[0, 1, 2], [0, 2, 3]], dtype=np.int32)
    hinge = np.array([[1, 3, 0, 2], [2, -1, 0, 1]], dtype=np.int32)
    no_bend = gc.construct_trimesh_graph_edges(tris)
    np.testing.assert_array_equal(no_bend, [[0, 1], [0, 2], [0, 3], [1, 2], [2, 3]])
    with_bend = gc.construct_trimesh_graph_edges(tris, None, hinge, None)
    np.testing.assert_array_equal(with_bend, [[0, 1], [0, 2], [0, 3], [1, 2], [1, 3], [2, 3]])
    masked = gc.construct_trimesh_graph_edges(tris, np.array([True, False]), hinge, np.array([False, True]))
    np.testing.assert_array_equal(masked, [[0, 1], [0, 2], [1, 2]])


def test_tet_graph_edges_known_answer():
    tets = np.array([[3, 1, 2, 0]], dtype=np.int32)
    np.testing.assert_array_equal(gc.construct_tetmesh_graph_edges(tets),
                                  [[0, 1], [0, 2], [0, 3], [1, 2], [1, 3], [2, 3]])


def _color_lattice_grid(num_x, num_y):
    """Closed-form 4-coloring of test_coloring.py:80-98."""
    colors = [[] for _ in range(4)]
    for xi in range(num_x + 1):
        for yi in range(num_y + 1):
            colors[(yi % 2) * 2 + (xi % 2)].append(yi * (num_x + 1) + xi)
    return [np.array(g) for g in colors]


def test_lattice_closed_form_coloring_is_valid_for_cloth_without_bending():
    from newton_amd.sim.builder import cloth_grid_mesh

    nx, ny = 9, 6
    _, tris = cloth_grid_mesh(nx, ny, 0.1, 0.1)
    edges = gc.construct_trimesh_graph_edges(tris.reshape(-1, 3))
    colors = _colors_from_groups((nx + 1) * (ny + 1), _color_lattice_grid(nx, ny))
    assert gc.validate_coloring(edges, colors)


def test_cloth_coloring_with_bending_is_valid():
    from newton_amd import scenes

    model, _, _ = scenes.cloth_drop_on_sphere("cpu", verts_per_side=40)
    edges = gc.construct_particle_graph(model.tri_indices.numpy(), None, model.edge_indices.numpy(), None, None, None)
    colors = model.particle_colors.numpy()
    assert gc.validate_coloring(edges, colors)
    sizes = np.array([len(g) for g in model.particle_color_groups])
    assert sizes.max() / sizes.min() <= 1.1 + 1e-6


def test_tet_coloring_is_valid():
    from newton_amd import scenes

    model, _, _ = scenes.soft_cube("cpu", cells=4)
    edges = gc.construct_tetmesh_graph_edges(model.tet_indices.numpy())
    assert gc.validate_coloring(edges, model.particle_colors.numpy())


def test_combine_plan_known_answer():
    """graph_coloring.py:382-410: the longer list ascending, the other descending, merged by rank."""
    a = [(5, ["a5"]), (2, ["a2"]), (7, ["a7"])]
    b = [(1, ["b1"]), (4, ["b4"])]
    plan = gc.combine_independent_coloring_plan(a, b)
    assert plan == [(2 + 4, ["a2", "b4"]), (5 + 1, ["a5", "b1"]), (7, ["a7"])]
    # a shorter first list is swapped to second place
    plan2 = gc.combine_independent_coloring_plan(b, a)
    assert plan2 == plan
    assert gc.combine_independent_coloring_plan([], b) == b


def test_world_merge_keeps_colorings_valid():
    """Merged groups of independent worlds stay a valid coloring (test_coloring.py:236-323)."""
    from newton_amd.sim.builder import cloth_grid_mesh

    scene = ModelBuilder(up_axis="Y")
    for k, n in enumerate([5, 8, 3, 11]):
        w = ModelBuilder(up_axis="Y")
        verts, tris = cloth_grid_mesh(n, n, 0.1, 0.1)
        w.add_cloth_mesh(pos=(float(k), 0.0, 0.0), rot=(0.0, 0.0, 0.0, 1.0), scale=1.0, vel=(0.0, 0.0, 0.0),
                         vertices=verts, indices=tris, density=0.1)
        w.color(include_bending=True)
        scene.add_world(w)
    model = scene.finalize("cpu")
    edges = gc.construct_particle_graph(model.tri_indices.numpy(), None, model.edge_indices.numpy(), None, None, None)
    colors = _colors_from_groups(model.particle_count, [g.numpy() for g in model.particle_color_groups])
    assert gc.validate_coloring(edges, colors)
    np.testing.assert_array_equal(model.particle_colors.numpy(), colors)
    assert model.world_count == 4
    assert set(np.unique(model.particle_world.numpy())) == {0, 1, 2, 3}


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_graph_edges_equal_set_of_sorted_pairs(seed):
    """Canonical edges == sorted set of (min, max) pairs without negatives (graph_coloring.py:45-51)."""
    rng = np.random.default_rng(seed)
    pairs = rng.integers(-1, 40, size=(500, 2))
    expect = sorted({(min(a, b), max(a, b)) for a, b in pairs.tolist() if a >= 0 and b >= 0})
    np.testing.assert_array_equal(gc._pairs_to_graph_edges(pairs), np.array(expect, dtype=np.int32))
    hinges = rng.integers(-1, 30, size=(60, 4)).astype(np.int32)
    expect = sorted({(min(h[i], h[j]), max(h[i], h[j])) for h in hinges.tolist() for i in range(4)
                     for j in range(i + 1, 4) if h[i] >= 0 and h[j] >= 0})
    np.testing.assert_array_equal(gc.construct_trimesh_graph_edges(None, None, hinges, None),
                                  np.array(expect, dtype=np.int32).reshape(-1, 2))
