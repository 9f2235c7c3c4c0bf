"""ModelBuilder subset: element pre-computations and model layout (builder.py restatement)."""

import math

import numpy as np
import pytest

from newton_amd import ModelBuilder, ParticleFlags, scenes
from newton_amd.sim.builder import cloth_grid_mesh, quat_between_vectors, quat_rotate


def test_cloth_grid_mesh_topology():
    """add_cloth_grid triangulation (builder.py:8999-9018): quad (v0,v1,v2,v3) -> [v0,v1,v3], [v1,v2,v3]."""
    verts, tris = cloth_grid_mesh(2, 1, 0.5, 0.25)
    np.testing.assert_array_equal(tris.reshape(-1, 3), [[0, 1, 3], [1, 4, 3], [1, 2, 4], [2, 5, 4]])
    np.testing.assert_allclose(verts[4], [0.5, 0.25, 0.0])
    _, rtris = cloth_grid_mesh(1, 1, 1.0, 1.0, reverse_winding=True)
    np.testing.assert_array_equal(rtris.reshape(-1, 3), [[0, 1, 3], [0, 3, 2]])


def test_lumped_mass_and_total_mass():
    model, _, _ = scenes.cloth_hanging("cpu")
    mass = model.particle_mass.numpy().astype(np.float64)
    flags = model.particle_flags.numpy()
    n = 31
    cell = 1.0 / n
    # interior vertex: 6 triangles of area cell^2/2, each gives density*area/3
    assert mass[16 * 32 + 16] == pytest.approx(0.1 * 6 * 0.5 * cell * cell / 3.0, rel=1e-6)
    # the diagonal runs (x, y-1)-(x-1, y): corner (0, 0) touches 1 triangle, corner (31, 0) touches 2
    assert mass[0] == pytest.approx(0.1 * 1 * 0.5 * cell * cell / 3.0, rel=1e-6)
    assert mass[31] == pytest.approx(0.1 * 2 * 0.5 * cell * cell / 3.0, rel=1e-6)
    pinned = np.where(flags & int(ParticleFlags.ACTIVE) == 0)[0]
    np.testing.assert_array_equal(pinned, [31 * 32, 31 * 32 + 31])
    assert np.all(mass[pinned] == 0.0) and np.all(model.particle_inv_mass.numpy()[pinned] == 0.0)
    assert mass.sum() + 0.1 * 0.5 * cell * cell == pytest.approx(0.1 * 1.0, rel=1e-5)


def test_rest_pose_gives_identity_metric():
    """tri_poses is the inverse 2-D rest basis: F^T F = I at rest, and the area matches."""
    model, _, _ = scenes.cloth_drop_on_sphere("cpu", verts_per_side=8)
    q = model.particle_q.numpy().astype(np.float64)
    tri = model.tri_indices.numpy()
    pose = model.tri_poses.numpy().astype(np.float64)
    for t in range(0, model.tri_count, 7):
        x0, x1, x2 = q[tri[t]]
        F = np.stack([x1 - x0, x2 - x0], axis=1) @ pose[t]
        np.testing.assert_allclose(F.T @ F, np.eye(2), atol=1e-5)
        area = 0.5 * np.linalg.norm(np.cross(x1 - x0, x2 - x0))
        assert model.tri_areas.numpy()[t] == pytest.approx(area, rel=1e-5)


def test_flat_cloth_rest_angles_and_lengths():
    model, _, _ = scenes.cloth_drop_on_sphere("cpu", verts_per_side=6)
    edges = model.edge_indices.numpy()
    q = model.particle_q.numpy().astype(np.float64)
    interior = (edges[:, 0] >= 0) & (edges[:, 1] >= 0)
    assert np.abs(model.edge_rest_angle.numpy()[interior]).max() < 1e-6
    lengths = np.linalg.norm(q[edges[:, 3]] - q[edges[:, 2]], axis=1)
    np.testing.assert_allclose(model.edge_rest_length.numpy(), lengths, rtol=1e-6)
    np.testing.assert_allclose(model.edge_bending_properties.numpy(), np.tile([1.0, 0.0], (model.edge_count, 1)))


def test_folded_hinge_rest_angle():
    """Rest angle of a hinge folded by 90 degrees (builder.py:8838-8867)."""
    b = ModelBuilder()
    for p in ((0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.5, 1.0, 0.0), (0.5, 0.0, 1.0)):
        b.add_particle(p, (0.0, 0.0, 0.0), 1.0)
    b.add_edge(2, 3, 0, 1)
    angle = float(b._edge_rest.array()[0])
    assert abs(abs(angle) - math.pi / 2) < 1e-9


def test_add_cloth_grid_overwrites_mass_and_fixes_flags():
    b = ModelBuilder()
    b.add_cloth_grid(pos=(0.0, 0.0, 0.0), rot=(0.0, 0.0, 0.0, 1.0), vel=(0.0, 0.0, 0.0), dim_x=3, dim_y=2,
                     cell_x=0.1, cell_y=0.1, mass=0.25, fix_left=True)
    mass = b.particle_mass
    flags = b.particle_flags
    left = np.arange(0, 12, 4)
    assert np.all(mass[left] == 0.0) and np.all(flags[left] == 0)
    others = np.setdiff1d(np.arange(12), left)
    assert np.all(mass[others] == 0.25) and np.all(flags[others] == int(ParticleFlags.ACTIVE))


def test_soft_grid_tets_and_masses():
    b = ModelBuilder(up_axis="Y")
    b.add_soft_grid(pos=(0.0, 1.0, 0.0), rot=(0.0, 0.0, 0.0, 1.0), vel=(0.0, 0.0, 0.0), dim_x=2, dim_y=3, dim_z=1,
                    cell_x=0.1, cell_y=0.1, cell_z=0.1, density=1000.0, k_mu=1e4, k_lambda=2e4, k_damp=0.0)
    assert b.tet_count == 5 * 2 * 3 * 1
    assert b.particle_count == 3 * 4 * 2
    np.testing.assert_allclose(b.particle_mass, 1000.0 * 0.001)
    model = b.finalize("cpu")
    pose = model.tet_poses.numpy().astype(np.float64)
    vol = 1.0 / (6.0 * np.linalg.det(pose))
    assert np.all(vol > 0)
    assert vol.sum() == pytest.approx(0.2 * 0.3 * 0.1, rel=1e-5)
    # the surface triangles carry zero stiffness and are kept out of the coloring graph
    assert np.all(model.tri_materials.numpy()[:, :2] == 0.0)


def test_ground_plane_transform_maps_local_z_to_up():
    q = quat_between_vectors((0.0, 0.0, 1.0), (0.0, 1.0, 0.0))
    np.testing.assert_allclose(quat_rotate(q, np.array([0.0, 0.0, 1.0], np.float32)), [0.0, 1.0, 0.0], atol=1e-6)
    b = ModelBuilder(up_axis="Y")
    s = b.add_ground_plane(cfg=ModelBuilder.ShapeConfig(ke=123.0, kd=4.0, mu=0.7))
    model = b.finalize("cpu")
    assert model.shape_type.numpy()[s] == 1
    np.testing.assert_allclose(model.shape_scale.numpy()[s], [0.0, 0.0, 0.0])
    assert model.shape_material_ke.numpy()[s] == 123.0
    assert model.shape_flags.numpy()[s] & 4


def test_model_layout_dtypes_and_shapes():
    model, state, _ = scenes.cloth_drop_on_sphere("cpu", verts_per_side=5)
    n, t, e = model.particle_count, model.tri_count, model.edge_count
    expect = {
        "particle_q": ((n, 3), "float32"), "particle_qd": ((n, 3), "float32"), "particle_mass": ((n,), "float32"),
        "particle_inv_mass": ((n,), "float32"), "particle_radius": ((n,), "float32"),
        "particle_flags": ((n,), "int32"), "particle_colors": ((n,), "int32"), "tri_indices": ((t, 3), "int32"),
        "tri_poses": ((t, 2, 2), "float32"), "tri_materials": ((t, 5), "float32"), "tri_areas": ((t,), "float32"),
        "edge_indices": ((e, 4), "int32"), "edge_rest_angle": ((e,), "float32"),
        "edge_rest_length": ((e,), "float32"), "edge_bending_properties": ((e, 2), "float32"),
        "gravity": ((1, 3), "float32"),
    }
    for name, (shape, dtype) in expect.items():
        arr = getattr(model, name)
        assert tuple(arr.shape) == shape, name
        assert str(arr.dtype).endswith(dtype), name
    assert tuple(state.particle_f.shape) == (n, 3) and float(state.particle_f.abs().max()) == 0.0
    np.testing.assert_allclose(model.gravity.numpy()[0], [0.0, -9.81, 0.0], rtol=1e-6)


def test_world_color_merge_equals_sequential_combination():
    """Lazy plan merging of many worlds gives the groups of merging (and materialising) one world at a time."""
    from newton_amd.sim import graph_coloring as gc

    rng = np.random.default_rng(5)
    templates = []
    for s in (4, 6, 7):
        t = ModelBuilder(up_axis="Y")
        verts, tris = cloth_grid_mesh(s - 1, s - 1, 0.1, 0.1)
        t.add_cloth_mesh(pos=(0.0, 1.0, 0.0), rot=(0.0, 0.0, 0.0, 1.0), scale=1.0, vel=(0.0, 0.0, 0.0),
                         vertices=verts, indices=tris, density=0.1, tri_ke=1e3, tri_ka=1e3, tri_kd=0.0, edge_ke=1.0)
        t.color(include_bending=True)
        templates.append(t)
    b = ModelBuilder(up_axis="Y")
    expect = []
    off = 0
    for w in range(11):
        t = templates[int(rng.integers(0, 3))]
        b.add_world(t)
        src = [g + off for g in t.particle_color_groups]
        expect = gc.combine_independent_particle_coloring(expect, src) if expect else src
        off += t.particle_count
    got = b.particle_color_groups
    assert len(got) == len(expect)
    for g, e in zip(got, expect):
        np.testing.assert_array_equal(g, e)
    allv = np.sort(np.concatenate(got))
    np.testing.assert_array_equal(allv, np.arange(b.particle_count))


def test_cloth_batched_scene_structure():
    from newton_amd import scenes

    model, state, info = scenes.cloth_batched("cpu", worlds=12)
    sides = info["sides"]
    assert model.particle_count == int((sides.astype(np.int64) ** 2).sum())
    assert model.tri_count == int((2 * (sides - 1) ** 2).sum())
    assert model.world_count == 12 and model.shape_count == 1
    world = model.particle_world.numpy()
    ps = info["world_particle_start"]
    for w in range(12):
        assert (world[ps[w]:ps[w + 1]] == w).all()
    ke = model.tri_materials[:, 0].numpy()
    ts = info["world_tri_start"]
    for w in range(12):
        assert np.unique(ke[ts[w]:ts[w + 1]]).size == 1
        assert 5e3 <= ke[ts[w]] <= 2e4
    pinned = (model.particle_mass.numpy() == 0)
    per_world = np.add.reduceat(pinned.astype(np.int64), ps[:-1])
    assert set(per_world.tolist()) <= {0, 1, 2}
    # every world's colors are valid on its own mesh: no triangle or hinge joins two same-colored vertices
    colors = model.particle_colors.numpy()
    tri = model.tri_indices.numpy()
    for a, c in ((0, 1), (1, 2), (0, 2)):
        assert (colors[tri[:, a]] != colors[tri[:, c]]).all()


def test_soft_cubes_batched_scene_structure():
    from newton_amd import scenes

    model, state, info = scenes.soft_cubes_batched("cpu", worlds=6)
    cells = info["cells"]
    assert model.particle_count == int(((cells + 1) ** 3).sum())
    assert model.tet_count == int((5 * cells ** 3).sum())
    ts = info["world_tet_start"]
    mat = model.tet_materials.numpy()
    for w in range(6):
        mu, lam = mat[ts[w]:ts[w + 1], 0], mat[ts[w]:ts[w + 1], 1]
        assert np.unique(mu).size == 1 and np.unique(lam).size == 1
        nu = lam[0] / (2.0 * (lam[0] + mu[0]))  # Poisson ratio from the Lame parameters
        e = mu[0] * 2.0 * (1.0 + nu)
        assert 0.3 - 1e-5 <= nu <= 0.45 + 1e-5 and 1e4 * (1 - 1e-5) <= e <= 1e5 * (1 + 1e-5)
    # surface triangles carry zero material: the solve skips them
    assert model.tri_count > 0 and float(np.abs(model.tri_materials.numpy()[:, :2]).max()) == 0.0
    # a valid coloring: no tet joins two particles of one color
    colors = model.particle_colors.numpy()
    tet = model.tet_indices.numpy()
    for a in range(4):
        for b in range(a + 1, 4):
            assert (colors[tet[:, a]] != colors[tet[:, b]]).all()


def test_cloth_springs_cover_edges_and_hinges():
    """add_springs: one spring per mesh edge plus one across every interior hinge, rest = distance."""
    b = ModelBuilder(up_axis="Y")
    verts, tris = cloth_grid_mesh(3, 2, 0.1, 0.2)
    b.add_cloth_mesh(pos=(0.0, 1.0, 0.0), rot=(0.0, 0.0, 0.0, 1.0), scale=1.0, vel=(0.0, 0.0, 0.0), vertices=verts,
                     indices=tris, density=0.1, add_springs=True, spring_ke=55.0, spring_kd=0.5)
    edges = b.edge_indices
    interior = int(((edges[:, 0] >= 0) & (edges[:, 1] >= 0)).sum())
    assert b.spring_count == len(edges) + interior
    pairs = {tuple(sorted(p)) for p in b.spring_indices.tolist()}
    assert len(pairs) == b.spring_count
    for i, j, k, l in edges.tolist():
        assert (min(k, l), max(k, l)) in pairs
        if i >= 0 and j >= 0:
            assert (min(i, j), max(i, j)) in pairs
    model = b.finalize("cpu")
    q = model.particle_q.numpy()
    si = model.spring_indices.numpy().reshape(-1, 2)
    rest = np.linalg.norm(q[si[:, 0]] - q[si[:, 1]], axis=1)
    np.testing.assert_allclose(model.spring_rest_length.numpy(), rest, rtol=1e-6)
    assert (model.spring_stiffness.numpy() == 55.0).all() and (model.spring_damping.numpy() == 0.5).all()
    # per-vertex spring CSR in spring index order
    adj = model.soft_mesh_adjacency
    v = 0
    ids = adj.v_adj_springs[adj.v_adj_springs_offsets[v]:adj.v_adj_springs_offsets[v + 1]]
    np.testing.assert_array_equal(ids, np.nonzero((si == v).any(axis=1))[0])
