"""GPU parity with springs: libvbd_hip against the FP32 CPU oracle, bit-exact.

Springs (accumulate_spring_force_and_hessian, particle_vbd_kernels.py:1733-1783) are
added to a vertex after its contacts, in spring-index order, evaluated at the
positions of the color pass start.  The cases cover springs whose endpoints have
different colors (evaluated inside the color pass), springs whose endpoints share
a color (evaluated by the pre-pass), springs on a vertex handled by the generic
kernel, springs together with the serial contact mode, both work-group shapes, and
the force residual.
"""

import math

import numpy as np
import pytest

from _sim import gpu_run, initial, oracle_run

pytestmark = pytest.mark.gpu


def _color_of(model):
    col = np.full(model.particle_count, -1, np.int64)
    for c, g in enumerate(model.particle_color_groups):
        col[g.cpu().numpy()] = c
    return col


def _cloth_with_springs(dev, n=11, extra=40, seed=3, planes=1, radius=0.02, kd=1.0e-2):
    """An n x n cloth with a spring on every mesh edge and hinge, plus `extra` random springs."""
    from newton_amd import ModelBuilder
    from newton_amd.sim.builder import cloth_grid_mesh, quat_from_axis_angle

    b = ModelBuilder(up_axis="Y", gravity=-9.81)
    for k in range(planes):
        b.add_ground_plane(cfg=ModelBuilder.ShapeConfig(ke=1.0e3 + 10.0 * k, kd=0.5, mu=0.3 + 0.01 * k))
    verts, tris = cloth_grid_mesh(n - 1, n - 1, 0.04, 0.04)
    b.add_cloth_mesh(pos=(0.0, 0.05, 0.0), rot=quat_from_axis_angle((1.0, 0.0, 0.0), math.pi / 2), scale=1.0,
                     vel=(0.0, -0.3, 0.0), vertices=verts, indices=tris, density=0.1, tri_ke=1.0e3, tri_ka=1.0e3,
                     tri_kd=1.0e-4, edge_ke=0.1, edge_kd=1.0e-4, particle_radius=radius, add_springs=True,
                     spring_ke=50.0, spring_kd=kd)
    rng = np.random.default_rng(seed)
    nv = b.particle_count
    i = rng.integers(0, nv, extra)
    j = (i + rng.integers(1, nv, extra)) % nv
    b.add_springs(i, j, rng.uniform(10.0, 80.0, extra), rng.uniform(0.0, 2e-2, extra))
    b.particle_mass[0] = 0.0
    b.particle_flags[0] = 0
    b.color(include_bending=True)
    model = b.finalize(dev)
    model.soft_contact_ke, model.soft_contact_kd, model.soft_contact_mu = 1.0e3, 0.5, 0.5
    state = model.state()
    info = {"dt": 1.0 / 600.0, "substeps": 10, "iterations": 3}
    return model, state, info


def _same_color_springs(model):
    col = _color_of(model)
    idx = model.spring_indices.cpu().numpy().reshape(-1, 2)
    return int(np.sum(col[idx[:, 0]] == col[idx[:, 1]]))


@pytest.mark.parametrize("block", [256, 128])
def test_cloth_springs_bitexact(cuda_device, block):
    model, state, info = _cloth_with_springs(cuda_device)
    assert model.spring_count > 300
    assert _same_color_springs(model) > 0
    q0, qd0 = initial(state)
    q_gpu, qd_gpu, _ = gpu_run(model, state, info, 12, block_shape=block)
    q_ref, qd_ref, _ = oracle_run(model, q0, qd0, info, 12)
    np.testing.assert_array_equal(q_gpu, q_ref)
    np.testing.assert_array_equal(qd_gpu, qd_ref)


def test_springs_with_serial_contacts_bitexact(cuda_device):
    """Forty coincident planes put the batches past the parallel contact rows (about 30 items and 40
    contacts per vertex): the serial path adds the springs after the contacts."""
    model, state, info = _cloth_with_springs(cuda_device, n=16, extra=60, planes=40, radius=0.06)
    q0, qd0 = initial(state)
    q_gpu, qd_gpu, _ = gpu_run(model, state, info, 4)
    q_ref, qd_ref, _ = oracle_run(model, q0, qd0, info, 4)
    np.testing.assert_array_equal(q_gpu, q_ref)
    np.testing.assert_array_equal(qd_gpu, qd_ref)


def _fan_with_springs(dev, segments=160):
    """A triangle fan whose centre (> 256 items) has springs to every rim vertex and to a free particle
    placed in the centre's color, so the generic kernel sees both in-pass and pre-pass springs."""
    from newton_amd import ModelBuilder

    b = ModelBuilder(up_axis="Y", gravity=-9.81)
    b.add_ground_plane(cfg=ModelBuilder.ShapeConfig(ke=1.0e4, kd=0.0, mu=0.5))
    verts = [(0.0, 0.0, 0.0)] + [(0.5 * math.cos(2 * math.pi * i / segments), 0.5 * math.sin(2 * math.pi * i / segments),
                                  0.0) for i in range(segments)]
    tris = []
    for i in range(segments):
        tris += [0, 1 + i, 1 + (i + 1) % segments]
    b.add_cloth_mesh(pos=(0.0, 0.3, 0.0), rot=(0.7071068, 0.0, 0.0, 0.7071068), scale=1.0, vel=(0.0, 0.0, 0.0),
                     vertices=verts, indices=tris, density=0.1, tri_ke=1.0e3, tri_ka=1.0e3, tri_kd=1.0e-3,
                     edge_ke=0.1, edge_kd=1.0e-4, particle_radius=0.02)
    free = b.add_particle((0.05, 0.45, 0.02), (0.0, 0.0, 0.0), 0.01, radius=0.02)
    b.add_springs(np.zeros(segments, np.int64), np.arange(1, segments + 1), 20.0, 1.0e-3)
    b.add_spring(0, free, 40.0, 5.0e-3)
    b.add_spring(free, 0, 30.0, 0.0)
    b.color(include_bending=True)
    groups = [np.asarray(g) for g in b.particle_color_groups]
    groups = [g[g != free] for g in groups]
    c0 = next(k for k, g in enumerate(groups) if 0 in g)
    groups[c0] = np.concatenate([groups[c0], [free]])
    b.set_coloring([g for g in groups if len(g)])
    model = b.finalize(dev)
    model.soft_contact_ke, model.soft_contact_kd, model.soft_contact_mu = 1.0e4, 0.0, 0.5
    state = model.state()
    info = {"dt": 1.0 / 600.0, "substeps": 10, "iterations": 4}
    return model, state, info


def test_generic_vertex_springs_bitexact(cuda_device):
    from newton_amd.solvers import SolverVBD

    model, state, info = _fan_with_springs(cuda_device)
    assert SolverVBD(model, iterations=1).layout_stats()["generic_vertices"] == 1
    assert _same_color_springs(model) == 2
    q0, qd0 = initial(state)
    q_gpu, qd_gpu, _ = gpu_run(model, state, info, 8)
    q_ref, qd_ref, _ = oracle_run(model, q0, qd0, info, 8)
    np.testing.assert_array_equal(q_gpu, q_ref)
    np.testing.assert_array_equal(qd_gpu, qd_ref)


def test_spring_residual_matches_oracle(cuda_device):
    """The force residual (mode-1 pass over every color) with springs, contacts and damping."""
    import torch

    import oracle
    from newton_amd import CollisionPipeline
    from newton_amd.solvers import SolverVBD

    model, state, info = _cloth_with_springs(cuda_device, n=9, extra=30)
    solver = SolverVBD(model, iterations=info["iterations"])
    pipe = CollisionPipeline(model, soft_contact_margin=0.05)
    contacts = pipe.contacts()
    pipe.collide(state, contacts)
    solver.begin_step(state, contacts, info["dt"])
    solver.iterate(state, info["dt"], 1)
    g = torch.zeros_like(state.particle_q)
    r2 = solver.residual(state, info["dt"], g)
    torch.cuda.synchronize()
    om = oracle.OracleModel(model)
    oc = oracle.Contacts.from_device(contacts)
    g_ref, n_ref = oracle.residual(om, oracle.params(info["iterations"]), oc, state.particle_q.cpu().numpy(),
                                   solver.scratch("q_prev").cpu().numpy(), solver.scratch("inertia").cpu().numpy(),
                                   info["dt"], precision="f32")
    assert np.any(g_ref != 0.0)
    np.testing.assert_array_equal(g.cpu().numpy(), g_ref)
    assert abs(float(r2.item()) ** 0.5 - n_ref) <= 1e-6 * max(n_ref, 1.0)


def test_spring_only_rope_bitexact(cuda_device):
    """Particles joined only by springs (no triangles, edges or tets): the coloring graph has no edges,
    so every spring joins two vertices of one color and goes through the pre-pass."""
    from newton_amd import ModelBuilder

    b = ModelBuilder(up_axis="Y", gravity=-9.81)
    b.add_ground_plane(cfg=ModelBuilder.ShapeConfig(ke=1.0e4, kd=1.0, mu=0.4))
    n = 40
    for k in range(n):
        b.add_particle((0.02 * k, 0.3 + 0.01 * math.sin(k), 0.0), (0.0, 0.0, 0.0), 0.0 if k == 0 else 0.01,
                       radius=0.01)
    b.add_springs(np.arange(n - 1), np.arange(1, n), 200.0, 0.05)
    b.add_springs(np.arange(n - 2), np.arange(2, n), 50.0, 0.0)  # bending-like second neighbours
    b.color()
    model = b.finalize(cuda_device)
    model.soft_contact_ke, model.soft_contact_kd, model.soft_contact_mu = 1.0e4, 1.0, 0.4
    assert model.tri_count == 0 and model.edge_count == 0 and model.spring_count == 2 * n - 3
    assert _same_color_springs(model) > 0
    state = model.state()
    info = {"dt": 1.0 / 600.0, "substeps": 10, "iterations": 5}
    q0, qd0 = initial(state)
    q_gpu, qd_gpu, _ = gpu_run(model, state, info, 60)
    q_ref, qd_ref, _ = oracle_run(model, q0, qd0, info, 60)
    np.testing.assert_array_equal(q_gpu, q_ref)
    np.testing.assert_array_equal(qd_gpu, qd_ref)
    assert np.abs(q_gpu - q0).max() > 0.01
