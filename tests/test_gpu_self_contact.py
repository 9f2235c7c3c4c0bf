"""GPU parity of cloth self-contact: libvbd_hip against the FP32 CPU oracle.

Detection (vertex-triangle and edge-edge lists, counts, minimum distances) must equal
the oracle's brute-force detection exactly; steps with self-contact (detection, per-color
contact rows, planar truncation, isotropic clamp) must give bit-identical positions and
velocities; the force residual includes the self-contact terms.
"""

import math

import numpy as np
import pytest

import oracle
from newton_amd.solvers.tri_mesh_collision import build_filters

pytestmark = pytest.mark.gpu


def _layers(dev, n=8, gap=0.012, vel=-2.0, worlds=False, ground=False):
    """Two n x n cloth layers, the upper one moving down onto the lower one."""
    from newton_amd import ModelBuilder
    from newton_amd.sim.builder import cloth_grid_mesh, quat_from_axis_angle

    def layer(b, y, v):
        verts, tris = cloth_grid_mesh(n - 1, n - 1, 0.02, 0.02)
        rot = quat_from_axis_angle((1.0, 0.0, 0.0), math.pi / 2)
        b.add_cloth_mesh(pos=(0.003 * (y > 0), y, 0.004 * (y > 0)), rot=rot, scale=1.0, vel=(0.0, v, 0.0),
                         vertices=verts, indices=tris, density=0.1, tri_ke=1e3, tri_ka=1e3, tri_kd=1e-4,
                         edge_ke=0.01, edge_kd=1e-4)

    b = ModelBuilder(up_axis="Y", gravity=-9.81 if ground else 0.0)
    if ground:
        b.add_ground_plane(cfg=ModelBuilder.ShapeConfig(ke=1.0e3, kd=0.0, mu=0.5))
    if worlds:
        for w in range(2):
            sub = ModelBuilder(up_axis="Y", gravity=0.0)
            layer(sub, 0.0 if w == 0 else gap, 0.0 if w == 0 else vel)
            sub.color(include_bending=True)
            b.add_world(sub)
    else:
        layer(b, 0.0, 0.0)
        layer(b, gap, vel)
    b.particle_mass[0] = 0.0
    b.particle_flags[0] = 0
    if not worlds:
        b.color(include_bending=True)
    model = b.finalize(dev)
    model.soft_contact_ke, model.soft_contact_kd, model.soft_contact_mu = 2.0e3, 1.0, 0.3
    return model


SC = dict(particle_self_contact_radius=0.004, particle_self_contact_margin=0.006,
          particle_vertex_contact_buffer_size=16, particle_edge_contact_buffer_size=32)


def _oracle_self(model, solver_kw, threshold=2, vt_map=None):
    vt, ee = build_filters(model, threshold, vt_map)
    return oracle.SelfContact(
        model.particle_count, model.tri_count, model.edge_count,
        radius=solver_kw["particle_self_contact_radius"], margin=solver_kw["particle_self_contact_margin"],
        relaxation=solver_kw.get("particle_conservative_bound_relaxation", 0.85),
        vt_capacity=solver_kw["particle_vertex_contact_buffer_size"],
        ee_capacity=solver_kw["particle_edge_contact_buffer_size"],
        detection_interval=solver_kw.get("particle_collision_detection_interval", 0),
        rest_exclusion_radius=solver_kw.get("particle_rest_shape_contact_exclusion_radius", 0.0),
        q_rest=model.particle_q.cpu().numpy(),
        vt_filter=None if vt is None else vt[0], vt_filter_offsets=None if vt is None else vt[1],
        ee_filter=None if ee is None else ee[0], ee_filter_offsets=None if ee is None else ee[1])


@pytest.mark.parametrize("worlds,variant", [(False, "default"), (True, "default"), (False, "filters")])
def test_detection_matches_oracle(cuda_device, worlds, variant):
    """Detection lists against the oracle's brute force; the "filters" variant adds a 3-ring topological
    filter, an external vertex-triangle map, a rest-shape exclusion radius and small buffers that
    overflow (both keep the smallest ids)."""
    import torch

    from newton_amd.solvers import SolverVBD

    model = _layers(cuda_device, worlds=worlds)
    kw = dict(SC, particle_self_contact_margin=0.016, particle_self_contact_radius=0.01)
    ext = None
    if variant == "filters":
        kw.update(particle_rest_shape_contact_exclusion_radius=0.004, particle_vertex_contact_buffer_size=3,
                  particle_edge_contact_buffer_size=5, particle_topological_contact_filter_threshold=3)
        ext = {70: [10, 11, 12], 71: list(range(90, 110))}
        kw["particle_external_vertex_contact_filtering_map"] = ext
    solver = SolverVBD(model, iterations=2, particle_enable_self_contact=True, **kw)
    state = model.state()
    rng = np.random.default_rng(3)
    q = state.particle_q.cpu().numpy() + rng.normal(0, 0.002, (model.particle_count, 3)).astype(np.float32)
    state.particle_q.copy_(torch.from_numpy(q))
    solver.collision_detection(state)
    info = {k: v.cpu().numpy() for k, v in solver.collision_info().items()}
    assert not solver.self_contact_overflow()
    om = oracle.OracleModel(model)
    sc = _oracle_self(model, kw, threshold=kw.get("particle_topological_contact_filter_threshold", 2),
                      vt_map=ext)
    oracle.self_detect(om, sc, q)
    np.testing.assert_array_equal(info["vertex_colliding_triangles_count"], sc.vt_count)
    np.testing.assert_array_equal(info["vertex_colliding_triangles"], sc.vt_list)
    np.testing.assert_array_equal(info["edge_colliding_edges_count"], sc.ee_count)
    np.testing.assert_array_equal(info["edge_colliding_edges"], sc.ee_list)
    np.testing.assert_array_equal(info["vertex_colliding_triangles_min_dist"], sc.vt_min_dist.astype(np.float32))
    np.testing.assert_array_equal(info["edge_colliding_edges_min_dist"], sc.ee_min_dist.astype(np.float32))
    np.testing.assert_array_equal(info["triangle_colliding_vertices_min_dist"], sc.tri_min_dist.astype(np.float32))
    if worlds:
        assert sc.vt_count.sum() == 0  # the layers are in different worlds
    else:
        assert sc.vt_count.sum() > 20 and sc.ee_count.sum() > 20
    if variant == "filters":
        assert (sc.vt_count > kw["particle_vertex_contact_buffer_size"]).any()
        assert (sc.ee_count > kw["particle_edge_contact_buffer_size"]).any()


def _run_gpu(model, kw, substeps, dt, iterations=4):
    import torch

    from newton_amd import CollisionPipeline
    from newton_amd.solvers import SolverVBD

    solver = SolverVBD(model, iterations=iterations, particle_enable_self_contact=True, **kw)
    pipe = CollisionPipeline(model, soft_contact_margin=0.01) if model.shape_count else None
    contacts = pipe.contacts() if pipe else None
    s0, s1 = model.state(), model.state()
    for _ in range(substeps):
        s0.clear_forces()
        if pipe:
            pipe.collide(s0, contacts)
        solver.step(s0, s1, None, contacts, dt)
        s0, s1 = s1, s0
    torch.cuda.synchronize()
    return s0.particle_q.cpu().numpy(), s0.particle_qd.cpu().numpy()


def _run_oracle(model, kw, substeps, dt, iterations=4):
    om = oracle.OracleModel(model)
    sc = _oracle_self(model, kw)
    prm = oracle.params(iterations, self_contact=sc)
    q = model.particle_q.cpu().numpy()
    qd = model.particle_qd.cpu().numpy()
    if model.shape_count:
        return oracle.simulate(om, prm, q, qd, substeps, dt, 0.01, len(om.pairs) // 2)[:2]
    for _ in range(substeps):
        _, q, qd = oracle.step(om, prm, q, qd, None, None, dt)
    return q, qd


@pytest.mark.parametrize("interval", [0, 2, -1])
def test_two_layer_steps_bitexact(cuda_device, interval):
    model = _layers(cuda_device)
    kw = dict(SC, particle_collision_detection_interval=interval)
    dt = 1.0 / 600.0
    q_gpu, qd_gpu = _run_gpu(model, kw, 20, dt)
    q_ref, qd_ref = _run_oracle(model, kw, 20, dt)
    np.testing.assert_array_equal(q_gpu, q_ref)
    np.testing.assert_array_equal(qd_gpu, qd_ref)
    n = model.particle_count // 2
    assert q_gpu[n:, 1].mean() > q_gpu[:n, 1].mean() + 1e-3  # the upper layer stays on top, pushing the lower


def test_layers_on_ground_with_body_contacts_bitexact(cuda_device):
    """Self-contact together with ground contacts and gravity (both contact kinds in one pass)."""
    model = _layers(cuda_device, ground=True, gap=0.01)
    dt = 1.0 / 600.0
    q_gpu, qd_gpu = _run_gpu(model, SC, 12, dt)
    q_ref, qd_ref = _run_oracle(model, SC, 12, dt)
    np.testing.assert_array_equal(q_gpu, q_ref)
    np.testing.assert_array_equal(qd_gpu, qd_ref)


def test_self_contact_residual_matches_oracle(cuda_device):
    import torch

    from newton_amd.solvers import SolverVBD

    model = _layers(cuda_device)
    solver = SolverVBD(model, iterations=2, particle_enable_self_contact=True, **SC)
    state = model.state()
    dt = 1.0 / 600.0
    solver.begin_step(state, None, dt)
    solver.iterate(state, dt, 1)
    solver.collision_detection(state)
    g = torch.zeros_like(state.particle_q)
    r2 = solver.residual(state, dt, g)
    torch.cuda.synchronize()
    om = oracle.OracleModel(model)
    prm = oracle.params(2, self_contact=_oracle_self(model, SC))
    g_ref, n_ref = oracle.residual(om, prm, None, state.particle_q.cpu().numpy(),
                                   solver.scratch("q_prev").cpu().numpy(), solver.scratch("inertia").cpu().numpy(),
                                   dt, precision="f32")
    np.testing.assert_array_equal(g.cpu().numpy(), g_ref)
    assert abs(float(r2.item()) ** 0.5 - n_ref) <= 1e-6 * max(n_ref, 1.0)


def test_graph_capture_replay_matches_oracle(cuda_device):
    """A captured substep with self-contact (detection, scans, rows, truncation) replayed several times,
    eager work and synchronisations between replays; every replay equals the oracle's substep."""
    import torch

    from newton_amd.solvers import SolverVBD

    model = _layers(cuda_device)
    dt = 1.0 / 600.0
    solver = SolverVBD(model, iterations=3, particle_enable_self_contact=True, **SC)
    s0, s1 = model.state(), model.state()

    def substep():
        solver.step(s0, s1, None, None, dt)
        s0.particle_q.copy_(s1.particle_q)
        s0.particle_qd.copy_(s1.particle_qd)

    q_init, qd_init = s0.particle_q.clone(), s0.particle_qd.clone()
    substep()  # warm-up (eager) step
    torch.cuda.synchronize()
    s0.particle_q.copy_(q_init)
    s0.particle_qd.copy_(qd_init)
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(stream):
        with torch.cuda.graph(graph, stream=stream):
            substep()
    torch.cuda.synchronize()
    s0.particle_q.copy_(q_init)
    s0.particle_qd.copy_(qd_init)
    torch.cuda.synchronize()
    replays = []
    for _ in range(6):
        graph.replay()
        torch.cuda.synchronize()
        _ = (s1.particle_q * 2.0).sum().item()  # unrelated eager work between replays
        replays.append(s0.particle_q.cpu().numpy().copy())
    om = oracle.OracleModel(model)
    prm = oracle.params(3, self_contact=_oracle_self(model, SC))
    q, qd = q_init.cpu().numpy(), qd_init.cpu().numpy()
    for k in range(6):
        _, q, qd = oracle.step(om, prm, q, qd, None, None, dt)
        np.testing.assert_array_equal(replays[k], q)
