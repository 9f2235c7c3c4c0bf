"""GPU parity of contacts against shapes on externally driven bodies and of full-surface edge/face contacts.

libvbd_hip against the FP32 CPU oracle, bit for bit: the soft-contact records (particle records, then
edge and face records, in pair order) at a body pose, and whole trajectories -- a cloth on a sphere and a
plane carried by bodies that translate and rotate every substep (SolverVBD with
integrate_with_external_rigid_solver=True reads state_out.body_q / state_in.body_q), a coarse cloth held
up by edge/face contacts on a sphere smaller than its cells, the same with the penalty ramp, the force
residual with edge/face records, and the allocation-free collision entry point with full-surface pairs.
"""

import ctypes
import math

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _quat(axis, angle):
    axis = np.asarray(axis, np.float64)
    axis = axis / np.linalg.norm(axis)
    return np.concatenate([axis * math.sin(0.5 * angle), [math.cos(0.5 * angle)]]).astype(np.float32)


def _body_scene(dev, n=12):
    """An n x n cloth above a sphere on body 0 and a ground plane on body 1."""
    from newton_amd import ModelBuilder
    from newton_amd.sim.builder import cloth_grid_mesh

    b = ModelBuilder(up_axis="Y", gravity=-9.81)
    b0 = b.add_body(xform=((0.05, 0.0, -0.03), _quat((0.2, 1.0, 0.1), 0.4)))
    b1 = b.add_body(xform=((0.0, 0.03, 0.0), (0.0, 0.0, 0.0, 1.0)))
    b.add_shape_sphere(b0, xform=((0.02, 0.01, 0.0), _quat((1.0, 0.0, 0.0), 0.3)), radius=0.12,
                       cfg=ModelBuilder.ShapeConfig(ke=2.0e3, kd=20.0, mu=0.4, margin=0.002))
    b.add_shape_plane(plane=(0.0, 1.0, 0.0, 0.0), width=0.0, length=0.0, body=b1,
                      cfg=ModelBuilder.ShapeConfig(ke=1.5e3, kd=10.0, mu=0.6))
    verts, tris = cloth_grid_mesh(n - 1, n - 1, 0.04, 0.04)
    b.add_cloth_mesh(pos=(-0.22, 0.16, 0.22), rot=_quat((1.0, 0.0, 0.0), -math.pi / 2), scale=1.0,
                     vel=(0.0, -0.4, 0.0), vertices=verts, indices=tris, density=0.2, tri_ke=1.0e3, tri_ka=1.0e3,
                     tri_kd=1.0e-4, edge_ke=0.2, edge_kd=1.0e-4, particle_radius=0.02)
    b.color()
    model = b.finalize(dev)
    model.soft_contact_ke, model.soft_contact_kd, model.soft_contact_mu = 1.0e3, 5.0, 0.5
    return model


def _poses(model, k, dt):
    """Body poses at substep k: the sphere body spins about y and drifts in x, the plane body slides."""
    base = model.body_q.cpu().numpy().astype(np.float64)
    t = k * dt
    out = base.copy()
    out[0, 0] += 0.3 * t
    spin = _quat((0.0, 1.0, 0.0), 2.0 * t)
    q0 = base[0, 3:]
    # spin * q0 (Hamilton product), the body turning about the world y axis
    ax, ay, az, aw = spin
    bx, by, bz, bw = q0
    out[0, 3:] = [aw * bx + bw * ax + ay * bz - by * az, aw * by + bw * ay + az * bx - bz * ax,
                  aw * bz + bw * az + ax * by - bx * ay, aw * bw - ax * bx - ay * by - az * bz]
    out[1, 0] += 0.2 * t
    out[1, 2] -= 0.1 * t
    return out.astype(np.float32)


def test_body_shape_records_bitexact(cuda_device):
    import torch

    from newton_amd import CollisionPipeline

    model = _body_scene(cuda_device)
    state = model.state()
    state.body_q.copy_(torch.as_tensor(_poses(model, 7, 1.0 / 600.0), device=cuda_device))
    pipe = CollisionPipeline(model, soft_contact_margin=0.05)
    contacts = pipe.contacts()
    pipe.collide(state, contacts)
    torch.cuda.synchronize()
    got = oracle.Contacts.from_device(contacts)
    om = oracle.OracleModel(model)
    ref = oracle.Contacts(pipe.soft_contact_max)
    n = oracle.create_soft_contacts(om, state.particle_q.cpu().numpy(), 0.05, ref,
                                    body_q=state.body_q.cpu().numpy())
    assert int(got.count[0]) == n > 20
    for name in ("particle", "indices", "shape", "body_pos", "body_vel", "normal", "bary"):
        np.testing.assert_array_equal(getattr(got, name)[:n], getattr(ref, name)[:n], err_msg=name)


@pytest.mark.parametrize("beta", [0.0, 1.0e5])
def test_moving_bodies_trajectory_bitexact(cuda_device, beta):
    import torch

    from newton_amd import CollisionPipeline
    from newton_amd.solvers import SolverVBD

    dt, substeps, iters, margin = 1.0 / 600.0, 60, 4, 0.02
    model = _body_scene(cuda_device)
    q0, qd0 = model.particle_q.cpu().numpy().copy(), model.particle_qd.cpu().numpy().copy()
    solver = SolverVBD(model, iterations=iters, integrate_with_external_rigid_solver=True, rigid_avbd_beta=beta)
    pipe = CollisionPipeline(model, soft_contact_margin=margin)
    contacts = pipe.contacts()
    s0, s1 = model.state(), model.state()
    n_contacts = []
    for k in range(substeps):
        s0.body_q.copy_(torch.as_tensor(_poses(model, k, dt), device=cuda_device))
        s1.body_q.copy_(torch.as_tensor(_poses(model, k + 1, dt), device=cuda_device))
        s0.clear_forces()
        pipe.collide(s0, contacts)
        n_contacts.append(int(contacts.soft_contact_count.item()))
        solver.step(s0, s1, None, contacts, dt)
        s0, s1 = s1, s0
    torch.cuda.synchronize()
    q_gpu = s0.particle_q.cpu().numpy()
    # the oracle, same poses
    om = oracle.OracleModel(model)
    prm = oracle.params(iters, avbd_beta=beta, k_start=-1.0 if beta == 0.0 else solver.rigid_contact_k_start_value)
    q, qd = q0.astype(np.float32), qd0.astype(np.float32)
    shapes = set()
    for k in range(substeps):
        p0, p1 = _poses(model, k, dt), _poses(model, k + 1, dt)
        c = oracle.Contacts(pipe.soft_contact_max)
        assert oracle.create_soft_contacts(om, q, margin, c, body_q=p0) == n_contacts[k]
        shapes |= set(c.shape[: n_contacts[k]].tolist())
        oracle.set_bodies(prm, p1, p0)
        _, q, qd = oracle.step(om, prm, q, qd, None, c, dt)
    assert max(n_contacts) > 30 and shapes == {0, 1}  # the sphere's body and the plane's
    assert np.isfinite(q_gpu).all()
    np.testing.assert_array_equal(q_gpu, q)


def _coarse_on_sphere(dev, n=5, cell=0.25, sphere_r=0.08, up=0.12):
    """An n x n-cell cloth whose middle cell lies over a sphere smaller than the cell."""
    from newton_amd import ModelBuilder

    b = ModelBuilder(up_axis="Y", gravity=-9.81)
    b.add_shape_sphere(-1, radius=sphere_r, cfg=ModelBuilder.ShapeConfig(ke=3.0e3, kd=5.0, mu=0.3))
    half = 0.5 * n * cell
    b.add_cloth_grid(pos=(-half + 0.01, up, half - 0.015), rot=_quat((1.0, 0.0, 0.0), -math.pi / 2),
                     vel=(0.0, 0.0, 0.0), dim_x=n, dim_y=n, cell_x=cell, cell_y=cell, mass=0.02, tri_ke=1.0e3,
                     tri_ka=1.0e3, tri_kd=1.0e-4, edge_ke=0.5, particle_radius=0.01)
    b.color()
    model = b.finalize(dev)
    model.soft_contact_ke, model.soft_contact_kd, model.soft_contact_mu = 2.0e4, 1.0, 0.3
    return model


def _full_oracle(model, pipe):
    om = oracle.OracleModel(model)
    om.set_full_surface_pairs(pipe.soft_edge_rigid_pairs_host, pipe.soft_face_rigid_pairs_host)
    return om


def test_full_surface_records_bitexact(cuda_device):
    import torch

    from newton_amd import CollisionPipeline

    model = _coarse_on_sphere(cuda_device, up=0.095)
    pipe = CollisionPipeline(model, soft_contact_margin=0.01, enable_rigid_soft_full_surface_contact=True)
    assert len(pipe.soft_edge_rigid_pairs_host) == model.edge_count
    assert len(pipe.soft_face_rigid_pairs_host) == model.tri_count
    contacts = pipe.contacts()
    state = model.state()
    pipe.collide(state, contacts)
    torch.cuda.synchronize()
    got = oracle.Contacts.from_device(contacts)
    om = _full_oracle(model, pipe)
    ref = oracle.Contacts(pipe.soft_contact_max)
    n = oracle.create_soft_contacts(om, state.particle_q.cpu().numpy(), 0.01, ref)
    assert int(got.count[0]) == n
    kinds = ref.indices[:n]
    assert np.any(kinds[:, 1] >= 0) and np.any(kinds[:, 2] >= 0)  # edge and face records
    for name in ("particle", "indices", "shape", "body_pos", "body_vel", "normal", "bary"):
        np.testing.assert_array_equal(getattr(got, name)[:n], getattr(ref, name)[:n], err_msg=name)


@pytest.mark.parametrize("beta", [0.0, 1.0e5])
def test_full_surface_trajectory_bitexact(cuda_device, beta):
    """The coarse cloth dropped on the small sphere: held up by edge/face contacts, bit-exact."""
    import torch

    from newton_amd import CollisionPipeline
    from newton_amd.solvers import SolverVBD

    dt, substeps, iters, margin = 1.0 / 600.0, 120, 6, 0.01
    model = _coarse_on_sphere(cuda_device)
    q0, qd0 = model.particle_q.cpu().numpy().copy(), model.particle_qd.cpu().numpy().copy()
    pipe = CollisionPipeline(model, soft_contact_margin=margin, enable_rigid_soft_full_surface_contact=True)
    contacts = pipe.contacts()
    solver = SolverVBD(model, iterations=iters, rigid_avbd_beta=beta)
    s0, s1 = model.state(), model.state()
    counts = []
    for _ in range(substeps):
        s0.clear_forces()
        pipe.collide(s0, contacts)
        counts.append(int(contacts.soft_contact_count.item()))
        solver.step(s0, s1, None, contacts, dt)
        s0, s1 = s1, s0
    torch.cuda.synchronize()
    q_gpu = s0.particle_q.cpu().numpy()
    om = _full_oracle(model, pipe)
    prm = oracle.params(iters, avbd_beta=beta, k_start=-1.0 if beta == 0.0 else solver.rigid_contact_k_start_value)
    q, qd = q0.astype(np.float32), qd0.astype(np.float32)
    for k in range(substeps):
        c = oracle.Contacts(pipe.soft_contact_max)
        assert oracle.create_soft_contacts(om, q, margin, c) == counts[k]
        _, q, qd = oracle.step(om, prm, q, qd, None, c, dt)
    assert max(counts) > 0
    np.testing.assert_array_equal(q_gpu, q)
    assert q_gpu[:, 1].max() > 0.05  # resting on the sphere (top at 0.08), not fallen through


def test_full_surface_residual_bitexact(cuda_device):
    import torch

    from newton_amd import CollisionPipeline
    from newton_amd.solvers import SolverVBD

    model = _coarse_on_sphere(cuda_device, up=0.085)
    pipe = CollisionPipeline(model, soft_contact_margin=0.01, enable_rigid_soft_full_surface_contact=True)
    contacts = pipe.contacts()
    state = model.state()
    pipe.collide(state, contacts)
    solver = SolverVBD(model, iterations=3)
    dt = 1.0 / 600.0
    solver.begin_step(state, contacts, dt)
    solver.iterate(state, dt, 2)
    g = torch.empty_like(state.particle_q)
    norm_sq = float(solver.residual(state, dt, g).item())
    torch.cuda.synchronize()
    c = oracle.Contacts.from_device(contacts)
    assert np.any(c.indices[: int(c.count[0]), 1] >= 0)
    om = _full_oracle(model, pipe)
    g_ref, norm_ref = oracle.residual(om, oracle.params(3), c, state.particle_q.cpu().numpy(),
                                      solver.scratch("q_prev").cpu().numpy(), solver.scratch("inertia").cpu().numpy(),
                                      dt, precision="f32")
    np.testing.assert_array_equal(g.cpu().numpy(), g_ref)
    assert norm_sq == pytest.approx(norm_ref**2, rel=1e-12)


def test_noalloc_full_surface_matches_allocating_entry_point(cuda_device):
    import torch

    from newton_amd import CollisionPipeline, _native
    from newton_amd.sim.collide import contacts_out
    from newton_amd.sim.model import Contacts

    model = _coarse_on_sphere(cuda_device, up=0.095)
    pipe = CollisionPipeline(model, soft_contact_margin=0.01, enable_rigid_soft_full_surface_contact=True)
    state = model.state()
    ref = pipe.contacts()
    pipe.collide(state, ref)
    got = Contacts(pipe.soft_contact_max, cuda_device)
    got.soft_contact_indices.fill_(-7)
    got.soft_contact_barycentric.fill_(float("nan"))
    stream = torch.cuda.current_stream(cuda_device).cuda_stream
    rc = _native.hip_lib().vbd_collide_generate_noalloc(ctypes.byref(pipe.desc), stream,
                                                        _native.ptr(state.particle_q), 0.01,
                                                        ctypes.byref(contacts_out(got)))
    _native.check(rc, "vbd_collide_generate_noalloc")
    torch.cuda.synchronize()
    n = int(ref.soft_contact_count.item())
    assert int(got.soft_contact_count.item()) == n > 0
    for name in ("soft_contact_indices", "soft_contact_barycentric", "soft_contact_body_pos", "soft_contact_normal",
                 "soft_contact_shape"):
        assert torch.equal(getattr(got, name)[:n], getattr(ref, name)[:n]), name
