"""GPU checks of the plugin boundary: SolverBase.integrate_particles against the oracle (bit-exact),
and notify_model_changed keeping the solver's context, pass shape and captured graphs valid."""

import numpy as np
import pytest

import oracle
from _sim import initial

pytestmark = pytest.mark.gpu


def test_integrate_particles_bitexact(cuda_device):
    import torch

    from newton_amd import ModelBuilder
    from newton_amd.solvers import SolverBase

    rng = np.random.default_rng(11)
    n = 1000
    b = ModelBuilder(up_axis="Y")
    flags = (rng.random(n) > 0.1).astype(np.int32)
    mass = np.where(rng.random(n) < 0.1, 0.0, rng.uniform(0.01, 2.0, n))
    b.add_particles(rng.normal(size=(n, 3)), 3.0 * rng.normal(size=(n, 3)), mass, None, flags)
    model = b.finalize(cuda_device)
    model.gravity = torch.tensor([[0.0, -9.81, 0.0], [0.0, 0.0, -2.0], [0.5, 0.1, 0.0]], device=cuda_device)
    world = rng.integers(-1, 3, n).astype(np.int32)
    model.particle_world = torch.as_tensor(world, device=cuda_device)
    model.particle_max_velocity = 6.0
    s0, s1 = model.state(), model.state()
    s0.particle_f = torch.as_tensor(rng.normal(size=(n, 3)).astype(np.float32), device=cuda_device)
    s1.particle_qd.fill_(123.0)  # inactive particles leave it unwritten
    SolverBase(model).integrate_particles(model, s0, s1, 1.0 / 60.0)
    torch.cuda.synchronize()
    x_ref, v_ref = oracle.integrate_particles(s0.particle_q.cpu().numpy(), s0.particle_qd.cpu().numpy(),
                                              s0.particle_f.cpu().numpy(), model.particle_inv_mass.cpu().numpy(),
                                              flags, world, model.gravity.cpu().numpy(), 1.0 / 60.0, 6.0)
    active = flags == 1
    np.testing.assert_array_equal(s1.particle_q.cpu().numpy(), x_ref)
    np.testing.assert_array_equal(s1.particle_qd.cpu().numpy()[active], v_ref[active])
    assert np.all(s1.particle_qd.cpu().numpy()[~active] == 123.0)
    assert np.abs(np.linalg.norm(v_ref[active], axis=1) - 6.0).min() < 1e-5  # the limit was hit


def test_notify_model_changed_keeps_context_shape_and_graphs(cuda_device):
    """A material edit is written into the existing layout: same context, same pass shape, and a graph
    captured before the edit replays with the new materials (bit-exact against the oracle)."""
    import torch

    from newton_amd import scenes
    from newton_amd.solvers import SolverVBD

    model, state, info = scenes.cloth_strip_pinned(cuda_device)
    q0, qd0 = initial(state)
    dt = info["dt"]
    solver = SolverVBD(model, iterations=info["iterations"])
    solver.set_pass_shape(192, 7, True)
    handle = solver._ctx.value
    s1 = model.state()
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(stream):
        solver.step(state, s1, None, None, dt)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=stream):
            solver.step(state, s1, None, None, dt)
    torch.cuda.synchronize()
    model.tri_materials[:, 0] *= 3.0
    model.tri_materials[:, 2] *= 0.5
    model.edge_bending_properties[:, 0] *= 0.5
    solver.notify_model_changed()
    assert solver._ctx.value == handle
    assert solver.pass_shape() == (192, 7, 1)
    om = oracle.OracleModel(model)
    _, q_ref, qd_ref = oracle.step(om, oracle.params(info["iterations"]), q0, qd0, None, None, dt)
    for _ in range(2):
        state.particle_q.copy_(torch.as_tensor(q0, device=cuda_device))
        state.particle_qd.copy_(torch.as_tensor(qd0, device=cuda_device))
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(s1.particle_q.cpu().numpy(), q_ref)
        np.testing.assert_array_equal(s1.particle_qd.cpu().numpy(), qd_ref)


def test_pass_shape_survives_context_regrowth(cuda_device):
    """A larger soft-contact buffer re-creates the context; the chosen pass shape is re-applied."""
    from newton_amd import CollisionPipeline, scenes
    from newton_amd.solvers import SolverVBD

    model, state, info = scenes.cloth_strip_pinned(cuda_device)
    solver = SolverVBD(model, iterations=info["iterations"], soft_contact_capacity=4)
    solver.set_pass_shape(128, 1, True)
    pipe = CollisionPipeline(model, soft_contact_margin=0.05)
    contacts = pipe.contacts()
    assert contacts.soft_contact_max > 4
    pipe.collide(state, contacts)
    solver.step(state, model.state(), None, contacts, info["dt"])
    assert solver.pass_shape() == (128, 1, 1)


def test_contact_scalars_are_read_per_step(cuda_device):
    """Changing model.soft_contact_ke between steps takes effect without re-creating the context."""
    import torch

    from newton_amd import CollisionPipeline, scenes
    from newton_amd.solvers import SolverVBD

    model, state, info = scenes.cloth_strip_pinned(cuda_device)
    q0, qd0 = initial(state)
    solver = SolverVBD(model, iterations=info["iterations"])
    handle = solver._ctx.value
    pipe = CollisionPipeline(model, soft_contact_margin=0.05)
    contacts = pipe.contacts()
    model.soft_contact_ke = 7.5e3
    model.soft_contact_mu = 0.2
    s1 = model.state()
    pipe.collide(state, contacts)
    solver.step(state, s1, None, contacts, info["dt"])
    torch.cuda.synchronize()
    assert solver._ctx.value == handle
    om = oracle.OracleModel(model)
    oc = oracle.Contacts(pipe.soft_contact_max)
    oracle.create_soft_contacts(om, q0, 0.05, oc)
    _, q_ref, _ = oracle.step(om, oracle.params(info["iterations"]), q0, qd0, None, oc, info["dt"])
    np.testing.assert_array_equal(s1.particle_q.cpu().numpy(), q_ref)
