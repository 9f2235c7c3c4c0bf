"""GPU: iterate-to-residual driver and the warm-start hook against the FP32 CPU oracle.

The oracle's step records the force-residual norm after every iteration; the
GPU driver must stop at the first iteration whose residual is within the
relative tolerance and its result must equal the oracle's step with exactly
that many iterations (bit-identical positions and velocities).
"""

import numpy as np
import pytest

import oracle
from _sim import initial

pytestmark = pytest.mark.gpu


def _drop_scene(dev, verts=24):
    from newton_amd import scenes

    # starts 0.01 m inside the sphere's contact distance (particle radius 0.1): contacts from the first substep
    return scenes.cloth_drop_on_sphere(dev, verts_per_side=verts, size=1.0, height=0.49, iterations=10)


def _prepare(model, state, margin=0.01):
    from newton_amd import CollisionPipeline

    pipe = CollisionPipeline(model, soft_contact_margin=margin)
    contacts = pipe.contacts()
    state.clear_forces()
    pipe.collide(state, contacts)
    return pipe, contacts


@pytest.mark.parametrize("rel_tol,check_every", [(1e-2, 1), (3e-2, 1), (1e-2, 4)])
def test_iterate_to_residual_matches_oracle(cuda_device, rel_tol, check_every):
    import torch

    from newton_amd.solvers import SolverVBD

    model, state, info = _drop_scene(cuda_device)
    q0, qd0 = initial(state)
    _, contacts = _prepare(model, state)
    oc = oracle.Contacts.from_device(contacts)
    assert int(oc.count[0]) > 0
    om = oracle.OracleModel(model)
    cap = 120
    _, _, _, hist = oracle.step(om, oracle.params(cap), q0, qd0, None, oc, info["dt"], residual_history=True)
    # the driver evaluates the residual after 0, c, 2c, ... iterations and stops at the first within tolerance
    checks = np.arange(0, cap + 1, check_every)
    hit = checks[hist[checks] <= rel_tol * hist[0]]
    assert hit.size, "oracle never reached the tolerance within the cap"
    k_expect = int(hit[0])
    k_first = int(np.nonzero(hist <= rel_tol * hist[0])[0][0])

    solver = SolverVBD(model, iterations=info["iterations"])
    out = model.state()
    res = solver.step_to_residual(state, out, None, contacts, info["dt"], rel_tol=rel_tol, max_iterations=cap,
                                  check_every=check_every)
    torch.cuda.synchronize()
    assert res["iterations"] == k_expect, (res, k_first, hist[: k_first + 2])
    assert res["r"] <= rel_tol * res["r0"]
    assert abs(res["r0"] - hist[0]) <= 1e-6 * hist[0]
    _, q_ref, qd_ref = oracle.step(om, oracle.params(k_expect), q0, qd0, None, oc, info["dt"])
    np.testing.assert_array_equal(out.particle_q.cpu().numpy(), q_ref)
    np.testing.assert_array_equal(out.particle_qd.cpu().numpy(), qd_ref)


def test_residual_norm_is_reproducible(cuda_device):
    import torch

    from newton_amd.solvers import SolverVBD

    model, state, info = _drop_scene(cuda_device, verts=40)
    _, contacts = _prepare(model, state)
    solver = SolverVBD(model, iterations=2)
    solver.begin_step(state, contacts, info["dt"])
    solver.iterate(state, info["dt"], 2)
    a = [solver.residual(state, info["dt"]).item() for _ in range(3)]
    torch.cuda.synchronize()
    assert a[0] == a[1] == a[2]


def test_warm_start_accepts_better_and_rejects_worse(cuda_device):
    """A converged guess is accepted (few iterations left); a displaced guess is rejected and changes nothing."""
    import torch

    from newton_amd.solvers import SolverVBD

    model, state, info = _drop_scene(cuda_device)
    q0, qd0 = initial(state)
    dev = model.device

    def fresh():
        s = model.state()
        s.particle_q.copy_(torch.as_tensor(q0, device=dev))
        s.particle_qd.copy_(torch.as_tensor(qd0, device=dev))
        return s

    solver = SolverVBD(model, iterations=info["iterations"])
    s = fresh()
    _, contacts = _prepare(model, s)
    out_plain = model.state()
    plain = solver.step_to_residual(s, out_plain, None, contacts, info["dt"], rel_tol=1e-2, max_iterations=150)
    s = fresh()
    _, contacts = _prepare(model, s)
    tight = model.state()
    solver.step_to_residual(s, tight, None, contacts, info["dt"], rel_tol=1e-4, max_iterations=400)
    converged = tight.particle_q.clone()

    s = fresh()
    _, contacts = _prepare(model, s)
    out_good = model.state()
    good = solver.step_to_residual(s, out_good, None, contacts, info["dt"], rel_tol=1e-2, max_iterations=150,
                                   warm_start=lambda solver_, st, dt: converged)
    assert good["warm_start_accepted"]
    assert good["r_warm_start"] < good["r0"]
    assert good["iterations"] < plain["iterations"]

    s = fresh()
    _, contacts = _prepare(model, s)
    out_bad = model.state()
    bad = solver.step_to_residual(s, out_bad, None, contacts, info["dt"], rel_tol=1e-2, max_iterations=150,
                                  warm_start=lambda solver_, st, dt: st.particle_q + 0.25)
    torch.cuda.synchronize()
    assert not bad["warm_start_accepted"]
    assert bad["iterations"] == plain["iterations"]
    np.testing.assert_array_equal(out_bad.particle_q.cpu().numpy(), out_plain.particle_q.cpu().numpy())
    np.testing.assert_array_equal(out_bad.particle_qd.cpu().numpy(), out_plain.particle_qd.cpu().numpy())


def test_set_iterate_trial_keeps_pinned_and_solver_state(cuda_device):
    """commit=False fills only the trial buffer; pinned particles keep x0 in it."""
    import torch

    from newton_amd import scenes
    from newton_amd.solvers import SolverVBD

    model, state, info = scenes.cloth_hanging(cuda_device)
    _, contacts = _prepare(model, state)
    solver = SolverVBD(model, iterations=2)
    solver.begin_step(state, contacts, info["dt"])
    q_before = state.particle_q.clone()
    disp_before = solver.scratch("displacement")
    guess = state.particle_q + 0.01
    trial = solver.set_iterate(state, guess, commit=False)
    torch.cuda.synchronize()
    assert torch.equal(state.particle_q, q_before)
    assert torch.equal(solver.scratch("displacement"), disp_before)
    pinned = (model.particle_inv_mass == 0).cpu().numpy()
    assert pinned.sum() == 2
    x0 = solver.scratch("q_prev").cpu().numpy()
    t = trial.cpu().numpy()
    np.testing.assert_array_equal(t[pinned], x0[pinned])
    g = guess.cpu().numpy()
    np.testing.assert_array_equal(t[~pinned], x0[~pinned] + (g[~pinned] - x0[~pinned]))
    solver.set_iterate(state, guess, commit=True)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(state.particle_q.cpu().numpy(), t)
    np.testing.assert_array_equal(solver.scratch("displacement").cpu().numpy()[~pinned],
                                  g[~pinned] - x0[~pinned])
