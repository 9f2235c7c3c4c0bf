"""Step-level checks of the CPU oracle (SolverVBD.step restatement), mirroring the reference's integration tests.

* forward step / velocity update closed form: a free particle follows the
  discrete free-fall recursion exactly (test_cloth.py:1116-1137);
* internal forces do not perturb a freely falling cloth (test_cloth.py:1047-1099);
* pinned vertices stay bit-identical (test_cloth.py:955-970);
* run-to-run and OpenMP-vs-serial bitwise equality (determinism, tests/determinism);
* the FP32 oracle tracks the FP64 oracle;
* the incremental-potential force residual decreases over VBD iterations.
"""

import numpy as np
import pytest

import oracle
from newton_amd import ModelBuilder, scenes


def test_free_particle_follows_discrete_free_fall():
    b = ModelBuilder(up_axis="Y", gravity=-9.81)
    b.add_particle((0.0, 2.0, 0.0), (0.3, 0.0, 0.0), 1.0)
    b.color()
    model = b.finalize("cpu")
    om = oracle.OracleModel(model)
    prm = oracle.params(3)
    dt = np.float32(1.0 / 600.0)
    g = np.float32(-9.81)
    q = np.array([[0.0, 2.0, 0.0]], dtype=np.float32)
    v = np.array([[0.3, 0.0, 0.0]], dtype=np.float32)
    q_ref, v_ref = q.copy(), v.copy()
    for _ in range(50):
        _, q, v = oracle.step(om, prm, q, v, None, None, float(dt))
        vn = v_ref + np.array([[0.0, g, 0.0]], dtype=np.float32) * dt
        q_new = q_ref + vn * dt
        v_ref = (q_new - q_ref) / dt
        q_ref = q_new
        np.testing.assert_array_equal(q, q_ref)
        np.testing.assert_array_equal(v, v_ref)
    t = 50 * float(dt)
    # FP32 roundoff of v = (q - q_prev) / dt accumulates; the reference's own bound is 0.2 (test_cloth.py:1116-1137)
    assert abs(float(q[0, 1]) - (2.0 - 0.5 * 9.81 * t * (t + float(dt)))) < 2e-4


def test_unpinned_cloth_falls_like_a_point_mass():
    model, state, info = scenes.cloth_hanging("cpu")
    flags = model.particle_flags.numpy().copy()
    # unpin: restore ACTIVE and give the corners the neighbouring mass
    mass = model.particle_mass.numpy()
    pinned = np.where(flags == 0)[0]
    mass[pinned] = mass[pinned - 1]
    model.particle_inv_mass[:] = 1.0 / model.particle_mass
    model.particle_flags[pinned] = 1
    om = oracle.OracleModel(model)
    prm = oracle.params(info["iterations"], num_threads=4)
    q0 = state.particle_q.numpy()
    q, qd, _ = oracle.simulate(om, prm, q0, state.particle_qd.numpy(), 60, info["dt"], 0.01, 0)
    t = 60 * info["dt"]
    drop = (q0[:, 1] - q[:, 1]).mean()
    expected = 0.5 * 9.81 * t * (t + info["dt"])
    assert abs(drop - expected) < 0.02 * expected + 1e-4
    assert np.abs(q[:, [0, 2]] - q0[:, [0, 2]]).max() < 1e-3


def test_pinned_vertices_are_bit_identical():
    model, state, info = scenes.cloth_hanging("cpu")
    om = oracle.OracleModel(model)
    q0 = state.particle_q.numpy()
    q, _, _ = oracle.simulate(om, oracle.params(info["iterations"], num_threads=4), q0, state.particle_qd.numpy(),
                              40, info["dt"], 0.01, len(om.pairs) // 2)
    pinned = np.where(model.particle_flags.numpy() == 0)[0]
    assert len(pinned) == 2
    np.testing.assert_array_equal(q[pinned], q0[pinned])


def test_run_to_run_and_thread_count_bitwise_equal():
    model, state, info = scenes.cloth_strip_pinned("cpu")
    om = oracle.OracleModel(model)
    q0, qd0 = state.particle_q.numpy(), state.particle_qd.numpy()
    runs = []
    for threads in (1, 1, 4):
        q, qd, _ = oracle.simulate(om, oracle.params(info["iterations"], num_threads=threads), q0, qd0, 30,
                                   info["dt"], 0.05, len(om.pairs) // 2)
        runs.append((q, qd))
    for q, qd in runs[1:]:
        np.testing.assert_array_equal(q, runs[0][0])
        np.testing.assert_array_equal(qd, runs[0][1])


@pytest.mark.parametrize("maker", [scenes.cloth_strip_pinned, scenes.soft_cube])
def test_fp32_oracle_tracks_fp64(maker):
    model, state, info = maker("cpu")
    om = oracle.OracleModel(model)
    q0, qd0 = state.particle_q.numpy(), state.particle_qd.numpy()
    prm = oracle.params(info["iterations"], num_threads=4)
    q32, _, _ = oracle.simulate(om, prm, q0, qd0, 30, info["dt"], 0.05, len(om.pairs) // 2)
    q64, _, _ = oracle.simulate(om, prm, q0, qd0, 30, info["dt"], 0.05, len(om.pairs) // 2, precision="f64")
    assert np.abs(q32 - q64).max() < 1e-4


def test_libm_atan2_variant_stays_close():
    """The portable atan2 (shared with the HIP kernels) versus the C library's atan2f: same trajectory to 1e-6."""
    model, state, info = scenes.cloth_hanging("cpu")
    om = oracle.OracleModel(model)
    q0, qd0 = state.particle_q.numpy(), state.particle_qd.numpy()
    qa, _, _ = oracle.simulate(om, oracle.params(info["iterations"], num_threads=4), q0, qd0, 60, info["dt"], 0.01,
                               len(om.pairs) // 2)
    qb, _, _ = oracle.simulate(om, oracle.params(info["iterations"], num_threads=4, use_libm_atan2=True), q0, qd0,
                               60, info["dt"], 0.01, len(om.pairs) // 2)
    assert np.abs(qa - qb).max() < 1e-6


def test_force_residual_decreases_with_iterations():
    model, state, info = scenes.cloth_drop_on_sphere("cpu", verts_per_side=24)
    om = oracle.OracleModel(model)
    q0, qd0 = state.particle_q.numpy(), state.particle_qd.numpy()
    prm = oracle.params(30)
    _, _, _, hist = oracle.step(om, prm, q0, qd0, None, None, info["dt"], precision="f64", residual_history=True)
    assert hist[0] > 0
    # Gauss-Seidel VBD on a flat 24x24 cloth: about 0.036 after 30 iterations
    assert hist[-1] < 0.1 * hist[0]
    assert hist[10] < hist[0] and hist[-1] < hist[10]


def test_inactive_and_zero_mass_particles_do_not_move():
    model, state, info = scenes.cloth_strip_pinned("cpu")
    model.particle_mass[5] = 0.0
    model.particle_inv_mass[5] = 0.0
    om = oracle.OracleModel(model)
    q0 = state.particle_q.numpy()
    q, qd, _ = oracle.simulate(om, oracle.params(info["iterations"]), q0, state.particle_qd.numpy(), 10, info["dt"],
                               0.05, len(om.pairs) // 2)
    np.testing.assert_array_equal(q[[0, 5]], q0[[0, 5]])
    np.testing.assert_array_equal(qd[[0, 5]], 0.0)


def test_fp64_block_solve_changes_and_improves_ill_conditioned_steps():
    """Nearly massless beads on stiff springs: blocks with condition numbers >> 1e6.  The FP64 solve
    takes another path than the all-FP32 one and lands far closer to the FP64 oracle of the same
    iteration; on a well-conditioned cloth both modes give identical bits."""
    from newton_amd import scenes

    model, state, info = scenes.bead_rope_light("cpu")
    om = oracle.OracleModel(model)
    q0 = state.particle_q.numpy()
    qd0 = state.particle_qd.numpy()
    cmax = len(om.pairs) // 2
    runs = {}
    for fp64 in (False, True):
        runs[fp64], _, _ = oracle.simulate(om, oracle.params(info["iterations"], fp64_ill_conditioned=fp64), q0, qd0,
                                           10, info["dt"], 0.01, cmax)
    ref, _, _ = oracle.simulate(om, oracle.params(info["iterations"]), q0.astype(np.float64),
                                qd0.astype(np.float64), 10, info["dt"], 0.01, cmax, precision="f64")
    err_fp32 = np.abs(runs[False] - ref).max()
    err_fp64 = np.abs(runs[True] - ref).max()
    assert np.abs(runs[True] - runs[False]).max() > 1e-4
    assert err_fp64 < 1e-5 and err_fp32 > 100 * err_fp64, (err_fp32, err_fp64)
    model, state, info = scenes.cloth_hanging("cpu")
    om = oracle.OracleModel(model)
    q0 = state.particle_q.numpy()
    outs = [oracle.simulate(om, oracle.params(info["iterations"], fp64_ill_conditioned=f), q0,
                            state.particle_qd.numpy(), 3, info["dt"], 0.01, len(om.pairs) // 2)[0] for f in (False, True)]
    np.testing.assert_array_equal(outs[0], outs[1])


def test_integrate_particles_closed_forms():
    """SolverBase.integrate_particles (solver.py:20-56): free fall v + g dt, x + v' dt; forces scale by 1/m;
    massless particles ignore gravity; the speed limit caps |v'| at v_max; inactive particles stay put."""
    rng = np.random.default_rng(3)
    n = 64
    x = rng.normal(size=(n, 3)).astype(np.float32)
    v = rng.normal(size=(n, 3)).astype(np.float32)
    f = np.zeros((n, 3), np.float32)
    w = np.full(n, 2.0, np.float32)
    flags = np.ones(n, np.int32)
    world = rng.integers(-1, 3, n)
    gravity = np.array([[0, -9.81, 0], [0, 0, -3.0], [1.0, 0, 0]], np.float32)
    dt = 1.0 / 60.0
    x1, v1 = oracle.integrate_particles(x, v, f, w, flags, world, gravity, dt, 1e5)
    g = gravity[np.maximum(world, 0)].astype(np.float64)
    np.testing.assert_allclose(v1, v + g * dt, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(x1, x + (v + g * dt) * dt, rtol=1e-6, atol=1e-6)
    f = rng.normal(size=(n, 3)).astype(np.float32)
    w[:8] = 0.0
    flags[8:12] = 0
    x1, v1 = oracle.integrate_particles(x, v, f, w, flags, world, gravity, dt, 1e5)
    np.testing.assert_array_equal(v1[:8], v[:8])  # no mass: neither force nor gravity
    np.testing.assert_array_equal(x1[8:12], x[8:12])
    np.testing.assert_allclose(v1[12:], v[12:] + (2.0 * f[12:] + g[12:]) * dt, rtol=1e-5, atol=1e-6)
    x1, v1 = oracle.integrate_particles(x, 100.0 * v, f, w, np.ones(n, np.int32), world, gravity, dt, 5.0)
    speed = np.linalg.norm(v1.astype(np.float64), axis=1)
    np.testing.assert_allclose(speed, 5.0, rtol=1e-6)
    np.testing.assert_allclose(x1, x + v1.astype(np.float64) * dt, rtol=1e-6, atol=1e-6)


def test_residual_stop_ends_at_first_iteration_within_tolerance():
    """``residual_stop`` (iterate to a relative tolerance): the step stops at the first iteration of
    the full residual history that is within the tolerance, with the result of a step of exactly
    that many iterations; the residual norm does not depend on the thread count."""
    import torch

    model, state, info = scenes.cloth_wind("cpu", verts_per_side=24)
    om = oracle.OracleModel(model)
    q0, qd0 = state.particle_q.numpy(), state.particle_qd.numpy()
    f = info["wind"].force(torch.from_numpy(q0), model.particle_mass).numpy()
    cap = 80
    _, _, _, hist = oracle.step(om, oracle.params(cap), q0, qd0, f, None, info["dt"], residual_history=True)
    k = int(np.nonzero(hist <= 0.7 * hist[0])[0][0])
    assert 0 < k < cap
    _, q_s, qd_s, h_s = oracle.step(om, oracle.params(cap, num_threads=4, residual_stop=0.7), q0, qd0, f, None,
                                    info["dt"], residual_history=True)
    np.testing.assert_array_equal(h_s[: k + 1], hist[: k + 1])
    assert (h_s[k + 1:] == 0).all()
    _, q_k, qd_k = oracle.step(om, oracle.params(k), q0, qd0, f, None, info["dt"])
    np.testing.assert_array_equal(q_s, q_k)
    np.testing.assert_array_equal(qd_s, qd_k)
