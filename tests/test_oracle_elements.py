"""Pins the CPU oracle's element functions with checks that a plausible mistake would fail.

* forces are minus the gradient of the energies the reference documents
  (membrane particle_vbd_kernels.py:514-518, bending (k/2)(theta-theta0)^2 with
  k = ke * rest_length :720-721, tet Smith et al. 2018 :148-230), by central
  finite differences of independent float64 numpy energies;
* Hessians: the tet per-vertex block and the stretched (s >= 0) membrane block
  equal the finite-difference Jacobian of the force; the bending block is the
  Gauss-Newton k * grad(theta) grad(theta)^T; compressed membranes stay PSD;
* the reference's known answers: absolute contact damping [0, 0, 2]
  (test_solver_vbd.py:1333-1444), zero bending for a collapsed anchor
  (:1560-1590), zero damping force under 100 seeded rigid motions (:1593-1646);
* the 3x3 block solve against numpy.linalg.solve, with the 1e-8 determinant guard;
* the portable atan2 against the correctly rounded value.
"""

import math

import numpy as np
import pytest

import oracle

RNG_SEED = 1234


# ---------------------------------------------------------------- energies (float64 numpy)
def _f64(*vals):
    """float32-rounded parameters promoted to Python floats, so the energy is evaluated in float64."""
    return [float(np.float32(v)) for v in vals]


def membrane_energy(X, pose, area, mu, lmbd):
    area, mu, lmbd = _f64(area, mu, lmbd)
    x0, x1, x2 = X
    Ds = np.stack([x1 - x0, x2 - x0], axis=1)
    F = Ds @ np.asarray(pose, dtype=np.float64).reshape(2, 2)
    f0, f1 = F[:, 0], F[:, 1]
    Ic = f0 @ f0 + f1 @ f1
    Js = math.sqrt(max((f0 @ f0) * (f1 @ f1) - (f0 @ f1) ** 2, 1e-20))
    lam = lmbd + mu
    alpha = 1.0 + mu / lam
    return area * (0.5 * mu * (Ic - 2.0) + 0.5 * lam * (Js - alpha) ** 2)


def dihedral_angle(X):
    x0, x1, x2, x3 = X
    n1 = np.cross(x2 - x0, x3 - x0)
    n2 = np.cross(x3 - x1, x2 - x1)
    n1 /= np.linalg.norm(n1)
    n2 /= np.linalg.norm(n2)
    e = (x3 - x2) / np.linalg.norm(x3 - x2)
    return math.atan2(np.cross(n1, n2) @ e, n1 @ n2)


def bending_energy(X, rest_angle, rest_length, ke):
    rest_angle, rest_length, ke = _f64(rest_angle, rest_length, ke)
    return 0.5 * ke * rest_length * (dihedral_angle(X) - rest_angle) ** 2


def tet_energy(X, dm_inv, mu, lmbd):
    mu, lmbd = _f64(mu, lmbd)
    x0, x1, x2, x3 = X
    Ds = np.stack([x1 - x0, x2 - x0, x3 - x0], axis=1)
    Dm_inv = np.asarray(dm_inv, dtype=np.float64).reshape(3, 3)
    F = Ds @ Dm_inv
    V0 = 1.0 / (np.linalg.det(Dm_inv) * 6.0)
    lam = lmbd + mu
    alpha = 1.0 + mu / lam
    return V0 * (0.5 * mu * (np.sum(F * F) - 3.0) + 0.5 * lam * (np.linalg.det(F) - alpha) ** 2)


def fd_grad(fun, X, v, h=1e-6):
    g = np.zeros(3)
    for k in range(3):
        Xp, Xm = X.copy(), X.copy()
        Xp[v, k] += h
        Xm[v, k] -= h
        g[k] = (fun(Xp) - fun(Xm)) / (2 * h)
    return g


def fd_jacobian(force_fn, X, v, h=1e-6):
    J = np.zeros((3, 3))
    for k in range(3):
        Xp, Xm = X.copy(), X.copy()
        Xp[v, k] += h
        Xm[v, k] -= h
        J[:, k] = -(force_fn(Xp) - force_fn(Xm)) / (2 * h)
    return J


def random_triangle(rng, stretch):
    rest = np.array([[0.0, 0.0, 0.0], [1.0, 0.1, 0.0], [0.2, 0.9, 0.0]]) + rng.normal(0, 0.05, (3, 3)) * [1, 1, 0]
    qp, rp = rest[1] - rest[0], rest[2] - rest[0]
    n = np.cross(qp, rp)
    n /= np.linalg.norm(n)
    e1 = qp / np.linalg.norm(qp)
    e2 = np.cross(n, e1)
    D = np.array([e1, e2]) @ np.array([qp, rp]).T
    pose = np.linalg.inv(D).astype(np.float32)
    area = float(np.linalg.det(D) / 2.0)
    R, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    X = rest @ R.T * stretch + rng.normal(0, 0.02, (3, 3))
    return X, pose, area


@pytest.mark.parametrize("stretch", [0.8, 1.0, 1.3])
def test_membrane_force_is_negative_energy_gradient(stretch):
    rng = np.random.default_rng(RNG_SEED)
    for _ in range(10):
        X, pose, area = random_triangle(rng, stretch)
        mu, lmbd = 1.0e4 * rng.uniform(0.5, 2), 1.0e4 * rng.uniform(0.5, 2)
        for v in range(3):
            f, _ = oracle.membrane(X, X, pose, area, mu, lmbd, 0.0, 1 / 600, v)
            g = fd_grad(lambda Y: membrane_energy(Y, pose, np.float32(area), np.float32(mu), np.float32(lmbd)), X, v)
            np.testing.assert_allclose(f, -g, rtol=1e-5, atol=1e-6 * np.abs(g).max())


def test_membrane_rest_state_has_zero_force():
    rng = np.random.default_rng(RNG_SEED)
    X, pose, area = random_triangle(rng, 1.0)
    # rest configuration itself (no rotation noise): rebuild exactly from the pose
    D = np.linalg.inv(pose.astype(np.float64))
    X = np.array([[0, 0, 0], [D[0, 0], D[1, 0], 0], [D[0, 1], D[1, 1], 0]], dtype=np.float64)
    for v in range(3):
        f, _ = oracle.membrane(X, X, pose, area, 1e4, 2e4, 0.0, 1 / 600, v)
        assert np.abs(f).max() < 1e-6


def test_membrane_hessian_exact_when_stretched_and_psd_when_compressed():
    rng = np.random.default_rng(RNG_SEED + 1)
    for _ in range(8):
        X, pose, area = random_triangle(rng, 1.6)  # J_s > alpha = 1.5 -> s > 0: exact Hessian
        for v in range(3):
            _, H = oracle.membrane(X, X, pose, area, 1e4, 1e4, 0.0, 1 / 600, v)
            J = fd_jacobian(lambda Y: oracle.membrane(Y, Y, pose, area, 1e4, 1e4, 0.0, 1 / 600, v)[0], X, v)
            np.testing.assert_allclose(H, J, rtol=1e-4, atol=1e-5 * np.abs(J).max())
        Xc, pose_c, area_c = random_triangle(rng, 0.7)
        for v in range(3):
            _, H = oracle.membrane(Xc, Xc, pose_c, area_c, 1e4, 1e4, 0.0, 1 / 600, v)
            assert np.linalg.eigvalsh(0.5 * (H + H.T)).min() > -1e-8 * np.abs(H).max()


def random_hinge(rng):
    x2 = np.zeros(3)
    x3 = np.array([1.0, 0.0, 0.0]) + rng.normal(0, 0.05, 3)
    x0 = np.array([0.4, 0.8, 0.1]) + rng.normal(0, 0.1, 3)
    x1 = np.array([0.6, -0.7, 0.3]) + rng.normal(0, 0.1, 3)
    return np.stack([x0, x1, x2, x3])


def test_bending_force_is_negative_energy_gradient():
    rng = np.random.default_rng(RNG_SEED)
    for _ in range(10):
        X = random_hinge(rng)
        theta0, L0, ke = float(rng.uniform(-0.3, 0.3)), 1.3, 2.5
        for v in range(4):
            f, _ = oracle.bending(X, X, theta0, L0, ke, 0.0, 1 / 600, v)
            g = fd_grad(lambda Y: bending_energy(Y, np.float32(theta0), np.float32(L0), np.float32(ke)), X, v)
            np.testing.assert_allclose(f, -g, rtol=1e-5, atol=1e-7)


def test_bending_hessian_is_gauss_newton():
    rng = np.random.default_rng(RNG_SEED + 2)
    X = random_hinge(rng)
    L0, ke = 1.3, 2.5
    for v in range(4):
        _, H = oracle.bending(X, X, 0.1, L0, ke, 0.0, 1 / 600, v)
        gth = fd_grad(dihedral_angle, X, v)
        k = np.float64(np.float32(ke)) * np.float64(np.float32(L0))
        np.testing.assert_allclose(H, k * np.outer(gth, gth), rtol=1e-5, atol=1e-9)


def test_bending_skips_boundary_and_degenerate_anchor():
    """test_solver_vbd.py:1560-1590: collapsed previous geometry gives zero damping contribution."""
    X = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [1.0, 1.0, 0.1]])
    anchor = np.zeros((4, 3))
    for v in range(4):
        f, H = oracle.bending(X, anchor, 0.0, 1.0, 0.0, 20.0, 0.1, v, precision="f32")
        total = np.linalg.norm(f) + np.linalg.norm(H, axis=1).sum()
        assert np.isfinite(total) and total <= 1e-6


def test_tet_force_and_hessian_match_energy_derivatives():
    """test_softbody.py:387-529 (30 seeded tets), with finite differences in place of autodiff."""
    rng = np.random.default_rng(42)
    for _ in range(30):
        rest = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=np.float64) + rng.normal(0, 0.1, (4, 3))
        Dm = np.stack([rest[1] - rest[0], rest[2] - rest[0], rest[3] - rest[0]], axis=1)
        if np.linalg.det(Dm) <= 0.05:
            continue
        dm_inv = np.linalg.inv(Dm).astype(np.float32)
        X = rest + rng.normal(0, 0.1, (4, 3))
        mu, lmbd = float(rng.uniform(1e3, 1e4)), float(rng.uniform(1e3, 1e4))
        for v in range(4):
            f, H = oracle.tet(X, X, dm_inv, mu, lmbd, 0.0, 1 / 600, v)
            g = fd_grad(lambda Y: tet_energy(Y, dm_inv, np.float32(mu), np.float32(lmbd)), X, v)
            np.testing.assert_allclose(f, -g, rtol=1e-5, atol=1e-6 * np.abs(g).max())
            J = fd_jacobian(lambda Y: oracle.tet(Y, Y, dm_inv, mu, lmbd, 0.0, 1 / 600, v)[0], X, v)
            np.testing.assert_allclose(H, J, rtol=1e-4, atol=1e-5 * np.abs(J).max())


def _rotations(count, seed):
    """Rotation matrices from normalised Gaussian quaternions (test_solver_vbd.py:62-93)."""
    rng = np.random.default_rng(seed)
    q = rng.normal(size=(count, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    x, y, z, w = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    R = np.empty((count, 3, 3), dtype=np.float32)
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def test_elastic_damping_vanishes_under_rigid_motion():
    """test_solver_vbd.py:172-242 and :1593-1646: 100 seeded rigid motions, atol 1e-4 (FP32 oracle)."""
    rest = np.array([[0.0, 0.0, 0.0], [1.2, 0.1, 0.0], [0.2, 1.1, 0.1], [0.1, 0.2, 1.3]], dtype=np.float32)
    R = _rotations(100, 17)
    t = np.random.default_rng(18).uniform(-1.0, 1.0, size=(100, 3)).astype(np.float32)
    qp, rp = rest[1] - rest[0], rest[2] - rest[0]
    n = np.cross(qp, rp)
    n /= np.linalg.norm(n)
    e1 = qp / np.linalg.norm(qp)
    e2 = np.cross(n, e1)
    e2 /= np.linalg.norm(e2)
    tri_D = np.array((e1, e2), dtype=np.float32) @ np.array((qp, rp), dtype=np.float32).T
    tri_pose = np.linalg.inv(tri_D).astype(np.float32)
    tri_area = float(np.float32(np.linalg.det(tri_D) * 0.5))
    tet_pose = np.linalg.inv(np.array((rest[1] - rest[0], rest[2] - rest[0], rest[3] - rest[0]),
                                      dtype=np.float32).T).astype(np.float32)
    x1, x2, x3, x4 = rest
    n1 = np.cross(x3 - x1, x4 - x1)
    n1 /= np.linalg.norm(n1)
    n2 = np.cross(x4 - x2, x3 - x2)
    n2 /= np.linalg.norm(n2)
    ed = (x4 - x3) / np.linalg.norm(x4 - x3)
    rest_angle = float(np.float32(math.atan2(np.dot(np.cross(n1, n2), ed), np.dot(n1, n2))))
    rest_len = float(np.float32(np.linalg.norm(x4 - x3)))
    worst = np.zeros(3)
    for s in range(100):
        pos = (rest @ R[s].T + t[s]).astype(np.float32)
        for v in range(3):
            f, _ = oracle.membrane(pos[:3], rest[:3], tri_pose, tri_area, 0.0, 1.0, 20.0, 0.1, v, precision="f32")
            worst[0] = max(worst[0], np.linalg.norm(f))
        for v in range(4):
            f, _ = oracle.bending(pos, rest, rest_angle, rest_len, 0.0, 20.0, 0.1, v, precision="f32")
            worst[1] = max(worst[1], np.linalg.norm(f))
            f, _ = oracle.tet(pos, rest, tet_pose, 0.0, 1.0, 20.0, 0.1, v, precision="f32")
            worst[2] = max(worst[2], np.linalg.norm(f))
    np.testing.assert_allclose(worst, 0.0, atol=1e-4)


def test_damping_is_nonzero_for_deforming_motion():
    """Counterpart of the rigid-motion test: a stretching motion does produce damping forces."""
    rest = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    pose = np.eye(2, dtype=np.float32)
    X = rest * 1.1
    f_damped, _ = oracle.membrane(X, rest, pose, 0.5, 0.0, 1.0, 20.0, 0.1, 1)
    f_plain, _ = oracle.membrane(X, X, pose, 0.5, 0.0, 1.0, 0.0, 0.1, 1)
    assert np.linalg.norm(f_damped - f_plain) > 1.0


def test_contact_damping_is_absolute():
    """test_solver_vbd.py:1333-1374 and :1377-1444: damping force [0, 0, 2] regardless of ke."""
    forces = []
    for ke, kd in ((400.0, 20.0), (400.0, 0.0), (100.0, 20.0), (100.0, 0.0)):
        f, _ = oracle.body_contact([0.0, 0.0, 0.04], [0.0, 0.0, 0.05], [0, 0, 0], [0, 0, 0], [0, 0, 1], 0.1, 0.0,
                                   ke, kd, 0.0, 0.01, 0.1, precision="f32")
        forces.append(f)
    np.testing.assert_allclose(forces[0] - forces[1], forces[2] - forces[3], rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(forces[2] - forces[3], [0.0, 0.0, 2.0], rtol=1e-6, atol=1e-6)


def test_contact_penalty_and_friction_closed_form():
    n = np.array([0.0, 0.0, 1.0])
    ke, mu, dt, eps = 1e4, 0.5, 1 / 600, 1e-2
    x = np.array([0.3, 0.2, 0.05])
    xp = x - np.array([0.01, -0.02, 0.0])  # tangential slip, no normal motion
    f, H = oracle.body_contact(x, xp, [0, 0, 0], [0, 0, 0], n, 0.1, 0.0, ke, 0.0, mu, eps, dt)
    pen = 0.1 - 0.05
    fn = ke * pen
    u = x - xp
    ut = u - n * (n @ u)
    assert np.linalg.norm(ut) > eps * dt
    expected_f = n * fn - mu * fn * ut / np.linalg.norm(ut)
    expected_H = ke * np.outer(n, n) + mu * fn / np.linalg.norm(ut) * (np.eye(3) - np.outer(n, n))
    np.testing.assert_allclose(f, expected_f, rtol=1e-12)
    np.testing.assert_allclose(H, expected_H, rtol=1e-12)
    # separated particle: no contact
    f0, H0 = oracle.body_contact([0, 0, 0.2], [0, 0, 0.2], [0, 0, 0], [0, 0, 0], n, 0.1, 0.0, ke, 0.0, mu, eps, dt)
    assert np.all(f0 == 0) and np.all(H0 == 0)


def test_contact_friction_smooth_stage():
    """Slip below eps_u: scale (-|u|/eps + 2)/eps (rigid_vbd_kernels.py:1630-1636)."""
    n = np.array([0.0, 0.0, 1.0])
    ke, mu, dt, eps = 1e4, 0.5, 1 / 600, 1e-2
    x = np.array([0.0, 0.0, 0.05])
    xp = x - np.array([1e-6, 0.0, 0.0])
    f, _ = oracle.body_contact(x, xp, [0, 0, 0], [0, 0, 0], n, 0.1, 0.0, ke, 0.0, mu, eps, dt)
    fn = ke * 0.05
    eu = eps * dt
    un = 1e-6
    scale = mu * fn * (-un / eu + 2.0) / eu
    np.testing.assert_allclose(f, [-scale * 1e-6, 0.0, fn], rtol=1e-9)


def test_block_solve_matches_numpy_and_guards_singular():
    rng = np.random.default_rng(RNG_SEED)
    for _ in range(50):
        A = rng.normal(size=(3, 3))
        H = A @ A.T + 0.1 * np.eye(3)
        f = rng.normal(size=3)
        ok, dx = oracle.block_solve(H, f, precision="f64")
        assert ok
        np.testing.assert_allclose(dx, np.linalg.solve(H, f), rtol=1e-10, atol=1e-12)
        ok32, dx32 = oracle.block_solve(H, f, precision="f32")
        assert ok32
        np.testing.assert_allclose(dx32, np.linalg.solve(H, f), rtol=1e-3, atol=1e-4 * np.abs(dx).max())
    ok, dx = oracle.block_solve(np.diag([1e-3, 1e-3, 1e-3]), [1.0, 2.0, 3.0], precision="f32")  # det 1e-9
    assert not ok and np.all(dx == 0)


def test_fp64_fallback_improves_ill_conditioned_blocks():
    H = np.array([[1.0, 1.0, 0.0], [1.0, 1.0 + 1e-6, 0.0], [0.0, 0.0, 1.0]]) * 1e3
    f = np.array([1.0, -1.0, 0.5])
    exact = np.linalg.solve(H, f)
    _, plain = oracle.block_solve(H.astype(np.float32), f, fp64_ill_conditioned=False, precision="f32")
    _, fb = oracle.block_solve(H.astype(np.float32), f, fp64_ill_conditioned=True, precision="f32")
    exact32 = np.linalg.solve(H.astype(np.float32).astype(np.float64), f.astype(np.float32))
    assert np.abs(fb - exact32).max() <= np.abs(plain - exact32).max()
    assert np.abs(fb - exact32).max() / np.abs(exact32).max() < 1e-6
    assert np.isfinite(exact).all()


def test_portable_atan2_accuracy_and_special_cases():
    rng = np.random.default_rng(RNG_SEED)
    ys = np.concatenate([rng.normal(size=20000), rng.uniform(-1e-3, 1e-3, 2000)]).astype(np.float32)
    xs = np.concatenate([rng.normal(size=20000), rng.uniform(-1.0, 1.0, 2000)]).astype(np.float32)
    worst = 0.0
    for y, x in zip(ys, xs):
        got = np.float32(oracle.atan2(float(y), float(x)))
        ref = math.atan2(float(y), float(x))
        ulp = float(np.spacing(np.float32(abs(ref)))) if ref != 0 else 1e-45
        worst = max(worst, abs(float(got) - ref) / ulp)
    assert worst <= 3.0, worst
    pi = np.float32(math.pi)
    for y, x, expect in ((0.0, 1.0, 0.0), (-0.0, 1.0, -0.0), (0.0, -1.0, pi), (-0.0, -1.0, -pi),
                         (1.0, 0.0, np.float32(math.pi / 2)), (-1.0, 0.0, -np.float32(math.pi / 2)),
                         (0.0, 0.0, 0.0), (0.0, -0.0, pi)):
        got = oracle.atan2(y, x)
        assert got == expect and math.copysign(1, got) == math.copysign(1, expect), (y, x, got)


def test_spring_force_is_minus_gradient_and_hessian_is_exact():
    """Spring (particle_vbd_kernels.py:1685-1735): f = -dE/dx of E = k/2 (|x0 - x1| - L0)^2; H = d^2E/dx0^2."""
    rng = np.random.default_rng(21)
    for _ in range(20):
        x = rng.normal(0, 0.3, (2, 3))
        rest = float(np.float32(rng.uniform(0.05, 0.6)))
        ke = float(np.float32(rng.uniform(10.0, 1e3)))
        f0, f1, h = oracle.spring(x, x, rest, ke, 0.0, 1.0 / 600.0, precision="f64")

        def energy(xx):
            d = xx[0] - xx[1]
            return 0.5 * ke * (np.sqrt(d @ d) - rest) ** 2

        eps = 1e-6
        for v, f in ((0, f0), (1, f1)):
            g = np.zeros(3)
            for k in range(3):
                xp, xm = x.copy(), x.copy()
                xp[v, k] += eps
                xm[v, k] -= eps
                g[k] = (energy(xp) - energy(xm)) / (2 * eps)
            np.testing.assert_allclose(f, -g, rtol=1e-6, atol=1e-6 * ke)
        jac = np.zeros((3, 3))
        for k in range(3):
            xp, xm = x.copy(), x.copy()
            xp[0, k] += eps
            xm[0, k] -= eps
            jac[:, k] = -(oracle.spring(xp, xp, rest, ke, 0.0, 1.0 / 600.0)[0] -
                          oracle.spring(xm, xm, rest, ke, 0.0, 1.0 / 600.0)[0]) / (2 * eps)
        np.testing.assert_allclose(h, jac, rtol=1e-5, atol=1e-5 * ke)


def test_spring_damping_known_answer():
    """Stretching by dL adds -kd/dt * dL along the spring direction on x0 (the opposite on x1) and kd/dt d d^T to H;
    at the rest length the elastic Hessian is k d d^T (its transverse part l0/L - 1 vanishes)."""
    dt = 0.01
    x = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 2.0]])
    xa = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.5]])
    f0, f1, h = oracle.spring(x, xa, 2.0, 50.0, 3.0, dt, precision="f64")
    # at rest length: no elastic force; damping: dir = (x0 - x1)/L = (0, 0, -1), dL = 0.5
    np.testing.assert_allclose(f0, [0.0, 0.0, 3.0 / dt * 0.5], atol=1e-9)
    np.testing.assert_allclose(f1, -f0, atol=1e-9)
    np.testing.assert_allclose(h, np.diag([0.0, 0.0, 50.0 + 3.0 / dt]), atol=1e-9)


def test_zero_length_spring_is_finite():
    x = np.array([[0.1, 0.1, 0.1], [0.1, 0.1, 0.1]])
    f0, f1, h = oracle.spring(x, x, 0.2, 10.0, 0.0, 1.0 / 600.0, precision="f32")
    assert np.isfinite(f0).all() and np.isfinite(h).all()
