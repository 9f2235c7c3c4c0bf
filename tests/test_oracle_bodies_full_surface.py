"""Pins the oracle's contacts against shapes on moving bodies and its full-surface edge/face contacts.

Every expected value is computed here in float64 from the mathematics, independently of the oracle:

* body-attached shapes (geometry/kernels.py:1061-1081, rigid_vbd_kernels.py:1365-1439 with body_q_prev,
  the one-way coupling of solver_vbd.py:2523-2526): contacts of a sphere on a rotated, translated body
  match the world-frame sphere (distance, normal, X_wb * body_pos = closest surface point, with the
  rotation built by scipy); a cloth sliding on a plane carried by a body in uniform motion is the
  cloth on the resting plane shifted by the motion (Galilean invariance pins the body point, its
  velocity (x - x_prev)/dt and its sign);
* full-surface records (soft_contacts_sdf.py:136-551): an edge's record sits at the closest point of
  the segment to the sphere (closed form of the projection, golden-section accuracy), a face's within
  the Frank-Wolfe tolerance of a dense barycentric search; against a plane the minimum is at the lower
  endpoint / corner; only features within the threshold get records, particle records come first;
* the edge/face force (rigid_vbd_kernels.py:1440-1536, particle_vbd_kernels.py:2379-2401): the force on
  each corner is minus the gradient of 1/2 ke pen(sum w_i x_i)^2 (central differences), and one Newton
  iteration of a free corner uses the block w^2 ke n n^T (known answer with numpy.linalg.solve), with
  and without the penalty ramp evaluated at the barycentric point;
* what full-surface contact is for: a coarse cloth dropped on a sphere smaller than its cells sinks
  over it with particle contacts only and stays on top with edge/face contacts.
"""

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import oracle
from newton_amd import ModelBuilder
from newton_amd.sim.collide import full_surface_capable_shapes, world_compatible_pairs


def _pose(p, q):
    return np.concatenate([np.asarray(p, np.float32), np.asarray(q, np.float32)])


def _xf(pose, x):
    """X * x for pose (p, q xyzw), float64."""
    return Rotation.from_quat(np.asarray(pose[3:], np.float64)).apply(x) + np.asarray(pose[:3], np.float64)


# ------------------------------------------------------------------ shapes on bodies
def test_body_sphere_contacts_follow_the_body_pose():
    rng = np.random.default_rng(5)
    b = ModelBuilder(up_axis="Y", gravity=0.0)
    body_pose = _pose((0.4, 0.9, -0.3), Rotation.from_rotvec([0.3, -0.8, 0.5]).as_quat())
    body = b.add_body(xform=(body_pose[:3], body_pose[3:]))
    shape_pose = _pose((0.1, -0.2, 0.05), Rotation.from_rotvec([0.0, 0.4, 0.2]).as_quat())
    radius = 0.35
    b.add_shape_sphere(body, xform=(shape_pose[:3], shape_pose[3:]), radius=radius,
                       cfg=ModelBuilder.ShapeConfig(margin=0.002))
    center = _xf(body_pose, _xf(shape_pose, np.zeros(3)))
    pts = center + rng.normal(0.0, 1.0, (80, 3)) * 0.3
    for p in pts:
        b.add_particle(tuple(p), (0.0, 0.0, 0.0), 0.1, radius=0.02)
    b.color()
    m = b.finalize(device="cpu")
    om = oracle.OracleModel(m)
    q = m.particle_q.numpy()
    margin = 0.01
    thr = margin + 0.002 + 0.02
    d = np.linalg.norm(q.astype(np.float64) - center, axis=1) - radius
    assert np.all(np.abs(d - thr) > 1e-5)
    c = oracle.Contacts(m.particle_count, "f64")
    count = oracle.create_soft_contacts(om, q, margin, c, precision="f64", body_q=m.body_q.numpy())
    hit = np.nonzero(d < thr)[0]
    assert count == len(hit) > 10
    np.testing.assert_array_equal(c.particle[:count], hit)
    for k, pi in enumerate(hit):
        n = (q[pi] - center) / np.linalg.norm(q[pi] - center)
        np.testing.assert_allclose(c.normal[k], n, atol=1e-6)
        # the record's body point is in the body frame: X_wb * body_pos is the closest surface point
        np.testing.assert_allclose(_xf(body_pose, c.body_pos[k]), center + radius * n, atol=1e-6)
    # without the body's pose the shape cannot be placed
    with pytest.raises(RuntimeError):
        oracle.create_soft_contacts(om, q, margin, c, precision="f64")


def _sliding_cloth(moving: bool, speed=(0.6, 0.0, 0.25), slide=(0.3, 0.0, -0.1)):
    """A 5x5 cloth resting on a plane carried by body 0 (normal +y), sliding with `slide`; with
    `moving` the body and the cloth additionally move with `speed`."""
    b = ModelBuilder(up_axis="Y")
    body = b.add_body(xform=((0.0, 0.0, 0.0), (0.0, 0.0, 0.0, 1.0)))
    b.add_shape_plane(plane=(0.0, 1.0, 0.0, 0.0), width=0.0, length=0.0, body=body,
                      cfg=ModelBuilder.ShapeConfig(ke=1e4, kd=50.0, mu=0.6, margin=0.0))
    v = np.asarray(slide) + (np.asarray(speed) if moving else 0.0)
    rot = Rotation.from_rotvec([-np.pi / 2, 0.0, 0.0]).as_quat()
    b.add_cloth_grid(pos=(0.0, 0.0499, 0.0), rot=rot, vel=tuple(v), dim_x=5, dim_y=5, cell_x=0.1, cell_y=0.1,
                     mass=0.01, tri_ke=1e3, tri_ka=1e3, tri_kd=1e-4, edge_ke=0.5, particle_radius=0.05)
    b.color()
    m = b.finalize(device="cpu")
    m.soft_contact_ke, m.soft_contact_kd, m.soft_contact_mu = 1e4, 50.0, 0.5
    return m


def test_plane_on_a_moving_body_is_galilean_invariant():
    speed = np.array([0.6, 0.0, 0.25])
    dt, steps, margin = 1.0 / 600.0, 40, 0.01
    runs = {}
    for moving in (False, True):
        m = _sliding_cloth(moving)
        om = oracle.OracleModel(m)
        prm = oracle.params(6)
        q, qd = m.particle_q.numpy().astype(np.float64), m.particle_qd.numpy().astype(np.float64)
        base = m.body_q.numpy()[0].astype(np.float64)
        hist = []
        n_contacts = 0
        for k in range(steps):
            pose0, pose1 = base.copy(), base.copy()
            if moving:
                pose0[:3] += speed * dt * k
                pose1[:3] += speed * dt * (k + 1)
            c = oracle.Contacts(m.particle_count, "f64")
            # the collision sees the bodies where the step starts; the step moves them to pose1
            n_contacts = oracle.create_soft_contacts(om, q, margin, c, precision="f64", body_q=pose0[None])
            oracle.set_bodies(prm, pose1[None], pose0[None])
            _, q, qd = oracle.step(om, prm, q, qd, None, c, dt, precision="f64")
            hist.append(q - (speed * dt * (k + 1) if moving else 0.0))
        assert n_contacts == m.particle_count  # every particle rests on the plane
        runs[moving] = np.stack(hist)
    # float64 rounding of the shifted coordinates, amplified by the stiff contact, stays ~1e-9; a wrong
    # body point or body velocity would be off by the motion itself (~1e-3 per substep)
    np.testing.assert_allclose(runs[True], runs[False], rtol=0, atol=1e-7)
    # and friction did act: the sliding cloth slowed down relative to the plane
    slid = runs[False][-1, :, 0].mean() - runs[False][0, :, 0].mean()
    assert 0.0 < slid < 0.3 * dt * (steps - 1)


def test_body_contacts_need_the_body_poses():
    m = _sliding_cloth(False)
    om = oracle.OracleModel(m)
    c = oracle.Contacts(m.particle_count, "f64")
    oracle.create_soft_contacts(om, m.particle_q.numpy(), 0.01, c, precision="f64", body_q=m.body_q.numpy())
    with pytest.raises(RuntimeError):
        oracle.step(om, oracle.params(2), m.particle_q.numpy(), m.particle_qd.numpy(), None, c, 1e-3,
                    precision="f64")


# ------------------------------------------------------------------ full-surface records
def _coarse_cloth(up=0.12, size=2, cell=0.5, sphere_r=0.1, plane=False, tilt=0.0, shift=(0.013, -0.021)):
    """A size x size cloth of `cell` cells at height `up` above a sphere of radius `sphere_r` at the
    origin (or a ground plane), optionally tilted about x, its corner shifted by `shift` in x / z."""
    b = ModelBuilder(up_axis="Y")
    if plane:
        b.add_ground_plane(cfg=ModelBuilder.ShapeConfig(margin=0.0))
    else:
        b.add_shape_sphere(-1, radius=sphere_r, cfg=ModelBuilder.ShapeConfig(margin=0.0))
    rot = (Rotation.from_rotvec([tilt, 0.0, 0.0]) * Rotation.from_rotvec([-np.pi / 2, 0.0, 0.0])).as_quat()
    half = 0.5 * size * cell
    b.add_cloth_grid(pos=(-half + shift[0], up, half + shift[1]), rot=rot, vel=(0.0, 0.0, 0.0), dim_x=size, dim_y=size,
                     cell_x=cell, cell_y=cell, mass=0.05, tri_ke=1e3, tri_ka=1e3, tri_kd=1e-4, edge_ke=0.5,
                     particle_radius=0.01)
    b.color()
    return b.finalize(device="cpu")


def _full_surface_oracle(m):
    om = oracle.OracleModel(m)
    cap = full_surface_capable_shapes(m)
    pw, sw = m.particle_world.numpy(), m.shape_world.numpy()
    ep = world_compatible_pairs(pw[m.edge_indices.numpy()[:, 2]], sw, m.world_count, cap)
    fp = world_compatible_pairs(pw[m.tri_indices.numpy()[:, 0]], sw, m.world_count, cap)
    om.set_full_surface_pairs(ep, fp)
    return om, ep, fp


def _closest_on_segment(p, q, c):
    u = np.clip(np.dot(c - p, q - p) / np.dot(q - p, q - p), 0.0, 1.0)
    return u, p + u * (q - p)


def test_edge_and_face_records_against_a_sphere():
    # one 0.5 m cell 0.11 m above a 0.1 m sphere: its diagonal and both triangles pass within the
    # threshold, its corners are far away
    m = _coarse_cloth(up=0.11, size=1, cell=0.5)
    om, ep, fp = _full_surface_oracle(m)
    q = m.particle_q.numpy()
    margin, r_s, rad = 0.01, 0.1, 0.01
    thr = margin + rad
    cap = len(om.pairs) // 2 + len(ep) + len(fp)
    c = oracle.Contacts(cap, "f64")
    count = oracle.create_soft_contacts(om, q, margin, c, precision="f64")
    x = q.astype(np.float64)
    # particles are far from the small sphere: no particle record
    assert np.all(np.linalg.norm(x, axis=1) - r_s > thr + 0.05)
    # expected edge records: closest point of each segment to the centre
    exp_edges = []
    for e, s in ep:
        v0, v1 = m.edge_indices.numpy()[e, 2:]
        u, pt = _closest_on_segment(x[v0], x[v1], np.zeros(3))
        phi = np.linalg.norm(pt) - r_s
        assert abs(phi - thr) > 1e-3
        if phi < thr:
            exp_edges.append((v0, v1, u, pt))
    # face minima by a dense barycentric search of each triangle's distance to the centre.  Frank-Wolfe
    # approaches the minimum from above: within ~3e-3 when it lies inside the triangle, more slowly when
    # it lies on an edge (the zig-zag of the method), so a face whose minimum is on its boundary may
    # stop short of the threshold
    g = np.linspace(0.0, 1.0, 401)
    wa, wb = np.meshgrid(g, g, indexing="ij")
    keep = wa + wb <= 1.0
    wa, wb = wa[keep], wb[keep]
    faces = []
    for t, s in fp:
        ia, ib, ic = m.tri_indices.numpy()[t]
        pts = wa[:, None] * x[ia] + wb[:, None] * x[ib] + (1.0 - wa - wb)[:, None] * x[ic]
        d = np.linalg.norm(pts, axis=1)
        k = int(np.argmin(d))
        interior = min(wa[k], wb[k], 1.0 - wa[k] - wb[k]) > 0.02
        faces.append(((ia, ib, ic), d[k] - r_s, interior))
    assert exp_edges and any(f[2] and f[1] < thr - 5e-3 for f in faces)
    recs = [(tuple(c.indices[k]), c.bary[k], c.body_pos[k], c.normal[k]) for k in range(count)]
    for k, (v0, v1, u, pt) in enumerate(exp_edges):
        idx, bary, bp, n = recs[k]
        assert idx == (v0, v1, -1)
        assert abs(bary[1] - u) < 1e-4 and bary[0] + bary[1] == pytest.approx(1.0, abs=1e-12) and bary[2] == 0.0
        n_exp = pt / np.linalg.norm(pt)
        np.testing.assert_allclose(n, n_exp, atol=1e-4)
        np.testing.assert_allclose(bp, r_s * n_exp, atol=1e-4)  # the surface point below the edge
    face_recs = {r[0]: r for r in recs[len(exp_edges):]}
    assert len(face_recs) == count - len(exp_edges)
    for idx, phi, interior in faces:
        if phi > thr + 1e-3:
            assert idx not in face_recs
            continue
        if interior and phi < thr - 5e-3:
            assert idx in face_recs
        if idx not in face_recs:
            continue
        _, bary, bp, n = face_recs[idx]
        assert np.all(bary >= -1e-9) and bary.sum() == pytest.approx(1.0, abs=1e-9)
        pt = bary[0] * x[idx[0]] + bary[1] * x[idx[1]] + bary[2] * x[idx[2]]
        phi_rec = np.linalg.norm(pt) - r_s
        assert phi - 1e-6 <= phi_rec <= phi + (3e-3 if interior else 1e-2)
        assert phi_rec < thr
        np.testing.assert_allclose(bp, r_s * pt / np.linalg.norm(pt), atol=1e-6)
        np.testing.assert_allclose(n, pt / np.linalg.norm(pt), atol=1e-6)
    # every record names the sphere; particle pairs exist but produced nothing
    assert np.all(c.shape[:count] == 0)


def test_edge_and_face_records_against_a_plane_sit_at_the_lowest_corner():
    m = _coarse_cloth(up=0.004, size=1, cell=0.4, plane=True, tilt=0.3)
    om, ep, fp = _full_surface_oracle(m)
    q = m.particle_q.numpy()
    x = q.astype(np.float64)
    c = oracle.Contacts(len(om.pairs) // 2 + len(ep) + len(fp), "f64")
    count = oracle.create_soft_contacts(om, q, 0.01, c, precision="f64")
    thr = 0.01 + 0.01
    n_particle = int(np.sum(x[:, 1] < thr))
    assert n_particle >= 1
    assert np.all(c.indices[:n_particle, 1] == -1)
    for k in range(n_particle, count):
        idx = [i for i in c.indices[k] if i >= 0]
        w = c.bary[k][: len(idx)]
        h = x[idx, 1]
        # the plane's distance is linear: the optimiser ends at (or within golden-section
        # tolerance of) the lowest corner -- any of them when several are equally low -- and the
        # record's height is that corner's
        assert w[h > h.min() + 1e-6].sum() < 1e-4
        assert float(w @ h) == pytest.approx(h.min(), abs=1e-4)
        np.testing.assert_allclose(c.normal[k], [0.0, 1.0, 0.0], atol=1e-6)  # float32 plane rotation
        assert h.min() < thr


# ------------------------------------------------------------------ edge/face force and Hessian
def _one_triangle(ke=1e4, pinned=True):
    """Corner 0 free (mass 0.02), corners 1 and 2 pinned (or free); no elastic stiffness; a ground plane
    with damping and friction off."""
    b = ModelBuilder(up_axis="Y")
    b.add_ground_plane(cfg=ModelBuilder.ShapeConfig(ke=ke, kd=0.0, mu=0.0, margin=0.001))
    m12 = 0.0 if pinned else 0.03
    b.add_particle((0.05, 0.003, 0.02), (0.0, 0.0, 0.0), 0.02, radius=0.004)
    b.add_particle((0.31, 0.012, -0.07), (0.0, 0.0, 0.0), m12, radius=0.006)
    b.add_particle((0.12, 0.009, 0.27), (0.0, 0.0, 0.0), m12, radius=0.005)
    b.add_triangle(0, 1, 2, tri_ke=0.0, tri_ka=0.0, tri_kd=0.0)
    b.color()
    m = b.finalize(device="cpu")
    m.soft_contact_ke, m.soft_contact_kd, m.soft_contact_mu = ke, 0.0, 0.0
    return m


def _face_record(bary, body_pos=(0.1, 0.0, 0.05)):
    c = oracle.Contacts(1, "f64")
    c.count[0] = 1
    c.indices[0] = (0, 1, 2)
    c.bary[0] = bary
    c.shape[0] = 0
    c.body_pos[0] = body_pos
    c.normal[0] = (0.0, 1.0, 0.0)
    return c


def test_face_contact_force_is_minus_the_energy_gradient():
    m = _one_triangle(pinned=False)
    om = oracle.OracleModel(m)
    prm = oracle.params(1, k_start=-1.0)
    w = np.array([0.5, 0.3, 0.2])
    c = _face_record(w)
    rng = np.random.default_rng(3)
    x = m.particle_q.numpy().astype(np.float64) + rng.normal(0.0, 1e-3, (3, 3))
    ke = 1e4  # mean of the particle and shape ke
    radius, margin = 0.006, 0.001  # largest corner radius, shape margin

    def energy(xx):
        pen = -((xx.T @ w)[1] - 0.0 - radius - margin)
        return 0.5 * ke * max(pen, 0.0) ** 2

    assert energy(x) > 0.0
    g_c, _ = oracle.residual(om, prm, c, x, x, x, 1e-3)
    g_0, _ = oracle.residual(om, prm, None, x, x, x, 1e-3)
    f = g_c - g_0
    h = 1e-7
    for i in range(3):
        for j in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, j] += h
            xm[i, j] -= h
            assert f[i, j] == pytest.approx(-(energy(xp) - energy(xm)) / (2 * h), rel=1e-6, abs=1e-6)
    # the force is the contact force at the point, split by the weights
    np.testing.assert_allclose(f[1], f[0] * w[1] / w[0], rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("beta", [0.0, 2e5])
def test_one_iteration_with_a_face_contact_known_answer(beta):
    """x0 at rest, forward step y = x0 + g h^2, then one Newton step of the free corner:
    (m/h^2 I + w0^2 k n n^T) dx = m/h^2 (y - x) + w0 k pen n at x = y, with the penalty ramped first
    (k = min(k_start + beta pen, ke)) when beta > 0."""
    ke, k_start = 1e4, 3e3
    m = _one_triangle(ke)
    om = oracle.OracleModel(m)
    w = np.array([0.6, 0.25, 0.15])
    c = _face_record(w)
    dt = 1.0 / 300.0
    prm = oracle.params(1, avbd_beta=beta, k_start=k_start if beta > 0 else -1.0)
    x0 = m.particle_q.numpy().astype(np.float64)
    _, q1, _ = oracle.step(om, prm, x0, np.zeros((3, 3)), None, c, dt, precision="f64")
    # the model's float32 constants, exactly as stored
    grav = m.gravity.numpy()[0].astype(np.float64)
    radius = float(m.particle_radius.numpy().max())
    margin = float(m.shape_margin.numpy()[0])
    mass = float(m.particle_mass.numpy()[0])
    y = x0.copy()
    y[0] = x0[0] + grav * dt * dt
    n = np.array([0.0, 1.0, 0.0])
    pen = -(n @ (w @ y) - radius - margin)
    assert pen > 0
    k = min(k_start + beta * pen, ke) if beta > 0 else ke
    hm = mass / dt**2 * np.eye(3) + w[0] ** 2 * k * np.outer(n, n)
    f = mass / dt**2 * (y[0] - y[0]) + w[0] * k * pen * n
    expect = y[0] + np.linalg.solve(hm, f)
    np.testing.assert_allclose(q1[0], expect, rtol=0, atol=1e-14)
    np.testing.assert_array_equal(q1[1:], x0[1:])


# ------------------------------------------------------------------ what full surface is for
def test_full_surface_contact_keeps_a_coarse_cloth_on_a_small_sphere():
    """A 4x4 cloth of 0.25 m cells dropped on a 0.08 m sphere that sits under the middle of a cell:
    particle contacts alone never touch it and the cloth falls through; edge and face contacts hold
    the cloth on top of it."""
    out = {}
    for full in (False, True):
        m = _coarse_cloth(up=0.12, size=4, cell=0.25, sphere_r=0.08, shift=(0.135, -0.118))
        m.soft_contact_ke, m.soft_contact_kd, m.soft_contact_mu = 2e4, 1.0, 0.3
        om, ep, fp = _full_surface_oracle(m) if full else (oracle.OracleModel(m), [], [])
        prm = oracle.params(8)
        q, qd = m.particle_q.numpy().astype(np.float32), m.particle_qd.numpy().astype(np.float32)
        dt = 1.0 / 600.0
        cap = len(om.pairs) // 2 + len(ep) + len(fp)
        for _ in range(200):
            c = oracle.Contacts(cap, "f32")
            oracle.create_soft_contacts(om, q, 0.01, c)
            _, q, qd = oracle.step(om, prm, q, qd, None, c, dt)
        assert np.isfinite(q).all()
        out[full] = float(q[:, 1].max())
    assert out[False] < -0.2  # fell past the sphere (free fall for 1/3 s is 0.55 m)
    assert out[True] > 0.05   # held up by the sphere (top at 0.08)
