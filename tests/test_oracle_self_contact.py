"""Pins of the oracle's cloth self-contact restatement (CPU).

The reference's own self-contact tests need the Warp example asset bunny.usd and
a CUDA BVH (newton/tests/test_collision_cloth.py:549-560), neither of which exists
here, so every check below is a mathematical one chosen so that a dropped term, a
wrong sign or a swapped role fails it:

* closest points (triangle, segment pair) never exceed a dense sampled minimum and
  reproduce their own distance; every triangle feature region is exercised;
* the contact-force norm is the derivative of its piecewise (quadratic / log-barrier /
  quadratic-extension) energy and is C1 at the joins;
* vertex-triangle and edge-edge contact forces equal minus the finite-difference
  gradient of E(distance) on every vertex they act on (anchors = positions, so no
  friction or damping), the normal Hessian block is d2E * b^2 * n n^T;
* friction opposes the tangential slip with magnitude mu * f_n in the constant stage and
  the IPC smooth profile below eps_u; damping acts only while the gap closes;
* the planar truncation factors stop every non-dummy vertex at or before the division
  plane, within [0, 1];
* detection equals an independent numpy brute force (point-triangle distance by plane
  projection / edge distances; segment distance by sampling refinement) with the
  adjacency, filter and world rules;
* a two-layer drop: without self-contact the layers pass through each other, with it
  they stay apart.
"""

import math

import numpy as np
import pytest

import oracle

RNG = np.random.default_rng(7)


def _seg_point_dist(p, a, b):
    ab = b - a
    t = np.clip(np.dot(p - a, ab) / max(np.dot(ab, ab), 1e-300), 0.0, 1.0)
    return np.linalg.norm(p - (a + t * ab))


def _point_tri_dist(p, a, b, c):
    """Independent point-triangle distance: plane projection when inside, else the three edges."""
    n = np.cross(b - a, c - a)
    nn = np.dot(n, n)
    if nn > 1e-300:
        proj = p - np.dot(p - a, n) / nn * n
        v0, v1, v2 = b - a, c - a, proj - a
        d00, d01, d11 = np.dot(v0, v0), np.dot(v0, v1), np.dot(v1, v1)
        d20, d21 = np.dot(v2, v0), np.dot(v2, v1)
        den = d00 * d11 - d01 * d01
        v = (d11 * d20 - d01 * d21) / den
        w = (d00 * d21 - d01 * d20) / den
        if v >= 0 and w >= 0 and v + w <= 1:
            return np.linalg.norm(p - proj)
    return min(_seg_point_dist(p, a, b), _seg_point_dist(p, b, c), _seg_point_dist(p, a, c))


def _seg_seg_dist(p1, q1, p2, q2):
    """Independent segment distance: the interior stationary point when it lies inside both
    segments, else the best of the four endpoint-to-segment distances."""
    d1, d2, r = q1 - p1, q2 - p2, p1 - p2
    a, e, b, c, f = d1 @ d1, d2 @ d2, d1 @ d2, d1 @ r, d2 @ r
    cands = [_seg_point_dist(p1, p2, q2), _seg_point_dist(q1, p2, q2), _seg_point_dist(p2, p1, q1),
             _seg_point_dist(q2, p1, q1)]
    den = a * e - b * b
    if den > 1e-14 * a * e:
        s = (b * f - c * e) / den
        t = (a * f - b * c) / den
        if 0.0 <= s <= 1.0 and 0.0 <= t <= 1.0:
            cands.append(np.linalg.norm(p1 + s * d1 - p2 - t * d2))
    return float(min(cands))


def test_triangle_closest_point_is_minimal_and_covers_every_feature():
    seen = set()
    for _ in range(400):
        a, b, c = RNG.normal(size=(3, 3))
        p = RNG.normal(size=3) * 2.0
        cp, bary, feat = oracle.triangle_closest(a, b, c, p)
        seen.add(feat)
        np.testing.assert_allclose(bary.sum(), 1.0, atol=1e-12)
        assert (bary >= -1e-12).all()
        np.testing.assert_allclose(bary[0] * a + bary[1] * b + bary[2] * c, cp, atol=1e-12)
        assert abs(np.linalg.norm(cp - p) - _point_tri_dist(p, a, b, c)) < 1e-9
    assert seen == set(range(7))


def test_closest_edge_edge_is_minimal_including_parallel_and_degenerate():
    cases = [RNG.normal(size=(4, 3)) for _ in range(60)]
    d = np.array([1.0, 0.5, -0.2])
    cases.append(np.array([[0, 0, 0], d, [0.1, 1.0, 0.3], [0.1, 1.0, 0.3] + 2 * d]))  # parallel
    cases.append(np.array([[0, 0, 0], [0, 0, 0], [1, 1, 0], [2, -1, 0.5]]))  # first a point
    cases.append(np.array([[0, 0, 0], [1, 0, 0], [0.5, 1, 0], [0.5, 1, 0]]))  # second a point
    cases.append(np.array([[0, 0, 0], [0, 0, 0], [1, 1, 1], [1, 1, 1]]))  # both points
    for p1, q1, p2, q2 in cases:
        s, t, dist = oracle.closest_edge_edge(p1, q1, p2, q2, 1e-12)
        assert 0.0 <= s <= 1.0 and 0.0 <= t <= 1.0
        np.testing.assert_allclose(np.linalg.norm(p1 + s * (q1 - p1) - p2 - t * (q2 - p2)), dist, atol=1e-12)
        assert dist <= _seg_seg_dist(p1, q1, p2, q2) + 1e-9
    # perpendicular skew segments: closed form
    s, t, dist = oracle.closest_edge_edge(np.array([-1.0, 0, 0]), np.array([1.0, 0, 0]), np.array([0.5, -1, 0.3]),
                                          np.array([0.5, 1, 0.3]), 1e-12)
    np.testing.assert_allclose([s, t, dist], [0.75, 0.5, 0.3], atol=1e-15)


def _barrier_energy(D, r, k):
    tau, dmin = 0.5 * r, 1e-5
    k2 = tau * tau * k
    if D >= tau:
        return 0.5 * k * (r - D) ** 2
    c1 = 0.5 * k * tau * tau + k2 * math.log(tau)
    if D > dmin:
        return -k2 * math.log(D) + c1
    e_dmin = -k2 * math.log(dmin) + c1
    # quadratic extension below d_min: E' = k2 (D - 2 dmin) / dmin^2
    return e_dmin + k2 * ((D * D / 2 - 2 * dmin * D) - (dmin * dmin / 2 - 2 * dmin * dmin)) / (dmin * dmin)


def test_self_force_norm_is_energy_derivative_and_c1():
    r, k = 0.02, 1e4
    for D in [2e-6, 8e-6, 3e-4, 4e-3, 9.99e-3, 1.01e-2, 0.015, 0.0199]:
        dE, d2E = oracle.self_force_norm(D, r, k)
        h = D * 1e-6
        fd = (_barrier_energy(D + h, r, k) - _barrier_energy(D - h, r, k)) / (2 * h)
        np.testing.assert_allclose(dE, fd, rtol=1e-6)
        fd2 = (oracle.self_force_norm(D + h, r, k)[0] - oracle.self_force_norm(D - h, r, k)[0]) / (2 * h)
        np.testing.assert_allclose(d2E, fd2, rtol=1e-5)
    for join in (0.5 * r, 1e-5):
        lo, hi = oracle.self_force_norm(join * (1 - 1e-9), r, k), oracle.self_force_norm(join * (1 + 1e-9), r, k)
        np.testing.assert_allclose(lo[0], hi[0], rtol=1e-6)
        np.testing.assert_allclose(lo[1], hi[1], rtol=1e-6)
    assert oracle.self_force_norm(0.5 * r, r, k)[0] < 0  # repulsive


def _vt_energy(x, r, k):
    cp, _, _ = oracle.triangle_closest(x[0], x[1], x[2], x[3])
    return _barrier_energy(np.linalg.norm(x[3] - cp), r, k)


def test_vt_contact_force_is_minus_gradient_of_distance_energy():
    r, k = 0.05, 2e3
    hits = 0
    for trial in range(40):
        tri = np.array([[0.0, 0, 0], [1.0, 0, 0], [0.0, 1.0, 0]]) + RNG.normal(0, 0.05, (3, 3))
        bary = RNG.dirichlet([2, 2, 2])
        p = bary @ tri + np.array([0, 0, RNG.uniform(0.004, 0.045)]) * (1 if trial % 2 else -1)
        x = np.vstack([tri, p])
        has, f, h = oracle.vt_contact(x, x, r, k, 0.0, 0.0, 1e-2, 1e-3)
        assert has
        hits += 1
        for role in range(4):
            g = np.zeros(3)
            for c in range(3):
                e = 1e-7
                xp, xm = x.copy(), x.copy()
                xp[role, c] += e
                xm[role, c] -= e
                g[c] = (_vt_energy(xp, r, k) - _vt_energy(xm, r, k)) / (2 * e)
            np.testing.assert_allclose(f[role], -g, rtol=2e-5, atol=1e-6 * max(1, np.abs(g).max()))
            # normal block: d2E * b^2 * n n^T
            cp, b3, _ = oracle.triangle_closest(*x)
            n = (p - cp) / np.linalg.norm(p - cp)
            bw = [-b3[0], -b3[1], -b3[2], 1.0][role]
            d2E = oracle.self_force_norm(np.linalg.norm(p - cp), r, k)[1]
            np.testing.assert_allclose(h[role], d2E * bw * bw * np.outer(n, n), rtol=1e-9, atol=1e-9)
    assert hits == 40
    far = np.vstack([np.eye(3), [[0.2, 0.2, 0.3]]])
    assert not oracle.vt_contact(far, far, r, k, 0.0, 0.5, 1e-2, 1e-3)[0]


def _ee_energy(x, r, k):
    return _barrier_energy(oracle.closest_edge_edge(*x, 1e-5)[2], r, k)


def test_ee_contact_force_is_minus_gradient_on_e1():
    r, k = 0.05, 2e3
    for _ in range(30):
        e1 = np.array([[-0.5, 0, 0], [0.5, 0, 0]]) + RNG.normal(0, 0.05, (2, 3))
        gap = RNG.uniform(0.004, 0.045)
        e2 = np.array([[0.1, -0.5, gap], [0.0, 0.5, gap]]) + RNG.normal(0, 0.02, (2, 3))
        x = np.vstack([e1, e2])
        has, f, _ = oracle.ee_contact(x, x, r, k, 0.0, 0.0, 1e-2, 1e-3)
        if not has:
            continue
        for role in range(2):
            g = np.zeros(3)
            for c in range(3):
                e = 1e-7
                xp, xm = x.copy(), x.copy()
                xp[role, c] += e
                xm[role, c] -= e
                g[c] = (_ee_energy(xp, r, k) - _ee_energy(xm, r, k)) / (2 * e)
            np.testing.assert_allclose(f[role], -g, rtol=2e-5, atol=1e-6 * max(1, np.abs(g).max()))


def test_friction_and_damping_known_answers():
    """Vertex p above the interior of a fixed triangle, n = +z, slipping along +x."""
    r, k, mu, eps, dt = 0.05, 1e3, 0.4, 1e-2, 1e-3
    tri = np.array([[0.0, 0, 0], [1.0, 0, 0], [0.0, 1.0, 0]])
    p = np.array([0.25, 0.25, 0.03])
    fn = -oracle.self_force_norm(0.03, r, k)[0]
    for slip in (1e-4, 5e-6):  # eps_u = 1e-5: constant stage, then the smooth stage
        anchor = np.vstack([tri, p - np.array([slip, 0, 0])])
        has, f, h = oracle.vt_contact(np.vstack([tri, p]), anchor, r, k, 0.0, mu, eps, dt)
        assert has
        eps_u = eps * dt
        f1 = 1 / slip if slip > eps_u else (2 - slip / eps_u) / eps_u
        np.testing.assert_allclose(f[3], [-mu * fn * f1 * slip, 0.0, fn], rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(h[3][:2, :2], mu * fn * f1 * np.eye(2), rtol=1e-9)
    # damping: p approaching the face (gap closing) at 0.2 m/s
    kd = 50.0
    anchor = np.vstack([tri, p + np.array([0, 0, 0.2 * dt])])
    _, f_d, h_d = oracle.vt_contact(np.vstack([tri, p]), anchor, r, k, kd, 0.0, eps, dt)
    np.testing.assert_allclose(f_d[3], [0, 0, fn + kd * 0.2], rtol=1e-9)
    np.testing.assert_allclose(h_d[3][2, 2], -oracle.self_force_norm(0.03, r, k)[1] * -1 + kd / dt, rtol=1e-9)
    # separating: no damping
    anchor = np.vstack([tri, p - np.array([0, 0, 0.2 * dt])])
    _, f_s, _ = oracle.vt_contact(np.vstack([tri, p]), anchor, r, k, kd, 0.0, eps, dt)
    np.testing.assert_allclose(f_s[3], [0, 0, fn], rtol=1e-9)


@pytest.mark.parametrize("kind", [0, 1])
def test_truncation_never_crosses_the_division_plane(kind):
    crossings = 0
    for _ in range(300):
        if kind == 1:
            tri = np.array([[0.0, 0, 0], [1.0, 0, 0], [0.0, 1.0, 0]]) + RNG.normal(0, 0.05, (3, 3))
            p = RNG.dirichlet([2, 2, 2]) @ tri + np.array([0, 0, RNG.uniform(0.01, 0.05)])
            x = np.vstack([p, tri])
        else:
            e1 = np.array([[-0.5, 0, 0.03], [0.5, 0, 0.03]]) + RNG.normal(0, 0.01, (2, 3))
            e2 = np.array([[0.0, -0.5, 0.0], [0.0, 0.5, 0.0]]) + RNG.normal(0, 0.01, (2, 3))
            x = np.vstack([e1, e2])
        delta = RNG.normal(0, 0.04, (4, 3))
        t = oracle.truncation_ts(kind, x, delta, 1e-5, 0.85)
        assert ((t >= 0) & (t <= 1)).all()
        # after truncation the two sides stay apart along z (the separating direction here)
        moved = x + t[:, None] * delta
        top = moved[:1] if kind == 1 else moved[:2]
        bottom = moved[1:] if kind == 1 else moved[2:]
        if kind == 1:
            cp, _, _ = oracle.triangle_closest(bottom[0], bottom[1], bottom[2], top[0])
            sep = np.linalg.norm(top[0] - cp)
        else:
            sep = oracle.closest_edge_edge(top[0], top[1], bottom[0], bottom[1], 1e-12)[2]
        full = x + delta
        if kind == 1:
            cpf, _, _ = oracle.triangle_closest(full[1], full[2], full[3], full[0])
            crossed = np.dot(full[0] - cpf, np.array([0, 0, 1.0])) < 0
        else:
            crossed = False
        crossings += int(crossed)
        assert sep > 0.0
    if kind == 1:
        assert crossings > 10  # the untruncated motions do penetrate


def _model_two_layers(dev="cpu", n=6, gap=0.012, vel=-1.0):
    from newton_amd import ModelBuilder
    from newton_amd.sim.builder import cloth_grid_mesh, quat_from_axis_angle

    b = ModelBuilder(up_axis="Y", gravity=0.0)
    verts, tris = cloth_grid_mesh(n - 1, n - 1, 0.02, 0.02)
    rot = quat_from_axis_angle((1.0, 0.0, 0.0), math.pi / 2)
    b.add_cloth_mesh(pos=(0.0, 0.0, 0.0), rot=rot, scale=1.0, vel=(0.0, 0.0, 0.0), vertices=verts, indices=tris,
                     density=0.1, tri_ke=1e3, tri_ka=1e3, tri_kd=0.0, edge_ke=0.01, edge_kd=0.0)
    b.add_cloth_mesh(pos=(0.003, gap, 0.004), rot=rot, scale=1.0, vel=(0.0, vel, 0.0), vertices=verts, indices=tris,
                     density=0.1, tri_ke=1e3, tri_ka=1e3, tri_kd=0.0, edge_ke=0.01, edge_kd=0.0)
    b.color(include_bending=True)
    return b.finalize(dev)


def _oracle_self(model, **kw):
    return oracle.SelfContact(model.particle_count, model.tri_count, model.edge_count,
                              q_rest=model.particle_q.cpu().numpy(), **kw)


def test_detection_matches_independent_brute_force():
    model = _model_two_layers()
    om = oracle.OracleModel(model)
    q = model.particle_q.cpu().numpy().astype(np.float64)
    q += RNG.normal(0, 0.002, q.shape)
    r = 0.015
    sc = _oracle_self(model, radius=0.01, margin=r, vt_capacity=64, ee_capacity=128)
    oracle.self_detect(om, sc, q, precision="f64")
    tri = model.tri_indices.cpu().numpy().reshape(-1, 3)
    edges = model.edge_indices.cpu().numpy().reshape(-1, 4)
    n_checked = 0
    for v in range(model.particle_count):
        want = []
        for t, (a, b, c) in enumerate(tri):
            if v in (a, b, c):
                continue
            d = _point_tri_dist(q[v], q[a], q[b], q[c])
            if abs(d - r) < 1e-9:
                pytest.skip("boundary case")
            if d < r:
                want.append(t)
        got = sc.vt_list[v][: sc.vt_count[v]]
        assert sc.vt_count[v] == len(want) and list(got) == want, v
        n_checked += len(want)
    assert n_checked > 50
    n_checked = 0
    for e in range(model.edge_count):
        a0, a1 = edges[e, 2], edges[e, 3]
        want = []
        for g in range(model.edge_count):
            b0, b1 = edges[g, 2], edges[g, 3]
            if len({a0, a1} & {b0, b1}):
                continue
            d = _seg_seg_dist(q[a0], q[a1], q[b0], q[b1])
            if d < r - 1e-7:
                want.append(g)
            elif d < r + 1e-7:
                pytest.skip("boundary case")
        got = list(sc.ee_list[e][: min(sc.ee_count[e], 128)])
        assert got == want, e
        n_checked += len(want)
    assert n_checked > 50


def test_detection_filter_and_rest_exclusion():
    model = _model_two_layers()
    om = oracle.OracleModel(model)
    q = model.particle_q.cpu().numpy()
    sc = _oracle_self(model, radius=0.01, margin=0.015)
    oracle.self_detect(om, sc, q)
    base = sc.vt_count.copy()
    v = int(np.argmax(base))
    assert base[v] >= 2
    drop = np.sort(sc.vt_list[v][:2])
    offs = np.zeros(model.particle_count + 1, np.int32)
    offs[v + 1:] = 2
    sc2 = _oracle_self(model, radius=0.01, margin=0.015, vt_filter=drop, vt_filter_offsets=offs)
    oracle.self_detect(om, sc2, q)
    assert sc2.vt_count[v] == base[v] - 2 and not set(drop) & set(sc2.vt_list[v][: sc2.vt_count[v]])
    assert (np.delete(sc2.vt_count, v) == np.delete(base, v)).all()
    # rest exclusion: the rest shape is this very configuration, so every pair is within 0.02 at rest
    sc3 = _oracle_self(model, radius=0.01, margin=0.015, rest_exclusion_radius=0.02)
    oracle.self_detect(om, sc3, q)
    assert sc3.vt_count.sum() == 0 and sc3.ee_count.sum() == 0


def _run_layers(self_contact: bool, substeps=30):
    model = _model_two_layers(vel=-2.0)
    om = oracle.OracleModel(model)
    sc = _oracle_self(model, radius=0.004, margin=0.006, detection_interval=1) if self_contact else None
    prm = oracle.params(5, self_contact=sc)
    q = model.particle_q.cpu().numpy()
    qd = model.particle_qd.cpu().numpy()
    dt = 1.0 / 600.0
    for _ in range(substeps):
        _, q, qd = oracle.step(om, prm, q, qd, None, None, dt)
    return model, q


def test_two_layer_drop_self_contact_prevents_pass_through():
    n = 36
    model, q_free = _run_layers(False)
    _, q_sc = _run_layers(True)
    lower, upper = q_free[:n, 1], q_free[n:, 1]
    assert upper.mean() < lower.mean()  # without self-contact the top layer falls through
    lower, upper = q_sc[:n, 1], q_sc[n:, 1]
    assert (upper > lower.max() - 1e-3).all() and upper.mean() > lower.mean()
    assert np.isfinite(q_sc).all()
