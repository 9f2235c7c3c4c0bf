"""CPU pins of the learned warm start's oracle (oracle/mlp_oracle.py).

* BF16 rounding equals torch's float32 -> bfloat16 conversion (round to nearest even);
* the backward pass equals finite differences of the float64 loss (rounding off);
* with rounding on, the gradient stays close to the exact one (BF16 noise only);
* Adam's first step moves every parameter by lr * sign(g);
* features: translation invariance, the bias column, neighbour statistics on a tiny mesh.
"""

import numpy as np
import pytest
import torch

import mlp_oracle as mo


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([
        rng.normal(0, 1, 4000), rng.normal(0, 1e-30, 500), rng.normal(0, 1e30, 500),
        np.array([0.0, -0.0, 1.0, -1.0, 1.00390625, 1.005859375, 1.0078125, 3.0e38, 1e-45, np.inf, -np.inf]),
    ]).astype(np.float32)
    # exact ties: mantissa ...1000 0000 0000 0000 (round to even both ways)
    ties = (np.arange(1, 200, dtype=np.uint32) << 16 | 0x8000).view(np.float32)
    x = np.concatenate([x, ties, -ties])
    ref = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    np.testing.assert_array_equal(mo.bf16_bits(x), ref)
    np.testing.assert_array_equal(mo.bf16_round(x), torch.from_numpy(x).to(torch.bfloat16).float().numpy())


def _params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    return rng.normal(0, scale, mo.N_PARAMS)


def test_backward_equals_finite_differences():
    rng = np.random.default_rng(1)
    x = rng.normal(0, 1, (7, mo.N_IN))
    t = rng.normal(0, 1, (7, 3))
    w = rng.uniform(0.5, 2.0, 7)
    p = _params(2)
    _, g = mo.loss_and_grad(x, p, t, w, bf16=False)
    idx = np.concatenate([rng.choice(mo.N_PARAMS, 60, replace=False),
                          np.array([0, 6143, 6144 + 5, mo.N_PARAMS - 1, mo.N_PARAMS - 4])])
    eps = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += eps
        pm[k] -= eps
        fd = (mo.loss_and_grad(x, pp, t, w, bf16=False)[0] - mo.loss_and_grad(x, pm, t, w, bf16=False)[0]) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_bf16_gradient_close_to_exact():
    rng = np.random.default_rng(3)
    x = mo.bf16_round(rng.normal(0, 1, (256, mo.N_IN)).astype(np.float32))
    t = rng.normal(0, 1, (256, 3))
    p = _params(4, 0.1)
    _, g_exact = mo.loss_and_grad(x, p, t, bf16=False)
    _, g_bf = mo.loss_and_grad(x, p, t, bf16=True)
    rel = np.linalg.norm(g_bf - g_exact) / np.linalg.norm(g_exact)
    assert 0.0 < rel < 0.05


def test_adam_first_step_is_signed_lr():
    rng = np.random.default_rng(5)
    p = rng.normal(0, 1, 100).astype(np.float32)
    g = rng.normal(0, 1, 100).astype(np.float32)
    p1, m1, v1 = mo.adam(p, g, np.zeros(100), np.zeros(100), 1e-2, 0.9, 0.999, 1e-12, 1)
    np.testing.assert_allclose(p - p1, 1e-2 * np.sign(g), rtol=1e-4)
    np.testing.assert_allclose(m1, 0.1 * g, rtol=1e-6)
    np.testing.assert_allclose(v1, 0.001 * g * g, rtol=1e-4)  # 1 - 0.999 in float32


def _tiny_mesh():
    # two triangles (0,1,2), (1,3,2): CSR 1-ring
    nbr = {0: [1, 2], 1: [0, 2, 3], 2: [0, 1, 3], 3: [1, 2]}
    off = np.zeros(5, dtype=np.int32)
    off[1:] = np.cumsum([len(nbr[i]) for i in range(4)])
    return off, np.concatenate([nbr[i] for i in range(4)]).astype(np.int32)


def test_features_known_values_and_translation_invariance():
    off, nbr = _tiny_mesh()
    rng = np.random.default_rng(6)
    x0 = rng.normal(0, 0.1, (4, 3)).astype(np.float32)
    y = (x0 + rng.normal(0, 1e-3, (4, 3))).astype(np.float32)
    corr = rng.normal(0, 1e-5, (4, 3)).astype(np.float32)
    acc = rng.normal(0, 2, (4, 3)).astype(np.float32)
    mass = np.array([0.0, 1.0, 2.0, 1.0], np.float32)
    area = np.array([1.0, 2.0, 2.0, 1.0], np.float32)
    sc = [1e-3, 1e-5, 0.1, 1.0, 1.5]
    dt = 1.0 / 600.0
    f = mo.features(off, nbr, x0, y, corr, acc, mass, area, dt, sc)
    np.testing.assert_array_equal(f[:, 47], 1.0)
    np.testing.assert_allclose(f[:, 0:3], (y - x0) / np.float32(1e-3), rtol=1e-6)
    np.testing.assert_allclose(f[:, 40], [2 / 6, 3 / 6, 3 / 6, 2 / 6], rtol=1e-7)
    # vertex 1's neighbours are 0, 2, 3; vertex 0 is pinned
    np.testing.assert_allclose(f[1, 39], 1.0 / 3.0, rtol=1e-6)
    np.testing.assert_allclose(f[1, 44:47], (x0[0] - x0[1]) / 3 / np.float32(0.1), rtol=1e-5)
    a = (y - x0) / np.float32(1e-3)
    np.testing.assert_allclose(f[1, 6:9], a[[0, 2, 3]].mean(axis=0) - a[1], rtol=1e-4, atol=1e-6)
    np.testing.assert_allclose(f[1, 12:15], a[[0, 2, 3]].max(axis=0) - a[1], rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(f[2, 30:33], acc[2] * np.float32(dt * dt) / np.float32(1e-3), rtol=1e-5)
    shift = np.float32([0.25, -0.5, 0.125])  # exact in binary: translation changes no feature by more than rounding
    g = mo.features(off, nbr, x0 + shift, y + shift, corr, acc, mass, area, dt, sc)
    np.testing.assert_allclose(g, f, atol=2e-3)


@pytest.mark.parametrize("bf16", [False, True])
def test_forward_shapes_and_zero_last_layer(bf16):
    rng = np.random.default_rng(7)
    x = rng.normal(0, 1, (5, mo.N_IN))
    P = mo.unpack(_params(8))
    P["W4"][:] = 0.0
    P["b4"][:] = [0.5, -1.0, 2.0]
    out, hs, zs = mo.forward(x, mo.pack(P), bf16)
    assert out.shape == (5, 3) and len(hs) == 3 and hs[0].shape == (5, 128)
    np.testing.assert_allclose(out, np.tile([0.5, -1.0, 2.0], (5, 1)))


def test_smooth_oracle_is_the_masked_ring_mean_operator():
    """The oracle's low-pass equals k applications of the dense operator P = D^-1 (I + A M) with the
    inactive rows zeroed (A: adjacency, M: active mask, D: 1 + degree), in float64 up to float32
    rounding; it keeps a constant field on a mesh without inactive particles constant and leaves
    inactive particles at 0."""
    rng = np.random.default_rng(2)
    n = 30
    pairs = {(int(a), int(b)) for a, b in rng.integers(0, n, size=(70, 2)) if a != b}
    pairs |= {(b, a) for a, b in pairs}
    adj = [sorted(b for a, b in pairs if a == i) for i in range(n)]
    off = np.concatenate([[0], np.cumsum([len(x) for x in adj])]).astype(np.int32)
    nbr = np.array([j for x in adj for j in x], dtype=np.int32)
    active = (rng.random(n) > 0.2).astype(np.float32)
    v = rng.normal(size=(n, 6)).astype(np.float32)
    A = np.zeros((n, n))
    for i, x in enumerate(adj):
        A[i, x] = 1.0
    P = (np.eye(n) + A * active[None, :]) / (1.0 + A.sum(axis=1))[:, None]
    P[active == 0] = 0.0
    ref = v.astype(np.float64)
    for _ in range(5):
        ref = P @ ref
    got = mo.smooth(off, nbr, active, v, 5)
    np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-6)
    assert (got[active == 0] == 0).all()
    ones = mo.smooth(off, nbr, np.ones(n, np.float32), np.ones((n, 3), np.float32), 7)
    np.testing.assert_allclose(ones, 1.0, rtol=1e-6)
