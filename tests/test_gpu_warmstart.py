"""GPU: libvbd_ws.so (features, BF16 MFMA MLP forward / backward, Adam) against the CPU oracle.

BF16 rounding points are the same on both sides; the device accumulates in FP32
(MFMA) where the oracle uses float64, so a rounding boundary can flip one BF16
ulp (2^-8 relative) of an activation now and then.  Tolerances:
* features: at most 1 BF16 ulp per entry;
* forward output: |dev - ref| <= 1e-2 * (1 + |ref|) per entry, relative Frobenius error <= 5e-3;
* gradient: relative Frobenius error <= 2e-2 per parameter block, cosine >= 0.999;
* Adam: float32 step equal to the oracle's within 1e-6 relative.
"""

import numpy as np
import pytest

import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def _rand_feat(rng, n):
    return mo.bf16_round(rng.normal(0, 1, (n, mo.N_IN)).astype(np.float32))


def _to_dev_bf16(x, dev):
    import torch

    return torch.as_tensor(mo.bf16_bits(x).view(np.int16), device=dev)


def test_forward_matches_oracle(cuda_device):
    import torch

    from newton_amd.warmstart import WarmStartMLP

    rng = np.random.default_rng(0)
    mlp = WarmStartMLP(cuda_device, seed=1)
    p = rng.normal(0, 0.15, mo.N_PARAMS).astype(np.float32)
    mlp.load_params(p)
    # 40,000 rows: more 128-row groups than work-groups, so each work-group loops
    for n in (1, 15, 64, 200, 1000, 40000):
        x = _rand_feat(rng, n)
        out = mlp.forward(_to_dev_bf16(x, cuda_device))
        torch.cuda.synchronize()
        got = out.cpu().numpy().astype(np.float64)
        ref, _, _ = mo.forward(x, p.astype(np.float64), bf16=True)
        assert np.all(np.abs(got - ref) <= 1e-2 * (1.0 + np.abs(ref))), n
        assert np.linalg.norm(got - ref) <= 5e-3 * np.linalg.norm(ref), n


def test_packed_weights_are_bf16_of_params(cuda_device):
    import torch

    from newton_amd.warmstart import WarmStartMLP

    mlp = WarmStartMLP(cuda_device, seed=3)
    torch.cuda.synchronize()
    P = mo.unpack(mlp.params.cpu().numpy())
    pk = mlp.packed.cpu().numpy().view(np.uint16)
    o = 0
    for l, (fi, fo) in enumerate([(48, 128), (128, 128), (128, 128)], start=1):
        W = P[f"W{l}"]
        np.testing.assert_array_equal(pk[o:o + fi * fo].reshape(fo, fi), mo.bf16_bits(W.T.copy()))
        o += fi * fo
        np.testing.assert_array_equal(pk[o:o + fi * fo].reshape(fi, fo), mo.bf16_bits(W))
        o += fi * fo
    wt4 = pk[o:o + 16 * 128].reshape(16, 128)
    np.testing.assert_array_equal(wt4[:3], mo.bf16_bits(P["W4"].T.copy()))
    assert not wt4[3:].any()
    o += 16 * 128
    w4 = pk[o:o + 128 * 16].reshape(128, 16)
    np.testing.assert_array_equal(w4[:, :3], mo.bf16_bits(P["W4"]))
    assert not w4[:, 3:].any()


@pytest.mark.parametrize("n,weighted", [(64, False), (300, True), (1000, False)])
def test_gradient_matches_oracle(cuda_device, n, weighted):
    import torch

    from newton_amd.warmstart import WarmStartMLP

    rng = np.random.default_rng(n)
    mlp = WarmStartMLP(cuda_device, seed=2)
    p = rng.normal(0, 0.15, mo.N_PARAMS).astype(np.float32)
    mlp.load_params(p)
    x = _rand_feat(rng, n)
    t = rng.normal(0, 1, (n, 3)).astype(np.float32)
    w = rng.uniform(0.0, 2.0, n).astype(np.float32) if weighted else None
    loss, grad = mlp.gradient(_to_dev_bf16(x, cuda_device), torch.as_tensor(t, device=cuda_device),
                              None if w is None else torch.as_tensor(w, device=cuda_device))
    torch.cuda.synchronize()
    g = grad.cpu().numpy().astype(np.float64)
    ref_loss, ref = mo.loss_and_grad(x, p.astype(np.float64), t, w, bf16=True)
    assert abs(float(loss.item()) - ref_loss) <= 1e-3 * ref_loss
    G, R = mo.unpack(g), mo.unpack(ref)
    for name, _ in mo.SHAPES:
        a, b = G[name].reshape(-1), R[name].reshape(-1)
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        cos = float(a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30))
        assert rel <= 2e-2 and cos >= 0.999, (name, rel, cos)


def test_adam_matches_oracle(cuda_device):
    import torch

    from newton_amd.warmstart import WarmStartMLP

    rng = np.random.default_rng(9)
    mlp = WarmStartMLP(cuda_device, seed=4)
    p0 = mlp.params.cpu().numpy().copy()
    m = np.zeros_like(p0)
    v = np.zeros_like(p0)
    x = _rand_feat(rng, 128)
    t = rng.normal(0, 1, (128, 3)).astype(np.float32)
    xd, td = _to_dev_bf16(x, cuda_device), torch.as_tensor(t, device=cuda_device)
    p = p0
    for step in (1, 2, 3):
        _, grad = mlp.gradient(xd, td)
        g = grad.cpu().numpy().copy()
        mlp.params.copy_(torch.as_tensor(p, device=cuda_device))
        mlp.pack()
        mlp.train_step(xd, td, lr=1e-3)
        torch.cuda.synchronize()
        p, m, v = mo.adam(p, g, m, v, 1e-3, 0.9, 0.999, 1e-8, step)
        np.testing.assert_allclose(mlp.params.cpu().numpy(), p, rtol=1e-6, atol=1e-9)
        np.testing.assert_allclose(mlp.m.cpu().numpy(), m, rtol=1e-6, atol=1e-12)


@pytest.mark.parametrize("verts", [12, 128, "cubes"])
def test_features_match_oracle(cuda_device, verts):
    """Features of the wind cloth, also at its stated size (128 x 128, 16,384 particles), and of 24
    tet cubes (the batched tet workload's 1-ring: tet and surface-triangle edges, tet rest volumes);
    then the MLP's forward pass over all rows of the 128 x 128 cloth against the oracle's."""
    import torch

    from newton_amd import _native
    from newton_amd.sim.collide import current_stream_handle
    from newton_amd.warmstart import vertex_neighbours
    from newton_amd import scenes

    if verts == "cubes":
        model, state, info = scenes.soft_cubes_batched(cuda_device, worlds=24, cells_min=2, cells_max=5)
        assert model.tet_count > 0
    else:
        model, state, info = scenes.cloth_wind(cuda_device, verts_per_side=verts)
    off, nbr = vertex_neighbours(model)
    n = model.particle_count
    rng = np.random.default_rng(11)
    x0 = state.particle_q.cpu().numpy()
    y = (x0 + rng.normal(0, 1e-3, x0.shape)).astype(np.float32)
    corr = rng.normal(0, 3e-5, x0.shape).astype(np.float32)
    acc = rng.normal(0, 2, x0.shape).astype(np.float32)
    mass = model.particle_mass.cpu().numpy()
    if verts == "cubes":
        from newton_amd.warmstart import vertex_areas

        area = vertex_areas(model)
        assert (area > 0).all()
    else:
        area = rng.uniform(0.5, 1.5, n).astype(np.float32)
    scales = np.array([1.0 / 600, 9.81 / 360000, 0.09, 1e-3, 7e-3], np.float32)
    dt = 1.0 / 600
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=cuda_device)  # noqa: E731
    feat = torch.empty((n, 48), dtype=torch.int16, device=cuda_device)
    args = [dev(off), dev(nbr), dev(x0), dev(y), dev(corr), dev(acc), dev(mass), dev(area)]
    sc = dev(scales)
    _native.ws_check(_native.ws_lib().vbd_ws_features(current_stream_handle(cuda_device), n,
                                                      *[a.data_ptr() for a in args], dt, sc.data_ptr(),
                                                      feat.data_ptr()), "vbd_ws_features")
    torch.cuda.synchronize()
    got = mo.bf16_to_f32(feat.cpu().numpy().view(np.uint16))
    ref = mo.bf16_round(mo.features(off, nbr, x0, y, corr, acc, mass, area, dt, scales))
    ulp = np.abs(ref) * 2.0 ** -7 + 1e-30
    bad = np.abs(got - ref) > ulp
    assert not bad.any(), (np.argwhere(bad)[:5], got[bad][:5], ref[bad][:5])
    if verts == 128:
        from newton_amd.warmstart import WarmStartMLP

        mlp = WarmStartMLP(cuda_device, seed=5)
        out = mlp.forward(feat)
        torch.cuda.synchronize()
        fwd = out.cpu().numpy().astype(np.float64)
        fref, _, _ = mo.forward(got, mlp.params.cpu().numpy().astype(np.float64), bf16=True)
        assert np.all(np.abs(fwd - fref) <= 1e-2 * (1.0 + np.abs(fref)))
        assert np.linalg.norm(fwd - fref) <= 5e-3 * np.linalg.norm(fref)


def test_training_reduces_loss(cuda_device):
    """A smooth synthetic map is learned: the loss falls by 10x within 300 Adam steps."""
    import torch

    from newton_amd.warmstart import WarmStartMLP

    rng = np.random.default_rng(12)
    n = 4096
    x = _rand_feat(rng, n)
    A = rng.normal(0, 0.3, (mo.N_IN, 3))
    t = np.tanh(x @ A).astype(np.float32)
    mlp = WarmStartMLP(cuda_device, seed=5)
    xd, td = _to_dev_bf16(x, cuda_device), torch.as_tensor(t, device=cuda_device)
    first = float(mlp.train_step(xd, td, lr=3e-3).item())
    for _ in range(300):
        last = mlp.train_step(xd, td, lr=3e-3)
    torch.cuda.synchronize()
    assert float(last.item()) < 0.1 * first, (first, float(last.item()))


def test_trained_network_gains_over_reuse(cuda_device):
    """What the trained network gains over the method without it: on held-out substeps of the wind
    cloth, the learned start y + c + MLP lies closer to the converged solution x* than the
    network-free start y + c (the correction carried over from the previous substep).

    Trained on the GPU as bench.py trains it (deployed trajectory: each substep starts from the better
    of y + c and y, iterates to 0.3 r0, then on to 150 iterations for the target), at 32 x 32."""
    import torch

    from newton_amd import scenes
    from newton_amd.solvers import SolverVBD
    from newton_amd.warmstart import LearnedWarmStart

    model, s0, info = scenes.cloth_wind(cuda_device, verts_per_side=32)
    wind, dt, sub = info["wind"], info["dt"], info["substeps"]
    s1 = model.state()
    solver = SolverVBD(model, iterations=10)
    ws = LearnedWarmStart(model, dt, seed=0, blend=1.0)  # the network's full prediction
    mass = model.particle_mass

    def substep(record):
        s0.particle_f.copy_(wind.force(s0.particle_q, mass))
        solver.begin_step(s0, None, dt)
        feat = ws.features(solver, s0).clone()
        cands = [c.clone() for c in ws(solver, s0, dt)]
        y = s0.particle_q.clone()
        r0 = float(solver.residual(s0, dt).item())
        trial = solver.set_iterate(s0, cands[0], commit=False)
        if float(solver.residual(type("S", (), {"particle_q": trial})(), dt).item()) < r0:
            solver.set_iterate(s0, cands[0], commit=True)
        its, _ = solver.iterate_to_residual(s0, dt, 0.09 * r0, 150)
        x_tol = s0.particle_q.clone()
        solver.iterate(s0, dt, 150 - its)
        record(feat, ws.target(solver, s0), cands, s0.particle_q.clone(), y)
        solver.set_iterate(s0, x_tol, commit=True)
        ws.observe(solver, s0)
        solver.end_step(s0, s1, dt)
        s0.particle_q.copy_(s1.particle_q)
        s0.particle_qd.copy_(s1.particle_qd)

    feats, targs = [], []
    for frame in range(24):
        wind.redraw(frame)
        for _ in range(sub):
            substep(lambda f, t, *_: (feats.append(f), targs.append(t)))
    X, T = torch.cat(feats), torch.cat(targs).contiguous()
    W = ws.active.float().repeat(len(feats))
    gen = torch.Generator(device=cuda_device)
    gen.manual_seed(0)
    for step in range(600):
        idx = torch.randint(0, X.shape[0], (16384,), device=cuda_device, generator=gen)
        ws.mlp.train_step(X[idx].contiguous(), T[idx].contiguous(), W[idx].contiguous(),
                          lr=max(1e-3 * 0.5 * (1.0 + np.cos(np.pi * step / 600)), 1e-5))
    act = ws.active
    err = {"reuse": [], "learned": [], "inertial": []}

    def measure(f, t, cands, x_star, y):
        for name, c in (("reuse", cands[0]), ("learned", cands[1]), ("inertial", y)):
            err[name].append(float(torch.linalg.norm((c - x_star)[act]).item()))

    for frame in range(24, 30):
        wind.redraw(frame)
        for _ in range(sub):
            substep(measure)
    e = {k: float(np.mean(v)) for k, v in err.items()}
    print(f"held-out distance to x*: inertial {e['inertial']:.4e}, y + c {e['reuse']:.4e}, learned {e['learned']:.4e}")
    assert e["reuse"] < e["inertial"]
    assert e["learned"] < 0.9 * e["reuse"], e


def test_smooth_matches_oracle(cuda_device):
    """vbd_ws_smooth on the 128x128 wind cloth's 1-ring (pinned top row inactive), two fields side by
    side, 64 sweeps: bit-identical to the oracle's float32 restatement."""
    import torch

    from newton_amd import scenes
    from newton_amd.warmstart import LearnedWarmStart

    model, state, info = scenes.cloth_wind(cuda_device)
    ws = LearnedWarmStart(model, info["dt"], smooth_sweeps=64)
    rng = np.random.default_rng(4)
    v = rng.normal(0, 1e-4, (model.particle_count, 6)).astype(np.float32)
    got = ws.smooth(torch.as_tensor(v, device=cuda_device)).cpu().numpy()
    torch.cuda.synchronize()
    ref = mo.smooth(ws.nbr_off.cpu().numpy(), ws.nbr.cpu().numpy(), ws._act.cpu().numpy()[:, 0], v, 64)
    np.testing.assert_array_equal(got, ref)
