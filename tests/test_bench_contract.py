"""bench.py's contract (CPU): one GPU only, and the batched wind-cloth scene's layout."""

import json
import subprocess
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 2
    msg = json.loads(p.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_batched_wind_worlds_are_contiguous_copies_with_their_own_wind():
    import torch

    from newton_amd import scenes

    model, state, info = scenes.cloth_wind_batched("cpu", worlds=3, verts_per_side=8)
    one, one_state, one_info = scenes.cloth_wind("cpu", verts_per_side=8)
    n = one.particle_count
    assert model.particle_count == 3 * n and model.world_count == 3
    pw = model.particle_world.numpy()
    np.testing.assert_array_equal(pw, np.repeat(np.arange(3), n))
    for w in range(3):
        assert torch.equal(state.particle_q[w * n:(w + 1) * n], one_state.particle_q)
    f = info["wind"].force(state.particle_q, model.particle_mass)
    f1 = one_info["wind"].force(one_state.particle_q, one.particle_mass)
    assert torch.allclose(f[:n], f1, rtol=1e-6, atol=1e-7)  # world 0 has the single-world wind (seed 3)
    assert not torch.allclose(f[n:2 * n], f1)  # world 1 its own (seed 4)
    # every world keeps the pinned top row of the single-world scene
    assert torch.equal(model.particle_mass[:n], one.particle_mass)
    assert torch.equal(model.particle_mass[2 * n:], one.particle_mass)
