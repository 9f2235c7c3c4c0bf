"""GPU: batched worlds (many independent meshes in one launch per color pass) against the FP32 oracle.

Small instances of the 4096-world cloth and the 512-cube tet workloads: the
oracle steps the same merged model (world-filtered contact pairs, merged color
groups), and positions / velocities must be bit-identical after several
substeps.  At full size the checks are size-independent: every world stays
finite, and pinned particles do not move.
"""

import numpy as np
import pytest

from _sim import gpu_run, initial, oracle_run

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("worlds", [3, 17])
def test_cloth_worlds_bitexact(cuda_device, worlds):
    from newton_amd import scenes

    model, state, info = scenes.cloth_batched(cuda_device, worlds=worlds, side_min=6, side_max=14)
    q0, qd0 = initial(state)
    q_gpu, qd_gpu, _ = gpu_run(model, state, info, 8)
    q_ref, qd_ref, _ = oracle_run(model, q0, qd0, info, 8)
    np.testing.assert_array_equal(q_gpu, q_ref)
    np.testing.assert_array_equal(qd_gpu, qd_ref)


def test_cloth_worlds_wave_aligned_with_contacts_bitexact(cuda_device):
    """Wave-aligned item ranges on uneven worlds resting on the ground (contact lanes after the padded
    static ranges)."""
    from newton_amd import scenes

    model, state, info = scenes.cloth_batched(cuda_device, worlds=9, side_min=6, side_max=14, height=0.12)
    q0, qd0 = initial(state)
    q_gpu, qd_gpu, _ = gpu_run(model, state, info, 8, aligned=True, waves=7)
    q_ref, qd_ref, _ = oracle_run(model, q0, qd0, info, 8)
    np.testing.assert_array_equal(q_gpu, q_ref)
    np.testing.assert_array_equal(qd_gpu, qd_ref)


def test_soft_cube_worlds_bitexact(cuda_device):
    from newton_amd import scenes

    model, state, info = scenes.soft_cubes_batched(cuda_device, worlds=4, cells_min=2, cells_max=4)
    q0, qd0 = initial(state)
    q_gpu, qd_gpu, _ = gpu_run(model, state, info, 6)
    q_ref, qd_ref, _ = oracle_run(model, q0, qd0, info, 6)
    np.testing.assert_array_equal(q_gpu, q_ref)
    np.testing.assert_array_equal(qd_gpu, qd_ref)


def test_full_size_batched_cloth_invariants(cuda_device):
    """The 4096-world cloth for one frame: finite, pinned corners fixed, every world moved."""
    import torch

    from newton_amd import scenes

    model, state, info = scenes.cloth_batched(cuda_device)
    q0 = state.particle_q.clone()
    q, qd, _ = gpu_run(model, state, info, info["substeps"])
    assert np.isfinite(q).all() and np.isfinite(qd).all()
    pinned = (model.particle_inv_mass == 0).cpu().numpy()
    np.testing.assert_array_equal(q[pinned], q0.cpu().numpy()[pinned])
    moved = np.abs(q - q0.cpu().numpy()).max(axis=1) > 0
    ps = info["world_particle_start"]
    per_world = np.add.reduceat(moved.astype(np.int64), ps[:-1])
    assert (per_world > 0).all()
    torch.cuda.synchronize()
