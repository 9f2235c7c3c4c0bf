"""Full-size GPU checks: the 256x256 cloth-on-sphere workload against the FP32 oracle, bit for bit,
and the reference's hanging-cloth self-contact example run for 100 frames against the example's own
end-state checks (newton/examples/cloth/example_cloth_hanging.py:180-196)."""

import numpy as np
import pytest

import oracle
from _sim import gpu_run, initial, oracle_run

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("height", [1.5, 0.41])
def test_cloth_256_one_frame_bitexact(cuda_device, height):
    """65,536 vertices, one frame (10 substeps x 10 iterations); at 0.41 m the cloth rests on the sphere."""
    from newton_amd import scenes

    model, state, info = scenes.cloth_drop_on_sphere(cuda_device, height=height)
    q0, qd0 = initial(state)
    q_gpu, qd_gpu, _ = gpu_run(model, state, info, info["substeps"])
    q_ref, qd_ref, _ = oracle_run(model, q0, qd0, info, info["substeps"], threads=16)
    np.testing.assert_array_equal(q_gpu, q_ref)
    np.testing.assert_array_equal(qd_gpu, qd_ref)
    if height < 1.0:
        from newton_amd import CollisionPipeline

        pipe = CollisionPipeline(model, soft_contact_margin=0.01)
        contacts = pipe.contacts()
        pipe.collide(model.state(), contacts)
        assert int(contacts.soft_contact_count.item()) > 100


def _hanging_solver(model, info):
    from newton_amd.solvers import SolverVBD

    return SolverVBD(model, iterations=info["iterations"], **info["solver_kwargs"])


def test_hanging_self_contact_example_matches_oracle_then_stays_valid(cuda_device):
    import torch

    from newton_amd import CollisionPipeline, scenes
    from newton_amd.solvers.tri_mesh_collision import build_filters

    model, state, info = scenes.cloth_hanging_self_contact(cuda_device)
    q0, qd0 = initial(state)
    solver = _hanging_solver(model, info)
    frames_exact = 3
    q_gpu, qd_gpu, _ = gpu_run(model, state, info, frames_exact * info["substeps"], solver=solver)
    vt, ee = build_filters(model, 2)
    kw = info["solver_kwargs"]
    sc = oracle.SelfContact(model.particle_count, model.tri_count, model.edge_count,
                            radius=kw["particle_self_contact_radius"], margin=kw["particle_self_contact_margin"],
                            q_rest=q0, vt_filter=vt[0], vt_filter_offsets=vt[1], ee_filter=ee[0],
                            ee_filter_offsets=ee[1])
    q_ref, qd_ref, _ = oracle_run(model, q0, qd0, info, frames_exact * info["substeps"], threads=16,
                                  self_contact=sc)
    np.testing.assert_array_equal(q_gpu, q_ref)
    np.testing.assert_array_equal(qd_gpu, qd_ref)
    # 100 frames in total, then the example's end-state checks
    state.particle_q.copy_(torch.from_numpy(q_gpu))
    state.particle_qd.copy_(torch.from_numpy(qd_gpu))
    q, qd, _ = gpu_run(model, state, info, 97 * info["substeps"], solver=solver)
    assert np.isfinite(q).all() and np.isfinite(qd).all()
    assert (q[:, 2] > 0.0).all(), "particles are above the ground"
    lo = np.array([-64 * 0.11, -4.0, -1.8])
    hi = np.array([0.1, 7.0, 4.0])
    assert ((q >= lo) & (q <= hi)).all(), "particles are within a reasonable volume"
