"""Helpers shared by the parity tests: drive the GPU path and the CPU oracle on the same scene."""

from __future__ import annotations

import numpy as np

import oracle


def gpu_run(model, state, info, substeps: int, record_every: int = 0, margin: float = 0.01, solver_kwargs=None,
            contact_max=None, block_shape=None, solver=None, waves=None, aligned=None):
    """clear_forces -> collide -> SolverVBD.step -> swap, on the GPU; returns (q, qd, recorded q)."""
    import torch

    from newton_amd import CollisionPipeline
    from newton_amd.solvers import SolverVBD

    if solver is None:
        solver = SolverVBD(model, iterations=info["iterations"], **(solver_kwargs or {}))
    if block_shape is not None:
        solver.set_block_shape(block_shape)
    if waves is not None:
        solver.set_pass_waves(waves)
    if aligned is not None:
        solver.set_aligned_items(aligned)
    pipe = CollisionPipeline(model, soft_contact_margin=margin, soft_contact_max=contact_max)
    contacts = pipe.contacts()
    s0, s1 = state, model.state()
    rec = []
    for i in range(substeps):
        s0.clear_forces()
        pipe.collide(s0, contacts)
        solver.step(s0, s1, None, contacts, info["dt"])
        s0, s1 = s1, s0
        if record_every and (i + 1) % record_every == 0:
            rec.append(s0.particle_q.cpu().numpy())
    torch.cuda.synchronize()
    return s0.particle_q.cpu().numpy(), s0.particle_qd.cpu().numpy(), np.asarray(rec)


def initial(state):
    """Host copies of the initial positions / velocities (the GPU step mutates state_in)."""
    return state.particle_q.cpu().numpy().copy(), state.particle_qd.cpu().numpy().copy()


def oracle_run(model, q0, qd0, info, substeps: int, record_every: int = 0, margin: float = 0.01, precision="f32",
               threads: int = 16, contact_max=None, **param_kwargs):
    """The same frame loop in the CPU oracle, from host arrays q0 / qd0."""
    om = oracle.OracleModel(model)
    prm = oracle.params(info["iterations"], num_threads=threads, **param_kwargs)
    cmax = len(om.pairs) // 2 if contact_max is None else contact_max
    return oracle.simulate(om, prm, q0, qd0, substeps, info["dt"], margin, cmax, record_every=record_every,
                           precision=precision)
