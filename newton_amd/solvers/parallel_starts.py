"""One world's substep solved from several starts at once.

A 16 k-vertex cloth leaves most of an MI355X idle: a color pass is one work-group's critical path, so
a model holding K copies of the world costs about as much per iteration as the world alone.
:class:`ParallelStarts` uses that capacity for the warm start: copy k of the world starts from start
k -- copy 0 from the inertial guess (plain VBD), copy 1 from ``y + c`` (the previous correction
carried over), copy 2 from the learned guess -- all copies iterate together, each judged by its own
force residual, and the first copy to reach the tolerance gives the substep's result.  On a tie the
lower copy wins, so a warm start is taken only when its residual reaches the tolerance while the
inertial start's is still above it; otherwise the result is the plain VBD iterate (copy 0 starts
from the forward step's guess untouched; the copies are independent worlds).  No counterpart in the
reference.
"""

from __future__ import annotations

import numpy as np
import torch

from ..sim.model import Model, State


class ParallelStarts:
    """``model_k`` holds ``copies`` worlds identical to the single world (particles of copy k at
    ``k * n .. (k + 1) * n``), ``solver_k`` is a :class:`SolverVBD` of it and ``warm_start_k`` a
    :class:`~newton_amd.warmstart.LearnedWarmStart` of it (candidates ``[y + c, learned]``); copy k >= 1
    starts from candidate ``candidates[k - 1]``."""

    def __init__(self, model_k: Model, solver_k, warm_start_k, copies: int = 3, candidates=(0, 1)):
        if model_k.particle_count % copies:
            raise ValueError("model_k must hold `copies` equal worlds")
        if len(candidates) != copies - 1:
            raise ValueError("one warm-start candidate index per copy after the inertial one")
        self.model_k, self.solver_k, self.ws_k, self.copies = model_k, solver_k, warm_start_k, copies
        self.candidates = tuple(int(k) for k in candidates)
        self.n = model_k.particle_count // copies
        self.state_k = model_k.state()
        self.out_k = model_k.state()

    def step(self, state_in: State, state_out: State, corr: torch.Tensor, dt: float, rel_tol: float,
             max_iterations: int = 400) -> dict:
        """One substep of the single world: ``state_in`` (q, qd, f of the world), ``corr`` its previous
        correction ``x* - y`` ([n, 3]).  Writes the winning copy's iterate to ``state_in.particle_q`` and
        its result to ``state_out``; returns the winner, its iterations and its new correction."""
        n, K, sk = self.n, self.copies, self.state_k
        for t_k, t in ((sk.particle_q, state_in.particle_q), (sk.particle_qd, state_in.particle_qd),
                       (sk.particle_f, state_in.particle_f), (self.ws_k.corr, corr)):
            t_k.view(K, n, 3).copy_(t.view(1, n, 3).expand(K, n, 3))
        s = self.solver_k
        s.world_offsets()
        s.begin_step(sk, None, dt)
        r0 = s.residual_worlds(sk, dt).cpu().numpy()
        starts = self.ws_k(s, sk, dt)
        for k in range(1, K):
            s.set_iterate_range(sk, starts[self.candidates[k - 1]], k * n, (k + 1) * n)
        target = (rel_tol * rel_tol) * r0
        its, r_sq, wits = s.iterate_to_residual_worlds(sk, dt, target, max_iterations, per_world=True,
                                                       stop_at_live=K - 1)
        reached = r_sq <= target
        win = int(np.argmin(np.where(reached, wits, np.iinfo(np.int32).max))) if reached.any() else 0
        y = s.scratch("inertia")
        s.end_step(sk, self.out_k, dt)
        sl = slice(win * n, (win + 1) * n)
        state_in.particle_q.copy_(sk.particle_q[sl])
        state_out.particle_q.copy_(self.out_k.particle_q[sl])
        state_out.particle_qd.copy_(self.out_k.particle_qd[sl])
        return {"winner": win, "iterations": int(wits[win]), "iterations_run": its, "r0": float(np.sqrt(r0[0])),
                "r": float(np.sqrt(r_sq[win])), "reached": bool(reached[win]),
                "correction": sk.particle_q[sl] - y[sl]}


__all__ = ["ParallelStarts"]
