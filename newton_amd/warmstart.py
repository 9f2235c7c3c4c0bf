"""Learned warm start for the VBD iteration (no counterpart in the reference).

Newton's SolverVBD starts every substep from the inertial guess ``y``
(``forward_step``, ``particle_vbd_kernels.py:1350-1375``).  Here a small
per-particle MLP predicts the correction ``x* - y`` of the converged iterate
from 1-ring features, as a change of the previous substep's correction
``c_prev`` (guess = ``y + c_prev + MLP``, so an untrained network starts from
carrying ``c_prev`` over); :meth:`SolverVBD.step_to_residual` evaluates the force
residual of the predicted iterate and keeps it only if that residual is lower
than the inertial guess's, so plain VBD iterations always follow and the
result never gets worse than starting from ``y``.

Everything numeric runs in ``libvbd_ws.so`` (features, BF16 MFMA forward and
backward, Adam); this module owns the buffers (torch tensors used as device
memory) and the training loop.  There is no CPU fallback.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _native
from .sim.collide import current_stream_handle
from .sim.model import Model

N_IN, N_HIDDEN, N_OUT = 48, 128, 3


def vertex_neighbours(model: Model) -> tuple[np.ndarray, np.ndarray]:
    """1-ring CSR (offsets [N+1], neighbours) from triangle and tet edges, neighbours ascending."""
    n = model.particle_count
    pairs = []
    if model.tri_count:
        t = model.tri_indices.cpu().numpy().astype(np.int64)
        pairs += [t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]]
    if model.tet_count:
        t = model.tet_indices.cpu().numpy().astype(np.int64)
        pairs += [t[:, [a, b]] for a, b in ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))]
    if not pairs:
        return np.zeros(n + 1, dtype=np.int32), np.zeros(0, dtype=np.int32)
    e = np.concatenate(pairs)
    e = np.concatenate([e, e[:, ::-1]])
    e = e[e[:, 0] != e[:, 1]]
    key = np.unique(e[:, 0] * n + e[:, 1])
    src, dst = key // n, key % n
    off = np.zeros(n + 1, dtype=np.int64)
    np.add.at(off, src + 1, 1)
    return np.cumsum(off).astype(np.int32), dst.astype(np.int32)


def vertex_areas(model: Model) -> np.ndarray:
    """Rest area per particle: a third of each incident triangle's area (tet meshes: a quarter of |Dm|/6)."""
    n = model.particle_count
    area = np.zeros(n, dtype=np.float64)
    if model.tri_count:
        t = model.tri_indices.cpu().numpy()
        a = model.tri_areas.cpu().numpy().astype(np.float64)
        for k in range(3):
            np.add.at(area, t[:, k], a / 3.0)
    if model.tet_count:
        t = model.tet_indices.cpu().numpy()
        inv = model.tet_poses.cpu().numpy().astype(np.float64)
        vol = 1.0 / (6.0 * np.abs(np.linalg.det(inv)))
        for k in range(4):
            np.add.at(area, t[:, k], vol ** (2.0 / 3.0) / 4.0)
    return area.astype(np.float32)


class WarmStartMLP:
    """Parameters, BF16 packing, inference and training steps of the 48-128-128-128-3 SiLU MLP."""

    def __init__(self, device, seed: int = 0):
        self.device = torch.device(device)
        sizes = (ctypes.c_int64 * 3)()
        lib = _native.ws_lib()
        _native.ws_check(lib.vbd_ws_sizes(sizes), "vbd_ws_sizes")
        self.n_params, self.n_packed, self.rows_per_wg = int(sizes[0]), int(sizes[1]), int(sizes[2])
        self.params = torch.as_tensor(self.initial_params(seed), device=self.device)
        self.packed = torch.empty(self.n_packed, dtype=torch.int16, device=self.device)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.grad = torch.zeros_like(self.params)
        self.loss = torch.zeros(1, dtype=torch.float32, device=self.device)
        self.step_count = 0
        self._train_cap = 0
        self.pack()

    @staticmethod
    def initial_params(seed: int) -> np.ndarray:
        """Xavier-uniform weights, zero biases, the last layer scaled by 0.1 (starts near the inertial guess)."""
        rng = np.random.default_rng(seed)
        out = []
        dims = [(N_IN, N_HIDDEN), (N_HIDDEN, N_HIDDEN), (N_HIDDEN, N_HIDDEN), (N_HIDDEN, N_OUT)]
        for li, (fi, fo) in enumerate(dims):
            lim = math.sqrt(6.0 / (fi + fo)) * (0.1 if li == 3 else 1.0)
            out.append(rng.uniform(-lim, lim, size=(fi, fo)).reshape(-1))
            out.append(np.zeros(fo))
        return np.concatenate(out).astype(np.float32)

    def _s(self):
        return current_stream_handle(self.device)

    def pack(self) -> None:
        _native.ws_check(_native.ws_lib().vbd_ws_pack(self._s(), self.params.data_ptr(), self.packed.data_ptr()),
                         "vbd_ws_pack")

    def load_params(self, params) -> None:
        self.params.copy_(torch.as_tensor(np.asarray(params, dtype=np.float32), device=self.device))
        self.pack()

    def forward(self, feat: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        n = feat.shape[0]
        out = torch.empty((n, N_OUT), dtype=torch.float32, device=self.device) if out is None else out
        _native.ws_check(
            _native.ws_lib().vbd_ws_forward(self._s(), n, feat.data_ptr(), self.packed.data_ptr(),
                                            self.params.data_ptr(), out.data_ptr()),
            "vbd_ws_forward",
        )
        return out

    def _ensure_train_buffers(self, n: int) -> None:
        if n <= self._train_cap:
            return
        cap = max(n, 2 * self._train_cap)
        parts = -(-cap // self.rows_per_wg)
        self._out = torch.empty((cap, N_OUT), dtype=torch.float32, device=self.device)
        self._h = torch.empty((cap, 3, N_HIDDEN), dtype=torch.int16, device=self.device)
        self._z = torch.empty((cap, 3, N_HIDDEN), dtype=torch.float32, device=self.device)
        self._pgrad = torch.empty((parts, self.n_params), dtype=torch.float32, device=self.device)
        self._ploss = torch.empty(parts, dtype=torch.float32, device=self.device)
        self._train_cap = cap

    def gradient(self, feat: torch.Tensor, target: torch.Tensor, row_weight: torch.Tensor | None = None):
        """Loss (device [1]) and gradient (device [n_params]) of the weighted MSE at the current parameters."""
        n = feat.shape[0]
        if n <= 0:
            raise ValueError("empty training batch")
        self._ensure_train_buffers(n)
        lib = _native.ws_lib()
        s = self._s()
        _native.ws_check(lib.vbd_ws_forward_train(s, n, feat.data_ptr(), self.packed.data_ptr(),
                                                  self.params.data_ptr(), self._out.data_ptr(), self._h.data_ptr(),
                                                  self._z.data_ptr()), "vbd_ws_forward_train")
        _native.ws_check(lib.vbd_ws_backward(s, n, feat.data_ptr(), target.data_ptr(),
                                             None if row_weight is None else row_weight.data_ptr(),
                                             self.packed.data_ptr(), self._out.data_ptr(), self._h.data_ptr(),
                                             self._z.data_ptr(), self._pgrad.data_ptr(), self._ploss.data_ptr()),
                         "vbd_ws_backward")
        _native.ws_check(lib.vbd_ws_reduce(s, n, self._pgrad.data_ptr(), self._ploss.data_ptr(),
                                           self.grad.data_ptr(), self.loss.data_ptr()), "vbd_ws_reduce")
        return self.loss, self.grad

    def train_step(self, feat, target, row_weight=None, lr: float = 1e-3, beta1: float = 0.9, beta2: float = 0.999,
                   eps: float = 1e-8) -> torch.Tensor:
        """One Adam step on the weighted MSE; returns the loss before the step (device [1])."""
        loss, _ = self.gradient(feat, target, row_weight)
        self.step_count += 1
        _native.ws_check(_native.ws_lib().vbd_ws_adam(self._s(), self.params.data_ptr(), self.grad.data_ptr(),
                                                      self.m.data_ptr(), self.v.data_ptr(), float(lr), float(beta1),
                                                      float(beta2), float(eps), self.step_count), "vbd_ws_adam")
        self.pack()
        return loss


class LearnedWarmStart:
    """Feature gather + MLP prediction of ``x* - y``, as a ``warm_start`` callable for ``step_to_residual``.

    Call :meth:`observe` after every substep (with the solver still holding
    that substep's inertial guess) so the next substep's features see the
    correction the solver made; :meth:`reset` forgets it.
    """

    def __init__(self, model: Model, dt: float, mlp: WarmStartMLP | None = None, seed: int = 0,
                 offer_reuse: bool = True, blend: float = 0.5, smooth_sweeps: int = 0, delta_sweeps: int = 0):
        self.model = model
        self.offer_reuse = offer_reuse
        self.blend = float(blend)
        self.smooth_sweeps = int(smooth_sweeps)
        self.delta_sweeps = int(delta_sweeps)
        self.device = model.device
        self.dt = float(dt)
        n = model.particle_count
        off, nbr = vertex_neighbours(model)
        self.nbr_off = torch.as_tensor(off, device=self.device)
        self.nbr = torch.as_tensor(nbr if nbr.size else np.zeros(1, np.int32), device=self.device)
        area = vertex_areas(model)
        self.area = torch.as_tensor(area, device=self.device)
        mass = model.particle_mass.cpu().numpy()
        edge_len = float(np.mean(np.linalg.norm(
            model.particle_q.cpu().numpy()[nbr] - model.particle_q.cpu().numpy()[np.repeat(np.arange(n), np.diff(off))],
            axis=1))) if nbr.size else 1.0
        g = float(np.linalg.norm(model.gravity[0].cpu().numpy())) or 9.81
        self.scale_disp = self.dt * 1.0              # displacement of 1 m/s over one substep
        self.scale_corr = g * self.dt * self.dt      # gravity's displacement over one substep
        self.scales = torch.tensor([self.scale_disp, self.scale_corr, edge_len,
                                    float(mass[mass > 0].mean()) if (mass > 0).any() else 1.0,
                                    float(area[area > 0].mean()) if (area > 0).any() else 1.0],
                                   dtype=torch.float32, device=self.device)
        self.mlp = mlp if mlp is not None else WarmStartMLP(self.device, seed)
        self.active = (model.particle_flags & 1).bool() & (model.particle_inv_mass != 0)
        self.corr = torch.zeros((n, 3), dtype=torch.float32, device=self.device)
        self.feat = torch.empty((n, N_IN), dtype=torch.int16, device=self.device)
        self.pred = torch.empty((n, 3), dtype=torch.float32, device=self.device)
        self.acc = torch.zeros((n, 3), dtype=torch.float32, device=self.device)
        self._tmp = torch.empty((n, 3), dtype=torch.float32, device=self.device)
        # 1-ring sums for the smoothness term of the training loss, from a padded neighbour table:
        # each row is summed in neighbour order, so training is bit-reproducible (a scatter-add
        # would add in the order its atomics land)
        deg = (self.nbr_off[1:] - self.nbr_off[:-1]).long()
        cols = torch.arange(int(deg.max().item()) if n else 0, device=self.device)
        valid = cols[None, :] < deg[:, None]
        idx = (self.nbr_off[:-1, None].long() + cols[None, :]).clamp(max=max(self.nbr.numel() - 1, 0))
        self._ring_idx = torch.where(valid, self.nbr.long()[idx], torch.zeros_like(idx))
        self._ring_w = valid.float()[:, :, None]
        self._deg = deg.clamp(min=1).float()[:, None]
        self._act = self.active.float()[:, None]

    def _ring_sum(self, e: torch.Tensor) -> torch.Tensor:
        return (e[self._ring_idx] * self._ring_w).sum(dim=1)

    def _ring_mean(self, e: torch.Tensor) -> torch.Tensor:
        return self._ring_sum(e) / self._deg

    def train_substep(self, feat: torch.Tensor, target: torch.Tensor, lr: float, smooth: float = 100.0):
        """One Adam step on one substep's rows (all particles) with a smoothness-weighted loss
        ``|e|^2 + smooth * |L e|^2`` over the solved particles, e = MLP(feat) - target and L = I - (1-ring
        mean): the residual that judges a start weighs rough (neighbour-to-neighbour) error by the
        membrane stiffness, so the network is trained to keep its error smooth.  The MLP kernels train
        on a weighted MSE, so the loss gradient g enters as the target ``prediction - g / 2``.
        Returns the MSE part (device [1])."""
        pred = self.mlp.forward(feat)
        e = (pred - target) * self._act
        g = e
        if smooth > 0.0:
            le = (e - self._ring_mean(e)) * self._act
            w = le / self._deg
            lt = le - self._ring_sum(w)  # (I - A)^T (L e)
            g = e + smooth * lt * self._act
        self.mlp.train_step(feat, (pred - 0.5 * g).contiguous(), self._act[:, 0].contiguous(), lr=lr)
        return (e * e).sum() / (3.0 * self._act.sum().clamp(min=1.0))

    def reset(self) -> None:
        self.corr.zero_()

    def features(self, solver, state_in) -> torch.Tensor:
        """Features of the current substep (after ``begin_step``): BF16 bit patterns [N, 48]."""
        x0 = solver.scratch("q_prev")
        y = solver.scratch("inertia")
        torch.mul(state_in.particle_f, self.model.particle_inv_mass[:, None], out=self.acc)
        _native.ws_check(
            _native.ws_lib().vbd_ws_features(current_stream_handle(self.device), self.model.particle_count,
                                             self.nbr_off.data_ptr(), self.nbr.data_ptr(), x0.data_ptr(),
                                             y.data_ptr(), self.corr.data_ptr(), self.acc.data_ptr(),
                                             self.model.particle_mass.data_ptr(), self.area.data_ptr(),
                                             float(self.dt), self.scales.data_ptr(), self.feat.data_ptr()),
            "vbd_ws_features",
        )
        self._y = y
        return self.feat

    def smooth(self, v: torch.Tensor, sweeps: int | None = None) -> torch.Tensor:
        """Low-pass of correction fields ([N, 3k] floats, in place): ``sweeps`` sweeps of the 1-ring
        mean (``vbd_ws_smooth``) over the solved particles."""
        k = self.smooth_sweeps if sweeps is None else int(sweeps)
        if k > 0:
            width = v.shape[1]
            if self._tmp.numel() < v.numel():
                self._tmp = torch.empty_like(v)
            _native.ws_check(_native.ws_lib().vbd_ws_smooth(current_stream_handle(self.device),
                                                            self.model.particle_count, width,
                                                            self.nbr_off.data_ptr(), self.nbr.data_ptr(),
                                                            self._act.data_ptr(), k, v.data_ptr(),
                                                            self._tmp.data_ptr()),
                             "vbd_ws_smooth")
        return v

    def __call__(self, solver, state_in, dt):
        """Candidate iterates for ``step_to_residual``: the learned one, y + S(c_prev + blend * scale *
        MLP(features)) (the network predicts how this substep's correction differs from the previous
        substep's), and, when ``offer_reuse`` is set, first the network-free y + S(c_prev); the residual
        picks among them.  S is ``smooth_sweeps`` sweeps of the 1-ring mean (identity at 0): the
        per-particle network leaves rough, neighbour-to-neighbour error that the stiff membrane turns
        into large forces, and the low-pass removes it while keeping the smooth correction the
        Gauss-Seidel sweeps are slow to find.  ``delta_sweeps`` low-passes the network's change alone
        (before it is added to c_prev), which keeps the carried correction's detail."""
        feat = self.features(solver, state_in)
        self.mlp.forward(feat, self.pred)
        delta = self.pred * (self.blend * self.scale_corr)
        if self.delta_sweeps > 0:
            self.smooth(delta, self.delta_sweeps)
        learned = self.corr + delta
        if not self.offer_reuse:
            return self._y + self.smooth(learned)
        both = self.smooth(torch.cat([self.corr, learned], dim=1))  # one set of sweeps for both fields
        return [self._y + both[:, :3], self._y + both[:, 3:]]

    def reuse(self, solver, state_in, dt) -> torch.Tensor:
        """The no-network baseline: y + S(c_prev) (the previous substep's correction carried over,
        low-passed as the learned candidate is)."""
        return solver.scratch("inertia") + self.smooth(self.corr.clone())

    def target(self, solver, state_in) -> torch.Tensor:
        """Training target (x - y - c_prev) / scale of the current iterate (call after iterating,
        before :meth:`observe`)."""
        return (state_in.particle_q - solver.scratch("inertia") - self.corr) / self.scale_corr

    def observe(self, solver, state_in) -> None:
        """Remember this substep's correction x* - y for the next substep's features."""
        torch.sub(state_in.particle_q, solver.scratch("inertia"), out=self.corr)


__all__ = ["LearnedWarmStart", "WarmStartMLP", "vertex_areas", "vertex_neighbours"]
