"""SolverBase: the plugin parent every solver of this package derives from.

Mirrors ``newton/_src/solvers/solver.py:187-431`` for the particle path: the
``device`` property, ``integrate_particles`` (a HIP kernel, ``vbd_integrate_particles``),
``integrate_bodies``, ``reset``, ``step``, ``notify_model_changed``,
``update_contacts(contacts, state=None)`` and ``register_custom_attributes``, with the
reference's defaults (no-op ``reset`` / ``notify_model_changed`` /
``register_custom_attributes``; ``step`` and ``update_contacts`` raise
``NotImplementedError``).  Rigid bodies are not part of this package, so
``integrate_bodies`` does nothing for a body-free model and raises otherwise.
"""

from __future__ import annotations

import torch

from .. import _native
from ..sim.collide import current_stream_handle
from ..sim.model import Contacts, Control, Model, State


class SolverBase:
    """Generic base class for solvers (``solver.py:187-431``)."""

    def __init__(self, model: Model):
        self.model = model

    @property
    def device(self):
        """The device the solver runs on (the model's)."""
        return self.model.device

    def integrate_bodies(self, model: Model, state_in: State, state_out: State, dt: float,
                         angular_damping: float = 0.0) -> None:
        """Rigid-body semi-implicit Euler (``solver.py:258-296``); this package has no rigid bodies."""
        if model.body_count:
            raise NotImplementedError("rigid bodies are not part of this package")

    def integrate_particles(self, model: Model, state_in: State, state_out: State, dt: float) -> None:
        """One semi-implicit Euler step of every particle (``solver.py:297-330``, kernel ``:20-56``):
        ``v' = v + (f / m + g) dt`` (gravity only for particles with mass), ``|v'|`` limited to
        ``model.particle_max_velocity``, ``x' = x + v' dt``; inactive particles keep ``x`` and their
        ``state_out.particle_qd`` is not written."""
        n = model.particle_count
        if not n:
            return
        gravity = model.gravity.reshape(-1, 3)
        world = getattr(model, "particle_world", None)
        _native.check(
            _native.hip_lib().vbd_integrate_particles(
                current_stream_handle(model.device), n, _native.ptr(state_in.particle_q),
                _native.ptr(state_in.particle_qd), _native.ptr(state_in.particle_f),
                _native.ptr(model.particle_inv_mass), _native.ptr(model.particle_flags),
                _native.ptr(world.to(torch.int32).contiguous() if world is not None else None),
                _native.ptr(gravity), int(gravity.shape[0]), float(dt), float(model.particle_max_velocity),
                _native.ptr(state_out.particle_q), _native.ptr(state_out.particle_qd)),
            "SolverBase.integrate_particles",
        )

    def reset(self, state: State, world_mask=None, flags=None) -> None:
        """Reset solver-internal state (``solver.py:332-356``); the default keeps none."""
        return None

    def step(self, state_in: State, state_out: State, control: Control | None, contacts: Contacts | None,
             dt: float) -> None:
        """Advance the model by ``dt`` (``solver.py:358-373``); each solver implements it."""
        raise NotImplementedError()

    def notify_model_changed(self, flags) -> None:
        """The caller edited parts of the model named by ``flags`` (``solver.py:375-410``); the default
        caches nothing."""
        return None

    def update_contacts(self, contacts: Contacts, state: State | None = None) -> None:
        """Write solver contact forces into ``contacts`` (``solver.py:412-421``); not provided by default."""
        raise NotImplementedError()

    @classmethod
    def register_custom_attributes(cls, builder) -> None:
        """Declare solver-specific model attributes on ``builder`` (``solver.py:423-431``); none by default."""
        return None


__all__ = ["SolverBase"]
