"""newton_amd: MI355X-native particle Vertex Block Descent (drop-in for Newton's SolverVBD path).

Public names mirror the reference's ``newton`` package for this path:
``ModelBuilder``, ``Model``, ``State``, ``Control``, ``Contacts``,
``CollisionPipeline``, ``ParticleFlags``, ``ShapeFlags``, ``GeoType``,
``ModelFlags``, ``StateFlags`` and ``newton_amd.solvers.SolverBase`` / ``SolverVBD``.
"""

from . import solvers
from .geometry import GeoType, ParticleFlags, ShapeFlags
from .sim import CollisionPipeline, Contacts, Control, Model, ModelBuilder, ModelFlags, State, StateFlags

__version__ = "0.1.0"

__all__ = [
    "CollisionPipeline",
    "Contacts",
    "Control",
    "GeoType",
    "Model",
    "ModelBuilder",
    "ModelFlags",
    "ParticleFlags",
    "ShapeFlags",
    "State",
    "StateFlags",
    "solvers",
]
