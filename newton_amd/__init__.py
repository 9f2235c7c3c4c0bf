"""newton_amd: MI355X-native particle Vertex Block Descent (drop-in for Newton's SolverVBD path).

Public names mirror the reference's ``newton`` package for this path:
``ModelBuilder``, ``Model``, ``State``, ``Control``, ``Contacts``,
``CollisionPipeline``, ``ParticleFlags``, ``ShapeFlags``, ``GeoType`` and
``newton_amd.solvers.SolverVBD``.
"""

from . import solvers
from .geometry import GeoType, ParticleFlags, ShapeFlags
from .sim import CollisionPipeline, Contacts, Control, Model, ModelBuilder, State

__version__ = "0.1.0"

__all__ = [
    "CollisionPipeline",
    "Contacts",
    "Control",
    "GeoType",
    "Model",
    "ModelBuilder",
    "ParticleFlags",
    "ShapeFlags",
    "State",
    "solvers",
]
