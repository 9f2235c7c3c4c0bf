"""Simulation data model and input construction for the particle VBD path."""

from .builder import ModelBuilder
from .collide import CollisionPipeline
from .enums import ModelFlags, StateFlags
from .model import Contacts, Control, Model, State

__all__ = ["CollisionPipeline", "Contacts", "Control", "Model", "ModelBuilder", "ModelFlags", "State", "StateFlags"]
