"""Solvers (``newton.solvers`` counterpart): the plugin parent and the particle VBD solver."""

from .solver import SolverBase
from .solver_vbd import SolverVBD

__all__ = ["SolverBase", "SolverVBD"]
