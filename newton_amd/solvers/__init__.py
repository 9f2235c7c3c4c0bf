"""Solvers (``newton.solvers`` counterpart): the particle VBD solver."""

from .solver_vbd import SolverVBD

__all__ = ["SolverVBD"]
