"""Mesh utilities."""

from .mesh import MeshAdjacency, MeshAdjacencyData

__all__ = ["MeshAdjacency", "MeshAdjacencyData"]
