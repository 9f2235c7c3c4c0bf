"""Geometry enums used by the particle VBD path.

Values match the reference's ``newton/_src/geometry/flags.py:8-40`` and
``newton/_src/geometry/types.py:70-100`` so that flags written by either
framework mean the same thing.
"""

from enum import IntEnum


class ParticleFlags(IntEnum):
    """Per-particle flags (``flags.py:8-23``)."""

    ACTIVE = 1 << 0
    PROXY = 1 << 1


class ShapeFlags(IntEnum):
    """Per-shape flags (``flags.py:27-40``)."""

    VISIBLE = 1 << 0
    COLLIDE_SHAPES = 1 << 1
    COLLIDE_PARTICLES = 1 << 2


class GeoType(IntEnum):
    """Shape geometry types (``types.py:70-100``); only PLANE and SPHERE collide with particles here."""

    NONE = 0
    PLANE = 1
    HFIELD = 2
    SPHERE = 3
    CAPSULE = 4
    ELLIPSOID = 5
    CYLINDER = 6
    BOX = 7


__all__ = ["GeoType", "ParticleFlags", "ShapeFlags"]
