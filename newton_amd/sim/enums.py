"""Model / state change flags (``newton/_src/sim/enums.py:8-110``): the same names and bit values, so
callers pass them to :meth:`SolverBase.notify_model_changed` and :meth:`SolverBase.reset` unchanged."""

from __future__ import annotations

from enum import IntEnum


class ModelFlags(IntEnum):
    """Which parts of a :class:`Model` were edited after the solver was created (``enums.py:8-52``)."""

    JOINT_PROPERTIES = 1 << 0
    JOINT_DOF_PROPERTIES = 1 << 1
    BODY_PROPERTIES = 1 << 2
    BODY_INERTIAL_PROPERTIES = 1 << 3
    SHAPE_PROPERTIES = 1 << 4
    MODEL_PROPERTIES = 1 << 5
    CONSTRAINT_PROPERTIES = 1 << 6
    TENDON_PROPERTIES = 1 << 7
    ACTUATOR_PROPERTIES = 1 << 8
    ALL = (
        JOINT_PROPERTIES | JOINT_DOF_PROPERTIES | BODY_PROPERTIES | BODY_INERTIAL_PROPERTIES | SHAPE_PROPERTIES
        | MODEL_PROPERTIES | CONSTRAINT_PROPERTIES | TENDON_PROPERTIES | ACTUATOR_PROPERTIES
    )


class StateFlags(IntEnum):
    """Which state attributes were updated or should be reset (``enums.py:55-114``)."""

    NONE = 0
    JOINT_Q = 1 << 0
    JOINT_QD = 1 << 1
    BODY_Q = 1 << 2
    BODY_QD = 1 << 3
    PARTICLE_Q = 1 << 4
    PARTICLE_QD = 1 << 5
    BODY_F = 1 << 6
    PARTICLE_F = 1 << 7
    JOINT_F = 1 << 8
    BODY = BODY_Q | BODY_QD
    PARTICLE = PARTICLE_Q | PARTICLE_QD
    JOINT = JOINT_Q | JOINT_QD
    FORCE = BODY_F | PARTICLE_F | JOINT_F
    ALL = BODY | PARTICLE | JOINT | FORCE


__all__ = ["ModelFlags", "StateFlags"]
