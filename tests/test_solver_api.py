"""The plugin boundary on the host: SolverBase's method set and defaults (solvers/solver.py:187-431) and
SolverVBD's keyword list, validation and deprecation behaviour (solver_vbd.py:221-423, :648-669,
:1828-1935), checked without a GPU (a particle-free model needs no device context)."""

import inspect
import warnings

import numpy as np
import pytest
import torch

from newton_amd import ModelBuilder, ModelFlags, StateFlags
from newton_amd.solvers import SolverBase, SolverVBD


def _empty_model():
    return ModelBuilder().finalize(device="cpu")


def _cloth_model():
    b = ModelBuilder(up_axis="Y")
    b.add_cloth_grid(pos=(0.0, 1.0, 0.0), rot=(0.0, 0.0, 0.0, 1.0), vel=(0.0, 0.0, 0.0), dim_x=2, dim_y=2,
                     cell_x=0.1, cell_y=0.1, mass=0.1)
    b.color()
    return b.finalize(device="cpu")


def test_solver_vbd_is_a_solver_base():
    assert issubclass(SolverVBD, SolverBase)
    s = SolverVBD(_empty_model())
    assert isinstance(s, SolverBase)
    assert s.device == torch.device("cpu")


def test_solver_base_defaults():
    m = _empty_model()
    base = SolverBase(m)
    with pytest.raises(NotImplementedError):
        base.step(None, None, None, None, 0.01)
    with pytest.raises(NotImplementedError):
        base.update_contacts(None)
    assert list(inspect.signature(SolverBase.update_contacts).parameters) == ["self", "contacts", "state"]
    assert base.reset(m.state()) is None
    assert base.notify_model_changed(ModelFlags.ALL) is None
    assert SolverBase.register_custom_attributes(ModelBuilder()) is None
    base.integrate_particles(m, m.state(), m.state(), 0.01)  # no particles: nothing to do
    base.integrate_bodies(m, m.state(), m.state(), 0.01)  # no bodies: nothing to do


def test_keyword_list_is_the_references():
    params = inspect.signature(SolverVBD.__init__).parameters
    names = list(params)
    ref = ["iterations", "friction_epsilon", "integrate_with_external_rigid_solver", "particle_enable_self_contact",
           "particle_self_contact_radius", "particle_self_contact_margin", "particle_conservative_bound_relaxation",
           "particle_vertex_contact_buffer_size", "particle_edge_contact_buffer_size",
           "particle_collision_detection_interval", "particle_edge_parallel_epsilon", "particle_enable_tile_solve",
           "particle_topological_contact_filter_threshold", "particle_rest_shape_contact_exclusion_radius",
           "particle_external_vertex_contact_filtering_map", "particle_external_edge_contact_filtering_map",
           "rigid_avbd_alpha", "rigid_avbd_joint_alpha", "rigid_avbd_contact_alpha", "rigid_avbd_beta",
           "rigid_avbd_linear_beta", "rigid_avbd_angular_beta", "rigid_avbd_gamma", "rigid_contact_hard",
           "rigid_contact_history", "rigid_contact_stick_motion_eps", "rigid_contact_stick_freeze_translation_eps",
           "rigid_contact_stick_freeze_angular_eps", "rigid_contact_k_start", "rigid_body_contact_buffer_size",
           "rigid_body_particle_contact_buffer_size", "rigid_joint_linear_ke", "rigid_joint_angular_ke",
           "rigid_joint_linear_k_start", "rigid_joint_angular_k_start", "rigid_joint_linear_kd",
           "rigid_joint_angular_kd", "deterministic"]
    assert names[:2] == ["self", "model"]
    assert names[2:2 + len(ref)] == ref
    assert all(params[n].kind == inspect.Parameter.KEYWORD_ONLY for n in ref)


def test_unknown_keyword_raises_type_error():
    with pytest.raises(TypeError):
        SolverVBD(_empty_model(), particle_self_contact_radiuss=0.1)
    with pytest.raises(TypeError):
        SolverVBD(_empty_model(), rigid_unknown=1.0)


def test_positional_keywords_warn_and_bind_in_order():
    with pytest.warns(DeprecationWarning, match="'iterations', 'friction_epsilon'"):
        s = SolverVBD(_empty_model(), 7, 0.5)
    assert s.iterations == 7 and s.friction_epsilon == 0.5
    with pytest.raises(TypeError, match="multiple values"):
        SolverVBD(_empty_model(), 7, iterations=3)


@pytest.mark.parametrize("kw,msg", [
    ({"rigid_avbd_beta": -1.0}, "rigid_avbd_beta must be >= 0"),
    ({"rigid_avbd_alpha": 1.5}, "rigid_avbd_alpha must be in"),
    ({"rigid_avbd_joint_alpha": -0.1}, "rigid_avbd_joint_alpha must be in"),
    ({"rigid_avbd_contact_alpha": 2.0}, "rigid_avbd_contact_alpha must be in"),
    ({"rigid_avbd_linear_beta": -1.0}, "rigid_avbd_linear_beta must be >= 0"),
    ({"rigid_avbd_angular_beta": -1.0}, "rigid_avbd_angular_beta must be >= 0"),
    ({"rigid_avbd_gamma": 1.1}, "rigid_avbd_gamma must be in"),
    ({"rigid_contact_k_start": -1.0}, "rigid_contact_k_start must be >= 0"),
    ({"rigid_joint_linear_k_start": -1.0}, "rigid_joint_linear_k_start must be >= 0"),
    ({"rigid_joint_angular_k_start": -1.0}, "rigid_joint_angular_k_start must be >= 0"),
    ({"rigid_joint_linear_ke": -1.0}, "rigid_joint_linear_ke must be >= 0"),
    ({"rigid_joint_angular_ke": -1.0}, "rigid_joint_angular_ke must be >= 0"),
])
def test_hyper_parameter_validation(kw, msg):
    with pytest.raises(ValueError, match=msg):
        SolverVBD(_empty_model(), **kw)


def test_deprecated_stick_keywords_warn():
    with pytest.warns(DeprecationWarning, match="deprecated and ignored"):
        SolverVBD(_empty_model(), rigid_contact_stick_motion_eps=1e-3)


def test_particle_model_errors():
    m = _cloth_model()
    with pytest.raises(ValueError, match="particle_self_contact_margin is smaller"):
        SolverVBD(m, particle_enable_self_contact=True, particle_self_contact_radius=0.3,
                  particle_self_contact_margin=0.1)
    m.particle_color_groups = []
    with pytest.raises(ValueError, match="particle_color_groups is empty"):
        SolverVBD(m)
    m = _cloth_model()
    m.soft_mesh_adjacency_device = None
    with pytest.raises(ValueError, match="soft_mesh_adjacency_device is missing"):
        SolverVBD(m)
    with pytest.raises(RuntimeError, match="GPU"):
        SolverVBD(_cloth_model())  # a cpu model has no HIP path


def test_reset_checks_like_the_reference():
    m = _empty_model()
    m.world_count = 3
    s = SolverVBD(m)
    with pytest.raises(ValueError, match="'state' argument is required"):
        s.reset(None)
    with pytest.raises(ValueError, match="one-dimensional boolean"):
        s.reset(m.state(), world_mask=torch.zeros(3, dtype=torch.int32))
    with pytest.raises(ValueError, match="length 5"):
        s.reset(m.state(), world_mask=torch.zeros(5, dtype=torch.bool))
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        s.reset(m.state(), world_mask=torch.ones(4, dtype=torch.bool), flags=StateFlags.PARTICLE_Q)  # no particles
    assert SolverVBD.register_custom_attributes(ModelBuilder(), dahl_defaults_enabled=True) is None
    with pytest.raises(NotImplementedError):
        s.update_contacts(None, state=None)


def test_flag_values_are_the_references():
    assert int(ModelFlags.SHAPE_PROPERTIES) == 16 and int(ModelFlags.ALL) == 511
    assert int(StateFlags.PARTICLE_Q) == 16 and int(StateFlags.PARTICLE) == 48 and int(StateFlags.ALL) == 511
    assert np.all([int(f) & (int(f) - 1) == 0 for f in (StateFlags.BODY_F, StateFlags.JOINT_F)])
