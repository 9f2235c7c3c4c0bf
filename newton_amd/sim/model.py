"""Model / State / Control / Contacts for the particle VBD path.

Field names, shapes and dtypes follow the reference (``newton/_src/sim/model.py:745-1311``,
``state.py:57-200``, ``control.py``, ``contacts.py:327-371``); arrays are torch
tensors on the model's device (vec3 -> [n, 3] float32, mat22 -> [n, 2, 2],
mat33 -> [n, 3, 3], indices int32).
"""

from __future__ import annotations

import numpy as np
import torch


class State:
    """Time-varying particle state (``state.py:57-125``)."""

    def __init__(self):
        self.particle_q: torch.Tensor | None = None
        self.particle_qd: torch.Tensor | None = None
        self.particle_f: torch.Tensor | None = None
        self.body_q = None
        self.body_qd = None
        self.body_f = None

    @property
    def particle_count(self) -> int:
        return 0 if self.particle_q is None else int(self.particle_q.shape[0])

    def clear_forces(self) -> None:
        """Zero the per-particle external forces (``state.py:189-200``)."""
        if self.particle_f is not None:
            self.particle_f.zero_()

    def assign(self, other: "State") -> None:
        for name in ("particle_q", "particle_qd", "particle_f", "body_q", "body_qd", "body_f"):
            a, b = getattr(self, name), getattr(other, name)
            if a is not None and b is not None:
                a.copy_(b)


class Control:
    """Control inputs; the particle VBD path reads none (``control.py``)."""

    def __init__(self):
        self.tri_activations = None
        self.tet_activations = None


class Contacts:
    """Soft-contact buffers (``contacts.py:327-371``); rigid contacts are not on this path."""

    def __init__(self, soft_contact_max: int, device):
        n = max(int(soft_contact_max), 0)
        self.soft_contact_max = n
        self.soft_contact_count = torch.zeros(1, dtype=torch.int32, device=device)
        self.soft_contact_particle = torch.full((n,), -1, dtype=torch.int32, device=device)
        self.soft_contact_indices = torch.full((n, 3), -1, dtype=torch.int32, device=device)
        self.soft_contact_barycentric = torch.zeros((n, 3), dtype=torch.float32, device=device)
        self.soft_contact_shape = torch.zeros(n, dtype=torch.int32, device=device)
        self.soft_contact_body_pos = torch.zeros((n, 3), dtype=torch.float32, device=device)
        self.soft_contact_body_vel = torch.zeros((n, 3), dtype=torch.float32, device=device)
        self.soft_contact_normal = torch.zeros((n, 3), dtype=torch.float32, device=device)
        self.soft_contact_tids = torch.zeros(0, dtype=torch.int32, device=device)
        self.device = torch.device(device)
        # set by the CollisionPipeline that allocates the buffer: edge/face records may appear
        self._enable_rigid_soft_full_surface_contact = False

    def clear(self) -> None:
        self.soft_contact_count.zero_()


class Model:
    """Static particle/element/shape arrays (``model.py:745-1311``)."""

    def __init__(self, device):
        self.device = torch.device(device)
        self.requires_grad = False
        self.world_count = 0
        self.up_axis = 1
        self.particle_count = 0
        self.particle_q = None
        self.particle_qd = None
        self.particle_mass = None
        self.particle_inv_mass = None
        self.particle_radius = None
        self.particle_flags = None
        self.particle_world = None
        self.particle_max_radius = 0.0
        self.particle_max_velocity = 1e5
        self.particle_colors = None
        self.particle_color_groups: list[torch.Tensor] = []
        self.tri_count = 0
        self.tri_indices = None
        self.tri_poses = None
        self.tri_activations = None
        self.tri_materials = None
        self.tri_areas = None
        self.edge_count = 0
        self.edge_indices = None
        self.edge_rest_angle = None
        self.edge_rest_length = None
        self.edge_bending_properties = None
        self.tet_count = 0
        self.tet_indices = None
        self.tet_poses = None
        self.tet_activations = None
        self.tet_materials = None
        self.spring_count = 0
        self.spring_indices = None
        self.spring_rest_length = None
        self.spring_stiffness = None
        self.spring_damping = None
        self.spring_control = None
        self.soft_mesh_adjacency = None
        self.soft_mesh_adjacency_device = None
        self.gravity = None
        self.soft_contact_ke = 1.0e3
        self.soft_contact_kd = 10.0
        self.soft_contact_mu = 0.5
        self.soft_contact_restitution = 0.0
        self.shape_count = 0
        self.shape_type = None
        self.shape_body = None
        self.shape_flags = None
        self.shape_world = None
        self.shape_transform = None
        self.shape_scale = None
        self.shape_margin = None
        self.shape_material_ke = None
        self.shape_material_kd = None
        self.shape_material_mu = None
        self.body_count = 0
        self.body_q = None
        self.body_qd = None
        self.body_com = None
        self.body_mass = None
        self.body_inv_mass = None
        self.body_inertia = None
        self.body_flags = None
        self.body_world = None
        self.joint_count = 0

    def state(self, requires_grad: bool | None = None) -> State:
        """New state holding clones of the initial particle and body configuration (``model.py:1828-1870``)."""
        s = State()
        if self.particle_count:
            s.particle_q = self.particle_q.clone()
            s.particle_qd = self.particle_qd.clone()
            s.particle_f = torch.zeros_like(self.particle_qd)
        if self.body_count:
            s.body_q = self.body_q.clone()
            s.body_qd = self.body_qd.clone()
            s.body_f = torch.zeros_like(self.body_qd)
        return s

    def control(self, clone_variables: bool = True) -> Control:
        return Control()

    def set_gravity(self, gravity, world: int | None = None) -> None:
        """Set gravity for every world or one world (``model.py:1958``)."""
        g = torch.as_tensor(np.asarray(gravity, dtype=np.float32).reshape(3), device=self.device)
        if world is None:
            self.gravity[:] = g
        else:
            self.gravity[world] = g


__all__ = ["Contacts", "Control", "Model", "State"]
