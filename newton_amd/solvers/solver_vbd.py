"""SolverVBD: drop-in particle Vertex Block Descent solver running on libvbd_hip.

Mirrors the reference's ``newton/_src/solvers/vbd/solver_vbd.py`` for the
particle path: the constructor signature (keyword-only after ``model``,
``:221-269``), ``step(state_in, state_out, control, contacts, dt)``
(``:1773-1825``), the in-place mutation of ``state_in.particle_q`` with the
iterate, and the exception types raised at construction (``:403-404``,
``:558-559``, ``:610-614``).  The compute path is the HIP library; there is no
Python or CPU fallback.
"""

from __future__ import annotations

import ctypes
import math
import warnings

import numpy as np
import torch

from .. import _native
from ..sim.collide import current_stream_handle
from ..sim.enums import StateFlags
from ..sim.model import Contacts, Control, Model, State
from ..utils.deprecation import deprecate_nonkeyword_arguments
from .solver import SolverBase


class _TrialState:
    """Positions-only stand-in for a State (residual of a trial iterate)."""

    def __init__(self, q: torch.Tensor):
        self.particle_q = q


def color_layout(model: Model):
    """``(color_offsets [C+1], color_particles [N])`` int32 arrays: the concatenated color groups."""
    groups = [g.cpu().numpy().astype(np.int32) for g in model.particle_color_groups]
    color_offsets = np.zeros(len(groups) + 1, dtype=np.int32)
    color_offsets[1:] = np.cumsum([len(g) for g in groups])
    color_particles = np.ascontiguousarray(np.concatenate(groups) if groups else np.empty(0, np.int32), dtype=np.int32)
    return color_offsets, color_particles


class SolverVBD(SolverBase):
    """Vertex Block Descent for particles (cloth triangles, bending edges, tetrahedra) with soft contact.

    The keyword list is the reference's (``solver_vbd.py:221-269``); the ``rigid_*`` keywords are
    validated as the reference validates them (``:403-423``, ``:648-669``) and otherwise only matter
    for rigid bodies, which this package does not simulate.  Two keywords extend the list:
    ``fp64_ill_conditioned`` (default on: a vertex block whose Frobenius condition number, evaluated
    in FP64, exceeds 1e6 is solved in FP64; off gives the reference's all-FP32 solve, and the two
    agree bit for bit on every block below that condition number) and ``soft_contact_capacity``
    (largest ``Contacts.soft_contact_max`` the solver accepts without re-allocating; default
    particles x shapes).
    """

    @deprecate_nonkeyword_arguments
    def __init__(
        self,
        model: Model,
        *,
        iterations: int = 10,
        friction_epsilon: float = 1e-2,
        integrate_with_external_rigid_solver: bool = False,
        particle_enable_self_contact: bool = False,
        particle_self_contact_radius: float = 0.2,
        particle_self_contact_margin: float = 0.2,
        particle_conservative_bound_relaxation: float = 0.85,
        particle_vertex_contact_buffer_size: int = 32,
        particle_edge_contact_buffer_size: int = 64,
        particle_collision_detection_interval: int = 0,
        particle_edge_parallel_epsilon: float = 1e-5,
        particle_enable_tile_solve: bool = True,
        particle_topological_contact_filter_threshold: int = 2,
        particle_rest_shape_contact_exclusion_radius: float = 0.0,
        particle_external_vertex_contact_filtering_map: dict | None = None,
        particle_external_edge_contact_filtering_map: dict | None = None,
        rigid_avbd_alpha: float = 0.95,
        rigid_avbd_joint_alpha: float | None = None,
        rigid_avbd_contact_alpha: float | None = None,
        rigid_avbd_beta: float = 0.0,
        rigid_avbd_linear_beta: float | None = None,
        rigid_avbd_angular_beta: float | None = None,
        rigid_avbd_gamma: float = 0.999,
        rigid_contact_hard: bool = True,
        rigid_contact_history: bool = False,
        rigid_contact_stick_motion_eps: float | None = None,
        rigid_contact_stick_freeze_translation_eps: float | None = None,
        rigid_contact_stick_freeze_angular_eps: float | None = None,
        rigid_contact_k_start: float = 1.0e2,
        rigid_body_contact_buffer_size: int = 64,
        rigid_body_particle_contact_buffer_size: int = 256,
        rigid_joint_linear_ke: float = 1.0e5,
        rigid_joint_angular_ke: float = 1.0e5,
        rigid_joint_linear_k_start: float = 1.0e2,
        rigid_joint_angular_k_start: float = 1.0e1,
        rigid_joint_linear_kd: float = 0.0,
        rigid_joint_angular_kd: float = 0.0,
        deterministic=None,
        fp64_ill_conditioned: bool = True,
        soft_contact_capacity: int | None = None,
    ):
        if rigid_avbd_beta < 0:
            raise ValueError(f"rigid_avbd_beta must be >= 0, got {rigid_avbd_beta}")
        linear_beta = rigid_avbd_beta if rigid_avbd_linear_beta is None else rigid_avbd_linear_beta
        angular_beta = rigid_avbd_beta if rigid_avbd_angular_beta is None else rigid_avbd_angular_beta
        if (rigid_contact_stick_motion_eps is not None or rigid_contact_stick_freeze_translation_eps is not None
                or rigid_contact_stick_freeze_angular_eps is not None):
            warnings.warn(
                "SolverVBD rigid_contact_stick_motion_eps, rigid_contact_stick_freeze_translation_eps, and "
                "rigid_contact_stick_freeze_angular_eps are deprecated and ignored, and will be removed in a "
                "future release. Use CollisionPipeline(contact_matching='sticky', "
                "contact_matching_pos_threshold=...) for persistent contact geometry. The SolverVBD body-level "
                "contact deadzone was removed.",
                DeprecationWarning,
                stacklevel=3,
            )
        super().__init__(model)
        if model.body_count > 0 and not integrate_with_external_rigid_solver:
            raise NotImplementedError(
                "SolverVBD here solves particles only: rigid bodies are not integrated (AVBD is not part of this "
                "path). Pass integrate_with_external_rigid_solver=True and move the bodies with another solver "
                "(state_in.body_q / state_out.body_q are then read for particle contact)."
            )
        self.iterations = int(iterations)
        self.friction_epsilon = float(friction_epsilon)
        self.integrate_with_external_rigid_solver = integrate_with_external_rigid_solver
        self.particle_enable_self_contact = bool(particle_enable_self_contact)
        self.particle_self_contact_radius = float(particle_self_contact_radius)
        self.particle_self_contact_margin = float(particle_self_contact_margin)
        self.particle_conservative_bound_relaxation = float(particle_conservative_bound_relaxation)
        self.particle_vertex_contact_buffer_size = int(particle_vertex_contact_buffer_size)
        self.particle_edge_contact_buffer_size = int(particle_edge_contact_buffer_size)
        self.particle_collision_detection_interval = int(particle_collision_detection_interval)
        self.particle_edge_parallel_epsilon = float(particle_edge_parallel_epsilon)
        self.particle_topological_contact_filter_threshold = int(particle_topological_contact_filter_threshold)
        self.particle_rest_shape_contact_exclusion_radius = float(particle_rest_shape_contact_exclusion_radius)
        self._external_vt_map = particle_external_vertex_contact_filtering_map
        self._external_ee_map = particle_external_edge_contact_filtering_map
        self.particle_enable_tile_solve = particle_enable_tile_solve
        self.deterministic = deterministic  # every reduction here has a fixed order already
        self.fp64_ill_conditioned = bool(fp64_ill_conditioned)
        self._full_surface = False  # per-corner contact buffers for edge/face records (set on first use)
        self._ctx = None
        self._pass_shape = None  # (block, waves, aligned) chosen by autotune() / set_*(), kept across rebuilds
        if model.particle_count > 0:
            if model.soft_mesh_adjacency_device is None or model.soft_mesh_adjacency is None:
                raise ValueError("model.soft_mesh_adjacency_device is missing; finalize the model with ModelBuilder.")
            if self.particle_enable_self_contact and particle_self_contact_margin < particle_self_contact_radius:
                raise ValueError(
                    "particle_self_contact_margin is smaller than particle_self_contact_radius, this will result in "
                    "missing contacts and cause instability.\nIt is advisable to make particle_self_contact_margin "
                    "1.5-2 times larger than particle_self_contact_radius."
                )
            if len(model.particle_color_groups) == 0:
                raise ValueError(
                    "model.particle_color_groups is empty! When using the SolverVBD you must call "
                    "ModelBuilder.color() or ModelBuilder.set_coloring() before calling ModelBuilder.finalize()."
                )
        # rigid AVBD hyper-parameters (solver_vbd.py:648-669)
        if not (0.0 <= rigid_avbd_alpha <= 1.0):
            raise ValueError(f"rigid_avbd_alpha must be in [0, 1], got {rigid_avbd_alpha}")
        if rigid_avbd_joint_alpha is not None and not (0.0 <= rigid_avbd_joint_alpha <= 1.0):
            raise ValueError(f"rigid_avbd_joint_alpha must be in [0, 1], got {rigid_avbd_joint_alpha}")
        if rigid_avbd_contact_alpha is not None and not (0.0 <= rigid_avbd_contact_alpha <= 1.0):
            raise ValueError(f"rigid_avbd_contact_alpha must be in [0, 1], got {rigid_avbd_contact_alpha}")
        if linear_beta < 0:
            raise ValueError(f"rigid_avbd_linear_beta must be >= 0, got {linear_beta}")
        if angular_beta < 0:
            raise ValueError(f"rigid_avbd_angular_beta must be >= 0, got {angular_beta}")
        if not (0.0 <= rigid_avbd_gamma <= 1.0):
            raise ValueError(f"rigid_avbd_gamma must be in [0, 1], got {rigid_avbd_gamma}")
        if rigid_contact_k_start < 0:
            raise ValueError(f"rigid_contact_k_start must be >= 0, got {rigid_contact_k_start}")
        if rigid_joint_linear_k_start < 0:
            raise ValueError(f"rigid_joint_linear_k_start must be >= 0, got {rigid_joint_linear_k_start}")
        if rigid_joint_angular_k_start < 0:
            raise ValueError(f"rigid_joint_angular_k_start must be >= 0, got {rigid_joint_angular_k_start}")
        if rigid_joint_linear_ke < 0:
            raise ValueError(f"rigid_joint_linear_ke must be >= 0, got {rigid_joint_linear_ke}")
        if rigid_joint_angular_ke < 0:
            raise ValueError(f"rigid_joint_angular_ke must be >= 0, got {rigid_joint_angular_ke}")
        self.rigid_avbd_gamma = float(rigid_avbd_gamma)
        self.rigid_linear_beta = float(linear_beta)
        # fixed penalty (k = averaged ke) when the linear beta is 0 (solver_vbd.py:671)
        self.rigid_contact_k_start_value = -1.0 if linear_beta == 0.0 else float(rigid_contact_k_start)
        if model.particle_count == 0:
            return
        if model.device.type != "cuda":
            raise RuntimeError("SolverVBD runs on the GPU through libvbd_hip; finalize the model on a cuda device")
        co, cp = color_layout(model)
        if int(co[-1]) != model.particle_count or np.unique(cp).size != model.particle_count:
            raise ValueError("particle_color_groups must contain every particle exactly once")
        self.color_offsets = co
        self.color_particles = torch.as_tensor(cp, device=model.device)
        if soft_contact_capacity is None:
            soft_contact_capacity = model.particle_count * max(model.shape_count, 1)
        self._capacity = int(soft_contact_capacity)
        if self.particle_enable_self_contact:
            self._prepare_self_contact()
        self._create_context()

    def _prepare_self_contact(self):
        """Rest positions and topological contact filters (``solver_vbd.py:556-606``,
        ``tri_mesh_collision.py:289-313``), uploaded once."""
        from .tri_mesh_collision import build_filters

        m = self.model
        dev = m.device
        self._q_rest = m.particle_q.detach().clone().contiguous()
        vt, ee = build_filters(m, self.particle_topological_contact_filter_threshold, self._external_vt_map,
                               self._external_ee_map)

        def up(a):
            return None if a is None else torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32), device=dev)

        self._vt_filter = None if vt is None else (up(vt[0]), up(vt[1]))
        self._ee_filter = None if ee is None else (up(ee[0]), up(ee[1]))
        self._particle_world = None
        if getattr(m, "particle_world", None) is not None and int(getattr(m, "world_count", 1)) > 1:
            self._particle_world = m.particle_world.to(torch.int32).contiguous()

    # ------------------------------------------------------------- internals
    def _create_context(self):
        m = self.model
        adj = m.soft_mesh_adjacency_device
        d = _native.VbdModelDesc()
        d.particle_count = m.particle_count
        d.tri_count = m.tri_count
        d.edge_count = m.edge_count
        d.tet_count = m.tet_count
        p = _native.ptr
        d.particle_mass = p(m.particle_mass)
        d.particle_inv_mass = p(m.particle_inv_mass)
        d.particle_radius = p(m.particle_radius)
        d.particle_flags = p(m.particle_flags)
        d.tri_indices = p(m.tri_indices)
        d.tri_poses = p(m.tri_poses)
        d.tri_materials = p(m.tri_materials)
        d.tri_areas = p(m.tri_areas)
        d.edge_indices = p(m.edge_indices)
        d.edge_rest_angle = p(m.edge_rest_angle)
        d.edge_rest_length = p(m.edge_rest_length)
        d.edge_bending_properties = p(m.edge_bending_properties)
        d.tet_indices = p(m.tet_indices)
        d.tet_poses = p(m.tet_poses)
        d.tet_materials = p(m.tet_materials)
        d.v_adj_tris = p(adj.v_adj_tris)
        d.v_adj_tris_offsets = p(adj.v_adj_tris_offsets)
        d.v_adj_edges = p(adj.v_adj_edges)
        d.v_adj_edges_offsets = p(adj.v_adj_edges_offsets)
        d.v_adj_tets = p(adj.v_adj_tets)
        d.v_adj_tets_offsets = p(adj.v_adj_tets_offsets)
        d.color_count = len(self.color_offsets) - 1
        d.color_offsets_host = self.color_offsets.ctypes.data
        d.color_particles = p(self.color_particles)
        d.gravity = p(m.gravity)
        d.shape_count = m.shape_count
        d.shape_margin = p(m.shape_margin)
        d.shape_material_ke = p(m.shape_material_ke)
        d.shape_material_kd = p(m.shape_material_kd)
        d.shape_material_mu = p(m.shape_material_mu)
        d.soft_contact_ke = float(m.soft_contact_ke)
        d.soft_contact_kd = float(m.soft_contact_kd)
        d.soft_contact_mu = float(m.soft_contact_mu)
        d.spring_count = m.spring_count
        if m.spring_count:
            d.spring_indices = p(m.spring_indices)
            d.spring_rest_length = p(m.spring_rest_length)
            d.spring_stiffness = p(m.spring_stiffness)
            d.spring_damping = p(m.spring_damping)
        if m.shape_count and m.body_count and bool((m.shape_body >= 0).any()):
            d.shape_body = p(m.shape_body)
        d.body_count = m.body_count
        prm = _native.VbdSolverParams()
        prm.soft_contact_full_surface = int(self._full_surface)
        prm.iterations = self.iterations
        prm.friction_epsilon = self.friction_epsilon
        prm.avbd_beta = self.rigid_linear_beta
        prm.k_start = self.rigid_contact_k_start_value
        prm.fp64_ill_conditioned = int(self.fp64_ill_conditioned)
        prm.soft_contact_capacity = self._capacity
        if self.particle_enable_self_contact:
            prm.self_contact = 1
            prm.self_contact_radius = self.particle_self_contact_radius
            prm.self_contact_margin = self.particle_self_contact_margin
            prm.conservative_bound_relaxation = self.particle_conservative_bound_relaxation
            # the reference multiplies the Python floats in double, then launches with float32
            prm.max_displacement = float(np.float32(self.particle_self_contact_margin
                                                    * self.particle_conservative_bound_relaxation * 0.5))
            prm.vertex_contact_buffer_size = self.particle_vertex_contact_buffer_size
            prm.edge_contact_buffer_size = self.particle_edge_contact_buffer_size
            prm.collision_detection_interval = self.particle_collision_detection_interval
            prm.edge_parallel_epsilon = self.particle_edge_parallel_epsilon
            prm.rest_shape_contact_exclusion_radius = self.particle_rest_shape_contact_exclusion_radius
            prm.particle_q_rest = p(self._q_rest)
            prm.particle_world = p(self._particle_world)
            if self._vt_filter is not None and self._vt_filter[0].numel():
                prm.vt_filter, prm.vt_filter_offsets = p(self._vt_filter[0]), p(self._vt_filter[1])
            if self._ee_filter is not None and self._ee_filter[0].numel():
                prm.ee_filter, prm.ee_filter_offsets = p(self._ee_filter[0]), p(self._ee_filter[1])
        self._desc, self._params = d, prm
        lib = _native.hip_lib()
        h = ctypes.c_void_p()
        with torch.cuda.device(m.device):
            _native.check(lib.vbd_create(ctypes.byref(d), ctypes.byref(prm), ctypes.byref(h)), "vbd_create")
        self._ctx = h
        self._push_scalars()

    def _contact_material_key(self):
        m = self.model
        return (float(m.soft_contact_ke), float(m.soft_contact_kd), float(m.soft_contact_mu), self.iterations,
                self.friction_epsilon)

    def _destroy_context(self):
        if self._ctx is not None and self._ctx.value:
            _native.hip_lib().vbd_destroy(self._ctx)
        self._ctx = None

    def __del__(self):
        try:
            self._destroy_context()
        except Exception:
            pass

    def _push_scalars(self):
        """Hands the solver's and the model's scalar settings (iterations, friction epsilon, soft-contact
        ke / kd / mu) to the context; they are read per launch, as the reference reads them."""
        m = self.model
        _native.check(_native.hip_lib().vbd_set_scalars(self._ctx, self.iterations, self.friction_epsilon,
                                                         float(m.soft_contact_ke), float(m.soft_contact_kd),
                                                         float(m.soft_contact_mu)), "vbd_set_scalars")
        self._material_key = self._contact_material_key()

    def _apply_pass_shape(self):
        if self._pass_shape is not None:
            self.set_pass_shape(*self._pass_shape)

    def _ensure_context(self, contacts: Contacts | None):
        need = contacts.soft_contact_max if contacts is not None else 0
        full = bool(contacts is not None and getattr(contacts, "_enable_rigid_soft_full_surface_contact", False))
        if need > self._capacity or (full and not self._full_surface):
            if torch.cuda.is_current_stream_capturing():
                raise RuntimeError(
                    "SolverVBD: the soft-contact buffers would need to grow (larger soft_contact_max, or edge/face "
                    "records) during graph capture; run one step before capturing"
                )
            torch.cuda.synchronize(self.model.device)
            self._capacity = max(need, self._capacity)
            self._full_surface = self._full_surface or full
            self._destroy_context()
            self._create_context()
            self._apply_pass_shape()
        elif self._contact_material_key() != self._material_key:
            self._push_scalars()

    def _contacts_desc(self, contacts: Contacts | None, state_in: State | None = None,
                       state_out: State | None = None, dt: float = 0.0):
        """The contacts and, for shapes on bodies, the external rigid solver's poses: ``state_out.body_q`` at
        the step's end and ``state_in.body_q`` at its start (``solver_vbd.py:2523-2526``)."""
        if contacts is None or contacts.soft_contact_max <= 0:
            return None
        c = _native.VbdContactsDesc()
        c.soft_contact_count = _native.ptr(contacts.soft_contact_count)
        c.soft_contact_max = contacts.soft_contact_max
        c.soft_contact_indices = _native.ptr(contacts.soft_contact_indices)
        c.soft_contact_shape = _native.ptr(contacts.soft_contact_shape)
        c.soft_contact_body_pos = _native.ptr(contacts.soft_contact_body_pos)
        c.soft_contact_body_vel = _native.ptr(contacts.soft_contact_body_vel)
        c.soft_contact_normal = _native.ptr(contacts.soft_contact_normal)
        if getattr(contacts, "_enable_rigid_soft_full_surface_contact", False):
            c.soft_contact_barycentric = _native.ptr(contacts.soft_contact_barycentric)
        if self.model.body_count:
            cur = state_out if state_out is not None and state_out.body_q is not None else state_in
            if cur is None or cur.body_q is None or state_in is None or state_in.body_q is None:
                raise ValueError("the model has bodies: state_in.body_q and state_out.body_q are required")
            c.body_q = _native.ptr(cur.body_q)
            c.body_q_prev = _native.ptr(state_in.body_q)
        c.dt = float(dt)
        return c

    # ------------------------------------------------------------ public API
    def step(self, state_in: State, state_out: State, control: Control | None, contacts: Contacts | None,
             dt: float) -> None:
        """One VBD step (``solver_vbd.py:1773-1825``).

        ``state_in.particle_q`` is overwritten with the iterate;
        ``state_out.particle_q`` / ``particle_qd`` receive the result.
        """
        if self.model.particle_count == 0:
            return
        self._ensure_context(contacts)
        cd = self._contacts_desc(contacts, state_in, state_out, dt)
        f_ext = state_in.particle_f
        rc = _native.hip_lib().vbd_step(
            self._ctx,
            current_stream_handle(self.model.device),
            _native.ptr(state_in.particle_q),
            _native.ptr(state_in.particle_qd),
            _native.ptr(f_ext),
            _native.ptr(state_out.particle_q),
            _native.ptr(state_out.particle_qd),
            ctypes.byref(cd) if cd is not None else None,
            float(dt),
        )
        _native.check(rc, "SolverVBD.step")

    # pieces of step(), for drivers that interleave residual checks
    def begin_step(self, state_in: State, contacts: Contacts | None, dt: float, state_out: State | None = None) -> None:
        """Contact preparation and the inertial forward step (``solver_vbd.py:2065-2100``, ``:2493-2514``);
        ``state_out.body_q`` (default: ``state_in.body_q``) poses the bodies at the step's end."""
        self._ensure_context(contacts)
        cd = self._contacts_desc(contacts, state_in, state_out, dt)
        lib = _native.hip_lib()
        s = current_stream_handle(self.model.device)
        _native.check(lib.vbd_contact_prepare(self._ctx, s, ctypes.byref(cd) if cd is not None else None),
                      "vbd_contact_prepare")
        _native.check(
            lib.vbd_forward(self._ctx, s, _native.ptr(state_in.particle_q), _native.ptr(state_in.particle_qd),
                            _native.ptr(state_in.particle_f), float(dt)),
            "vbd_forward",
        )

    def iterate(self, state_in: State, dt: float, iterations: int = 1) -> None:
        """Run ``iterations`` VBD iterations (all colors) on ``state_in.particle_q``."""
        lib = _native.hip_lib()
        s = current_stream_handle(self.model.device)
        for _ in range(int(iterations)):
            _native.check(lib.vbd_iteration(self._ctx, s, _native.ptr(state_in.particle_q), float(dt)),
                          "vbd_iteration")

    def color_pass(self, state_in: State, color: int, dt: float) -> None:
        _native.check(
            _native.hip_lib().vbd_color_pass(self._ctx, current_stream_handle(self.model.device), int(color),
                                             _native.ptr(state_in.particle_q), float(dt)),
            "vbd_color_pass",
        )

    def end_step(self, state_in: State, state_out: State, dt: float) -> None:
        """Final copy and velocity update (``solver_vbd.py:3158-3169``)."""
        _native.check(
            _native.hip_lib().vbd_finalize(self._ctx, current_stream_handle(self.model.device),
                                           _native.ptr(state_in.particle_q), _native.ptr(state_out.particle_q),
                                           _native.ptr(state_out.particle_qd), float(dt)),
            "vbd_finalize",
        )

    def residual(self, state_in: State, dt: float, g_out: torch.Tensor | None = None) -> torch.Tensor:
        """Squared force-residual norm (device double tensor of shape [1]) at ``state_in.particle_q``."""
        out = torch.empty(1, dtype=torch.float64, device=self.model.device)
        _native.check(
            _native.hip_lib().vbd_residual(self._ctx, current_stream_handle(self.model.device),
                                           _native.ptr(state_in.particle_q), float(dt), _native.ptr(g_out),
                                           _native.ptr(out)),
            "vbd_residual",
        )
        return out

    # ------------------------------------------------------------ self-contact
    def collision_detection(self, state: State) -> None:
        """Self-contact detection at ``state.particle_q`` (``solver_vbd.py:3221-3237``); the positions
        become the base of the displacements, as in the reference."""
        if not self.particle_enable_self_contact:
            raise RuntimeError("self-contact is off (particle_enable_self_contact=False)")
        _native.check(_native.hip_lib().vbd_self_detect(self._ctx, current_stream_handle(self.model.device),
                                                        _native.ptr(state.particle_q)), "vbd_self_detect")

    def collision_info(self) -> dict:
        """The last detection's lists as device tensors: ``vertex_colliding_triangles`` [N, vertex buffer]
        (-1 padded, ascending), ``..._count``, ``..._min_dist``, ``triangle_colliding_vertices_min_dist``,
        ``edge_colliding_edges`` [E, edge buffer], ``..._count``, ``..._min_dist``."""
        if not self.particle_enable_self_contact:
            raise RuntimeError("self-contact is off (particle_enable_self_contact=False)")
        m = self.model
        dev = m.device
        i32, f32 = torch.int32, torch.float32
        out = {
            "vertex_colliding_triangles": torch.empty((m.particle_count, self.particle_vertex_contact_buffer_size),
                                                      dtype=i32, device=dev),
            "vertex_colliding_triangles_count": torch.empty(m.particle_count, dtype=i32, device=dev),
            "vertex_colliding_triangles_min_dist": torch.empty(m.particle_count, dtype=f32, device=dev),
            "triangle_colliding_vertices_min_dist": torch.empty(m.tri_count, dtype=f32, device=dev),
            "edge_colliding_edges": torch.empty((m.edge_count, self.particle_edge_contact_buffer_size), dtype=i32,
                                                device=dev),
            "edge_colliding_edges_count": torch.empty(m.edge_count, dtype=i32, device=dev),
            "edge_colliding_edges_min_dist": torch.empty(m.edge_count, dtype=f32, device=dev),
        }
        _native.check(
            _native.hip_lib().vbd_self_copy_lists(self._ctx, current_stream_handle(dev),
                                                  *[_native.ptr(t) for t in out.values()]),
            "vbd_self_copy_lists",
        )
        return out

    def rebuild_bvh(self, state: State) -> None:
        """``solver_vbd.py:3239-3249``: the reference rebuilds its self-contact BVHs from ``state``.  The
        detection grid here is rebuilt from the current positions at every detection, so there is
        nothing to refresh; kept for callers of the reference's API."""
        if not self.particle_enable_self_contact:
            return

    def self_contact_overflow(self) -> bool:
        """True when a primitive spanned more detection-grid cells than a detection lists (synchronises)."""
        flag = ctypes.c_int32(0)
        _native.check(_native.hip_lib().vbd_self_status(self._ctx, ctypes.byref(flag)), "vbd_self_status")
        return bool(flag.value)

    def set_iterate(self, state_in: State, guess: torch.Tensor, commit: bool = True,
                    out: torch.Tensor | None = None) -> torch.Tensor:
        """Iterate from ``guess`` ([N,3] positions) after :meth:`begin_step`.

        Solved particles get ``x0 + (guess - x0)``, the others keep ``x0``.
        ``commit`` writes the solver's iterate (``state_in.particle_q`` and its
        displacement); otherwise ``out`` (or a new tensor) receives the trial
        positions, e.g. to judge the guess with :meth:`residual` first.
        """
        guess = guess.contiguous()
        if guess.shape != (self.model.particle_count, 3) or guess.dtype != torch.float32:
            raise ValueError("guess must be a float32 tensor of shape [particle_count, 3]")
        dst = state_in.particle_q if commit else (out if out is not None else torch.empty_like(guess))
        _native.check(
            _native.hip_lib().vbd_set_iterate(self._ctx, current_stream_handle(self.model.device), _native.ptr(dst),
                                              _native.ptr(guess), 1 if commit else 0),
            "vbd_set_iterate",
        )
        return dst

    def set_iterate_range(self, state_in: State, guess: torch.Tensor, p_begin: int, p_end: int) -> None:
        """:meth:`set_iterate` (committed) for the particles ``p_begin <= i < p_end`` only; the others keep
        their iterate (``guess`` is the full [N,3] array)."""
        guess = guess.contiguous()
        if guess.shape != (self.model.particle_count, 3) or guess.dtype != torch.float32:
            raise ValueError("guess must be a float32 tensor of shape [particle_count, 3]")
        _native.check(
            _native.hip_lib().vbd_set_iterate_range(self._ctx, current_stream_handle(self.model.device),
                                                    _native.ptr(state_in.particle_q), _native.ptr(guess), 1,
                                                    int(p_begin), int(p_end)),
            "vbd_set_iterate_range",
        )

    def iterate_to_residual(self, state_in: State, dt: float, target_norm_sq: float, max_iterations: int = 200,
                            check_every: int = 1) -> tuple[int, float]:
        """Iterate until ``|g|^2 <= target_norm_sq`` (checked every ``check_every`` iterations).

        Host-synchronous.  Returns ``(iterations run, final |g|^2)``.
        """
        its = ctypes.c_int32(0)
        r = ctypes.c_double(0.0)
        _native.check(
            _native.hip_lib().vbd_iterate_to_residual(self._ctx, current_stream_handle(self.model.device),
                                                      _native.ptr(state_in.particle_q), float(dt),
                                                      float(target_norm_sq), int(max_iterations), int(check_every),
                                                      ctypes.byref(its), ctypes.byref(r)),
            "vbd_iterate_to_residual",
        )
        return int(its.value), float(r.value)

    # ------------------------------------------------------------ per-world residuals
    def world_offsets(self) -> torch.Tensor:
        """Device int32 [world_count + 1]: world w owns particles ``off[w] <= i < off[w+1]``.

        Requires every particle in a world and the worlds in contiguous, ascending particle
        ranges (the layout :meth:`ModelBuilder.add_world` produces); raises ``ValueError`` otherwise.
        """
        if getattr(self, "_world_off", None) is None:
            m = self.model
            pw = m.particle_world.cpu().numpy().astype(np.int64) if m.particle_count else np.zeros(0, np.int64)
            nw = max(int(m.world_count), 1)
            if pw.size and (pw.min() < 0 or (np.diff(pw) < 0).any() or pw.max() >= nw):
                raise ValueError("per-world residuals need every particle in a world, in ascending world order")
            off = np.searchsorted(pw, np.arange(nw + 1), side="left").astype(np.int32)
            self._world_off = torch.as_tensor(off, device=m.device)
            self._world_of_particle = torch.as_tensor(pw, device=m.device)
            self._world_of_particle_i32 = self._world_of_particle.to(torch.int32).contiguous()
            self._world_scratch_dev = torch.empty(2 * nw, dtype=torch.float64, device=m.device)
            self._world_state_dev = torch.empty(2 * nw + 1, dtype=torch.int32, device=m.device)
        return self._world_off

    def residual_worlds(self, state_in: State, dt: float) -> torch.Tensor:
        """Squared force-residual norm of every world (device double tensor [world_count])."""
        off = self.world_offsets()
        out = torch.empty(off.numel() - 1, dtype=torch.float64, device=self.model.device)
        _native.check(
            _native.hip_lib().vbd_residual_worlds(self._ctx, current_stream_handle(self.model.device),
                                                  _native.ptr(state_in.particle_q), float(dt), off.numel() - 1,
                                                  _native.ptr(off), _native.ptr(out)),
            "vbd_residual_worlds",
        )
        return out

    def iterate_to_residual_worlds(self, state_in: State, dt: float, target_norm_sq, max_iterations: int = 200,
                                   check_every: int = 1, per_world: bool = False, stop_at_live: int = 0):
        """Iterate every world until its ``|g_w|^2 <= target_norm_sq[w]``, then freeze it (its vertices
        are no longer solved, so its iterate is the one a solve of that world alone stops at);
        host-synchronous.

        Returns ``(iterations run, last squared norms [world_count])``, plus each world's iteration
        count when ``per_world``.  ``stop_at_live`` > 0 stops once at most that many worlds are live
        (e.g. ``world_count - 1``: at the first world that reaches its target).
        """
        off = self.world_offsets()
        nw = off.numel() - 1
        tgt = np.ascontiguousarray(np.broadcast_to(np.asarray(target_norm_sq, dtype=np.float64), (nw,)))
        out = np.zeros(nw, dtype=np.float64)
        wits = np.zeros(nw, dtype=np.int32)
        its = ctypes.c_int32(0)
        dptr = ctypes.POINTER(ctypes.c_double)
        _native.check(
            _native.hip_lib().vbd_iterate_to_residual_worlds(
                self._ctx, current_stream_handle(self.model.device), _native.ptr(state_in.particle_q), float(dt), nw,
                _native.ptr(off), _native.ptr(self._world_of_particle_i32), tgt.ctypes.data_as(dptr),
                int(max_iterations), int(check_every), int(stop_at_live), _native.ptr(self._world_scratch_dev),
                _native.ptr(self._world_state_dev), ctypes.byref(its), wits.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                out.ctypes.data_as(dptr)),
            "vbd_iterate_to_residual_worlds",
        )
        if per_world:
            return int(its.value), out, wits
        return int(its.value), out

    def step_to_residual_worlds(self, state_in: State, state_out: State, control: Control | None,
                                contacts: Contacts | None, dt: float, rel_tol: float = 1.0e-3,
                                max_iterations: int = 200, check_every: int = 1, warm_start=None,
                                trial_iterations: int = 0, abs_tol=0.0) -> dict:
        """:meth:`step_to_residual` for a model of several worlds, each judged by its own residual.

        Each world is frozen once ``r_w <= rel_tol * r0_w`` (its vertices are no longer solved, and a
        batch of frozen vertices costs one header read), so each world's result is the iterate a solve
        of that world alone stops at; the batch runs until all worlds are frozen (or
        ``max_iterations``).  ``world_iterations`` reports each world's count.  Warm-start candidates are
        chosen per world: after up to ``trial_iterations`` iterations from each start (0: at the start
        itself), each world keeps the start with its lowest residual (a candidate must beat the
        inertial start strictly) and continues from that start's iterate.  The trial iterations are the
        first iterations of an ordinary per-world solve: a world within its tolerance stops there (a
        world whose inertial start is already within it runs no trial iterations), and each world's
        count includes the trial iterations every start ran for it.

        ``abs_tol`` (a norm in force units, scalar or one per world) floors each world's target at
        ``abs_tol**2``: a world whose inertial guess is already at round-off (free fall of an unstressed
        body) stops there instead of chasing a relative tolerance its FP32 residual cannot reach.
        """
        self.world_offsets()
        self.begin_step(state_in, contacts, dt, state_out)
        r0 = self.residual_worlds(state_in, dt).cpu().numpy()
        nw = r0.size
        target = np.maximum((rel_tol * rel_tol) * r0, np.square(np.broadcast_to(np.asarray(abs_tol, np.float64),
                                                                                  (nw,))))
        n_cand = 0
        chosen = np.full(nw, -1, dtype=np.int64)
        trial_its = 0
        trial_world = np.zeros(nw, dtype=np.int64)
        if warm_start is not None:
            guesses = warm_start(self, state_in, dt)
            if guesses is not None and not isinstance(guesses, (list, tuple)):
                guesses = [guesses]
            guesses = list(guesses or [])
            n_cand = len(guesses)
            if guesses:
                wp = self._world_of_particle
                if trial_iterations > 0:
                    # each start runs up to k iterations of the per-world solve; its iterate is kept
                    # for the worlds it wins
                    # residuals before and after the k iterations only (a check after every trial
                    # iteration would cost a residual pass each)
                    its, best_r, wits = self.iterate_to_residual_worlds(state_in, dt, target, trial_iterations,
                                                                        trial_iterations, per_world=True)
                    trial_its += its
                    trial_world += wits
                    best_q = state_in.particle_q.clone()
                    best_d = self.scratch("displacement")
                    # worlds the inertial start already brought within tolerance keep it; the candidates
                    # are tried on the others only (an infinite target freezes a world at once)
                    done = best_r <= target
                    cand_target = np.where(done, np.inf, target)
                    for k, g in enumerate(guesses):
                        self.set_iterate(state_in, g, commit=True)
                        its, r, wits = self.iterate_to_residual_worlds(state_in, dt, cand_target, trial_iterations,
                                                                       trial_iterations, per_world=True)
                        trial_its += its
                        trial_world += wits
                        better = (r < best_r) & ~done
                        if better.any():
                            best_r = np.where(better, r, best_r)
                            chosen[better] = k
                            take = torch.as_tensor(better, device=wp.device)[wp][:, None]
                            best_q = torch.where(take, state_in.particle_q, best_q)
                            best_d = torch.where(take, self.scratch("displacement"), best_d)
                    # every world continues from its winning start's iterate exactly (positions and
                    # displacement as saved; a world that kept the inertial start equals a plain solve)
                    state_in.particle_q.copy_(best_q)
                    self.set_displacement(best_d)
                else:
                    # judged at the starts themselves; the solver state is untouched unless a
                    # candidate wins somewhere, and then the mixed guess restarts every world
                    best_r = r0.copy()
                    best_g = state_in.particle_q.clone()
                    d_inertial = self.scratch("displacement")
                    trial = None
                    for k, g in enumerate(guesses):
                        trial = self.set_iterate(state_in, g, commit=False, out=trial)
                        r = self.residual_worlds(_TrialState(trial), dt).cpu().numpy()
                        better = r < best_r
                        if better.any():
                            best_r = np.where(better, r, best_r)
                            chosen[better] = k
                            take = torch.as_tensor(better, device=wp.device)[wp]
                            best_g = torch.where(take[:, None], g, best_g)
                    if (chosen >= 0).any():
                        self.set_iterate(state_in, best_g, commit=True)
                        # worlds that keep the inertial start keep its displacement bit for bit
                        keep = torch.as_tensor(chosen < 0, device=wp.device)[wp][:, None]
                        self.set_displacement(torch.where(keep, d_inertial, self.scratch("displacement")))
        its, r_sq, wits = self.iterate_to_residual_worlds(state_in, dt, target,
                                                          max(max_iterations - trial_iterations, 0) if trial_iterations
                                                          and n_cand else max_iterations, check_every, per_world=True)
        self.end_step(state_in, state_out, dt)
        return {"iterations": its + trial_its, "world_iterations": wits + trial_world, "r0": np.sqrt(r0),
                "r": np.sqrt(r_sq),
                "reached": int((r_sq <= target).sum()), "worlds": nw,
                "candidate_chosen": [int((chosen == k).sum()) for k in range(n_cand)]}

    def step_to_residual(self, state_in: State, state_out: State, control: Control | None,
                         contacts: Contacts | None, dt: float, rel_tol: float = 1.0e-3, max_iterations: int = 200,
                         check_every: int = 1, warm_start=None, trial_iterations: int = 0) -> dict:
        """One substep iterated to a relative force residual instead of a fixed count.

        The residual is normalised by its value at the inertial guess ``r0``;
        iterations stop once ``r <= rel_tol * r0`` (or after ``max_iterations``).
        ``warm_start(solver, state_in, dt)`` may return candidate positions
        ([N,3] float32) or a list of them; the candidate with the lowest
        residual replaces the inertial guess only if that residual is strictly
        lower than ``r0`` (ties keep the earlier candidate), otherwise the
        inertial guess is kept.  ``warm_start_accepted`` reports the index of the
        chosen candidate (None: the inertial guess).

        With ``trial_iterations`` = k > 0 every start (the inertial guess first, then each
        candidate) is judged by its residual after k VBD iterations instead: Gauss-Seidel removes a
        candidate's rough error in a few sweeps, so the comparison sees what the start leaves for the
        remaining iterations.  The winning start's iterate continues (a candidate must beat the
        inertial start strictly); every trial iteration is counted in ``iterations``.
        """
        self.begin_step(state_in, contacts, dt, state_out)
        r0_dev = self.residual(state_in, dt)
        chosen = None
        r_ws = None
        r_candidates = []
        if warm_start is not None and trial_iterations > 0:
            guesses = warm_start(self, state_in, dt)
            if guesses is not None and not isinstance(guesses, (list, tuple)):
                guesses = [guesses]
            guesses = list(guesses or [])
            r0_sq = float(r0_dev.item())
            self.iterate(state_in, dt, trial_iterations)
            best_q = state_in.particle_q.clone()
            best_d = self.scratch("displacement")
            best_r = float(self.residual(state_in, dt).item())
            r_candidates = [math.sqrt(best_r)]
            for k, g in enumerate(guesses):
                self.set_iterate(state_in, g, commit=True)
                self.iterate(state_in, dt, trial_iterations)
                r = float(self.residual(state_in, dt).item())
                r_candidates.append(math.sqrt(r))
                if r < best_r:
                    best_r, chosen = r, k
                    best_q.copy_(state_in.particle_q)
                    best_d = self.scratch("displacement")
            # the winning start's iterate continues exactly: positions and displacement restored as saved
            state_in.particle_q.copy_(best_q)
            self.set_displacement(best_d)
            r_ws = math.sqrt(best_r)
            its, r_sq = self.iterate_to_residual(state_in, dt, (rel_tol * rel_tol) * r0_sq,
                                                 max(max_iterations - trial_iterations, 0), check_every)
            self.end_step(state_in, state_out, dt)
            return {"iterations": its + trial_iterations * (1 + len(guesses)), "r0": math.sqrt(r0_sq),
                    "r": math.sqrt(r_sq), "warm_start_accepted": chosen is not None, "warm_start_index": chosen,
                    "r_warm_start": r_ws, "r_candidates": r_candidates}
        if warm_start is not None:
            guesses = warm_start(self, state_in, dt)
            if guesses is not None and not isinstance(guesses, (list, tuple)):
                guesses = [guesses]
            if guesses:
                trial = None
                norms = []
                for g in guesses:
                    trial = self.set_iterate(state_in, g, commit=False, out=trial)
                    norms.append(self.residual(_TrialState(trial), dt))
                r0_sq = float(r0_dev.item())
                r_candidates = [math.sqrt(float(n.item())) for n in norms]
                best = min(range(len(guesses)), key=lambda k: (r_candidates[k], k))
                r_ws = r_candidates[best]
                if r_ws * r_ws < r0_sq:
                    self.set_iterate(state_in, guesses[best], commit=True)
                    chosen = best
        r0_sq = float(r0_dev.item())
        its, r_sq = self.iterate_to_residual(state_in, dt, (rel_tol * rel_tol) * r0_sq, max_iterations, check_every)
        self.end_step(state_in, state_out, dt)
        return {"iterations": its, "r0": math.sqrt(r0_sq), "r": math.sqrt(r_sq),
                "warm_start_accepted": chosen is not None, "warm_start_index": chosen, "r_warm_start": r_ws,
                "r_candidates": r_candidates}

    def set_displacement(self, disp: torch.Tensor) -> None:
        """Restore a displacement saved with ``scratch("displacement")`` together with the positions it
        belongs to (``vbd_set_displacement``), so that saved iterate continues bit for bit."""
        _native.check(_native.hip_lib().vbd_set_displacement(self._ctx, current_stream_handle(self.model.device),
                                                             _native.ptr(disp.contiguous())),
                      "vbd_set_displacement")

    def scratch(self, which: str) -> torch.Tensor:
        """Copy of the current substep's ``inertia`` / ``q_prev`` / ``displacement`` array ([N, 3])."""
        idx = {"inertia": 0, "q_prev": 1, "displacement": 2}[which]
        out = torch.empty((self.model.particle_count, 3), dtype=torch.float32, device=self.model.device)
        _native.check(
            _native.hip_lib().vbd_copy_scratch(self._ctx, current_stream_handle(self.model.device), idx,
                                               _native.ptr(out)),
            "vbd_copy_scratch",
        )
        return out

    def notify_model_changed(self, flags=0) -> None:
        """Re-read the model after the caller edited it (``solver.py:375-410``, ``solver_vbd.py:851-856``).

        Element arrays (materials, rest data), masses and flags are re-read into the color-pass
        layout now, into the same device buffers when the layout's shape is unchanged (so a graph
        captured before stays valid); the soft-contact scalars are handed over now as well.  The
        solver keeps its context and its pass shape."""
        if self._ctx is not None:
            torch.cuda.synchronize(self.model.device)
            _native.check(_native.hip_lib().vbd_rebuild(self._ctx), "vbd_rebuild")
            self._push_scalars()
        self._world_off = None  # particle worlds may have changed

    def autotune(self, state_in: State, contacts: Contacts | None, dt: float) -> dict:
        """Pick the color-pass shape by timing the candidates on this model: 256-lane work-groups at the
        compiler's register budget, 256 lanes at seven waves per SIMD, 128 lanes, 256 lanes with
        wave-aligned item ranges at either budget, and 192 lanes wave-aligned at either budget.

        Runs on a copy of ``state_in`` (prepared with contacts and the forward
        step); ``state_in`` is left untouched.  Results never depend on the shape.
        """
        self._ensure_context(contacts)
        trial = State()
        trial.particle_q = state_in.particle_q.clone()
        trial.particle_qd = state_in.particle_qd.clone()
        trial.particle_f = state_in.particle_f.clone()
        trial.body_q = state_in.body_q
        self.begin_step(trial, contacts, dt)
        shape = (ctypes.c_int32 * 3)()
        ms = (ctypes.c_float * 7)()
        _native.check(_native.hip_lib().vbd_autotune(self._ctx, current_stream_handle(self.model.device),
                                                     _native.ptr(trial.particle_q), float(dt), shape, ms),
                      "vbd_autotune")
        self._pass_shape = (int(shape[0]), int(shape[1]), int(shape[2]))
        return {"block": int(shape[0]), "waves": int(shape[1]), "aligned": int(shape[2]), "ms_256": float(ms[0]),
                "ms_256_w7": float(ms[1]), "ms_128": float(ms[2]), "ms_256_aligned": float(ms[3]),
                "ms_256_w7_aligned": float(ms[4]), "ms_192_aligned": float(ms[5]),
                "ms_192_w7_aligned": float(ms[6])}

    def pass_shape(self) -> tuple[int, int, int]:
        """The color pass's current shape: (lanes per work-group, waves target, wave-aligned items)."""
        out = (ctypes.c_int32 * 3)()
        _native.check(_native.hip_lib().vbd_get_pass_shape(self._ctx, out), "vbd_get_pass_shape")
        return int(out[0]), int(out[1]), int(out[2])

    def set_pass_shape(self, block: int, waves: int = 1, aligned: bool = False) -> None:
        """Use ``block``-lane color-pass work-groups (256, 192 or 128) at an occupancy target of ``waves``
        (1: the compiler's register budget, 7: seven waves per SIMD) with item ranges packed or each on a
        wave boundary (``aligned``): what :meth:`autotune` chooses between.  Rebuilds the layout once."""
        _native.check(_native.hip_lib().vbd_set_pass_shape(self._ctx, current_stream_handle(self.model.device),
                                                           int(block), int(waves), 1 if aligned else 0),
                      "vbd_set_pass_shape")
        self._pass_shape = (int(block), int(waves), 1 if aligned else 0)

    def set_aligned_items(self, aligned: bool) -> None:
        """Item ranges of a color-pass batch packed, or each on a wave boundary (a wave evaluates one
        element kind only; fewer vertices per batch); what :meth:`autotune` also chooses between."""
        block, waves, _ = self.pass_shape()
        self.set_pass_shape(block, waves, aligned)

    def set_pass_waves(self, waves: int) -> None:
        """Occupancy target of the color pass: 1 (compiler's register budget) or 7 waves per SIMD (what
        :meth:`autotune` also chooses between)."""
        block, _, aligned = self.pass_shape()
        self.set_pass_shape(block, waves, bool(aligned))

    def set_block_shape(self, block: int) -> None:
        """Use 256-, 192- or 128-lane color-pass work-groups (what :meth:`autotune` chooses between)."""
        _, waves, aligned = self.pass_shape()
        self.set_pass_shape(block, waves, bool(aligned))

    def layout_stats(self) -> dict:
        """Color-pass layout: packed batches, vertices on the generic path, packed element items, colors."""
        out = (ctypes.c_int64 * 4)()
        _native.check(_native.hip_lib().vbd_layout_stats(self._ctx, out), "vbd_layout_stats")
        return {"batches": out[0], "generic_vertices": out[1], "items": out[2], "colors": out[3]}

    def reset(self, state: State, world_mask=None, flags=None) -> None:
        """``solver_vbd.py:1828-1935`` for a body-free model: the arguments are checked as the reference
        checks them; the particle path keeps no history across steps, and, as in the reference, a
        request to reset particle positions or velocities is ignored with a warning."""
        if state is None:
            raise ValueError("'state' argument is required.")
        model = self.model
        if world_mask is not None:
            if not isinstance(world_mask, torch.Tensor) or world_mask.ndim != 1 or world_mask.dtype != torch.bool:
                raise ValueError("world_mask must be a one-dimensional boolean tensor.")
            n = int(world_mask.shape[0])
            if n not in (model.world_count, model.world_count + 1):
                raise ValueError(f"world_mask has length {n}, expected {model.world_count} or {model.world_count + 1}.")
            if world_mask.device != torch.device(self.device):
                raise ValueError(f"world_mask is on device {world_mask.device}, expected solver device {self.device}.")
        flags_value = int(StateFlags.ALL if flags is None else flags)
        if model.particle_count > 0 and (flags_value & (int(StateFlags.PARTICLE_Q) | int(StateFlags.PARTICLE_QD))):
            warnings.warn(
                "SolverVBD.reset() does not yet support particle resets; StateFlags.PARTICLE_Q and "
                "StateFlags.PARTICLE_QD are ignored; particle and body-particle solver history is unchanged.",
                stacklevel=2,
            )

    @classmethod
    def register_custom_attributes(cls, builder, *, dahl_defaults_enabled: bool = False) -> None:
        """``solver_vbd.py:1534-1560`` registers joint and cable attributes (``vbd:joint_is_hard``,
        ``vbd:dahl_*``); this package has no joints or cables, so there is nothing to register."""
        return None


__all__ = ["SolverVBD", "color_layout"]
