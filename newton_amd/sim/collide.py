"""Soft-contact generation between particles / mesh features and plane or sphere shapes.

``CollisionPipeline`` mirrors the soft part of the reference's
``newton/_src/sim/collide.py`` (``CollisionPipeline`` :710, soft-contact sizing
:1110-1136, ``collide`` :1236, launches :1553-1610, the full-surface edge / face
passes of ``geometry/soft_contacts_sdf.py``).  Candidate pairs are built on the
host once, in the order of ``_world_compatible_pairs`` (:480-546); the contacts
are generated on the GPU by ``libvbd_hip`` (``vbd_collide_generate`` /
``vbd_collide_generate_bodies``), which writes records in pair order (a stable
compaction: particle records, then edge records, then face records), where the
reference takes slots from an atomic counter.
"""

from __future__ import annotations

import ctypes
import warnings

import numpy as np
import torch

from .. import _native
from .model import Contacts, Model, State


def world_compatible_pairs(feature_world: np.ndarray, shape_world: np.ndarray, world_count: int,
                           shape_ok: np.ndarray | None = None) -> np.ndarray:
    """``(feature, shape)`` pairs whose worlds are compatible (``collide.py:480-546``); ``shape_ok`` (per
    shape) drops the pairs of shapes that cannot take part, keeping the order of the others."""
    pairs = _world_pairs(feature_world, shape_world, world_count)
    if shape_ok is not None and len(pairs):
        pairs = pairs[np.asarray(shape_ok, dtype=bool)[pairs[:, 1]]]
    return np.ascontiguousarray(pairs, dtype=np.int32)


def full_surface_capable_shapes(model: Model) -> np.ndarray:
    """Shapes that can generate edge/face contacts (``collide.py:607-640``): spheres and infinite planes
    (width = length = 0) among this path's shape types."""
    if not model.shape_count:
        return np.zeros(0, dtype=bool)
    types = model.shape_type.cpu().numpy()
    scale = model.shape_scale.cpu().numpy()
    return (types == 3) | ((types == 1) & (scale[:, 0] == 0.0) & (scale[:, 1] == 0.0))


def _world_pairs(feature_world: np.ndarray, shape_world: np.ndarray, world_count: int) -> np.ndarray:
    feature_world = np.asarray(feature_world)
    shape_world = np.asarray(shape_world)
    n_features, n_shapes = len(feature_world), len(shape_world)
    if n_features == 0 or n_shapes == 0:
        return np.empty((0, 2), dtype=np.int32)
    features = np.arange(n_features)
    shapes = np.arange(n_shapes)
    f_local = (feature_world >= 0) & (feature_world < world_count)
    s_local = (shape_world >= 0) & (shape_world < world_count)
    f_cols, s_cols = [], []
    global_features = features[feature_world < 0]
    if len(global_features):
        f_cols.append(np.repeat(global_features, len(shapes)))
        s_cols.append(np.tile(shapes, len(global_features)))
    local_features = features[f_local]
    global_shapes = shapes[shape_world < 0]
    if len(local_features) and len(global_shapes):
        f_cols.append(np.repeat(local_features, len(global_shapes)))
        s_cols.append(np.tile(global_shapes, len(local_features)))
    local_feature_world = feature_world[f_local]
    shapes_per_world = np.bincount(shape_world[s_local], minlength=world_count)
    reps = shapes_per_world[local_feature_world] if len(local_feature_world) else np.zeros(0, np.intp)
    if reps.sum():
        shapes_by_world = shapes[s_local][np.argsort(shape_world[s_local], kind="stable")]
        world_start = np.cumsum(shapes_per_world) - shapes_per_world
        within = np.arange(reps.sum()) - np.repeat(np.cumsum(reps) - reps, reps)
        f_cols.append(np.repeat(local_features, reps))
        s_cols.append(shapes_by_world[np.repeat(world_start[local_feature_world], reps) + within])
    if not f_cols:
        return np.empty((0, 2), dtype=np.int32)
    return np.column_stack((np.concatenate(f_cols), np.concatenate(s_cols))).astype(np.int32)


def current_stream_handle(device) -> int:
    """Raw hipStream_t of torch's current stream on ``device``."""
    return torch.cuda.current_stream(device).cuda_stream


class CollisionPipeline:
    """Soft contacts between particles (and, with ``enable_rigid_soft_full_surface_contact``, mesh edges
    and triangles) and plane / sphere shapes, static or attached to bodies (``collide.py:710-1592``)."""

    def __init__(self, model: Model, *, soft_contact_max: int | None = None, soft_contact_margin: float = 0.01,
                 enable_rigid_soft_full_surface_contact: bool = False):
        self.model = model
        self.device = model.device
        pworld = model.particle_world.cpu().numpy() if model.particle_count else np.empty(0, np.int32)
        sworld = model.shape_world.cpu().numpy() if model.shape_count else np.empty(0, np.int32)
        pairs = world_compatible_pairs(pworld, sworld, model.world_count)
        if model.shape_count:
            types = model.shape_type.cpu().numpy()
            if np.any((types != 1) & (types != 3)):
                raise NotImplementedError("particle contact supports plane and sphere shapes only")
            body = model.shape_body.cpu().numpy()
            if np.any(body >= model.body_count):
                raise ValueError("shape_body names a body the model does not have")
        self.has_bodies = bool(model.shape_count and np.any(model.shape_body.cpu().numpy() >= 0))
        self.enable_rigid_soft_full_surface_contact = bool(enable_rigid_soft_full_surface_contact)
        edge_pairs = face_pairs = np.empty((0, 2), dtype=np.int32)
        if self.enable_rigid_soft_full_surface_contact and model.shape_count:
            capable = full_surface_capable_shapes(model)
            fallback = np.where(((model.shape_flags.cpu().numpy() & 4) != 0) & ~capable)[0]
            if fallback.size:
                warnings.warn(
                    "enable_rigid_soft_full_surface_contact=True: these participating shapes cannot generate "
                    f"edge/face contacts and fall back to per-particle soft contact only -- finite planes "
                    f"{[model.shape_label[i] if getattr(model, 'shape_label', None) else i for i in fallback]} "
                    "(only infinite planes are supported). Full-surface contacts still apply to the rest of the "
                    "scene.", stacklevel=2)
            if model.edge_count:
                ew = pworld[model.edge_indices.cpu().numpy()[:, 2]]
                edge_pairs = world_compatible_pairs(ew, sworld, model.world_count, capable)
            if model.tri_count:
                fw = pworld[model.tri_indices.cpu().numpy()[:, 0]]
                face_pairs = world_compatible_pairs(fw, sworld, model.world_count, capable)
        self.soft_rigid_contact_pairs_host = pairs
        self.soft_rigid_contact_pairs = torch.as_tensor(np.ascontiguousarray(pairs), device=self.device)
        self.soft_rigid_contact_pair_count = int(pairs.shape[0])
        self.soft_edge_rigid_pairs_host = edge_pairs
        self.soft_face_rigid_pairs_host = face_pairs
        self.soft_edge_rigid_pairs = torch.as_tensor(np.ascontiguousarray(edge_pairs), device=self.device)
        self.soft_face_rigid_pairs = torch.as_tensor(np.ascontiguousarray(face_pairs), device=self.device)
        if soft_contact_max is None:
            # one record per candidate pair (collide.py:1131-1134)
            soft_contact_max = self.soft_rigid_contact_pair_count + len(edge_pairs) + len(face_pairs)
        self.soft_contact_max = int(soft_contact_max)
        self.soft_contact_margin = float(soft_contact_margin)
        self._handle = None
        self.desc = None
        total = self.soft_rigid_contact_pair_count + len(edge_pairs) + len(face_pairs)
        if self.device.type == "cuda" and total > 0:
            lib = _native.hip_lib()
            d = _native.VbdCollideDesc()
            d.particle_count = model.particle_count
            d.pair_count = self.soft_rigid_contact_pair_count
            d.pairs = _native.ptr(self.soft_rigid_contact_pairs)
            d.particle_radius = _native.ptr(model.particle_radius)
            d.particle_flags = _native.ptr(model.particle_flags)
            d.particle_world = _native.ptr(model.particle_world)
            d.shape_count = model.shape_count
            d.shape_type = _native.ptr(model.shape_type)
            d.shape_flags = _native.ptr(model.shape_flags)
            d.shape_world = _native.ptr(model.shape_world)
            d.shape_transform = _native.ptr(model.shape_transform)
            d.shape_scale = _native.ptr(model.shape_scale)
            d.shape_margin = _native.ptr(model.shape_margin)
            d.shape_body = _native.ptr(model.shape_body) if self.has_bodies else None
            d.edge_pair_count = len(edge_pairs)
            d.edge_pairs = _native.ptr(self.soft_edge_rigid_pairs) if len(edge_pairs) else None
            d.edge_indices = _native.ptr(model.edge_indices) if len(edge_pairs) else None
            d.face_pair_count = len(face_pairs)
            d.face_pairs = _native.ptr(self.soft_face_rigid_pairs) if len(face_pairs) else None
            d.tri_indices = _native.ptr(model.tri_indices) if len(face_pairs) else None
            h = ctypes.c_void_p()
            _native.check(lib.vbd_collide_create(ctypes.byref(d), ctypes.byref(h)), "vbd_collide_create")
            self._handle = h
            self.desc = d  # the pair / shape arrays as the C ABI reads them

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            try:
                _native.hip_lib().vbd_collide_destroy(h)
            except Exception:
                pass
            self._handle = None

    def contacts(self) -> Contacts:
        """Allocate a :class:`Contacts` buffer sized for this pipeline."""
        c = Contacts(self.soft_contact_max, self.device)
        c._enable_rigid_soft_full_surface_contact = self.enable_rigid_soft_full_surface_contact
        return c

    def collide(self, state: State, contacts: Contacts, *, soft_contact_margin: float | None = None) -> None:
        """Generate soft contacts for ``state.particle_q`` (shapes on bodies placed by ``state.body_q``)."""
        if self.device.type != "cuda":
            raise RuntimeError("CollisionPipeline.collide runs on the GPU (libvbd_hip); the model is on the CPU")
        margin = self.soft_contact_margin if soft_contact_margin is None else float(soft_contact_margin)
        if contacts.soft_contact_max > self.soft_contact_max:
            raise ValueError("contacts buffer is larger than the pipeline's soft_contact_max")
        if self.enable_rigid_soft_full_surface_contact:
            contacts._enable_rigid_soft_full_surface_contact = True
        if self._handle is None:
            contacts.soft_contact_count.zero_()
            return
        lib = _native.hip_lib()
        s = current_stream_handle(self.device)
        if self.has_bodies:
            if state.body_q is None:
                raise ValueError("shapes are attached to bodies: state.body_q is required")
            rc = lib.vbd_collide_generate_bodies(self._handle, s, _native.ptr(state.particle_q),
                                                 _native.ptr(state.body_q), int(state.body_q.shape[0]), margin,
                                                 ctypes.byref(contacts_out(contacts)))
        else:
            rc = lib.vbd_collide_generate(self._handle, s, _native.ptr(state.particle_q), margin,
                                          ctypes.byref(contacts_out(contacts)))
        _native.check(rc, "vbd_collide_generate")


def contacts_out(contacts: Contacts) -> _native.VbdContactsOut:
    """The writable soft-contact arrays of ``contacts`` as the C ABI's ``VbdContactsOut``."""
    out = _native.VbdContactsOut()
    out.soft_contact_count = _native.ptr(contacts.soft_contact_count)
    out.soft_contact_max = contacts.soft_contact_max
    out.soft_contact_particle = _native.ptr(contacts.soft_contact_particle)
    out.soft_contact_indices = _native.ptr(contacts.soft_contact_indices)
    out.soft_contact_barycentric = _native.ptr(contacts.soft_contact_barycentric)
    out.soft_contact_shape = _native.ptr(contacts.soft_contact_shape)
    out.soft_contact_body_pos = _native.ptr(contacts.soft_contact_body_pos)
    out.soft_contact_body_vel = _native.ptr(contacts.soft_contact_body_vel)
    out.soft_contact_normal = _native.ptr(contacts.soft_contact_normal)
    return out


__all__ = ["CollisionPipeline", "contacts_out", "full_surface_capable_shapes", "world_compatible_pairs"]
