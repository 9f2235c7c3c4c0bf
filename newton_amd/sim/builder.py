"""ModelBuilder subset for the particle VBD path.

Restates the parts of the reference's ``newton/_src/sim/builder.py`` that the
VBD path consumes: particles, triangles (2-D rest basis and areas in float64),
bending edges (rest angle / length), tetrahedra, cloth grid/mesh helpers,
soft tet grids, static plane / sphere shapes, coloring, world merging and
``finalize``.  Storage is chunked numpy so multi-world batches build quickly;
the float64 element pre-computations use the same numpy operations as the
reference so the resulting float32 arrays match it.
"""

from __future__ import annotations

import copy
import math
from dataclasses import dataclass

import numpy as np
import torch

from ..geometry import GeoType, ParticleFlags, ShapeFlags
from ..utils.mesh import MeshAdjacency, compute_edge_adjacency
from . import graph_coloring as gc
from .model import Model

_AXES = {"X": 0, "Y": 1, "Z": 2}


def _axis_index(axis) -> int:
    if isinstance(axis, str):
        return _AXES[axis.upper()]
    return int(axis)


def quat_from_axis_angle(axis, angle: float) -> np.ndarray:
    """Unit quaternion (x, y, z, w) in float32 for a rotation of ``angle`` about ``axis``."""
    a = np.asarray(axis, dtype=np.float32)
    a = a / np.float32(np.linalg.norm(a))
    h = np.float32(angle * 0.5)
    s = np.float32(math.sin(float(h)))
    return np.array([a[0] * s, a[1] * s, a[2] * s, np.float32(math.cos(float(h)))], dtype=np.float32)


def quat_to_matrix(q) -> np.ndarray:
    """Rotation matrix of a unit quaternion (x, y, z, w): columns are the rotated basis vectors (float32)."""
    return np.stack([quat_rotate(q, e) for e in np.eye(3, dtype=np.float32)], axis=1).astype(np.float32)


def quat_rotate(q, v) -> np.ndarray:
    """Rotate float32 vectors ``v`` ([n,3] or [3]) by the unit quaternion ``q`` (x, y, z, w)."""
    v = np.asarray(v, dtype=np.float32)
    qx, qy, qz, qw = (np.float32(c) for c in q)
    two, one = np.float32(2.0), np.float32(1.0)
    c = two * qw * qw - one
    x, y, zz = v[..., 0], v[..., 1], v[..., 2]
    d = two * (qx * x + qy * y + qz * zz)
    return np.stack(
        [
            x * c + qx * d + (qy * zz - qz * y) * qw * two,
            y * c + qy * d + (qz * x - qx * zz) * qw * two,
            zz * c + qz * d + (qx * y - qy * x) * qw * two,
        ],
        axis=-1,
    ).astype(np.float32)


def quat_between_vectors(a, b) -> np.ndarray:
    """Shortest-arc unit quaternion rotating ``a`` onto ``b`` (float32)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    a = a / np.linalg.norm(a)
    b = b / np.linalg.norm(b)
    c = np.cross(a, b)
    d = float(np.dot(a, b))
    if d < -1.0 + 1e-12:
        # 180 degrees: any axis orthogonal to a
        axis = np.cross(a, [1.0, 0.0, 0.0])
        if np.linalg.norm(axis) < 1e-6:
            axis = np.cross(a, [0.0, 1.0, 0.0])
        axis = axis / np.linalg.norm(axis)
        return np.array([axis[0], axis[1], axis[2], 0.0], dtype=np.float32)
    q = np.array([c[0], c[1], c[2], 1.0 + d])
    return (q / np.linalg.norm(q)).astype(np.float32)


class _Col:
    """Append-only column of numpy chunks."""

    def __init__(self, dtype, width=()):
        self.dtype = dtype
        self.width = tuple(width) if isinstance(width, tuple) else ((width,) if width else ())
        self.chunks: list[np.ndarray] = []
        self.n = 0
        self._cache = None

    def extend(self, arr) -> None:
        a = np.asarray(arr, dtype=self.dtype).reshape((-1,) + self.width)
        if a.shape[0] == 0:
            return
        self.chunks.append(a)
        self.n += a.shape[0]
        self._cache = None

    def array(self) -> np.ndarray:
        if self._cache is None:
            if not self.chunks:
                self._cache = np.empty((0,) + self.width, dtype=self.dtype)
            else:
                self._cache = np.concatenate(self.chunks, axis=0)
                self.chunks = [self._cache]
        return self._cache

    def set_rows(self, idx, values) -> None:
        a = self.array()
        a[idx] = values

    def __len__(self) -> int:
        return self.n


class ModelBuilder:
    """Builds a particle :class:`Model` (``builder.py:1104-1164`` defaults)."""

    @dataclass(kw_only=True)
    class ShapeConfig:
        """Per-shape contact material (``builder.py:452-490``)."""

        density: float = 1000.0
        ke: float = 2.5e3
        kd: float = 100.0
        kf: float = 1000.0
        ka: float = 0.0
        mu: float = 1.0
        restitution: float = 0.0
        margin: float = 0.0
        has_shape_collision: bool = True
        has_particle_collision: bool = True
        is_visible: bool = True

        def copy(self):
            return copy.copy(self)

    def __init__(self, up_axis="Z", gravity=None):
        self.up_axis = _axis_index(up_axis)
        up = np.zeros(3, dtype=np.float32)
        up[self.up_axis] = 1.0
        self.up_vector = tuple(float(v) for v in up)
        if gravity is None:
            gravity = -9.81
        if np.isscalar(gravity):
            self.gravity = tuple(float(gravity) * v for v in self.up_vector)
        else:
            self.gravity = tuple(float(v) for v in gravity)
        self.world_count = 0
        self.current_world = -1
        self.default_shape_cfg = ModelBuilder.ShapeConfig()
        self.default_particle_radius = 0.1
        self.particle_max_velocity = 1e5
        self.default_tri_ke = 100.0
        self.default_tri_ka = 100.0
        self.default_tri_kd = 10.0
        self.default_tri_drag = 0.0
        self.default_tri_lift = 0.0
        self.default_edge_ke = 100.0
        self.default_edge_kd = 0.0
        self.default_spring_ke = 100.0
        self.default_spring_kd = 0.0
        self.default_tet_k_mu = 1.0e3
        self.default_tet_k_lambda = 1.0e3
        self.default_tet_k_damp = 0.0
        # particles
        self._pq = _Col(np.float64, 3)
        self._pqd = _Col(np.float64, 3)
        self._pmass = _Col(np.float64)
        self._pradius = _Col(np.float64)
        self._pflags = _Col(np.int32)
        self._pworld = _Col(np.int32)
        # triangles
        self._tri = _Col(np.int32, 3)
        self._tri_pose = _Col(np.float64, (2, 2))
        self._tri_mat = _Col(np.float64, 5)
        self._tri_area = _Col(np.float64)
        # edges
        self._edge = _Col(np.int32, 4)
        self._edge_rest = _Col(np.float64)
        self._edge_len = _Col(np.float64)
        self._edge_bend = _Col(np.float64, 2)
        # springs
        self._spring = _Col(np.int32, 2)
        self._spring_rest = _Col(np.float64)
        self._spring_ke = _Col(np.float64)
        self._spring_kd = _Col(np.float64)
        self._spring_control = _Col(np.float64)
        # tets
        self._tet = _Col(np.int32, 4)
        self._tet_pose = _Col(np.float64, (3, 3))
        self._tet_mat = _Col(np.float64, 3)
        # shapes (static only on this path)
        self.shape_type: list[int] = []
        self.shape_body: list[int] = []
        self.shape_transform: list[tuple] = []
        self.shape_scale: list[tuple] = []
        self.shape_flags: list[int] = []
        self.shape_world: list[int] = []
        self.shape_margin: list[float] = []
        self.shape_material_ke: list[float] = []
        self.shape_material_kd: list[float] = []
        self.shape_material_mu: list[float] = []
        self.shape_label: list[str] = []
        # color groups as a lazy plan of (size, [(group, offset|None), ...]) items,
        # materialised on first read (merging W worlds one by one stays O(total))
        self._color_plan: list = []
        self._color_groups_cache: list[np.ndarray] | None = []
        # rigid bodies: poses only (this package does not integrate them; an external rigid solver or the
        # caller moves them and the particle solver reads their poses, one-way coupling)
        self.body_q: list[tuple] = []
        self.body_com: list[tuple] = []
        self.body_mass: list[float] = []
        self.body_inertia: list[np.ndarray] = []
        self.body_flags: list[int] = []
        self.body_world: list[int] = []
        self.body_label: list[str] = []
        self.joint_count = 0

    @property
    def body_count(self) -> int:
        return len(self.body_q)

    # ------------------------------------------------------------------ counts
    @property
    def particle_color_groups(self) -> list[np.ndarray]:
        if self._color_groups_cache is None:
            merged = []
            for _, chunks in self._color_plan:
                parts = [g if o is None else g + o for g, o in chunks]
                merged.append(parts[0] if len(parts) == 1 else np.concatenate(parts))
            self._color_groups_cache = merged
            self._color_plan = [(len(g), [(g, None)]) for g in merged]
        return self._color_groups_cache

    @particle_color_groups.setter
    def particle_color_groups(self, groups) -> None:
        groups = [np.asarray(g, dtype=np.int64) for g in groups]
        self._color_plan = [(len(g), [(g, None)]) for g in groups]
        self._color_groups_cache = groups

    @property
    def particle_count(self) -> int:
        return len(self._pq)

    @property
    def tri_count(self) -> int:
        return len(self._tri)

    @property
    def edge_count(self) -> int:
        return len(self._edge)

    @property
    def spring_count(self) -> int:
        return len(self._spring)

    @property
    def spring_indices(self) -> np.ndarray:
        return self._spring.array()

    @property
    def tet_count(self) -> int:
        return len(self._tet)

    @property
    def shape_count(self) -> int:
        return len(self.shape_type)

    @property
    def particle_q(self) -> np.ndarray:
        return self._pq.array()

    @property
    def particle_mass(self) -> np.ndarray:
        return self._pmass.array()

    @property
    def particle_flags(self) -> np.ndarray:
        return self._pflags.array()

    @property
    def tri_indices(self) -> np.ndarray:
        return self._tri.array()

    @property
    def tri_materials(self) -> np.ndarray:
        return self._tri_mat.array()

    @property
    def edge_indices(self) -> np.ndarray:
        return self._edge.array()

    @property
    def edge_bending_properties(self) -> np.ndarray:
        return self._edge_bend.array()

    @property
    def tet_indices(self) -> np.ndarray:
        return self._tet.array()

    # --------------------------------------------------------------- particles
    def add_particle(self, pos, vel, mass: float, radius: float | None = None, flags: int = ParticleFlags.ACTIVE) -> int:
        """``builder.py:8294-8337``."""
        self.add_particles([pos], [vel], [mass], None if radius is None else [radius], [int(flags)])
        return self.particle_count - 1

    def add_particles(self, pos, vel, mass, radius=None, flags=None) -> None:
        """``builder.py:8339-8406``."""
        pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
        n = pos.shape[0]
        vel = np.asarray(vel, dtype=np.float64).reshape(-1, 3)
        mass = np.asarray(mass, dtype=np.float64).reshape(-1)
        if vel.shape[0] != n:
            raise ValueError(f"vel length mismatch: expected {n} values to match pos, got {vel.shape[0]}")
        if mass.shape[0] != n:
            raise ValueError(f"mass length mismatch: expected {n} values to match pos, got {mass.shape[0]}")
        if radius is None:
            radius = np.full(n, self.default_particle_radius)
        radius = np.asarray(radius, dtype=np.float64).reshape(-1)
        if radius.shape[0] != n:
            raise ValueError(f"radius length mismatch: expected {n} values to match pos, got {radius.shape[0]}")
        if flags is None:
            flags = np.full(n, int(ParticleFlags.ACTIVE))
        flags = np.asarray(flags, dtype=np.int32).reshape(-1)
        if flags.shape[0] != n:
            raise ValueError(f"flags length mismatch: expected {n} values to match pos, got {flags.shape[0]}")
        self._pq.extend(pos)
        self._pqd.extend(vel)
        self._pmass.extend(mass)
        self._pradius.extend(radius)
        self._pflags.extend(flags)
        self._pworld.extend(np.full(n, self.current_world, dtype=np.int32))

    # --------------------------------------------------------------- triangles
    def add_triangles(self, i, j, k, *, tri_ke=None, tri_ka=None, tri_kd=None, tri_drag=None, tri_lift=None):
        """Triangles with float64 2-D rest bases (``builder.py:8543-8655``); returns areas."""
        q_ = self.particle_q
        i_ = np.asarray(i, dtype=np.int64).reshape(-1)
        j_ = np.asarray(j, dtype=np.int64).reshape(-1)
        k_ = np.asarray(k, dtype=np.int64).reshape(-1)
        p, q, r = q_[i_], q_[j_], q_[k_]
        qp = q - p
        rp = r - p

        def normalized(a):
            l = np.linalg.norm(a, axis=-1, keepdims=True)
            l[l == 0] = 1.0
            return a / l

        n = normalized(np.cross(qp, rp))
        e1 = normalized(qp)
        e2 = normalized(np.cross(n, e1))
        R = np.concatenate((e1[..., None], e2[..., None]), axis=-1)
        M = np.concatenate((qp[..., None], rp[..., None]), axis=-1)
        D = np.matmul(R.transpose(0, 2, 1), M)
        areas = np.linalg.det(D) / 2.0
        areas[areas < 0.0] = 0.0
        valid = (areas > 0.0).nonzero()[0]
        if len(valid) < len(areas):
            print("inverted or degenerate triangle elements")
        D[areas == 0.0] = np.eye(2)[None, ...]
        inv_D = np.linalg.inv(D)
        count = len(areas)

        def col(v, default):
            if v is None:
                return np.full(count, default, dtype=np.float64)
            return np.asarray(v, dtype=np.float64).reshape(-1)

        mats = np.stack(
            [
                col(tri_ke, self.default_tri_ke),
                col(tri_ka, self.default_tri_ka),
                col(tri_kd, self.default_tri_kd),
                col(tri_drag, self.default_tri_drag),
                col(tri_lift, self.default_tri_lift),
            ],
            axis=1,
        )
        self._tri.extend(np.stack([i_[valid], j_[valid], k_[valid]], axis=1))
        self._tri_pose.extend(inv_D[valid])
        self._tri_mat.extend(mats[valid])
        # the reference stores every area (including zeros) but only valid triangles
        self._tri_area.extend(areas[valid])
        return areas.tolist()

    def add_triangle(self, i: int, j: int, k: int, *, tri_ke=None, tri_ka=None, tri_kd=None, tri_drag=None,
                     tri_lift=None) -> float:
        """Single triangle with a float32 rest basis (``builder.py:8458-8540``)."""
        q_ = self.particle_q
        p = q_[i].astype(np.float32)
        q = q_[j].astype(np.float32)
        r = q_[k].astype(np.float32)
        qp = q - p
        rp = r - p

        def normalize(a):
            return (a / np.float32(np.sqrt(np.dot(a, a)))).astype(np.float32)

        n = normalize(np.cross(qp, rp).astype(np.float32))
        e1 = normalize(qp)
        e2 = normalize(np.cross(n, e1).astype(np.float32))
        R = np.array((e1, e2))
        M = np.array((qp, rp))
        D = R @ M.T
        area = np.linalg.det(D) / 2.0
        if area <= 0.0:
            print("inverted or degenerate triangle element")
            return 0.0
        self._tri.extend([[i, j, k]])
        self._tri_pose.extend(np.linalg.inv(D)[None])
        self._tri_mat.extend(
            [[
                self.default_tri_ke if tri_ke is None else tri_ke,
                self.default_tri_ka if tri_ka is None else tri_ka,
                self.default_tri_kd if tri_kd is None else tri_kd,
                self.default_tri_drag if tri_drag is None else tri_drag,
                self.default_tri_lift if tri_lift is None else tri_lift,
            ]]
        )
        self._tri_area.extend([area])
        return float(area)

    # ------------------------------------------------------------------ edges
    def add_edges(self, i, j, k, l, *, rest=None, edge_ke=None, edge_kd=None) -> None:
        """Bending hinges [o0, o1, v0, v1] with float64 rest angle / length (``builder.py:8798-8894``)."""
        i_ = np.asarray(i, dtype=np.int64).reshape(-1)
        j_ = np.asarray(j, dtype=np.int64).reshape(-1)
        k_ = np.asarray(k, dtype=np.int64).reshape(-1)
        l_ = np.asarray(l, dtype=np.int64).reshape(-1)
        pq = self.particle_q
        x3 = pq[k_]
        x4 = pq[l_]
        x4_minus_x3 = x4 - x3
        if rest is None:
            rest = np.zeros_like(i_, dtype=float)
            valid = (i_ != -1) & (j_ != -1)
            x1v = pq[i_[valid]]
            x2v = pq[j_[valid]]
            x3v = pq[k_[valid]]
            x4v = pq[l_[valid]]

            def normalized(a):
                ln = np.linalg.norm(a, axis=-1, keepdims=True)
                ln[ln == 0] = 1.0
                return a / ln

            n1 = normalized(np.cross(x3v - x1v, x4v - x1v))
            n2 = normalized(np.cross(x4v - x2v, x3v - x2v))
            e = normalized(x4v - x3v)
            cos_theta = np.clip((n1 * n2).sum(axis=-1), -1.0, 1.0)
            sin_theta = (np.cross(n1, n2) * e).sum(axis=-1)
            rest[valid] = np.arctan2(sin_theta, cos_theta)
        n = len(i_)
        ke = np.full(n, self.default_edge_ke) if edge_ke is None else np.asarray(edge_ke, dtype=np.float64).reshape(-1)
        kd = np.full(n, self.default_edge_kd) if edge_kd is None else np.asarray(edge_kd, dtype=np.float64).reshape(-1)
        self._edge.extend(np.stack([i_, j_, k_, l_], axis=1))
        self._edge_rest.extend(np.asarray(rest, dtype=np.float64))
        self._edge_len.extend(np.linalg.norm(x4_minus_x3, axis=1))
        self._edge_bend.extend(np.stack([ke, kd], axis=1))

    def add_edge(self, i: int, j: int, k: int, l: int, rest: float | None = None, edge_ke: float | None = None,
                 edge_kd: float | None = None) -> None:
        """Single bending hinge (``builder.py:8725-8796``)."""
        self.add_edges([i], [j], [k], [l], rest=None if rest is None else [rest],
                       edge_ke=None if edge_ke is None else [edge_ke], edge_kd=None if edge_kd is None else [edge_kd])

    def _add_soft_mesh_edges_from_triangles(self, start_tri: int, end_tri: int, *, edge_ke=None, edge_kd=None) -> range:
        """Register the triangle range's unique edges as bending hinges (``builder.py:8909-8941``)."""
        start = self.edge_count
        if end_tri > start_tri:
            edges, _, _ = compute_edge_adjacency(self.tri_indices[start_tri:end_tri])
            ne = edges.shape[0]
            if ne:
                self.add_edges(
                    edges[:, 0], edges[:, 1], edges[:, 2], edges[:, 3],
                    edge_ke=None if edge_ke is None else np.broadcast_to(np.float32(edge_ke), (ne,)),
                    edge_kd=None if edge_kd is None else np.broadcast_to(np.float32(edge_kd), (ne,)),
                )
        return range(start, self.edge_count)

    # ------------------------------------------------------------------- tets
    def add_spring(self, i: int, j: int, ke: float, kd: float, control: float = 0.0) -> None:
        """A spring between particles i and j, rest length from their current positions (``builder.py:8409-8446``)."""
        self.add_springs([i], [j], [ke], [kd], [control])

    def add_springs(self, i, j, ke, kd, control=None) -> None:
        """Vectorised :meth:`add_spring`: rest length sqrt(dot(p_i - p_j, p_i - p_j)) in float32."""
        i = np.asarray(i, dtype=np.int64).reshape(-1)
        j = np.asarray(j, dtype=np.int64).reshape(-1)
        n = len(i)
        if n == 0:
            return
        pos = self.particle_q
        p = pos[i].astype(np.float32)
        q = pos[j].astype(np.float32)
        delta = np.subtract(p, q)
        # np.dot(delta, delta) of the reference, summed in index order
        rest = np.sqrt((delta[:, 0] * delta[:, 0] + delta[:, 1] * delta[:, 1]) + delta[:, 2] * delta[:, 2])
        self._spring.extend(np.stack([i, j], axis=1))
        self._spring_rest.extend(rest)
        self._spring_ke.extend(np.broadcast_to(np.asarray(ke, dtype=np.float64), (n,)))
        self._spring_kd.extend(np.broadcast_to(np.asarray(kd, dtype=np.float64), (n,)))
        self._spring_control.extend(np.broadcast_to(np.asarray(0.0 if control is None else control,
                                                               dtype=np.float64), (n,)))

    def add_tetrahedron(self, i: int, j: int, k: int, l: int, k_mu: float = 1.0e3, k_lambda: float = 1.0e3,
                        k_damp: float = 0.0) -> float:
        """``builder.py:8657-8722``."""
        return float(self.add_tetrahedra([i], [j], [k], [l], k_mu, k_lambda, k_damp)[0])

    def add_tetrahedra(self, i, j, k, l, k_mu, k_lambda, k_damp) -> np.ndarray:
        """Vectorised :meth:`add_tetrahedron`: float64 Dm^-1 per element, inverted ones dropped."""
        pq = self.particle_q
        i_, j_, k_, l_ = (np.asarray(a, dtype=np.int64).reshape(-1) for a in (i, j, k, l))
        p, q, r, s = pq[i_], pq[j_], pq[k_], pq[l_]
        Dm = np.stack([q - p, r - p, s - p], axis=2)  # columns
        vol = np.linalg.det(Dm) / 6.0
        ok = vol > 0.0
        if not np.all(ok):
            print("inverted tetrahedral element")
        n = len(i_)
        mats = np.stack(
            [np.broadcast_to(np.float64(k_mu), (n,)), np.broadcast_to(np.float64(k_lambda), (n,)),
             np.broadcast_to(np.float64(k_damp), (n,))], axis=1)
        if np.any(ok):
            self._tet.extend(np.stack([i_[ok], j_[ok], k_[ok], l_[ok]], axis=1))
            self._tet_pose.extend(np.linalg.inv(Dm[ok]))
            self._tet_mat.extend(mats[ok])
        return vol

    # ------------------------------------------------------------ cloth helpers
    def add_cloth_mesh(self, *, pos, rot, scale: float, vel, vertices, indices, density: float, tri_ke=None,
                       tri_ka=None, tri_kd=None, tri_drag=None, tri_lift=None, edge_ke=None, edge_kd=None,
                       particle_radius=None, add_springs: bool = False, spring_ke: float | None = None,
                       spring_kd: float | None = None) -> None:
        """``builder.py:9071-9218``: particles, triangles, lumped mass density*area/3, bending edges
        and, with ``add_springs``, a spring on every mesh edge and across every hinge."""
        tri_ke = self.default_tri_ke if tri_ke is None else tri_ke
        tri_ka = self.default_tri_ka if tri_ka is None else tri_ka
        tri_kd = self.default_tri_kd if tri_kd is None else tri_kd
        tri_drag = self.default_tri_drag if tri_drag is None else tri_drag
        tri_lift = self.default_tri_lift if tri_lift is None else tri_lift
        edge_ke = self.default_edge_ke if edge_ke is None else edge_ke
        edge_kd = self.default_edge_kd if edge_kd is None else edge_kd
        particle_radius = self.default_particle_radius if particle_radius is None else particle_radius
        verts = np.asarray(vertices, dtype=np.float32).reshape(-1, 3) * np.float32(scale)
        rot_mat = quat_to_matrix(rot)
        verts_3d = (np.dot(verts, rot_mat.T) + np.asarray(pos, dtype=np.float32)).astype(np.float32)
        num_verts = verts_3d.shape[0]
        ind = np.asarray(indices, dtype=np.int64).reshape(-1, 3)
        num_tris = ind.shape[0]
        start_vertex = self.particle_count
        start_tri = self.tri_count
        self.add_particles(verts_3d.astype(np.float64), np.tile(np.asarray(vel, dtype=np.float32), (num_verts, 1)),
                           np.zeros(num_verts), np.full(num_verts, particle_radius))
        inds = start_vertex + ind
        areas = self.add_triangles(inds[:, 0], inds[:, 1], inds[:, 2], tri_ke=np.full(num_tris, tri_ke),
                                   tri_ka=np.full(num_tris, tri_ka), tri_kd=np.full(num_tris, tri_kd),
                                   tri_drag=np.full(num_tris, tri_drag), tri_lift=np.full(num_tris, tri_lift))
        # lumped mass: each triangle adds density*area/3 to its three vertices, in triangle order
        mass = self._pmass.array()
        contrib = density * np.asarray(areas, dtype=np.float64) / 3.0
        # np.add.at accumulates unbuffered in index order: per vertex, triangle order, as the reference loop
        np.add.at(mass, inds.reshape(-1), np.repeat(contrib, 3))
        self._pmass._cache = mass
        end_tri = self.tri_count
        edge_range = self._add_soft_mesh_edges_from_triangles(start_tri, end_tri, edge_ke=edge_ke, edge_kd=edge_kd)
        if add_springs:
            self._add_cloth_springs(edge_range, self.default_spring_ke if spring_ke is None else spring_ke,
                                    self.default_spring_kd if spring_kd is None else spring_kd)

    def _add_cloth_springs(self, edge_range: range, spring_ke: float, spring_kd: float) -> None:
        """Springs of ``builder.py:9204-9218``: every hinge adds its edge (k, l), the four opposite-to-edge
        pairs and (i, j), as sorted pairs in a Python set, then the set is walked in its own order
        (CPython's iteration order for the same insertions, so the spring numbering matches)."""
        rows = self.edge_indices[edge_range.start:edge_range.stop]
        pairs = set()
        for i, j, k, l in rows.tolist():
            pairs.add((min(k, l), max(k, l)))
            if i != -1:
                pairs.add((min(i, k), max(i, k)))
                pairs.add((min(i, l), max(i, l)))
            if j != -1:
                pairs.add((min(j, k), max(j, k)))
                pairs.add((min(j, l), max(j, l)))
            if i != -1 and j != -1:
                pairs.add((min(i, j), max(i, j)))
        if pairs:
            ij = np.asarray(list(pairs), dtype=np.int64)
            self.add_springs(ij[:, 0], ij[:, 1], spring_ke, spring_kd, 0.0)

    def add_cloth_grid(self, *, pos, rot, vel, dim_x: int, dim_y: int, cell_x: float, cell_y: float, mass: float,
                       reverse_winding: bool = False, fix_left: bool = False, fix_right: bool = False,
                       fix_top: bool = False, fix_bottom: bool = False, tri_ke=None, tri_ka=None, tri_kd=None,
                       tri_drag=None, tri_lift=None, edge_ke=None, edge_kd=None, particle_radius=None,
                       add_springs: bool = False, spring_ke: float | None = None,
                       spring_kd: float | None = None) -> None:
        """``builder.py:8944-9068``; per-particle mass is then overwritten with ``mass``."""
        verts, indices = cloth_grid_mesh(dim_x, dim_y, cell_x, cell_y, reverse_winding)
        start_vertex = self.particle_count
        total_mass = mass * (dim_x + 1) * (dim_x + 1)
        total_area = cell_x * cell_y * dim_x * dim_y
        density = total_mass / total_area
        self.add_cloth_mesh(pos=pos, rot=rot, scale=1.0, vel=vel, vertices=verts, indices=indices, density=density,
                            tri_ke=tri_ke, tri_ka=tri_ka, tri_kd=tri_kd, tri_drag=tri_drag, tri_lift=tri_lift,
                            edge_ke=edge_ke, edge_kd=edge_kd, particle_radius=particle_radius,
                            add_springs=add_springs, spring_ke=spring_ke, spring_kd=spring_kd)
        ys, xs = np.meshgrid(np.arange(dim_y + 1), np.arange(dim_x + 1), indexing="ij")
        xs, ys = xs.reshape(-1), ys.reshape(-1)
        fixed = (xs == 0) & fix_left | (xs == dim_x) & fix_right | (ys == 0) & fix_bottom | (ys == dim_y) & fix_top
        idx = start_vertex + np.arange(len(xs))
        m = self._pmass.array()
        f = self._pflags.array()
        m[idx] = np.where(fixed, 0.0, mass)
        f[idx] = np.where(fixed, int(ParticleFlags.ACTIVE) & ~int(ParticleFlags.ACTIVE), int(ParticleFlags.ACTIVE))

    def add_soft_grid(self, *, pos, rot, vel, dim_x: int, dim_y: int, dim_z: int, cell_x: float, cell_y: float,
                      cell_z: float, density: float, k_mu: float, k_lambda: float, k_damp: float,
                      fix_left: bool = False, fix_right: bool = False, fix_top: bool = False,
                      fix_bottom: bool = False, tri_ke: float = 0.0, tri_ka: float = 0.0, tri_kd: float = 0.0,
                      tri_drag: float = 0.0, tri_lift: float = 0.0, add_surface_mesh_edges: bool = True,
                      edge_ke: float = 0.0, edge_kd: float = 0.0, particle_radius=None) -> None:
        """5-tet-per-hex soft grid (``builder.py:9325-9494``)."""
        start_vertex = self.particle_count
        mass = cell_x * cell_y * cell_z * density
        zs, ys, xs = np.meshgrid(np.arange(dim_z + 1), np.arange(dim_y + 1), np.arange(dim_x + 1), indexing="ij")
        xs, ys, zs = xs.reshape(-1), ys.reshape(-1), zs.reshape(-1)
        local = np.stack([xs * np.float32(cell_x), ys * np.float32(cell_y), zs * np.float32(cell_z)], axis=1)
        local = local.astype(np.float32)
        p = (quat_rotate(rot, local) + np.asarray(pos, dtype=np.float32)).astype(np.float32)
        m = np.full(len(xs), mass, dtype=np.float64)
        m[(xs == 0) & fix_left] = 0.0
        m[(xs == dim_x) & fix_right] = 0.0
        m[(ys == dim_y) & fix_top] = 0.0
        m[(ys == 0) & fix_bottom] = 0.0
        self.add_particles(p.astype(np.float64), np.tile(np.asarray(vel, dtype=np.float32), (len(xs), 1)), m,
                           None if particle_radius is None else np.full(len(xs), particle_radius))

        def gi(x, y, z):
            return (dim_x + 1) * (dim_y + 1) * z + (dim_x + 1) * y + x

        tets = []
        for z in range(dim_z):
            for y in range(dim_y):
                for x in range(dim_x):
                    v0 = gi(x, y, z) + start_vertex
                    v1 = gi(x + 1, y, z) + start_vertex
                    v2 = gi(x + 1, y, z + 1) + start_vertex
                    v3 = gi(x, y, z + 1) + start_vertex
                    v4 = gi(x, y + 1, z) + start_vertex
                    v5 = gi(x + 1, y + 1, z) + start_vertex
                    v6 = gi(x + 1, y + 1, z + 1) + start_vertex
                    v7 = gi(x, y + 1, z + 1) + start_vertex
                    if (x & 1) ^ (y & 1) ^ (z & 1):
                        tets += [(v0, v1, v4, v3), (v2, v3, v6, v1), (v5, v4, v1, v6), (v7, v6, v3, v4), (v4, v1, v6, v3)]
                    else:
                        tets += [(v1, v2, v5, v0), (v3, v0, v7, v2), (v4, v7, v0, v5), (v6, v5, v2, v7), (v5, v2, v7, v0)]
        tets_np = np.asarray(tets, dtype=np.int64)
        self.add_tetrahedra(tets_np[:, 0], tets_np[:, 1], tets_np[:, 2], tets_np[:, 3], k_mu, k_lambda, k_damp)
        # open faces: faces seen once, keyed by sorted vertices, kept in first-insertion order
        faces: dict = {}
        for i, j, k, l in tets:
            for f in ((i, k, j), (j, k, l), (i, j, l), (i, l, k)):
                key = tuple(sorted(f))
                if key in faces:
                    del faces[key]
                else:
                    faces[key] = f
        start_tri = self.tri_count
        for v in faces.values():
            self.add_triangle(v[0], v[1], v[2], tri_ke=tri_ke, tri_ka=tri_ka, tri_kd=tri_kd, tri_drag=tri_drag,
                              tri_lift=tri_lift)
        end_tri = self.tri_count
        if add_surface_mesh_edges and end_tri > start_tri:
            self._add_soft_mesh_edges_from_triangles(start_tri, end_tri, edge_ke=edge_ke, edge_kd=edge_kd)

    def add_soft_mesh(self, *, pos, rot, scale: float, vel, vertices, indices, density: float | None = None,
                      k_mu=None, k_lambda=None, k_damp=None, tri_ke: float = 0.0, tri_ka: float = 0.0,
                      tri_kd: float = 0.0, tri_drag: float = 0.0, tri_lift: float = 0.0,
                      add_surface_mesh_edges: bool = True, edge_ke: float = 0.0, edge_kd: float = 0.0,
                      particle_radius=None) -> None:
        """Tetrahedral soft body from vertices / tet indices (``builder.py:9497-9700``).

        Mass is lumped as density * volume / 4 per tet corner; the surface
        triangles (faces used by exactly one tet, in first-seen order) get the
        given (default zero) triangle material and optional bending edges.
        """
        density = 1000.0 if density is None else density
        k_mu = self.default_tet_k_mu if k_mu is None else k_mu
        k_lambda = self.default_tet_k_lambda if k_lambda is None else k_lambda
        k_damp = self.default_tet_k_damp if k_damp is None else k_damp
        ind = np.asarray(indices, dtype=np.int64).reshape(-1, 4)
        verts = np.asarray(vertices, dtype=np.float32).reshape(-1, 3) * np.float32(scale)
        p = (quat_rotate(rot, verts) + np.asarray(pos, dtype=np.float32)).astype(np.float32)
        start = self.particle_count
        n = p.shape[0]
        self.add_particles(p.astype(np.float64), np.tile(np.asarray(vel, dtype=np.float32), (n, 1)), np.zeros(n),
                           None if particle_radius is None else np.full(n, particle_radius))
        t = ind + start
        vols = self.add_tetrahedra(t[:, 0], t[:, 1], t[:, 2], t[:, 3], np.float32(k_mu), np.float32(k_lambda),
                                   np.float32(k_damp))
        mass = self._pmass.array()
        ok = vols > 0.0
        contrib = density * vols[ok] / 4.0
        np.add.at(mass, t[ok].reshape(-1), np.repeat(contrib, 4))
        faces: dict = {}
        for i, j, k, l in ind.tolist():
            for f in ((i, k, j), (j, k, l), (i, j, l), (i, l, k)):
                key = tuple(sorted(f))
                if key in faces:
                    del faces[key]
                else:
                    faces[key] = f
        start_tri = self.tri_count
        for f in faces.values():
            self.add_triangle(start + f[0], start + f[1], start + f[2], tri_ke=tri_ke, tri_ka=tri_ka, tri_kd=tri_kd,
                              tri_drag=tri_drag, tri_lift=tri_lift)
        end_tri = self.tri_count
        if add_surface_mesh_edges and end_tri > start_tri:
            self._add_soft_mesh_edges_from_triangles(start_tri, end_tri, edge_ke=edge_ke, edge_kd=edge_kd)

    # ----------------------------------------------------------------- bodies
    def add_body(self, *, xform=None, com=None, inertia=None, mass: float = 0.0, label: str | None = None,
                 lock_inertia: bool = False, is_kinematic: bool = False, custom_attributes: dict | None = None) -> int:
        """A rigid body whose pose the caller (or an external rigid solver) drives (``builder.py:4430-4500``).

        Shapes attached to it (``add_shape_*(body=...)``) follow ``State.body_q``; ``SolverVBD`` with
        ``integrate_with_external_rigid_solver=True`` reads the poses for particle contact and never moves
        the body.  The joint and articulation the reference adds with a free body are not built here."""
        if xform is None:
            xform = (np.zeros(3, np.float32), np.array([0.0, 0.0, 0.0, 1.0], np.float32))
        p, q = xform
        self.body_q.append(tuple(float(v) for v in np.asarray(p, np.float32)) +
                           tuple(float(v) for v in np.asarray(q, np.float32)))
        self.body_com.append(tuple(float(v) for v in np.asarray(com if com is not None else (0.0, 0.0, 0.0),
                                                                np.float32)))
        self.body_mass.append(float(mass))
        self.body_inertia.append(np.asarray(inertia if inertia is not None else np.zeros((3, 3)), np.float32))
        self.body_flags.append(1 if is_kinematic else 0)
        self.body_world.append(self.current_world)
        self.body_label.append(label or f"body_{len(self.body_q) - 1}")
        return len(self.body_q) - 1

    # ----------------------------------------------------------------- shapes
    def _add_shape(self, *, type: int, xform, cfg, scale, body: int, label: str | None) -> int:
        if body < -1 or body >= self.body_count:
            raise ValueError(f"body {body} does not exist (body_count {self.body_count})")
        cfg = self.default_shape_cfg if cfg is None else cfg
        p, q = xform
        self.shape_type.append(int(type))
        self.shape_body.append(int(body))
        self.shape_transform.append(tuple(float(v) for v in np.asarray(p, np.float32)) +
                                    tuple(float(v) for v in np.asarray(q, np.float32)))
        self.shape_scale.append(tuple(float(v) for v in np.asarray(scale, np.float32)))
        flags = int(ShapeFlags.VISIBLE) if cfg.is_visible else 0
        if cfg.has_shape_collision:
            flags |= int(ShapeFlags.COLLIDE_SHAPES)
        if cfg.has_particle_collision:
            flags |= int(ShapeFlags.COLLIDE_PARTICLES)
        self.shape_flags.append(flags)
        self.shape_world.append(self.current_world)
        self.shape_margin.append(float(cfg.margin))
        self.shape_material_ke.append(float(cfg.ke))
        self.shape_material_kd.append(float(cfg.kd))
        self.shape_material_mu.append(float(cfg.mu))
        self.shape_label.append(label or f"shape_{len(self.shape_type) - 1}")
        return len(self.shape_type) - 1

    def add_shape_plane(self, plane=(0.0, 0.0, 1.0, 0.0), *, xform=None, width: float = 10.0, length: float = 10.0,
                        body: int = -1, cfg=None, label: str | None = None) -> int:
        """``builder.py:6599-6664``; width = length = 0 makes the plane infinite."""
        if xform is None:
            normal = np.array(plane[:3], dtype=np.float64)
            norm = np.linalg.norm(normal)
            normal /= norm
            d_normalized = plane[3] / norm
            pos = -d_normalized * normal
            rot = quat_between_vectors((0.0, 0.0, 1.0), normal)
            xform = (pos.astype(np.float32), rot)
        return self._add_shape(type=GeoType.PLANE, xform=xform, cfg=cfg, scale=(width, length, 0.0), body=body,
                               label=label)

    def add_ground_plane(self, *, height: float = 0.0, cfg=None, label: str | None = None) -> int:
        """``builder.py:6666-6692``."""
        return self.add_shape_plane(plane=(*self.up_vector, -height), width=0.0, length=0.0, cfg=cfg,
                                    label=label or "ground_plane")

    def add_shape_sphere(self, body: int, *, xform=None, radius: float = 1.0, cfg=None, label: str | None = None) -> int:
        """``builder.py:6695-6730``."""
        if xform is None:
            xform = (np.zeros(3, np.float32), np.array([0.0, 0.0, 0.0, 1.0], np.float32))
        return self._add_shape(type=GeoType.SPHERE, xform=xform, cfg=cfg, scale=(radius, 0.0, 0.0), body=body,
                               label=label)

    # --------------------------------------------------------------- coloring
    def set_coloring(self, particle_color_groups) -> None:
        """``builder.py:10130-10142``."""
        self.particle_color_groups = [np.asarray(g, dtype=np.int64) for g in particle_color_groups]

    def color(self, include_bending: bool = False, balance_colors: bool = True, target_max_min_color_ratio: float = 1.1,
              coloring_algorithm=gc.ColoringAlgorithm.MCS) -> None:
        """Graph + coloring (``builder.py:10144-10236``)."""
        if self.particle_count == 0:
            return
        tri_mat = self.tri_materials
        tet_mat = self._tet_mat.array()
        bend_idx = bend_mask = None
        if include_bending and self.edge_count:
            bend_idx = self.edge_indices
            props = self.edge_bending_properties
            bend_mask = (props[:, 0] != 0.0) | (props[:, 1] != 0.0)
        edges = gc.construct_particle_graph(
            self.tri_indices if self.tri_count else None,
            tri_mat[:, 0] * tri_mat[:, 1] if len(tri_mat) else None,
            bend_idx,
            bend_mask,
            self.tet_indices if self.tet_count else None,
            tet_mat[:, 0] * tet_mat[:, 1] if len(tet_mat) else None,
        )
        if len(edges) > 0:
            self.particle_color_groups = gc.color_graph(self.particle_count, edges, balance_colors,
                                                        target_max_min_color_ratio, coloring_algorithm)
        else:
            self.particle_color_groups = [np.arange(self.particle_count, dtype=np.int64)]

    # ------------------------------------------------------------------ worlds
    def add_builder(self, builder: "ModelBuilder", xform=None, world: int | None = None) -> None:
        """Append another builder's particles/elements/shapes (``builder.py:4265``).

        Color groups are merged with the independent-coloring plan of
        ``builder.py:2924-2939``.  ``xform`` (p, q) rotates and translates the
        appended particles and shapes.
        """
        w = self.current_world if world is None else world
        off = self.particle_count
        pos = builder.particle_q
        if xform is not None:
            p, q = xform
            pos = (quat_rotate(q, pos.astype(np.float32)) + np.asarray(p, np.float32)).astype(np.float64)
        self._pq.extend(pos)
        self._pqd.extend(builder._pqd.array())
        self._pmass.extend(builder._pmass.array())
        self._pradius.extend(builder._pradius.array())
        self._pflags.extend(builder._pflags.array())
        self._pworld.extend(np.full(builder.particle_count, w, dtype=np.int32))
        if builder.tri_count:
            self._tri.extend(builder.tri_indices + off)
            self._tri_pose.extend(builder._tri_pose.array())
            self._tri_mat.extend(builder.tri_materials)
            self._tri_area.extend(builder._tri_area.array())
        if builder.edge_count:
            e = builder.edge_indices.copy()
            e[e >= 0] += off
            self._edge.extend(e)
            self._edge_rest.extend(builder._edge_rest.array())
            self._edge_len.extend(builder._edge_len.array())
            self._edge_bend.extend(builder.edge_bending_properties)
        if builder.spring_count:
            self._spring.extend(builder.spring_indices + off)
            self._spring_rest.extend(builder._spring_rest.array())
            self._spring_ke.extend(builder._spring_ke.array())
            self._spring_kd.extend(builder._spring_kd.array())
            self._spring_control.extend(builder._spring_control.array())
        if builder.tet_count:
            self._tet.extend(builder.tet_indices + off)
            self._tet_pose.extend(builder._tet_pose.array())
            self._tet_mat.extend(builder._tet_mat.array())
        body_off = self.body_count
        for b in range(builder.body_count):
            bq = builder.body_q[b]
            if xform is not None:
                p, q = xform
                bp = quat_rotate(q, np.asarray(bq[:3], np.float32)) + np.asarray(p, np.float32)
                br = _quat_mul(np.asarray(q, np.float32), np.asarray(bq[3:], np.float32))
                bq = tuple(float(v) for v in bp) + tuple(float(v) for v in br)
            self.body_q.append(bq)
            self.body_com.append(builder.body_com[b])
            self.body_mass.append(builder.body_mass[b])
            self.body_inertia.append(builder.body_inertia[b])
            self.body_flags.append(builder.body_flags[b])
            self.body_world.append(w)
            self.body_label.append(builder.body_label[b])
        for s in range(builder.shape_count):
            tf = builder.shape_transform[s]
            body = builder.shape_body[s]
            if xform is not None and body < 0:  # a body's shapes stay in its frame
                p, q = xform
                sp = quat_rotate(q, np.asarray(tf[:3], np.float32)) + np.asarray(p, np.float32)
                sq = _quat_mul(np.asarray(q, np.float32), np.asarray(tf[3:], np.float32))
                tf = tuple(float(v) for v in sp) + tuple(float(v) for v in sq)
            self.shape_type.append(builder.shape_type[s])
            self.shape_body.append(body + body_off if body >= 0 else -1)
            self.shape_transform.append(tf)
            self.shape_scale.append(builder.shape_scale[s])
            self.shape_flags.append(builder.shape_flags[s])
            self.shape_world.append(w)
            self.shape_margin.append(builder.shape_margin[s])
            self.shape_material_ke.append(builder.shape_material_ke[s])
            self.shape_material_kd.append(builder.shape_material_kd[s])
            self.shape_material_mu.append(builder.shape_material_mu[s])
            self.shape_label.append(builder.shape_label[s])
        src_groups = builder.particle_color_groups
        if src_groups:
            self._color_plan = gc.combine_independent_coloring_plan(
                self._color_plan, [(len(g), [(np.asarray(g, dtype=np.int64), off)]) for g in src_groups]
            )
            self._color_groups_cache = None

    def add_world(self, builder: "ModelBuilder", xform=None) -> int:
        """Add ``builder`` as a new world (``builder.py:4016-4065``); returns the world index."""
        w = self.world_count
        self.world_count += 1
        self.add_builder(builder, xform=xform, world=w)
        return w

    # ---------------------------------------------------------------- finalize
    def finalize(self, device="cuda:0", requires_grad: bool = False) -> Model:
        """Build the :class:`Model` (``builder.py:10962``; particles :11033-11070, elements :11799-11840)."""
        dev = torch.device(device)
        m = Model(dev)
        m.requires_grad = requires_grad
        m.world_count = self.world_count
        m.up_axis = self.up_axis
        n = self.particle_count
        m.particle_count = n

        def f32(a, shape=None):
            arr = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
            if shape is not None:
                arr = arr.reshape(shape)
            return torch.as_tensor(arr, device=dev)

        def i32(a, shape=None):
            arr = np.ascontiguousarray(np.asarray(a, dtype=np.int32))
            if shape is not None:
                arr = arr.reshape(shape)
            return torch.as_tensor(arr, device=dev)

        ms = self._pmass.array().astype(np.float32)
        inv = np.divide(np.float32(1.0), ms, out=np.zeros_like(ms), where=ms != 0.0)
        m.particle_q = f32(self.particle_q, (-1, 3))
        m.particle_qd = f32(self._pqd.array(), (-1, 3))
        m.particle_mass = f32(ms)
        m.particle_inv_mass = f32(inv)
        m.particle_radius = f32(self._pradius.array())
        m.particle_flags = i32(self._pflags.array())
        m.particle_world = i32(self._pworld.array())
        m.particle_max_radius = float(np.max(self._pradius.array())) if n else 0.0
        m.particle_max_velocity = float(self.particle_max_velocity)
        colors = np.full(n, -1, dtype=np.int32)
        for c, g in enumerate(self.particle_color_groups):
            colors[np.asarray(g, dtype=np.int64)] = c
        m.particle_colors = i32(colors)
        m.particle_color_groups = [i32(g) for g in self.particle_color_groups]
        m.tri_count = self.tri_count
        m.tri_indices = i32(self.tri_indices, (-1, 3))
        m.tri_poses = f32(self._tri_pose.array(), (-1, 2, 2))
        m.tri_activations = torch.zeros(self.tri_count, dtype=torch.float32, device=dev)
        m.tri_materials = f32(self.tri_materials, (-1, 5))
        m.tri_areas = f32(self._tri_area.array())
        m.edge_count = self.edge_count
        m.edge_indices = i32(self.edge_indices, (-1, 4))
        m.edge_rest_angle = f32(self._edge_rest.array())
        m.edge_rest_length = f32(self._edge_len.array())
        m.edge_bending_properties = f32(self.edge_bending_properties, (-1, 2))
        m.tet_count = self.tet_count
        m.tet_indices = i32(self.tet_indices, (-1, 4))
        m.tet_poses = f32(self._tet_pose.array(), (-1, 3, 3))
        m.tet_activations = torch.zeros(self.tet_count, dtype=torch.float32, device=dev)
        m.tet_materials = f32(self._tet_mat.array(), (-1, 3))
        m.spring_count = self.spring_count
        m.spring_indices = i32(self.spring_indices.reshape(-1))
        m.spring_rest_length = f32(self._spring_rest.array())
        m.spring_stiffness = f32(self._spring_ke.array())
        m.spring_damping = f32(self._spring_kd.array())
        m.spring_control = f32(self._spring_control.array())
        adj = MeshAdjacency(
            tri_indices=self.tri_indices,
            edge_indices=self.edge_indices if self.edge_count else np.empty((0, 4), dtype=np.int32),
            spring_indices=self.spring_indices.reshape(-1) if self.spring_count else None,
            tet_indices=self.tet_indices if self.tet_count else None,
        )
        adj.init_vertex_adjacency(n)
        m.soft_mesh_adjacency = adj
        m.soft_mesh_adjacency_device = adj.to(dev)
        nw = max(self.world_count, 1)
        m.gravity = f32(np.tile(np.asarray(self.gravity, dtype=np.float32), (nw, 1)), (-1, 3))
        m.shape_count = self.shape_count
        m.shape_type = i32(self.shape_type)
        m.shape_body = i32(self.shape_body)
        m.shape_flags = i32(self.shape_flags)
        m.shape_world = i32(self.shape_world)
        m.shape_transform = f32(self.shape_transform, (-1, 7))
        m.shape_scale = f32(self.shape_scale, (-1, 3))
        m.shape_margin = f32(self.shape_margin)
        m.shape_material_ke = f32(self.shape_material_ke)
        m.shape_material_kd = f32(self.shape_material_kd)
        m.shape_material_mu = f32(self.shape_material_mu)
        m.shape_label = list(self.shape_label)
        m.body_count = self.body_count
        nb = self.body_count
        m.body_q = f32(self.body_q if nb else np.zeros((0, 7)), (-1, 7))
        m.body_qd = torch.zeros((nb, 6), dtype=torch.float32, device=dev)
        m.body_com = f32(self.body_com if nb else np.zeros((0, 3)), (-1, 3))
        bm = np.asarray(self.body_mass, dtype=np.float32)
        m.body_mass = f32(bm)
        m.body_inv_mass = f32(np.divide(np.float32(1.0), bm, out=np.zeros_like(bm), where=bm != 0.0))
        m.body_inertia = f32(np.stack(self.body_inertia) if nb else np.zeros((0, 3, 3)), (-1, 3, 3))
        m.body_flags = i32(self.body_flags)
        m.body_world = i32(self.body_world)
        m.body_label = list(self.body_label)
        return m


def _quat_mul(a, b) -> np.ndarray:
    ax, ay, az, aw = (np.float32(v) for v in a)
    bx, by, bz, bw = (np.float32(v) for v in b)
    return np.array(
        [
            aw * bx + bw * ax + ay * bz - by * az,
            aw * by + bw * ay + az * bx - bz * ax,
            aw * bz + bw * az + ax * by - bx * ay,
            aw * bw - ax * bx - ay * by - az * bz,
        ],
        dtype=np.float32,
    )


def cloth_grid_mesh(dim_x: int, dim_y: int, cell_x: float, cell_y: float, reverse_winding: bool = False):
    """Grid vertices and triangles exactly as ``add_cloth_grid`` builds them (``builder.py:8999-9018``).

    Vertex (x, y) is at (x*cell_x, y*cell_y, 0) rounded to float32; each quad
    (v0 = (x-1, y-1), v1 = (x, y-1), v2 = (x, y), v3 = (x-1, y)) gives
    [v0, v1, v3], [v1, v2, v3] (or [v0, v1, v2], [v0, v2, v3] reversed).
    """
    ys, xs = np.meshgrid(np.arange(dim_y + 1), np.arange(dim_x + 1), indexing="ij")
    verts = np.stack(
        [
            (xs.reshape(-1) * cell_x).astype(np.float32),
            (ys.reshape(-1) * cell_y).astype(np.float32),
            np.zeros(xs.size, dtype=np.float32),
        ],
        axis=1,
    )
    qy, qx = np.meshgrid(np.arange(1, dim_y + 1), np.arange(1, dim_x + 1), indexing="ij")
    qx, qy = qx.reshape(-1), qy.reshape(-1)
    w = dim_x + 1
    v0 = (qy - 1) * w + (qx - 1)
    v1 = (qy - 1) * w + qx
    v2 = qy * w + qx
    v3 = qy * w + (qx - 1)
    if reverse_winding:
        tris = np.stack([np.stack([v0, v1, v2], 1), np.stack([v0, v2, v3], 1)], axis=1)
    else:
        tris = np.stack([np.stack([v0, v1, v3], 1), np.stack([v1, v2, v3], 1)], axis=1)
    return verts, tris.reshape(-1).astype(np.int64)


__all__ = ["ModelBuilder", "cloth_grid_mesh", "quat_between_vectors", "quat_from_axis_angle", "quat_rotate",
           "quat_to_matrix"]
