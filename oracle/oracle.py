"""ctypes front end of the CPU oracle (``oracle/libvbd_oracle.so``) -- TEST INFRASTRUCTURE.

Only tests, the smoke check and the benchmark's CPU-baseline leg use this
module; the product path never imports it.  ``OracleModel.from_model`` copies a
:class:`newton_amd.Model` (on any device) into host arrays laid out as the
reference's Model, and the ``step`` / ``simulate`` / element functions run the
C restatement in FP32 ("f32") or FP64 ("f64").
"""

from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libvbd_oracle.so"

i32 = ctypes.c_int32
f32 = ctypes.c_float
vp = ctypes.c_void_p


class _OracleModel(ctypes.Structure):
    _fields_ = [
        ("particle_count", i32),
        ("tri_count", i32),
        ("edge_count", i32),
        ("tet_count", i32),
        ("particle_mass", vp),
        ("particle_inv_mass", vp),
        ("particle_radius", vp),
        ("particle_flags", vp),
        ("particle_colors", vp),
        ("particle_world", vp),
        ("tri_indices", vp),
        ("tri_poses", vp),
        ("tri_materials", vp),
        ("tri_areas", vp),
        ("edge_indices", vp),
        ("edge_rest_angle", vp),
        ("edge_rest_length", vp),
        ("edge_bending_properties", vp),
        ("tet_indices", vp),
        ("tet_poses", vp),
        ("tet_materials", vp),
        ("v_adj_tris", vp),
        ("v_adj_tris_offsets", vp),
        ("v_adj_edges", vp),
        ("v_adj_edges_offsets", vp),
        ("v_adj_tets", vp),
        ("v_adj_tets_offsets", vp),
        ("color_count", i32),
        ("color_offsets", vp),
        ("color_particles", vp),
        ("gravity", f32 * 3),
        ("shape_count", i32),
        ("shape_type", vp),
        ("shape_body", vp),
        ("shape_flags", vp),
        ("shape_world", vp),
        ("shape_transform", vp),
        ("shape_scale", vp),
        ("shape_margin", vp),
        ("shape_material_ke", vp),
        ("shape_material_kd", vp),
        ("shape_material_mu", vp),
        ("world_count", i32),
        ("soft_contact_ke", f32),
        ("soft_contact_kd", f32),
        ("soft_contact_mu", f32),
        ("spring_count", i32),
        ("spring_indices", vp),
        ("spring_rest_length", vp),
        ("spring_stiffness", vp),
        ("spring_damping", vp),
    ]


class _OracleParams(ctypes.Structure):
    _fields_ = [
        ("iterations", i32),
        ("friction_epsilon", f32),
        ("avbd_beta", f32),
        ("k_start", f32),
        ("use_libm_atan2", i32),
        ("fp64_ill_conditioned", i32),
        ("num_threads", i32),
        ("self_contact", vp),
        ("residual_stop", ctypes.c_double),
        ("body_q", vp),
        ("body_q_prev", vp),
        ("body_count", i32),
    ]


class _OracleSelf(ctypes.Structure):
    _fields_ = [
        ("radius", f32),
        ("margin", f32),
        ("relaxation", f32),
        ("max_displacement", f32),
        ("vt_capacity", i32),
        ("ee_capacity", i32),
        ("detection_interval", i32),
        ("edge_parallel_eps", f32),
        ("rest_exclusion_radius", f32),
        ("q_rest", vp),
        ("vt_filter", vp),
        ("vt_filter_offsets", vp),
        ("ee_filter", vp),
        ("ee_filter_offsets", vp),
        ("vt_list", vp),
        ("vt_count", vp),
        ("vt_min_dist", vp),
        ("tri_min_dist", vp),
        ("ee_list", vp),
        ("ee_count", vp),
        ("ee_min_dist", vp),
    ]


class _OraclePairs(ctypes.Structure):
    _fields_ = [("pair_count", i32), ("pairs", vp), ("edge_pair_count", i32), ("edge_pairs", vp),
                ("face_pair_count", i32), ("face_pairs", vp)]


_lib = None


def build() -> Path:
    """Compile the oracle with its Makefile (gcc, no FMA contraction)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        for sfx in ("f32", "f64"):
            for name in ("step", "create_soft_contacts", "simulate", "residual", "block_solve", "triangle_closest",
                         "vt_contact", "ee_contact", "self_detect"):
                getattr(L, f"vbd_oracle_{name}_{sfx}").restype = i32
            for name in ("membrane", "bending", "tet", "body_contact", "spring", "closest_edge_edge",
                         "self_force_norm", "truncation_ts"):
                getattr(L, f"vbd_oracle_{name}_{sfx}").restype = None
        L.vbd_oracle_atan2_f32.restype = f32
        L.vbd_oracle_atan2_f32.argtypes = [f32, f32]
        L.vbd_oracle_atan2_f64.restype = ctypes.c_double
        L.vbd_oracle_atan2_f64.argtypes = [ctypes.c_double, ctypes.c_double]
        _lib = L
    return _lib


def _np(t, dtype):
    if t is None:
        return np.zeros(0, dtype=dtype)
    if hasattr(t, "detach"):
        t = t.detach().cpu().numpy()
    return np.ascontiguousarray(np.asarray(t), dtype=dtype)


def _p(a: np.ndarray):
    """Address of a numpy array as a ctypes pointer argument (NULL when empty)."""
    return ctypes.c_void_p(a.ctypes.data) if a.size else None


def _addr(a: np.ndarray):
    """Address for a c_void_p Structure field."""
    return a.ctypes.data if a.size else None


def _real(precision: str):
    return (np.float32, ctypes.c_float) if precision == "f32" else (np.float64, ctypes.c_double)


class OracleModel:
    """Host copy of a model in the oracle's layout (keeps the numpy arrays alive)."""

    def __init__(self, model, color_groups=None):
        adj = model.soft_mesh_adjacency
        self.N = model.particle_count
        a = {}
        a["mass"] = _np(model.particle_mass, np.float32)
        a["inv_mass"] = _np(model.particle_inv_mass, np.float32)
        a["radius"] = _np(model.particle_radius, np.float32)
        a["flags"] = _np(model.particle_flags, np.int32)
        a["world"] = _np(model.particle_world, np.int32)
        groups = color_groups if color_groups is not None else [_np(g, np.int32) for g in model.particle_color_groups]
        groups = [np.asarray(g, dtype=np.int32) for g in groups]
        colors = np.full(self.N, -1, dtype=np.int32)
        for c, g in enumerate(groups):
            colors[g] = c
        a["colors"] = colors
        a["color_offsets"] = np.concatenate([[0], np.cumsum([len(g) for g in groups])]).astype(np.int32)
        a["color_particles"] = np.ascontiguousarray(np.concatenate(groups) if groups else np.zeros(0), dtype=np.int32)
        a["tri_indices"] = _np(model.tri_indices, np.int32)
        a["tri_poses"] = _np(model.tri_poses, np.float32)
        a["tri_materials"] = _np(model.tri_materials, np.float32)
        a["tri_areas"] = _np(model.tri_areas, np.float32)
        a["edge_indices"] = _np(model.edge_indices, np.int32)
        a["edge_rest_angle"] = _np(model.edge_rest_angle, np.float32)
        a["edge_rest_length"] = _np(model.edge_rest_length, np.float32)
        a["edge_bend"] = _np(model.edge_bending_properties, np.float32)
        a["tet_indices"] = _np(model.tet_indices, np.int32)
        a["tet_poses"] = _np(model.tet_poses, np.float32)
        a["tet_materials"] = _np(model.tet_materials, np.float32)
        for key in ("tris", "edges", "tets"):
            vals = getattr(adj, f"v_adj_{key}")
            offs = getattr(adj, f"v_adj_{key}_offsets")
            a[f"adj_{key}"] = np.ascontiguousarray(vals if vals is not None else np.zeros(0), dtype=np.int32)
            offs = offs if offs is not None and len(offs) else np.zeros(self.N + 1)
            a[f"adj_{key}_off"] = np.ascontiguousarray(offs, dtype=np.int32)
        a["shape_type"] = _np(model.shape_type, np.int32)
        a["shape_body"] = _np(model.shape_body, np.int32)
        a["shape_flags"] = _np(model.shape_flags, np.int32)
        a["shape_world"] = _np(model.shape_world, np.int32)
        a["shape_transform"] = _np(model.shape_transform, np.float32)
        a["shape_scale"] = _np(model.shape_scale, np.float32)
        a["shape_margin"] = _np(model.shape_margin, np.float32)
        a["shape_ke"] = _np(model.shape_material_ke, np.float32)
        a["shape_kd"] = _np(model.shape_material_kd, np.float32)
        a["shape_mu"] = _np(model.shape_material_mu, np.float32)
        ns = int(getattr(model, "spring_count", 0) or 0)
        a["spring_indices"] = _np(model.spring_indices, np.int32) if ns else np.zeros(2, np.int32)
        for key in ("spring_rest_length", "spring_stiffness", "spring_damping"):
            a[key] = _np(getattr(model, key), np.float32) if ns else np.zeros(1, np.float32)
        self.arrays = a
        s = _OracleModel()
        s.particle_count = self.N
        s.tri_count = model.tri_count
        s.edge_count = model.edge_count
        s.tet_count = model.tet_count
        s.particle_mass = _addr(a["mass"])
        s.particle_inv_mass = _addr(a["inv_mass"])
        s.particle_radius = _addr(a["radius"])
        s.particle_flags = _addr(a["flags"])
        s.particle_colors = _addr(a["colors"])
        s.particle_world = _addr(a["world"])
        s.tri_indices = _addr(a["tri_indices"])
        s.tri_poses = _addr(a["tri_poses"])
        s.tri_materials = _addr(a["tri_materials"])
        s.tri_areas = _addr(a["tri_areas"])
        s.edge_indices = _addr(a["edge_indices"])
        s.edge_rest_angle = _addr(a["edge_rest_angle"])
        s.edge_rest_length = _addr(a["edge_rest_length"])
        s.edge_bending_properties = _addr(a["edge_bend"])
        s.tet_indices = _addr(a["tet_indices"])
        s.tet_poses = _addr(a["tet_poses"])
        s.tet_materials = _addr(a["tet_materials"])
        s.v_adj_tris = _addr(a["adj_tris"])
        s.v_adj_tris_offsets = _addr(a["adj_tris_off"])
        s.v_adj_edges = _addr(a["adj_edges"])
        s.v_adj_edges_offsets = _addr(a["adj_edges_off"])
        s.v_adj_tets = _addr(a["adj_tets"])
        s.v_adj_tets_offsets = _addr(a["adj_tets_off"])
        s.color_count = len(groups)
        s.color_offsets = _addr(a["color_offsets"])
        s.color_particles = _addr(a["color_particles"])
        g = _np(model.gravity, np.float32).reshape(-1, 3)[0]
        s.gravity = (f32 * 3)(*[float(v) for v in g])
        s.shape_count = model.shape_count
        s.shape_type = _addr(a["shape_type"])
        s.shape_body = _addr(a["shape_body"])
        s.shape_flags = _addr(a["shape_flags"])
        s.shape_world = _addr(a["shape_world"])
        s.shape_transform = _addr(a["shape_transform"])
        s.shape_scale = _addr(a["shape_scale"])
        s.shape_margin = _addr(a["shape_margin"])
        s.shape_material_ke = _addr(a["shape_ke"])
        s.shape_material_kd = _addr(a["shape_kd"])
        s.shape_material_mu = _addr(a["shape_mu"])
        s.world_count = model.world_count
        s.soft_contact_ke = float(model.soft_contact_ke)
        s.soft_contact_kd = float(model.soft_contact_kd)
        s.soft_contact_mu = float(model.soft_contact_mu)
        s.spring_count = ns
        s.spring_indices = _addr(a["spring_indices"])
        s.spring_rest_length = _addr(a["spring_rest_length"])
        s.spring_stiffness = _addr(a["spring_stiffness"])
        s.spring_damping = _addr(a["spring_damping"])
        self.struct = s
        self.world_count = model.world_count
        from newton_amd.sim.collide import world_compatible_pairs  # noqa: PLC0415

        pairs = world_compatible_pairs(a["world"], a["shape_world"], model.world_count)
        self.pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1)
        self.pairs_struct = _OraclePairs(int(pairs.shape[0]), _addr(self.pairs), 0, None, 0, None)

    def set_full_surface_pairs(self, edge_pairs, face_pairs) -> None:
        """Candidate (edge, shape) and (triangle, shape) pairs of the full-surface passes (empty: off)."""
        self.edge_pairs = np.ascontiguousarray(np.asarray(edge_pairs, dtype=np.int32).reshape(-1))
        self.face_pairs = np.ascontiguousarray(np.asarray(face_pairs, dtype=np.int32).reshape(-1))
        self.pairs_struct.edge_pair_count = self.edge_pairs.size // 2
        self.pairs_struct.edge_pairs = _addr(self.edge_pairs) if self.edge_pairs.size else None
        self.pairs_struct.face_pair_count = self.face_pairs.size // 2
        self.pairs_struct.face_pairs = _addr(self.face_pairs) if self.face_pairs.size else None


def params(iterations: int, friction_epsilon: float = 1e-2, avbd_beta: float = 0.0, k_start: float = -1.0,
           use_libm_atan2: bool = False, fp64_ill_conditioned: bool = True, num_threads: int = 1,
           self_contact: "SelfContact | None" = None, residual_stop: float = 0.0, body_q=None,
           body_q_prev=None) -> _OracleParams:
    """Solver settings of the oracle.  ``residual_stop`` > 0 ends a step with a residual history once
    the residual norm is <= ``residual_stop`` x its value at the inertial guess (iterate to a relative
    tolerance, checked after every iteration); later history entries stay 0.  ``body_q`` /
    ``body_q_prev`` ([B,7] float32, p then q xyzw) are the poses of the bodies shapes are attached to at
    the step's end and start (an external rigid solver's state_out / state_in)."""
    prm = _OracleParams(int(iterations), float(friction_epsilon), float(avbd_beta), float(k_start),
                        int(use_libm_atan2), int(fp64_ill_conditioned), int(num_threads),
                        None if self_contact is None else ctypes.addressof(self_contact.struct),
                        float(residual_stop))
    prm._self_contact = self_contact  # keeps the settings and their arrays alive
    set_bodies(prm, body_q, body_q_prev)
    return prm


def set_bodies(prm: _OracleParams, body_q, body_q_prev=None) -> None:
    """Body poses of the next oracle calls ([B,7] float32 each; ``body_q_prev`` defaults to ``body_q``)."""
    if body_q is None:
        prm.body_q = prm.body_q_prev = None
        prm.body_count = 0
        prm._bodies = None
        return
    bq = np.ascontiguousarray(np.asarray(body_q, dtype=np.float32).reshape(-1, 7))
    bp = bq if body_q_prev is None else np.ascontiguousarray(np.asarray(body_q_prev, dtype=np.float32).reshape(-1, 7))
    prm._bodies = (bq, bp)
    prm.body_q, prm.body_q_prev, prm.body_count = _addr(bq), _addr(bp), int(bq.shape[0])


class SelfContact:
    """Cloth self-contact settings (SolverVBD ``particle_*`` arguments) and the outputs of the
    oracle's last collision detection (triangle / edge lists in ascending primitive order)."""

    def __init__(self, particle_count: int, tri_count: int, edge_count: int, radius: float = 0.2,
                 margin: float = 0.2, relaxation: float = 0.85, vt_capacity: int = 32, ee_capacity: int = 64,
                 detection_interval: int = 0, edge_parallel_eps: float = 1e-5, rest_exclusion_radius: float = 0.0,
                 q_rest=None, vt_filter=None, vt_filter_offsets=None, ee_filter=None, ee_filter_offsets=None):
        n, t, e = int(particle_count), int(tri_count), int(edge_count)
        self._keep = []

        def arr(x, dtype):
            if x is None:
                return None
            a = np.ascontiguousarray(np.asarray(x), dtype=dtype)
            self._keep.append(a)
            return a

        self.q_rest = arr(None if q_rest is None else np.asarray(q_rest).reshape(-1, 3), np.float32)
        self.vt_filter = arr(vt_filter, np.int32)
        self.vt_filter_offsets = arr(vt_filter_offsets, np.int32)
        self.ee_filter = arr(ee_filter, np.int32)
        self.ee_filter_offsets = arr(ee_filter_offsets, np.int32)
        self.vt_list = np.full((max(n, 1), vt_capacity), -1, dtype=np.int32)
        self.vt_count = np.zeros(max(n, 1), dtype=np.int32)
        self.vt_min_dist = np.zeros(max(n, 1), dtype=np.float64)
        self.tri_min_dist = np.zeros(max(t, 1), dtype=np.float64)
        self.ee_list = np.full((max(e, 1), ee_capacity), -1, dtype=np.int32)
        self.ee_count = np.zeros(max(e, 1), dtype=np.int32)
        self.ee_min_dist = np.zeros(max(e, 1), dtype=np.float64)
        # the isotropic clamp is computed by the reference in Python double, then passed as float32
        max_disp = float(np.float32(float(margin) * float(relaxation) * 0.5))
        self.struct = _OracleSelf(
            float(radius), float(margin), float(relaxation), max_disp, int(vt_capacity), int(ee_capacity),
            int(detection_interval), float(edge_parallel_eps), float(rest_exclusion_radius),
            *(None if x is None or x.size == 0 else x.ctypes.data
              for x in (self.q_rest, self.vt_filter, self.vt_filter_offsets, self.ee_filter, self.ee_filter_offsets,
                        self.vt_list, self.vt_count, self.vt_min_dist, self.tri_min_dist, self.ee_list,
                        self.ee_count, self.ee_min_dist)))


def self_detect(om: "OracleModel", sc: SelfContact, q, precision: str = "f32") -> None:
    """Vertex-triangle and edge-edge detection at q into sc's output arrays."""
    rt, _ = _real(precision)
    q = np.ascontiguousarray(q, dtype=rt).reshape(-1, 3)
    rc = getattr(lib(), f"vbd_oracle_self_detect_{precision}")(ctypes.byref(om.struct), ctypes.byref(sc.struct),
                                                                 _p(q))
    if rc != 0:
        raise RuntimeError(f"oracle self_detect failed ({rc})")


class Contacts:
    """Host soft-contact arrays in the oracle's REAL type."""

    def __init__(self, capacity: int, precision: str = "f32"):
        dt, _ = _real(precision)
        self.max = int(capacity)
        n = max(self.max, 1)
        self.count = np.zeros(1, dtype=np.int32)
        self.particle = np.zeros(n, dtype=np.int32)
        self.indices = np.full((n, 3), -1, dtype=np.int32)
        self.shape = np.zeros(n, dtype=np.int32)
        self.body_pos = np.zeros((n, 3), dtype=dt)
        self.body_vel = np.zeros((n, 3), dtype=dt)
        self.normal = np.zeros((n, 3), dtype=dt)
        self.bary = np.zeros((n, 3), dtype=dt)

    @classmethod
    def from_device(cls, contacts, precision: str = "f32"):
        c = cls(contacts.soft_contact_max, precision)
        dt, _ = _real(precision)
        n = contacts.soft_contact_max
        if n:
            c.count[:] = contacts.soft_contact_count.cpu().numpy()
            c.particle[:n] = contacts.soft_contact_particle.cpu().numpy()
            c.indices[:n] = contacts.soft_contact_indices.cpu().numpy()
            c.shape[:n] = contacts.soft_contact_shape.cpu().numpy()
            c.body_pos[:n] = contacts.soft_contact_body_pos.cpu().numpy().astype(dt)
            c.body_vel[:n] = contacts.soft_contact_body_vel.cpu().numpy().astype(dt)
            c.normal[:n] = contacts.soft_contact_normal.cpu().numpy().astype(dt)
            c.bary[:n] = contacts.soft_contact_barycentric.cpu().numpy().astype(dt)
        return c


def create_soft_contacts(om: OracleModel, q: np.ndarray, margin: float, contacts: Contacts, precision: str = "f32",
                         body_q=None):
    """Particle records (and, with full-surface pairs set on ``om``, edge and face records after them)
    in pair order; ``body_q`` ([B,7]) places shapes attached to bodies."""
    dt, ct = _real(precision)
    q = np.ascontiguousarray(q, dtype=dt)
    bq = None if body_q is None else np.ascontiguousarray(np.asarray(body_q, dtype=np.float32).reshape(-1, 7))
    fn = getattr(lib(), f"vbd_oracle_create_soft_contacts_{precision}")
    rc = fn(ctypes.byref(om.struct), ctypes.byref(om.pairs_struct), _p(q), None if bq is None else _p(bq),
            i32(0 if bq is None else bq.shape[0]), ct(margin), i32(contacts.max), _p(contacts.count),
            _p(contacts.particle), _p(contacts.indices), _p(contacts.bary), _p(contacts.shape),
            _p(contacts.body_pos), _p(contacts.body_vel), _p(contacts.normal))
    if rc != 0:
        raise RuntimeError(f"oracle create_soft_contacts failed ({rc})")
    return int(contacts.count[0])


def step(om: OracleModel, prm: _OracleParams, q, qd, f_ext, contacts: Contacts | None, dt: float,
         precision: str = "f32", residual_history: bool = False):
    """One SolverVBD.step; returns (q_iterate, q_out, qd_out[, residual norms per iteration])."""
    rt, ct = _real(precision)
    q = np.array(q, dtype=rt, copy=True).reshape(-1, 3)
    qd = np.ascontiguousarray(qd, dtype=rt).reshape(-1, 3)
    fe = None if f_ext is None else np.ascontiguousarray(f_ext, dtype=rt).reshape(-1, 3)
    q_out = np.zeros_like(q)
    qd_out = np.zeros_like(q)
    hist = np.zeros(prm.iterations + 1, dtype=np.float64) if residual_history else None
    if contacts is None:
        count, cmax, ci, cs, cbp, cbv, cn, cb = 0, 0, None, None, None, None, None, None
    else:
        count, cmax = int(contacts.count[0]), contacts.max
        ci, cs = _p(contacts.indices), _p(contacts.shape)
        keep = [np.ascontiguousarray(x, dtype=rt)
                for x in (contacts.body_pos, contacts.body_vel, contacts.normal, contacts.bary)]
        cbp, cbv, cn, cb = (_p(x) for x in keep)
    fn = getattr(lib(), f"vbd_oracle_step_{precision}")
    rc = fn(ctypes.byref(om.struct), ctypes.byref(prm), i32(count), i32(cmax), ci, cs, cbp, cbv, cn, cb, _p(q),
            _p(qd), None if fe is None else _p(fe), _p(q_out), _p(qd_out), ct(dt), None if hist is None else _p(hist))
    if rc != 0:
        raise RuntimeError(f"oracle step failed ({rc})")
    if residual_history:
        return q, q_out, qd_out, hist
    return q, q_out, qd_out


def simulate(om: OracleModel, prm: _OracleParams, q, qd, substeps: int, dt: float, margin: float, contact_max: int,
             record_every: int = 0, f_ext=None, precision: str = "f32"):
    """Frame loop clear_forces -> collide -> step -> swap; returns (q, qd, recorded q every `record_every`)."""
    rt, ct = _real(precision)
    q = np.array(q, dtype=rt, copy=True).reshape(-1, 3)
    qd = np.array(qd, dtype=rt, copy=True).reshape(-1, 3)
    fe = None if f_ext is None else np.ascontiguousarray(f_ext, dtype=rt).reshape(-1, 3)
    nrec = substeps // record_every if record_every > 0 else 0
    rec = np.zeros((max(nrec, 1), om.N, 3), dtype=rt)
    fn = getattr(lib(), f"vbd_oracle_simulate_{precision}")
    rc = fn(ctypes.byref(om.struct), ctypes.byref(prm), ctypes.byref(om.pairs_struct), _p(q), _p(qd),
            None if fe is None else _p(fe), i32(substeps), ct(dt), ct(margin), i32(contact_max), i32(record_every),
            _p(rec))
    if rc != 0:
        raise RuntimeError(f"oracle simulate failed ({rc})")
    return q, qd, rec[:nrec]


def residual(om: OracleModel, prm: _OracleParams, contacts: Contacts | None, q, q_prev, inertia, dt: float,
             precision: str = "f64"):
    rt, ct = _real(precision)
    q, q_prev, inertia = (np.ascontiguousarray(x, dtype=rt).reshape(-1, 3) for x in (q, q_prev, inertia))
    g = np.zeros_like(q)
    norm = np.zeros(1, dtype=np.float64)
    if contacts is None:
        args = (i32(0), i32(0), None, None, None, None, None, None)
    else:
        keep = [np.ascontiguousarray(x, dtype=rt)
                for x in (contacts.body_pos, contacts.body_vel, contacts.normal, contacts.bary)]
        args = (i32(int(contacts.count[0])), i32(contacts.max), _p(contacts.indices), _p(contacts.shape),
                *(_p(x) for x in keep))
    fn = getattr(lib(), f"vbd_oracle_residual_{precision}")
    rc = fn(ctypes.byref(om.struct), ctypes.byref(prm), *args, _p(q), _p(q_prev), _p(inertia), ct(dt), _p(g),
            _p(norm))
    if rc != 0:
        raise RuntimeError(f"oracle residual failed ({rc})")
    return g, float(norm[0])


# ---- element-level entry points (for unit tests) ---------------------------
def membrane(x, x_anchor, pose, area, mu, lmbd, kd, dt, v_order, precision="f64"):
    rt, ct = _real(precision)
    x = np.ascontiguousarray(x, dtype=rt).reshape(3, 3)
    xa = np.ascontiguousarray(x_anchor, dtype=rt).reshape(3, 3)
    pose = np.ascontiguousarray(pose, dtype=np.float32).reshape(4)
    f, h = np.zeros(3, dtype=rt), np.zeros(9, dtype=rt)
    getattr(lib(), f"vbd_oracle_membrane_{precision}")(_p(x), _p(xa), _p(pose), f32(area), f32(mu), f32(lmbd),
                                                       f32(kd), ct(dt), i32(v_order), _p(f), _p(h))
    return f, h.reshape(3, 3)


def bending(x, x_anchor, rest_angle, rest_length, ke, kd, dt, v_order, precision="f64", use_libm_atan2=False):
    rt, ct = _real(precision)
    x = np.ascontiguousarray(x, dtype=rt).reshape(4, 3)
    xa = np.ascontiguousarray(x_anchor, dtype=rt).reshape(4, 3)
    f, h = np.zeros(3, dtype=rt), np.zeros(9, dtype=rt)
    getattr(lib(), f"vbd_oracle_bending_{precision}")(_p(x), _p(xa), f32(rest_angle), f32(rest_length), f32(ke),
                                                      f32(kd), ct(dt), i32(v_order), i32(int(use_libm_atan2)),
                                                      _p(f), _p(h))
    return f, h.reshape(3, 3)


def tet(x, x_prev, dm_inv, mu, lmbd, kd, dt, v_order, precision="f64"):
    rt, ct = _real(precision)
    x = np.ascontiguousarray(x, dtype=rt).reshape(4, 3)
    xp = np.ascontiguousarray(x_prev, dtype=rt).reshape(4, 3)
    dm = np.ascontiguousarray(dm_inv, dtype=np.float32).reshape(9)
    f, h = np.zeros(3, dtype=rt), np.zeros(9, dtype=rt)
    getattr(lib(), f"vbd_oracle_tet_{precision}")(_p(x), _p(xp), _p(dm), f32(mu), f32(lmbd), f32(kd), ct(dt),
                                                  i32(v_order), _p(f), _p(h))
    return f, h.reshape(3, 3)


def body_contact(x, x_prev, body_pos, body_vel, normal, radius, margin, ke, kd, mu, friction_epsilon, dt,
                 precision="f64"):
    rt, ct = _real(precision)
    arrs = [np.ascontiguousarray(v, dtype=rt).reshape(3) for v in (x, x_prev, body_pos, body_vel, normal)]
    f, h = np.zeros(3, dtype=rt), np.zeros(9, dtype=rt)
    getattr(lib(), f"vbd_oracle_body_contact_{precision}")(*[_p(a) for a in arrs], ct(radius), ct(margin), ct(ke),
                                                           ct(kd), ct(mu), ct(friction_epsilon), ct(dt), _p(f), _p(h))
    return f, h.reshape(3, 3)


def spring(x, x_anchor, rest, ke, kd, dt, precision="f64"):
    """evaluate_spring_force_and_hessian_both_vertices: (force on x0, force on x1, shared Hessian)."""
    rt, ct = _real(precision)
    x = np.ascontiguousarray(x, dtype=rt).reshape(2, 3)
    xa = np.ascontiguousarray(x_anchor, dtype=rt).reshape(2, 3)
    f, h = np.zeros(6, dtype=rt), np.zeros(9, dtype=rt)
    getattr(lib(), f"vbd_oracle_spring_{precision}")(_p(x), _p(xa), ct(rest), ct(ke), ct(kd), ct(dt), _p(f), _p(h))
    return f[:3], f[3:], h.reshape(3, 3)


def block_solve(h, f, fp64_ill_conditioned=False, precision="f32"):
    rt, _ = _real(precision)
    h = np.ascontiguousarray(h, dtype=rt).reshape(9)
    f = np.ascontiguousarray(f, dtype=rt).reshape(3)
    dx = np.zeros(3, dtype=rt)
    ok = getattr(lib(), f"vbd_oracle_block_solve_{precision}")(_p(h), _p(f), i32(int(fp64_ill_conditioned)), _p(dx))
    return bool(ok), dx


def atan2(y, x, precision="f32"):
    return getattr(lib(), f"vbd_oracle_atan2_{precision}")(y, x)


# ---- self-contact element-level entry points ---------------------------------
def triangle_closest(a, b, c, p, precision="f64"):
    """triangle_closest_point: (closest point, barycentrics, feature 0..6)."""
    rt, _ = _real(precision)
    x = np.ascontiguousarray(np.stack([a, b, c, p]), dtype=rt)
    cp, bary = np.zeros(3, dtype=rt), np.zeros(3, dtype=rt)
    feat = getattr(lib(), f"vbd_oracle_triangle_closest_{precision}")(_p(x), _p(cp), _p(bary))
    return cp, bary, int(feat)


def closest_edge_edge(p1, q1, p2, q2, eps=1e-5, precision="f64"):
    """(s, t, distance) of the closest points of segments p1q1 and p2q2."""
    rt, ct = _real(precision)
    x = np.ascontiguousarray(np.stack([p1, q1, p2, q2]), dtype=rt)
    out = np.zeros(3, dtype=rt)
    getattr(lib(), f"vbd_oracle_closest_edge_edge_{precision}")(_p(x), ct(eps), _p(out))
    return out


def self_force_norm(dis, radius, k, precision="f64"):
    rt, ct = _real(precision)
    out = np.zeros(2, dtype=rt)
    getattr(lib(), f"vbd_oracle_self_force_norm_{precision}")(ct(dis), ct(radius), ct(k), _p(out))
    return out


def vt_contact(abcp, anchor, radius, ke, kd, mu, friction_epsilon, dt, precision="f64"):
    """Vertex-triangle contact: (has_contact, forces [4,3] on a, b, c, p, Hessians [4,3,3])."""
    rt, ct = _real(precision)
    x = np.ascontiguousarray(abcp, dtype=rt).reshape(4, 3)
    xa = np.ascontiguousarray(anchor, dtype=rt).reshape(4, 3)
    f, h = np.zeros(12, dtype=rt), np.zeros(36, dtype=rt)
    has = getattr(lib(), f"vbd_oracle_vt_contact_{precision}")(_p(x), _p(xa), ct(radius), ct(ke), ct(kd), ct(mu),
                                                               ct(friction_epsilon), ct(dt), _p(f), _p(h))
    return bool(has), f.reshape(4, 3), h.reshape(4, 3, 3)


def ee_contact(x, anchor, radius, ke, kd, mu, friction_epsilon, dt, parallel_eps=1e-5, precision="f64"):
    """Edge-edge contact on e1's vertices: (has_contact, forces [2,3], Hessians [2,3,3])."""
    rt, ct = _real(precision)
    x = np.ascontiguousarray(x, dtype=rt).reshape(4, 3)
    xa = np.ascontiguousarray(anchor, dtype=rt).reshape(4, 3)
    f, h = np.zeros(6, dtype=rt), np.zeros(18, dtype=rt)
    has = getattr(lib(), f"vbd_oracle_ee_contact_{precision}")(_p(x), _p(xa), ct(radius), ct(ke), ct(kd), ct(mu),
                                                               ct(friction_epsilon), ct(dt), ct(parallel_eps),
                                                               _p(f), _p(h))
    return bool(has), f.reshape(2, 3), h.reshape(2, 3, 3)


def truncation_ts(kind, x, delta, parallel_eps=1e-5, gamma=0.85, precision="f64"):
    """Planar truncation factors of one collision's four vertices (kind 0 edge-edge, 1 vertex-triangle)."""
    rt, ct = _real(precision)
    x = np.ascontiguousarray(x, dtype=rt).reshape(4, 3)
    d = np.ascontiguousarray(delta, dtype=rt).reshape(4, 3)
    t = np.zeros(4, dtype=rt)
    getattr(lib(), f"vbd_oracle_truncation_ts_{precision}")(i32(kind), _p(x), _p(d), ct(parallel_eps), ct(gamma),
                                                            _p(t))
    return t


# ---- SolverBase.integrate_particles (solvers/solver.py:20-56) -------------------------------
def integrate_particles(x, v, f, inv_mass, flags, world, gravity, dt: float, v_max: float):
    """Semi-implicit Euler of every particle in float32, one IEEE rounding per operation in the
    reference kernel's order: ``v1 = v0 + (f * w + g * step(-w)) * dt``; ``|v1| = sqrt((x*x + y*y) + z*z)``;
    above ``v_max`` ``v1 *= v_max / |v1|``; ``x1 = x0 + v1 * dt``.  Inactive particles keep x and v
    (the kernel leaves their v_new unwritten).  ``world`` -1 selects gravity row 0.  Returns (x1, v1)."""
    f32 = np.float32
    x = np.asarray(x, f32).reshape(-1, 3)
    v = np.asarray(v, f32).reshape(-1, 3)
    f = np.asarray(f, f32).reshape(-1, 3)
    w = np.asarray(inv_mass, f32)
    g = np.asarray(gravity, f32).reshape(-1, 3)
    wi = np.clip(np.asarray(world, np.int64), 0, g.shape[0] - 1)
    dt, v_max = f32(dt), f32(v_max)
    active = (np.asarray(flags) & 1) != 0
    has_mass = (-w < f32(0.0)).astype(f32)
    acc = f * w[:, None] + g[wi] * has_mass[:, None]
    v1 = v + acc * dt
    mag = np.sqrt((v1[:, 0] * v1[:, 0] + v1[:, 1] * v1[:, 1]) + v1[:, 2] * v1[:, 2])
    over = mag > v_max
    scale = np.where(over, v_max / np.where(over, mag, f32(1.0)), f32(1.0)).astype(f32)
    v1 = np.where(over[:, None], v1 * scale[:, None], v1)
    x1 = x + v1 * dt
    return np.where(active[:, None], x1, x).astype(f32), np.where(active[:, None], v1, v).astype(f32)
