"""ctypes bindings for the native libraries built by ``newton_amd.build``.

``libvbd_hip.so`` is the product compute path; it is loaded lazily and any
failure to find or load it raises -- there is no fallback implementation.
The Structure layouts mirror ``include/vbd_hip.h`` field for field.
"""

from __future__ import annotations

import ctypes
import threading
from pathlib import Path

LIB_DIR = Path(__file__).resolve().parent / "lib"

_lock = threading.Lock()
_hip = None
_host = None
_ws = None

c_int32 = ctypes.c_int32
c_float = ctypes.c_float
c_void_p = ctypes.c_void_p


class VbdModelDesc(ctypes.Structure):
    _fields_ = [
        ("particle_count", c_int32),
        ("tri_count", c_int32),
        ("edge_count", c_int32),
        ("tet_count", c_int32),
        ("particle_mass", c_void_p),
        ("particle_inv_mass", c_void_p),
        ("particle_radius", c_void_p),
        ("particle_flags", c_void_p),
        ("tri_indices", c_void_p),
        ("tri_poses", c_void_p),
        ("tri_materials", c_void_p),
        ("tri_areas", c_void_p),
        ("edge_indices", c_void_p),
        ("edge_rest_angle", c_void_p),
        ("edge_rest_length", c_void_p),
        ("edge_bending_properties", c_void_p),
        ("tet_indices", c_void_p),
        ("tet_poses", c_void_p),
        ("tet_materials", c_void_p),
        ("v_adj_tris", c_void_p),
        ("v_adj_tris_offsets", c_void_p),
        ("v_adj_edges", c_void_p),
        ("v_adj_edges_offsets", c_void_p),
        ("v_adj_tets", c_void_p),
        ("v_adj_tets_offsets", c_void_p),
        ("color_count", c_int32),
        ("color_offsets_host", c_void_p),
        ("color_particles", c_void_p),
        ("gravity", c_void_p),
        ("shape_count", c_int32),
        ("shape_margin", c_void_p),
        ("shape_material_ke", c_void_p),
        ("shape_material_kd", c_void_p),
        ("shape_material_mu", c_void_p),
        ("soft_contact_ke", c_float),
        ("soft_contact_kd", c_float),
        ("soft_contact_mu", c_float),
        ("spring_count", c_int32),
        ("spring_indices", c_void_p),
        ("spring_rest_length", c_void_p),
        ("spring_stiffness", c_void_p),
        ("spring_damping", c_void_p),
        ("shape_body", c_void_p),
        ("body_count", c_int32),
    ]


class VbdSolverParams(ctypes.Structure):
    _fields_ = [
        ("iterations", c_int32),
        ("friction_epsilon", c_float),
        ("avbd_beta", c_float),
        ("k_start", c_float),
        ("fp64_ill_conditioned", c_int32),
        ("soft_contact_capacity", c_int32),
        ("self_contact", c_int32),
        ("self_contact_radius", c_float),
        ("self_contact_margin", c_float),
        ("conservative_bound_relaxation", c_float),
        ("max_displacement", c_float),
        ("vertex_contact_buffer_size", c_int32),
        ("edge_contact_buffer_size", c_int32),
        ("collision_detection_interval", c_int32),
        ("edge_parallel_epsilon", c_float),
        ("rest_shape_contact_exclusion_radius", c_float),
        ("particle_q_rest", c_void_p),
        ("particle_world", c_void_p),
        ("vt_filter", c_void_p),
        ("vt_filter_offsets", c_void_p),
        ("ee_filter", c_void_p),
        ("ee_filter_offsets", c_void_p),
        ("soft_contact_full_surface", c_int32),
    ]


class VbdContactsDesc(ctypes.Structure):
    _fields_ = [
        ("soft_contact_count", c_void_p),
        ("soft_contact_max", c_int32),
        ("soft_contact_indices", c_void_p),
        ("soft_contact_shape", c_void_p),
        ("soft_contact_body_pos", c_void_p),
        ("soft_contact_body_vel", c_void_p),
        ("soft_contact_normal", c_void_p),
        ("soft_contact_barycentric", c_void_p),
        ("body_q", c_void_p),
        ("body_q_prev", c_void_p),
        ("dt", c_float),
    ]


class VbdContactsOut(ctypes.Structure):
    _fields_ = [
        ("soft_contact_count", c_void_p),
        ("soft_contact_max", c_int32),
        ("soft_contact_particle", c_void_p),
        ("soft_contact_indices", c_void_p),
        ("soft_contact_barycentric", c_void_p),
        ("soft_contact_shape", c_void_p),
        ("soft_contact_body_pos", c_void_p),
        ("soft_contact_body_vel", c_void_p),
        ("soft_contact_normal", c_void_p),
    ]


class VbdCollideDesc(ctypes.Structure):
    _fields_ = [
        ("particle_count", c_int32),
        ("pair_count", c_int32),
        ("pairs", c_void_p),
        ("particle_radius", c_void_p),
        ("particle_flags", c_void_p),
        ("particle_world", c_void_p),
        ("shape_count", c_int32),
        ("shape_type", c_void_p),
        ("shape_flags", c_void_p),
        ("shape_world", c_void_p),
        ("shape_transform", c_void_p),
        ("shape_scale", c_void_p),
        ("shape_margin", c_void_p),
        ("shape_body", c_void_p),
        ("edge_pair_count", c_int32),
        ("edge_pairs", c_void_p),
        ("edge_indices", c_void_p),
        ("face_pair_count", c_int32),
        ("face_pairs", c_void_p),
        ("tri_indices", c_void_p),
    ]


# entry points of include/vbd_hip.h: name -> (restype, argtypes)
HIP_EXPORTS = {
    "vbd_last_error": (ctypes.c_char_p, []),
    "vbd_version": (c_int32, []),
    "vbd_create": (c_int32, [ctypes.POINTER(VbdModelDesc), ctypes.POINTER(VbdSolverParams), ctypes.POINTER(c_void_p)]),
    "vbd_destroy": (c_int32, [c_void_p]),
    "vbd_rebuild": (c_int32, [c_void_p]),
    "vbd_layout_stats": (c_int32, [c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    "vbd_autotune": (c_int32, [c_void_p, c_void_p, c_void_p, c_float, c_void_p, c_void_p]),
    "vbd_set_block_shape": (c_int32, [c_void_p, c_void_p, c_int32]),
    "vbd_set_pass_waves": (c_int32, [c_void_p, c_int32]),
    "vbd_set_aligned_items": (c_int32, [c_void_p, c_void_p, c_int32]),
    "vbd_set_pass_shape": (c_int32, [c_void_p, c_void_p, c_int32, c_int32, c_int32]),
    "vbd_get_pass_shape": (c_int32, [c_void_p, c_void_p]),
    "vbd_set_scalars": (c_int32, [c_void_p, c_int32, c_float, c_float, c_float, c_float]),
    "vbd_step": (
        c_int32,
        [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.POINTER(VbdContactsDesc), c_float],
    ),
    "vbd_contact_prepare": (c_int32, [c_void_p, c_void_p, ctypes.POINTER(VbdContactsDesc)]),
    "vbd_forward": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_float]),
    "vbd_iteration": (c_int32, [c_void_p, c_void_p, c_void_p, c_float]),
    "vbd_color_pass": (c_int32, [c_void_p, c_void_p, c_int32, c_void_p, c_float]),
    "vbd_finalize": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_float]),
    "vbd_integrate_particles": (
        c_int32,
        [c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_float,
         c_float, c_void_p, c_void_p],
    ),
    "vbd_residual": (c_int32, [c_void_p, c_void_p, c_void_p, c_float, c_void_p, c_void_p]),
    "vbd_set_iterate": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_int32]),
    "vbd_set_iterate_range": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_int32]),
    "vbd_iterate_to_residual": (
        c_int32,
        [c_void_p, c_void_p, c_void_p, c_float, ctypes.c_double, c_int32, c_int32, ctypes.POINTER(c_int32),
         ctypes.POINTER(ctypes.c_double)],
    ),
    "vbd_residual_worlds": (c_int32, [c_void_p, c_void_p, c_void_p, c_float, c_int32, c_void_p, c_void_p]),
    "vbd_iterate_to_residual_worlds": (
        c_int32,
        [c_void_p, c_void_p, c_void_p, c_float, c_int32, c_void_p, c_void_p, ctypes.POINTER(ctypes.c_double), c_int32,
         c_int32, c_int32, c_void_p, c_void_p, ctypes.POINTER(c_int32), ctypes.POINTER(c_int32),
         ctypes.POINTER(ctypes.c_double)],
    ),
    "vbd_copy_scratch": (c_int32, [c_void_p, c_void_p, c_int32, c_void_p]),
    "vbd_set_displacement": (c_int32, [c_void_p, c_void_p, c_void_p]),
    "vbd_self_detect": (c_int32, [c_void_p, c_void_p, c_void_p]),
    "vbd_self_copy_lists": (c_int32, [c_void_p, c_void_p] + [c_void_p] * 7),
    "vbd_self_status": (c_int32, [c_void_p, ctypes.POINTER(c_int32)]),
    "vbd_collide_create": (c_int32, [ctypes.POINTER(VbdCollideDesc), ctypes.POINTER(c_void_p)]),
    "vbd_collide_destroy": (c_int32, [c_void_p]),
    "vbd_collide_generate": (c_int32, [c_void_p, c_void_p, c_void_p, c_float, ctypes.POINTER(VbdContactsOut)]),
    "vbd_collide_generate_bodies": (
        c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_float, ctypes.POINTER(VbdContactsOut)]),
    "vbd_collide_generate_noalloc": (
        c_int32, [ctypes.POINTER(VbdCollideDesc), c_void_p, c_void_p, c_float, ctypes.POINTER(VbdContactsOut)]),
}

HOST_EXPORTS = {
    "vbd_host_number_edges": (
        c_int32,
        [ctypes.c_int64, c_void_p, c_int32, c_void_p, c_void_p, c_void_p, ctypes.POINTER(ctypes.c_int64),
         ctypes.POINTER(ctypes.c_int64)],
    ),
    "vbd_host_edge_tri_maps": (c_int32, [ctypes.c_int64, c_void_p, ctypes.c_int64, c_void_p, c_void_p, c_void_p]),
    "vbd_host_vertex_csr": (
        c_int32,
        [ctypes.c_int64, c_int32, c_void_p, c_int32, c_void_p, c_int32, c_int32, c_void_p, c_void_p, ctypes.c_int64],
    ),
    "vbd_host_color_graph": (
        c_int32,
        [
            c_int32,
            ctypes.c_int64,
            ctypes.POINTER(c_int32),
            c_int32,
            c_int32,
            c_float,
            ctypes.POINTER(c_int32),
            ctypes.POINTER(c_float),
        ],
    ),
}


c_u16p = c_void_p  # BF16 arrays are passed as device addresses

WS_EXPORTS = {
    "vbd_ws_last_error": (ctypes.c_char_p, []),
    "vbd_ws_sizes": (c_int32, [ctypes.POINTER(ctypes.c_int64)]),
    "vbd_ws_features": (
        c_int32,
        [c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_float,
         c_void_p, c_u16p],
    ),
    "vbd_ws_smooth": (c_int32, [c_void_p, c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_int32, c_void_p,
                                c_void_p]),
    "vbd_ws_pack": (c_int32, [c_void_p, c_void_p, c_u16p]),
    "vbd_ws_forward": (c_int32, [c_void_p, c_int32, c_u16p, c_u16p, c_void_p, c_void_p]),
    "vbd_ws_forward_train": (c_int32, [c_void_p, c_int32, c_u16p, c_u16p, c_void_p, c_void_p, c_u16p, c_void_p]),
    "vbd_ws_backward": (
        c_int32,
        [c_void_p, c_int32, c_u16p, c_void_p, c_void_p, c_u16p, c_void_p, c_u16p, c_void_p, c_void_p, c_void_p],
    ),
    "vbd_ws_reduce": (c_int32, [c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_void_p]),
    "vbd_ws_adam": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_float, c_float, c_float,
                              c_int32]),
}


def _load(path: Path, exports: dict):
    if not path.exists():
        raise RuntimeError(
            f"native library {path} is missing; build it with `python -m newton_amd.build` "
            "(there is no non-native fallback)"
        )
    lib = ctypes.CDLL(str(path))
    for name, (restype, argtypes) in exports.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    return lib


def hip_lib():
    """The HIP compute library (raises if it is not built)."""
    global _hip
    with _lock:
        if _hip is None:
            _hip = _load(LIB_DIR / "libvbd_hip.so", HIP_EXPORTS)
        return _hip


def host_lib():
    """The host model-construction library (raises if it is not built)."""
    global _host
    with _lock:
        if _host is None:
            _host = _load(LIB_DIR / "libvbd_host.so", HOST_EXPORTS)
        return _host


def ws_lib():
    """The learned warm-start library (raises if it is not built)."""
    global _ws
    with _lock:
        if _ws is None:
            _ws = _load(LIB_DIR / "libvbd_ws.so", WS_EXPORTS)
        return _ws


def ws_check(rc: int, what: str) -> None:
    if rc == 0:
        return
    msg = ws_lib().vbd_ws_last_error()
    text = msg.decode() if msg else ""
    if rc == -1:
        raise ValueError(f"{what}: {text}")
    raise VbdError(f"{what} failed ({rc}): {text}")


class VbdError(RuntimeError):
    """A libvbd_hip call returned an error status."""


def check(rc: int, what: str) -> None:
    """Turns a non-zero libvbd_hip status into the reference's exception types."""
    if rc == 0:
        return
    msg = hip_lib().vbd_last_error()
    text = msg.decode() if msg else ""
    if rc == -1:
        raise ValueError(f"{what}: {text}")
    if rc == -3:
        raise RuntimeError(f"{what}: {text}")
    raise VbdError(f"{what} failed ({rc}): {text}")


def ptr(t) -> int | None:
    """Device address of a torch tensor (None for None / empty tensors)."""
    if t is None or t.numel() == 0:
        return None
    return t.data_ptr()
