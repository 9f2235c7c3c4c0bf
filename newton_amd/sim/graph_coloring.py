"""Particle graph construction, coloring and per-world color merging.

Graph construction restates the reference's numpy pipeline exactly
(``newton/_src/sim/graph_coloring.py:45-240``): canonical ``(min, max)``
unique edges from active triangles, all six vertex pairs of active bending
hinges (when ``include_bending``), and all six pairs of active tetrahedra.
The coloring itself runs in this package's native host library
(``csrc/vbd_host.cpp``: MCS ordering + greedy assignment + balancing); the
reference delegates that step to warp-lang's native code, so the color of each
vertex is this implementation's own choice and every consumer of the model
uses the same groups.  The per-world merge follows
``combine_independent_coloring_plan`` (``graph_coloring.py:382-410``).
"""

from __future__ import annotations

import ctypes
from enum import Enum

import numpy as np

from .. import _native


class ColoringAlgorithm(Enum):
    MCS = 0
    GREEDY = 1


def _canonicalize_edges_np(edges_np: np.ndarray) -> np.ndarray:
    """Sorted endpoints, duplicates dropped (``graph_coloring.py:45-51``)."""
    if edges_np.size == 0:
        return np.empty((0, 2), dtype=np.int32)
    edges_sorted = np.sort(edges_np, axis=1)
    return np.unique(edges_sorted, axis=0).astype(np.int32)


def _apply_mask(arr: np.ndarray, mask) -> np.ndarray:
    if mask is None:
        return arr
    mask_arr = np.asarray(mask, dtype=bool)
    if mask_arr.ndim == 0:
        return arr if bool(mask_arr) else arr[:0]
    return arr[mask_arr]


def construct_tetmesh_graph_edges(tet_indices, tet_active_mask=None) -> np.ndarray:
    """Unique graph edges of tetrahedra (``graph_coloring.py:54-96``)."""
    if tet_indices is None:
        return np.empty((0, 2), dtype=np.int32)
    tet_np = _apply_mask(np.asarray(tet_indices, dtype=np.int32).reshape(-1, 4), tet_active_mask)
    if tet_np.size == 0:
        return np.empty((0, 2), dtype=np.int32)
    v0, v1, v2, v3 = tet_np[:, 0], tet_np[:, 1], tet_np[:, 2], tet_np[:, 3]
    edges = np.stack(
        [
            np.stack([v0, v1], axis=1),
            np.stack([v0, v2], axis=1),
            np.stack([v0, v3], axis=1),
            np.stack([v1, v2], axis=1),
            np.stack([v1, v3], axis=1),
            np.stack([v2, v3], axis=1),
        ],
        axis=0,
    ).reshape(-1, 2)
    return _canonicalize_edges_np(edges)


def construct_trimesh_graph_edges(
    tri_indices, tri_active_mask=None, bending_edge_indices=None, bending_edge_active_mask=None
) -> np.ndarray:
    """Unique graph edges of triangles plus active bending hinges (``graph_coloring.py:99-197``)."""
    parts = []
    if tri_indices is not None:
        tris = _apply_mask(np.asarray(tri_indices, dtype=np.int32).reshape(-1, 3), tri_active_mask)
        if tris.size > 0:
            v0, v1, v2 = tris[:, 0], tris[:, 1], tris[:, 2]
            parts.append(
                np.stack(
                    [np.stack([v0, v1], axis=1), np.stack([v0, v2], axis=1), np.stack([v1, v2], axis=1)], axis=0
                ).reshape(-1, 2)
            )
    if bending_edge_indices is not None:
        bend = _apply_mask(np.asarray(bending_edge_indices, dtype=np.int32).reshape(-1, 4), bending_edge_active_mask)
        if bend.size > 0:
            c = [bend[:, k : k + 1] for k in range(4)]
            pairs = np.concatenate(
                [
                    np.concatenate([c[0], c[1]], axis=1),
                    np.concatenate([c[0], c[2]], axis=1),
                    np.concatenate([c[0], c[3]], axis=1),
                    np.concatenate([c[1], c[2]], axis=1),
                    np.concatenate([c[1], c[3]], axis=1),
                    np.concatenate([c[2], c[3]], axis=1),
                ],
                axis=0,
            )
            pairs = pairs[np.all(pairs >= 0, axis=1)]
            if pairs.size > 0:
                parts.append(pairs)
    edges = np.concatenate(parts, axis=0) if parts else np.empty((0, 2), dtype=np.int32)
    return _canonicalize_edges_np(edges)


def construct_particle_graph(
    tri_indices, tri_active_mask, bending_edge_indices, bending_edge_active_mask, tet_indices, tet_active_mask
) -> np.ndarray:
    """Unified particle graph (``graph_coloring.py:200-240``)."""
    tri_edges = construct_trimesh_graph_edges(tri_indices, tri_active_mask, bending_edge_indices, bending_edge_active_mask)
    tet_edges = construct_tetmesh_graph_edges(tet_indices, tet_active_mask)
    return _canonicalize_edges_np(np.vstack([tri_edges, tet_edges]).astype(np.int32))


def color_graph(
    num_nodes: int,
    graph_edge_indices: np.ndarray,
    balance_colors: bool = True,
    target_max_min_color_ratio: float = 1.1,
    algorithm: ColoringAlgorithm = ColoringAlgorithm.MCS,
) -> list[np.ndarray]:
    """Color groups of a graph: one int array of vertex ids (ascending) per color."""
    if num_nodes == 0:
        return []
    edges = np.ascontiguousarray(np.asarray(graph_edge_indices, dtype=np.int32).reshape(-1, 2))
    if edges.ndim != 2:
        raise ValueError("graph_edge_indices must be a 2 dimensional array!")
    lib = _native.host_lib()
    colors = np.empty(num_nodes, dtype=np.int32)
    ratio = ctypes.c_float(0.0)
    alg = algorithm.value if isinstance(algorithm, ColoringAlgorithm) else int(algorithm)
    nc = lib.vbd_host_color_graph(
        int(num_nodes),
        int(edges.shape[0]),
        edges.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
        int(alg),
        int(bool(balance_colors)),
        float(target_max_min_color_ratio),
        colors.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
        ctypes.byref(ratio),
    )
    if nc < 0:
        raise ValueError(f"graph coloring failed with code {nc}")
    return convert_to_color_groups(nc, colors)


def convert_to_color_groups(num_colors: int, particle_colors: np.ndarray) -> list[np.ndarray]:
    """Per-color ascending vertex lists."""
    particle_colors = np.asarray(particle_colors)
    order = np.argsort(particle_colors, kind="stable")
    counts = np.bincount(particle_colors, minlength=num_colors)
    bounds = np.concatenate([[0], np.cumsum(counts)])
    return [order[bounds[c] : bounds[c + 1]].astype(np.int64) for c in range(num_colors)]


def validate_coloring(graph_edges: np.ndarray, particle_colors: np.ndarray) -> bool:
    """True when no graph edge joins two vertices of one color (``graph_coloring.py:26-31``)."""
    e = np.asarray(graph_edges).reshape(-1, 2)
    c = np.asarray(particle_colors)
    return bool(np.all(c[e[:, 0]] != c[e[:, 1]])) if e.size else True


def combine_independent_coloring_plan(sized_groups_1, sized_groups_2) -> list:
    """Pair the groups of two independent colorings (``graph_coloring.py:382-410``).

    Items are ``(size, chunks)``; the longer list is sorted ascending by size,
    the other descending, and groups with equal rank are merged (first
    coloring's chunks first).
    """
    if len(sized_groups_1) == 0:
        return sized_groups_2
    if len(sized_groups_2) == 0:
        return sized_groups_1
    if len(sized_groups_1) < len(sized_groups_2):
        sized_groups_1, sized_groups_2 = sized_groups_2, sized_groups_1
    groups_1 = sorted(sized_groups_1, key=lambda group: group[0])
    groups_2 = sorted(sized_groups_2, key=lambda group: -group[0])
    combined = []
    for i, (size_1, chunks_1) in enumerate(groups_1):
        if i < len(groups_2):
            size_2, chunks_2 = groups_2[i]
            combined.append((size_1 + size_2, chunks_1 + chunks_2))
        else:
            combined.append((size_1, chunks_1))
    return combined


def combine_independent_particle_coloring(color_groups_1, color_groups_2) -> list[np.ndarray]:
    """Merged groups of two independent colorings (``graph_coloring.py:413-431``)."""
    plan = combine_independent_coloring_plan(
        [(len(g), [g]) for g in color_groups_1], [(len(g), [g]) for g in color_groups_2]
    )
    return [chunks[0] if len(chunks) == 1 else np.concatenate(chunks) for _, chunks in plan]


__all__ = [
    "ColoringAlgorithm",
    "color_graph",
    "combine_independent_coloring_plan",
    "combine_independent_particle_coloring",
    "construct_particle_graph",
    "construct_tetmesh_graph_edges",
    "construct_trimesh_graph_edges",
    "convert_to_color_groups",
    "validate_coloring",
]
