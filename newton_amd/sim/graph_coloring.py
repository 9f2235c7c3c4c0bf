"""Particle graph construction, coloring and per-world color merging.

Graph construction gives the reference's graph exactly
(``newton/_src/sim/graph_coloring.py:45-240``): canonical ``(min, max)``
unique edges from active triangles, all six vertex pairs of active bending
hinges (when ``include_bending``), and all six pairs of active tetrahedra,
built here from int64 pair keys.
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


def _pairs_to_graph_edges(pairs: np.ndarray) -> np.ndarray:
    """Undirected graph edges from vertex pairs: each pair as ``(min, max)``, duplicates and pairs
    with a negative vertex dropped, rows in ascending ``(min, max)`` order -- the canonical form the
    reference's ``_canonicalize_edges_np`` produces (``graph_coloring.py:45-51``).

    Each pair is encoded as one int64 key ``min * 2^32 + max``; one ``np.unique`` over the keys sorts
    and de-duplicates."""
    pairs = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    pairs = pairs[(pairs >= 0).all(axis=1)]
    if pairs.shape[0] == 0:
        return np.empty((0, 2), dtype=np.int32)
    keys = np.unique((pairs.min(axis=1) << 32) | pairs.max(axis=1))
    out = np.empty((keys.size, 2), dtype=np.int32)
    out[:, 0] = keys >> 32
    out[:, 1] = keys & 0xFFFFFFFF
    return out


def _element_pairs(rows: np.ndarray) -> np.ndarray:
    """Every pair of columns of every element row, as an (M, 2) array (order irrelevant: the caller
    canonicalises)."""
    k = rows.shape[1]
    return np.concatenate([rows[:, [a, b]] for a in range(k) for b in range(a + 1, k)], axis=0)


def _active_rows(indices, width: int, mask) -> np.ndarray:
    rows = np.asarray(indices, dtype=np.int32).reshape(-1, width)
    if mask is None:
        return rows
    m = np.asarray(mask, dtype=bool)
    if m.ndim == 0:
        return rows if bool(m) else rows[:0]
    return rows[m]


def construct_tetmesh_graph_edges(tet_indices, tet_active_mask=None) -> np.ndarray:
    """Unique graph edges of active tetrahedra: all six vertex pairs (``graph_coloring.py:54-96``)."""
    if tet_indices is None:
        return np.empty((0, 2), dtype=np.int32)
    tets = _active_rows(tet_indices, 4, tet_active_mask)
    return _pairs_to_graph_edges(_element_pairs(tets)) if tets.size else np.empty((0, 2), dtype=np.int32)


def construct_trimesh_graph_edges(
    tri_indices, tri_active_mask=None, bending_edge_indices=None, bending_edge_active_mask=None
) -> np.ndarray:
    """Unique graph edges of active triangles (three sides) plus all six vertex pairs of active bending
    hinges (``graph_coloring.py:99-197``); hinge pairs with a missing (-1) vertex are dropped."""
    parts = [np.empty((0, 2), dtype=np.int32)]
    if tri_indices is not None:
        tris = _active_rows(tri_indices, 3, tri_active_mask)
        if tris.size:
            parts.append(_element_pairs(tris))
    if bending_edge_indices is not None:
        hinges = _active_rows(bending_edge_indices, 4, bending_edge_active_mask)
        if hinges.size:
            parts.append(_element_pairs(hinges))
    return _pairs_to_graph_edges(np.concatenate(parts, axis=0))


def construct_particle_graph(
    tri_indices, tri_active_mask, bending_edge_indices, bending_edge_active_mask, tet_indices, tet_active_mask
) -> np.ndarray:
    """Unified particle graph (``graph_coloring.py:200-240``)."""
    tri_edges = construct_trimesh_graph_edges(tri_indices, tri_active_mask, bending_edge_indices, bending_edge_active_mask)
    tet_edges = construct_tetmesh_graph_edges(tet_indices, tet_active_mask)
    return _pairs_to_graph_edges(np.vstack([tri_edges, tet_edges]))


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
