"""Soft-mesh adjacency: edge numbering, edge/triangle maps and per-vertex CSR.

Reproduces the integer results of the reference's ``newton/_src/utils/mesh.py``
bit for bit, with linear-time C++ routines (``libvbd_host.so``) so that
multi-million-vertex batches build in seconds:

* edge numbering = order of first occurrence while walking each triangle's
  local edges (v0,v1), (v1,v2), (v2,v0); row ``[o0, o1, v0, v1]`` takes
  ``o0, v0, v1`` from the first triangle and ``o1`` from the second
  (``mesh.py:464-529``), numbered in C++;
* edge/triangle maps against pre-numbered edges (``mesh.py:531-576``), in C++;
* per-vertex CSR of ``(element, slot)`` int32 pairs whose order is the serial
  fill order of the reference's count/fill kernels (``mesh.py:674-857``):
  elements in increasing id; within one bending edge the fill order is
  v0 (slot 2), v1 (slot 3), o0 (slot 0), o1 (slot 1).
"""

from __future__ import annotations

import warnings
from dataclasses import dataclass

import numpy as np


def _int_rows(data, width: int) -> np.ndarray:
    if data is None:
        return np.empty((0, width), dtype=np.int32)
    if hasattr(data, "detach"):
        data = data.detach().cpu().numpy()
    arr = np.asarray(data, dtype=np.int32)
    if arr.size == 0:
        return np.empty((0, width), dtype=np.int32)
    return arr.reshape(-1, width)


def compute_edge_adjacency(tri_indices, tri_start: int = 0):
    """Edge table, edge->triangle and triangle->edge maps from triangles (``mesh.py:464-529``), in C++
    (``vbd_host_number_edges``: sides bucketed by their smaller vertex, numbered in one walk).

    Returns ``(edge_indices [E,4], edge_tri_indices [E,2], tri_edge_indices [T,3])``; warns once if
    a vertex pair is shared by more than two triangle sides (the extra sides are ignored).
    """
    import ctypes  # noqa: PLC0415

    from .. import _native  # noqa: PLC0415

    tris = np.ascontiguousarray(_int_rows(tri_indices, 3))
    tri_count = tris.shape[0]
    edges = np.empty((max(3 * tri_count, 1), 4), dtype=np.int32)
    edge_tri = np.empty((max(3 * tri_count, 1), 2), dtype=np.int32)
    tri_edge = np.empty((tri_count, 3), dtype=np.int32)
    n_edges = ctypes.c_int64(0)
    n_extra = ctypes.c_int64(0)
    rc = _native.host_lib().vbd_host_number_edges(tri_count, tris.ctypes.data, int(tri_start), edges.ctypes.data,
                                                  edge_tri.ctypes.data, tri_edge.ctypes.data, ctypes.byref(n_edges),
                                                  ctypes.byref(n_extra))
    if rc == -2:
        raise ValueError("triangle indices must be non-negative")
    if rc != 0:
        raise ValueError(f"vbd_host_number_edges failed ({rc})")
    if n_extra.value:
        warnings.warn("Detected non-manifold edge", stacklevel=2)
    e = n_edges.value
    return edges[:e].copy(), edge_tri[:e].copy(), tri_edge


def _vertex_csr_np(rows: np.ndarray, slots_in_fill_order, particle_count: int, values_per_entry: int = 2):
    """CSR of (element, slot) pairs in the serial count/fill order (``mesh.py:674-926``), numpy restatement.

    ``slots_in_fill_order`` lists the element columns in the order the fill
    kernel visits them; the stored slot is the column index.
    """
    n_elem = rows.shape[0]
    if n_elem == 0:
        return np.empty(0, dtype=np.int32), np.empty(0, dtype=np.int32)
    cols = np.asarray(slots_in_fill_order, dtype=np.int32)
    verts = rows[:, cols].reshape(-1)  # element-major, fill order within element
    elems = np.repeat(np.arange(n_elem, dtype=np.int32), len(cols))
    slots = np.tile(cols, n_elem)
    valid = verts != -1
    verts, elems, slots = verts[valid], elems[valid], slots[valid]
    if np.any(verts < 0) or np.any(verts >= particle_count):
        raise ValueError("element references a particle index out of range")
    order = np.argsort(verts, kind="stable")
    counts = np.bincount(verts, minlength=particle_count).astype(np.int64)
    offsets = np.empty(particle_count + 1, dtype=np.int32)
    offsets[0] = 0
    offsets[1:] = np.cumsum(values_per_entry * counts)
    if values_per_entry == 1:
        return elems[order].astype(np.int32), offsets
    values = np.empty(2 * len(order), dtype=np.int32)
    values[0::2] = elems[order]
    values[1::2] = slots[order]
    return values, offsets


def build_edge_tri_maps(edge_indices, tri_indices):
    """Edge/triangle maps for pre-numbered edges (``mesh.py:531-576``), in C++ (``vbd_host_edge_tri_maps``)."""
    from .. import _native  # noqa: PLC0415

    edges = np.ascontiguousarray(_int_rows(edge_indices, 4))
    tris = np.ascontiguousarray(_int_rows(tri_indices, 3))
    edge_tri = np.empty((edges.shape[0], 2), dtype=np.int32)
    tri_edge = np.empty((tris.shape[0], 3), dtype=np.int32)
    rc = _native.host_lib().vbd_host_edge_tri_maps(edges.shape[0], edges.ctypes.data, tris.shape[0], tris.ctypes.data,
                                                   edge_tri.ctypes.data, tri_edge.ctypes.data)
    if rc != 0:
        raise ValueError(f"vbd_host_edge_tri_maps failed ({rc})")
    return edge_tri, tri_edge


def _vertex_csr(rows: np.ndarray, slots_in_fill_order, particle_count: int, values_per_entry: int = 2):
    """CSR of (element, slot) pairs in the serial count/fill order (``mesh.py:674-926``), in C++."""
    from .. import _native  # noqa: PLC0415

    rows = np.ascontiguousarray(rows, dtype=np.int32)
    if rows.shape[0] == 0:
        return np.empty(0, dtype=np.int32), np.empty(0, dtype=np.int32)
    cols = np.ascontiguousarray(slots_in_fill_order, dtype=np.int32)
    n_valid = int(np.count_nonzero(rows[:, cols] != -1))
    offsets = np.empty(particle_count + 1, dtype=np.int32)
    values = np.empty(max(values_per_entry * n_valid, 1), dtype=np.int32)
    rc = _native.host_lib().vbd_host_vertex_csr(rows.shape[0], rows.shape[1], rows.ctypes.data, len(cols),
                                                cols.ctypes.data, particle_count, values_per_entry,
                                                offsets.ctypes.data, values.ctypes.data, values.size)
    if rc == -2:
        raise ValueError("element references a particle index out of range")
    if rc != 0:
        raise ValueError(f"vbd_host_vertex_csr failed ({rc})")
    return values[: values_per_entry * n_valid], offsets


@dataclass
class MeshAdjacencyData:
    """Device-facing adjacency arrays (``mesh.py:183-201``), as torch int32 tensors."""

    edge_tri_indices: object = None
    tri_edge_indices: object = None
    v_adj_tris: object = None
    v_adj_tris_offsets: object = None
    v_adj_edges: object = None
    v_adj_edges_offsets: object = None
    v_adj_springs: object = None
    v_adj_springs_offsets: object = None
    v_adj_tets: object = None
    v_adj_tets_offsets: object = None


class MeshAdjacency:
    """Connectivity tables of a soft mesh (``mesh.py:248-654``).

    ``MeshAdjacency(tri_indices, edge_indices=None, spring_indices=None, tet_indices=None)``
    keeps a caller's edge numbering when ``edge_indices`` is given and only
    derives the edge/triangle maps; otherwise it numbers the edges itself.
    """

    def __init__(self, tri_indices=None, edge_indices=None, spring_indices=None, tet_indices=None):
        self.indices = None if tri_indices is None else _int_rows(tri_indices, 3).copy()
        self.spring_indices = None if spring_indices is None else np.array(spring_indices, dtype=np.int32).reshape(-1)
        self.tet_indices = None if tet_indices is None else _int_rows(tet_indices, 4).copy()
        if edge_indices is not None:
            self.edge_indices = _int_rows(edge_indices, 4).copy()
            tris = self.indices if self.indices is not None else np.empty((0, 3), dtype=np.int32)
            self.edge_tri_indices, self.tri_edge_indices = build_edge_tri_maps(self.edge_indices, tris)
        elif tri_indices is not None:
            self.edge_indices, self.edge_tri_indices, self.tri_edge_indices = compute_edge_adjacency(self.indices)
        else:
            self.edge_indices = np.empty((0, 4), dtype=np.int32)
            self.edge_tri_indices = np.empty((0, 2), dtype=np.int32)
            self.tri_edge_indices = np.empty((0, 3), dtype=np.int32)
        self.v_adj_tris = self.v_adj_tris_offsets = None
        self.v_adj_edges = self.v_adj_edges_offsets = None
        self.v_adj_springs = self.v_adj_springs_offsets = None
        self.v_adj_tets = self.v_adj_tets_offsets = None
        self.vertex_adjacency_initialized = False

    def init_vertex_adjacency(self, particle_count: int) -> "MeshAdjacency":
        """Per-vertex CSR tables (``mesh.py:595-654``); idempotent."""
        if self.vertex_adjacency_initialized:
            return self
        empty = (np.empty(0, dtype=np.int32), np.empty(0, dtype=np.int32))
        if self.edge_indices is not None and self.edge_indices.size:
            # fill order per edge: v0 (slot 2), v1 (slot 3), o0 (slot 0), o1 (slot 1)
            self.v_adj_edges, self.v_adj_edges_offsets = _vertex_csr(self.edge_indices, [2, 3, 0, 1], particle_count)
        else:
            self.v_adj_edges, self.v_adj_edges_offsets = empty
        if self.indices is not None and self.indices.size:
            self.v_adj_tris, self.v_adj_tris_offsets = _vertex_csr(self.indices, [0, 1, 2], particle_count)
        else:
            self.v_adj_tris, self.v_adj_tris_offsets = empty
        if self.tet_indices is not None and self.tet_indices.size:
            self.v_adj_tets, self.v_adj_tets_offsets = _vertex_csr(self.tet_indices, [0, 1, 2, 3], particle_count)
        else:
            self.v_adj_tets, self.v_adj_tets_offsets = empty
        if self.spring_indices is not None and self.spring_indices.size:
            self.v_adj_springs, self.v_adj_springs_offsets = _vertex_csr(
                self.spring_indices.reshape(-1, 2), [0, 1], particle_count, values_per_entry=1
            )
        else:
            self.v_adj_springs, self.v_adj_springs_offsets = empty
        self.vertex_adjacency_initialized = True
        return self

    def to(self, device) -> MeshAdjacencyData:
        """Upload the adjacency to ``device`` as torch int32 tensors (``mesh.py:424-462``)."""
        import torch  # noqa: PLC0415

        def t(a):
            return torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32), device=device)

        data = MeshAdjacencyData(edge_tri_indices=t(self.edge_tri_indices), tri_edge_indices=t(self.tri_edge_indices))
        if not self.vertex_adjacency_initialized:
            warnings.warn("MeshAdjacency.to(): vertex adjacency not initialized", stacklevel=2)
            return data
        data.v_adj_tris = t(self.v_adj_tris)
        data.v_adj_tris_offsets = t(self.v_adj_tris_offsets)
        data.v_adj_edges = t(self.v_adj_edges)
        data.v_adj_edges_offsets = t(self.v_adj_edges_offsets)
        data.v_adj_springs = t(self.v_adj_springs)
        data.v_adj_springs_offsets = t(self.v_adj_springs_offsets)
        data.v_adj_tets = t(self.v_adj_tets)
        data.v_adj_tets_offsets = t(self.v_adj_tets_offsets)
        return data


def _csr_offsets_or_zeros(offsets: np.ndarray, particle_count: int) -> np.ndarray:
    """Offsets array of length N+1 (an empty table means zero entries for every vertex)."""
    if offsets is None or offsets.size == 0:
        return np.zeros(particle_count + 1, dtype=np.int32)
    return offsets


__all__ = ["MeshAdjacency", "MeshAdjacencyData", "build_edge_tri_maps", "compute_edge_adjacency"]
