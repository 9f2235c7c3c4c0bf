"""Cloth self-contact: topological contact filters and the collision-detector front end.

The filters follow ``newton/_src/solvers/vbd/tri_mesh_collision.py:106-233``: a
vertex ignores the triangles of the vertices within ``n - 1`` edge rings of it
(its own triangles are never candidates anyway), an edge ignores the edges having
an endpoint within ``n - 1`` rings of either of its endpoints, except edges that
touch its endpoints in any role.  The reference walks Python sets per vertex; here
the same sets come from sparse boolean products (ring reachability R = (A + I)^(n-1)
without the diagonal, A the endpoint adjacency of the edges), so million-vertex
meshes build in seconds.  Lists are sorted per row (``set_to_csr``), which the
detection kernels binary-search.
"""

from __future__ import annotations

import numpy as np

try:
    import scipy.sparse as sp
except ImportError:  # pragma: no cover - scipy is part of the image
    sp = None


def set_to_csr(list_of_sets, dtype=np.int32):
    """Per-row integer sets to (flat sorted values, offsets) (``tri_mesh_collision.py:106-120``)."""
    offsets = np.zeros(len(list_of_sets) + 1, dtype=dtype)
    offsets[1:] = np.cumsum([len(s) for s in list_of_sets])
    flat = np.empty(int(offsets[-1]), dtype=dtype)
    for i, row in enumerate(list_of_sets):
        flat[offsets[i]:offsets[i + 1]] = np.fromiter(sorted(row), count=len(row), dtype=dtype)
    return flat, offsets


def _csr_pairs(values, offsets, n_rows):
    """(row, element, slot) triples of a per-vertex (element, slot) CSR."""
    values = np.asarray(values, dtype=np.int64).reshape(-1)
    offsets = np.asarray(offsets, dtype=np.int64)
    if values.size == 0:
        z = np.zeros(0, dtype=np.int64)
        return z, z, z
    counts = (offsets[1:] - offsets[:-1]) // 2
    rows = np.repeat(np.arange(n_rows, dtype=np.int64), counts)
    return rows, values[0::2], values[1::2]


def _ring_reach(n_rings, particle_count, edge_indices):
    """Boolean sparse R[v, u]: u != v within n_rings edge rings of v (edges by their v0, v1)."""
    e = np.asarray(edge_indices, dtype=np.int64).reshape(-1, 4)
    v0, v1 = e[:, 2], e[:, 3]
    ok = (v0 >= 0) & (v1 >= 0) & (v0 != v1)
    rows = np.concatenate([v0[ok], v1[ok]])
    cols = np.concatenate([v1[ok], v0[ok]])
    A = sp.csr_matrix((np.ones(rows.size, dtype=np.int8), (rows, cols)), shape=(particle_count, particle_count))
    A.data[:] = 1
    step = (A + sp.identity(particle_count, dtype=np.int8, format="csr")).astype(bool).astype(np.int32)
    R = step.copy()
    for _ in range(n_rings - 1):
        R = (R @ step).astype(bool).astype(np.int32)
    R = R.tolil()
    R.setdiag(0)
    R = R.tocsr()
    R.eliminate_zeros()
    return R.astype(bool)


def build_vertex_n_ring_tris_collision_filter(n, particle_count, edge_indices, v_adj_edges, v_adj_edges_offsets,
                                              v_adj_tris, v_adj_tris_offsets):
    """Per-vertex sorted triangle filter as (flat, offsets), or None when n <= 1
    (``tri_mesh_collision.py:163-190``)."""
    if n <= 1:
        return None
    N = int(particle_count)
    R = _ring_reach(n - 1, N, edge_indices)
    vr, tri, _ = _csr_pairs(v_adj_tris, v_adj_tris_offsets, N)
    T = int(tri.max()) + 1 if tri.size else 0
    B = sp.csr_matrix((np.ones(vr.size, dtype=np.int32), (vr, tri)), shape=(N, max(T, 1))).astype(bool)
    F = (R.astype(np.int32) @ B.astype(np.int32)).astype(bool)
    F = (F.astype(np.int8) - F.multiply(B).astype(np.int8)).astype(bool)
    F = F.tocsr()
    F.sort_indices()
    return F.indices.astype(np.int32), F.indptr.astype(np.int32)


def build_edge_n_ring_edge_collision_filter(n, edge_indices, v_adj_edges, v_adj_edges_offsets):
    """Per-edge sorted edge filter as (flat, offsets), or None when n <= 1
    (``tri_mesh_collision.py:193-232``)."""
    if n <= 1:
        return None
    e = np.asarray(edge_indices, dtype=np.int64).reshape(-1, 4)
    E = e.shape[0]
    N = int(max(int(e[:, 2:].max()) + 1 if E else 0, np.asarray(v_adj_edges_offsets).size - 1))
    R = _ring_reach(n - 1, N, edge_indices).astype(np.int32)
    v0, v1 = e[:, 2], e[:, 3]
    S0 = sp.csr_matrix((np.ones(E, dtype=np.int32), (np.arange(E), v0)), shape=(E, N))
    S1 = sp.csr_matrix((np.ones(E, dtype=np.int32), (np.arange(E), v1)), shape=(E, N))
    # neighbours of either endpoint, minus the endpoints themselves
    C = ((S0 @ R) + (S1 @ R)).astype(bool).astype(np.int32)
    C = C - C.multiply(S0 + S1).astype(bool).astype(np.int32)
    C.eliminate_zeros()
    vr, ed, slot = _csr_pairs(v_adj_edges, v_adj_edges_offsets, N)
    end = slot >= 2
    D = sp.csr_matrix((np.ones(int(end.sum()), dtype=np.int32), (vr[end], ed[end])), shape=(N, E))
    Adj = sp.csr_matrix((np.ones(vr.size, dtype=np.int32), (vr, ed)), shape=(N, E)).astype(bool).astype(np.int32)
    F = (C @ D).astype(bool)
    touching = ((S0 @ Adj) + (S1 @ Adj)).astype(bool)
    F = (F.astype(np.int8) - F.multiply(touching).astype(np.int8)).astype(bool).tocsr()
    F.eliminate_zeros()
    F.sort_indices()
    return F.indices.astype(np.int32), F.indptr.astype(np.int32)


def merge_filter(csr, external_map, rows):
    """Union of a (flat, offsets) filter and an external {row: iterable} map, as sorted CSR."""
    if external_map is None:
        return csr
    sets = [set() for _ in range(rows)]
    if csr is not None:
        flat, offs = csr
        for i in range(rows):
            sets[i].update(int(x) for x in flat[offs[i]:offs[i + 1]])
    for k, vals in external_map.items():
        sets[int(k)].update(int(x) for x in vals)
    return set_to_csr(sets)


def build_filters(model, threshold: int, external_vertex_map=None, external_edge_map=None):
    """The detector's vertex-triangle and edge-edge filters for a model (``tri_mesh_collision.py:289-313``)."""
    adj = model.soft_mesh_adjacency
    vt = ee = None
    if model.tri_count > 0:
        if threshold >= 2 and model.edge_count > 0:
            e = model.edge_indices.cpu().numpy()
            vt = build_vertex_n_ring_tris_collision_filter(threshold, model.particle_count, e, adj.v_adj_edges,
                                                           adj.v_adj_edges_offsets, adj.v_adj_tris,
                                                           adj.v_adj_tris_offsets)
            ee = build_edge_n_ring_edge_collision_filter(threshold, e, adj.v_adj_edges, adj.v_adj_edges_offsets)
        vt = merge_filter(vt, external_vertex_map, model.particle_count)
        ee = merge_filter(ee, external_edge_map, model.edge_count)
    return vt, ee


__all__ = ["build_edge_n_ring_edge_collision_filter", "build_filters", "build_vertex_n_ring_tris_collision_filter",
           "merge_filter", "set_to_csr"]
