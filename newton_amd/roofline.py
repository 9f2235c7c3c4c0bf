"""Algorithmic HBM bytes of the VBD color pass, computed exactly from a model and its coloring.

Per color pass c every distinct datum the pass needs is counted once (4-byte
FP32 / INT32 words):

* per vertex of the color: list id 4 + flags 4 + mass 4; for an active vertex
  additionally x 12 + y 12 + x0 12 + displacement read/write 24 + x write 12,
  8 bytes of CSR offsets per element type present, 8 per CSR entry, and x_prev
  12 when any of its elements or contacts needs the previous position;
* per element touched by the pass: triangle 44 (indices 12, pose 16, mu/lambda/kd 12, area 4),
  bending edge 32 (indices 16, rest angle 4, rest length 4, ke/kd 8),
  tetrahedron 64 (indices 16, Dm^-1 36, mu/lambda/kd 12);
* per distinct vertex outside the pass whose position an element reads: 12,
  plus 12 for its previous position when that element is damped;
* per contact record of the pass: 64 (penalty, kd, mu, shape, body position,
  normal, indices) plus 8 of per-particle list offsets;
* per spring with an active endpoint in the color: 20 (indices 8, rest length 4,
  ke/kd 8); its other endpoint counts as a neighbour read (12, plus 12 for both
  endpoints' previous positions when kd > 0).
"""

from __future__ import annotations

import numpy as np


def _csr(offsets, values, n):
    if offsets is None or len(offsets) == 0:
        return np.zeros(n + 1, dtype=np.int64), np.zeros(0, dtype=np.int64)
    return np.asarray(offsets, dtype=np.int64), np.asarray(values, dtype=np.int64)


def color_pass_bytes(model, color_groups=None, contacts_per_particle: np.ndarray | None = None) -> np.ndarray:
    """Algorithmic bytes of each color pass (one entry per color)."""
    n = model.particle_count
    adj = model.soft_mesh_adjacency
    flags = model.particle_flags.cpu().numpy()
    mass = model.particle_mass.cpu().numpy()
    active = ((flags & 1) != 0) & (mass != 0.0)
    tri_idx = model.tri_indices.cpu().numpy().reshape(-1, 3)
    tri_mat = model.tri_materials.cpu().numpy().reshape(-1, 5)
    edge_idx = model.edge_indices.cpu().numpy().reshape(-1, 4)
    edge_bp = model.edge_bending_properties.cpu().numpy().reshape(-1, 2)
    tet_idx = model.tet_indices.cpu().numpy().reshape(-1, 4)
    tet_mat = model.tet_materials.cpu().numpy().reshape(-1, 3)
    groups = color_groups if color_groups is not None else [g.cpu().numpy() for g in model.particle_color_groups]
    cpp = np.zeros(n, dtype=np.int64) if contacts_per_particle is None else np.asarray(contacts_per_particle)
    kinds = []
    for key, elem_bytes, idx, damped in (
        ("tris", 44, tri_idx, tri_mat[:, 2] > 0 if len(tri_mat) else np.zeros(0, bool)),
        ("edges", 32, edge_idx, edge_bp[:, 1] > 0 if len(edge_bp) else np.zeros(0, bool)),
        ("tets", 64, tet_idx, tet_mat[:, 2] > 0 if len(tet_mat) else np.zeros(0, bool)),
    ):
        off, val = _csr(getattr(adj, f"v_adj_{key}_offsets"), getattr(adj, f"v_adj_{key}"), n)
        kinds.append((off, val, elem_bytes, idx, damped, len(idx) > 0))
    S = int(getattr(model, "spring_count", 0) or 0)
    if S:
        sp_idx = model.spring_indices.cpu().numpy().reshape(-1, 2).astype(np.int64)
        sp_damped = model.spring_damping.cpu().numpy() > 0
    out = np.zeros(len(groups), dtype=np.int64)
    for c, g in enumerate(groups):
        g = np.asarray(g, dtype=np.int64)
        in_color = np.zeros(n, dtype=bool)
        in_color[g] = True
        act = g[active[g]]
        b = 12 * len(g)  # list id, flags, mass
        b += (12 + 12 + 12 + 24 + 12) * len(act)
        needs_prev = np.zeros(n, dtype=bool)
        neighbor = np.zeros(n, dtype=bool)
        neighbor_prev = np.zeros(n, dtype=bool)
        for off, val, elem_bytes, idx, damped, present in kinds:
            if not present:
                continue
            b += 8 * len(act)  # CSR offsets
            counts = (off[act + 1] - off[act]) >> 1
            b += 8 * int(counts.sum())
            if counts.sum() == 0:
                continue
            starts = off[act]
            pos = np.repeat(starts, counts) + 2 * (np.arange(counts.sum()) - np.repeat(np.cumsum(counts) - counts, counts))
            elems = np.unique(val[pos])
            b += elem_bytes * len(elems)
            verts = idx[elems]
            verts_flat = verts[verts >= 0]
            neighbor[verts_flat] = True
            if damped.any():
                dv = idx[elems[damped[elems]]]
                dv = dv[dv >= 0]
                neighbor_prev[dv] = True
                needs_prev[dv] = True
        if S:
            act_in = in_color & active
            touched = act_in[sp_idx[:, 0]] | act_in[sp_idx[:, 1]]
            b += 20 * int(touched.sum())
            ends = sp_idx[touched].reshape(-1)
            neighbor[ends] = True
            dends = sp_idx[touched & sp_damped].reshape(-1)
            neighbor_prev[dends] = True
            needs_prev[dends] = True
        own_prev = needs_prev & in_color
        own_prev[act[cpp[act] > 0]] = True
        b += 12 * int(own_prev.sum())
        outside = neighbor & ~in_color
        b += 12 * int(outside.sum()) + 12 * int((neighbor_prev & ~in_color).sum())
        b += int((64 * cpp[act]).sum()) + 8 * len(act)
        out[c] = b
    return out


__all__ = ["color_pass_bytes"]
