"""Edge numbering, edge/triangle maps and per-vertex CSR adjacency (bit-exact integer pipelines).

Known answers come from the reference's own tests (test_model.py:532-648,
test_softbody.py:190-359); the rest is pinned against a literal serial
restatement of the reference's count/fill kernels (mesh.py:674-857) and of
its edge-numbering loop (mesh.py:464-529) on random meshes.
"""

import warnings

import numpy as np
import pytest

from newton_amd import ModelBuilder
from newton_amd.utils.mesh import MeshAdjacency, compute_edge_adjacency

QUAT_IDENTITY = (0.0, 0.0, 0.0, 1.0)
SQUARE = [(0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (1.0, 1.0, 0.0), (0.0, 1.0, 0.0)]

# test_softbody.py:192-238
PYRAMID_TET_INDICES = np.array(
    [[0, 1, 3, 9], [1, 4, 3, 13], [1, 3, 9, 13], [3, 9, 13, 12], [1, 9, 10, 13], [1, 2, 4, 10], [2, 5, 4, 14],
     [2, 4, 10, 14], [4, 10, 14, 13], [2, 10, 11, 14], [3, 4, 6, 12], [4, 7, 6, 16], [4, 6, 12, 16],
     [6, 12, 16, 15], [4, 12, 13, 16], [4, 5, 7, 13], [5, 8, 7, 17], [5, 7, 13, 17], [7, 13, 17, 16],
     [5, 13, 14, 17]], dtype=np.int32)
PYRAMID_PARTICLES = [(0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (2.0, 0.0, 0.0), (0.0, 1.0, 0.0), (1.0, 1.0, 0.0),
                     (2.0, 1.0, 0.0), (0.0, 2.0, 0.0), (1.0, 2.0, 0.0), (2.0, 2.0, 0.0), (0.0, 0.0, 1.0),
                     (1.0, 0.0, 1.0), (2.0, 0.0, 1.0), (0.0, 1.0, 1.0), (1.0, 1.0, 1.0), (2.0, 1.0, 1.0),
                     (0.0, 2.0, 1.0), (1.0, 2.0, 1.0), (2.0, 2.0, 1.0)]


def _square_builder():
    b = ModelBuilder()
    b.add_cloth_mesh(pos=(0.0, 0.0, 0.0), rot=QUAT_IDENTITY, scale=1.0, vel=(0.0, 0.0, 0.0), vertices=SQUARE,
                     indices=[0, 1, 2, 0, 2, 3], density=1.0)
    return b


def test_cloth_mesh_edge_maps_known_answer():
    """test_model.py:532-560."""
    adj = _square_builder().finalize(device="cpu").soft_mesh_adjacency
    np.testing.assert_array_equal(adj.tri_edge_indices, [[0, 1, 2], [2, 3, 4]])
    np.testing.assert_array_equal(adj.edge_tri_indices, [[0, -1], [0, -1], [0, 1], [1, -1], [1, -1]])
    assert adj.v_adj_tris is not None and adj.v_adj_edges_offsets is not None


def test_placeholder_edge_stays_unlinked():
    """test_model.py:562-575: a bare triangle and an (o0 == o1 == -1) edge are not linked."""
    b = ModelBuilder()
    b.add_particle((0.0, 0.0, 0.0), (0.0, 0.0, 0.0), 1.0)
    b.add_particle((1.0, 0.0, 0.0), (0.0, 0.0, 0.0), 1.0)
    b.add_particle((0.0, 1.0, 0.0), (0.0, 0.0, 0.0), 1.0)
    b.add_triangle(0, 1, 2)
    b.add_edge(-1, -1, 0, 1)
    adj = b.finalize(device="cpu").soft_mesh_adjacency
    np.testing.assert_array_equal(adj.tri_edge_indices, [[-1, -1, -1]])
    np.testing.assert_array_equal(adj.edge_tri_indices, [[-1, -1]])


def test_add_builder_offsets_known_answer():
    """test_model.py:577-605."""
    base = _square_builder()
    combined = ModelBuilder()
    combined.add_builder(base)
    combined.add_builder(base)
    base_adj = base.finalize(device="cpu").soft_mesh_adjacency
    comb = combined.finalize(device="cpu").soft_mesh_adjacency
    np.testing.assert_array_equal(comb.tri_edge_indices[:2], base_adj.tri_edge_indices)
    np.testing.assert_array_equal(comb.edge_tri_indices[:5], base_adj.edge_tri_indices)
    np.testing.assert_array_equal(comb.tri_edge_indices[2:], base_adj.tri_edge_indices + 5)
    np.testing.assert_array_equal(comb.edge_tri_indices[5:], [[2, -1], [2, -1], [2, 3], [3, -1], [3, -1]])


def test_cloth_plus_bare_triangle_known_answer():
    """test_model.py:607-640."""
    b = _square_builder()
    p0 = b.add_particle((2.0, 0.0, 0.0), (0.0, 0.0, 0.0), 1.0)
    p1 = b.add_particle((3.0, 0.0, 0.0), (0.0, 0.0, 0.0), 1.0)
    p2 = b.add_particle((2.0, 1.0, 0.0), (0.0, 0.0, 0.0), 1.0)
    b.add_triangle(p0, p1, p2)
    adj = b.finalize(device="cpu").soft_mesh_adjacency
    np.testing.assert_array_equal(adj.tri_edge_indices, [[0, 1, 2], [2, 3, 4], [-1, -1, -1]])
    np.testing.assert_array_equal(adj.edge_tri_indices, [[0, -1], [0, -1], [0, 1], [1, -1], [1, -1]])


def _expected_tet_adjacency(particle_count, tet_indices):
    """Closed form of test_softbody.py:264-276: (tet_id, local_order) appended per vertex in tet order."""
    buckets = [[] for _ in range(particle_count)]
    for tet_id, tet in enumerate(tet_indices):
        for local_order, v in enumerate(tet):
            buckets[int(v)].append((tet_id, local_order))
    offsets, flat = [0], []
    for bkt in buckets:
        offsets.append(offsets[-1] + 2 * len(bkt))
        for tet_id, order in bkt:
            flat.extend([tet_id, order])
    return np.array(offsets, dtype=np.int32), np.array(flat, dtype=np.int32)


@pytest.mark.parametrize("case", ["single", "pyramid"])
def test_tet_adjacency_known_answer(case):
    """test_softbody.py:329-359 on the single-tet and pyramid fixtures."""
    if case == "single":
        tets = np.array([[0, 1, 2, 3]], dtype=np.int32)
        parts = [(0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0)]
    else:
        tets, parts = PYRAMID_TET_INDICES, PYRAMID_PARTICLES
    b = ModelBuilder()
    with warnings.catch_warnings():
        warnings.filterwarnings("ignore", message="Detected non-manifold edge")
        b.add_soft_mesh(pos=(0.0, 0.0, 0.0), rot=QUAT_IDENTITY, scale=1.0, vel=(0.0, 0.0, 0.0), vertices=parts,
                        indices=tets.flatten().tolist(), density=1.0, k_mu=1.0, k_lambda=1.0, k_damp=0.0)
    b.color()
    adj = b.finalize(device="cpu").soft_mesh_adjacency
    exp_off, exp_flat = _expected_tet_adjacency(len(parts), tets)
    np.testing.assert_array_equal(adj.v_adj_tets_offsets, exp_off)
    np.testing.assert_array_equal(adj.v_adj_tets, exp_flat)
    for v in range(len(parts)):
        entries = adj.v_adj_tets[adj.v_adj_tets_offsets[v]:adj.v_adj_tets_offsets[v + 1]].reshape(-1, 2)
        for tet_id, order in entries:
            assert tets[tet_id, order] == v


def _serial_csr(elements, order, n):
    """Literal restatement of the reference's serial count/fill kernels (mesh.py:674-857)."""
    counts = [0] * n
    for row in elements:
        for col in order:
            v = int(row[col])
            if v != -1:
                counts[v] += 1
    offsets = [0]
    for c in counts:
        offsets.append(offsets[-1] + 2 * c)
    values = [0] * offsets[-1]
    fill = [0] * n
    for eid, row in enumerate(elements):
        for col in order:
            v = int(row[col])
            if v == -1:
                continue
            values[offsets[v] + 2 * fill[v]] = eid
            values[offsets[v] + 2 * fill[v] + 1] = col
            fill[v] += 1
    return np.array(values, dtype=np.int32), np.array(offsets, dtype=np.int32)


def _serial_edge_numbering(tris):
    """Literal restatement of the edge-numbering loop (mesh.py:464-529) with a dict instead of np.unique."""
    first = {}
    edges, edge_tris = [], []
    tri_edge = np.full((len(tris), 3), -1, dtype=np.int32)
    for t, (a, b, c) in enumerate(tris.tolist()):
        for slot, (v0, v1, opp) in enumerate(((a, b, c), (b, c, a), (c, a, b))):
            key = (min(v0, v1), max(v0, v1))
            if key not in first:
                first[key] = len(edges)
                edges.append([opp, -1, v0, v1])
                edge_tris.append([t, -1])
            else:
                e = first[key]
                if edges[e][1] == -1 and edge_tris[e][1] == -1:
                    edges[e][1] = opp
                    edge_tris[e][1] = t
            tri_edge[t, slot] = first[key]
    return np.array(edges, dtype=np.int32), np.array(edge_tris, dtype=np.int32), tri_edge


def _random_manifold_mesh(rng, nx, ny):
    """A jittered grid triangulated with random diagonals, vertices relabelled at random."""
    verts = nx * ny
    perm = rng.permutation(verts)
    tris = []
    for y in range(ny - 1):
        for x in range(nx - 1):
            v0, v1, v2, v3 = y * nx + x, y * nx + x + 1, (y + 1) * nx + x + 1, (y + 1) * nx + x
            if rng.random() < 0.5:
                tris += [[v0, v1, v3], [v1, v2, v3]]
            else:
                tris += [[v0, v1, v2], [v0, v2, v3]]
    tris = perm[np.array(tris)]
    return tris[rng.permutation(len(tris))].astype(np.int32), verts


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_edge_numbering_matches_serial_loop(seed):
    rng = np.random.default_rng(seed)
    tris, _ = _random_manifold_mesh(rng, 7, 6)
    e, et, te = compute_edge_adjacency(tris)
    e_ref, et_ref, te_ref = _serial_edge_numbering(tris)
    np.testing.assert_array_equal(e, e_ref)
    np.testing.assert_array_equal(et, et_ref)
    np.testing.assert_array_equal(te, te_ref)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_vertex_csr_matches_serial_fill(seed):
    rng = np.random.default_rng(seed)
    tris, n = _random_manifold_mesh(rng, 8, 5)
    adj = MeshAdjacency(tri_indices=tris)
    tets = rng.integers(0, n, size=(11, 4)).astype(np.int32)
    adj2 = MeshAdjacency(tri_indices=tris, edge_indices=adj.edge_indices, tet_indices=tets)
    adj2.init_vertex_adjacency(n)
    v, o = _serial_csr(adj.edge_indices, [2, 3, 0, 1], n)
    np.testing.assert_array_equal(adj2.v_adj_edges, v)
    np.testing.assert_array_equal(adj2.v_adj_edges_offsets, o)
    v, o = _serial_csr(tris, [0, 1, 2], n)
    np.testing.assert_array_equal(adj2.v_adj_tris, v)
    np.testing.assert_array_equal(adj2.v_adj_tris_offsets, o)
    v, o = _serial_csr(tets, [0, 1, 2, 3], n)
    np.testing.assert_array_equal(adj2.v_adj_tets, v)
    np.testing.assert_array_equal(adj2.v_adj_tets_offsets, o)


def test_csr_entries_point_back_at_their_vertex():
    from newton_amd import scenes

    model, _, _ = scenes.cloth_drop_on_sphere("cpu", verts_per_side=24)
    adj = model.soft_mesh_adjacency
    tri = model.tri_indices.numpy()
    edge = model.edge_indices.numpy()
    for vals, offs, elems in ((adj.v_adj_tris, adj.v_adj_tris_offsets, tri),
                              (adj.v_adj_edges, adj.v_adj_edges_offsets, edge)):
        pairs = vals.reshape(-1, 2)
        owner = np.repeat(np.arange(model.particle_count), np.diff(offs) // 2)
        np.testing.assert_array_equal(elems[pairs[:, 0], pairs[:, 1]], owner)
        # per vertex, element ids strictly increase (serial fill order)
        same = owner[1:] == owner[:-1]
        assert np.all(pairs[1:, 0][same] > pairs[:-1, 0][same])


def test_grid_edge_counts():
    from newton_amd import scenes

    model, _, _ = scenes.cloth_hanging("cpu")
    n = 31
    assert model.particle_count == 32 * 32
    assert model.tri_count == 2 * n * n
    assert model.edge_count == 3 * n * n + 2 * n
    adj = model.soft_mesh_adjacency
    # interior vertex of the regular grid: 6 triangles, 12 hinge memberships (6 as edge end, 6 as opposite)
    v = 16 * 32 + 16
    assert (adj.v_adj_tris_offsets[v + 1] - adj.v_adj_tris_offsets[v]) // 2 == 6
    assert (adj.v_adj_edges_offsets[v + 1] - adj.v_adj_edges_offsets[v]) // 2 == 12


def _random_soup(rng, n_verts, n_tris):
    tris = np.stack([rng.choice(n_verts, 3, replace=False) for _ in range(n_tris)]).astype(np.int32)
    return tris


def _loop_edge_tri_maps(edges, tris):
    """The linking rule of mesh.py:531-576 as plain loops: each triangle side takes the lowest-numbered
    edge on its vertex pair, and links only if its opposite vertex is that edge's o0 (side 0) or o1
    (side 1); a later side overwrites an earlier one on the same (edge, side)."""
    edge_tri = np.full((len(edges), 2), -1, dtype=np.int32)
    tri_edge = np.full((len(tris), 3), -1, dtype=np.int32)
    lowest = {}
    for e in range(len(edges) - 1, -1, -1):
        a, b = int(edges[e, 2]), int(edges[e, 3])
        if a >= 0 and b >= 0:
            lowest[frozenset((a, b))] = e
    for t, tri in enumerate(tris.tolist()):
        for k in range(3):
            e = lowest.get(frozenset((tri[k], tri[(k + 1) % 3])))
            if e is None:
                continue
            opp = tri[(k + 2) % 3]
            for side in (0, 1):
                if opp == edges[e, side]:
                    tri_edge[t, k] = e
                    edge_tri[e, side] = t
                    break
    return edge_tri, tri_edge


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_native_edge_tri_maps_equal_loop_rule(seed):
    """C++ vbd_host_edge_tri_maps == the rule of mesh.py:531-576 as loops, incl. duplicates and placeholders."""
    from newton_amd.utils.mesh import build_edge_tri_maps

    rng = np.random.default_rng(seed)
    tris = _random_soup(rng, 12, 40)  # dense soup: non-manifold edges, repeated pairs
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        edges, _, _ = compute_edge_adjacency(tris)
    edges = edges.copy()
    # duplicate some edges (same vertex pair, other opposites) and add placeholder rows
    dup = edges[rng.choice(len(edges), 6, replace=False)].copy()
    dup[:, :2] = rng.integers(-1, 12, size=(6, 2))
    ph = edges[rng.choice(len(edges), 3, replace=False)].copy()
    ph[:, :2] = -1
    edges = np.concatenate([edges, dup, ph])
    perm = rng.permutation(len(edges))
    edges = np.ascontiguousarray(edges[perm])
    a = build_edge_tri_maps(edges, tris)
    b = _loop_edge_tri_maps(edges, tris)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


@pytest.mark.parametrize("seed", [0, 1])
def test_edge_numbering_non_manifold_soup_matches_serial_loop(seed):
    """A dense soup (pairs shared by three or more sides): the extra sides change nothing and warn."""
    rng = np.random.default_rng(seed)
    tris = _random_soup(rng, 9, 30)
    with pytest.warns(UserWarning, match="non-manifold"):
        e, et, te = compute_edge_adjacency(tris, tri_start=5)
    e_ref, et_ref, te_ref = _serial_edge_numbering(tris)
    np.testing.assert_array_equal(e, e_ref)
    np.testing.assert_array_equal(et, np.where(et_ref >= 0, et_ref + 5, -1))
    np.testing.assert_array_equal(te, te_ref)


def test_edge_numbering_empty_and_negative():
    e, et, te = compute_edge_adjacency(np.empty((0, 3), dtype=np.int32))
    assert e.shape == (0, 4) and et.shape == (0, 2) and te.shape == (0, 3)
    with pytest.raises(ValueError):
        compute_edge_adjacency(np.array([[0, 1, -2]], dtype=np.int32))


@pytest.mark.parametrize("cols,vpe", [([0, 1, 2], 2), ([2, 3, 0, 1], 2), ([0, 1], 1)])
def test_native_vertex_csr_equals_numpy_restatement(cols, vpe):
    from newton_amd.utils.mesh import _vertex_csr, _vertex_csr_np

    rng = np.random.default_rng(len(cols) * 10 + vpe)
    width = max(cols) + 1
    rows = rng.integers(-1, 50, size=(300, width)).astype(np.int32)
    v_a, o_a = _vertex_csr(rows, cols, 50, values_per_entry=vpe)
    v_b, o_b = _vertex_csr_np(rows, cols, 50, values_per_entry=vpe)
    np.testing.assert_array_equal(o_a, o_b)
    np.testing.assert_array_equal(v_a, v_b)
    with pytest.raises(ValueError):
        bad = rows.copy()
        bad[7, cols[0]] = 50
        _vertex_csr(bad, cols, 50, values_per_entry=vpe)
