// Host-side model-construction routines for the particle VBD path.
//
// * Graph coloring of the particle graph: maximum-cardinality-search (MCS)
//   ordering with greedy color assignment (Pereira & Palsberg 2005), or
//   degree-ordered greedy coloring, followed by a balancing pass that moves
//   vertices from the largest color to smaller ones until the max/min size
//   ratio reaches a target.  This stands in for the reference's calls into
//   warp-lang's native graph_coloring_assign / graph_coloring_balance
//   (newton/_src/sim/graph_coloring.py:243-311); the exact color of each
//   vertex is this library's own and is fed to every consumer (GPU solver and
//   CPU oracle alike).
//
// Plain C ABI; called through ctypes.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct Csr {
    std::vector<int64_t> off;
    std::vector<int32_t> adj;
};

Csr build_csr(int32_t n, int64_t m, const int32_t* edges) {
    Csr g;
    g.off.assign((size_t)n + 1, 0);
    for (int64_t e = 0; e < m; ++e) {
        int32_t a = edges[2 * e], b = edges[2 * e + 1];
        if (a == b) continue;
        g.off[a + 1]++;
        g.off[b + 1]++;
    }
    for (int32_t i = 0; i < n; ++i) g.off[i + 1] += g.off[i];
    g.adj.resize((size_t)g.off[n]);
    std::vector<int64_t> fill(g.off.begin(), g.off.end() - 1);
    for (int64_t e = 0; e < m; ++e) {
        int32_t a = edges[2 * e], b = edges[2 * e + 1];
        if (a == b) continue;
        g.adj[fill[a]++] = b;
        g.adj[fill[b]++] = a;
    }
    return g;
}

// smallest color not used by already-colored neighbours
int32_t first_free_color(const Csr& g, int32_t v, const std::vector<int32_t>& color, std::vector<int32_t>& mark,
                         int32_t stamp) {
    for (int64_t k = g.off[v]; k < g.off[v + 1]; ++k) {
        int32_t c = color[g.adj[k]];
        if (c >= 0) {
            if ((size_t)c >= mark.size()) mark.resize((size_t)c + 1, -1);
            mark[c] = stamp;
        }
    }
    int32_t c = 0;
    while ((size_t)c < mark.size() && mark[c] == stamp) ++c;
    return c;
}

// MCS order: repeatedly visit the unvisited vertex with the most visited
// neighbours (ties: lowest index among the bucket head), coloring greedily.
int32_t color_mcs(const Csr& g, int32_t n, std::vector<int32_t>& color) {
    std::vector<int32_t> weight((size_t)n, 0), next((size_t)n, -1), prev((size_t)n, -1);
    std::vector<int32_t> head(1, -1);
    std::vector<char> visited((size_t)n, 0);
    // bucket 0 holds all vertices, inserted in reverse so the head is vertex 0
    for (int32_t v = n - 1; v >= 0; --v) {
        next[v] = head[0];
        if (head[0] >= 0) prev[head[0]] = v;
        head[0] = v;
    }
    auto unlink = [&](int32_t v) {
        int32_t w = weight[v];
        if (prev[v] >= 0)
            next[prev[v]] = next[v];
        else
            head[w] = next[v];
        if (next[v] >= 0) prev[next[v]] = prev[v];
        prev[v] = next[v] = -1;
    };
    auto push = [&](int32_t v) {
        int32_t w = weight[v];
        if ((size_t)w >= head.size()) head.resize((size_t)w + 1, -1);
        next[v] = head[w];
        prev[v] = -1;
        if (head[w] >= 0) prev[head[w]] = v;
        head[w] = v;
    };
    std::vector<int32_t> mark;
    int32_t max_w = 0, num_colors = 0;
    for (int32_t it = 0; it < n; ++it) {
        while (max_w > 0 && head[max_w] < 0) --max_w;
        int32_t v = head[max_w];
        unlink(v);
        visited[v] = 1;
        int32_t c = first_free_color(g, v, color, mark, it);
        color[v] = c;
        num_colors = std::max(num_colors, c + 1);
        for (int64_t k = g.off[v]; k < g.off[v + 1]; ++k) {
            int32_t u = g.adj[k];
            if (visited[u]) continue;
            unlink(u);
            weight[u]++;
            push(u);
            max_w = std::max(max_w, weight[u]);
        }
    }
    return num_colors;
}

// degree-ordered greedy (largest degree first, ties by index)
int32_t color_greedy(const Csr& g, int32_t n, std::vector<int32_t>& color) {
    std::vector<int32_t> order((size_t)n);
    for (int32_t i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        return (g.off[a + 1] - g.off[a]) > (g.off[b + 1] - g.off[b]);
    });
    std::vector<int32_t> mark;
    int32_t num_colors = 0, stamp = 0;
    for (int32_t v : order) {
        int32_t c = first_free_color(g, v, color, mark, stamp++);
        color[v] = c;
        num_colors = std::max(num_colors, c + 1);
    }
    return num_colors;
}

float size_ratio(const std::vector<int64_t>& sizes) {
    int64_t mx = 0, mn = INT64_MAX;
    for (int64_t s : sizes) {
        mx = std::max(mx, s);
        mn = std::min(mn, s);
    }
    if (mn <= 0) return 1e30f;
    return (float)mx / (float)mn;
}

// Move vertices out of the largest color into smaller colors (smallest first)
// when none of their neighbours already has the target color.
float balance(const Csr& g, int32_t n, int32_t num_colors, float target, std::vector<int32_t>& color) {
    if (num_colors <= 1) return 1.0f;
    std::vector<int64_t> sizes((size_t)num_colors, 0);
    for (int32_t v = 0; v < n; ++v) sizes[color[v]]++;
    std::vector<char> blocked((size_t)num_colors);
    for (int round = 0; round < 4 * num_colors + 16; ++round) {
        float ratio = size_ratio(sizes);
        if (ratio <= target) return ratio;
        int32_t big = (int32_t)(std::max_element(sizes.begin(), sizes.end()) - sizes.begin());
        std::vector<int32_t> targets;
        for (int32_t c = 0; c < num_colors; ++c)
            if (c != big) targets.push_back(c);
        std::stable_sort(targets.begin(), targets.end(), [&](int32_t a, int32_t b) { return sizes[a] < sizes[b]; });
        bool moved_any = false;
        for (int32_t v = 0; v < n && sizes[big] > 0; ++v) {
            if (color[v] != big) continue;
            std::fill(blocked.begin(), blocked.end(), 0);
            for (int64_t k = g.off[v]; k < g.off[v + 1]; ++k) blocked[color[g.adj[k]]] = 1;
            for (int32_t c : targets) {
                if (sizes[c] + 1 >= sizes[big]) break;  // sorted ascending: no better target
                if (blocked[c]) continue;
                color[v] = c;
                sizes[big]--;
                sizes[c]++;
                moved_any = true;
                break;
            }
            if (size_ratio(sizes) <= target) break;
        }
        if (!moved_any) return size_ratio(sizes);
    }
    return size_ratio(sizes);
}

}  // namespace

extern "C" {

// Colors the graph (n nodes, m undirected edges as int32 pairs).
// algorithm: 0 = MCS, 1 = degree-ordered greedy.  Returns the number of colors
// (>= 1 when n > 0) or a negative code on bad arguments.
int32_t vbd_host_color_graph(int32_t n, int64_t m, const int32_t* edges, int32_t algorithm, int32_t balance_colors,
                             float target_ratio, int32_t* colors_out, float* ratio_out) {
    if (n < 0 || m < 0 || (m > 0 && !edges) || !colors_out) return -1;
    if (n == 0) return 0;
    for (int64_t e = 0; e < 2 * m; ++e)
        if (edges[e] < 0 || edges[e] >= n) return -2;
    Csr g = build_csr(n, m, edges);
    std::vector<int32_t> color((size_t)n, -1);
    int32_t nc = algorithm == 1 ? color_greedy(g, n, color) : color_mcs(g, n, color);
    float ratio = 1.0f;
    if (balance_colors) ratio = balance(g, n, nc, target_ratio, color);
    std::memcpy(colors_out, color.data(), sizeof(int32_t) * (size_t)n);
    if (ratio_out) *ratio_out = ratio;
    return nc;
}

// Edge->triangle and triangle->edge maps for pre-numbered edges
// (MeshAdjacency._build_maps, newton/_src/utils/mesh.py:531-576).
// Triangle slot k spans (v_k, v_{k+1}) with opposite vertex v_{k+2}.  A slot is
// matched to the lowest-numbered edge with the same unordered vertex pair and
// linked only if its opposite vertex equals the edge's o0 (side 0) or, failing
// that, o1 (side 1).  When several slots link the same (edge, side), the last
// in (triangle, slot) order wins.  Edge rows are (o0, o1, v0, v1).
// Vertex ids may be any int32 >= 0 (edges with negative v0/v1 never match).
// Linear time: edges are bucketed by their smaller vertex in id order.
int32_t vbd_host_edge_tri_maps(int64_t edge_count, const int32_t* edges, int64_t tri_count, const int32_t* tris,
                               int32_t* edge_tri_out, int32_t* tri_edge_out) {
    if (edge_count < 0 || tri_count < 0 || (edge_count > 0 && !edges) || (tri_count > 0 && !tris) ||
        (edge_count > 0 && !edge_tri_out) || (tri_count > 0 && !tri_edge_out))
        return -1;
    for (int64_t i = 0; i < 2 * edge_count; ++i) edge_tri_out[i] = -1;
    for (int64_t i = 0; i < 3 * tri_count; ++i) tri_edge_out[i] = -1;
    if (edge_count == 0 || tri_count == 0) return 0;
    int64_t vmax = -1;
    for (int64_t e = 0; e < edge_count; ++e)
        vmax = std::max<int64_t>(vmax, std::max(edges[4 * e + 2], edges[4 * e + 3]));
    if (vmax < 0) return 0;
    const int64_t nb = vmax + 1;
    std::vector<int64_t> off((size_t)nb + 1, 0);
    for (int64_t e = 0; e < edge_count; ++e) {
        int32_t a = edges[4 * e + 2], b = edges[4 * e + 3];
        if (a < 0 || b < 0) continue;
        off[(size_t)std::min(a, b) + 1]++;
    }
    for (int64_t i = 0; i < nb; ++i) off[i + 1] += off[i];
    std::vector<int32_t> hi((size_t)off[nb]), id((size_t)off[nb]);
    {
        std::vector<int64_t> fill(off.begin(), off.end() - 1);
        for (int64_t e = 0; e < edge_count; ++e) {
            int32_t a = edges[4 * e + 2], b = edges[4 * e + 3];
            if (a < 0 || b < 0) continue;
            int64_t k = fill[std::min(a, b)]++;
            hi[k] = std::max(a, b);
            id[k] = (int32_t)e;
        }
    }
    for (int64_t t = 0; t < tri_count; ++t) {
        const int32_t* v = tris + 3 * t;
        for (int k = 0; k < 3; ++k) {
            int32_t a = v[k], b = v[(k + 1) % 3], opp = v[(k + 2) % 3];
            int32_t lo = std::min(a, b), h = std::max(a, b);
            if (lo < 0 || lo > vmax) continue;
            int32_t cand = -1;
            for (int64_t j = off[lo]; j < off[lo + 1]; ++j)
                if (hi[j] == h) {
                    cand = id[j];  // bucket is in edge-id order: first hit is the lowest id
                    break;
                }
            if (cand < 0) continue;
            if (opp == edges[4 * (int64_t)cand + 0]) {
                tri_edge_out[3 * t + k] = cand;
                edge_tri_out[2 * (int64_t)cand + 0] = (int32_t)t;
            } else if (opp == edges[4 * (int64_t)cand + 1]) {
                tri_edge_out[3 * t + k] = cand;
                edge_tri_out[2 * (int64_t)cand + 1] = (int32_t)t;
            }
        }
    }
    return 0;
}

// Edge numbering of a triangle mesh (MeshAdjacency._compute_edge_adjacency,
// newton/_src/utils/mesh.py:464-529).  The rule: walk the triangles in order
// and, inside each, its sides k = 0, 1, 2 spanning (v_k, v_{k+1}) with opposite
// vertex v_{k+2}; an unordered vertex pair gets the next edge id the first time
// it is met.  Its row (o0, o1, v0, v1) takes v0, v1 and o0 from that first
// side and o1 from the second; edge_tri holds the two triangles (+ tri_start);
// a third or later side on one pair counts as non-manifold and changes nothing.
//
// Two linear passes: the sides are bucketed by their smaller vertex (counting
// sort, so each bucket lists its sides in walk order); then, in walk order,
// each side looks for the first side of its bucket with the same larger
// vertex -- the side that introduced the pair -- and takes over its edge.
// edges_out needs room for 3 * tri_count rows (the worst case); the number of
// edges is returned in *edge_count_out and the number of ignored non-manifold
// sides in *nonmanifold_out.  Returns -1 on a bad argument, -2 on a negative
// vertex id.
int32_t vbd_host_number_edges(int64_t tri_count, const int32_t* tris, int32_t tri_start, int32_t* edges_out,
                              int32_t* edge_tri_out, int32_t* tri_edge_out, int64_t* edge_count_out,
                              int64_t* nonmanifold_out) {
    if (tri_count < 0 || (tri_count > 0 && (!tris || !edges_out || !edge_tri_out || !tri_edge_out)) ||
        !edge_count_out || !nonmanifold_out)
        return -1;
    *edge_count_out = 0;
    *nonmanifold_out = 0;
    if (tri_count == 0) return 0;
    const int64_t sides = 3 * tri_count;
    int32_t vmax = -1;
    for (int64_t i = 0; i < sides; ++i) {
        if (tris[i] < 0) return -2;
        vmax = std::max(vmax, tris[i]);
    }
    // bucket start per smaller vertex, then the sides of each bucket in walk order
    std::vector<int64_t> start((size_t)vmax + 2, 0);
    for (int64_t s = 0; s < sides; ++s) {
        const int64_t t = s / 3;
        const int k = (int)(s % 3);
        start[(size_t)std::min(tris[3 * t + k], tris[3 * t + (k + 1) % 3]) + 1]++;
    }
    for (int32_t v = 0; v <= vmax; ++v) start[(size_t)v + 1] += start[(size_t)v];
    std::vector<int64_t> bucket((size_t)sides);
    {
        std::vector<int64_t> cursor(start.begin(), start.end() - 1);
        for (int64_t s = 0; s < sides; ++s) {
            const int64_t t = s / 3;
            const int k = (int)(s % 3);
            bucket[(size_t)cursor[(size_t)std::min(tris[3 * t + k], tris[3 * t + (k + 1) % 3])]++] = s;
        }
    }
    std::vector<int32_t> side_edge((size_t)sides, -1);
    std::vector<int32_t> uses;  // sides met so far per edge
    int32_t next_edge = 0;
    for (int64_t s = 0; s < sides; ++s) {
        const int64_t t = s / 3;
        const int k = (int)(s % 3);
        const int32_t a = tris[3 * t + k], b = tris[3 * t + (k + 1) % 3], opp = tris[3 * t + (k + 2) % 3];
        const int32_t lo = std::min(a, b), hi = std::max(a, b);
        int64_t first = s;
        for (int64_t j = start[(size_t)lo]; j < start[(size_t)lo + 1]; ++j) {
            const int64_t u = bucket[(size_t)j];
            const int64_t tu = u / 3;
            const int ku = (int)(u % 3);
            if (std::max(tris[3 * tu + ku], tris[3 * tu + (ku + 1) % 3]) == hi) {
                first = u;  // buckets are in walk order: the first match introduced the pair
                break;
            }
        }
        int32_t e;
        if (first == s) {
            e = next_edge++;
            uses.push_back(0);
            int32_t* row = edges_out + 4 * (int64_t)e;
            row[0] = opp;
            row[1] = -1;
            row[2] = a;
            row[3] = b;
            edge_tri_out[2 * (int64_t)e] = tri_start + (int32_t)t;
            edge_tri_out[2 * (int64_t)e + 1] = -1;
        } else {
            e = side_edge[(size_t)first];
            if (uses[(size_t)e] == 1) {
                edges_out[4 * (int64_t)e + 1] = opp;
                edge_tri_out[2 * (int64_t)e + 1] = tri_start + (int32_t)t;
            } else {
                ++*nonmanifold_out;
            }
        }
        uses[(size_t)e]++;
        side_edge[(size_t)s] = e;
        tri_edge_out[s] = e;
    }
    *edge_count_out = next_edge;
    return 0;
}

// Per-vertex CSR of (element, slot) pairs (values_per_entry 2) or element ids
// (values_per_entry 1) in the serial count/fill order of the reference's
// adjacency kernels (newton/_src/utils/mesh.py:674-857): elements in increasing
// id, within an element the columns in fill_cols order; entries equal to -1 are
// skipped.  rows is [n_elem, width]; offsets_out [n+1] counts ints.  Returns -2
// if an entry is outside [0, n) (and not -1), -3 if the table exceeds int32.
int32_t vbd_host_vertex_csr(int64_t n_elem, int32_t width, const int32_t* rows, int32_t ncols,
                            const int32_t* fill_cols, int32_t n, int32_t values_per_entry, int32_t* offsets_out,
                            int32_t* values_out, int64_t values_capacity) {
    if (n_elem < 0 || width <= 0 || ncols <= 0 || n < 0 || !fill_cols || !offsets_out ||
        (values_per_entry != 1 && values_per_entry != 2) || (n_elem > 0 && !rows))
        return -1;
    for (int32_t c = 0; c < ncols; ++c)
        if (fill_cols[c] < 0 || fill_cols[c] >= width) return -1;
    std::vector<int64_t> cnt((size_t)n + 1, 0);
    for (int64_t e = 0; e < n_elem; ++e)
        for (int32_t c = 0; c < ncols; ++c) {
            int32_t v = rows[e * width + fill_cols[c]];
            if (v == -1) continue;
            if (v < 0 || v >= n) return -2;
            cnt[(size_t)v + 1]++;
        }
    for (int32_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    const int64_t total = cnt[n] * values_per_entry;
    if (total > INT32_MAX) return -3;
    for (int32_t i = 0; i <= n; ++i) offsets_out[i] = (int32_t)(cnt[i] * values_per_entry);
    if (total > values_capacity || (total > 0 && !values_out)) return -1;
    std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
    for (int64_t e = 0; e < n_elem; ++e)
        for (int32_t c = 0; c < ncols; ++c) {
            int32_t v = rows[e * width + fill_cols[c]];
            if (v == -1) continue;
            int64_t k = fill[v]++;
            if (values_per_entry == 2) {
                values_out[2 * k] = (int32_t)e;
                values_out[2 * k + 1] = fill_cols[c];
            } else {
                values_out[k] = (int32_t)e;
            }
        }
    return 0;
}

}  // extern "C"
