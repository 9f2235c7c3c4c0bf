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

}  // extern "C"
