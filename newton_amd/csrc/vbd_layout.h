// Host-side "pass layout" for the color-pass kernel (built once per model in vbd_create).
//
// For every color, the color's vertices (in color-group order) are cut into
// batches: runs of at most max_v consecutive vertices whose element counts sum
// to at most cap (the work-group shape: 256 lanes -> cap 256, max_v 21; 128
// lanes -> cap 128, max_v 10).  A batch is one workgroup.  Its element
// contributions are stored as packed records in evaluation order -- all
// triangle items of the batch, then all bending-edge items, then all tet
// items -- each carrying the canonical row (the position of that contribution
// in its vertex's serial sum: the vertex's triangles in CSR order, then its
// edges, then its tets) and the vertex's local slot in the element.
//
// Elements the reference's solve skips are left out: triangles and tets with
// mu <= 0 and lambda <= 0, edges with ke <= 0 (particle_vbd_kernels.py:2636,
// :2658, :2679), and boundary hinges (o0 or o1 == -1), whose contribution is
// an exact zero (:682-683).  Inactive particles (flag bit 0 clear or mass 0)
// get no items.  A vertex with more than cap items goes to the "slow"
// list, handled by the generic kernel that walks the CSR in chunks.
//
// Springs (particle_vbd_kernels.py:1733-1783) join the batch as items of their
// own: each active endpoint of a spring gets one item, in spring-index order
// per vertex, whose row follows the vertex's element rows (the solve adds them
// to the contact sum, after the contacts).  A spring whose endpoints share a
// color cannot be evaluated inside the pass (the other endpoint moves during
// it); its items are evaluated by a pre-pass over the color into global rows
// ("pre items") and the pass item only points at its row.
#pragma once

#include <cstdint>
#include <cstring>
#include <vector>

namespace vbd_layout {

constexpr int kLayoutCap = 256;  // static items (rows) per batch of the 256-lane shape
constexpr int kLayoutMaxV = 21;  // vertices per batch of the 256-lane shape (12 component lanes each)

// record widths in 32-bit words
constexpr int kTriWords = 12;   // i0 i1 i2 meta | D00 D01 D10 D11 | area mu lambda kd
constexpr int kEdgeWords = 8;   // o0 o1 v0 v1 | rest_angle rest_length ke kd      (+ edge_meta)
constexpr int kTetWords = 16;   // i0 i1 i2 i3 | Dm^-1 (9) | mu lambda kd         (+ tet_meta)
constexpr int kSpringWords = 8; // v0 v1 (row | endpoint << 16) pre_row | rest ke kd 0
constexpr int kBatchWords = 12; // vbegin nv tri_begin ntri edge_begin nedge tet_begin ntet spr_begin nspr 0 0
constexpr int kVertWords = 4;   // v row_start nrows (local | active << 16 | nsprings << 17)
// batch block: header (kBatchWords) then max_v vertex records (unused ones zero),
// fixed size so a workgroup reads its header and vertex records in one round trip
inline int block_words(int max_v) { return kBatchWords + max_v * kVertWords; }  // a multiple of 4 words

inline int32_t meta(int row, int slot) { return (int32_t)(row | (slot << 16)); }

struct HostModel {
    int N = 0, T = 0, E = 0, K = 0;
    std::vector<float> mass;
    std::vector<int32_t> flags;
    std::vector<int32_t> tri_idx, edge_idx, tet_idx;
    std::vector<float> tri_pose, tri_mat, tri_area;
    std::vector<float> edge_rest, edge_len, edge_bend;
    std::vector<float> tet_pose, tet_mat;
    std::vector<int32_t> adj_tri, adj_tri_off, adj_edge, adj_edge_off, adj_tet, adj_tet_off;
    std::vector<int32_t> color_off, color_particles;
    int S = 0;
    std::vector<int32_t> spring_idx;
    std::vector<float> spring_rest, spring_ke, spring_kd;
};

struct PassLayout {
    std::vector<int32_t> fast_color_off;  // [C+1] into batch records
    std::vector<int32_t> slow_color_off;  // [C+1] into slow_pos
    std::vector<int32_t> batch_blk;       // [B * block_words(max_v)]: header + vertex records
    int cap = kLayoutCap, max_v = kLayoutMaxV, blk_words = block_words(kLayoutMaxV);
    std::vector<int32_t> vert_rec;        // [N * kVertWords], indexed by color position
    std::vector<int32_t> tri_rec, edge_rec, edge_meta, tet_rec, tet_meta;
    std::vector<int32_t> slow_pos;        // color positions of oversized vertices
    std::vector<int32_t> spr_rec;         // [*, kSpringWords] spring items of the batches
    std::vector<int32_t> pre_item;        // [P, 2] (spring, endpoint) of same-color spring items
    std::vector<int32_t> pre_color_off;   // [C+1] into pre_item
    std::vector<int32_t> slow_spr_off;    // [slow + 1] into slow_spr
    std::vector<int32_t> slow_spr;        // [*, 4] (spring, endpoint, pre_row or -1, 0) of slow vertices
    int64_t fast_items = 0;
};

inline int32_t fbits(float f) {
    int32_t i;
    std::memcpy(&i, &f, 4);
    return i;
}

// Builds the layout; returns an empty string or an error message.
// Lanes an item range takes when ranges start on wave (64-lane) boundaries.
inline int wave_align(int n) { return (n + 63) & ~63; }

// aligned: each kind's items start on a wave boundary of the work-group (a wave then evaluates one
// kind only); the batch's lane budget counts the padded ranges.
inline std::string build_pass_layout(const HostModel& m, PassLayout& L, int cap = kLayoutCap,
                                     int max_v = kLayoutMaxV, bool aligned = false) {
    const int N = m.N;
    const int C = (int)m.color_off.size() - 1;
    std::vector<char> active((size_t)N), tri_ok((size_t)m.T), edge_ok((size_t)m.E), tet_ok((size_t)m.K);
    for (int v = 0; v < N; ++v) active[v] = (m.flags[v] & 1) && m.mass[v] != 0.0f;
    for (int t = 0; t < m.T; ++t) tri_ok[t] = m.tri_mat[5 * t] > 0.0f || m.tri_mat[5 * t + 1] > 0.0f;
    for (int e = 0; e < m.E; ++e)
        edge_ok[e] = m.edge_bend[2 * e] > 0.0f && m.edge_idx[4 * e] != -1 && m.edge_idx[4 * e + 1] != -1;
    for (int k = 0; k < m.K; ++k) tet_ok[k] = m.tet_mat[3 * k] > 0.0f || m.tet_mat[3 * k + 1] > 0.0f;
    std::vector<int32_t> color_of((size_t)N, -1);
    for (int c = 0; c < C; ++c)
        for (int k = m.color_off[c]; k < m.color_off[c + 1]; ++k) color_of[m.color_particles[k]] = c;
    // per-vertex spring items in spring-index order: 2 * spring + endpoint
    std::vector<int32_t> vs_off((size_t)N + 1, 0), vs_item;
    for (int s = 0; s < m.S; ++s)
        for (int e = 0; e < 2; ++e) {
            int v = m.spring_idx[2 * s + e];
            if (v < 0 || v >= N) return "spring_indices out of range";
            ++vs_off[v + 1];
        }
    for (int v = 0; v < N; ++v) vs_off[v + 1] += vs_off[v];
    vs_item.assign((size_t)vs_off[N], 0);
    {
        std::vector<int32_t> fill(vs_off.begin(), vs_off.end() - 1);
        for (int s = 0; s < m.S; ++s)
            for (int e = 0; e < 2; ++e) vs_item[fill[m.spring_idx[2 * s + e]]++] = 2 * s + e;
    }
    auto same_color = [&](int s) { return color_of[m.spring_idx[2 * s]] == color_of[m.spring_idx[2 * s + 1]]; };

    auto count_ok = [&](const std::vector<int32_t>& vals, const std::vector<int32_t>& off, const std::vector<char>& ok,
                        int v) {
        if (off.empty()) return 0;
        int c = 0;
        for (int p = off[v]; p < off[v + 1]; p += 2) c += ok[vals[p]] ? 1 : 0;
        return c;
    };
    std::vector<int32_t> ntri((size_t)N, 0), nedge((size_t)N, 0), ntet((size_t)N, 0), nspr((size_t)N, 0);
    for (int v = 0; v < N; ++v) {
        if (!active[v]) continue;
        if (m.T > 0) ntri[v] = count_ok(m.adj_tri, m.adj_tri_off, tri_ok, v);
        if (m.E > 0) nedge[v] = count_ok(m.adj_edge, m.adj_edge_off, edge_ok, v);
        if (m.K > 0) ntet[v] = count_ok(m.adj_tet, m.adj_tet_off, tet_ok, v);
        nspr[v] = vs_off[v + 1] - vs_off[v];
    }
    auto nitems = [&](int v) { return ntri[v] + nedge[v] + ntet[v] + nspr[v]; };

    L = PassLayout();
    L.cap = cap;
    L.max_v = max_v;
    L.blk_words = block_words(max_v);
    L.vert_rec.assign((size_t)N * kVertWords, 0);
    L.fast_color_off.assign((size_t)C + 1, 0);
    L.slow_color_off.assign((size_t)C + 1, 0);
    L.pre_color_off.assign((size_t)C + 1, 0);
    L.slow_spr_off.push_back(0);
    // pre rows of same-color spring items, per color in color-position order
    std::vector<int32_t> pre_row((size_t)vs_off[N], -1);
    for (int c = 0; c < C; ++c) {
        L.pre_color_off[c] = (int32_t)(L.pre_item.size() / 2);
        for (int k = m.color_off[c]; k < m.color_off[c + 1]; ++k) {
            int v = m.color_particles[k];
            if (!active[v]) continue;
            for (int p = vs_off[v]; p < vs_off[v + 1]; ++p) {
                int s = vs_item[p] >> 1;
                if (!same_color(s)) continue;
                pre_row[p] = (int32_t)(L.pre_item.size() / 2);
                L.pre_item.push_back(s);
                L.pre_item.push_back(vs_item[p] & 1);
            }
        }
    }
    L.pre_color_off[C] = (int32_t)(L.pre_item.size() / 2);
    std::vector<int> rows;
    for (int c = 0; c < C; ++c) {
        L.fast_color_off[c] = (int32_t)(L.batch_blk.size() / L.blk_words);
        L.slow_color_off[c] = (int32_t)L.slow_pos.size();
        int k = m.color_off[c];
        const int kend = m.color_off[c + 1];
        while (k < kend) {
            int v0 = m.color_particles[k];
            int items0 = nitems(v0);
            const int lanes0 = wave_align(ntri[v0]) + wave_align(nedge[v0]) + wave_align(ntet[v0]) +
                               wave_align(nspr[v0]);
            if (items0 > cap || (aligned && lanes0 > cap)) {  // oversized vertex: generic kernel
                L.slow_pos.push_back(k);
                if (active[v0])
                    for (int p = vs_off[v0]; p < vs_off[v0 + 1]; ++p) {
                        const int32_t r[4] = {vs_item[p] >> 1, vs_item[p] & 1, pre_row[p], 0};
                        L.slow_spr.insert(L.slow_spr.end(), r, r + 4);
                    }
                L.slow_spr_off.push_back((int32_t)(L.slow_spr.size() / 4));
                ++k;
                continue;
            }
            // greedy run of vertices that fit one workgroup
            int vb = k, nv = 0, items = 0, kt = 0, ke = 0, kk = 0, ks = 0;
            while (k < kend && nv < max_v) {
                int v = m.color_particles[k];
                int it = nitems(v);
                if (it > cap || items + it > cap) break;
                if (aligned && active[v]) {
                    const int lanes = wave_align(kt + ntri[v]) + wave_align(ke + nedge[v]) +
                                      wave_align(kk + ntet[v]) + wave_align(ks + nspr[v]);
                    if (lanes > cap) break;
                }
                if (active[v]) {
                    kt += ntri[v];
                    ke += nedge[v];
                    kk += ntet[v];
                    ks += nspr[v];
                }
                items += it;
                ++nv;
                ++k;
            }
            // canonical rows and vertex records
            rows.assign((size_t)nv, 0);
            int acc = 0;
            for (int s = 0; s < nv; ++s) {
                int v = m.color_particles[vb + s];
                rows[s] = acc;
                int nr = ntri[v] + nedge[v] + ntet[v];
                int32_t* vr = &L.vert_rec[(size_t)(vb + s) * kVertWords];
                vr[0] = v;
                vr[1] = acc;
                vr[2] = nr;
                vr[3] = s | ((active[v] ? 1 : 0) << 16) | (nspr[v] << 17);
                acc += nr + nspr[v];
            }
            int32_t rec[kBatchWords];
            rec[0] = vb;
            rec[1] = nv;
            rec[2] = (int32_t)(L.tri_rec.size() / kTriWords);
            for (int s = 0; s < nv; ++s) {
                int v = m.color_particles[vb + s];
                if (!active[v] || m.T == 0) continue;
                int j = 0;
                for (int p = m.adj_tri_off[v]; p < m.adj_tri_off[v + 1]; p += 2) {
                    int t = m.adj_tri[p], slot = m.adj_tri[p + 1];
                    if (!tri_ok[t]) continue;
                    int32_t r[kTriWords];
                    r[0] = m.tri_idx[3 * t];
                    r[1] = m.tri_idx[3 * t + 1];
                    r[2] = m.tri_idx[3 * t + 2];
                    r[3] = meta(rows[s] + j, slot);
                    for (int q = 0; q < 4; ++q) r[4 + q] = fbits(m.tri_pose[4 * t + q]);
                    r[8] = fbits(m.tri_area[t]);
                    r[9] = fbits(m.tri_mat[5 * t]);
                    r[10] = fbits(m.tri_mat[5 * t + 1]);
                    r[11] = fbits(m.tri_mat[5 * t + 2]);
                    L.tri_rec.insert(L.tri_rec.end(), r, r + kTriWords);
                    ++j;
                }
            }
            rec[3] = (int32_t)(L.tri_rec.size() / kTriWords) - rec[2];
            rec[4] = (int32_t)L.edge_meta.size();
            for (int s = 0; s < nv; ++s) {
                int v = m.color_particles[vb + s];
                if (!active[v] || m.E == 0) continue;
                int j = 0;
                for (int p = m.adj_edge_off[v]; p < m.adj_edge_off[v + 1]; p += 2) {
                    int e = m.adj_edge[p], slot = m.adj_edge[p + 1];
                    if (!edge_ok[e]) continue;
                    int32_t r[kEdgeWords];
                    for (int q = 0; q < 4; ++q) r[q] = m.edge_idx[4 * e + q];
                    r[4] = fbits(m.edge_rest[e]);
                    r[5] = fbits(m.edge_len[e]);
                    r[6] = fbits(m.edge_bend[2 * e]);
                    r[7] = fbits(m.edge_bend[2 * e + 1]);
                    L.edge_rec.insert(L.edge_rec.end(), r, r + kEdgeWords);
                    L.edge_meta.push_back(meta(rows[s] + ntri[v] + j, slot));
                    ++j;
                }
            }
            rec[5] = (int32_t)L.edge_meta.size() - rec[4];
            rec[6] = (int32_t)L.tet_meta.size();
            for (int s = 0; s < nv; ++s) {
                int v = m.color_particles[vb + s];
                if (!active[v] || m.K == 0) continue;
                int j = 0;
                for (int p = m.adj_tet_off[v]; p < m.adj_tet_off[v + 1]; p += 2) {
                    int t = m.adj_tet[p], slot = m.adj_tet[p + 1];
                    if (!tet_ok[t]) continue;
                    int32_t r[kTetWords];
                    for (int q = 0; q < 4; ++q) r[q] = m.tet_idx[4 * t + q];
                    for (int q = 0; q < 9; ++q) r[4 + q] = fbits(m.tet_pose[9 * t + q]);
                    r[13] = fbits(m.tet_mat[3 * t]);
                    r[14] = fbits(m.tet_mat[3 * t + 1]);
                    r[15] = fbits(m.tet_mat[3 * t + 2]);
                    L.tet_rec.insert(L.tet_rec.end(), r, r + kTetWords);
                    L.tet_meta.push_back(meta(rows[s] + ntri[v] + nedge[v] + j, slot));
                    ++j;
                }
            }
            rec[7] = (int32_t)L.tet_meta.size() - rec[6];
            rec[8] = (int32_t)(L.spr_rec.size() / kSpringWords);
            for (int s = 0; s < nv; ++s) {
                int v = m.color_particles[vb + s];
                if (!active[v]) continue;
                const int row0 = rows[s] + ntri[v] + nedge[v] + ntet[v];
                for (int p = vs_off[v]; p < vs_off[v + 1]; ++p) {
                    const int sp = vs_item[p] >> 1, end = vs_item[p] & 1;
                    int32_t r[kSpringWords];
                    r[0] = m.spring_idx[2 * sp];
                    r[1] = m.spring_idx[2 * sp + 1];
                    r[2] = meta(row0 + (p - vs_off[v]), end);
                    r[3] = pre_row[p];
                    r[4] = fbits(m.spring_rest[sp]);
                    r[5] = fbits(m.spring_ke[sp]);
                    r[6] = fbits(m.spring_kd[sp]);
                    r[7] = 0;
                    L.spr_rec.insert(L.spr_rec.end(), r, r + kSpringWords);
                }
            }
            rec[9] = (int32_t)(L.spr_rec.size() / kSpringWords) - rec[8];
            rec[10] = 0;
            rec[11] = 0;
            L.fast_items += rec[3] + rec[5] + rec[7] + rec[9];
            const size_t at = L.batch_blk.size();
            L.batch_blk.resize(at + L.blk_words, 0);
            std::memcpy(&L.batch_blk[at], rec, sizeof(rec));
            std::memcpy(&L.batch_blk[at + kBatchWords], &L.vert_rec[(size_t)vb * kVertWords],
                        sizeof(int32_t) * kVertWords * (size_t)nv);
        }
    }
    L.fast_color_off[C] = (int32_t)(L.batch_blk.size() / L.blk_words);
    L.slow_color_off[C] = (int32_t)L.slow_pos.size();
    return std::string();
}

}  // namespace vbd_layout
