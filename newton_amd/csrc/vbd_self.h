// Cloth self-contact on the GPU (SolverVBD particle_enable_self_contact).
// Included by vbd_hip.hip after the vector helpers and store_row.
//
// Reference: newton/_src/solvers/vbd/particle_vbd_kernels.py (K), solver_vbd.py (S),
// newton/_src/geometry/kernels.py (G), newton/_src/math/__init__.py (M).
//
// Detection.  The reference queries a grouped BVH (G:1329-1637).  Here each
// detection builds a hashed uniform grid: every triangle is listed in the cells
// its box, grown by the query radius, overlaps (cell edge = twice the largest
// triangle half-extent plus twice the radius, so at most two cells per axis),
// every edge in the cells of its box grown by the radius plus the largest edge
// half-extent.  A vertex then reads the one cell holding it, an edge the one
// cell holding its midpoint, so no pair is seen twice and no sorting pass is
// needed.  Each candidate goes through the BVH's own leaf test (query box vs
// the primitive's vertex min/max box) and the reference's distance test, so the
// contact set is the reference's.  A vertex's (edge's) list keeps the smallest
// `capacity` triangle (edge) ids in ascending order; the count and minimum
// distance cover every match, as in the reference.
//
// Forces.  After each detection the contacts are listed per color position
// (k_self_items_*), ordered edge-edge terms by (e1, slot) then vertex-triangle
// terms by (vertex, slot).  Before a color pass, k_self_rows evaluates every
// item of the color at the pass-start positions into a global 12-float row
// (contacts may join two vertices of one color, so they cannot be evaluated
// inside the pass); the pass adds a vertex's rows after its springs.
//
// Truncation.  After each color pass every collision computes its division
// plane and the four vertices' truncation factors (K:2133-2263), combined per
// vertex by an integer atomic minimum (exact and order-free for factors in
// [0, 1]); k_trunc_apply then scales and clamps each displacement and writes
// q = pos_cd + d (K:2266-2286).
#pragma once

struct SelfParams {
    int N, T, E, W;              // particles, triangles, edges, worlds
    int world_hash;              // 1: the grid key includes the world (no global particles)
    const int* tri_idx;          // [T,3]
    const int* edge_idx;         // [E,4]
    const int* tri_world;        // [T]
    const int* edge_world;       // [E]
    const int* pworld;           // [N]
    const int* pos_of;           // [N] color position of each particle
    const int* color_particles;  // [N]
    float radius, margin, relax, max_disp, par_eps, rest_excl;
    float ke, kd, mu, friction_eps;
    int vt_cap, ee_cap;
    const float* q_rest;         // [N,3] or null
    const int* vt_filter;        // sorted per vertex, or null
    const int* vt_filter_off;    // [N+1]
    const int* ee_filter;
    const int* ee_filter_off;    // [E+1]
    // detection output
    int* vt_list;                // [N*vt_cap], -1 padded
    int* vt_count;               // [N]
    float* vt_min;               // [N]
    int* tri_min;                // [T] float bits (atomic minimum)
    int* ee_list;                // [E*ee_cap]
    int* ee_count;               // [E]
    float* ee_min;               // [E]
    // grid
    int* ext;                    // [2] float bits: largest triangle / edge half-extent
    int nbucket;                 // power of two
    int* g_count;                // [nbucket + 1] zero between detections (counted up, then back down)
    int* g_off;                  // [nbucket + 1]
    int4* g_ent;                 // [g_cap] (prim, ix, iy, iz)
    float4* g_box;               // [2 * g_cap] the entry's primitive box: (lo.xyz, -), (hi.xyz, -)
    int g_cap;
    int* g_overflow;             // [1]
    // per color position item lists and rows
    int* it_cnt;                 // [N+1]
    int* it_off;                 // [N+1]
    int* it_cur;                 // [N]
    int* ee_n;                   // [N] edge-edge items of each color position
    const int* adj_edge;         // per-particle (edge, slot) CSR, ascending edge ids
    const int* adj_edge_off;     // [N+1]
    int2* items;                 // [it_cap] (key, role)
    float* rows;                 // [it_cap * 12]
    // displacement base and truncation
    float* pos_cd;               // [N,3]
    int* trunc;                  // [N] float bits
    int* coll;                   // [E*ee_cap + N*vt_cap] recorded collision slots (any order)
    int* ncoll;                  // [1]
};

constexpr int kSelfGrid = 2048;  // work-groups of the self-contact grid-stride kernels (a contact-rich pass has a few hundred thousand items: one per lane)

// ---------------------------------------------------------------------------
// geometry and forces, in the oracle's (reference) operation order
// ---------------------------------------------------------------------------
__device__ __forceinline__ float sc_clamp(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }
__device__ __forceinline__ float sc_min(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float sc_max(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ V3 sc_vmin(V3 a, V3 b) { return v3(sc_min(a.x, b.x), sc_min(a.y, b.y), sc_min(a.z, b.z)); }
__device__ __forceinline__ V3 sc_vmax(V3 a, V3 b) { return v3(sc_max(a.x, b.x), sc_max(a.y, b.y), sc_max(a.z, b.z)); }

// triangle_closest_point (G:80-144)
__device__ V3 sc_tri_closest(V3 a, V3 b, V3 c, V3 p, V3& bary) {
    const V3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
    const float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) {
        bary = v3(1.0f, 0.0f, 0.0f);
        return a;
    }
    const V3 bp = sub(p, b);
    const float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) {
        bary = v3(0.0f, 1.0f, 0.0f);
        return b;
    }
    const V3 cp = sub(p, c);
    const float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) {
        bary = v3(0.0f, 0.0f, 1.0f);
        return c;
    }
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        const float v = d1 / (d1 - d3);
        bary = v3(1.0f - v, v, 0.0f);
        return add(a, scl(ab, v));
    }
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        const float v = d2 / (d2 - d6);
        bary = v3(1.0f - v, 0.0f, v);
        return add(a, scl(ac, v));
    }
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        const float v = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        bary = v3(0.0f, 1.0f - v, v);
        return add(b, scl(sub(c, b), v));
    }
    const float denom = 1.0f / ((va + vb) + vc);
    const float v = vb * denom, w = vc * denom;
    bary = v3((1.0f - v) - w, v, w);
    return add(add(a, scl(ab, v)), scl(ac, w));
}

// closest points of two segments (Ericson, Real-Time Collision Detection 5.1.9): (s, t, distance)
__device__ V3 sc_closest_edge_edge(V3 p1, V3 q1, V3 p2, V3 q2, float eps) {
    const V3 d1 = sub(q1, p1), d2 = sub(q2, p2), r = sub(p1, p2);
    const float a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    float s = 0.0f, t = 0.0f;
    if (a <= eps && e <= eps) return v3(0.0f, 0.0f, length(sub(p1, p2)));
    if (a <= eps) {
        s = 0.0f;
        t = sc_clamp(f / e, 0.0f, 1.0f);
    } else {
        const float c = dot(d1, r);
        if (e <= eps) {
            t = 0.0f;
            s = sc_clamp(-c / a, 0.0f, 1.0f);
        } else {
            const float b = dot(d1, d2);
            const float denom = a * e - b * b;
            s = denom != 0.0f ? sc_clamp((b * f - c * e) / denom, 0.0f, 1.0f) : 0.0f;
            t = (b * s + f) / e;
            if (t < 0.0f) {
                t = 0.0f;
                s = sc_clamp(-c / a, 0.0f, 1.0f);
            } else if (t > 1.0f) {
                t = 1.0f;
                s = sc_clamp((b - c) / a, 0.0f, 1.0f);
            }
        }
    }
    const V3 c1 = add(p1, scl(d1, s)), c2 = add(p2, scl(d2, t));
    return v3(s, t, length(sub(c1, c2)));
}

// evaluate_self_contact_force_norm (K:819-848)
__device__ __forceinline__ void sc_force_norm(float dis, float radius, float k, float& dEdD, float& d2E) {
    const float pen = radius - dis;
    const float tau = radius * 0.5f;
    const float d_min = 1.0e-5f;
    if (tau > dis && dis > d_min) {
        const float k2 = (tau * tau) * k;
        dEdD = -k2 / dis;
        d2E = k2 / (dis * dis);
    } else if (dis <= d_min) {
        const float k2 = (tau * tau) * k;
        const float dmsq = d_min * d_min;
        dEdD = (k2 * (dis - 2.0f * d_min)) / dmsq;
        d2E = k2 / dmsq;
    } else {
        dEdD = -k * pen;
        d2E = k;
    }
}

// orthonormal_basis (M:236-273)
__device__ __forceinline__ void sc_basis(V3 n, V3& b1, V3& b2) {
    if (n.z < 0.0f) {
        const float a = 1.0f / (1.0f - n.z);
        const float b = (n.x * n.y) * a;
        b1 = v3(1.0f - (n.x * n.x) * a, -b, n.x);
        b2 = v3(b, (n.y * n.y) * a - 1.0f, -n.y);
    } else {
        const float a = 1.0f / (1.0f + n.z);
        const float b = ((-n.x) * n.y) * a;
        b1 = v3(1.0f - (n.x * n.x) * a, b, -n.x);
        b2 = v3(b, 1.0f - (n.y * n.y) * a, -n.y);
    }
}

// compute_friction (K:1316-1347), T = [b1 b2]
__device__ void sc_friction(float mu, float fn, V3 b1, V3 b2, float u0, float u1, float eps_u, V3& f, M3& h) {
    const float un = sqrtf(u0 * u0 + u1 * u1);
    if (!(un > 0.0f)) {
        f = v3(0.0f, 0.0f, 0.0f);
        h = mzero();
        return;
    }
    const float f1 = un > eps_u ? 1.0f / un : (-un / eps_u + 2.0f) / eps_u;
    const float sf = (-mu) * fn;
    const float w0 = f1 * u0, w1 = f1 * u1;
    const float T[3][2] = {{b1.x, b2.x}, {b1.y, b2.y}, {b1.z, b2.z}};
    f = v3((sf * T[0][0]) * w0 + (sf * T[0][1]) * w1, (sf * T[1][0]) * w0 + (sf * T[1][1]) * w1,
           (sf * T[2][0]) * w0 + (sf * T[2][1]) * w1);
    const float sh = mu * fn;
    float A[3][2];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float m0 = sh * T[i][0], m1 = sh * T[i][1];
        A[i][0] = m0 * f1 + m1 * 0.0f;
        A[i][1] = m0 * 0.0f + m1 * f1;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) h.m[i][j] = A[i][0] * T[j][0] + A[i][1] * T[j][1];
}

// one role's share of a contact: normal force/Hessian scaled by b, damp_collision (K:851-866), friction
__device__ __forceinline__ void sc_role(V3 cf, const M3& ch, V3 ff, const M3& fh, V3 n, float gap_rate, float b,
                                        float kd, float dt, V3& f, M3& h) {
    f = scl(cf, b);
    h = mscl(mscl(ch, b), b);
    V3 df = v3(0.0f, 0.0f, 0.0f);
    M3 dh = mzero();
    if (gap_rate < -1.0e-6f) {
        df = scl(n, ((-kd) * gap_rate) * b);
        dh = mscl(outer(n, n), ((kd / dt) * b) * b);
    }
    f = add(f, add(df, scl(ff, b)));
    h = madd(h, madd(dh, mscl(fh, b * b)));
}

// evaluate_vertex_triangle_collision_force_hessian_4_vertices (K:1195-1313), the row of one role
// (0..2: triangle vertices a, b, c; 3: the colliding vertex p); false when out of contact
__device__ bool sc_vt_role(V3 a, V3 b, V3 c, V3 p, V3 aa, V3 ab, V3 ac, V3 ap, float radius, float ke, float kd,
                           float mu, float fric_eps, float dt, int role, V3& f, M3& h) {
    V3 bary;
    const V3 cp = sc_tri_closest(a, b, c, p, bary);
    const V3 diff = sub(p, cp);
    const float dis = length(diff);
    const V3 n = divs(diff, dis);
    if (!(0.0f < dis && dis < radius)) return false;
    const float bs = role == 0 ? -bary.x : (role == 1 ? -bary.y : (role == 2 ? -bary.z : 1.0f));
    float dEdD, d2E;
    sc_force_norm(dis, radius, ke, dEdD, d2E);
    const V3 cf = scl(n, -dEdD);
    const M3 ch = mscl(outer(n, n), d2E);
    const V3 dx_v = sub(p, ap);
    const V3 cp_prev = add(add(scl(aa, bary.x), scl(ab, bary.y)), scl(ac, bary.z));
    const V3 dx = sub(dx_v, sub(cp, cp_prev));
    V3 e0, e1;
    sc_basis(n, e0, e1);
    const float u0 = (e0.x * dx.x + e0.y * dx.y) + e0.z * dx.z;
    const float u1 = (e1.x * dx.x + e1.y * dx.y) + e1.z * dx.z;
    V3 ff;
    M3 fh;
    sc_friction(mu, -dEdD, e0, e1, u0, u1, fric_eps * dt, ff, fh);
    const float gap_rate = dot(n, dx) / dt;
    sc_role(cf, ch, ff, fh, n, gap_rate, bs, kd, dt, f, h);
    return true;
}

// evaluate_edge_edge_contact_2_vertices (K:990-1111), the row of e1's vertex `role` (0 or 1)
__device__ bool sc_ee_role(V3 p1, V3 q1, V3 p2, V3 q2, V3 ap1, V3 aq1, V3 ap2, V3 aq2, float radius, float ke,
                           float kd, float mu, float fric_eps, float dt, float par_eps, int role, V3& f, M3& h) {
    const V3 st = sc_closest_edge_edge(p1, q1, p2, q2, par_eps);
    const float s = st.x, t = st.y;
    const V3 c1 = add(p1, scl(sub(q1, p1), s)), c2 = add(p2, scl(sub(q2, p2), t));
    const V3 diff = sub(c1, c2);
    const float dis = st.z;
    const V3 n = divs(diff, dis);
    if (!(0.0f < dis && dis < radius)) return false;
    const float bs = role == 0 ? 1.0f - s : s;
    float dEdD, d2E;
    sc_force_norm(dis, radius, ke, dEdD, d2E);
    const V3 cf = scl(n, -dEdD);
    const M3 ch = mscl(outer(n, n), d2E);
    const V3 c1p = add(ap1, scl(sub(aq1, ap1), s));
    const V3 c2p = add(ap2, scl(sub(aq2, ap2), t));
    const V3 dx = sub(sub(c1, c1p), sub(c2, c2p));
    V3 e0, e1;
    sc_basis(n, e0, e1);
    const float u0 = (e0.x * dx.x + e0.y * dx.y) + e0.z * dx.z;
    const float u1 = (e1.x * dx.x + e1.y * dx.y) + e1.z * dx.z;
    V3 ff;
    M3 fh;
    sc_friction(mu, -dEdD, e0, e1, u0, u1, fric_eps * dt, ff, fh);
    const float gap_rate = dot(n, dx) / dt;
    sc_role(cf, ch, ff, fh, n, gap_rate, bs, kd, dt, f, h);
    return true;
}

// ---------------------------------------------------------------------------
// division planes and truncation factors (K:1864-2130)
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool sc_seg_plane(V3 v, V3 dv, V3 n, V3 d, float eps_par, float eps_near, float eps_far) {
    const float nv = dot(n, dv);
    const float num = -dot(n, sub(v, d));
    if (fabsf(nv) < eps_par) return false;
    const float t = num / nv;
    return t >= eps_near && t <= 1.0f + eps_far;
}

// wp.max over vector components: the first, replaced by each strictly larger one
__device__ __forceinline__ float sc_vecmax3(float a, float b, float c) {
    float r = a;
    if (b > r) r = b;
    if (c > r) r = c;
    return r;
}

__device__ void sc_vt_plane(V3 v, V3 dv, V3 t1, V3 dt1, V3 t2, V3 dt2, V3 t3, V3 dt3, bool dummy[4], V3& n_out,
                            V3& d_out) {
    V3 bary;
    const V3 cp = sc_tri_closest(t1, t2, t3, v, bary);
    const V3 n_hat = sub(v, cp);
    if (length(n_hat) < 1e-12f) {
        dummy[0] = dummy[1] = dummy[2] = dummy[3] = false;
        n_out = v3(0.0f, 0.0f, 0.0f);
        d_out = v;
        return;
    }
    const V3 n = divs(n_hat, length(n_hat));
    const float dvn = sc_max(-dot(n, dv), 0.0f);
    float dtn = dot(n, dt1);
    const float x2 = dot(n, dt2), x3 = dot(n, dt3);
    if (x2 > dtn) dtn = x2;
    if (x3 > dtn) dtn = x3;
    if (0.0f > dtn) dtn = 0.0f;
    V3 d;
    if (dtn + dvn == 0.0f) {
        d = add(cp, scl(n_hat, 0.5f));
    } else {
        const float l = sc_clamp(dtn / (dtn + dvn), 0.05f, 0.95f);
        d = add(cp, scl(n_hat, l));
    }
    dummy[0] = dvn == 0.0f ? true : !sc_seg_plane(v, dv, n, d, 1e-6f, -1e-8f, 1e-8f);
    if (dtn == 0.0f) {
        dummy[1] = dummy[2] = dummy[3] = true;
    } else {
        dummy[1] = !sc_seg_plane(t1, dt1, n, d, 1e-6f, -1e-8f, 1e-8f);
        dummy[2] = !sc_seg_plane(t2, dt2, n, d, 1e-6f, -1e-8f, 1e-8f);
        dummy[3] = !sc_seg_plane(t3, dt3, n, d, 1e-6f, -1e-8f, 1e-8f);
    }
    n_out = n;
    d_out = d;
}

// robust_edge_pair_normal (K:1950-2016)
__device__ V3 sc_robust_normal(V3 a0, V3 a1, V3 b0, V3 b1) {
    const float eps = 1.0e-6f;
    V3 dir0 = sub(a1, a0), dir1 = sub(b1, b0);
    const float l0 = length(dir0), l1 = length(dir1);
    dir0 = l0 > eps ? divs(dir0, l0) : v3(0.0f, 0.0f, 0.0f);
    dir1 = l1 > eps ? divs(dir1, l1) : v3(0.0f, 0.0f, 0.0f);
    V3 n = cross(dir0, dir1);
    float ln = length(n);
    if (ln > eps) return divs(n, ln);
    V3 ref = dir0;
    if (length(ref) <= eps) ref = dir1;
    if (length(ref) <= eps) return v3(1.0f, 0.0f, 0.0f);
    const V3 bridge = scl(sub(add(b0, b1), add(a0, a1)), 0.5f);
    const float bl = length(bridge);
    if (bl > eps) {
        n = cross(ref, divs(bridge, bl));
        ln = length(n);
        if (ln > eps) return divs(n, ln);
    }
    V3 axis = v3(1.0f, 0.0f, 0.0f);
    if (fabsf(dot(ref, axis)) > 0.9f) axis = v3(0.0f, 1.0f, 0.0f);
    n = cross(ref, axis);
    ln = length(n);
    if (ln > eps) return divs(n, ln);
    axis = v3(0.0f, 0.0f, 1.0f);
    n = cross(ref, axis);
    ln = length(n);
    if (ln > eps) return divs(n, ln);
    return v3(1.0f, 0.0f, 0.0f);
}

__device__ void sc_ee_plane(V3 a0, V3 da0, V3 a1, V3 da1, V3 b0, V3 db0, V3 b1, V3 db1, bool dummy[4], V3& n_out,
                            V3& d_out) {
    const V3 st = sc_closest_edge_edge(a0, a1, b0, b1, 1e-6f);
    const V3 c1 = add(a0, scl(sub(a1, a0), st.x));
    const V3 c2 = add(b0, scl(sub(b1, b0), st.y));
    const V3 n_hat = sub(c1, c2);
    if (length(n_hat) < 1e-12f) {
        dummy[0] = dummy[1] = dummy[2] = dummy[3] = false;
        n_out = sc_robust_normal(a0, a1, b0, b1);
        d_out = add(scl(c1, 0.5f), scl(c2, 0.5f));
        return;
    }
    const V3 n = divs(n_hat, length(n_hat));
    const float de0 = sc_vecmax3(-dot(n, da0), -dot(n, da1), 0.0f);
    const float de1 = sc_vecmax3(dot(n, db0), dot(n, db1), 0.0f);
    V3 d;
    if (de0 + de1 == 0.0f) {
        d = add(c2, scl(n_hat, 0.5f));
    } else {
        const float l = sc_clamp(de1 / (de1 + de0), 0.05f, 0.95f);
        d = add(c2, scl(n_hat, l));
    }
    if (de0 == 0.0f) {
        dummy[0] = dummy[1] = true;
    } else {
        dummy[0] = !sc_seg_plane(a0, da0, n, d, 1e-6f, -1e-8f, 1e-6f);
        dummy[1] = !sc_seg_plane(a1, da1, n, d, 1e-6f, -1e-8f, 1e-6f);
    }
    if (de1 == 0.0f) {
        dummy[2] = dummy[3] = true;
    } else {
        dummy[2] = !sc_seg_plane(b0, db0, n, d, 1e-6f, -1e-8f, 1e-6f);
        dummy[3] = !sc_seg_plane(b1, db1, n, d, 1e-6f, -1e-8f, 1e-6f);
    }
    n_out = n;
    d_out = d;
}

// planar_truncation_t (K:2113-2130), gamma_min = 1e-3
__device__ __forceinline__ float sc_trunc_t(V3 v, V3 dv, V3 n, V3 d, float eps, float gamma) {
    const float denom = dot(n, dv);
    if (fabsf(denom) < eps) return 1.0f;
    const float t = dot(n, sub(d, v)) / denom;
    if (t < 0.0f) return 1.0f;
    return sc_clamp(sc_min(t * gamma, t - 1e-3f), 0.0f, 1.0f);
}

// ---------------------------------------------------------------------------
// detection grid
// ---------------------------------------------------------------------------
__device__ __forceinline__ int sc_cell(float x, float h) {
    float c = floorf(x / h);
    c = fminf(fmaxf(c, -1048576.0f), 1048576.0f);
    return (int)c;
}
__device__ __forceinline__ int sc_hash(int ix, int iy, int iz, int world, int nbucket) {
    unsigned h = (unsigned)ix * 73856093u ^ (unsigned)iy * 19349663u ^ (unsigned)iz * 83492791u ^
                 (unsigned)(world + 1) * 2654435761u;
    return (int)(h & (unsigned)(nbucket - 1));
}
// mode 0: triangles (cell edge 2 Ht + 2 r, grown by r); mode 1: edges (4 He + 2 r, grown by r + He)
__device__ __forceinline__ float sc_cell_size(const SelfParams& S, int mode) {
    const float ht = __int_as_float(S.ext[0]), he = __int_as_float(S.ext[1]);
    const float r = S.margin;
    const float h = mode == 0 ? 2.0f * ht + 2.0f * r : 4.0f * he + 2.0f * r;
    return fmaxf(h * 1.01f, 1e-6f);
}
__device__ __forceinline__ float sc_grow(const SelfParams& S, int mode) {
    const float he = __int_as_float(S.ext[1]);
    const float g = mode == 0 ? S.margin : S.margin + he;
    return g * 1.001f + 1e-7f;
}

__device__ __forceinline__ void sc_prim_box(const SelfParams& S, const float* q, int mode, int p, V3& lo, V3& hi) {
    if (mode == 0) {
        const int* t = S.tri_idx + 3 * p;
        const V3 a = ld3(q, t[0]), b = ld3(q, t[1]), c = ld3(q, t[2]);
        lo = sc_vmin(sc_vmin(a, b), c);
        hi = sc_vmax(sc_vmax(a, b), c);
    } else {
        const int* e = S.edge_idx + 4 * p;
        const V3 a = ld3(q, e[2]), b = ld3(q, e[3]);
        lo = sc_vmin(a, b);
        hi = sc_vmax(a, b);
    }
}

// largest half-extents, reduced per work-group (256 lanes) before one atomic each
__global__ __launch_bounds__(256) void k_sc_extents(SelfParams S, const float* __restrict__ q) {
    __shared__ float s_t[256], s_e[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float ht = 0.0f, he = 0.0f;
    if (i < S.T) {
        V3 lo, hi;
        sc_prim_box(S, q, 0, i, lo, hi);
        ht = 0.5f * fmaxf(fmaxf(hi.x - lo.x, hi.y - lo.y), hi.z - lo.z);
    }
    if (i < S.E) {
        V3 lo, hi;
        sc_prim_box(S, q, 1, i, lo, hi);
        he = 0.5f * fmaxf(fmaxf(hi.x - lo.x, hi.y - lo.y), hi.z - lo.z);
    }
    s_t[threadIdx.x] = ht;
    s_e[threadIdx.x] = he;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s_t[threadIdx.x] = fmaxf(s_t[threadIdx.x], s_t[threadIdx.x + w]);
            s_e[threadIdx.x] = fmaxf(s_e[threadIdx.x], s_e[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicMax(S.ext + 0, __float_as_int(s_t[0]));
        atomicMax(S.ext + 1, __float_as_int(s_e[0]));
    }
}

// pass 0: count entries per bucket; pass 1: write them, each bucket filled from its end as its count
// goes back down to zero (the order within a bucket does not matter: a query keeps the smallest ids
// by rank, and counts and minima are order-free)
__global__ void k_sc_grid_insert(SelfParams S, const float* __restrict__ q, int mode, int pass) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = mode == 0 ? S.T : S.E;
    if (p >= n) return;
    const float h = sc_cell_size(S, mode), g = sc_grow(S, mode);
    V3 lo, hi;
    sc_prim_box(S, q, mode, p, lo, hi);
    const int x0 = sc_cell(lo.x - g, h), x1 = sc_cell(hi.x + g, h);
    const int y0 = sc_cell(lo.y - g, h), y1 = sc_cell(hi.y + g, h);
    const int z0 = sc_cell(lo.z - g, h), z1 = sc_cell(hi.z + g, h);
    const int w = S.world_hash ? (mode == 0 ? S.tri_world[p] : S.edge_world[p]) : 0;
    const long sx = (long)x1 - x0 + 1, sy = (long)y1 - y0 + 1, sz = (long)z1 - z0 + 1;
    if (sx > 27 || sy > 27 || sz > 27 || sx * sy * sz > 27) {
        S.g_overflow[0] = 1;
        return;
    }
    for (int ix = x0; ix <= x1; ++ix)
        for (int iy = y0; iy <= y1; ++iy)
            for (int iz = z0; iz <= z1; ++iz) {
                const int b = sc_hash(ix, iy, iz, w, S.nbucket);
                if (pass == 0) {
                    atomicAdd(S.g_count + b, 1);
                } else {
                    const int slot = S.g_off[b] + atomicSub(S.g_count + b, 1) - 1;
                    if (slot < S.g_cap) {
                        S.g_ent[slot] = make_int4(p, ix, iy, iz);
                        S.g_box[2 * slot] = make_float4(lo.x, lo.y, lo.z, 0.0f);
                        S.g_box[2 * slot + 1] = make_float4(hi.x, hi.y, hi.z, 0.0f);
                    }
                }
            }
}

__device__ __forceinline__ bool sc_world_ok(int a, int b) { return a < 0 || b < 0 || a == b; }

__device__ __forceinline__ bool sc_filtered(const int* list, const int* off, int row, int key) {
    if (!list) return false;
    int lo = off[row], hi = off[row + 1];
    const int b = lo;
    if (hi <= lo) return false;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (list[mid] <= key)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo > b && list[lo - 1] == key;
}

__device__ __forceinline__ bool sc_box_overlap(V3 la, V3 ha, V3 lb, V3 hb) {
    return !(la.x > hb.x || la.y > hb.y || la.z > hb.z || lb.x > ha.x || lb.y > ha.y || lb.z > ha.z);
}

// keeps the `cap` smallest ids of a row in ascending order
__device__ __forceinline__ void sc_keep_smallest(int* row, int cap, int have, int id) {
    int k = have < cap ? have : cap;
    if (k == cap) {
        if (cap == 0 || id >= row[cap - 1]) return;
        --k;
    }
    while (k > 0 && row[k - 1] > id) {
        row[k] = row[k - 1];
        --k;
    }
    row[k] = id;
}

// Candidate tests of one query (the BVH leaf test, adjacency, filter, rest-shape exclusion and
// the distance test of G:1329-1637).  true and the distance when the candidate collides.
__device__ __forceinline__ bool sc_vt_test(const SelfParams& S, const float* __restrict__ q, int v, V3 x, int wv,
                                           V3 lo, V3 hi, bool use_rest, int t, float& dist) {
    if (!sc_world_ok(wv, S.tri_world[t])) return false;
    const int* ti = S.tri_idx + 3 * t;
    const int i0 = ti[0], i1 = ti[1], i2 = ti[2];
    const V3 A = ld3(q, i0), B = ld3(q, i1), C = ld3(q, i2);
    if (!sc_box_overlap(lo, hi, sc_vmin(sc_vmin(A, B), C), sc_vmax(sc_vmax(A, B), C))) return false;
    if (v == i0 || v == i1 || v == i2) return false;
    // the distance test before the filter list search and the rest-shape test (the tests are
    // independent predicates, so their order leaves the result unchanged)
    V3 bary;
    const V3 cp = sc_tri_closest(A, B, C, x, bary);
    dist = length(sub(cp, x));
    if (!(dist < S.margin)) return false;
    if (sc_filtered(S.vt_filter, S.vt_filter_off, v, t)) return false;
    if (use_rest) {
        const V3 rp = ld3(S.q_rest, v);
        const V3 cr = sc_tri_closest(ld3(S.q_rest, i0), ld3(S.q_rest, i1), ld3(S.q_rest, i2), rp, bary);
        if (length(sub(cr, rp)) < S.rest_excl) return false;
    }
    return true;
}

__device__ __forceinline__ bool sc_ee_test(const SelfParams& S, const float* __restrict__ q, int e, int a0, int a1,
                                           V3 P, V3 Q, int we, V3 lo, V3 hi, bool use_rest, int g, float& dist) {
    if (!sc_world_ok(we, S.edge_world[g])) return false;
    const int b0 = S.edge_idx[4 * g + 2], b1 = S.edge_idx[4 * g + 3];
    const V3 P2 = ld3(q, b0), Q2 = ld3(q, b1);
    if (!sc_box_overlap(lo, hi, sc_vmin(P2, Q2), sc_vmax(P2, Q2))) return false;
    if (a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1) return false;
    const V3 st = sc_closest_edge_edge(P, Q, P2, Q2, S.par_eps);
    dist = st.z;
    if (!(dist < S.margin)) return false;
    if (sc_filtered(S.ee_filter, S.ee_filter_off, e, g)) return false;
    if (use_rest) {
        const V3 sr = sc_closest_edge_edge(ld3(S.q_rest, a0), ld3(S.q_rest, a1), ld3(S.q_rest, b0), ld3(S.q_rest, b1),
                                           S.par_eps);
        if (sr.z < S.rest_excl) return false;
    }
    return true;
}

// A query runs on kQL lanes that stride over its cell's bucket; matches are appended to an LDS
// list (kQCap per query), counts and minima are reduced across the lanes, and the list's
// `capacity` smallest ids are placed by rank (ids are distinct: a cell lists a primitive once).
// A query with more than kQCap matches is redone by its first lane alone (sc_keep_smallest).
constexpr int kQL = 16;
constexpr int kQCap = 128;
constexpr int kQPerBlock = 256 / kQL;

__device__ __forceinline__ void sc_place_by_rank(int* __restrict__ row, int cap, const int* buf, int n, int lane) {
    for (int i = lane; i < n; i += kQL) {
        const int x = buf[i];
        int rank = 0;
        for (int k = 0; k < n; ++k) rank += buf[k] < x ? 1 : 0;
        if (rank < cap) row[rank] = x;
    }
    for (int k = (n < cap ? n : cap) + lane; k < cap; k += kQL) row[k] = -1;
}

// vertex_triangle_collision_detection_kernel (G:1329-1497)
__global__ __launch_bounds__(256) void k_sc_vt_query(SelfParams S, const float* __restrict__ q) {
    __shared__ int s_buf[kQPerBlock][kQCap];
    __shared__ int s_n[kQPerBlock];
    const int lane = threadIdx.x % kQL, qi = threadIdx.x / kQL;
    const int v = blockIdx.x * kQPerBlock + qi;
    if (lane == 0) s_n[qi] = 0;
    __syncthreads();
    const bool valid = v < S.N;
    const float r = S.margin;
    const bool use_rest = S.q_rest && S.rest_excl > 0.0f;
    int cnt = 0;
    float mind = r;
    V3 x = v3(0.0f, 0.0f, 0.0f), lo = x, hi = x;
    int wv = 0, jb = 0, je = 0;
    if (valid) {
        x = ld3(q, v);
        lo = v3(x.x - r, x.y - r, x.z - r);
        hi = v3(x.x + r, x.y + r, x.z + r);
        wv = S.pworld ? S.pworld[v] : 0;
        const float h = sc_cell_size(S, 0);
        const int ix = sc_cell(x.x, h), iy = sc_cell(x.y, h), iz = sc_cell(x.z, h);
        const int b = sc_hash(ix, iy, iz, S.world_hash ? wv : 0, S.nbucket);
        jb = S.g_off[b];
        je = S.g_off[b + 1];
        for (int j = jb + lane; j < je; j += kQL) {
            const int4 en = S.g_ent[j];
            const float4 bl = S.g_box[2 * j], bh = S.g_box[2 * j + 1];
            if (en.y != ix || en.z != iy || en.w != iz) continue;
            // the primitive's box test (the first geometric test of the candidate) from the entry
            if (!sc_box_overlap(lo, hi, v3(bl.x, bl.y, bl.z), v3(bh.x, bh.y, bh.z))) continue;
            float dist;
            if (!sc_vt_test(S, q, v, x, wv, lo, hi, use_rest, en.x, dist)) continue;
            mind = sc_min(mind, dist);
            ++cnt;
            atomicMin(S.tri_min + en.x, __float_as_int(dist));
            const int k = atomicAdd(&s_n[qi], 1);
            if (k < kQCap) s_buf[qi][k] = en.x;
        }
    }
#pragma unroll
    for (int o = kQL / 2; o > 0; o >>= 1) {
        cnt += __shfl_xor(cnt, o, kQL);
        mind = fminf(mind, __shfl_xor(mind, o, kQL));
    }
    __syncthreads();
    if (!valid) return;
    int* row = S.vt_list + (size_t)v * S.vt_cap;
    const int n = s_n[qi];
    if (n <= kQCap) {
        sc_place_by_rank(row, S.vt_cap, s_buf[qi], n, lane);
    } else if (lane == 0) {
        for (int k = 0; k < S.vt_cap; ++k) row[k] = -1;
        const float h = sc_cell_size(S, 0);
        const int ix = sc_cell(x.x, h), iy = sc_cell(x.y, h), iz = sc_cell(x.z, h);
        int have = 0;
        for (int j = jb; j < je; ++j) {
            const int4 en = S.g_ent[j];
            const float4 bl = S.g_box[2 * j], bh = S.g_box[2 * j + 1];
            if (en.y != ix || en.z != iy || en.w != iz) continue;
            // the primitive's box test (the first geometric test of the candidate) from the entry
            if (!sc_box_overlap(lo, hi, v3(bl.x, bl.y, bl.z), v3(bh.x, bh.y, bh.z))) continue;
            float dist;
            if (!sc_vt_test(S, q, v, x, wv, lo, hi, use_rest, en.x, dist)) continue;
            sc_keep_smallest(row, S.vt_cap, have, en.x);
            ++have;
        }
    }
    if (lane == 0) {
        S.vt_count[v] = cnt;
        S.vt_min[v] = mind;
    }
}

// edge_colliding_edges_detection_kernel (G:1500-1637)
__global__ __launch_bounds__(256) void k_sc_ee_query(SelfParams S, const float* __restrict__ q) {
    __shared__ int s_buf[kQPerBlock][kQCap];
    __shared__ int s_n[kQPerBlock];
    const int lane = threadIdx.x % kQL, qi = threadIdx.x / kQL;
    const int e = blockIdx.x * kQPerBlock + qi;
    if (lane == 0) s_n[qi] = 0;
    __syncthreads();
    const bool valid = e < S.E;
    const float r = S.margin;
    const bool use_rest = S.q_rest && S.rest_excl > 0.0f;
    int cnt = 0;
    float mind = r;
    V3 P = v3(0.0f, 0.0f, 0.0f), Q = P, lo = P, hi = P;
    int a0 = 0, a1 = 0, we = 0, jb = 0, je = 0, ix = 0, iy = 0, iz = 0;
    if (valid) {
        a0 = S.edge_idx[4 * e + 2];
        a1 = S.edge_idx[4 * e + 3];
        P = ld3(q, a0);
        Q = ld3(q, a1);
        lo = sc_vmin(P, Q);
        hi = sc_vmax(P, Q);
        lo = v3(lo.x - r, lo.y - r, lo.z - r);
        hi = v3(hi.x + r, hi.y + r, hi.z + r);
        we = S.edge_world[e];
        const float h = sc_cell_size(S, 1);
        const V3 mid = scl(add(P, Q), 0.5f);
        ix = sc_cell(mid.x, h);
        iy = sc_cell(mid.y, h);
        iz = sc_cell(mid.z, h);
        const int b = sc_hash(ix, iy, iz, S.world_hash ? we : 0, S.nbucket);
        jb = S.g_off[b];
        je = S.g_off[b + 1];
        for (int j = jb + lane; j < je; j += kQL) {
            const int4 en = S.g_ent[j];
            const float4 bl = S.g_box[2 * j], bh = S.g_box[2 * j + 1];
            if (en.y != ix || en.z != iy || en.w != iz) continue;
            // the primitive's box test (the first geometric test of the candidate) from the entry
            if (!sc_box_overlap(lo, hi, v3(bl.x, bl.y, bl.z), v3(bh.x, bh.y, bh.z))) continue;
            float dist;
            if (!sc_ee_test(S, q, e, a0, a1, P, Q, we, lo, hi, use_rest, en.x, dist)) continue;
            mind = sc_min(mind, dist);
            ++cnt;
            const int k = atomicAdd(&s_n[qi], 1);
            if (k < kQCap) s_buf[qi][k] = en.x;
        }
    }
#pragma unroll
    for (int o = kQL / 2; o > 0; o >>= 1) {
        cnt += __shfl_xor(cnt, o, kQL);
        mind = fminf(mind, __shfl_xor(mind, o, kQL));
    }
    __syncthreads();
    if (!valid) return;
    int* row = S.ee_list + (size_t)e * S.ee_cap;
    const int n = s_n[qi];
    if (n <= kQCap) {
        sc_place_by_rank(row, S.ee_cap, s_buf[qi], n, lane);
    } else if (lane == 0) {
        for (int k = 0; k < S.ee_cap; ++k) row[k] = -1;
        int have = 0;
        for (int j = jb; j < je; ++j) {
            const int4 en = S.g_ent[j];
            const float4 bl = S.g_box[2 * j], bh = S.g_box[2 * j + 1];
            if (en.y != ix || en.z != iy || en.w != iz) continue;
            // the primitive's box test (the first geometric test of the candidate) from the entry
            if (!sc_box_overlap(lo, hi, v3(bl.x, bl.y, bl.z), v3(bh.x, bh.y, bh.z))) continue;
            float dist;
            if (!sc_ee_test(S, q, e, a0, a1, P, Q, we, lo, hi, use_rest, en.x, dist)) continue;
            sc_keep_smallest(row, S.ee_cap, have, en.x);
            ++have;
        }
    }
    if (lane == 0) {
        S.ee_count[e] = cnt;
        S.ee_min[e] = mind;
    }
}

// ---------------------------------------------------------------------------
// per color position item lists: key = e1 * ee_cap + slot (edge-edge) or
// E * ee_cap + p * vt_cap + slot (vertex-triangle); role = vertex's place in the contact
// ---------------------------------------------------------------------------
// Per color position: the edge-edge items first, in (e1, slot) order -- written by a walk over the
// particle's edges (its CSR lists them in ascending id) -- then the vertex-triangle items in
// (p, slot) order, appended by atomics and sorted (short lists: a vertex's own contacts and
// those of the few triangles around it).
// Edge-edge item counts of each color position, by the walk over the particle's edges that
// k_sc_items_ee writes the items with (no atomics: a contact-rich detection has a few hundred
// thousand edge-edge contacts, two items each).
__global__ void k_sc_items_count_ee(SelfParams S) {
    const int pos = blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= S.N) return;
    const int v = S.color_particles[pos];
    int n = 0;
    if (S.E > 0) {
        for (int c = S.adj_edge_off[v]; c < S.adj_edge_off[v + 1]; c += 2) {
            if (S.adj_edge[c + 1] < 2) continue;
            const int e1 = S.adj_edge[c];
            n += S.ee_count[e1] < S.ee_cap ? S.ee_count[e1] : S.ee_cap;
        }
    }
    S.it_cnt[pos] = n;
    S.ee_n[pos] = n;
}

// vertex-triangle item counts (added to the edge-edge counts): one per corner and one for the vertex
__global__ void k_sc_items_count(SelfParams S) {
    const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= (long)S.N * S.vt_cap) return;
    const int p = (int)(j / S.vt_cap), k = (int)(j - (long)p * S.vt_cap);
    if (k >= S.vt_count[p]) return;
    const int* t = S.tri_idx + 3 * S.vt_list[j];
    atomicAdd(S.it_cnt + S.pos_of[t[0]], 1);
    atomicAdd(S.it_cnt + S.pos_of[t[1]], 1);
    atomicAdd(S.it_cnt + S.pos_of[t[2]], 1);
    atomicAdd(S.it_cnt + S.pos_of[p], 1);
}

// edge-edge items of each color position by walking the particle's edges; sets the cursor of its
// vertex-triangle items
__global__ void k_sc_items_ee(SelfParams S) {
    const int pos = blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= S.N) return;
    const int v = S.color_particles[pos];
    int j = S.it_off[pos];
    if (S.E > 0) {
        for (int c = S.adj_edge_off[v]; c < S.adj_edge_off[v + 1]; c += 2) {
            const int slot = S.adj_edge[c + 1];
            if (slot < 2) continue;
            const int e1 = S.adj_edge[c];
            const int n = S.ee_count[e1] < S.ee_cap ? S.ee_count[e1] : S.ee_cap;
            for (int k = 0; k < n; ++k) S.items[j++] = make_int2(e1 * S.ee_cap + k, slot - 2);
        }
    }
    S.it_cur[pos] = j;
}

__device__ __forceinline__ void sc_vt_put(const SelfParams& S, int v, int key, int role) {
    const int slot = atomicAdd(S.it_cur + S.pos_of[v], 1);
    S.items[slot] = make_int2(key, role);
}

__global__ void k_sc_items_vt(SelfParams S) {
    const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= (long)S.N * S.vt_cap) return;
    const int p = (int)(j / S.vt_cap), k = (int)(j - (long)p * S.vt_cap);
    if (k >= S.vt_count[p]) return;
    const int* t = S.tri_idx + 3 * S.vt_list[j];
    const int key = (int)((long)S.E * S.ee_cap + j);
    sc_vt_put(S, t[0], key, 0);
    sc_vt_put(S, t[1], key, 1);
    sc_vt_put(S, t[2], key, 2);
    sc_vt_put(S, p, key, 3);
}

// insertion sort of each color position's vertex-triangle items by key (keys of one position are
// distinct); lists of up to kSortCap items are sorted in the lane's LDS column, longer ones in place
constexpr int kSortCap = 32;
__global__ __launch_bounds__(256) void k_sc_items_sort(SelfParams S) {
    __shared__ int2 s_it[kSortCap * 256];  // element k of lane l at k * 256 + l
    const int pos = blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= S.N) return;
    const int b = S.it_off[pos] + S.ee_n[pos], e = S.it_off[pos + 1];
    const int n = e - b;
    if (n <= 1) return;
    if (n <= kSortCap) {
        int2* L = s_it + threadIdx.x;
        for (int i = 0; i < n; ++i) L[i * 256] = S.items[b + i];
        for (int i = 1; i < n; ++i) {
            const int2 x = L[i * 256];
            int j = i;
            while (j > 0 && L[(j - 1) * 256].x > x.x) {
                L[j * 256] = L[(j - 1) * 256];
                --j;
            }
            L[j * 256] = x;
        }
        for (int i = 0; i < n; ++i) S.items[b + i] = L[i * 256];
        return;
    }
    for (int i = b + 1; i < e; ++i) {
        const int2 x = S.items[i];
        int j = i;
        while (j > b && S.items[j - 1].x > x.x) {
            S.items[j] = S.items[j - 1];
            --j;
        }
        S.items[j] = x;
    }
}

// rows of the items of color positions [pos0, pos1) at the positions q (the pass start) and
// anchors q_prev; the item range is read on the device (grid-stride)
__device__ void sc_row(const SelfParams& S, const float* __restrict__ q, const float* __restrict__ q_prev, int j,
                       float dt);

__global__ void k_sc_rows(SelfParams S, const float* __restrict__ q, const float* __restrict__ q_prev, int pos0,
                          int pos1, float dt) {
    const int j0 = S.it_off[pos0], j1 = S.it_off[pos1];
    for (int j = j0 + blockIdx.x * blockDim.x + threadIdx.x; j < j1; j += gridDim.x * blockDim.x)
        sc_row(S, q, q_prev, j, dt);
}

__device__ void sc_row(const SelfParams& S, const float* __restrict__ q, const float* __restrict__ q_prev, int j,
                       float dt) {
    const int2 it = S.items[j];
    const long nee = (long)S.E * S.ee_cap;
    V3 f = v3(0.0f, 0.0f, 0.0f);
    M3 h = mzero();
    bool has;
    if (it.x < nee) {
        const int e1 = it.x / S.ee_cap;
        const int e2 = S.ee_list[it.x];
        const int* a = S.edge_idx + 4 * e1;
        const int* b = S.edge_idx + 4 * e2;
        has = sc_ee_role(ld3(q, a[2]), ld3(q, a[3]), ld3(q, b[2]), ld3(q, b[3]), ld3(q_prev, a[2]), ld3(q_prev, a[3]),
                         ld3(q_prev, b[2]), ld3(q_prev, b[3]), S.radius, S.ke, S.kd, S.mu, S.friction_eps, dt,
                         S.par_eps, it.y, f, h);
    } else {
        const long jj = it.x - nee;
        const int p = (int)(jj / S.vt_cap);
        const int* t = S.tri_idx + 3 * S.vt_list[jj];
        has = sc_vt_role(ld3(q, t[0]), ld3(q, t[1]), ld3(q, t[2]), ld3(q, p), ld3(q_prev, t[0]), ld3(q_prev, t[1]),
                         ld3(q_prev, t[2]), ld3(q_prev, p), S.radius, S.ke, S.kd, S.mu, S.friction_eps, dt, it.y, f, h);
    }
    if (!has) {
        f = v3(0.0f, 0.0f, 0.0f);
        h = mzero();
    }
    store_row(S.rows + 12 * (size_t)j, f, h);
}

// ---------------------------------------------------------------------------
// truncation
// ---------------------------------------------------------------------------
// compact list of the recorded collision slots (edge-edge slots first in the slot space)
__device__ __forceinline__ bool sc_slot_recorded(const SelfParams& S, long i) {
    const long nee = (long)S.E * S.ee_cap;
    if (i < nee) {
        const int e1 = (int)(i / S.ee_cap), k = (int)(i - (long)e1 * S.ee_cap);
        return k < S.ee_count[e1];
    }
    const long j = i - nee;
    if (j >= (long)S.N * S.vt_cap) return false;
    const int p = (int)(j / S.vt_cap), k = (int)(j - (long)p * S.vt_cap);
    return k < S.vt_count[p];
}

// kCollectPer slots per lane, one counter atomic per work-group (a contact-rich detection records a
// few hundred thousand slots, which one atomic each -- or one per wave, a wave mostly covering one
// edge's list -- would serialise on the counter)
constexpr int kCollectPer = 16;
__global__ __launch_bounds__(256) void k_sc_collect(SelfParams S) {
    __shared__ int s_cnt[256];
    __shared__ int s_base;
    const int t = threadIdx.x;
    const long nslots = (long)S.E * S.ee_cap + (long)S.N * S.vt_cap;
    const long b0 = (long)blockIdx.x * 256 * kCollectPer;
    int mine = 0;
    unsigned rec = 0u;
    for (int k = 0; k < kCollectPer; ++k) {
        const long i = b0 + (long)k * 256 + t;
        if (i < nslots && sc_slot_recorded(S, i)) {
            rec |= 1u << k;
            ++mine;
        }
    }
    // inclusive scan of the lanes' counts
    s_cnt[t] = mine;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        const int v = t >= o ? s_cnt[t - o] : 0;
        __syncthreads();
        s_cnt[t] += v;
        __syncthreads();
    }
    const int total = s_cnt[255];
    if (total == 0) return;  // uniform
    if (t == 0) s_base = atomicAdd(S.ncoll, total);
    __syncthreads();
    int w = s_base + s_cnt[t] - mine;
    for (int k = 0; k < kCollectPer; ++k)
        if ((rec >> k) & 1u) S.coll[w++] = (int)(b0 + (long)k * 256 + t);
}

__device__ void sc_trunc_one(const SelfParams& S, const float* __restrict__ disp, long i);

__global__ void k_sc_trunc_collisions(SelfParams S, const float* __restrict__ disp) {
    const int n = S.ncoll[0];
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x)
        sc_trunc_one(S, disp, S.coll[c]);
}

__device__ void sc_trunc_one(const SelfParams& S, const float* __restrict__ disp, long i) {
    const long nee = (long)S.E * S.ee_cap;
    int who[4];
    int kind;
    if (i < nee) {
        const int e1 = (int)(i / S.ee_cap);
        const int* a = S.edge_idx + 4 * e1;
        const int* b = S.edge_idx + 4 * S.ee_list[i];
        who[0] = a[2];
        who[1] = a[3];
        who[2] = b[2];
        who[3] = b[3];
        kind = 0;
    } else {
        const long j = i - nee;
        const int p = (int)(j / S.vt_cap);
        const int* t = S.tri_idx + 3 * S.vt_list[j];
        who[0] = p;
        who[1] = t[0];
        who[2] = t[1];
        who[3] = t[2];
        kind = 1;
    }
    V3 x[4], dx[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        x[r] = ld3(S.pos_cd, who[r]);
        dx[r] = ld3(disp, who[r]);
    }
    bool dummy[4];
    V3 n, d;
    if (kind == 0)
        sc_ee_plane(x[0], dx[0], x[1], dx[1], x[2], dx[2], x[3], dx[3], dummy, n, d);
    else
        sc_vt_plane(x[0], dx[0], x[1], dx[1], x[2], dx[2], x[3], dx[3], dummy, n, d);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        if (dummy[r]) continue;
        const float t = sc_trunc_t(x[r], dx[r], n, d, S.par_eps, S.relax);
        atomicMin(S.trunc + who[r], __float_as_int(t));
    }
}

// apply_truncation_ts (K:2266-2286) with the isotropic clamp
__global__ void k_sc_trunc_apply(SelfParams S, float* __restrict__ disp, float* __restrict__ q) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S.N) return;
    const float t = __int_as_float(S.trunc[i]);
    V3 d = scl(ld3(disp, i), t);
    const float len = length(d);
    if (len > S.max_disp) d = divs(scl(d, S.max_disp), len);
    st3(disp, i, d);
    st3(q, i, add(ld3(S.pos_cd, i), d));
    S.trunc[i] = __float_as_int(1.0f);
}

__global__ void k_sc_copy3(float* __restrict__ dst, const float* __restrict__ src, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 3 * n) dst[i] = src[i];
}
