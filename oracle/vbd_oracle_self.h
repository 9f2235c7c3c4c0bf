/*
 * CPU oracle, cloth self-contact part -- TEST INFRASTRUCTURE (the checker, never the product).
 * Included by vbd_oracle_impl.h (same REAL / SFX instantiation, same helpers).
 *
 * Citations (CSYSLAM/newton):
 *   K = newton/_src/solvers/vbd/particle_vbd_kernels.py
 *   S = newton/_src/solvers/vbd/solver_vbd.py
 *   G = newton/_src/geometry/kernels.py
 *   M = newton/_src/math/__init__.py
 *
 * Detection is a brute-force loop over all primitives with the BVH query's
 * box-overlap test applied explicitly (G:1207-1249: leaf boxes are the
 * primitives' vertex min/max, the query box the point or edge box grown by the
 * query radius), so the contact set equals the reference's; lists are kept in
 * ascending primitive order.  wp.closest_point_edge_edge is a Warp builtin
 * (absent): it is restated from Ericson, Real-Time Collision Detection 5.1.9
 * (ClosestPtSegmentSegment), `eps` taken as the degenerate squared-length
 * threshold -- parity with the builtin itself unpinned.
 */

typedef struct {
    REAL s, t, d;
} EEClosest;

static REAL rclamp(REAL x, REAL lo, REAL hi) { return x < lo ? lo : (x > hi ? hi : x); }
static V3 vmin3(V3 a, V3 b) { return v3(rmin(a.x, b.x), rmin(a.y, b.y), rmin(a.z, b.z)); }
static V3 vmax3(V3 a, V3 b) { return v3(rmax(a.x, b.x), rmax(a.y, b.y), rmax(a.z, b.z)); }

/* triangle_closest_point (G:80-144): closest point, barycentrics, feature (0..6) */
static V3 tri_closest(V3 a, V3 b, V3 c, V3 p, V3* bary, int* feature) {
    V3 ab = vsub(b, a), ac = vsub(c, a), ap = vsub(p, a);
    REAL d1 = vdot(ab, ap), d2 = vdot(ac, ap);
    if (d1 <= R(0.0) && d2 <= R(0.0)) {
        *bary = v3(R(1.0), R(0.0), R(0.0));
        *feature = 0;
        return a;
    }
    V3 bp = vsub(p, b);
    REAL d3 = vdot(ab, bp), d4 = vdot(ac, bp);
    if (d3 >= R(0.0) && d4 <= d3) {
        *bary = v3(R(0.0), R(1.0), R(0.0));
        *feature = 1;
        return b;
    }
    V3 cp = vsub(p, c);
    REAL d5 = vdot(ab, cp), d6 = vdot(ac, cp);
    if (d6 >= R(0.0) && d5 <= d6) {
        *bary = v3(R(0.0), R(0.0), R(1.0));
        *feature = 2;
        return c;
    }
    REAL vc = d1 * d4 - d3 * d2;
    if (vc <= R(0.0) && d1 >= R(0.0) && d3 <= R(0.0)) {
        REAL v = d1 / (d1 - d3);
        *bary = v3(R(1.0) - v, v, R(0.0));
        *feature = 3;
        return vadd(a, vscale(ab, v));
    }
    REAL vb = d5 * d2 - d1 * d6;
    if (vb <= R(0.0) && d2 >= R(0.0) && d6 <= R(0.0)) {
        REAL v = d2 / (d2 - d6);
        *bary = v3(R(1.0) - v, R(0.0), v);
        *feature = 4;
        return vadd(a, vscale(ac, v));
    }
    REAL va = d3 * d6 - d5 * d4;
    if (va <= R(0.0) && (d4 - d3) >= R(0.0) && (d5 - d6) >= R(0.0)) {
        REAL v = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        *bary = v3(R(0.0), R(1.0) - v, v);
        *feature = 5;
        return vadd(b, vscale(vsub(c, b), v));
    }
    REAL denom = R(1.0) / ((va + vb) + vc);
    REAL v = vb * denom, w = vc * denom;
    *bary = v3((R(1.0) - v) - w, v, w);
    *feature = 6;
    return vadd(vadd(a, vscale(ab, v)), vscale(ac, w));
}

/* wp.closest_point_edge_edge: (s, t, |c1 - c2|), Ericson 5.1.9 */
static EEClosest closest_edge_edge(V3 p1, V3 q1, V3 p2, V3 q2, REAL eps) {
    V3 d1 = vsub(q1, p1), d2 = vsub(q2, p2), r = vsub(p1, p2);
    REAL a = vdot(d1, d1), e = vdot(d2, d2), f = vdot(d2, r);
    REAL s = R(0.0), t = R(0.0);
    EEClosest o;
    if (a <= eps && e <= eps) {
        o.s = R(0.0);
        o.t = R(0.0);
        o.d = vlength(vsub(p1, p2));
        return o;
    }
    if (a <= eps) {
        s = R(0.0);
        t = rclamp(f / e, R(0.0), R(1.0));
    } else {
        REAL c = vdot(d1, r);
        if (e <= eps) {
            t = R(0.0);
            s = rclamp(-c / a, R(0.0), R(1.0));
        } else {
            REAL b = vdot(d1, d2);
            REAL denom = a * e - b * b;
            s = denom != R(0.0) ? rclamp((b * f - c * e) / denom, R(0.0), R(1.0)) : R(0.0);
            t = (b * s + f) / e;
            if (t < R(0.0)) {
                t = R(0.0);
                s = rclamp(-c / a, R(0.0), R(1.0));
            } else if (t > R(1.0)) {
                t = R(1.0);
                s = rclamp((b - c) / a, R(0.0), R(1.0));
            }
        }
    }
    V3 c1 = vadd(p1, vscale(d1, s)), c2 = vadd(p2, vscale(d2, t));
    o.s = s;
    o.t = t;
    o.d = vlength(vsub(c1, c2));
    return o;
}

/* evaluate_self_contact_force_norm (K:819-848): (dE/dD, d2E/dD2) */
static void self_force_norm(REAL dis, REAL radius, REAL k, REAL* dEdD, REAL* d2E) {
    REAL pen = radius - dis;
    REAL tau = radius * R(0.5);
    REAL d_min = R(1.0e-5);
    if (tau > dis && dis > d_min) {
        REAL k2 = (tau * tau) * k;
        *dEdD = -k2 / dis;
        *d2E = k2 / (dis * dis);
    } else if (dis <= d_min) {
        REAL k2 = (tau * tau) * k;
        REAL dmsq = d_min * d_min;
        *dEdD = (k2 * (dis - R(2.0) * d_min)) / dmsq;
        *d2E = k2 / dmsq;
    } else {
        *dEdD = -k * pen;
        *d2E = k;
    }
}

/* damp_collision (K:851-866) */
static void damp_collision(REAL gap_rate, REAL b, V3 n, REAL kd, REAL dt, V3* f, M3* h) {
    if (gap_rate < R(-1.0e-6)) {
        *f = vscale(n, ((-kd) * gap_rate) * b);
        *h = mscale(mouter(n, n), ((kd / dt) * b) * b);
    } else {
        *f = v3(R(0.0), R(0.0), R(0.0));
        *h = mzero();
    }
}

/* orthonormal_basis (M:236-273) */
static void orthonormal_basis(V3 n, V3* b1, V3* b2) {
    if (n.z < R(0.0)) {
        REAL a = R(1.0) / (R(1.0) - n.z);
        REAL b = (n.x * n.y) * a;
        *b1 = v3(R(1.0) - (n.x * n.x) * a, -b, n.x);
        *b2 = v3(b, (n.y * n.y) * a - R(1.0), -n.y);
    } else {
        REAL a = R(1.0) / (R(1.0) + n.z);
        REAL b = ((-n.x) * n.y) * a;
        *b1 = v3(R(1.0) - (n.x * n.x) * a, b, -n.x);
        *b2 = v3(b, R(1.0) - (n.y * n.y) * a, -n.y);
    }
}

/* compute_friction (K:1316-1347) with T = [b1 b2] (3x2, rows (b1_i, b2_i)) */
static void self_friction(REAL mu, REAL fn, V3 b1, V3 b2, REAL u0, REAL u1, REAL eps_u, V3* f, M3* h) {
    REAL un = RSQRT(u0 * u0 + u1 * u1);
    if (!(un > R(0.0))) {
        *f = v3(R(0.0), R(0.0), R(0.0));
        *h = mzero();
        return;
    }
    REAL f1 = un > eps_u ? R(1.0) / un : (-un / eps_u + R(2.0)) / eps_u;
    /* force = ((-mu * fn) * T) * (f1 * u) */
    REAL sf = (-mu) * fn;
    REAL w0 = f1 * u0, w1 = f1 * u1;
    REAL T[3][2] = {{b1.x, b2.x}, {b1.y, b2.y}, {b1.z, b2.z}};
    REAL Fr[3];
    for (int i = 0; i < 3; ++i) Fr[i] = (sf * T[i][0]) * w0 + (sf * T[i][1]) * w1;
    *f = v3(Fr[0], Fr[1], Fr[2]);
    /* hessian = (((mu * fn) * T) * (f1 * I2)) * T^T */
    REAL sh = mu * fn;
    REAL A[3][2];
    for (int i = 0; i < 3; ++i) {
        REAL m0 = sh * T[i][0], m1 = sh * T[i][1];
        A[i][0] = m0 * f1 + m1 * R(0.0);
        A[i][1] = m0 * R(0.0) + m1 * f1;
    }
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = A[i][0] * T[j][0] + A[i][1] * T[j][1];
    *h = r;
}

/* evaluate_vertex_triangle_collision_force_hessian_4_vertices (K:1195-1313): roles 0..2 the
 * triangle's vertices a, b, c, role 3 the colliding vertex p. */
static int vt_contact_eval(V3 a, V3 b, V3 c, V3 p, V3 aa, V3 ab, V3 ac, V3 ap, REAL radius, REAL ke, REAL kd,
                           REAL mu, REAL fric_eps, REAL dt, V3 f[4], M3 h[4]) {
    V3 bary;
    int feature;
    V3 cp = tri_closest(a, b, c, p, &bary, &feature);
    V3 diff = vsub(p, cp);
    REAL dis = vlength(diff);
    V3 n = vdivs(diff, dis);
    if (!(R(0.0) < dis && dis < radius)) return 0;
    REAL bs[4] = {-bary.x, -bary.y, -bary.z, R(1.0)};
    REAL dEdD, d2E;
    self_force_norm(dis, radius, ke, &dEdD, &d2E);
    V3 cf = vscale(n, -dEdD);
    M3 ch = mscale(mouter(n, n), d2E);
    V3 dx_v = vsub(p, ap);
    V3 cp_prev = vadd(vadd(vscale(aa, bary.x), vscale(ab, bary.y)), vscale(ac, bary.z));
    V3 dx = vsub(dx_v, vsub(cp, cp_prev));
    V3 e0, e1;
    orthonormal_basis(n, &e0, &e1);
    REAL u0 = (e0.x * dx.x + e0.y * dx.y) + e0.z * dx.z;
    REAL u1 = (e1.x * dx.x + e1.y * dx.y) + e1.z * dx.z;
    REAL eps_u = fric_eps * dt;
    V3 ff;
    M3 fh;
    self_friction(mu, -dEdD, e0, e1, u0, u1, eps_u, &ff, &fh);
    REAL gap_rate = vdot(n, dx) / dt;
    for (int k = 0; k < 4; ++k) {
        V3 df;
        M3 dh;
        f[k] = vscale(cf, bs[k]);
        h[k] = mscale(mscale(ch, bs[k]), bs[k]);
        damp_collision(gap_rate, bs[k], n, kd, dt, &df, &dh);
        f[k] = vadd(f[k], vadd(df, vscale(ff, bs[k])));
        h[k] = madd(h[k], madd(dh, mscale(fh, bs[k] * bs[k])));
    }
    return 1;
}

/* evaluate_edge_edge_contact_2_vertices (K:990-1111): forces on e1's two vertices */
static int ee_contact_eval(V3 p1, V3 q1, V3 p2, V3 q2, V3 ap1, V3 aq1, V3 ap2, V3 aq2, REAL radius, REAL ke,
                           REAL kd, REAL mu, REAL fric_eps, REAL dt, REAL par_eps, V3 f[2], M3 h[2]) {
    EEClosest st = closest_edge_edge(p1, q1, p2, q2, par_eps);
    REAL s = st.s, t = st.t;
    V3 e1v = vsub(q1, p1), e2v = vsub(q2, p2);
    V3 c1 = vadd(p1, vscale(e1v, s)), c2 = vadd(p2, vscale(e2v, t));
    V3 diff = vsub(c1, c2);
    REAL dis = st.d;
    V3 n = vdivs(diff, dis);
    if (!(R(0.0) < dis && dis < radius)) return 0;
    REAL bs[2] = {R(1.0) - s, s};
    REAL dEdD, d2E;
    self_force_norm(dis, radius, ke, &dEdD, &d2E);
    V3 cf = vscale(n, -dEdD);
    M3 ch = mscale(mouter(n, n), d2E);
    V3 c1p = vadd(ap1, vscale(vsub(aq1, ap1), s));
    V3 c2p = vadd(ap2, vscale(vsub(aq2, ap2), t));
    V3 dx = vsub(vsub(c1, c1p), vsub(c2, c2p));
    V3 e0, e1;
    orthonormal_basis(n, &e0, &e1);
    REAL u0 = (e0.x * dx.x + e0.y * dx.y) + e0.z * dx.z;
    REAL u1 = (e1.x * dx.x + e1.y * dx.y) + e1.z * dx.z;
    REAL eps_u = fric_eps * dt;
    V3 ff;
    M3 fh;
    self_friction(mu, -dEdD, e0, e1, u0, u1, eps_u, &ff, &fh);
    REAL gap_rate = vdot(n, dx) / dt;
    for (int k = 0; k < 2; ++k) {
        V3 df;
        M3 dh;
        f[k] = vscale(cf, bs[k]);
        h[k] = mscale(mscale(ch, bs[k]), bs[k]);
        damp_collision(gap_rate, bs[k], n, kd, dt, &df, &dh);
        f[k] = vadd(f[k], vadd(df, vscale(ff, bs[k])));
        h[k] = madd(h[k], madd(dh, mscale(fh, bs[k] * bs[k])));
    }
    return 1;
}

/* ---- detection (G:1329-1497 vertex-triangle, G:1500-1637 edge-edge) ------- */
static int world_ok(int wa, int wb) { return wa < 0 || wb < 0 || wa == wb; }

static int sorted_contains(const int32_t* list, int b, int e, int key) {
    if (!list || e <= b) return 0;
    int lo = b, hi = e;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (list[mid] <= key)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo > b && list[lo - 1] == key;
}

static int box_overlap(V3 lo_a, V3 hi_a, V3 lo_b, V3 hi_b) {
    return !(lo_a.x > hi_b.x || lo_a.y > hi_b.y || lo_a.z > hi_b.z || lo_b.x > hi_a.x || lo_b.y > hi_a.y ||
             lo_b.z > hi_a.z);
}

typedef struct {
    int vt_cap, ee_cap;
    int32_t* vt;     /* [N*vt_cap] */
    int32_t* vt_n;   /* [N] */
    int32_t* ee;     /* [E*ee_cap] */
    int32_t* ee_n;   /* [E] */
    double* vt_min;  /* [N] */
    double* tri_min; /* [T] */
    double* ee_min;  /* [E] */
} SelfLists;

static void self_lists_alloc(const OracleModel* m, const OracleSelf* sc, SelfLists* L) {
    int N = m->particle_count, T = m->tri_count, E = m->edge_count;
    L->vt_cap = sc->vt_capacity;
    L->ee_cap = sc->ee_capacity;
    L->vt = (int32_t*)malloc(sizeof(int32_t) * (size_t)(N > 0 ? N : 1) * (size_t)(L->vt_cap > 0 ? L->vt_cap : 1));
    L->vt_n = (int32_t*)calloc((size_t)(N > 0 ? N : 1), sizeof(int32_t));
    L->ee = (int32_t*)malloc(sizeof(int32_t) * (size_t)(E > 0 ? E : 1) * (size_t)(L->ee_cap > 0 ? L->ee_cap : 1));
    L->ee_n = (int32_t*)calloc((size_t)(E > 0 ? E : 1), sizeof(int32_t));
    L->vt_min = (double*)calloc((size_t)(N > 0 ? N : 1), sizeof(double));
    L->tri_min = (double*)calloc((size_t)(T > 0 ? T : 1), sizeof(double));
    L->ee_min = (double*)calloc((size_t)(E > 0 ? E : 1), sizeof(double));
}

static void self_lists_free(SelfLists* L) {
    free(L->vt);
    free(L->vt_n);
    free(L->ee);
    free(L->ee_n);
    free(L->vt_min);
    free(L->tri_min);
    free(L->ee_min);
}

static V3 rest_pos(const OracleSelf* sc, int i) {
    return v3(R(sc->q_rest[3 * i]), R(sc->q_rest[3 * i + 1]), R(sc->q_rest[3 * i + 2]));
}

/* TriMeshCollisionDetector.vertex_triangle_collision_detection + edge_edge_collision_detection
 * with query radius = margin and min_query_radius = the rest exclusion radius (S:3221-3237) */
static void self_detect(const OracleModel* m, const OracleSelf* sc, const REAL* pos, SelfLists* L) {
    const int N = m->particle_count, T = m->tri_count, E = m->edge_count;
    const REAL r = R(sc->margin), rmin_q = R(sc->rest_exclusion_radius);
    const int use_rest = sc->q_rest != NULL && rmin_q > R(0.0);
    for (int t = 0; t < T; ++t) L->tri_min[t] = (double)r;
    for (int v = 0; v < N; ++v) {
        V3 x = vld(pos, v);
        V3 lo = v3(x.x - r, x.y - r, x.z - r), hi = v3(x.x + r, x.y + r, x.z + r);
        int wv = m->particle_world ? m->particle_world[v] : 0;
        int cnt = 0;
        REAL mind = r;
        for (int k = 0; k < L->vt_cap; ++k) L->vt[(size_t)v * L->vt_cap + k] = -1;
        for (int t = 0; t < T; ++t) {
            const int32_t* ti = m->tri_indices + 3 * t;
            int wt = m->particle_world ? m->particle_world[ti[0]] : 0;
            if (!world_ok(wv, wt)) continue;
            V3 A = vld(pos, ti[0]), B = vld(pos, ti[1]), C = vld(pos, ti[2]);
            if (!box_overlap(lo, hi, vmin3(vmin3(A, B), C), vmax3(vmax3(A, B), C))) continue;
            if (v == ti[0] || v == ti[1] || v == ti[2]) continue;
            if (sc->vt_filter && sorted_contains(sc->vt_filter, sc->vt_filter_offsets[v], sc->vt_filter_offsets[v + 1], t))
                continue;
            V3 bary;
            int feature;
            V3 cp = tri_closest(A, B, C, x, &bary, &feature);
            REAL dist = vlength(vsub(cp, x));
            if (use_rest) {
                V3 rp = rest_pos(sc, v);
                V3 cr = tri_closest(rest_pos(sc, ti[0]), rest_pos(sc, ti[1]), rest_pos(sc, ti[2]), rp, &bary, &feature);
                if (vlength(vsub(cr, rp)) < rmin_q) continue;
            }
            if (dist < r) {
                mind = rmin(mind, dist);
                if (cnt < L->vt_cap) L->vt[(size_t)v * L->vt_cap + cnt] = t;
                ++cnt;
                if ((double)dist < L->tri_min[t]) L->tri_min[t] = (double)dist;
            }
        }
        L->vt_n[v] = cnt;
        L->vt_min[v] = (double)mind;
    }
    for (int e = 0; e < E; ++e) {
        const int32_t* ei = m->edge_indices + 4 * e;
        V3 P = vld(pos, ei[2]), Q = vld(pos, ei[3]);
        V3 lo = vmin3(P, Q), hi = vmax3(P, Q);
        lo = v3(lo.x - r, lo.y - r, lo.z - r);
        hi = v3(hi.x + r, hi.y + r, hi.z + r);
        int we = m->particle_world ? m->particle_world[ei[2]] : 0;
        int cnt = 0;
        REAL mind = r;
        for (int k = 0; k < L->ee_cap; ++k) L->ee[(size_t)e * L->ee_cap + k] = -1;
        for (int g = 0; g < E; ++g) {
            const int32_t* gi = m->edge_indices + 4 * g;
            int wg = m->particle_world ? m->particle_world[gi[2]] : 0;
            if (!world_ok(we, wg)) continue;
            V3 P2 = vld(pos, gi[2]), Q2 = vld(pos, gi[3]);
            if (!box_overlap(lo, hi, vmin3(P2, Q2), vmax3(P2, Q2))) continue;
            if (ei[2] == gi[2] || ei[2] == gi[3] || ei[3] == gi[2] || ei[3] == gi[3]) continue;
            if (sc->ee_filter && sorted_contains(sc->ee_filter, sc->ee_filter_offsets[e], sc->ee_filter_offsets[e + 1], g))
                continue;
            EEClosest st = closest_edge_edge(P, Q, P2, Q2, R(sc->edge_parallel_eps));
            if (use_rest) {
                EEClosest sr = closest_edge_edge(rest_pos(sc, ei[2]), rest_pos(sc, ei[3]), rest_pos(sc, gi[2]),
                                                 rest_pos(sc, gi[3]), R(sc->edge_parallel_eps));
                if (sr.d < rmin_q) continue;
            }
            if (st.d < r) {
                mind = rmin(mind, st.d);
                if (cnt < L->ee_cap) L->ee[(size_t)e * L->ee_cap + cnt] = g;
                ++cnt;
            }
        }
        L->ee_n[e] = cnt;
        L->ee_min[e] = (double)mind;
    }
}

static void self_publish(const OracleModel* m, const OracleSelf* sc, const SelfLists* L) {
    const int N = m->particle_count, T = m->tri_count, E = m->edge_count;
    if (sc->vt_list) memcpy(sc->vt_list, L->vt, sizeof(int32_t) * (size_t)N * L->vt_cap);
    if (sc->vt_count) memcpy(sc->vt_count, L->vt_n, sizeof(int32_t) * (size_t)N);
    if (sc->vt_min_dist) memcpy(sc->vt_min_dist, L->vt_min, sizeof(double) * (size_t)N);
    if (sc->tri_min_dist) memcpy(sc->tri_min_dist, L->tri_min, sizeof(double) * (size_t)T);
    if (sc->ee_list) memcpy(sc->ee_list, L->ee, sizeof(int32_t) * (size_t)E * L->ee_cap);
    if (sc->ee_count) memcpy(sc->ee_count, L->ee_n, sizeof(int32_t) * (size_t)E);
    if (sc->ee_min_dist) memcpy(sc->ee_min_dist, L->ee_min, sizeof(double) * (size_t)E);
}

/* accumulate_self_contact_force_and_hessian (K:1498-1641) for one color (color < 0: all):
 * per vertex, edge-edge terms in (e1, slot) order, then vertex-triangle terms in (p, slot) order,
 * continuing the contact/spring sums in pf / ph. */
static void accumulate_self_contacts(const OracleModel* m, const OracleParams* p, const SelfLists* L,
                                     const REAL* q, const REAL* q_prev, REAL dt, int color, REAL* pf, REAL* ph) {
    const OracleSelf* sc = p->self_contact;
    const int N = m->particle_count, E = m->edge_count;
    const REAL radius = R(sc->radius), ke = R(m->soft_contact_ke), kd = R(m->soft_contact_kd);
    const REAL mu = R(m->soft_contact_mu), feps = R(p->friction_epsilon), peps = R(sc->edge_parallel_eps);
#define IN_COLOR(i) (color < 0 || m->particle_colors[(i)] == color)
    for (int e1 = 0; e1 < E; ++e1) {
        const int32_t* a = m->edge_indices + 4 * e1;
        int n1 = L->ee_n[e1] < L->ee_cap ? L->ee_n[e1] : L->ee_cap;
        if (!(IN_COLOR(a[2]) || IN_COLOR(a[3]))) continue;
        for (int k = 0; k < n1; ++k) {
            int e2 = L->ee[(size_t)e1 * L->ee_cap + k];
            const int32_t* b = m->edge_indices + 4 * e2;
            V3 f[2];
            M3 h[2];
            if (!ee_contact_eval(vld(q, a[2]), vld(q, a[3]), vld(q, b[2]), vld(q, b[3]), vld(q_prev, a[2]),
                                 vld(q_prev, a[3]), vld(q_prev, b[2]), vld(q_prev, b[3]), radius, ke, kd, mu, feps,
                                 dt, peps, f, h))
                continue;
            for (int r = 0; r < 2; ++r) {
                int v = a[2 + r];
                if (!IN_COLOR(v)) continue;
                vst(pf, v, vadd(vld(pf, v), f[r]));
                mstore(ph + 9 * v, madd(mload(ph + 9 * v), h[r]));
            }
        }
    }
    for (int v = 0; v < N; ++v) {
        int nv = L->vt_n[v] < L->vt_cap ? L->vt_n[v] : L->vt_cap;
        for (int k = 0; k < nv; ++k) {
            int t = L->vt[(size_t)v * L->vt_cap + k];
            const int32_t* ti = m->tri_indices + 3 * t;
            if (!(IN_COLOR(v) || IN_COLOR(ti[0]) || IN_COLOR(ti[1]) || IN_COLOR(ti[2]))) continue;
            V3 f[4];
            M3 h[4];
            if (!vt_contact_eval(vld(q, ti[0]), vld(q, ti[1]), vld(q, ti[2]), vld(q, v), vld(q_prev, ti[0]),
                                 vld(q_prev, ti[1]), vld(q_prev, ti[2]), vld(q_prev, v), radius, ke, kd, mu, feps,
                                 dt, f, h))
                continue;
            const int who[4] = {ti[0], ti[1], ti[2], v};
            for (int r = 0; r < 4; ++r) {
                int u = who[r];
                if (!IN_COLOR(u)) continue;
                vst(pf, u, vadd(vld(pf, u), f[r]));
                mstore(ph + 9 * u, madd(mload(ph + 9 * u), h[r]));
            }
        }
    }
#undef IN_COLOR
}

/* ---- penetration-free truncation (K:1864-2287, S:2003-2063) --------------- */
/* segment_plane_intersects (K:1864-1886) */
static int segment_plane_hits(V3 v, V3 dv, V3 n, V3 d, REAL eps_par, REAL eps_near, REAL eps_far) {
    REAL nv = vdot(n, dv);
    REAL num = -vdot(n, vsub(v, d));
    if (RABS(nv) < eps_par) return 0; /* coplanar_counts = False */
    REAL t = num / nv;
    return t >= eps_near && t <= R(1.0) + eps_far;
}

/* wp.max over a vector's components: the first element, replaced by each strictly larger one */
static REAL rmax4(REAL a, REAL b, REAL c, REAL d) {
    REAL r = a;
    if (b > r) r = b;
    if (c > r) r = c;
    if (d > r) r = d;
    return r;
}
static REAL rmax3v(REAL a, REAL b, REAL c) {
    REAL r = a;
    if (b > r) r = b;
    if (c > r) r = c;
    return r;
}

/* create_vertex_triangle_division_plane_closest_pt (K:1889-1947); dummy[0] is the vertex */
static void vt_division_plane(V3 v, V3 dv, V3 t1, V3 dt1, V3 t2, V3 dt2, V3 t3, V3 dt3, int dummy[4], V3* n_out,
                              V3* d_out) {
    V3 bary;
    int feature;
    V3 cp = tri_closest(t1, t2, t3, v, &bary, &feature);
    V3 n_hat = vsub(v, cp);
    if (vlength(n_hat) < R(1e-12)) {
        dummy[0] = dummy[1] = dummy[2] = dummy[3] = 0;
        *n_out = v3(R(0.0), R(0.0), R(0.0));
        *d_out = v;
        return;
    }
    V3 n = vdivs(n_hat, vlength(n_hat));
    REAL dvn = rmax(-vdot(n, dv), R(0.0));
    REAL dtn = rmax4(vdot(n, dt1), vdot(n, dt2), vdot(n, dt3), R(0.0));
    V3 d;
    if (dtn + dvn == R(0.0)) {
        d = vadd(cp, vscale(n_hat, R(0.5)));
    } else {
        REAL l = rclamp(dtn / (dtn + dvn), R(0.05), R(0.95));
        d = vadd(cp, vscale(n_hat, l));
    }
    dummy[0] = dvn == R(0.0) ? 1 : !segment_plane_hits(v, dv, n, d, R(1e-6), R(-1e-8), R(1e-8));
    if (dtn == R(0.0)) {
        dummy[1] = dummy[2] = dummy[3] = 1;
    } else {
        dummy[1] = !segment_plane_hits(t1, dt1, n, d, R(1e-6), R(-1e-8), R(1e-8));
        dummy[2] = !segment_plane_hits(t2, dt2, n, d, R(1e-6), R(-1e-8), R(1e-8));
        dummy[3] = !segment_plane_hits(t3, dt3, n, d, R(1e-6), R(-1e-8), R(1e-8));
    }
    *n_out = n;
    *d_out = d;
}

/* robust_edge_pair_normal (K:1950-2016) */
static V3 robust_edge_pair_normal(V3 a0, V3 a1, V3 b0, V3 b1) {
    const REAL eps = R(1.0e-6);
    V3 dir0 = vsub(a1, a0), dir1 = vsub(b1, b0);
    REAL l0 = vlength(dir0), l1 = vlength(dir1);
    dir0 = l0 > eps ? vdivs(dir0, l0) : v3(R(0.0), R(0.0), R(0.0));
    dir1 = l1 > eps ? vdivs(dir1, l1) : v3(R(0.0), R(0.0), R(0.0));
    V3 n = vcross(dir0, dir1);
    REAL ln = vlength(n);
    if (ln > eps) return vdivs(n, ln);
    V3 ref = dir0;
    if (vlength(ref) <= eps) ref = dir1;
    if (vlength(ref) <= eps) return v3(R(1.0), R(0.0), R(0.0));
    V3 bridge = vscale(vsub(vadd(b0, b1), vadd(a0, a1)), R(0.5));
    REAL bl = vlength(bridge);
    if (bl > eps) {
        n = vcross(ref, vdivs(bridge, bl));
        ln = vlength(n);
        if (ln > eps) return vdivs(n, ln);
    }
    V3 axis = v3(R(1.0), R(0.0), R(0.0));
    if (RABS(vdot(ref, axis)) > R(0.9)) axis = v3(R(0.0), R(1.0), R(0.0));
    n = vcross(ref, axis);
    ln = vlength(n);
    if (ln > eps) return vdivs(n, ln);
    axis = v3(R(0.0), R(0.0), R(1.0));
    n = vcross(ref, axis);
    ln = vlength(n);
    if (ln > eps) return vdivs(n, ln);
    return v3(R(1.0), R(0.0), R(0.0));
}

/* create_edge_edge_division_plane_closest_pt (K:2019-2090) */
static void ee_division_plane(V3 a0, V3 da0, V3 a1, V3 da1, V3 b0, V3 db0, V3 b1, V3 db1, int dummy[4], V3* n_out,
                              V3* d_out) {
    EEClosest st = closest_edge_edge(a0, a1, b0, b1, R(1e-6));
    V3 c1 = vadd(a0, vscale(vsub(a1, a0), st.s));
    V3 c2 = vadd(b0, vscale(vsub(b1, b0), st.t));
    V3 n_hat = vsub(c1, c2);
    if (vlength(n_hat) < R(1e-12)) {
        dummy[0] = dummy[1] = dummy[2] = dummy[3] = 0;
        *n_out = robust_edge_pair_normal(a0, a1, b0, b1);
        *d_out = vadd(vscale(c1, R(0.5)), vscale(c2, R(0.5)));
        return;
    }
    V3 n = vdivs(n_hat, vlength(n_hat));
    REAL de0 = rmax3v(-vdot(n, da0), -vdot(n, da1), R(0.0));
    REAL de1 = rmax3v(vdot(n, db0), vdot(n, db1), R(0.0));
    V3 d;
    if (de0 + de1 == R(0.0)) {
        d = vadd(c2, vscale(n_hat, R(0.5)));
    } else {
        REAL l = rclamp(de1 / (de1 + de0), R(0.05), R(0.95));
        d = vadd(c2, vscale(n_hat, l));
    }
    if (de0 == R(0.0)) {
        dummy[0] = dummy[1] = 1;
    } else {
        dummy[0] = !segment_plane_hits(a0, da0, n, d, R(1e-6), R(-1e-8), R(1e-6));
        dummy[1] = !segment_plane_hits(a1, da1, n, d, R(1e-6), R(-1e-8), R(1e-6));
    }
    if (de1 == R(0.0)) {
        dummy[2] = dummy[3] = 1;
    } else {
        dummy[2] = !segment_plane_hits(b0, db0, n, d, R(1e-6), R(-1e-8), R(1e-6));
        dummy[3] = !segment_plane_hits(b1, db1, n, d, R(1e-6), R(-1e-8), R(1e-6));
    }
    *n_out = n;
    *d_out = d;
}

/* planar_truncation_t (K:2113-2130), gamma_min = 1e-3 */
static REAL planar_truncation_t(V3 v, V3 dv, V3 n, V3 d, REAL eps, REAL gamma) {
    REAL denom = vdot(n, dv);
    if (RABS(denom) < eps) return R(1.0);
    REAL t = vdot(n, vsub(d, v)) / denom;
    if (t < R(0.0)) return R(1.0);
    return rclamp(rmin(t * gamma, t - R(1e-3)), R(0.0), R(1.0));
}

/* truncation parameters of one collision: kind 0 edge-edge (x = e1.v0, e1.v1, e2.v0, e2.v1),
 * kind 1 vertex-triangle (x = p, a, b, c); t[k] = 1 for dummy roles */
static void collision_truncation_ts(int kind, const V3 x[4], const V3 dx[4], REAL peps, REAL gamma, REAL t[4]) {
    int dummy[4];
    V3 n, d;
    if (kind == 0)
        ee_division_plane(x[0], dx[0], x[1], dx[1], x[2], dx[2], x[3], dx[3], dummy, &n, &d);
    else
        vt_division_plane(x[0], dx[0], x[1], dx[1], x[2], dx[2], x[3], dx[3], dummy, &n, &d);
    for (int k = 0; k < 4; ++k) t[k] = dummy[k] ? R(1.0) : planar_truncation_t(x[k], dx[k], n, d, peps, gamma);
}

/* _penetration_free_truncation with self-contact on (S:2025-2063): per-vertex minimum of the
 * collision planes' truncation factors (order-free), then the isotropic clamp; q = pos_cd + d */
static void self_truncation(const OracleModel* m, const OracleSelf* sc, const SelfLists* L, const REAL* pos_cd,
                            REAL* disp, REAL* q, REAL* ts) {
    const int N = m->particle_count, E = m->edge_count;
    const REAL peps = R(sc->edge_parallel_eps), gamma = R(sc->relaxation);
    for (int i = 0; i < N; ++i) ts[i] = R(1.0);
    for (int e1 = 0; e1 < E; ++e1) {
        const int32_t* a = m->edge_indices + 4 * e1;
        int n1 = L->ee_n[e1] < L->ee_cap ? L->ee_n[e1] : L->ee_cap;
        for (int k = 0; k < n1; ++k) {
            const int32_t* b = m->edge_indices + 4 * L->ee[(size_t)e1 * L->ee_cap + k];
            const int who[4] = {a[2], a[3], b[2], b[3]};
            V3 x[4], dx[4];
            REAL t[4];
            for (int r = 0; r < 4; ++r) {
                x[r] = vld(pos_cd, who[r]);
                dx[r] = vld(disp, who[r]);
            }
            collision_truncation_ts(0, x, dx, peps, gamma, t);
            for (int r = 0; r < 4; ++r) ts[who[r]] = rmin(ts[who[r]], t[r]);
        }
    }
    for (int v = 0; v < N; ++v) {
        int nv = L->vt_n[v] < L->vt_cap ? L->vt_n[v] : L->vt_cap;
        for (int k = 0; k < nv; ++k) {
            const int32_t* ti = m->tri_indices + 3 * L->vt[(size_t)v * L->vt_cap + k];
            const int who[4] = {v, ti[0], ti[1], ti[2]};
            V3 x[4], dx[4];
            REAL t[4];
            for (int r = 0; r < 4; ++r) {
                x[r] = vld(pos_cd, who[r]);
                dx[r] = vld(disp, who[r]);
            }
            collision_truncation_ts(1, x, dx, peps, gamma, t);
            for (int r = 0; r < 4; ++r) ts[who[r]] = rmin(ts[who[r]], t[r]);
        }
    }
    const REAL maxd = R(sc->max_displacement);
    for (int i = 0; i < N; ++i) {
        V3 d = vscale(vld(disp, i), ts[i]);
        REAL len = vlength(d);
        if (len > maxd) d = vdivs(vscale(d, maxd), len);
        vst(disp, i, d);
        vst(q, i, vadd(vld(pos_cd, i), d));
    }
}

/* ---- standalone entry points (element-level checks) ---------------------- */
int32_t API(vbd_oracle_triangle_closest)(const REAL* x, REAL* closest, REAL* bary) {
    V3 b;
    int feature;
    V3 c = tri_closest(vld(x, 0), vld(x, 1), vld(x, 2), vld(x, 3), &b, &feature);
    vst(closest, 0, c);
    vst(bary, 0, b);
    return feature;
}

void API(vbd_oracle_closest_edge_edge)(const REAL* x, REAL eps, REAL* out) {
    EEClosest st = closest_edge_edge(vld(x, 0), vld(x, 1), vld(x, 2), vld(x, 3), eps);
    out[0] = st.s;
    out[1] = st.t;
    out[2] = st.d;
}

void API(vbd_oracle_self_force_norm)(REAL dis, REAL radius, REAL k, REAL* out) {
    self_force_norm(dis, radius, k, &out[0], &out[1]);
}

int32_t API(vbd_oracle_vt_contact)(const REAL* x, const REAL* anchor, REAL radius, REAL ke, REAL kd, REAL mu,
                                   REAL friction_epsilon, REAL dt, REAL* f, REAL* h) {
    V3 fv[4];
    M3 hm[4];
    int has = vt_contact_eval(vld(x, 0), vld(x, 1), vld(x, 2), vld(x, 3), vld(anchor, 0), vld(anchor, 1),
                              vld(anchor, 2), vld(anchor, 3), radius, ke, kd, mu, friction_epsilon, dt, fv, hm);
    for (int k = 0; k < 4; ++k) {
        vst(f, k, has ? fv[k] : v3(R(0.0), R(0.0), R(0.0)));
        mstore(h + 9 * k, has ? hm[k] : mzero());
    }
    return has;
}

int32_t API(vbd_oracle_ee_contact)(const REAL* x, const REAL* anchor, REAL radius, REAL ke, REAL kd, REAL mu,
                                   REAL friction_epsilon, REAL dt, REAL parallel_eps, REAL* f, REAL* h) {
    V3 fv[2];
    M3 hm[2];
    int has = ee_contact_eval(vld(x, 0), vld(x, 1), vld(x, 2), vld(x, 3), vld(anchor, 0), vld(anchor, 1),
                              vld(anchor, 2), vld(anchor, 3), radius, ke, kd, mu, friction_epsilon, dt, parallel_eps,
                              fv, hm);
    for (int k = 0; k < 2; ++k) {
        vst(f, k, has ? fv[k] : v3(R(0.0), R(0.0), R(0.0)));
        mstore(h + 9 * k, has ? hm[k] : mzero());
    }
    return has;
}

void API(vbd_oracle_truncation_ts)(int32_t kind, const REAL* x, const REAL* delta, REAL parallel_eps, REAL gamma,
                                   REAL* t) {
    V3 xv[4], dv[4];
    for (int k = 0; k < 4; ++k) {
        xv[k] = vld(x, k);
        dv[k] = vld(delta, k);
    }
    collision_truncation_ts(kind, xv, dv, parallel_eps, gamma, t);
}

int32_t API(vbd_oracle_self_detect)(const OracleModel* m, const OracleSelf* sc, const REAL* q) {
    if (!m || !sc) return -1;
    SelfLists L;
    self_lists_alloc(m, sc, &L);
    self_detect(m, sc, q, &L);
    self_publish(m, sc, &L);
    self_lists_free(&L);
    return 0;
}
