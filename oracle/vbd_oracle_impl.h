/*
 * CPU oracle body for the particle VBD step -- TEST INFRASTRUCTURE (the
 * checker, never the product).  Included twice, by vbd_oracle_f32.c and
 * vbd_oracle_f64.c, with REAL / SFX / ORACLE_IS_F32 defined.
 *
 * Every function restates the reference in its CPU (Warp `solve_elasticity`)
 * evaluation order.  Citations are to CSYSLAM/newton:
 *   K = newton/_src/solvers/vbd/particle_vbd_kernels.py
 *   R = newton/_src/solvers/vbd/rigid_vbd_kernels.py
 *   S = newton/_src/solvers/vbd/solver_vbd.py
 *   G = newton/_src/geometry/kernels.py
 *
 * Warp builtins the reference calls (warp-lang 1.16.0.dev20260716, absent
 * here) are restated from their textbook definitions; their exact FP32
 * operation order is "parity unpinned" (no reference output can be run):
 *   dot/cross/length/outer/skew, mat33 product and mat-vec (left-to-right
 *   sums), determinant = row0 . (row1 x row2), inverse = adjugate * (1/det),
 *   quat_rotate, atan2.
 * atan2: by default the FP32 instantiation uses the portable algorithm below
 * (Cephes atanf reduction, S. L. Moshier) so that the HIP kernels, which use
 * the same published method, can be compared bit for bit; use_libm_atan2=1
 * switches to the C library's atan2f, the function Warp's CPU backend calls.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#pragma STDC FP_CONTRACT OFF

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define API(name) CAT(name##_, SFX)

#ifdef ORACLE_IS_F32
#define R(x) ((float)(x))
#define RSQRT(x) sqrtf(x)
#define RABS(x) fabsf(x)
#else
#define R(x) ((double)(x))
#define RSQRT(x) sqrt(x)
#define RABS(x) fabs(x)
#endif

typedef struct { REAL x, y, z; } V3;
typedef struct { REAL m[3][3]; } M3;

static V3 v3(REAL x, REAL y, REAL z) { V3 r = {x, y, z}; return r; }
static V3 vld(const REAL* p, int i) { return v3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
static void vst(REAL* p, int i, V3 v) { p[3 * i] = v.x; p[3 * i + 1] = v.y; p[3 * i + 2] = v.z; }
static V3 vadd(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static V3 vsub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static V3 vneg(V3 a) { return v3(-a.x, -a.y, -a.z); }
/* scalar * vec and vec * scalar are the same per-component product */
static V3 vscale(V3 a, REAL s) { return v3(a.x * s, a.y * s, a.z * s); }
static V3 vdivs(V3 a, REAL s) { return v3(a.x / s, a.y / s, a.z / s); }
static REAL vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static V3 vcross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static REAL vlength(V3 a) { return RSQRT(vdot(a, a)); }
static REAL vcomp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

static M3 mzero(void) { M3 r; memset(&r, 0, sizeof(r)); return r; }
static M3 mident(void) { M3 r = mzero(); r.m[0][0] = R(1.0); r.m[1][1] = R(1.0); r.m[2][2] = R(1.0); return r; }
static M3 mouter(V3 a, V3 b) {
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = vcomp(a, i) * vcomp(b, j);
    return r;
}
static M3 madd(M3 a, M3 b) {
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] + b.m[i][j];
    return r;
}
static M3 msub(M3 a, M3 b) {
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] - b.m[i][j];
    return r;
}
static M3 mscale(M3 a, REAL s) {
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = s * a.m[i][j];
    return r;
}
static M3 mneg(M3 a) {
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = -a.m[i][j];
    return r;
}
static M3 mmul(M3 a, M3 b) {
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j];
    return r;
}
static M3 mtranspose(M3 a) {
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[j][i];
    return r;
}
static V3 mvec(M3 a, V3 v) {
    return v3(a.m[0][0] * v.x + a.m[0][1] * v.y + a.m[0][2] * v.z,
              a.m[1][0] * v.x + a.m[1][1] * v.y + a.m[1][2] * v.z,
              a.m[2][0] * v.x + a.m[2][1] * v.y + a.m[2][2] * v.z);
}
static M3 mskew(V3 v) {
    M3 r = mzero();
    r.m[0][1] = -v.z; r.m[0][2] = v.y;
    r.m[1][0] = v.z;  r.m[1][2] = -v.x;
    r.m[2][0] = -v.y; r.m[2][1] = v.x;
    return r;
}
static M3 mcols(V3 c0, V3 c1, V3 c2) {
    M3 r;
    r.m[0][0] = c0.x; r.m[0][1] = c1.x; r.m[0][2] = c2.x;
    r.m[1][0] = c0.y; r.m[1][1] = c1.y; r.m[1][2] = c2.y;
    r.m[2][0] = c0.z; r.m[2][1] = c1.z; r.m[2][2] = c2.z;
    return r;
}
/* determinant = row0 . (row1 x row2) */
static REAL mdet(M3 a) {
    V3 r0 = v3(a.m[0][0], a.m[0][1], a.m[0][2]);
    V3 r1 = v3(a.m[1][0], a.m[1][1], a.m[1][2]);
    V3 r2 = v3(a.m[2][0], a.m[2][1], a.m[2][2]);
    return vdot(r0, vcross(r1, r2));
}
/* inverse = adjugate * (1/det); zero matrix when det == 0 */
static M3 minverse(M3 a) {
    REAL det = mdet(a);
    M3 b = mzero();
    if (det == R(0.0)) return b;
    b.m[0][0] = a.m[1][1] * a.m[2][2] - a.m[1][2] * a.m[2][1];
    b.m[1][0] = a.m[1][2] * a.m[2][0] - a.m[1][0] * a.m[2][2];
    b.m[2][0] = a.m[1][0] * a.m[2][1] - a.m[1][1] * a.m[2][0];
    b.m[0][1] = a.m[0][2] * a.m[2][1] - a.m[0][1] * a.m[2][2];
    b.m[1][1] = a.m[0][0] * a.m[2][2] - a.m[0][2] * a.m[2][0];
    b.m[2][1] = a.m[0][1] * a.m[2][0] - a.m[0][0] * a.m[2][1];
    b.m[0][2] = a.m[0][1] * a.m[1][2] - a.m[0][2] * a.m[1][1];
    b.m[1][2] = a.m[0][2] * a.m[1][0] - a.m[0][0] * a.m[1][2];
    b.m[2][2] = a.m[0][0] * a.m[1][1] - a.m[0][1] * a.m[1][0];
    REAL inv_det = R(1.0) / det;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) b.m[i][j] = b.m[i][j] * inv_det;
    return b;
}
static REAL rmax(REAL a, REAL b) { return a > b ? a : b; }
static REAL rmin(REAL a, REAL b) { return a < b ? a : b; }
static REAL rsign(REAL a) { return a < R(0.0) ? R(-1.0) : R(1.0); }

/* ---- atan2 ------------------------------------------------------------- */
#ifdef ORACLE_IS_F32
/* Portable FP32 atan2: argument reduction to [0, 1] by |y|/|x| ordering,
 * then the Cephes atanf reduction (tan(pi/8) split) and its degree-4
 * polynomial in z = t^2.  Identical steps are used in the HIP kernels. */
static float portable_atan2f(float y, float x) {
    const float PI_F = 3.14159265358979323846f;
    const float PIO2_F = 1.57079632679489661923f;
    const float PIO4_F = 0.78539816339744830962f;
    float ax = fabsf(x), ay = fabsf(y);
    float r;
    if (ax == 0.0f && ay == 0.0f) {
        r = signbit(x) ? PI_F : 0.0f;
    } else {
        int swap = ay > ax;
        float t = swap ? (ax / ay) : (ay / ax);
        float base = 0.0f;
        if (t > 0.41421356237309504880f) {
            base = PIO4_F;
            t = (t - 1.0f) / (t + 1.0f);
        }
        float z = t * t;
        float poly = (((8.05374449538e-2f * z - 1.38776856032e-1f) * z + 1.99777106478e-1f) * z
                      - 3.33329491539e-1f) * z * t + t;
        r = base + poly;
        if (swap) r = PIO2_F - r;
        if (signbit(x)) r = PI_F - r;
    }
    return signbit(y) ? -r : r;
}
static int g_use_libm_atan2 = 0;
static REAL ratan2(REAL y, REAL x) { return g_use_libm_atan2 ? atan2f(y, x) : portable_atan2f(y, x); }
#else
static REAL ratan2(REAL y, REAL x) { return atan2(y, x); }
#endif

REAL API(vbd_oracle_atan2)(REAL y, REAL x) { return ratan2(y, x); }

/* ---- membrane: K:502-632 ------------------------------------------------ */
static void membrane_eval(V3 X0, V3 X1, V3 X2, V3 A0, V3 A1, V3 A2, const float* pose, REAL area,
                          REAL mu, REAL lmbd, REAL damping, REAL dt, int v_order, V3* f_out,
                          M3* h_out) {
    V3 x0 = X0;
    V3 x01 = vsub(X1, x0);
    V3 x02 = vsub(X2, x0);
    REAL D00 = R(pose[0]), D01 = R(pose[1]), D10 = R(pose[2]), D11 = R(pose[3]);
    V3 f0 = vadd(vscale(x01, D00), vscale(x02, D10));
    V3 f1 = vadd(vscale(x01, D01), vscale(x02, D11));
    REAL f0f0 = vdot(f0, f0), f1f1 = vdot(f1, f1), f0f1 = vdot(f0, f1);
    REAL Jsq = f0f0 * f1f1 - f0f1 * f0f1;
    Jsq = rmax(Jsq, R(1.0e-20));
    REAL Js = RSQRT(Jsq);
    REAL invJ = R(1.0) / Js;
    REAL mu_nh = mu;
    REAL l_nh = lmbd + mu;
    REAL l_safe = rsign(l_nh) * rmax(RABS(l_nh), R(1.0e-6));
    REAL alpha = R(1.0) + mu_nh / l_safe;
    V3 g0 = vscale(vsub(vscale(f0, f1f1), vscale(f1, f0f1)), invJ);
    V3 g1 = vscale(vsub(vscale(f1, f0f0), vscale(f0, f0f1)), invJ);
    REAL s = l_nh * (Js - alpha);
    V3 P0 = vadd(vscale(f0, mu_nh), vscale(g0, s));
    V3 P1 = vadd(vscale(f1, mu_nh), vscale(g1, s));
    REAL m0 = (REAL)(v_order == 0), m1 = (REAL)(v_order == 1), m2 = (REAL)(v_order == 2);
    REAL df0 = D00 * (m1 - m0) + D10 * (m2 - m0);
    REAL df1 = D01 * (m1 - m0) + D11 * (m2 - m0);
    V3 dpsi = vadd(vscale(P0, df0), vscale(P1, df1));
    V3 force = vneg(dpsi);
    REAL s_clamp = rmax(R(0.0), s);
    REAL r = s_clamp * invJ;
    REAL c1 = l_nh - r;
    REAL df0sq = df0 * df0, df1sq = df1 * df1;
    V3 dJ = vadd(vscale(g0, df0), vscale(g1, df1));
    V3 w = vsub(vscale(f1, df0), vscale(f0, df1));
    REAL I_coeff = mu_nh * (df0sq + df1sq) + r * (df0sq * f1f1 + df1sq * f0f0 - R(2.0) * df0 * df1 * f0f1);
    M3 hess = msub(madd(mscale(mident(), I_coeff), mscale(mouter(dJ, dJ), c1)), mscale(mouter(w, w), r));
    if (damping > R(0.0)) {
        REAL inv_dt = R(1.0) / dt;
        V3 x0p = A0;
        V3 x01p = vsub(A1, x0p);
        V3 x02p = vsub(A2, x0p);
        V3 f0p = vadd(vscale(x01p, D00), vscale(x02p, D10));
        V3 f1p = vadd(vscale(x01p, D01), vscale(x02p, D11));
        REAL c00 = (f0f0 - vdot(f0p, f0p)) * inv_dt;
        REAL c01 = (f0f1 - vdot(f0p, f1p)) * inv_dt;
        REAL c11 = (f1f1 - vdot(f1p, f1p)) * inv_dt;
        V3 dc00 = vscale(f0, R(2.0) * df0);
        V3 dc01 = vadd(vscale(f1, df0), vscale(f0, df1));
        V3 dc11 = vscale(f1, R(2.0) * df1);
        V3 f_damp = vadd(vadd(vscale(dc00, c00), vscale(dc01, R(2.0) * c01)), vscale(dc11, c11));
        force = vadd(force, vscale(f_damp, -damping));
        M3 sum = madd(madd(mouter(dc00, dc00), mscale(mouter(dc01, dc01), R(2.0))), mouter(dc11, dc11));
        hess = madd(hess, mscale(sum, damping * inv_dt));
    }
    *f_out = vscale(force, area);
    *h_out = mscale(hess, area);
}

/* ---- bending: K:635-816 -------------------------------------------------- */
static M3 normalized_vector_derivative(REAL len, V3 nhat, M3 dvec) {
    M3 proj = msub(mident(), mouter(nhat, nhat));
    return mmul(mscale(proj, R(1.0) / len), dvec);
}
static V3 angle_derivative(V3 n1h, V3 n2h, V3 eh, M3 dn1, M3 dn2, REAL sin_t, REAL cos_t, M3 skew_n1,
                           M3 skew_n2) {
    V3 dsin = mvec(mtranspose(msub(mmul(skew_n1, dn2), mmul(skew_n2, dn1))), eh);
    V3 dcos = vadd(mvec(mtranspose(dn1), n2h), mvec(mtranspose(dn2), n1h));
    return vsub(vscale(dsin, cos_t), vscale(dcos, sin_t));
}
static void bending_eval(V3 x0, V3 x1, V3 x2, V3 x3, V3 p0, V3 p1, V3 p2, V3 p3, int has_o0, int has_o1,
                         REAL rest_angle, REAL rest_length, REAL stiffness, REAL damping, REAL dt,
                         int v_order, V3* f_out, M3* h_out) {
    *f_out = v3(R(0.0), R(0.0), R(0.0));
    *h_out = mzero();
    if (!has_o0 || !has_o1) return;
    const REAL eps = R(1.0e-6);
    V3 x02 = vsub(x2, x0), x03 = vsub(x3, x0), x13 = vsub(x3, x1), x12 = vsub(x2, x1), e = vsub(x3, x2);
    V3 n1 = vcross(x02, x03), n2 = vcross(x13, x12);
    REAL n1n = vlength(n1), n2n = vlength(n2), en = vlength(e);
    if (n1n < eps || n2n < eps || en < eps) return;
    V3 n1h = vdivs(n1, n1n), n2h = vdivs(n2, n2n), eh = vdivs(e, en);
    REAL sin_t = vdot(vcross(n1h, n2h), eh);
    REAL cos_t = vdot(n1h, n2h);
    REAL theta = ratan2(sin_t, cos_t);
    REAL k = stiffness * rest_length;
    REAL dE = k * (theta - rest_angle);
    M3 skew_e = mskew(e), skew_x03 = mskew(x03), skew_x02 = mskew(x02), skew_x13 = mskew(x13),
       skew_x12 = mskew(x12), skew_n1 = mskew(n1h), skew_n2 = mskew(n2h);
    M3 dn1_0 = normalized_vector_derivative(n1n, n1h, skew_e);
    M3 dn2_0 = mzero();
    M3 dn1_1 = mzero();
    M3 dn2_1 = normalized_vector_derivative(n2n, n2h, mneg(skew_e));
    M3 dn1_2 = normalized_vector_derivative(n1n, n1h, mneg(skew_x03));
    M3 dn2_2 = normalized_vector_derivative(n2n, n2h, skew_x13);
    M3 dn1_3 = normalized_vector_derivative(n1n, n1h, skew_x02);
    M3 dn2_3 = normalized_vector_derivative(n2n, n2h, mneg(skew_x12));
    V3 d0 = angle_derivative(n1h, n2h, eh, dn1_0, dn2_0, sin_t, cos_t, skew_n1, skew_n2);
    V3 d1 = angle_derivative(n1h, n2h, eh, dn1_1, dn2_1, sin_t, cos_t, skew_n1, skew_n2);
    V3 d2 = angle_derivative(n1h, n2h, eh, dn1_2, dn2_2, sin_t, cos_t, skew_n1, skew_n2);
    V3 d3 = angle_derivative(n1h, n2h, eh, dn1_3, dn2_3, sin_t, cos_t, skew_n1, skew_n2);
    REAL k0 = (REAL)(v_order == 0), k1 = (REAL)(v_order == 1), k2 = (REAL)(v_order == 2),
         k3 = (REAL)(v_order == 3);
    V3 dth = vadd(vadd(vadd(vscale(d0, k0), vscale(d1, k1)), vscale(d2, k2)), vscale(d3, k3));
    V3 bf = vscale(dth, -dE);
    M3 bh = mscale(mouter(dth, dth), k);
    if (damping > R(0.0)) {
        REAL inv_dt = R(1.0) / dt;
        V3 x02p = vsub(p2, p0), x03p = vsub(p3, p0), x13p = vsub(p3, p1), x12p = vsub(p2, p1),
           ep = vsub(p3, p2);
        V3 n1pr = vcross(x02p, x03p), n2pr = vcross(x13p, x12p);
        REAL n1pn = vlength(n1pr), n2pn = vlength(n2pr), epn = vlength(ep);
        if (n1pn < eps || n2pn < eps || epn < eps) {
            *f_out = bf;
            *h_out = bh;
            return;
        }
        V3 n1p = vdivs(n1pr, n1pn), n2p = vdivs(n2pr, n2pn), ehp = vdivs(ep, epn);
        REAL sinp = vdot(vcross(n1p, n2p), ehp);
        REAL cosp = vdot(n1p, n2p);
        REAL thetap = ratan2(sinp, cosp);
        REAL dtheta = theta - thetap;
        if (dtheta > R(3.141592653589793))
            dtheta = dtheta - R(6.283185307179586);
        else if (dtheta < R(-3.141592653589793))
            dtheta = dtheta + R(6.283185307179586);
        REAL dtheta_dt = dtheta * inv_dt;
        V3 dfrc = vscale(dth, -damping * rest_length * dtheta_dt);
        M3 dhes = mscale(mouter(dth, dth), damping * rest_length * inv_dt);
        bf = vadd(bf, dfrc);
        bh = madd(bh, dhes);
    }
    *f_out = bf;
    *h_out = bh;
}

/* ---- volumetric Neo-Hookean tet: K:76-145, K:148-292, K:347-367 ---------- */
static M3 cofactor(M3 F) {
    REAL F11 = F.m[0][0], F21 = F.m[1][0], F31 = F.m[2][0];
    REAL F12 = F.m[0][1], F22 = F.m[1][1], F32 = F.m[2][1];
    REAL F13 = F.m[0][2], F23 = F.m[1][2], F33 = F.m[2][2];
    M3 c;
    c.m[0][0] = F22 * F33 - F23 * F32;
    c.m[0][1] = F23 * F31 - F21 * F33;
    c.m[0][2] = F21 * F32 - F22 * F31;
    c.m[1][0] = F13 * F32 - F12 * F33;
    c.m[1][1] = F11 * F33 - F13 * F31;
    c.m[1][2] = F12 * F31 - F11 * F32;
    c.m[2][0] = F12 * F23 - F13 * F22;
    c.m[2][1] = F13 * F21 - F11 * F23;
    c.m[2][2] = F11 * F22 - F12 * F21;
    return c;
}
static void flatten_colmajor(M3 a, REAL* v9) {
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) v9[3 * c + r] = a.m[r][c];
}
static void tet_eval(const V3* X, const V3* P, const float* dm_inv, REAL mu, REAL lmbd, REAL damping, REAL dt,
                     int v_order, V3* f_out, M3* h_out) {
    M3 Dm;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Dm.m[i][j] = R(dm_inv[3 * i + j]);
    REAL rest_volume = R(1.0) / (mdet(Dm) * R(6.0));
    M3 Ds = mcols(vsub(X[1], X[0]), vsub(X[2], X[0]), vsub(X[3], X[0]));
    M3 F = mmul(Ds, Dm);
    REAL f9[9];
    flatten_colmajor(F, f9);
    REAL J = mdet(F);
    REAL mu_nh = mu;
    REAL l_nh = lmbd + mu;
    REAL l_safe = rsign(l_nh) * rmax(RABS(l_nh), R(1.0e-6));
    REAL alpha = R(1.0) + mu_nh / l_safe;
    REAL c9[9];
    flatten_colmajor(cofactor(F), c9);
    REAL s = l_nh * (J - alpha);
    REAL Pv[9];
    for (int i = 0; i < 9; ++i) Pv[i] = rest_volume * (mu_nh * f9[i] + s * c9[i]);
    REAL H[9][9];
    for (int i = 0; i < 9; ++i)
        for (int j = 0; j < 9; ++j)
            H[i][j] = rest_volume * (mu_nh * (i == j ? R(1.0) : R(0.0)) + l_nh * (c9[i] * c9[j]));
    V3 mv;
    if (v_order == 0)
        mv = v3(-(Dm.m[0][0] + Dm.m[1][0] + Dm.m[2][0]), -(Dm.m[0][1] + Dm.m[1][1] + Dm.m[2][1]),
                -(Dm.m[0][2] + Dm.m[1][2] + Dm.m[2][2]));
    else if (v_order == 1)
        mv = v3(Dm.m[0][0], Dm.m[0][1], Dm.m[0][2]);
    else if (v_order == 2)
        mv = v3(Dm.m[1][0], Dm.m[1][1], Dm.m[1][2]);
    else
        mv = v3(Dm.m[2][0], Dm.m[2][1], Dm.m[2][2]);
    REAL m1 = mv.x, m2 = mv.y, m3 = mv.z;
    V3 force = v3(-(Pv[0] * m1 + Pv[3] * m2 + Pv[6] * m3), -(Pv[1] * m1 + Pv[4] * m2 + Pv[7] * m3),
                  -(Pv[2] * m1 + Pv[5] * m2 + Pv[8] * m3));
    M3 hess;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            hess.m[r][c] = m1 * (H[r][c] * m1 + H[r + 3][c] * m2 + H[r + 6][c] * m3) +
                           m2 * (H[r][c + 3] * m1 + H[r + 3][c + 3] * m2 + H[r + 6][c + 3] * m3) +
                           m3 * (H[r][c + 6] * m1 + H[r + 3][c + 6] * m2 + H[r + 6][c + 6] * m3);
    if (damping > R(0.0)) {
        REAL inv_dt = R(1.0) / dt;
        M3 Dsp = mcols(vsub(P[1], P[0]), vsub(P[2], P[0]), vsub(P[3], P[0]));
        M3 Fp = mmul(Dsp, Dm);
        V3 f0 = v3(F.m[0][0], F.m[1][0], F.m[2][0]);
        V3 f1 = v3(F.m[0][1], F.m[1][1], F.m[2][1]);
        V3 f2 = v3(F.m[0][2], F.m[1][2], F.m[2][2]);
        V3 f0p = v3(Fp.m[0][0], Fp.m[1][0], Fp.m[2][0]);
        V3 f1p = v3(Fp.m[0][1], Fp.m[1][1], Fp.m[2][1]);
        V3 f2p = v3(Fp.m[0][2], Fp.m[1][2], Fp.m[2][2]);
        REAL c00 = (vdot(f0, f0) - vdot(f0p, f0p)) * inv_dt;
        REAL c01 = (vdot(f0, f1) - vdot(f0p, f1p)) * inv_dt;
        REAL c02 = (vdot(f0, f2) - vdot(f0p, f2p)) * inv_dt;
        REAL c11 = (vdot(f1, f1) - vdot(f1p, f1p)) * inv_dt;
        REAL c12 = (vdot(f1, f2) - vdot(f1p, f2p)) * inv_dt;
        REAL c22 = (vdot(f2, f2) - vdot(f2p, f2p)) * inv_dt;
        V3 dc00 = vscale(f0, R(2.0) * m1);
        V3 dc01 = vadd(vscale(f1, m1), vscale(f0, m2));
        V3 dc02 = vadd(vscale(f2, m1), vscale(f0, m3));
        V3 dc11 = vscale(f1, R(2.0) * m2);
        V3 dc12 = vadd(vscale(f2, m2), vscale(f1, m3));
        V3 dc22 = vscale(f2, R(2.0) * m3);
        V3 fd = vscale(dc00, c00);
        fd = vadd(fd, vscale(dc01, R(2.0) * c01));
        fd = vadd(fd, vscale(dc02, R(2.0) * c02));
        fd = vadd(fd, vscale(dc11, c11));
        fd = vadd(fd, vscale(dc12, R(2.0) * c12));
        fd = vadd(fd, vscale(dc22, c22));
        force = vsub(force, vscale(fd, rest_volume * damping));
        M3 hs = mouter(dc00, dc00);
        hs = madd(hs, mscale(mouter(dc01, dc01), R(2.0)));
        hs = madd(hs, mscale(mouter(dc02, dc02), R(2.0)));
        hs = madd(hs, mouter(dc11, dc11));
        hs = madd(hs, mscale(mouter(dc12, dc12), R(2.0)));
        hs = madd(hs, mouter(dc22, dc22));
        hess = madd(hess, mscale(hs, rest_volume * damping * inv_dt));
    }
    *f_out = force;
    *h_out = hess;
}

/* ---- rigid transforms (p, q xyzw): warp transform_point / transform_vector / transform_inverse /
 * transform_multiply and the Hamilton quaternion product, restated from their textbook definitions */
typedef struct { V3 p; REAL qx, qy, qz, qw; } Xf;

static V3 quat_rotate(REAL qx, REAL qy, REAL qz, REAL qw, V3 x) {
    REAL c = R(2.0) * qw * qw - R(1.0);
    REAL d = R(2.0) * (qx * x.x + qy * x.y + qz * x.z);
    return v3(x.x * c + qx * d + (qy * x.z - qz * x.y) * qw * R(2.0),
              x.y * c + qy * d + (qz * x.x - qx * x.z) * qw * R(2.0),
              x.z * c + qz * d + (qx * x.y - qy * x.x) * qw * R(2.0));
}
static V3 xf_point(Xf t, V3 x) { return vadd(quat_rotate(t.qx, t.qy, t.qz, t.qw, x), t.p); }
static V3 xf_vector(Xf t, V3 x) { return quat_rotate(t.qx, t.qy, t.qz, t.qw, x); }
static Xf xf_inverse(Xf t) {
    Xf r;
    r.qx = -t.qx; r.qy = -t.qy; r.qz = -t.qz; r.qw = t.qw;
    r.p = vneg(quat_rotate(r.qx, r.qy, r.qz, r.qw, t.p));
    return r;
}
static Xf xf_load(const float* t) {
    Xf r;
    r.p = v3(R(t[0]), R(t[1]), R(t[2]));
    r.qx = R(t[3]); r.qy = R(t[4]); r.qz = R(t[5]); r.qw = R(t[6]);
    return r;
}
/* transform_multiply(a, b) = (quat_rotate(a.q, b.p) + a.p, a.q * b.q) */
static Xf xf_mul(Xf a, Xf b) {
    Xf r;
    r.p = vadd(quat_rotate(a.qx, a.qy, a.qz, a.qw, b.p), a.p);
    r.qx = a.qw * b.qx + b.qw * a.qx + a.qy * b.qz - b.qy * a.qz;
    r.qy = a.qw * b.qy + b.qw * a.qy + a.qz * b.qx - b.qz * a.qx;
    r.qz = a.qw * b.qz + b.qw * a.qz + a.qx * b.qy - b.qx * a.qy;
    r.qw = a.qw * b.qw - a.qx * b.qx - a.qy * b.qy - a.qz * b.qz;
    return r;
}

/* World contact point bx = X_wb * body_pos of a soft contact and the body's velocity there
 * (R:1365-1439 with body_q_prev, the one-way coupling to an external rigid solver, solver_vbd.py:2523-2526):
 * bv = (bx - X_wb_prev * body_pos) / dt + X_wb . body_vel.  A static shape (body < 0) takes the identity
 * transform: bx = body_pos, bv = (0 + cross(0, bx)) + body_vel (the same values up to signed zeros).
 * Returns 0, or -4 when the shape has a body but no body poses were given. */
static int contact_body_frame(const OracleModel* m, const OracleParams* p, int shape, V3 body_pos, V3 body_vel,
                              REAL dt, V3* bx, V3* bv) {
    int b = (shape >= 0 && shape < m->shape_count && m->shape_body) ? m->shape_body[shape] : -1;
    V3 zero = v3(R(0.0), R(0.0), R(0.0));
    if (b < 0) {
        *bx = body_pos;
        *bv = vadd(vadd(zero, vcross(zero, *bx)), body_vel);
        return 0;
    }
    if (!p->body_q || !p->body_q_prev || b >= p->body_count) return -4;
    Xf X = xf_load(p->body_q + 7 * (size_t)b);
    Xf Xp = xf_load(p->body_q_prev + 7 * (size_t)b);
    *bx = xf_point(X, body_pos);
    V3 bxp = xf_point(Xp, body_pos);
    *bv = vadd(vdivs(vsub(*bx, bxp), dt), xf_vector(X, body_vel));
    return 0;
}

/* ---- soft contact force law: R:1331-1363 (penalty + damping + projected isotropic friction), at the
 * contact point x (anchor x_prev) against the world body point bx moving at bv ---- */
static void contact_force_law(V3 x, V3 x_prev, V3 bx, V3 bv, V3 n, REAL radius, REAL margin, REAL ke, REAL kd,
                              REAL mu, REAL friction_epsilon, REAL dt, V3* f_out, M3* h_out) {
    REAL pen = -(vdot(n, vsub(x, bx)) - radius - margin);
    if (!(pen > R(0.0))) {
        *f_out = v3(R(0.0), R(0.0), R(0.0));
        *h_out = mzero();
        return;
    }
    V3 dx = vsub(x, x_prev);
    V3 zero = v3(R(0.0), R(0.0), R(0.0));
    V3 u = vsub(dx, vscale(bv, dt));
    REAL f_n = pen * ke;
    V3 force = vscale(n, f_n);
    M3 hess = mscale(mouter(n, n), ke);
    if (vdot(n, u) < R(0.0)) {
        M3 dh = mscale(mouter(n, n), kd / dt);
        hess = madd(hess, dh);
        force = vsub(force, mvec(dh, u));
    }
    REAL eps_u = friction_epsilon * dt;
    REAL dot_nu = vdot(n, u);
    V3 u_t = vsub(u, vscale(n, dot_nu));
    REAL u_norm = vlength(u_t);
    V3 ff = zero;
    M3 fh = mzero();
    if (u_norm > R(0.0)) {
        REAL f1;
        if (u_norm > eps_u)
            f1 = R(1.0) / u_norm;
        else
            f1 = (-u_norm / eps_u + R(2.0)) / eps_u;
        REAL scale = mu * f_n * f1;
        ff = vneg(vscale(u_t, scale));
        fh = mscale(msub(mident(), mouter(n, n)), scale);
    }
    *f_out = vadd(force, ff);
    *h_out = madd(hess, fh);
}

/* particle vs static shape (R:1365-1439 with the identity body transform) */
static void body_contact_eval(V3 x, V3 x_prev, V3 body_pos, V3 body_vel, V3 n, REAL radius, REAL margin,
                              REAL ke, REAL kd, REAL mu, REAL friction_epsilon, REAL dt, V3* f_out,
                              M3* h_out) {
    V3 zero = v3(R(0.0), R(0.0), R(0.0));
    V3 bx = body_pos;
    V3 bv = vadd(vadd(zero, vcross(zero, bx)), body_vel);
    contact_force_law(x, x_prev, bx, bv, n, radius, margin, ke, kd, mu, friction_epsilon, dt, f_out, h_out);
}

/* Guarded 3x3 solve (K:2701-2706): no update when |det H| <= 1e-8.  With fp64_ill, the Frobenius
 * condition number ||H||_F ||H^-1||_F = ||H||_F ||adj H||_F / |det H| is evaluated in double from
 * the stored entries (a single-precision estimate is unreliable exactly for ill-conditioned
 * blocks); above 1e6 the block is solved in double, otherwise in the reference's single-precision
 * adjugate order. */
static int block_solve(M3 h, V3 f, int fp64_ill, V3* dx) {
    REAL det = mdet(h);
    if (!(RABS(det) > R(1e-8))) return 0;
    if (fp64_ill) {
        double a[3][3], b[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) a[i][j] = (double)h.m[i][j];
        b[0][0] = a[1][1] * a[2][2] - a[1][2] * a[2][1];
        b[1][0] = a[1][2] * a[2][0] - a[1][0] * a[2][2];
        b[2][0] = a[1][0] * a[2][1] - a[1][1] * a[2][0];
        b[0][1] = a[0][2] * a[2][1] - a[0][1] * a[2][2];
        b[1][1] = a[0][0] * a[2][2] - a[0][2] * a[2][0];
        b[2][1] = a[0][1] * a[2][0] - a[0][0] * a[2][1];
        b[0][2] = a[0][1] * a[1][2] - a[0][2] * a[1][1];
        b[1][2] = a[0][2] * a[1][0] - a[0][0] * a[1][2];
        b[2][2] = a[0][0] * a[1][1] - a[0][1] * a[1][0];
        double det_d = a[0][0] * b[0][0] + a[0][1] * b[1][0] + a[0][2] * b[2][0];
        double n1 = 0.0, n2 = 0.0;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                n1 += a[i][j] * a[i][j];
                n2 += b[i][j] * b[i][j];
            }
        /* cond > 1e6  <=>  sqrt(n1 n2) > 1e6 |det|  (also taken when det_d is 0 or not finite) */
        if (!(sqrt(n1 * n2) <= 1.0e6 * fabs(det_d))) {
            double fd[3] = {(double)f.x, (double)f.y, (double)f.z};
            double inv = 1.0 / det_d;
            double r[3];
            for (int i = 0; i < 3; ++i)
                r[i] = (b[i][0] * inv) * fd[0] + (b[i][1] * inv) * fd[1] + (b[i][2] * inv) * fd[2];
            *dx = v3((REAL)r[0], (REAL)r[1], (REAL)r[2]);
            return 1;
        }
    }
    *dx = mvec(minverse(h), f);
    return 1;
}

/* ---- exported element functions (for unit tests) ------------------------ */
static void mstore(REAL* h, M3 a) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) h[3 * i + j] = a.m[i][j];
}
static M3 mload(const REAL* h) {
    M3 a;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) a.m[i][j] = h[3 * i + j];
    return a;
}

void API(vbd_oracle_membrane)(const REAL* x, const REAL* xa, const float* pose, float area, float mu,
                              float lmbd, float kd, REAL dt, int32_t v_order, REAL* f, REAL* h) {
    V3 fo;
    M3 ho;
    membrane_eval(vld(x, 0), vld(x, 1), vld(x, 2), vld(xa, 0), vld(xa, 1), vld(xa, 2), pose, R(area), R(mu),
                  R(lmbd), R(kd), dt, v_order, &fo, &ho);
    vst(f, 0, fo);
    mstore(h, ho);
}

void API(vbd_oracle_bending)(const REAL* x, const REAL* xa, float rest_angle, float rest_length, float ke,
                             float kd, REAL dt, int32_t v_order, int32_t use_libm_atan2, REAL* f, REAL* h) {
#ifdef ORACLE_IS_F32
    g_use_libm_atan2 = use_libm_atan2;
#else
    (void)use_libm_atan2;
#endif
    V3 fo;
    M3 ho;
    bending_eval(vld(x, 0), vld(x, 1), vld(x, 2), vld(x, 3), vld(xa, 0), vld(xa, 1), vld(xa, 2), vld(xa, 3), 1,
                 1, R(rest_angle), R(rest_length), R(ke), R(kd), dt, v_order, &fo, &ho);
    vst(f, 0, fo);
    mstore(h, ho);
}

void API(vbd_oracle_tet)(const REAL* x, const REAL* xp, const float* dm_inv, float mu, float lmbd, float kd,
                         REAL dt, int32_t v_order, REAL* f, REAL* h) {
    V3 X[4], P[4];
    for (int i = 0; i < 4; ++i) {
        X[i] = vld(x, i);
        P[i] = vld(xp, i);
    }
    V3 fo;
    M3 ho;
    tet_eval(X, P, dm_inv, R(mu), R(lmbd), R(kd), dt, v_order, &fo, &ho);
    vst(f, 0, fo);
    mstore(h, ho);
}

void API(vbd_oracle_body_contact)(const REAL* x, const REAL* xp, const REAL* body_pos, const REAL* body_vel,
                                  const REAL* normal, REAL radius, REAL margin, REAL ke, REAL kd, REAL mu,
                                  REAL friction_epsilon, REAL dt, REAL* f, REAL* h) {
    V3 fo;
    M3 ho;
    body_contact_eval(vld(x, 0), vld(xp, 0), vld(body_pos, 0), vld(body_vel, 0), vld(normal, 0), radius, margin,
                      ke, kd, mu, friction_epsilon, dt, &fo, &ho);
    vst(f, 0, fo);
    mstore(h, ho);
}

int32_t API(vbd_oracle_block_solve)(const REAL* h, const REAL* f, int32_t fp64_ill, REAL* dx) {
    V3 d = v3(R(0.0), R(0.0), R(0.0));
    int ok = block_solve(mload(h), vld(f, 0), fp64_ill, &d);
    vst(dx, 0, d);
    return ok;
}

/* ---- per-vertex force/Hessian assembly: K:2568-2702 (CPU order) --------- */
typedef struct {
    const OracleModel* m;
    const REAL* q;
    const REAL* q_prev;
    REAL dt;
} Ctx;

/* Returns 0 when the vertex is inactive (skipped by the solve). */
static int assemble_vertex(const Ctx* c, int i, const REAL* inertia, V3* f_out, M3* h_out) {
    const OracleModel* m = c->m;
    if (!(m->particle_flags[i] & 1) || m->particle_mass[i] == 0.0f) return 0;
    REAL dt = c->dt;
    REAL dtr = R(1.0) / (dt * dt);
    REAL mass = R(m->particle_mass[i]);
    V3 f = vscale(vsub(vld(inertia, i), vld(c->q, i)), mass);
    f = vscale(f, dtr);
    M3 h = mscale(mident(), mass * dtr);
    if (m->tri_count > 0) {
        int b = m->v_adj_tris_offsets[i], e = m->v_adj_tris_offsets[i + 1];
        int n = (e - b) >> 1;
        for (int k = 0; k < n; ++k) {
            int t = m->v_adj_tris[b + 2 * k], slot = m->v_adj_tris[b + 2 * k + 1];
            const float* mat = m->tri_materials + 5 * t;
            if (mat[0] > 0.0f || mat[1] > 0.0f) {
                const int32_t* ti = m->tri_indices + 3 * t;
                V3 ft;
                M3 ht;
                membrane_eval(vld(c->q, ti[0]), vld(c->q, ti[1]), vld(c->q, ti[2]), vld(c->q_prev, ti[0]),
                              vld(c->q_prev, ti[1]), vld(c->q_prev, ti[2]), m->tri_poses + 4 * t,
                              R(m->tri_areas[t]), R(mat[0]), R(mat[1]), R(mat[2]), dt, slot, &ft, &ht);
                f = vadd(f, ft);
                h = madd(h, ht);
            }
        }
    }
    if (m->edge_count > 0) {
        int b = m->v_adj_edges_offsets[i], e = m->v_adj_edges_offsets[i + 1];
        int n = (e - b) >> 1;
        for (int k = 0; k < n; ++k) {
            int ed = m->v_adj_edges[b + 2 * k], slot = m->v_adj_edges[b + 2 * k + 1];
            const float* bp = m->edge_bending_properties + 2 * ed;
            if (bp[0] > 0.0f) {
                const int32_t* ei = m->edge_indices + 4 * ed;
                int h0 = ei[0] != -1, h1 = ei[1] != -1;
                V3 z = v3(R(0.0), R(0.0), R(0.0));
                V3 x0 = h0 ? vld(c->q, ei[0]) : z, x1 = h1 ? vld(c->q, ei[1]) : z;
                V3 p0 = h0 ? vld(c->q_prev, ei[0]) : z, p1 = h1 ? vld(c->q_prev, ei[1]) : z;
                V3 fe;
                M3 he;
                bending_eval(x0, x1, vld(c->q, ei[2]), vld(c->q, ei[3]), p0, p1, vld(c->q_prev, ei[2]),
                             vld(c->q_prev, ei[3]), h0, h1, R(m->edge_rest_angle[ed]),
                             R(m->edge_rest_length[ed]), R(bp[0]), R(bp[1]), dt, slot, &fe, &he);
                f = vadd(f, fe);
                h = madd(h, he);
            }
        }
    }
    if (m->tet_count > 0) {
        int b = m->v_adj_tets_offsets[i], e = m->v_adj_tets_offsets[i + 1];
        int n = (e - b) >> 1;
        for (int k = 0; k < n; ++k) {
            int t = m->v_adj_tets[b + 2 * k], slot = m->v_adj_tets[b + 2 * k + 1];
            const float* mat = m->tet_materials + 3 * t;
            if (mat[0] > 0.0f || mat[1] > 0.0f) {
                const int32_t* ti = m->tet_indices + 4 * t;
                V3 X[4], P[4];
                for (int a = 0; a < 4; ++a) {
                    X[a] = vld(c->q, ti[a]);
                    P[a] = vld(c->q_prev, ti[a]);
                }
                V3 fk;
                M3 hk;
                tet_eval(X, P, m->tet_poses + 9 * t, R(mat[0]), R(mat[1]), R(mat[2]), dt, slot, &fk, &hk);
                f = vadd(f, fk);
                h = madd(h, hk);
            }
        }
    }
    *f_out = f;
    *h_out = h;
    return 1;
}

/* contact material averaging + penalty seed: R:799-815, R:3276-3321 */
static void init_contacts(const OracleModel* m, const OracleParams* p, int ncontact, const int32_t* c_shape,
                          REAL* pk, REAL* mke, REAL* mkd, REAL* mmu) {
    for (int i = 0; i < ncontact; ++i) {
        int s = c_shape[i];
        REAL ke = R(0.5) * (R(m->soft_contact_ke) + R(m->shape_material_ke[s]));
        REAL kd = R(0.5) * (R(m->soft_contact_kd) + R(m->shape_material_kd[s]));
        REAL mu = RSQRT(R(m->soft_contact_mu) * R(m->shape_material_mu[s]));
        mke[i] = ke;
        mkd[i] = kd;
        mmu[i] = mu;
        pk[i] = R(p->k_start) < R(0.0) ? ke : rmin(R(p->k_start), ke);
    }
}

static REAL shape_margin_of(const OracleModel* m, int s) {
    return (m->shape_count > 0 && m->shape_margin) ? R(m->shape_margin[s]) : R(0.0);
}

/* Soft-contact point of an edge/face record (corners (v0, v1, -1) / (v0, v1, v2), R:1466-1478,
 * R:4906-4919): x = sum_i bary[i] * pos[c_i] over the valid corners in corner order, the anchor likewise
 * from pos_prev (may be NULL), and the largest corner radius. */
static void ef_contact_point(const OracleModel* m, const int32_t* cr, V3 bary, const REAL* q, const REAL* q_prev,
                             V3* x, V3* x_prev, REAL* radius) {
    *x = vscale(vld(q, cr[0]), bary.x);
    if (q_prev) *x_prev = vscale(vld(q_prev, cr[0]), bary.x);
    *radius = R(m->particle_radius[cr[0]]);
    if (cr[1] >= 0) {
        *x = vadd(*x, vscale(vld(q, cr[1]), bary.y));
        if (q_prev) *x_prev = vadd(*x_prev, vscale(vld(q_prev, cr[1]), bary.y));
        *radius = rmax(*radius, R(m->particle_radius[cr[1]]));
    }
    if (cr[2] >= 0) {
        *x = vadd(*x, vscale(vld(q, cr[2]), bary.z));
        if (q_prev) *x_prev = vadd(*x_prev, vscale(vld(q_prev, cr[2]), bary.z));
        *radius = rmax(*radius, R(m->particle_radius[cr[2]]));
    }
}

/* per-vertex contact sum in contact-index order (the serial atomic_add order of K:2289-2401): a particle
 * record (p, -1, -1) adds its force to p; an edge/face record evaluates the force law at its barycentric
 * point and adds bary[i] * f and bary[i]^2 * H to each corner of the color (R:1440-1536).  Returns -4 for
 * a contact against a body whose poses were not given. */
static int accumulate_contacts(const OracleModel* m, const OracleParams* p, int ncontact, const int32_t* c_indices,
                               const int32_t* c_shape, const REAL* c_body_pos, const REAL* c_body_vel,
                               const REAL* c_normal, const REAL* c_bary, const REAL* pk, const REAL* mkd,
                               const REAL* mmu, const REAL* q, const REAL* q_prev, REAL dt, int color, REAL* pf,
                               REAL* ph) {
    for (int t = 0; t < ncontact; ++t) {
        const int32_t* cr = c_indices + 3 * t;
        V3 fc, bx, bv;
        M3 hc;
        if (cr[1] < 0) {
            int pi = cr[0];
            if (color >= 0 && m->particle_colors[pi] != color) continue;
            if (contact_body_frame(m, p, c_shape[t], vld(c_body_pos, t), vld(c_body_vel, t), dt, &bx, &bv)) return -4;
            contact_force_law(vld(q, pi), vld(q_prev, pi), bx, bv, vld(c_normal, t), R(m->particle_radius[pi]),
                              shape_margin_of(m, c_shape[t]), pk[t], mkd[t], mmu[t], R(p->friction_epsilon), dt,
                              &fc, &hc);
            vst(pf, pi, vadd(vld(pf, pi), fc));
            mstore(ph + 9 * pi, madd(mload(ph + 9 * pi), hc));
            continue;
        }
        if (!c_bary) return -5; /* edge/face records need their barycentric weights */
        V3 w = vld(c_bary, t);
        int any = color < 0;
        for (int i = 0; i < 3; ++i)
            if (cr[i] >= 0 && m->particle_colors[cr[i]] == color) any = 1;
        if (!any) continue; /* the force would be added to no corner */
        V3 x, xp;
        REAL radius;
        ef_contact_point(m, cr, w, q, q_prev, &x, &xp, &radius);
        if (contact_body_frame(m, p, c_shape[t], vld(c_body_pos, t), vld(c_body_vel, t), dt, &bx, &bv)) return -4;
        contact_force_law(x, xp, bx, bv, vld(c_normal, t), radius, shape_margin_of(m, c_shape[t]), pk[t], mkd[t],
                          mmu[t], R(p->friction_epsilon), dt, &fc, &hc);
        for (int i = 0; i < 3; ++i) {
            int ci = cr[i];
            if (ci < 0 || (color >= 0 && m->particle_colors[ci] != color)) continue;
            REAL wi = vcomp(w, i);
            vst(pf, ci, vadd(vld(pf, ci), vscale(fc, wi)));
            mstore(ph + 9 * ci, madd(mload(ph + 9 * ci), mscale(hc, wi * wi)));
        }
    }
    return 0;
}

/* evaluate_spring_force_and_hessian_both_vertices (K:1685-1735): the forces on v0 and v1 and the
 * shared Hessian, in the reference's operation order. */
static void spring_eval_both(const OracleModel* m, int s, const REAL* q, const REAL* q_anchor, REAL dt, V3* f0,
                             V3* f1, M3* h) {
    int v0 = m->spring_indices[2 * s], v1 = m->spring_indices[2 * s + 1];
    V3 diff = vsub(vld(q, v0), vld(q, v1));
    REAL len = rmax(vlength(diff), R(1e-8));
    REAL l0 = R(m->spring_rest_length[s]);
    REAL ke = R(m->spring_stiffness[s]), kd = R(m->spring_damping[s]);
    V3 base = vscale(diff, ke * (l0 - len) / len);
    M3 nd = mouter(diff, diff);
    M3 inner;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) inner.m[i][j] = (i == j ? R(1.0) : R(0.0)) - nd.m[i][j] / (len * len);
    M3 structural;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) structural.m[i][j] = (i == j ? R(1.0) : R(0.0)) - (l0 / len) * inner.m[i][j];
    M3 hs = mscale(structural, ke);
    V3 dir = vdivs(diff, len);
    V3 diff_a = vsub(vld(q_anchor, v0), vld(q_anchor, v1));
    REAL len_a = rmax(vlength(diff_a), R(1e-8));
    REAL dlen = len - len_a;
    REAL kdt = kd / dt;
    M3 hd;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) hd.m[i][j] = (vcomp(dir, i) * vcomp(dir, j)) * kdt;
    V3 fd0 = vscale(vneg(dir), kdt * dlen);
    V3 fd1 = vscale(dir, kdt * dlen);
    *f0 = vadd(base, fd0);
    *f1 = vadd(vneg(base), fd1);
    *h = madd(hs, hd);
}

void API(vbd_oracle_spring)(const REAL* x, const REAL* anchor, REAL rest, REAL ke, REAL kd, REAL dt, REAL* f,
                            REAL* h) {
    int32_t idx[2] = {0, 1};
    float rl = (float)rest, k = (float)ke, d = (float)kd;
    OracleModel m;
    memset(&m, 0, sizeof(m));
    m.spring_count = 1;
    m.spring_indices = idx;
    m.spring_rest_length = &rl;
    m.spring_stiffness = &k;
    m.spring_damping = &d;
    V3 f0, f1;
    M3 hm;
    spring_eval_both(&m, 0, x, anchor, dt, &f0, &f1, &hm);
    vst(f, 0, f0);
    vst(f, 1, f1);
    mstore(h, hm);
}

/* accumulate_spring_force_and_hessian (K:1738-1783): after the contacts, in spring index order, for
 * the endpoints of the current color (color < 0: every endpoint), positions of the pass start. */
static void accumulate_springs(const OracleModel* m, const REAL* q, const REAL* q_prev, REAL dt, int color,
                               REAL* pf, REAL* ph) {
    for (int s = 0; s < m->spring_count; ++s) {
        int v0 = m->spring_indices[2 * s], v1 = m->spring_indices[2 * s + 1];
        int c0 = m->particle_colors[v0] == color || color < 0;
        int c1 = m->particle_colors[v1] == color || color < 0;
        if (!c0 && !c1) continue;
        V3 f0, f1;
        M3 h;
        spring_eval_both(m, s, q, q_prev, dt, &f0, &f1, &h);
        if (c0) {
            vst(pf, v0, vadd(vld(pf, v0), f0));
            mstore(ph + 9 * v0, madd(mload(ph + 9 * v0), h));
        }
        if (c1) {
            vst(pf, v1, vadd(vld(pf, v1), f1));
            mstore(ph + 9 * v1, madd(mload(ph + 9 * v1), h));
        }
    }
}

#include "vbd_oracle_self.h"

int32_t API(vbd_oracle_residual)(const OracleModel* m, const OracleParams* p, int32_t contact_count,
                                 int32_t contact_max, const int32_t* c_indices, const int32_t* c_shape,
                                 const REAL* c_body_pos, const REAL* c_body_vel, const REAL* c_normal,
                                 const REAL* c_bary, const REAL* q, const REAL* q_prev, const REAL* inertia, REAL dt,
                                 REAL* g_out, double* norm_out) {
#ifdef ORACLE_IS_F32
    g_use_libm_atan2 = p->use_libm_atan2;
#endif
    int N = m->particle_count;
    int nc = contact_count < contact_max ? contact_count : contact_max;
    REAL* pk = (REAL*)calloc((size_t)(nc > 0 ? nc : 1) * 4, sizeof(REAL));
    REAL *mke = pk + nc, *mkd = pk + 2 * nc, *mmu = pk + 3 * nc;
    REAL* pf = (REAL*)calloc((size_t)N * 3, sizeof(REAL));
    REAL* ph = (REAL*)calloc((size_t)N * 9, sizeof(REAL));
    init_contacts(m, p, nc, c_shape, pk, mke, mkd, mmu);
    int crc = accumulate_contacts(m, p, nc, c_indices, c_shape, c_body_pos, c_body_vel, c_normal, c_bary, pk, mkd,
                                  mmu, q, q_prev, dt, -1, pf, ph);
    if (crc) {
        free(pk);
        free(pf);
        free(ph);
        return crc;
    }
    accumulate_springs(m, q, q_prev, dt, -1, pf, ph);
    if (p->self_contact) {
        /* self-contact terms of the contact set detected at q */
        SelfLists L;
        self_lists_alloc(m, p->self_contact, &L);
        self_detect(m, p->self_contact, q, &L);
        accumulate_self_contacts(m, p, &L, q, q_prev, dt, -1, pf, ph);
        self_lists_free(&L);
    }
    Ctx c = {m, q, q_prev, dt};
    REAL* gb = g_out ? g_out : (REAL*)calloc((size_t)N * 3, sizeof(REAL));
#ifdef _OPENMP
    int nt = p->num_threads > 1 ? p->num_threads : 1;
#pragma omp parallel for num_threads(nt) schedule(static) if (nt > 1)
#endif
    for (int i = 0; i < N; ++i) {
        V3 f;
        M3 h;
        V3 g = v3(R(0.0), R(0.0), R(0.0));
        if (assemble_vertex(&c, i, inertia, &f, &h)) g = vadd(f, vld(pf, i));
        vst(gb, i, g);
    }
    double acc = 0.0; /* serial, in particle order: the same sum for any thread count */
    for (int i = 0; i < N; ++i) {
        V3 g = vld(gb, i);
        acc += (double)g.x * (double)g.x + (double)g.y * (double)g.y + (double)g.z * (double)g.z;
    }
    *norm_out = sqrt(acc);
    if (!g_out) free(gb);
    free(pk);
    free(pf);
    free(ph);
    return 0;
}

/* ---- SolverVBD.step, particle-only: S:1773-1825 -------------------------- */
int32_t API(vbd_oracle_step)(const OracleModel* m, const OracleParams* p, int32_t contact_count,
                             int32_t contact_max, const int32_t* c_indices, const int32_t* c_shape,
                             const REAL* c_body_pos, const REAL* c_body_vel, const REAL* c_normal,
                             const REAL* c_bary, REAL* q, const REAL* qd, const REAL* f_ext, REAL* q_out,
                             REAL* qd_out, REAL dt, double* residual_hist) {
#ifdef ORACLE_IS_F32
    g_use_libm_atan2 = p->use_libm_atan2;
#endif
    int N = m->particle_count;
    if (N <= 0) return 0;
    int nc = contact_count < contact_max ? contact_count : contact_max;
    if (nc < 0) nc = 0;
    REAL* buf = (REAL*)calloc((size_t)N * (3 * 5 + 9), sizeof(REAL));
    if (!buf) return -1;
    REAL *q_prev = buf, *inertia = buf + 3 * N, *x0 = buf + 6 * N, *disp = buf + 9 * N, *pf = buf + 12 * N,
         *ph = buf + 15 * N;
    REAL* cbuf = (REAL*)calloc((size_t)(nc > 0 ? nc : 1) * 4, sizeof(REAL));
    REAL *pk = cbuf, *mke = cbuf + nc, *mkd = cbuf + 2 * nc, *mmu = cbuf + 3 * nc;

    /* _initialize_rigid_bodies -> init_body_particle_contacts (S:2493-2514) */
    init_contacts(m, p, nc, c_shape, pk, mke, mkd, mmu);

    /* self-contact: detection before initialization (S:2074-2075), positions of the last
     * detection (pos_cd, the base of the displacements) and truncation factors */
    const OracleSelf* sc = p->self_contact;
    SelfLists SL;
    REAL *pos_cd = NULL, *ts = NULL;
    if (sc) {
        self_lists_alloc(m, sc, &SL);
        pos_cd = (REAL*)malloc(sizeof(REAL) * 3 * (size_t)N);
        ts = (REAL*)malloc(sizeof(REAL) * (size_t)N);
        self_detect(m, sc, q, &SL);
        self_publish(m, sc, &SL);
        memcpy(pos_cd, q, sizeof(REAL) * 3 * (size_t)N);
    }

    /* _initialize_particles (S:2065-2100): x0 <- q, d <- 0, forward_step (K:1350-1375) */
    memcpy(x0, q, sizeof(REAL) * 3 * (size_t)N);
    for (int i = 0; i < N; ++i) {
        V3 x = vld(q, i);
        vst(q_prev, i, x);
        if (!(m->particle_flags[i] & 1) || m->particle_inv_mass[i] == 0.0f) {
            vst(inertia, i, x);
            vst(disp, i, v3(R(0.0), R(0.0), R(0.0)));
            continue;
        }
        V3 g = v3(R(m->gravity[0]), R(m->gravity[1]), R(m->gravity[2]));
        V3 fe = f_ext ? vld(f_ext, i) : v3(R(0.0), R(0.0), R(0.0));
        V3 vnew = vadd(vld(qd, i), vscale(vadd(g, vscale(fe, R(m->particle_inv_mass[i]))), dt));
        vst(inertia, i, vadd(x, vscale(vnew, dt)));
        vst(disp, i, vscale(vnew, dt));
    }
    /* apply_truncation_ts with t = 1, max = inf (K:2266-2286): q = x0 + d; with self-contact the
     * planar truncation of S:2025-2063 */
    if (sc)
        self_truncation(m, sc, &SL, pos_cd, disp, q, ts);
    else
        for (int i = 0; i < N; ++i) vst(q, i, vadd(vld(x0, i), vscale(vld(disp, i), R(1.0))));

    Ctx c = {m, q, q_prev, dt};
    int run = p->iterations;
    int rc = 0;
    if (residual_hist) {
        API(vbd_oracle_residual)(m, p, contact_count, contact_max, c_indices, c_shape, c_body_pos, c_body_vel,
                                 c_normal, c_bary, q, q_prev, inertia, dt, NULL, &residual_hist[0]);
        if (p->residual_stop > 0.0 && residual_hist[0] <= 0.0) run = 0;
    }
    for (int it = 0; it < run; ++it) {
        /* self-contact re-detection (S:2540-2545): positions and displacement base reset */
        if (sc && ((sc->detection_interval == 0 && it == 0) ||
                   (sc->detection_interval >= 1 && it % sc->detection_interval == 0))) {
            self_detect(m, sc, q, &SL);
            self_publish(m, sc, &SL);
            memcpy(pos_cd, q, sizeof(REAL) * 3 * (size_t)N);
            memset(disp, 0, sizeof(REAL) * 3 * (size_t)N);
        }
        /* update_duals_body_particle_contacts (R:4861-4926): the penetration of the particle, or of an
         * edge/face record's barycentric point with its largest corner radius, at the body point */
        for (int t = 0; t < nc; ++t) {
            const int32_t* cr = c_indices + 3 * t;
            V3 n = vld(c_normal, t), bx, bv, xc, xp;
            REAL radius;
            if (contact_body_frame(m, p, c_shape[t], vld(c_body_pos, t), vld(c_body_vel, t), dt, &bx, &bv)) {
                rc = -4;
                break;
            }
            if (cr[1] < 0) {
                xc = vld(q, cr[0]);
                radius = R(m->particle_radius[cr[0]]);
            } else {
                if (!c_bary) {
                    rc = -5;
                    break;
                }
                ef_contact_point(m, cr, vld(c_bary, t), q, NULL, &xc, &xp, &radius);
            }
            REAL pen = -(vdot(n, vsub(xc, bx)) - radius - shape_margin_of(m, c_shape[t]));
            pen = rmax(R(0.0), pen);
            pk[t] = rmin(pk[t] + R(p->avbd_beta) * pen, mke[t]);
        }
        if (rc) break;
        /* _solve_particle_iteration (S:2516-2704) */
        memset(pf, 0, sizeof(REAL) * 3 * (size_t)N);
        memset(ph, 0, sizeof(REAL) * 9 * (size_t)N);
        for (int color = 0; color < m->color_count; ++color) {
            rc = accumulate_contacts(m, p, nc, c_indices, c_shape, c_body_pos, c_body_vel, c_normal, c_bary, pk, mkd,
                                     mmu, q, q_prev, dt, color, pf, ph);
            if (rc) break;
            accumulate_springs(m, q, q_prev, dt, color, pf, ph);
            if (sc) accumulate_self_contacts(m, p, &SL, q, q_prev, dt, color, pf, ph);
            int cb = m->color_offsets[color], ce = m->color_offsets[color + 1];
#ifdef _OPENMP
            int nt = p->num_threads > 1 ? p->num_threads : 1;
#pragma omp parallel for num_threads(nt) schedule(static) if (nt > 1)
#endif
            for (int k = cb; k < ce; ++k) {
                int i = m->color_particles[k];
                V3 f;
                M3 h;
                if (!assemble_vertex(&c, i, inertia, &f, &h)) {
                    vst(disp, i, v3(R(0.0), R(0.0), R(0.0)));
                    continue;
                }
                h = madd(h, mload(ph + 9 * i));
                f = vadd(f, vld(pf, i));
                V3 dx;
                if (block_solve(h, f, p->fp64_ill_conditioned, &dx)) vst(disp, i, vadd(vld(disp, i), dx));
            }
            if (sc)
                self_truncation(m, sc, &SL, pos_cd, disp, q, ts);
            else
                for (int i = 0; i < N; ++i) vst(q, i, vadd(vld(x0, i), vscale(vld(disp, i), R(1.0))));
        }
        if (rc) break;
        memcpy(q_out, q, sizeof(REAL) * 3 * (size_t)N);
        if (residual_hist) {
            API(vbd_oracle_residual)(m, p, contact_count, contact_max, c_indices, c_shape, c_body_pos,
                                     c_body_vel, c_normal, c_bary, q, q_prev, inertia, dt, NULL,
                                     &residual_hist[it + 1]);
            if (p->residual_stop > 0.0 && residual_hist[it + 1] <= p->residual_stop * residual_hist[0]) break;
        }
    }
    if (run <= 0) memcpy(q_out, q, sizeof(REAL) * 3 * (size_t)N);
    /* _finalize_particles -> update_velocity (K:1468-1471) */
    for (int i = 0; i < N; ++i) vst(qd_out, i, vdivs(vsub(vld(q_out, i), vld(q_prev, i)), dt));
    if (sc) {
        self_lists_free(&SL);
        free(pos_cd);
        free(ts);
    }
    free(buf);
    free(cbuf);
    return rc;
}

/* ---- soft-contact generation: create_soft_contacts (G:1020-1184) and the full-surface edge / face
 * passes (soft_contacts_sdf.py:240-551), plane and sphere shapes, static or attached to a body ---- */

/* shape-local -> world (X_ws) and its inverse; X_bs alone for a static shape (transform_multiply of the
 * identity and X_bs is X_bs up to signed zeros) */
static int shape_frames(const OracleModel* m, const float* body_q, int32_t body_count, int s, Xf* X_bs, Xf* X_ws,
                        Xf* X_sw) {
    *X_bs = xf_load(m->shape_transform + 7 * (size_t)s);
    int b = m->shape_body ? m->shape_body[s] : -1;
    if (b >= 0) {
        if (!body_q || b >= body_count) return -2;
        *X_ws = xf_mul(xf_load(body_q + 7 * (size_t)b), *X_bs);
    } else {
        *X_ws = *X_bs;
    }
    *X_sw = xf_inverse(*X_ws);
    return 0;
}

/* eval_shape_sdf (soft_contacts_sdf.py:61-96) for the analytic shapes of this path: sphere |x| - r with the
 * gradient x / |x| ((0, 0, 1) at the centre, G:181-211), infinite plane x.z with gradient +z (G:584-620) */
static REAL shape_sdf(int type, const float* sc, V3 x, V3* grad) {
    if (type == 3) {
        REAL d = vlength(x) - R(sc[0]);
        REAL l = vlength(x);
        *grad = l > R(1.0e-8) ? vdivs(x, l) : v3(R(0.0), R(0.0), R(1.0));
        return d;
    }
    *grad = v3(R(0.0), R(0.0), R(1.0));
    return x.z;
}

/* optimize_edge_sdf (soft_contacts_sdf.py:136-175): golden-section search of the SDF along p -> q with a
 * fixed iteration count; returns u, the point, its distance and gradient */
static REAL opt_edge_sdf(int type, const float* sc, V3 p, V3 q, int n_iter, V3* x_out, REAL* phi_out, V3* g_out) {
    const REAL inv_phi = R(0.6180339887498949);
    REAL lo = R(0.0), hi = R(1.0);
    REAL c = hi - (hi - lo) * inv_phi;
    REAL d = lo + (hi - lo) * inv_phi;
    V3 g;
    REAL fc = shape_sdf(type, sc, vadd(vscale(p, R(1.0) - c), vscale(q, c)), &g);
    REAL fd = shape_sdf(type, sc, vadd(vscale(p, R(1.0) - d), vscale(q, d)), &g);
    for (int i = 0; i < n_iter; ++i) {
        if (fc < fd) {
            hi = d;
            d = c;
            fd = fc;
            c = hi - (hi - lo) * inv_phi;
            fc = shape_sdf(type, sc, vadd(vscale(p, R(1.0) - c), vscale(q, c)), &g);
        } else {
            lo = c;
            c = d;
            fc = fd;
            d = lo + (hi - lo) * inv_phi;
            fd = shape_sdf(type, sc, vadd(vscale(p, R(1.0) - d), vscale(q, d)), &g);
        }
    }
    REAL u = R(0.5) * (lo + hi);
    *x_out = vadd(vscale(p, R(1.0) - u), vscale(q, u));
    *phi_out = shape_sdf(type, sc, *x_out, g_out);
    return u;
}

static V3 bary_point(V3 a, V3 b, V3 c, V3 w) {
    return vadd(vadd(vscale(a, w.x), vscale(b, w.y)), vscale(c, w.z));
}

/* optimize_face_sdf (soft_contacts_sdf.py:178-219): Frank-Wolfe on the barycentric simplex from the
 * centroid, each step a golden-section line search toward the simplex vertex minimising grad . corner */
static V3 opt_face_sdf(int type, const float* sc, V3 a, V3 b, V3 c, int n_iter, int ls_iter, V3* x_out,
                       REAL* phi_out, V3* g_out) {
    REAL third = R(1.0) / R(3.0);
    V3 bary = v3(third, third, third);
    for (int i = 0; i < n_iter; ++i) {
        V3 x = bary_point(a, b, c, bary);
        V3 g;
        shape_sdf(type, sc, x, &g);
        REAL da = vdot(g, a), db = vdot(g, b), dc = vdot(g, c);
        V3 s = v3(R(1.0), R(0.0), R(0.0));
        if (db <= da && db <= dc)
            s = v3(R(0.0), R(1.0), R(0.0));
        else if (dc <= da && dc <= db)
            s = v3(R(0.0), R(0.0), R(1.0));
        V3 target = bary_point(a, b, c, s);
        V3 lx, lg;
        REAL lphi;
        REAL gamma = opt_edge_sdf(type, sc, x, target, ls_iter, &lx, &lphi, &lg);
        bary = vadd(vscale(bary, R(1.0) - gamma), vscale(s, gamma));
    }
    *x_out = bary_point(a, b, c, bary);
    *phi_out = shape_sdf(type, sc, *x_out, g_out);
    return bary;
}

/* the soft-contact counter: every hit counts, records past contact_max are dropped (G:986-1010) */
static void emit_record(int* count, int32_t contact_max, int32_t* c_particle, int32_t* c_indices, REAL* c_bary,
                        int32_t* c_shape, REAL* c_body_pos, REAL* c_body_vel, REAL* c_normal, int particle, int i0,
                        int i1, int i2, V3 bary, int s, V3 bp, V3 bv, V3 wn) {
    int index = (*count)++;
    if (index >= contact_max) return;
    c_shape[index] = s;
    vst(c_body_pos, index, bp);
    vst(c_body_vel, index, bv);
    if (c_particle) c_particle[index] = particle;
    c_indices[3 * index] = i0;
    c_indices[3 * index + 1] = i1;
    c_indices[3 * index + 2] = i2;
    if (c_bary) vst(c_bary, index, bary);
    vst(c_normal, index, wn);
}

int32_t API(vbd_oracle_create_soft_contacts)(const OracleModel* m, const OraclePairs* pairs, const REAL* q,
                                             const float* body_q, int32_t body_count, REAL margin,
                                             int32_t contact_max, int32_t* c_count, int32_t* c_particle,
                                             int32_t* c_indices, REAL* c_bary, int32_t* c_shape, REAL* c_body_pos,
                                             REAL* c_body_vel, REAL* c_normal) {
    int count = 0;
    V3 zero = v3(R(0.0), R(0.0), R(0.0));
    /* particle pass (G:1020-1184), pair order */
    for (int tid = 0; tid < pairs->pair_count; ++tid) {
        int pi = pairs->pairs[2 * tid], s = pairs->pairs[2 * tid + 1];
        if ((m->particle_flags[pi] & 1) == 0) continue;
        if ((m->shape_flags[s] & 4) == 0) continue;
        int pw = m->particle_world[pi], sw = m->shape_world[s];
        if (pw != -1 && sw != -1 && pw != sw) continue;
        V3 px = vld(q, pi);
        REAL radius = R(m->particle_radius[pi]);
        Xf X_bs, X_ws, X_sw;
        if (shape_frames(m, body_q, body_count, s, &X_bs, &X_ws, &X_sw)) return -2;
        V3 xl = xf_point(X_sw, px);
        const float* sc = m->shape_scale + 3 * s;
        REAL s_margin = shape_margin_of(m, s);
        REAL d = R(1.0e6);
        V3 n = zero;
        V3 v = zero;
        int type = m->shape_type[s];
        if (type == 3) {
            REAL pl = vlength(xl);
            d = pl - R(sc[0]);
            REAL pl2 = vlength(xl);
            n = pl2 > R(1.0e-8) ? vdivs(xl, pl2) : v3(R(0.0), R(0.0), R(1.0));
        } else if (type == 1) {
            REAL width = R(sc[0]) * R(0.5), length = R(sc[1]) * R(0.5);
            if (width > R(0.0) && length > R(0.0)) {
                REAL dd = rmax(RABS(xl.x) - width, RABS(xl.y) - length);
                d = rmax(dd, RABS(xl.z));
            } else {
                d = xl.z;
            }
            n = v3(R(0.0), R(0.0), R(1.0));
        } else {
            return -3; /* other shape types: not on this path */
        }
        if (d < margin + s_margin + radius)
            emit_record(&count, contact_max, c_particle, c_indices, c_bary, c_shape, c_body_pos, c_body_vel, c_normal,
                        pi, pi, -1, -1, v3(R(1.0), R(0.0), R(0.0)), s, xf_point(X_bs, vsub(xl, vscale(n, d))),
                        xf_vector(X_bs, v), xf_vector(X_ws, n));
    }
    /* edge pass (soft_contacts_sdf.py:380-469), edge-pair order */
    for (int tid = 0; tid < pairs->edge_pair_count; ++tid) {
        int e = pairs->edge_pairs[2 * tid], s = pairs->edge_pairs[2 * tid + 1];
        if ((m->shape_flags[s] & 4) == 0) continue;
        int type = m->shape_type[s];
        if (type != 1 && type != 3) return -3;
        int v0 = m->edge_indices[4 * e + 2], v1 = m->edge_indices[4 * e + 3];
        REAL radius = rmax(R(m->particle_radius[v0]), R(m->particle_radius[v1]));
        Xf X_bs, X_ws, X_sw;
        if (shape_frames(m, body_q, body_count, s, &X_bs, &X_ws, &X_sw)) return -2;
        V3 p_s = xf_point(X_sw, vld(q, v0));
        V3 q_s = xf_point(X_sw, vld(q, v1));
        const float* sc = m->shape_scale + 3 * s;
        REAL threshold = margin + shape_margin_of(m, s) + radius;
        V3 mid = vscale(vadd(p_s, q_s), R(0.5));
        V3 g;
        REAL phi_m = shape_sdf(type, sc, mid, &g);
        if (phi_m > threshold + R(0.5) * vlength(vsub(q_s, p_s))) continue;
        V3 x;
        REAL phi;
        REAL u = opt_edge_sdf(type, sc, p_s, q_s, 24, &x, &phi, &g);
        if (phi < threshold) {
            V3 y = vsub(x, vscale(g, phi));
            emit_record(&count, contact_max, c_particle, c_indices, c_bary, c_shape, c_body_pos, c_body_vel, c_normal,
                        -1, v0, v1, -1, v3(R(1.0) - u, u, R(0.0)), s, xf_point(X_bs, y), zero, xf_vector(X_ws, g));
        }
    }
    /* face pass (soft_contacts_sdf.py:277-377), face-pair order */
    for (int tid = 0; tid < pairs->face_pair_count; ++tid) {
        int t = pairs->face_pairs[2 * tid], s = pairs->face_pairs[2 * tid + 1];
        if ((m->shape_flags[s] & 4) == 0) continue;
        int type = m->shape_type[s];
        if (type != 1 && type != 3) return -3;
        int ia = m->tri_indices[3 * t], ib = m->tri_indices[3 * t + 1], ic = m->tri_indices[3 * t + 2];
        REAL radius = rmax(R(m->particle_radius[ia]), rmax(R(m->particle_radius[ib]), R(m->particle_radius[ic])));
        Xf X_bs, X_ws, X_sw;
        if (shape_frames(m, body_q, body_count, s, &X_bs, &X_ws, &X_sw)) return -2;
        V3 a_s = xf_point(X_sw, vld(q, ia));
        V3 b_s = xf_point(X_sw, vld(q, ib));
        V3 c_s = xf_point(X_sw, vld(q, ic));
        const float* sc = m->shape_scale + 3 * s;
        REAL threshold = margin + shape_margin_of(m, s) + radius;
        V3 cen = vdivs(vadd(vadd(a_s, b_s), c_s), R(3.0));
        V3 g;
        REAL phi_c = shape_sdf(type, sc, cen, &g);
        REAL reach = rmax(vlength(vsub(a_s, cen)), rmax(vlength(vsub(b_s, cen)), vlength(vsub(c_s, cen))));
        if (phi_c > threshold + reach) continue;
        V3 x;
        REAL phi;
        V3 bary = opt_face_sdf(type, sc, a_s, b_s, c_s, 24, 16, &x, &phi, &g);
        if (phi < threshold) {
            V3 y = vsub(x, vscale(g, phi));
            emit_record(&count, contact_max, c_particle, c_indices, c_bary, c_shape, c_body_pos, c_body_vel, c_normal,
                        -1, ia, ib, ic, bary, s, xf_point(X_bs, y), zero, xf_vector(X_ws, g));
        }
    }
    *c_count = count;
    return 0;
}

/* ---- frame loop: clear_forces -> collide -> step -> swap ----------------- */
int32_t API(vbd_oracle_simulate)(const OracleModel* m, const OracleParams* p, const OraclePairs* pairs, REAL* q,
                                 REAL* qd, const REAL* f_ext, int32_t substeps, REAL dt, REAL margin,
                                 int32_t contact_max, int32_t record_every, REAL* q_record) {
    int N = m->particle_count;
    int S = contact_max > 0 ? contact_max : 1;
    REAL* q1 = (REAL*)malloc(sizeof(REAL) * 3 * (size_t)N);
    REAL* qd1 = (REAL*)malloc(sizeof(REAL) * 3 * (size_t)N);
    int32_t* ci = (int32_t*)malloc(sizeof(int32_t) * 4 * (size_t)S);
    int32_t* cs = ci + 3 * S;
    REAL* cf = (REAL*)malloc(sizeof(REAL) * 12 * (size_t)S);
    REAL *cbp = cf, *cbv = cf + 3 * S, *cn = cf + 6 * S, *cb = cf + 9 * S;
    int rc = 0;
    for (int s = 0; s < substeps && rc == 0; ++s) {
        int32_t count = 0;
        if (pairs && pairs->pair_count + pairs->edge_pair_count + pairs->face_pair_count > 0 && contact_max > 0) {
            rc = API(vbd_oracle_create_soft_contacts)(m, pairs, q, p->body_q, p->body_count, margin, contact_max,
                                                      &count, NULL, ci, cb, cs, cbp, cbv, cn);
            if (rc) break;
        }
        rc = API(vbd_oracle_step)(m, p, count, contact_max, ci, cs, cbp, cbv, cn, cb, q, qd, f_ext, q1, qd1, dt,
                                  NULL);
        memcpy(q, q1, sizeof(REAL) * 3 * (size_t)N);
        memcpy(qd, qd1, sizeof(REAL) * 3 * (size_t)N);
        if (record_every > 0 && q_record && (s + 1) % record_every == 0)
            memcpy(q_record + (size_t)((s + 1) / record_every - 1) * 3 * (size_t)N, q, sizeof(REAL) * 3 * (size_t)N);
    }
    free(q1);
    free(qd1);
    free(ci);
    free(cf);
    return rc;
}
