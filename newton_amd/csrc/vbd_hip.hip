// libvbd_hip.so -- particle Vertex Block Descent step for MI355X (gfx950).
//
// Kernels:
//   k_forward          inertial predictor + first truncation   (particle_vbd_kernels.py:1350-1375, :2266-2286)
//   k_contact_*        soft-contact material averaging, penalty seed and a
//                      per-particle contact list in contact-index order
//                      (rigid_vbd_kernels.py:799-815, :3276-3321)
//   k_dual_update      per-iteration penalty ramp   (rigid_vbd_kernels.py:4861-4926)
//   k_solve_fast       one color pass over the packed pass layout (vbd_layout.h):
//                      element and contact forces/Hessians evaluated lane-parallel,
//                      serial canonical-order sums, guarded 3x3 solve, position
//                      update   (particle_vbd_kernels.py:2289-2401, :2568-2706)
//   k_solve_generic    the same for vertices with more than 256 elements (CSR walk)
//   k_finalize         velocity update   (particle_vbd_kernels.py:1468-1471)
//   k_integrate_particles  SolverBase.integrate_particles, semi-implicit Euler (solvers/solver.py:20-56)
//   k_collide_tile_*   particle vs static plane/sphere soft contacts (geometry/kernels.py:1020-1184)
//
// Numerics: every FP32 expression follows the reference's CPU kernel order with
// no FMA contraction (built with -ffp-contract=off), IEEE division and square
// root, and the portable atan2 below, so results are bit-identical to the FP32
// CPU oracle.  The per-vertex force is summed serially in CSR order (inertia,
// triangles, bending edges, tetrahedra, then the separately summed contacts),
// exactly as the reference's CPU solve does; the parallelism comes from
// evaluating each vertex's incident elements on separate lanes.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "vbd_hip.h"
#include "vbd_layout.h"

namespace {

thread_local std::string g_last_error;

int32_t fail(int32_t code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

#define VBD_HIP_TRY(expr)                                                                          \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(VBD_ERR_HIP, std::string(#expr " failed: ") + hipGetErrorString(e_));      \
    } while (0)

#define VBD_LAUNCH_CHECK(name)                                                                     \
    do {                                                                                           \
        hipError_t e_ = hipGetLastError();                                                         \
        if (e_ != hipSuccess)                                                                      \
            return fail(VBD_ERR_HIP, std::string("launch of " name " failed: ") + hipGetErrorString(e_)); \
    } while (0)

constexpr int kBlock = 256;     // threads per color-pass workgroup (4 wavefronts)
constexpr int kComp = 12;       // f (3) + H (9) per element contribution

// ---------------------------------------------------------------------------
// small vector algebra; each helper evaluates in the reference's order
// ---------------------------------------------------------------------------
struct V3 {
    float x, y, z;
};
struct M3 {
    float m[3][3];
};

__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 ld3(const float* __restrict__ p, int i) {
    return V3{p[3 * i], p[3 * i + 1], p[3 * i + 2]};
}
__device__ __forceinline__ void st3(float* __restrict__ p, int i, V3 v) {
    p[3 * i] = v.x;
    p[3 * i + 1] = v.y;
    p[3 * i + 2] = v.z;
}
__device__ __forceinline__ V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 neg(V3 a) { return V3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ V3 scl(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 divs(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float length(V3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ float comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

__device__ __forceinline__ M3 mzero() {
    M3 r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r.m[i][j] = 0.0f;
    return r;
}
__device__ __forceinline__ M3 madd(M3 a, M3 b) {
    M3 r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] + b.m[i][j];
    return r;
}
__device__ __forceinline__ M3 mscl(M3 a, float s) {
    M3 r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r.m[i][j] = s * a.m[i][j];
    return r;
}
__device__ __forceinline__ M3 outer(V3 a, V3 b) {
    M3 r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r.m[i][j] = comp(a, i) * comp(b, j);
    return r;
}
__device__ __forceinline__ V3 mvec(const M3& a, V3 v) {
    return V3{a.m[0][0] * v.x + a.m[0][1] * v.y + a.m[0][2] * v.z, a.m[1][0] * v.x + a.m[1][1] * v.y + a.m[1][2] * v.z,
              a.m[2][0] * v.x + a.m[2][1] * v.y + a.m[2][2] * v.z};
}
// transpose(a) * v
__device__ __forceinline__ V3 mtvec(const M3& a, V3 v) {
    return V3{a.m[0][0] * v.x + a.m[1][0] * v.y + a.m[2][0] * v.z, a.m[0][1] * v.x + a.m[1][1] * v.y + a.m[2][1] * v.z,
              a.m[0][2] * v.x + a.m[1][2] * v.y + a.m[2][2] * v.z};
}
// determinant = row0 . (row1 x row2)
__device__ __forceinline__ float mdet(const M3& a) {
    V3 r0 = v3(a.m[0][0], a.m[0][1], a.m[0][2]);
    V3 r1 = v3(a.m[1][0], a.m[1][1], a.m[1][2]);
    V3 r2 = v3(a.m[2][0], a.m[2][1], a.m[2][2]);
    return dot(r0, cross(r1, r2));
}

// Products with skew matrices.  skew(v) = [[0,-vz,vy],[vz,0,-vx],[-vy,vx,0]];
// the reference multiplies the full 3x3, whose zero entries add signed zeros
// only, so the two non-zero terms per entry are kept in the same order.
// A * skew(v):
__device__ __forceinline__ M3 mul_skew_right(const M3& A, V3 v) {
    M3 r;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        r.m[i][0] = A.m[i][1] * v.z + A.m[i][2] * (-v.y);
        r.m[i][1] = A.m[i][0] * (-v.z) + A.m[i][2] * v.x;
        r.m[i][2] = A.m[i][0] * v.y + A.m[i][1] * (-v.x);
    }
    return r;
}
// skew(a) * B:
__device__ __forceinline__ M3 mul_skew_left(V3 a, const M3& B) {
    M3 r;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        r.m[0][j] = (-a.z) * B.m[1][j] + a.y * B.m[2][j];
        r.m[1][j] = a.z * B.m[0][j] + (-a.x) * B.m[2][j];
        r.m[2][j] = (-a.y) * B.m[0][j] + a.x * B.m[1][j];
    }
    return r;
}

// Portable FP32 atan2 (Cephes atanf reduction, S. L. Moshier), same steps as the oracle's.
__device__ __forceinline__ float portable_atan2f(float y, float x) {
    const float PI_F = 3.14159265358979323846f;
    const float PIO2_F = 1.57079632679489661923f;
    const float PIO4_F = 0.78539816339744830962f;
    float ax = fabsf(x), ay = fabsf(y);
    float r;
    if (ax == 0.0f && ay == 0.0f) {
        r = signbit(x) ? PI_F : 0.0f;
    } else {
        bool swap = ay > ax;
        // one division with selected operands (two divisions under a divergent branch otherwise)
        float t = (swap ? ax : ay) / (swap ? ay : ax);
        float base = 0.0f;
        if (t > 0.41421356237309504880f) {
            base = PIO4_F;
            t = (t - 1.0f) / (t + 1.0f);
        }
        float z = t * t;
        float poly = (((8.05374449538e-2f * z - 1.38776856032e-1f) * z + 1.99777106478e-1f) * z - 3.33329491539e-1f) *
                         z * t +
                     t;
        r = base + poly;
        if (swap) r = PIO2_F - r;
        if (signbit(x)) r = PI_F - r;
    }
    return signbit(y) ? -r : r;
}

// ---------------------------------------------------------------------------
// element force / Hessian contributions for one (vertex, element) pair
// ---------------------------------------------------------------------------

// Stable Neo-Hookean membrane, particle_vbd_kernels.py:502-632.
__device__ void eval_membrane(V3 X0, V3 X1, V3 X2, const float* __restrict__ qp, int i0, int i1, int i2,
                              const float* __restrict__ pose, float area, float mu, float lmbd, float damping, float dt,
                              int v_order, V3& f_out, M3& h_out) {
    V3 x0 = X0;
    V3 x01 = sub(X1, x0);
    V3 x02 = sub(X2, x0);
    float D00 = pose[0], D01 = pose[1], D10 = pose[2], D11 = pose[3];
    V3 f0 = add(scl(x01, D00), scl(x02, D10));
    V3 f1 = add(scl(x01, D01), scl(x02, D11));
    float f0f0 = dot(f0, f0), f1f1 = dot(f1, f1), f0f1 = dot(f0, f1);
    float Jsq = f0f0 * f1f1 - f0f1 * f0f1;
    Jsq = fmaxf(Jsq, 1.0e-20f);
    float Js = sqrtf(Jsq);
    float invJ = 1.0f / Js;
    float mu_nh = mu;
    float l_nh = lmbd + mu;
    float l_safe = (l_nh < 0.0f ? -1.0f : 1.0f) * fmaxf(fabsf(l_nh), 1.0e-6f);
    float alpha = 1.0f + mu_nh / l_safe;
    V3 g0 = scl(sub(scl(f0, f1f1), scl(f1, f0f1)), invJ);
    V3 g1 = scl(sub(scl(f1, f0f0), scl(f0, f0f1)), invJ);
    float s = l_nh * (Js - alpha);
    V3 P0 = add(scl(f0, mu_nh), scl(g0, s));
    V3 P1 = add(scl(f1, mu_nh), scl(g1, s));
    float m0 = (float)(v_order == 0), m1 = (float)(v_order == 1), m2 = (float)(v_order == 2);
    float df0 = D00 * (m1 - m0) + D10 * (m2 - m0);
    float df1 = D01 * (m1 - m0) + D11 * (m2 - m0);
    V3 force = neg(add(scl(P0, df0), scl(P1, df1)));
    float s_clamp = fmaxf(0.0f, s);
    float r = s_clamp * invJ;
    float c1 = l_nh - r;
    float df0sq = df0 * df0, df1sq = df1 * df1;
    V3 dJ = add(scl(g0, df0), scl(g1, df1));
    V3 w = sub(scl(f1, df0), scl(f0, df1));
    float I_coeff = mu_nh * (df0sq + df1sq) + r * (df0sq * f1f1 + df1sq * f0f0 - 2.0f * df0 * df1 * f0f1);
    M3 hess;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            hess.m[i][j] = (I_coeff * (i == j ? 1.0f : 0.0f) + c1 * (comp(dJ, i) * comp(dJ, j))) -
                           r * (comp(w, i) * comp(w, j));
    if (damping > 0.0f) {
        float inv_dt = 1.0f / dt;
        V3 A0 = ld3(qp, i0), A1 = ld3(qp, i1), A2 = ld3(qp, i2);
        V3 x01p = sub(A1, A0);
        V3 x02p = sub(A2, A0);
        V3 f0p = add(scl(x01p, D00), scl(x02p, D10));
        V3 f1p = add(scl(x01p, D01), scl(x02p, D11));
        float c00 = (f0f0 - dot(f0p, f0p)) * inv_dt;
        float c01 = (f0f1 - dot(f0p, f1p)) * inv_dt;
        float c11 = (f1f1 - dot(f1p, f1p)) * inv_dt;
        V3 dc00 = scl(f0, 2.0f * df0);
        V3 dc01 = add(scl(f1, df0), scl(f0, df1));
        V3 dc11 = scl(f1, 2.0f * df1);
        V3 f_damp = add(add(scl(dc00, c00), scl(dc01, 2.0f * c01)), scl(dc11, c11));
        force = add(force, scl(f_damp, -damping));
        float dscale = damping * inv_dt;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                float sum = (comp(dc00, i) * comp(dc00, j) + 2.0f * (comp(dc01, i) * comp(dc01, j))) +
                            comp(dc11, i) * comp(dc11, j);
                hess.m[i][j] = hess.m[i][j] + dscale * sum;
            }
    }
    f_out = scl(force, area);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) h_out.m[i][j] = area * hess.m[i][j];
}

// ((1/len) * (I - n n^T)) * (sign * skew(v)), where sign*skew(v) == skew(sign*v)
__device__ __forceinline__ M3 nvd_skew(float len, V3 nh, V3 v) {
    float s = 1.0f / len;
    M3 P;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) P.m[i][j] = s * ((i == j ? 1.0f : 0.0f) - comp(nh, i) * comp(nh, j));
    return mul_skew_right(P, v);
}

// Dihedral-angle bending, particle_vbd_kernels.py:635-816.  Only the solved
// vertex's angle derivative is formed, in one branch-free form for all four
// slots: with s1, s2 the vectors whose skew matrices enter dn1/dx and dn2/dx
// (zero where the reference uses a zero matrix, :730-741) the reference's
// products are reproduced term for term; the terms it adds for the other
// slots, and its 0/1 mask selection, contribute signed zeros only.
__device__ void eval_bending(const float* __restrict__ q, const float* __restrict__ qp, int i0, int i1, int i2,
                             int i3, float rest_angle, float rest_length, float stiffness, float damping, float dt,
                             int v_order, V3& f_out, M3& h_out) {
    f_out = v3(0.0f, 0.0f, 0.0f);
    h_out = mzero();
    if (i0 == -1 || i1 == -1) return;
    const float eps = 1.0e-6f;
    V3 x0 = ld3(q, i0), x1 = ld3(q, i1), x2 = ld3(q, i2), x3 = ld3(q, i3);
    V3 x02 = sub(x2, x0), x03 = sub(x3, x0), x13 = sub(x3, x1), x12 = sub(x2, x1), e = sub(x3, x2);
    V3 n1 = cross(x02, x03), n2 = cross(x13, x12);
    float n1n = length(n1), n2n = length(n2), en = length(e);
    if (n1n < eps || n2n < eps || en < eps) return;
    V3 n1h = divs(n1, n1n), n2h = divs(n2, n2n), eh = divs(e, en);
    float sin_t = dot(cross(n1h, n2h), eh);
    float cos_t = dot(n1h, n2h);
    float theta = portable_atan2f(sin_t, cos_t);
    float k = stiffness * rest_length;
    float dE = k * (theta - rest_angle);
    // per-component selects (the compiler turns selects of whole vectors into branches)
    const bool o0 = v_order == 0, o1 = v_order == 1, o2 = v_order == 2;
    const V3 m03 = neg(x03), me = neg(e), m12 = neg(x12);
    V3 s1, s2;
    s1.x = o0 ? e.x : (o1 ? 0.0f : (o2 ? m03.x : x02.x));
    s1.y = o0 ? e.y : (o1 ? 0.0f : (o2 ? m03.y : x02.y));
    s1.z = o0 ? e.z : (o1 ? 0.0f : (o2 ? m03.z : x02.z));
    s2.x = o0 ? 0.0f : (o1 ? me.x : (o2 ? x13.x : m12.x));
    s2.y = o0 ? 0.0f : (o1 ? me.y : (o2 ? x13.y : m12.y));
    s2.z = o0 ? 0.0f : (o1 ? me.z : (o2 ? x13.z : m12.z));
    M3 dn1 = nvd_skew(n1n, n1h, s1);
    M3 dn2 = nvd_skew(n2n, n2h, s2);
    M3 a = mul_skew_left(n1h, dn2);
    M3 b = mul_skew_left(n2h, dn1);
    M3 Mm;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Mm.m[i][j] = a.m[i][j] - b.m[i][j];
    V3 dsin = mtvec(Mm, eh);
    V3 dcos = add(mtvec(dn1, n2h), mtvec(dn2, n1h));
    V3 dth = sub(scl(dsin, cos_t), scl(dcos, sin_t));
    V3 bf = scl(dth, -dE);
    M3 bh;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) bh.m[i][j] = k * (comp(dth, i) * comp(dth, j));
    if (damping > 0.0f) {
        float inv_dt = 1.0f / dt;
        V3 p0 = ld3(qp, i0), p1 = ld3(qp, i1), p2 = ld3(qp, i2), p3 = ld3(qp, i3);
        V3 x02p = sub(p2, p0), x03p = sub(p3, p0), x13p = sub(p3, p1), x12p = sub(p2, p1), ep = sub(p3, p2);
        V3 n1pr = cross(x02p, x03p), n2pr = cross(x13p, x12p);
        float n1pn = length(n1pr), n2pn = length(n2pr), epn = length(ep);
        if (!(n1pn < eps || n2pn < eps || epn < eps)) {
            V3 n1p = divs(n1pr, n1pn), n2p = divs(n2pr, n2pn), ehp = divs(ep, epn);
            float sinp = dot(cross(n1p, n2p), ehp);
            float cosp = dot(n1p, n2p);
            float thetap = portable_atan2f(sinp, cosp);
            float dtheta = theta - thetap;
            if (dtheta > 3.141592653589793f)
                dtheta = dtheta - 6.283185307179586f;
            else if (dtheta < -3.141592653589793f)
                dtheta = dtheta + 6.283185307179586f;
            float dtheta_dt = dtheta * inv_dt;
            bf = add(bf, scl(dth, -damping * rest_length * dtheta_dt));
            float hs = damping * rest_length * inv_dt;
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) bh.m[i][j] = bh.m[i][j] + hs * (comp(dth, i) * comp(dth, j));
        }
    }
    f_out = bf;
    h_out = bh;
}

// Volumetric stable Neo-Hookean tetrahedron, particle_vbd_kernels.py:76-292, :347-367.
__device__ void eval_tet(const float* __restrict__ q, const float* __restrict__ qp, const int* __restrict__ ti,
                         const float* __restrict__ dm, float mu, float lmbd, float damping, float dt, int v_order,
                         V3& f_out, M3& h_out) {
    M3 Dm;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Dm.m[i][j] = dm[3 * i + j];
    float rest_volume = 1.0f / (mdet(Dm) * 6.0f);
    V3 X0 = ld3(q, ti[0]), X1 = ld3(q, ti[1]), X2 = ld3(q, ti[2]), X3 = ld3(q, ti[3]);
    V3 c0 = sub(X1, X0), c1 = sub(X2, X0), c2 = sub(X3, X0);
    M3 Ds;
    Ds.m[0][0] = c0.x; Ds.m[0][1] = c1.x; Ds.m[0][2] = c2.x;
    Ds.m[1][0] = c0.y; Ds.m[1][1] = c1.y; Ds.m[1][2] = c2.y;
    Ds.m[2][0] = c0.z; Ds.m[2][1] = c1.z; Ds.m[2][2] = c2.z;
    M3 F;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) F.m[i][j] = Ds.m[i][0] * Dm.m[0][j] + Ds.m[i][1] * Dm.m[1][j] + Ds.m[i][2] * Dm.m[2][j];
    float f9[9];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) f9[3 * c + r] = F.m[r][c];
    float J = mdet(F);
    float mu_nh = mu;
    float l_nh = lmbd + mu;
    float l_safe = (l_nh < 0.0f ? -1.0f : 1.0f) * fmaxf(fabsf(l_nh), 1.0e-6f);
    float alpha = 1.0f + mu_nh / l_safe;
    float F11 = F.m[0][0], F21 = F.m[1][0], F31 = F.m[2][0];
    float F12 = F.m[0][1], F22 = F.m[1][1], F32 = F.m[2][1];
    float F13 = F.m[0][2], F23 = F.m[1][2], F33 = F.m[2][2];
    M3 cof;
    cof.m[0][0] = F22 * F33 - F23 * F32;
    cof.m[0][1] = F23 * F31 - F21 * F33;
    cof.m[0][2] = F21 * F32 - F22 * F31;
    cof.m[1][0] = F13 * F32 - F12 * F33;
    cof.m[1][1] = F11 * F33 - F13 * F31;
    cof.m[1][2] = F12 * F31 - F11 * F32;
    cof.m[2][0] = F12 * F23 - F13 * F22;
    cof.m[2][1] = F13 * F21 - F11 * F23;
    cof.m[2][2] = F11 * F22 - F12 * F21;
    float c9[9];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) c9[3 * c + r] = cof.m[r][c];
    float s = l_nh * (J - alpha);
    float Pv[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) Pv[i] = rest_volume * (mu_nh * f9[i] + s * c9[i]);
    float m1, m2, m3;
    if (v_order == 0) {
        m1 = -(Dm.m[0][0] + Dm.m[1][0] + Dm.m[2][0]);
        m2 = -(Dm.m[0][1] + Dm.m[1][1] + Dm.m[2][1]);
        m3 = -(Dm.m[0][2] + Dm.m[1][2] + Dm.m[2][2]);
    } else if (v_order == 1) {
        m1 = Dm.m[0][0]; m2 = Dm.m[0][1]; m3 = Dm.m[0][2];
    } else if (v_order == 2) {
        m1 = Dm.m[1][0]; m2 = Dm.m[1][1]; m3 = Dm.m[1][2];
    } else {
        m1 = Dm.m[2][0]; m2 = Dm.m[2][1]; m3 = Dm.m[2][2];
    }
    V3 force = v3(-(Pv[0] * m1 + Pv[3] * m2 + Pv[6] * m3), -(Pv[1] * m1 + Pv[4] * m2 + Pv[7] * m3),
                  -(Pv[2] * m1 + Pv[5] * m2 + Pv[8] * m3));
    M3 hess;
#define VBD_H9(i, j) (rest_volume * (mu_nh * ((i) == (j) ? 1.0f : 0.0f) + l_nh * (c9[i] * c9[j])))
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            hess.m[r][c] = m1 * (VBD_H9(r, c) * m1 + VBD_H9(r + 3, c) * m2 + VBD_H9(r + 6, c) * m3) +
                           m2 * (VBD_H9(r, c + 3) * m1 + VBD_H9(r + 3, c + 3) * m2 + VBD_H9(r + 6, c + 3) * m3) +
                           m3 * (VBD_H9(r, c + 6) * m1 + VBD_H9(r + 3, c + 6) * m2 + VBD_H9(r + 6, c + 6) * m3);
#undef VBD_H9
    if (damping > 0.0f) {
        float inv_dt = 1.0f / dt;
        V3 P0 = ld3(qp, ti[0]), P1 = ld3(qp, ti[1]), P2 = ld3(qp, ti[2]), P3 = ld3(qp, ti[3]);
        V3 d0 = sub(P1, P0), d1 = sub(P2, P0), d2 = sub(P3, P0);
        M3 Dsp;
        Dsp.m[0][0] = d0.x; Dsp.m[0][1] = d1.x; Dsp.m[0][2] = d2.x;
        Dsp.m[1][0] = d0.y; Dsp.m[1][1] = d1.y; Dsp.m[1][2] = d2.y;
        Dsp.m[2][0] = d0.z; Dsp.m[2][1] = d1.z; Dsp.m[2][2] = d2.z;
        M3 Fp;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                Fp.m[i][j] = Dsp.m[i][0] * Dm.m[0][j] + Dsp.m[i][1] * Dm.m[1][j] + Dsp.m[i][2] * Dm.m[2][j];
        V3 f0 = v3(F.m[0][0], F.m[1][0], F.m[2][0]);
        V3 f1 = v3(F.m[0][1], F.m[1][1], F.m[2][1]);
        V3 f2 = v3(F.m[0][2], F.m[1][2], F.m[2][2]);
        V3 f0p = v3(Fp.m[0][0], Fp.m[1][0], Fp.m[2][0]);
        V3 f1p = v3(Fp.m[0][1], Fp.m[1][1], Fp.m[2][1]);
        V3 f2p = v3(Fp.m[0][2], Fp.m[1][2], Fp.m[2][2]);
        float c00 = (dot(f0, f0) - dot(f0p, f0p)) * inv_dt;
        float c01 = (dot(f0, f1) - dot(f0p, f1p)) * inv_dt;
        float c02 = (dot(f0, f2) - dot(f0p, f2p)) * inv_dt;
        float c11 = (dot(f1, f1) - dot(f1p, f1p)) * inv_dt;
        float c12 = (dot(f1, f2) - dot(f1p, f2p)) * inv_dt;
        float c22 = (dot(f2, f2) - dot(f2p, f2p)) * inv_dt;
        V3 dc00 = scl(f0, 2.0f * m1);
        V3 dc01 = add(scl(f1, m1), scl(f0, m2));
        V3 dc02 = add(scl(f2, m1), scl(f0, m3));
        V3 dc11 = scl(f1, 2.0f * m2);
        V3 dc12 = add(scl(f2, m2), scl(f1, m3));
        V3 dc22 = scl(f2, 2.0f * m3);
        V3 fd = scl(dc00, c00);
        fd = add(fd, scl(dc01, 2.0f * c01));
        fd = add(fd, scl(dc02, 2.0f * c02));
        fd = add(fd, scl(dc11, c11));
        fd = add(fd, scl(dc12, 2.0f * c12));
        fd = add(fd, scl(dc22, c22));
        force = sub(force, scl(fd, rest_volume * damping));
        float hsc = rest_volume * damping * inv_dt;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                float hs = comp(dc00, i) * comp(dc00, j);
                hs = hs + 2.0f * (comp(dc01, i) * comp(dc01, j));
                hs = hs + 2.0f * (comp(dc02, i) * comp(dc02, j));
                hs = hs + comp(dc11, i) * comp(dc11, j);
                hs = hs + 2.0f * (comp(dc12, i) * comp(dc12, j));
                hs = hs + comp(dc22, i) * comp(dc22, j);
                hess.m[i][j] = hess.m[i][j] + hsc * hs;
            }
    }
    f_out = force;
    h_out = hess;
}

// Soft contact force law at the contact point x (anchor x_prev) against the world body point bx moving
// at bv (computed per substep by contact_body_frame when the contacts are packed): penalty + damping +
// IPC friction, rigid_vbd_kernels.py:1331-1439, :1603-1645.
__device__ void eval_contact(V3 x, V3 x_prev, V3 bx, V3 bv, V3 n, float radius, float margin, float ke, float kd,
                             float mu, float friction_epsilon, float dt, V3& f_out, M3& h_out) {
    float pen = -(dot(n, sub(x, bx)) - radius - margin);
    if (!(pen > 0.0f)) {
        f_out = v3(0.0f, 0.0f, 0.0f);
        h_out = mzero();
        return;
    }
    V3 dx = sub(x, x_prev);
    V3 zero = v3(0.0f, 0.0f, 0.0f);
    V3 u = sub(dx, scl(bv, dt));
    float f_n = pen * ke;
    V3 force = scl(n, f_n);
    M3 nn = outer(n, n);
    M3 hess = mscl(nn, ke);
    if (dot(n, u) < 0.0f) {
        M3 dh = mscl(nn, kd / dt);
        hess = madd(hess, dh);
        force = sub(force, mvec(dh, u));
    }
    float eps_u = friction_epsilon * dt;
    float dot_nu = dot(n, u);
    V3 u_t = sub(u, scl(n, dot_nu));
    float u_norm = length(u_t);
    if (u_norm > 0.0f) {
        float f1 = (u_norm > eps_u) ? 1.0f / u_norm : (-u_norm / eps_u + 2.0f) / eps_u;
        float scale = mu * f_n * f1;
        force = add(force, neg(scl(u_t, scale)));
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) hess.m[i][j] = hess.m[i][j] + scale * ((i == j ? 1.0f : 0.0f) - nn.m[i][j]);
    } else {
        force = add(force, zero);
        hess = madd(hess, mzero());
    }
    f_out = force;
    h_out = hess;
}

// Guarded 3x3 solve, particle_vbd_kernels.py:2701-2706: returns false when |det H| <= 1e-8.  With
// fp64_ill the Frobenius condition number ||H||_F ||adj H||_F / |det H| is evaluated in double from
// the stored entries (a single-precision estimate is unreliable exactly for ill-conditioned blocks);
// above 1e6 the block is solved in double, otherwise in the reference's single-precision adjugate order.
__device__ bool block_solve(const M3& h, V3 f, int fp64_ill, V3& dx) {
    float det = mdet(h);
    if (!(fabsf(det) > 1e-8f)) return false;
    if (fp64_ill) {
        double a[3][3], c[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) a[i][j] = (double)h.m[i][j];
        c[0][0] = a[1][1] * a[2][2] - a[1][2] * a[2][1];
        c[1][0] = a[1][2] * a[2][0] - a[1][0] * a[2][2];
        c[2][0] = a[1][0] * a[2][1] - a[1][1] * a[2][0];
        c[0][1] = a[0][2] * a[2][1] - a[0][1] * a[2][2];
        c[1][1] = a[0][0] * a[2][2] - a[0][2] * a[2][0];
        c[2][1] = a[0][1] * a[2][0] - a[0][0] * a[2][1];
        c[0][2] = a[0][1] * a[1][2] - a[0][2] * a[1][1];
        c[1][2] = a[0][2] * a[1][0] - a[0][0] * a[1][2];
        c[2][2] = a[0][0] * a[1][1] - a[0][1] * a[1][0];
        const double det_d = a[0][0] * c[0][0] + a[0][1] * c[1][0] + a[0][2] * c[2][0];
        double n1 = 0.0, n2 = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                n1 += a[i][j] * a[i][j];
                n2 += c[i][j] * c[i][j];
            }
        if (!(sqrt(n1 * n2) <= 1.0e6 * fabs(det_d))) {
            const double inv = 1.0 / det_d;
            const double fd0 = (double)f.x, fd1 = (double)f.y, fd2 = (double)f.z;
            const double r0 = (c[0][0] * inv) * fd0 + (c[0][1] * inv) * fd1 + (c[0][2] * inv) * fd2;
            const double r1 = (c[1][0] * inv) * fd0 + (c[1][1] * inv) * fd1 + (c[1][2] * inv) * fd2;
            const double r2 = (c[2][0] * inv) * fd0 + (c[2][1] * inv) * fd1 + (c[2][2] * inv) * fd2;
            dx = v3((float)r0, (float)r1, (float)r2);
            return true;
        }
    }
    M3 b;
    b.m[0][0] = h.m[1][1] * h.m[2][2] - h.m[1][2] * h.m[2][1];
    b.m[1][0] = h.m[1][2] * h.m[2][0] - h.m[1][0] * h.m[2][2];
    b.m[2][0] = h.m[1][0] * h.m[2][1] - h.m[1][1] * h.m[2][0];
    b.m[0][1] = h.m[0][2] * h.m[2][1] - h.m[0][1] * h.m[2][2];
    b.m[1][1] = h.m[0][0] * h.m[2][2] - h.m[0][2] * h.m[2][0];
    b.m[2][1] = h.m[0][1] * h.m[2][0] - h.m[0][0] * h.m[2][1];
    b.m[0][2] = h.m[0][1] * h.m[1][2] - h.m[0][2] * h.m[1][1];
    b.m[1][2] = h.m[0][2] * h.m[1][0] - h.m[0][0] * h.m[1][2];
    b.m[2][2] = h.m[0][0] * h.m[1][1] - h.m[0][1] * h.m[1][0];
    float inv_det = 1.0f / det;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) b.m[i][j] = b.m[i][j] * inv_det;
    dx = mvec(b, f);
    return true;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
struct KParams {
    int N;
    const float* mass;
    const float* inv_mass;
    const float* radius;
    const int* flags;
    const int* tri_idx;
    const float* tri_pose;
    const float* tri_mat;
    const float* tri_area;
    const int* edge_idx;
    const float* edge_rest_angle;
    const float* edge_rest_len;
    const float* edge_bend;
    const int* tet_idx;
    const float* tet_pose;
    const float* tet_mat;
    const int* adj_tri;
    const int* adj_tri_off;
    const int* adj_edge;
    const int* adj_edge_off;
    const int* adj_tet;
    const int* adj_tet_off;
    int has_tri, has_edge, has_tet;
    const int* color_particles;
    const float* gravity;
    float* q_prev;
    float* inertia;
    float* disp;
    // pass layout (vbd_layout.h)
    const int* batch_blk;
    const int* batch_list;  // when set: launches enumerate these batch indices (live worlds only)
    int aligned;            // item ranges of a batch start on wave boundaries (layout built so)
    const int* bat_cinfo;  // per batch {first packed contact, contact count} (k_batch_contacts)
    const int* bat_coff;   // per batch, max_v + 1 contact offsets relative to the first
    const int* vert_rec;
    const int* tri_rec;
    const int* edge_rec;
    const int* edge_meta;
    const int* tet_rec;
    const int* tet_meta;
    const int* slow_pos;
    // springs: model arrays, batch items, same-color pre items and their rows, slow-vertex lists
    int has_spring;
    const int* spring_idx;
    const float* spring_rest;
    const float* spring_ke;
    const float* spring_kd;
    const int* spr_rec;
    const int* pre_item;
    float* spr_rows;
    const int* slow_spr_off;
    const int* slow_spr;
    // self-contact rows per color position (vbd_self.h) and whether the pass writes q
    int has_self;
    const int* so_off;
    const float* so_rows;
    int write_q;
    // per-world freezing (batched iterate-to-tolerance): when w_active is set, vertices of
    // worlds with w_active[p_world[v]] == 0 are neither solved nor re-evaluated
    const int* p_world;
    const int* w_active;
    // per-contact state
    const float* c_pk;
    const float* c_kd;
    const float* c_mu;
    const int* c_shape;
    const float* c_body_pos;
    const float* c_body_vel;
    const float* c_normal;
    // contacts packed per color position (k_contact_sort, k_contact_pack)
    const int* pos_coff;
    const float* cpack;
    const int* cpack_ef;  // per packed entry of an edge/face record: {c0, c1, c2, corner} {w0, w1, w2, 0}
    float* crows;         // full-surface contexts: every packed entry's row (force 3 + Hessian 9), evaluated
                          // by k_contact_rows at the pass start and read back in the serial sum
    const float* shape_margin;
    int shape_count;
    float friction_epsilon;
    int fp64_ill;
};

// packed contact record (one per corner entry): 4 x float4
//   {particle, local slot | kContactEF, ke, kd} {mu, margin, radius, bx} {by, bz, bvx, bvy} {bvz, nx, ny, nz}
// with bx the world body point and bv the body's velocity there; an edge/face record's entries carry
// the largest corner radius and kContactEF, and their corners and weights in cpack_ef
constexpr int kContactWords = 16;
constexpr int kContactEF = 1 << 30;

__device__ __forceinline__ int wave_align(int n) { return (n + 63) & ~63; }

__device__ __forceinline__ bool particle_active(const KParams& P, int v) {
    return (P.flags[v] & 1) && P.mass[v] != 0.0f;
}

// Work-group -> batch mapping that keeps consecutive batches (neighbouring
// vertices of one color) on the same XCD, whose L2 then serves their shared
// neighbours.  Hardware dispatches work-group i to XCD i % 8.
__device__ __forceinline__ int xcd_remap(int bid, int nb) {
    const int kXcd = 8;
    int xcd = bid % kXcd, idx = bid / kXcd;
    int qn = nb / kXcd, rn = nb % kXcd;
    return xcd * qn + (xcd < rn ? xcd : rn) + idx;
}

__device__ __forceinline__ void store_row(float* __restrict__ row, V3 f, const M3& h) {
    row[0] = f.x;
    row[1] = f.y;
    row[2] = f.z;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) row[3 + 3 * i + j] = h.m[i][j];
}

__device__ __forceinline__ void store_force(float* __restrict__ row, V3 f) {
    row[0] = f.x;
    row[1] = f.y;
    row[2] = f.z;
}

// component c (0..11) of a row {f, H row-major}, by selects (no dynamic register indexing)
__device__ __forceinline__ float row_component(V3 f, const M3& h, int c) {
    const float r[12] = {f.x, f.y, f.z, h.m[0][0], h.m[0][1], h.m[0][2],
                         h.m[1][0], h.m[1][1], h.m[1][2], h.m[2][0], h.m[2][1], h.m[2][2]};
    float v = r[0];
#pragma unroll
    for (int k = 1; k < 12; ++k) v = c == k ? r[k] : v;
    return v;
}

// Packed contact entry rec (its four float4 A-D) at the current positions: a particle record at its
// particle; an edge/face record at x = sum_i w_i q[c_i] over its corners in corner order, its force and
// Hessian scaled by w and w^2 of this entry's corner (particle_vbd_kernels.py:2379-2401,
// rigid_vbd_kernels.py:1440-1536).
__device__ __forceinline__ void eval_contact_rec(const KParams& P, const float* __restrict__ q, int rec, float4 A,
                                                 float4 B, float4 C, float4 D, float dt, V3& f, M3& h) {
    const int p = __float_as_int(A.x);
    if (!(__float_as_int(A.y) & kContactEF)) {
        eval_contact(ld3(q, p), ld3(P.q_prev, p), v3(B.w, C.x, C.y), v3(C.z, C.w, D.x), v3(D.y, D.z, D.w), B.z, B.y,
                     A.z, A.w, B.x, P.friction_epsilon, dt, f, h);
        return;
    }
    const int4* e = reinterpret_cast<const int4*>(P.cpack_ef) + 2 * (size_t)rec;
    const int4 ci = e[0];
    const float4 w = __builtin_bit_cast(float4, e[1]);
    V3 x = scl(ld3(q, ci.x), w.x), xp = scl(ld3(P.q_prev, ci.x), w.x);
    if (ci.y >= 0) {
        x = add(x, scl(ld3(q, ci.y), w.y));
        xp = add(xp, scl(ld3(P.q_prev, ci.y), w.y));
    }
    if (ci.z >= 0) {
        x = add(x, scl(ld3(q, ci.z), w.z));
        xp = add(xp, scl(ld3(P.q_prev, ci.z), w.z));
    }
    V3 fc;
    M3 hc;
    eval_contact(x, xp, v3(B.w, C.x, C.y), v3(C.z, C.w, D.x), v3(D.y, D.z, D.w), B.z, B.y, A.z, A.w, B.x,
                 P.friction_epsilon, dt, fc, hc);
    const float wi = ci.w == 0 ? w.x : (ci.w == 1 ? w.y : w.z);
    f = scl(fc, wi);
    h = mscl(hc, wi * wi);
}

__device__ __forceinline__ void eval_packed_contact(const KParams& P, const float* __restrict__ q, int rec, float dt,
                                                    V3& f, M3& h) {
    const float4* c4 = reinterpret_cast<const float4*>(P.cpack) + 4 * (size_t)rec;
    eval_contact_rec(P, q, rec, c4[0], c4[1], c4[2], c4[3], dt, f, h);
}

// The rows of every packed contact entry of the color positions [p_lo, p_hi) at the positions q of
// the pass start (a grid-stride loop over the device-sized entry range): with full-surface contacts a
// vertex can carry a dozen entries and a batch's entries outnumber its LDS rows, so they are evaluated
// here, one lane each, and the pass sums the stored rows in order.  An entry's force depends on its
// vertex and on corners of other colors only, none of which moves before the vertex is solved, so
// the rows hold the values the pass would evaluate itself.
constexpr int kContactRowsGrid = 1024;  // work-groups of its grid-stride loop
__global__ __launch_bounds__(256) void k_contact_rows(KParams P, const float* __restrict__ q, int p_lo, int p_hi,
                                                      float dt) {
    const int j0 = P.pos_coff[p_lo], j1 = P.pos_coff[p_hi];
    for (int j = j0 + (int)(blockIdx.x * blockDim.x + threadIdx.x); j < j1; j += (int)(gridDim.x * blockDim.x)) {
        V3 f;
        M3 h;
        eval_packed_contact(P, q, j, dt, f, h);
        store_row(P.crows + (size_t)kComp * j, f, h);
    }
}

// evaluate_spring_force_and_hessian_both_vertices (particle_vbd_kernels.py:1685-1730), the
// contribution to one endpoint (0: v0, 1: v1); x* are the pass-start positions, a* the anchors.
__device__ __forceinline__ void eval_spring(V3 x0, V3 x1, V3 a0, V3 a1, float rest, float ke, float kd, float dt,
                                            int endpoint, V3& f, M3& h) {
    const V3 diff = sub(x0, x1);
    const float len = fmaxf(length(diff), 1.0e-8f);
    const V3 base = scl(diff, ke * (rest - len) / len);
    const M3 nd = outer(diff, diff);
    const float lsq = len * len, rl = rest / len;
    M3 hs;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const float id = i == j ? 1.0f : 0.0f;
            const float inner = id - nd.m[i][j] / lsq;
            hs.m[i][j] = ke * (id - rl * inner);
        }
    const V3 dir = divs(diff, len);
    const float len_a = fmaxf(length(sub(a0, a1)), 1.0e-8f);
    const float dlen = len - len_a;
    const float kdt = kd / dt;
    M3 hd;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) hd.m[i][j] = (comp(dir, i) * comp(dir, j)) * kdt;
    const float fdm = kdt * dlen;
    f = endpoint == 0 ? add(base, scl(neg(dir), fdm)) : add(neg(base), scl(dir, fdm));
    h = madd(hs, hd);
}

// Spring item: in-pass evaluation (pre_row < 0) or the pre-pass row.
__device__ __forceinline__ void spring_item(const KParams& P, const float* __restrict__ q, int v0, int v1,
                                            int endpoint, int pre_row, float rest, float ke, float kd, float dt,
                                            V3& f, M3& h) {
    if (pre_row >= 0) {
        const float4* r4 = reinterpret_cast<const float4*>(P.spr_rows) + 3 * (size_t)pre_row;
        const float4 A = r4[0], B = r4[1], C = r4[2];
        f = v3(A.x, A.y, A.z);
        h.m[0][0] = A.w; h.m[0][1] = B.x; h.m[0][2] = B.y;
        h.m[1][0] = B.z; h.m[1][1] = B.w; h.m[1][2] = C.x;
        h.m[2][0] = C.y; h.m[2][1] = C.z; h.m[2][2] = C.w;
    } else {
        eval_spring(ld3(q, v0), ld3(q, v1), ld3(P.q_prev, v0), ld3(P.q_prev, v1), rest, ke, kd, dt, endpoint, f, h);
    }
}

// Pre-pass of one color: rows of the spring items whose endpoints share the color, evaluated
// at the positions of the color pass start (accumulate_spring_force_and_hessian,
// particle_vbd_kernels.py:1733-1783, runs before the color's solve).
__global__ void k_spring_rows(KParams P, const float* __restrict__ q, int j0, int j1, float dt) {
    const int j = j0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= j1) return;
    const int2 it = reinterpret_cast<const int2*>(P.pre_item)[j];
    const int s = it.x;
    const int v0 = P.spring_idx[2 * s], v1 = P.spring_idx[2 * s + 1];
    V3 f;
    M3 h;
    eval_spring(ld3(q, v0), ld3(q, v1), ld3(P.q_prev, v0), ld3(P.q_prev, v1), P.spring_rest[s], P.spring_ke[s],
                P.spring_kd[s], dt, it.y, f, h);
    store_row(P.spr_rows + 12 * (size_t)j, f, h);
}

#include "vbd_self.h"

// Finish one vertex: the block solve and position update (mode 0) or the force output (mode 1).
// x0 and d_old are the vertex's q_prev and displacement (read by the caller).  With self-contact
// on, the pass writes only the displacement; the truncation pass that follows writes q.
__device__ __forceinline__ void finish_vertex(const KParams& P, float* __restrict__ q, int v, bool active, V3 x0,
                                              V3 d_old, V3 f, M3 h, float dt, int mode, float* __restrict__ g_out) {
    if (!active) {
        if (mode == 0) {
            V3 z = v3(0.0f, 0.0f, 0.0f);
            st3(P.disp, v, z);
            if (P.write_q) st3(q, v, add(x0, scl(z, 1.0f)));
        } else {
            st3(g_out, v, v3(0.0f, 0.0f, 0.0f));
        }
        return;
    }
    if (mode == 1) {
        st3(g_out, v, f);
        return;
    }
    V3 dx;
    if (block_solve(h, f, P.fp64_ill, dx)) {
        V3 d = add(d_old, dx);
        st3(P.disp, v, d);
        if (P.write_q) st3(q, v, add(x0, scl(d, 1.0f)));
    }
}

// ca + component cc of the global rows [r, re), added one at a time in row order; eight loads are in
// flight before their additions (the rows' latency is paid once per eight rather than per row)
__device__ __forceinline__ float add_rows_in_order(float ca, const float* __restrict__ rows, int r, int re, int cc) {
    for (; r + 8 <= re; r += 8) {
        float a[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) a[k] = rows[(size_t)kComp * (r + k) + cc];
#pragma unroll
        for (int k = 0; k < 8; ++k) ca = ca + a[k];
    }
    for (; r < re; ++r) ca = ca + rows[(size_t)kComp * r + cc];
    return ca;
}

// One color pass over packed batches (mode 0: solve and update; mode 1: write the assembled force g).
// The work-group's global reads form three dependent round trips:
//   1  the batch block (header + vertex records) and the batch's contact ranges (indexed by batch)
//   2  each lane's item record (triangle / edge / tet / packed contact) and, for lanes t < nv, the
//      vertex's mass, inertia target, position, q_prev and displacement
//   3  the positions (and q_prev when damped) of the items' vertices
// then, in LDS:
//   a  every lane evaluates its item into the item's canonical row
//   b  lane (vertex s, component c) sums component c of s's rows serially, in canonical order,
//      starting from the inertia term; contacts are summed from zero and added afterwards
//   c  the vertex's lane solves the 3x3 block and writes displacement and position
constexpr int kBatchHdr = vbd_layout::kBatchWords;

// BLOCK lanes per work-group: 256 (batches of <= 21 vertices / 256 items) or
// 128 (<= 10 vertices / 128 items).  LDS holds the static items plus 64 contact rows (17 KB at
// 256 lanes), so LDS does not bound occupancy; WAVES is the register budget's occupancy target:
// 1 leaves it to the compiler (77 VGPRs, six waves per SIMD), 7 fits seven waves per SIMD at the
// price of a few values kept in scratch -- faster on cloth, slower on tets (whose evaluation then
// spills more).  The autotuner picks the shape per model.
// FORCE_ONLY (the residual, mode 1): only the force rows are stored and summed, so the Hessian
// arithmetic of every item is dead code; the force bits are those of the solving pass.
template <int BLOCK, int WAVES, bool FORCE_ONLY = false>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(WAVES))) void k_solve_fast(KParams P, float* __restrict__ q, int batch_base, int nb,
                                                      float dt, int mode, float* __restrict__ g_out) {
    constexpr int MAXV = BLOCK / kComp;
    constexpr int ROWS = BLOCK + 64;  // static items (<= BLOCK) + contacts evaluated in parallel
    constexpr int BLK = kBatchHdr + 4 * MAXV;
    __shared__ int s_row[MAXV];
    __shared__ int s_nrow[MAXV];
    __shared__ int s_nspr[MAXV];
    __shared__ int s_cb[MAXV + 1];
    __shared__ int s_so[MAXV + 1];
    __shared__ __align__(16) float s_rows[ROWS * kComp];
    __shared__ float s_acc[MAXV * kComp];

    const int t = threadIdx.x;
    const int bi = batch_base + xcd_remap(blockIdx.x, nb);
    const int b = P.batch_list ? P.batch_list[bi] : bi;
    // ---- round trip 1
    const int* blk = P.batch_blk + (size_t)b * BLK;
    const int4 b0 = reinterpret_cast<const int4*>(blk)[0];
    const int4 b1 = reinterpret_cast<const int4*>(blk)[1];
    const int4 b2 = reinterpret_cast<const int4*>(blk)[2];
    const int2 ci = reinterpret_cast<const int2*>(P.bat_cinfo)[b];
    int4 vr = make_int4(0, 0, 0, 0);
    if (t < MAXV) vr = reinterpret_cast<const int4*>(blk + kBatchHdr)[t];
    int cbt = 0;
    if (t <= MAXV) cbt = P.bat_coff[(size_t)b * (MAXV + 1) + t];
    const int nv = b0.y, tri_b = b0.z, ntri = b0.w, edge_b = b1.x, nedge = b1.y, tet_b = b1.z, ntet = b1.w;
    // frozen worlds: a batch none of whose vertices is live ends here (uniformly)
    bool own_live = true;
    if (P.w_active) {
        int live = 0;
        if (t < nv) live = own_live = P.w_active[P.p_world[vr.x]] != 0;
        if (!__syncthreads_or(live)) return;
    }
    const int spr_b = b2.x, nspr = b2.y;
    const int nstatic = ntri + nedge + ntet + nspr;
    const int cbeg = ci.x, ncont = ci.y;
    const bool par_contacts = !P.crows && nstatic + ncont <= ROWS;
    const float dtr = 1.0f / (dt * dt);

    // ---- round trip 2: item record of this lane (items 0..255), own-vertex state
    int kind = 0;  // 1 triangle, 2 edge, 3 tet, 4 contact, 5 spring
    int4 r0 = make_int4(0, 0, 0, 0), r1 = r0, r2 = r0, r3 = r0;
    int meta = 0;
    // first lane of each item range: packed, or each range on a wave boundary (a wave then runs
    // one element kind only)
    const int L1 = P.aligned ? wave_align(ntri) : ntri;
    const int L2 = L1 + (P.aligned ? wave_align(nedge) : nedge);
    const int L3 = L2 + (P.aligned ? wave_align(ntet) : ntet);
    const int L4 = L3 + (P.aligned ? wave_align(nspr) : nspr);
    if (t < ntri) {
        kind = 1;
        const int4* r4 = reinterpret_cast<const int4*>(P.tri_rec) + 3 * (size_t)(tri_b + t);
        r0 = r4[0];
        r1 = r4[1];
        r2 = r4[2];
    } else if (t >= L1 && t < L1 + nedge) {
        kind = 2;
        const int k = t - L1;
        const int4* r4 = reinterpret_cast<const int4*>(P.edge_rec) + 2 * (size_t)(edge_b + k);
        r0 = r4[0];
        r1 = r4[1];
        meta = P.edge_meta[edge_b + k];
    } else if (t >= L2 && t < L2 + ntet) {
        kind = 3;
        const int k = t - L2;
        const int4* r4 = reinterpret_cast<const int4*>(P.tet_rec) + 4 * (size_t)(tet_b + k);
        r0 = r4[0];
        r1 = r4[1];
        r2 = r4[2];
        r3 = r4[3];
        meta = P.tet_meta[tet_b + k];
    } else if (t >= L3 && t < L3 + nspr) {
        kind = 5;
        const int4* r4 = reinterpret_cast<const int4*>(P.spr_rec) + 2 * (size_t)(spr_b + t - L3);
        r0 = r4[0];
        r1 = r4[1];
    } else if (par_contacts && t >= L4 && t - L4 < ncont) {
        kind = 4;
        const int4* c4 = reinterpret_cast<const int4*>(P.cpack) + 4 * (size_t)(cbeg + t - L4);
        r0 = c4[0];
        r1 = c4[1];
        r2 = c4[2];
        r3 = c4[3];
    }
    V3 own_x0 = v3(0.0f, 0.0f, 0.0f), own_d = own_x0;
    const int own_v = vr.x;
    const bool own_active = ((vr.w >> 16) & 1) != 0;
    if (t < nv) {
        const float m = P.mass[own_v];
        float* a = s_acc + t * kComp;
#pragma unroll
        for (int c = 0; c < 3; ++c) a[c] = (m * (P.inertia[3 * own_v + c] - q[3 * own_v + c])) * dtr;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) a[3 + 3 * r + c] = (m * dtr) * (r == c ? 1.0f : 0.0f);
        own_x0 = ld3(P.q_prev, own_v);
        own_d = ld3(P.disp, own_v);
        s_row[t] = vr.y;
        s_nrow[t] = vr.z;
        s_nspr[t] = vr.w >> 17;
    }
    if (t <= nv) {
        s_cb[t] = cbt;
        // self-contact row bounds of the batch's color positions (consecutive positions)
        if (P.has_self) s_so[t] = P.so_off[b0.x + t];
    }

    // ---- round trip 3 + phase a: evaluate the lane's item
    if (kind != 0) {
        V3 fe = v3(0.0f, 0.0f, 0.0f);
        M3 he = mzero();
        int row;
        if (kind == 1) {
            float4 pz = __builtin_bit_cast(float4, r1);
            float4 mt = __builtin_bit_cast(float4, r2);
            row = r0.w & 0xffff;
            float pose[4] = {pz.x, pz.y, pz.z, pz.w};
            eval_membrane(ld3(q, r0.x), ld3(q, r0.y), ld3(q, r0.z), P.q_prev, r0.x, r0.y, r0.z, pose, mt.x, mt.y,
                          mt.z, mt.w, dt, r0.w >> 16, fe, he);
        } else if (kind == 2) {
            float4 pr = __builtin_bit_cast(float4, r1);
            row = meta & 0xffff;
            eval_bending(q, P.q_prev, r0.x, r0.y, r0.z, r0.w, pr.x, pr.y, pr.z, pr.w, dt, meta >> 16, fe, he);
        } else if (kind == 3) {
            float4 d0 = __builtin_bit_cast(float4, r1);
            float4 d1 = __builtin_bit_cast(float4, r2);
            float4 d2 = __builtin_bit_cast(float4, r3);
            int ti[4] = {r0.x, r0.y, r0.z, r0.w};
            float dm[9] = {d0.x, d0.y, d0.z, d0.w, d1.x, d1.y, d1.z, d1.w, d2.x};
            row = meta & 0xffff;
            eval_tet(q, P.q_prev, ti, dm, d2.y, d2.z, d2.w, dt, meta >> 16, fe, he);
        } else if (kind == 5) {
            const float4 pr = __builtin_bit_cast(float4, r1);
            row = r0.z & 0xffff;
            spring_item(P, q, r0.x, r0.y, r0.z >> 16, r0.w, pr.x, pr.y, pr.z, dt, fe, he);
        } else {
            eval_contact_rec(P, q, cbeg + t - L4, __builtin_bit_cast(float4, r0), __builtin_bit_cast(float4, r1),
                             __builtin_bit_cast(float4, r2), __builtin_bit_cast(float4, r3), dt, fe, he);
            row = nstatic + (t - L4);  // contact rows follow the static rows
        }
        if constexpr (FORCE_ONLY)
            store_force(s_rows + row * kComp, fe);
        else
            store_row(s_rows + row * kComp, fe, he);
    }
    // contacts beyond the lanes left after the static items (only when L4 + ncont > BLOCK)
    const int npar = par_contacts ? ncont : 0;
    for (int ci = BLOCK - L4 + t; ci < npar; ci += BLOCK) {
        V3 fe;
        M3 he;
        eval_packed_contact(P, q, cbeg + ci, dt, fe, he);
        if constexpr (FORCE_ONLY)
            store_force(s_rows + (nstatic + ci) * kComp, fe);
        else
            store_row(s_rows + (nstatic + ci) * kComp, fe, he);
    }
    __syncthreads();

    // ---- phase b
    const int cs = t / kComp;
    const int cc = t - cs * kComp;
    if (cs < nv && (!FORCE_ONLY || cc < 3)) {
        float acc = s_acc[cs * kComp + cc];
        const int q0 = s_row[cs], q1 = q0 + s_nrow[cs];
        // the LDS reads of four rows are issued together; the additions stay serial, in row order
        int r = q0;
        for (; r + 4 <= q1; r += 4) {
            const float a0 = s_rows[r * kComp + cc], a1 = s_rows[(r + 1) * kComp + cc];
            const float a2 = s_rows[(r + 2) * kComp + cc], a3 = s_rows[(r + 3) * kComp + cc];
            acc = acc + a0;
            acc = acc + a1;
            acc = acc + a2;
            acc = acc + a3;
        }
        for (; r < q1; ++r) acc = acc + s_rows[r * kComp + cc];
        // contacts from zero -- their LDS rows, or, when they did not fit, evaluated here by each of
        // the vertex's 12 lanes (component cc kept) -- then the vertex's spring rows and
        // self-contact rows continue that sum, which is added to the element sum
        const int c0 = s_cb[cs], c1 = s_cb[cs + 1];
        float ca = 0.0f;
        if (par_contacts) {
            for (int r = c0; r < c1; ++r) ca = ca + s_rows[(nstatic + r) * kComp + cc];
        } else if (P.crows) {
            ca = add_rows_in_order(ca, P.crows + (size_t)kComp * cbeg, c0, c1, cc);
        } else {
            for (int r = c0; r < c1; ++r) {
                V3 fc;
                M3 hc;
                eval_packed_contact(P, q, cbeg + r, dt, fc, hc);
                ca = ca + row_component(fc, hc, cc);
            }
        }
        for (int r = q1, re = q1 + s_nspr[cs]; r < re; ++r) ca = ca + s_rows[r * kComp + cc];
        if (P.has_self) ca = add_rows_in_order(ca, P.so_rows, s_so[cs], s_so[cs + 1], cc);
        acc = acc + ca;
        s_acc[cs * kComp + cc] = acc;
    }
    __syncthreads();

    // ---- phase c
    if (t < nv) {
        const float* a = s_acc + t * kComp;
        V3 f = v3(a[0], a[1], a[2]);
        M3 h;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) h.m[i][j] = a[3 + 3 * i + j];
        if constexpr (FORCE_ONLY) {
            if (own_live) st3(g_out, own_v, own_active ? f : v3(0.0f, 0.0f, 0.0f));
        } else {
            if (own_live) finish_vertex(P, q, own_v, own_active, own_x0, own_d, f, h, dt, mode, g_out);
        }
    }
}

// Generic color pass for vertices with more than 256 element items: one
// vertex per work-group, the CSR walked in chunks of 256 items, same
// canonical summation order; contacts from the per-particle lists.
__global__ __launch_bounds__(kBlock) void k_solve_generic(KParams P, float* __restrict__ q, int slow_base, float dt,
                                                          int mode, float* __restrict__ g_out) {
    __shared__ int s_cnt[3];
    __shared__ int s_beg[3];
    __shared__ __align__(16) float s_item[kBlock * kComp];
    __shared__ float s_acc[kComp];

    const int t = threadIdx.x;
    const int v = P.color_particles[P.slow_pos[slow_base + blockIdx.x]];
    if (P.w_active && P.w_active[P.p_world[v]] == 0) return;  // frozen world (uniform)
    const bool active = particle_active(P, v);
    const float dtr = 1.0f / (dt * dt);
    if (t == 0) {
        int nt = 0, ne = 0, nk = 0, bt = 0, be = 0, bk = 0;
        if (active) {
            if (P.has_tri) {
                bt = P.adj_tri_off[v];
                nt = (P.adj_tri_off[v + 1] - bt) >> 1;
            }
            if (P.has_edge) {
                be = P.adj_edge_off[v];
                ne = (P.adj_edge_off[v + 1] - be) >> 1;
            }
            if (P.has_tet) {
                bk = P.adj_tet_off[v];
                nk = (P.adj_tet_off[v + 1] - bk) >> 1;
            }
        }
        s_cnt[0] = nt; s_cnt[1] = ne; s_cnt[2] = nk;
        s_beg[0] = bt; s_beg[1] = be; s_beg[2] = bk;
    }
    __syncthreads();
    const int total = s_cnt[0] + s_cnt[1] + s_cnt[2];
    float acc = 0.0f;
    if (t < kComp) {
        float m = P.mass[v];
        if (t < 3) {
            acc = (m * (P.inertia[3 * v + t] - q[3 * v + t])) * dtr;
        } else {
            int r = (t - 3) / 3, c = (t - 3) % 3;
            acc = (m * dtr) * (r == c ? 1.0f : 0.0f);
        }
    }
    for (int lo = 0; lo < total; lo += kBlock) {
        const int k = lo + t;
        if (k < total) {
            V3 fe = v3(0.0f, 0.0f, 0.0f);
            M3 he = mzero();
            if (k < s_cnt[0]) {
                int base = s_beg[0] + 2 * k;
                int tri = P.adj_tri[base], slot = P.adj_tri[base + 1];
                const float* mat = P.tri_mat + 5 * tri;
                if (mat[0] > 0.0f || mat[1] > 0.0f) {
                    const int* ti = P.tri_idx + 3 * tri;
                    eval_membrane(ld3(q, ti[0]), ld3(q, ti[1]), ld3(q, ti[2]), P.q_prev, ti[0], ti[1], ti[2],
                                  P.tri_pose + 4 * tri, P.tri_area[tri], mat[0], mat[1], mat[2], dt, slot, fe, he);
                }
            } else if (k < s_cnt[0] + s_cnt[1]) {
                int base = s_beg[1] + 2 * (k - s_cnt[0]);
                int ed = P.adj_edge[base], slot = P.adj_edge[base + 1];
                float ke = P.edge_bend[2 * ed];
                if (ke > 0.0f) {
                    const int* ei = P.edge_idx + 4 * ed;
                    eval_bending(q, P.q_prev, ei[0], ei[1], ei[2], ei[3], P.edge_rest_angle[ed], P.edge_rest_len[ed],
                                 ke, P.edge_bend[2 * ed + 1], dt, slot, fe, he);
                }
            } else {
                int base = s_beg[2] + 2 * (k - s_cnt[0] - s_cnt[1]);
                int tet = P.adj_tet[base], slot = P.adj_tet[base + 1];
                const float* mat = P.tet_mat + 3 * tet;
                if (mat[0] > 0.0f || mat[1] > 0.0f)
                    eval_tet(q, P.q_prev, P.tet_idx + 4 * tet, P.tet_pose + 9 * tet, mat[0], mat[1], mat[2], dt, slot,
                             fe, he);
            }
            store_row(s_item + t * kComp, fe, he);
        }
        __syncthreads();
        if (t < kComp) {
            int e = min(kBlock, total - lo);
            for (int i = 0; i < e; ++i) acc = acc + s_item[i * kComp + t];
        }
        __syncthreads();
    }
    if (t < kComp) s_acc[t] = acc;
    __syncthreads();
    if (t == 0) {
        V3 f = v3(s_acc[0], s_acc[1], s_acc[2]);
        M3 h;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) h.m[i][j] = s_acc[3 + 3 * i + j];
        if (active) {
            V3 cf = v3(0.0f, 0.0f, 0.0f);
            M3 ch = mzero();
            // the vertex's contacts as packed at its color position (contact-index order)
            const int vpos = P.slow_pos[slow_base + blockIdx.x];
            const int c0 = P.pos_coff[vpos], c1 = P.pos_coff[vpos + 1];
            for (int j = c0; j < c1; ++j) {
                V3 fc;
                M3 hc;
                eval_packed_contact(P, q, j, dt, fc, hc);
                cf = add(cf, fc);
                ch = madd(ch, hc);
            }
            if (P.has_spring) {
                const int si = slow_base + blockIdx.x;
                for (int j = P.slow_spr_off[si]; j < P.slow_spr_off[si + 1]; ++j) {
                    const int4 it = reinterpret_cast<const int4*>(P.slow_spr)[j];
                    const int sp = it.x;
                    V3 fs;
                    M3 hs;
                    spring_item(P, q, P.spring_idx[2 * sp], P.spring_idx[2 * sp + 1], it.y, it.z, P.spring_rest[sp],
                                P.spring_ke[sp], P.spring_kd[sp], dt, fs, hs);
                    cf = add(cf, fs);
                    ch = madd(ch, hs);
                }
            }
            if (P.has_self) {
                const int pos = P.slow_pos[slow_base + blockIdx.x];
                for (int j = P.so_off[pos], je = P.so_off[pos + 1]; j < je; ++j) {
                    const float* r = P.so_rows + (size_t)kComp * j;
                    cf = add(cf, v3(r[0], r[1], r[2]));
#pragma unroll
                    for (int a = 0; a < 3; ++a)
#pragma unroll
                        for (int b = 0; b < 3; ++b) ch.m[a][b] = ch.m[a][b] + r[3 + 3 * a + b];
                }
            }
            h = madd(h, ch);
            f = add(f, cf);
        }
        finish_vertex(P, q, v, active, ld3(P.q_prev, v), ld3(P.disp, v), f, h, dt, mode, g_out);
    }
}

// forward_step + apply_truncation_ts(t = 1): particle_vbd_kernels.py:1350-1375, :2266-2286
__global__ void k_forward(KParams P, float* __restrict__ q, const float* __restrict__ qd, const float* __restrict__ f_ext,
                          float dt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.N) return;
    V3 x = ld3(q, i);
    st3(P.q_prev, i, x);
    V3 y, d;
    if (!(P.flags[i] & 1) || P.inv_mass[i] == 0.0f) {
        y = x;
        d = v3(0.0f, 0.0f, 0.0f);
    } else {
        V3 g = v3(P.gravity[0], P.gravity[1], P.gravity[2]);
        V3 fe = f_ext ? ld3(f_ext, i) : v3(0.0f, 0.0f, 0.0f);
        V3 vnew = add(ld3(qd, i), scl(add(g, scl(fe, P.inv_mass[i])), dt));
        y = add(x, scl(vnew, dt));
        d = scl(vnew, dt);
    }
    st3(P.inertia, i, y);
    st3(P.disp, i, d);
    st3(q, i, add(x, scl(d, 1.0f)));
}

// update_velocity: particle_vbd_kernels.py:1468-1471 (plus the final q copy, solver_vbd.py:2704)
__global__ void k_finalize(KParams P, const float* __restrict__ q, float* __restrict__ q_out, float* __restrict__ qd_out,
                           float dt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.N) return;
    V3 x = ld3(q, i);
    if (q_out != q) st3(q_out, i, x);
    st3(qd_out, i, divs(sub(x, ld3(P.q_prev, i)), dt));
}

// SolverBase.integrate_particles: semi-implicit Euler with a speed limit (solver.py:20-56).  Inactive
// particles keep x and leave v_new unwritten; gravity is that of the particle's world (-1: world 0);
// world ids at or beyond the gravity table read its last row instead of memory outside it.
__global__ void k_integrate_particles(int n, const float* __restrict__ x, const float* __restrict__ v,
                                      const float* __restrict__ f, const float* __restrict__ w,
                                      const int* __restrict__ flags, const int* __restrict__ world,
                                      const float* __restrict__ gravity, int gravity_count, float dt, float v_max,
                                      float* __restrict__ x_new, float* __restrict__ v_new) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    V3 x0 = ld3(x, i);
    if (!(flags[i] & 1)) {
        st3(x_new, i, x0);
        return;
    }
    const float inv_mass = w[i];
    int wi = world ? world[i] : 0;
    wi = wi < 0 ? 0 : (wi >= gravity_count ? gravity_count - 1 : wi);
    const V3 g = ld3(gravity, wi);
    const float has_mass = -inv_mass < 0.0f ? 1.0f : 0.0f;  // wp.step(-inv_mass)
    V3 v1 = add(ld3(v, i), scl(add(scl(ld3(f, i), inv_mass), scl(g, has_mass)), dt));
    const float mag = length(v1);
    if (mag > v_max) v1 = scl(v1, v_max / mag);
    st3(x_new, i, add(x0, scl(v1, dt)));
    st3(v_new, i, v1);
}

// ---- rigid transforms (p, q xyzw): warp's transform_point / transform_vector / transform_inverse /
// transform_multiply and the Hamilton quaternion product, from their textbook definitions
__device__ __forceinline__ V3 quat_rotate(float qx, float qy, float qz, float qw, V3 x) {
    float c = 2.0f * qw * qw - 1.0f;
    float d = 2.0f * (qx * x.x + qy * x.y + qz * x.z);
    return v3(x.x * c + qx * d + (qy * x.z - qz * x.y) * qw * 2.0f, x.y * c + qy * d + (qz * x.x - qx * x.z) * qw * 2.0f,
              x.z * c + qz * d + (qx * x.y - qy * x.x) * qw * 2.0f);
}

struct Xf {
    V3 p;
    float qx, qy, qz, qw;
};

__device__ __forceinline__ Xf xf_load(const float* __restrict__ t) {
    return Xf{v3(t[0], t[1], t[2]), t[3], t[4], t[5], t[6]};
}
__device__ __forceinline__ V3 xf_point(const Xf& t, V3 x) { return add(quat_rotate(t.qx, t.qy, t.qz, t.qw, x), t.p); }
__device__ __forceinline__ V3 xf_vector(const Xf& t, V3 x) { return quat_rotate(t.qx, t.qy, t.qz, t.qw, x); }
__device__ __forceinline__ Xf xf_inverse(const Xf& t) {
    Xf r;
    r.qx = -t.qx;
    r.qy = -t.qy;
    r.qz = -t.qz;
    r.qw = t.qw;
    r.p = neg(quat_rotate(r.qx, r.qy, r.qz, r.qw, t.p));
    return r;
}
// transform_multiply(a, b) = (quat_rotate(a.q, b.p) + a.p, a.q * b.q)
__device__ __forceinline__ Xf xf_mul(const Xf& a, const Xf& b) {
    Xf r;
    r.p = add(quat_rotate(a.qx, a.qy, a.qz, a.qw, b.p), a.p);
    r.qx = a.qw * b.qx + b.qw * a.qx + a.qy * b.qz - b.qy * a.qz;
    r.qy = a.qw * b.qy + b.qw * a.qy + a.qz * b.qx - b.qz * a.qx;
    r.qz = a.qw * b.qz + b.qw * a.qz + a.qx * b.qy - b.qx * a.qy;
    r.qw = a.qw * b.qw - a.qx * b.qx - a.qy * b.qy - a.qz * b.qz;
    return r;
}

// World contact point bx = X_wb * body_pos of a soft contact and the body's velocity there: for a shape
// attached to a body moved by an external rigid solver (rigid_vbd_kernels.py:1365-1439 with body_q_prev,
// solver_vbd.py:2523-2526) bv = (bx - X_wb_prev * body_pos) / dt + X_wb . body_vel; a static shape takes
// the identity transform, bx = body_pos, bv = (0 + cross(0, bx)) + body_vel (the same values up to signed
// zeros).
struct BodyPoses {
    const int* shape_body;  // null: every shape static
    const float* body_q;
    const float* body_q_prev;
    int body_count;
    int shape_count;
};

__device__ __forceinline__ void contact_body_frame(const BodyPoses& B, int shape, V3 body_pos, V3 body_vel, float dt,
                                                   V3& bx, V3& bv) {
    const int b = (B.shape_body && shape >= 0 && shape < B.shape_count) ? B.shape_body[shape] : -1;
    if (b < 0 || !B.body_q || !B.body_q_prev || b >= B.body_count) {
        const V3 zero = v3(0.0f, 0.0f, 0.0f);
        bx = body_pos;
        bv = add(add(zero, cross(zero, bx)), body_vel);
        return;
    }
    const Xf X = xf_load(B.body_q + 7 * (size_t)b);
    const Xf Xp = xf_load(B.body_q_prev + 7 * (size_t)b);
    bx = xf_point(X, body_pos);
    const V3 bxp = xf_point(Xp, body_pos);
    bv = add(divs(sub(bx, bxp), dt), xf_vector(X, body_vel));
}

__global__ void k_fill_i32(int* __restrict__ p, int value, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

// up to kFillMax int arrays filled by one launch (grid-stride over each)
constexpr int kFillMax = 6;
struct FillList {
    int* p[kFillMax];
    int value[kFillMax];
    int n[kFillMax];
    int count;
};
__global__ void k_fill_many(FillList F) {
    for (int k = 0; k < F.count; ++k)
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < F.n[k]; i += gridDim.x * blockDim.x)
            F.p[k][i] = F.value[k];
}

// init_body_particle_contacts (rigid_vbd_kernels.py:3276-3321) + counts per color position: one entry
// for a particle record's particle, one for each corner of an edge/face record (when ef is set; without
// barycentric weights such records are skipped).  pos_cnt must be zero on entry; k_contact_place leaves it
// zero again.
__global__ void k_contact_init_pos(const int* __restrict__ count_ptr, int cmax, const int* __restrict__ c_indices,
                                   const int* __restrict__ c_shape, float soft_ke, float soft_kd, float soft_mu,
                                   const float* __restrict__ shape_ke, const float* __restrict__ shape_kd,
                                   const float* __restrict__ shape_mu, float k_start, float* __restrict__ pk,
                                   float* __restrict__ mke, float* __restrict__ mkd, float* __restrict__ mmu,
                                   const int* __restrict__ pos_of, int* __restrict__ pos_cnt, int n_particles,
                                   int n_shapes, int ef) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int nc = min(*count_ptr, cmax);
    if (i >= nc) return;
    int s = c_shape[i];
    if (s < 0 || s >= n_shapes) {
        // a record naming no shape of this model contributes nothing (it is never gathered)
        pk[i] = 0.0f;
        mke[i] = 0.0f;
        mkd[i] = 0.0f;
        mmu[i] = 0.0f;
        return;
    }
    float ke = 0.5f * (soft_ke + shape_ke[s]);
    float kd = 0.5f * (soft_kd + shape_kd[s]);
    float mu = sqrtf(soft_mu * shape_mu[s]);
    mke[i] = ke;
    mkd[i] = kd;
    mmu[i] = mu;
    pk[i] = k_start < 0.0f ? ke : fminf(k_start, ke);
    const int* cr = c_indices + 3 * i;
    if (cr[1] < 0) {
        if (cr[0] >= 0 && cr[0] < n_particles) atomicAdd(&pos_cnt[pos_of[cr[0]]], 1);
    } else if (ef) {
        for (int k = 0; k < 3; ++k)
            if (cr[k] >= 0 && cr[k] < n_particles) atomicAdd(&pos_cnt[pos_of[cr[k]]], 1);
    }
}

// Places every counted entry, key (contact id << 2 | corner), in its color position's range (any order
// within the range) by counting the position's counter back down to zero, ready for the next substep.
__global__ void k_contact_place(const int* __restrict__ count_ptr, int cmax, const int* __restrict__ c_indices,
                                const int* __restrict__ c_shape, const int* __restrict__ pos_of,
                                const int* __restrict__ pos_coff, int* __restrict__ pos_cnt, int* __restrict__ ids,
                                int n_particles, int n_shapes, int ef) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int nc = min(*count_ptr, cmax);
    if (i >= nc) return;
    int sh = c_shape[i];
    if (sh < 0 || sh >= n_shapes) return;
    const int* cr = c_indices + 3 * i;
    const int corners = cr[1] < 0 ? 1 : (ef ? 3 : 0);
    for (int k = 0; k < corners; ++k) {
        const int p = cr[k];
        if (p < 0 || p >= n_particles) continue;
        const int pos = pos_of[p];
        const int slot = atomicSub(&pos_cnt[pos], 1) - 1;
        ids[pos_coff[pos] + slot] = (i << 2) | k;
    }
}

// k_contact_sort sorts each color position's entry keys ascending (contact-index order: the serial
// accumulation order; lists of up to kSortCols keys in the lane's LDS column) and records each entry's
// position; k_contact_pack then packs every entry in place, so a batch's contacts are one contiguous,
// vertex-ordered range.  The body point and its velocity are resolved there once per substep (the
// body poses are fixed during a step).
constexpr int kSortCols = 32;
constexpr int kPackGrid = 1024;  // work-groups of the pack's grid-stride loop
struct PackArgs {
    const float* pk;
    const float* mkd;
    const float* mmu;
    const int* c_indices;
    const float* c_bary;
    const int* c_shape;
    const float* body_pos;
    const float* body_vel;
    const float* normal;
    const float* shape_margin;
    int shape_count;
    const float* radius;
    BodyPoses bodies;
    float dt;
};

__global__ __launch_bounds__(256) void k_contact_sort(int n, const int* __restrict__ pos_coff, int* __restrict__ ids,
                                                      int* __restrict__ ent_pos) {
    __shared__ int s_ids[kSortCols * 256];  // element i of lane l at i * 256 + l
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int b = pos_coff[k], e = pos_coff[k + 1];
    for (int j = b; j < e; ++j) ent_pos[j] = k;
    if (e - b <= 1) return;
    if (e - b <= kSortCols) {
        int* L = s_ids + threadIdx.x;
        const int m = e - b;
        for (int i = 0; i < m; ++i) L[i * 256] = ids[b + i];
        for (int i = 1; i < m; ++i) {
            const int key = L[i * 256];
            int j = i - 1;
            while (j >= 0 && L[j * 256] > key) {
                L[(j + 1) * 256] = L[j * 256];
                --j;
            }
            L[(j + 1) * 256] = key;
        }
        for (int i = 0; i < m; ++i) ids[b + i] = L[i * 256];
        return;
    }
    for (int i = b + 1; i < e; ++i) {
        int key = ids[i];
        int j = i - 1;
        while (j >= b && ids[j] > key) {
            ids[j + 1] = ids[j];
            --j;
        }
        ids[j + 1] = key;
    }
}

// one packed record per entry (grid-stride over the device-sized entry count pos_coff[n]); an entry's
// loads (contact, shape, body pose, corners) no longer wait on the entries before it at its position
__global__ __launch_bounds__(256) void k_contact_pack(int n, const int* __restrict__ color_particles,
                                                      const int* __restrict__ vert_rec,
                                                      const int* __restrict__ pos_coff, const int* __restrict__ ids,
                                                      const int* __restrict__ ent_pos, PackArgs A,
                                                      float* __restrict__ cpack, int* __restrict__ cpack_ef,
                                                      int* __restrict__ cid_to_pack) {
    const int total = pos_coff[n];
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < total; j += gridDim.x * blockDim.x) {
        const int k = ent_pos[j];
        const int v = color_particles[k];
        const int local = vert_rec[4 * k + 3] & 0xffff;
        const int key = ids[j];
        const int cid = key >> 2, corner = key & 3;
        const int sh = A.c_shape[cid];
        const int* cr = A.c_indices + 3 * cid;
        const bool ef = cr[1] >= 0;
        float rad = A.radius[v];
        if (ef) {
            rad = A.radius[cr[0]];
            if (cr[1] >= 0) rad = fmaxf(rad, A.radius[cr[1]]);
            if (cr[2] >= 0) rad = fmaxf(rad, A.radius[cr[2]]);
            int4* o = reinterpret_cast<int4*>(cpack_ef) + 2 * (size_t)j;
            o[0] = make_int4(cr[0], cr[1], cr[2], corner);
            o[1] = make_int4(__float_as_int(A.c_bary[3 * cid]), __float_as_int(A.c_bary[3 * cid + 1]),
                             __float_as_int(A.c_bary[3 * cid + 2]), 0);
        }
        V3 bx, bv;
        contact_body_frame(A.bodies, sh, ld3(A.body_pos, cid), ld3(A.body_vel, cid), A.dt, bx, bv);
        const V3 nrm = ld3(A.normal, cid);
        float4* o = reinterpret_cast<float4*>(cpack) + 4 * (size_t)j;
        o[0] = make_float4(__int_as_float(v), __int_as_float(local | (ef ? kContactEF : 0)), A.pk[cid], A.mkd[cid]);
        o[1] = make_float4(A.mmu[cid], A.shape_count > 0 ? A.shape_margin[sh] : 0.0f, rad, bx.x);
        o[2] = make_float4(bx.y, bx.z, bv.x, bv.y);
        o[3] = make_float4(bv.z, nrm.x, nrm.y, nrm.z);
        if (cid_to_pack) cid_to_pack[3 * (size_t)cid + corner] = j;
    }
}

// Contact range of every batch from the color-position offsets: the first packed
// contact and count (cinfo) and the max_v + 1 vertex offsets relative to the first
// (coff), so the color pass reads them indexed by batch in its first round trip.
__global__ void k_batch_contacts(int nbatch, const int* __restrict__ blk, int blk_words, int max_v,
                                 const int* __restrict__ pos_coff, int* __restrict__ cinfo, int* __restrict__ coff) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbatch) return;
    const int* h = blk + (size_t)b * blk_words;
    const int vb = h[0], nv = h[1];
    const int c0 = pos_coff[vb];
    for (int k = 0; k <= nv; ++k) coff[(size_t)b * (max_v + 1) + k] = pos_coff[vb + k] - c0;
    cinfo[2 * b] = c0;
    cinfo[2 * b + 1] = pos_coff[vb + nv] - c0;
}

// update_duals_body_particle_contacts, rigid_vbd_kernels.py:4861-4926: the penetration of the particle,
// or of an edge/face record's barycentric point with its largest corner radius (records skipped without
// barycentric weights), at the world body point; the ramped penalty goes to every packed entry.
__global__ void k_dual_update(const int* __restrict__ count_ptr, int cmax, const int* __restrict__ c_indices,
                              const float* __restrict__ c_bary, const int* __restrict__ c_shape,
                              const float* __restrict__ body_pos, const float* __restrict__ body_vel,
                              const float* __restrict__ normal, const float* __restrict__ q,
                              const float* __restrict__ radius, const float* __restrict__ shape_margin, int shape_count,
                              BodyPoses bodies, float dt, const float* __restrict__ mke, float beta,
                              float* __restrict__ pk, int n_particles, const int* __restrict__ cid_to_pack,
                              float* __restrict__ cpack) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int nc = min(*count_ptr, cmax);
    if (i >= nc) return;
    const int* cr = c_indices + 3 * i;
    const bool ef = cr[1] >= 0;
    if (ef && !c_bary) return;
    int p = cr[0];
    if (p < 0 || p >= n_particles) return;
    int s = c_shape[i];
    if (s < 0 || s >= shape_count) return;
    float margin = shape_count > 0 ? shape_margin[s] : 0.0f;
    V3 n = ld3(normal, i);
    V3 bx, bv;
    contact_body_frame(bodies, s, ld3(body_pos, i), ld3(body_vel, i), dt, bx, bv);
    V3 x;
    float rad;
    if (!ef) {
        x = ld3(q, p);
        rad = radius[p];
    } else {
        x = scl(ld3(q, cr[0]), c_bary[3 * i]);
        rad = radius[cr[0]];
        if (cr[1] >= 0) {
            x = add(x, scl(ld3(q, cr[1]), c_bary[3 * i + 1]));
            rad = fmaxf(rad, radius[cr[1]]);
        }
        if (cr[2] >= 0) {
            x = add(x, scl(ld3(q, cr[2]), c_bary[3 * i + 2]));
            rad = fmaxf(rad, radius[cr[2]]);
        }
    }
    float pen = -(dot(n, sub(x, bx)) - rad - margin);
    pen = fmaxf(0.0f, pen);
    float k = fminf(pk[i] + beta * pen, mke[i]);
    pk[i] = k;
    for (int c = 0; c < 3; ++c) {
        int slot = cid_to_pack[3 * (size_t)i + c];
        if (slot >= 0) cpack[16 * (size_t)slot + 2] = k;
    }
}

// Squared residual norm in two deterministic stages: per-workgroup partial
// sums (fixed shuffle tree), then one workgroup folds the partials in a fixed
// order.  Same input => same bits, so iterate-to-tolerance loops stop at the
// same iteration on every run.
constexpr int kNormBlock = 256;
constexpr int kNormItems = 4;  // particles per thread in the first stage

__device__ inline double block_sum_f64(double v, double* s_part) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) s_part[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += s_part[w];
    return s;
}

__global__ __launch_bounds__(kNormBlock) void k_residual_partial(int n, const float* __restrict__ g,
                                                                double* __restrict__ partial) {
    __shared__ double s_part[kNormBlock / 64];
    double v = 0.0;
    long base = (long)blockIdx.x * kNormBlock * kNormItems + threadIdx.x;
    for (int k = 0; k < kNormItems; ++k) {
        long i = base + (long)k * kNormBlock;
        if (i < n) {
            double a = g[3 * i], b = g[3 * i + 1], c = g[3 * i + 2];
            v += a * a + b * b + c * c;
        }
    }
    double s = block_sum_f64(v, s_part);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(kNormBlock) void k_residual_final(int nparts, const double* __restrict__ partial,
                                                              double* __restrict__ out) {
    __shared__ double s_part[kNormBlock / 64];
    double v = 0.0;
    for (int i = threadIdx.x; i < nparts; i += kNormBlock) v += partial[i];
    double s = block_sum_f64(v, s_part);
    if (threadIdx.x == 0) out[0] = s;
}

// Per-world norms, first stage: `split` work-groups per world each sum |g_p|^2 over one contiguous
// slice of the world's particle range [off[w], off[w+1]) (clipped to [0, n); thread-strided in
// double, then the fixed shuffle tree) into part[w * split + k]; k_world_norms (or the batched
// driver's freezing kernel) adds a world's slices in order.  Splitting keeps a few dozen large
// worlds from leaving most CUs idle.  With live flags (the batched driver), a frozen world --
// whose iterate no longer changes -- keeps the norm it was frozen with and is not summed again.
__global__ __launch_bounds__(kNormBlock) void k_residual_worlds(int n, const int* __restrict__ off,
                                                               const float* __restrict__ g, int split,
                                                               double* __restrict__ part,
                                                               const int* __restrict__ live) {
    __shared__ double s_part[kNormBlock / 64];
    const int w = blockIdx.x / split, k = blockIdx.x - w * split;
    if (live && !live[w]) return;  // uniform
    const int p0 = max(off[w], 0), p1 = max(min(off[w + 1], n), p0);
    const int chunk = (p1 - p0 + split - 1) / split;
    const int q0 = min(p0 + k * chunk, p1), q1 = min(q0 + chunk, p1);
    double v = 0.0;
    for (int i = q0 + (int)threadIdx.x; i < q1; i += kNormBlock) {
        double a = g[3 * (size_t)i], b = g[3 * (size_t)i + 1], c = g[3 * (size_t)i + 2];
        v += a * a + b * b + c * c;
    }
    double sum = block_sum_f64(v, s_part);
    if (threadIdx.x == 0) part[blockIdx.x] = sum;
}

// a world's squared norm: its slices' sums added in slice order
__device__ __forceinline__ double world_norm(const double* __restrict__ part, int split, int w) {
    double s = 0.0;
    for (int k = 0; k < split; ++k) s += part[(size_t)w * split + k];
    return s;
}

__global__ void k_world_norms(int nw, const double* __restrict__ part, int split, double* __restrict__ out) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w < nw) out[w] = world_norm(part, split, w);
}

// Iterate from a caller's guess: for solved particles (ACTIVE, inv_mass != 0)
// d = guess - x0, else d = 0; q_dst = x0 + d (the truncation form forward_step
// uses).  commit != 0 also stores d as the solver's displacement (q_dst is then
// the solver's iterate); commit == 0 only fills a trial buffer.
__global__ void k_set_iterate(KParams P, float* __restrict__ q_dst, const float* __restrict__ guess, int commit,
                              int i0, int i1) {
    int i = i0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= i1) return;
    bool solved = (P.flags[i] & 1) && P.inv_mass[i] != 0.0f;
    V3 x0 = ld3(P.q_prev, i);
    V3 d = solved ? sub(ld3(guess, i), x0) : v3(0.0f, 0.0f, 0.0f);
    if (commit && solved) st3(P.disp, i, d);
    st3(q_dst, i, add(x0, scl(d, 1.0f)));
}

// ---- soft-contact generation (geometry/kernels.py:1020-1184) and the full-surface edge / face passes
// (geometry/soft_contacts_sdf.py:240-551), plane and sphere shapes, static or attached to a body -------
// One index space of candidate pairs: [0, P) (particle, shape), [P, P + E) (edge, shape),
// [P + E, P + E + F) (triangle, shape); records are written in that order.
struct ColParams {
    int P, E, F;
    const int* pairs;
    const int* epairs;
    const int* fpairs;
    const int* edge_idx;  // [edges, 4] (o0, o1, v0, v1)
    const int* tri_idx;   // [T, 3]
    const float* radius;
    const int* pflags;
    const int* pworld;
    const int* stype;
    const int* sflags;
    const int* sworld;
    const int* sbody;     // null: every shape static
    const float* sxf;
    const float* sscale;
    const float* smargin;
    const float* body_q;  // this call's body poses (null without bodies)
    int body_count;
    float4* ef_cache;     // [E + F][3] edge/face results of the count kernel, read by the write kernel
                          // (null: the write kernel evaluates them again)
};

struct ContactGeo {
    bool hit;
    int particle, shape;
    int i0, i1, i2;
    V3 bary;
    V3 body_pos, body_vel, normal;
};

// shape-local -> body (X_bs), -> world (X_ws) and world -> shape-local (X_sw); a static shape has
// X_ws = X_bs (transform_multiply with the identity gives X_bs up to signed zeros)
__device__ __forceinline__ void shape_frames(const ColParams& C, int s, Xf& X_bs, Xf& X_ws, Xf& X_sw) {
    X_bs = xf_load(C.sxf + 7 * (size_t)s);
    const int b = C.sbody ? C.sbody[s] : -1;
    X_ws = (b >= 0 && C.body_q && b < C.body_count) ? xf_mul(xf_load(C.body_q + 7 * (size_t)b), X_bs) : X_bs;
    X_sw = xf_inverse(X_ws);
}

// eval_shape_sdf (soft_contacts_sdf.py:61-96) for this path's analytic shapes: sphere |x| - r with the
// gradient x / |x| ((0, 0, 1) at the centre), infinite plane x.z with gradient +z
__device__ __forceinline__ float shape_sdf(int type, const float* __restrict__ sc, V3 x, V3& grad) {
    if (type == 3) {
        const float d = length(x) - sc[0];
        const float l = length(x);
        grad = l > 1.0e-8f ? divs(x, l) : v3(0.0f, 0.0f, 1.0f);
        return d;
    }
    grad = v3(0.0f, 0.0f, 1.0f);
    return x.z;
}

__device__ __forceinline__ V3 seg_point(V3 p, V3 q, float u) { return add(scl(p, 1.0f - u), scl(q, u)); }

// optimize_edge_sdf (soft_contacts_sdf.py:136-175): golden-section search along p -> q, fixed iterations
__device__ float opt_edge_sdf(int type, const float* __restrict__ sc, V3 p, V3 q, int n_iter, V3& x_out,
                              float& phi_out, V3& g_out) {
    const float inv_phi = 0.6180339887498949f;
    float lo = 0.0f, hi = 1.0f;
    float c = hi - (hi - lo) * inv_phi;
    float d = lo + (hi - lo) * inv_phi;
    V3 g;
    float fc = shape_sdf(type, sc, seg_point(p, q, c), g);
    float fd = shape_sdf(type, sc, seg_point(p, q, d), g);
    for (int i = 0; i < n_iter; ++i) {
        if (fc < fd) {
            hi = d;
            d = c;
            fd = fc;
            c = hi - (hi - lo) * inv_phi;
            fc = shape_sdf(type, sc, seg_point(p, q, c), g);
        } else {
            lo = c;
            c = d;
            fc = fd;
            d = lo + (hi - lo) * inv_phi;
            fd = shape_sdf(type, sc, seg_point(p, q, d), g);
        }
    }
    const float u = 0.5f * (lo + hi);
    x_out = seg_point(p, q, u);
    phi_out = shape_sdf(type, sc, x_out, g_out);
    return u;
}

__device__ __forceinline__ V3 bary_point(V3 a, V3 b, V3 c, V3 w) {
    return add(add(scl(a, w.x), scl(b, w.y)), scl(c, w.z));
}

// optimize_face_sdf (soft_contacts_sdf.py:178-219): Frank-Wolfe on the barycentric simplex from the
// centroid; each step line-searches toward the corner minimising grad . corner
__device__ V3 opt_face_sdf(int type, const float* __restrict__ sc, V3 a, V3 b, V3 c, int n_iter, int ls_iter,
                           V3& x_out, float& phi_out, V3& g_out) {
    const float third = 1.0f / 3.0f;
    V3 bary = v3(third, third, third);
    for (int i = 0; i < n_iter; ++i) {
        const V3 x = bary_point(a, b, c, bary);
        V3 g;
        shape_sdf(type, sc, x, g);
        const float da = dot(g, a), db = dot(g, b), dc = dot(g, c);
        V3 s = v3(1.0f, 0.0f, 0.0f);
        if (db <= da && db <= dc)
            s = v3(0.0f, 1.0f, 0.0f);
        else if (dc <= da && dc <= db)
            s = v3(0.0f, 0.0f, 1.0f);
        const V3 target = bary_point(a, b, c, s);
        V3 lx, lg;
        float lphi;
        const float gamma = opt_edge_sdf(type, sc, x, target, ls_iter, lx, lphi, lg);
        bary = add(scl(bary, 1.0f - gamma), scl(s, gamma));
    }
    x_out = bary_point(a, b, c, bary);
    phi_out = shape_sdf(type, sc, x_out, g_out);
    return bary;
}

constexpr int kSdfEdgeIters = 24;  // soft_contacts_sdf.py:43-45
constexpr int kSdfFaceIters = 24;
constexpr int kSdfLsIters = 16;

__device__ ContactGeo particle_contact_geometry(const ColParams& C, const float* __restrict__ q, float margin,
                                                int tid) {
    ContactGeo g;
    g.hit = false;
    int p = C.pairs[2 * tid], s = C.pairs[2 * tid + 1];
    g.particle = p;
    g.shape = s;
    g.i0 = p;
    g.i1 = -1;
    g.i2 = -1;
    g.bary = v3(1.0f, 0.0f, 0.0f);
    if ((C.pflags[p] & 1) == 0) return g;
    if ((C.sflags[s] & 4) == 0) return g;
    int pw = C.pworld[p], sw = C.sworld[s];
    if (pw != -1 && sw != -1 && pw != sw) return g;
    V3 px = ld3(q, p);
    float radius = C.radius[p];
    Xf X_bs, X_ws, X_sw;
    shape_frames(C, s, X_bs, X_ws, X_sw);
    V3 xl = xf_point(X_sw, px);
    const float* sc = C.sscale + 3 * s;
    float s_margin = C.smargin[s];
    float dd = 1.0e6f;
    V3 n = v3(0.0f, 0.0f, 0.0f);
    int type = C.stype[s];
    if (type == 3) {
        float pl = length(xl);
        dd = pl - sc[0];
        n = pl > 1.0e-8f ? divs(xl, pl) : v3(0.0f, 0.0f, 1.0f);
    } else if (type == 1) {
        float width = sc[0] * 0.5f, len = sc[1] * 0.5f;
        if (width > 0.0f && len > 0.0f) {
            float a = fmaxf(fabsf(xl.x) - width, fabsf(xl.y) - len);
            dd = fmaxf(a, fabsf(xl.z));
        } else {
            dd = xl.z;
        }
        n = v3(0.0f, 0.0f, 1.0f);
    } else {
        return g;
    }
    if (dd < margin + s_margin + radius) {
        g.hit = true;
        g.body_pos = xf_point(X_bs, sub(xl, scl(n, dd)));
        g.body_vel = xf_vector(X_bs, v3(0.0f, 0.0f, 0.0f));
        g.normal = xf_vector(X_ws, n);
    }
    return g;
}

// create_soft_edge_contacts (soft_contacts_sdf.py:380-469)
__device__ ContactGeo edge_contact_geometry(const ColParams& C, const float* __restrict__ q, float margin, int k) {
    ContactGeo g;
    g.hit = false;
    g.particle = -1;
    const int e = C.epairs[2 * k], s = C.epairs[2 * k + 1];
    g.shape = s;
    if ((C.sflags[s] & 4) == 0) return g;
    const int type = C.stype[s];
    if (type != 1 && type != 3) return g;
    const int v0 = C.edge_idx[4 * e + 2], v1 = C.edge_idx[4 * e + 3];
    const float radius = fmaxf(C.radius[v0], C.radius[v1]);
    Xf X_bs, X_ws, X_sw;
    shape_frames(C, s, X_bs, X_ws, X_sw);
    const V3 p_s = xf_point(X_sw, ld3(q, v0));
    const V3 q_s = xf_point(X_sw, ld3(q, v1));
    const float* sc = C.sscale + 3 * s;
    const float threshold = margin + C.smargin[s] + radius;
    V3 gr;
    const float phi_m = shape_sdf(type, sc, scl(add(p_s, q_s), 0.5f), gr);
    if (phi_m > threshold + 0.5f * length(sub(q_s, p_s))) return g;
    V3 x;
    float phi;
    const float u = opt_edge_sdf(type, sc, p_s, q_s, kSdfEdgeIters, x, phi, gr);
    if (phi < threshold) {
        g.hit = true;
        g.i0 = v0;
        g.i1 = v1;
        g.i2 = -1;
        g.bary = v3(1.0f - u, u, 0.0f);
        g.body_pos = xf_point(X_bs, sub(x, scl(gr, phi)));
        g.body_vel = v3(0.0f, 0.0f, 0.0f);
        g.normal = xf_vector(X_ws, gr);
    }
    return g;
}

// create_soft_face_contacts (soft_contacts_sdf.py:277-377)
__device__ ContactGeo face_contact_geometry(const ColParams& C, const float* __restrict__ q, float margin, int k) {
    ContactGeo g;
    g.hit = false;
    g.particle = -1;
    const int t = C.fpairs[2 * k], s = C.fpairs[2 * k + 1];
    g.shape = s;
    if ((C.sflags[s] & 4) == 0) return g;
    const int type = C.stype[s];
    if (type != 1 && type != 3) return g;
    const int ia = C.tri_idx[3 * t], ib = C.tri_idx[3 * t + 1], ic = C.tri_idx[3 * t + 2];
    const float radius = fmaxf(C.radius[ia], fmaxf(C.radius[ib], C.radius[ic]));
    Xf X_bs, X_ws, X_sw;
    shape_frames(C, s, X_bs, X_ws, X_sw);
    const V3 a_s = xf_point(X_sw, ld3(q, ia));
    const V3 b_s = xf_point(X_sw, ld3(q, ib));
    const V3 c_s = xf_point(X_sw, ld3(q, ic));
    const float* sc = C.sscale + 3 * s;
    const float threshold = margin + C.smargin[s] + radius;
    const V3 cen = divs(add(add(a_s, b_s), c_s), 3.0f);
    V3 gr;
    const float phi_c = shape_sdf(type, sc, cen, gr);
    const float reach = fmaxf(length(sub(a_s, cen)), fmaxf(length(sub(b_s, cen)), length(sub(c_s, cen))));
    if (phi_c > threshold + reach) return g;
    V3 x;
    float phi;
    const V3 bary = opt_face_sdf(type, sc, a_s, b_s, c_s, kSdfFaceIters, kSdfLsIters, x, phi, gr);
    if (phi < threshold) {
        g.hit = true;
        g.i0 = ia;
        g.i1 = ib;
        g.i2 = ic;
        g.bary = bary;
        g.body_pos = xf_point(X_bs, sub(x, scl(gr, phi)));
        g.body_vel = v3(0.0f, 0.0f, 0.0f);
        g.normal = xf_vector(X_ws, gr);
    }
    return g;
}

__device__ __forceinline__ ContactGeo soft_contact_geometry(const ColParams& C, const float* __restrict__ q,
                                                            float margin, int tid) {
    if (tid < C.P) return particle_contact_geometry(C, q, margin, tid);
    if (tid < C.P + C.E) return edge_contact_geometry(C, q, margin, tid - C.P);
    return face_contact_geometry(C, q, margin, tid - C.P - C.E);
}

// An edge/face result as the count kernel leaves it for the write kernel: {hit, body_pos} {bary, nx} {ny, nz}.
__device__ __forceinline__ void ef_store(const ColParams& C, int tid, const ContactGeo& g) {
    float4* o = C.ef_cache + 3 * (size_t)(tid - C.P);
    if (!g.hit) {
        o[0] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    o[0] = make_float4(1.0f, g.body_pos.x, g.body_pos.y, g.body_pos.z);
    o[1] = make_float4(g.bary.x, g.bary.y, g.bary.z, g.normal.x);
    o[2] = make_float4(g.normal.y, g.normal.z, 0.0f, 0.0f);
}

__device__ __forceinline__ ContactGeo ef_load(const ColParams& C, int tid) {
    ContactGeo g;
    const float4* o = C.ef_cache + 3 * (size_t)(tid - C.P);
    const float4 a = o[0];
    g.hit = a.x != 0.0f;
    g.particle = -1;
    if (!g.hit) return g;
    const float4 b = o[1], c = o[2];
    const int k = tid - C.P;
    if (k < C.E) {
        const int e = C.epairs[2 * k];
        g.shape = C.epairs[2 * k + 1];
        g.i0 = C.edge_idx[4 * e + 2];
        g.i1 = C.edge_idx[4 * e + 3];
        g.i2 = -1;
    } else {
        const int t = C.fpairs[2 * (k - C.E)];
        g.shape = C.fpairs[2 * (k - C.E) + 1];
        g.i0 = C.tri_idx[3 * t];
        g.i1 = C.tri_idx[3 * t + 1];
        g.i2 = C.tri_idx[3 * t + 2];
    }
    g.body_pos = v3(a.y, a.z, a.w);
    g.bary = v3(b.x, b.y, b.z);
    g.normal = v3(b.w, c.x, c.y);
    g.body_vel = v3(0.0f, 0.0f, 0.0f);
    return g;
}

// ---- soft-contact generation from per-tile hit counts ------------------------
// Records in pair order (a stable compaction of the hitting pairs; the reference's atomic counter
// gives an arbitrary order), with no per-pair flag or index array: a tile of kColTile pairs is one
// 1024-lane work-group, one pair per lane (an edge or face pair near a shape runs a few thousand
// dependent SDF evaluations, so a lane must not carry several); hits are ranked with wave ballots
// and per-wave sums; one work-group scans the tile counts.
constexpr int kColTile = 1024;
constexpr int kColWaves = kColTile / 64;

__device__ __forceinline__ int wave_rank(bool hit, int lane, int* total) {
    unsigned long long m = __ballot(hit ? 1 : 0);
    *total = __popcll(m);
    return __popcll(m & ((1ull << lane) - 1ull));
}

__global__ __launch_bounds__(kColTile) void k_collide_tile_count(ColParams C, const float* __restrict__ q,
                                                                 float margin, int* __restrict__ tile_count) {
    __shared__ int wsum[kColWaves];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tid = blockIdx.x * kColTile + threadIdx.x;
    bool hit = false;
    if (tid < C.P + C.E + C.F) {
        const ContactGeo g = soft_contact_geometry(C, q, margin, tid);
        hit = g.hit;
        if (C.ef_cache && tid >= C.P) ef_store(C, tid, g);
    }
    int t;
    wave_rank(hit, lane, &t);
    if (lane == 0) wsum[wave] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        int sum = 0;
        for (int w = 0; w < kColWaves; ++w) sum += wsum[w];
        tile_count[blockIdx.x] = sum;
    }
}

// Exclusive scan of the tile counts in place (one 1024-lane work-group, each lane a run of
// consecutive tiles); the total is the contact count (may exceed soft_contact_max, as the
// reference's atomic counter does).
__global__ __launch_bounds__(1024) void k_collide_tile_scan(int tiles, int* __restrict__ tile_count,
                                                            int* __restrict__ count_out) {
    __shared__ int part[1024];
    const int per = (tiles + 1023) / 1024;
    const int lo = min(threadIdx.x * per, tiles), hi = min(lo + per, tiles);
    int s = 0;
    for (int i = lo; i < hi; ++i) s += tile_count[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int v = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    int run = part[threadIdx.x] - s;
    for (int i = lo; i < hi; ++i) {
        int c = tile_count[i];
        tile_count[i] = run;
        run += c;
    }
    if (threadIdx.x == 1023) *count_out = part[1023];
}

__global__ __launch_bounds__(kColTile) void k_collide_tile_write(ColParams C, const float* __restrict__ q,
                                                                 float margin, const int* __restrict__ tile_off,
                                                                 VbdContactsOut out) {
    __shared__ int wsum[kColWaves];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tid = blockIdx.x * kColTile + threadIdx.x;
    ContactGeo g;
    g.hit = false;
    if (tid < C.P + C.E + C.F)
        g = (C.ef_cache && tid >= C.P) ? ef_load(C, tid) : soft_contact_geometry(C, q, margin, tid);
    int t;
    const int rank = wave_rank(g.hit, lane, &t);
    if (lane == 0) wsum[wave] = t;
    __syncthreads();
    int before = 0;
    for (int w = 0; w < wave; ++w) before += wsum[w];
    const int k = tile_off[blockIdx.x] + before + rank;
    if (g.hit && k < out.soft_contact_max) {
        out.soft_contact_shape[k] = g.shape;
        st3(out.soft_contact_body_pos, k, g.body_pos);
        st3(out.soft_contact_body_vel, k, g.body_vel);
        if (out.soft_contact_particle) out.soft_contact_particle[k] = g.particle;
        out.soft_contact_indices[3 * k] = g.i0;
        out.soft_contact_indices[3 * k + 1] = g.i1;
        out.soft_contact_indices[3 * k + 2] = g.i2;
        if (out.soft_contact_barycentric) st3(out.soft_contact_barycentric, k, g.bary);
        st3(out.soft_contact_normal, k, g.normal);
    }
}

inline int blocks_for(long n, int bs) { return (int)((n + bs - 1) / bs); }
inline int norm_parts(long n) { return n > 0 ? blocks_for(n, kNormBlock * kNormItems) : 1; }

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
constexpr int kWorldPinned = 4096;  // worlds whose driver flags go through pinned staging

struct VbdContext {
    VbdModelDesc model;
    VbdSolverParams params;
    std::vector<int32_t> color_offsets;
    std::vector<int32_t> fast_color_off;
    std::vector<int32_t> slow_color_off;
    // batched drivers: batches of live worlds only (see vbd_iterate_to_residual_worlds)
    std::vector<int32_t> bat_pmin, bat_pmax;  // smallest / largest particle of each packed batch
    std::vector<int32_t> bat_wlo, bat_whi;    // their worlds, for the offsets in bw_key
    std::vector<int32_t> bw_key;
    std::vector<int32_t> live_off;            // per color, into h_batch_list
    std::vector<int32_t> h_batch_list;
    int* batch_list = nullptr;                // device copy (capacity: every packed batch)
    bool use_list = false;
    KParams kp;
    float* scratch3 = nullptr;  // q_prev | inertia | disp, each [N,3]
    int* c_cnt = nullptr;       // [N+1] contacts per color position; zero between substeps
    int* c_ids = nullptr;       // [entry_cap] entry keys (contact id << 2 | corner) in color-position order
    int* c_ent_pos = nullptr;   // [entry_cap] color position of each entry (k_contact_sort, for the pack)
    float* c_state = nullptr;   // pk | mke | mkd | mmu, each [cap]
    int* pos_coff = nullptr;    // [N+1]
    float* cpack = nullptr;     // [entry_cap * 16]
    int* cpack_ef = nullptr;    // [entry_cap * 8] corners and weights of edge/face entries (full surface only)
    float* crows = nullptr;     // [entry_cap * 12] contact rows of a pass (full surface only)
    int* cid_to_pack = nullptr; // [cap * 3] packed entry of each (contact, corner), -1 when none
    float* g_tmp = nullptr;     // [N,3] residual scratch
    double* r_part = nullptr;   // residual partial sums, one per first-stage workgroup
    double* r_dev = nullptr;    // [1] squared residual norm (iterate-to-tolerance driver)
    int* zero_world = nullptr;  // [N] zeros: every particle in world 0 (single-world freezing)
    int* one_state = nullptr;   // [3] active, iterations, live of the single-world driver
    double* r_host = nullptr;   // [1] pinned host copy of r_dev
    int32_t* w_host = nullptr;  // [kWorldPinned] pinned staging of the world driver's state
    void* cub_tmp = nullptr;
    size_t cub_bytes = 0;
    int capacity = 0;
    size_t entry_cap = 0;      // packed entries: capacity, or 3 x capacity with full-surface records
    int32_t max_body = -1;     // largest Model.shape_body entry (-1: every shape static)
    // pass layout on the device
    int* lay_batch = nullptr;
    int* lay_vert = nullptr;
    int* lay_tri = nullptr;
    int* lay_edge = nullptr;
    int* lay_edge_meta = nullptr;
    int* lay_tet = nullptr;
    int* lay_tet_meta = nullptr;
    int* lay_slow = nullptr;
    int* lay_spr = nullptr;
    int* lay_pre = nullptr;
    float* spr_rows = nullptr;
    int* lay_slow_spr_off = nullptr;
    int* lay_slow_spr = nullptr;
    std::vector<int32_t> pre_color_off;
    int* bat_cinfo = nullptr;  // [B, 2] per-substep contact range of each batch
    int* bat_coff = nullptr;   // [B, max_v + 1] per-substep contact offsets within the batch
    int block = 256;           // color-pass work-group shape (256 or 128 lanes)
    int waves = 1;             // occupancy target of the 256-lane pass (1: compiler's choice, or 7)
    int aligned = 0;           // item ranges of a batch on wave boundaries (layout and kernel)
    // cloth self-contact (vbd_self.h)
    bool self_on = false;
    SelfParams sp{};
    int iter = 0;              // iteration index within the current step (collision detection interval)
    int interval = 0;
    int* pos_of = nullptr;     // [N] color position of each particle
    int* sc_tri_world = nullptr;
    int* sc_edge_world = nullptr;
    void* sc_cub = nullptr;
    size_t sc_cub_bytes = 0;
    std::vector<void*> sc_alloc;  // every self-contact buffer (freed by vbd_destroy)
    int max_v = vbd_layout::kLayoutMaxV;
    int blk_words = vbd_layout::block_words(vbd_layout::kLayoutMaxV);
    int64_t fast_items = 0;
    std::vector<size_t> lay_sizes;  // element counts of the uploaded layout arrays (rebuild in place when equal)
    VbdContactsDesc contacts{};
    bool have_contacts = false;
};

struct VbdCollide {
    VbdCollideDesc desc;
    ColParams cp;
    int* tiles = nullptr;  // hit count per tile of kColTile pairs, then its exclusive scan
    float4* ef_cache = nullptr;  // edge/face results between the count and the write kernel
    bool has_bodies = false;  // some shape is attached to a body (its poses come with each call)
    int32_t max_body = -1;
};

namespace {

template <typename T>
std::string d2h(std::vector<T>& dst, const void* src, size_t count) {
    dst.resize(count);
    if (count == 0) return std::string();
    if (!src) return "null device array";
    hipError_t e = hipMemcpy(dst.data(), src, sizeof(T) * count, hipMemcpyDeviceToHost);
    return e == hipSuccess ? std::string() : std::string(hipGetErrorString(e));
}

template <typename T>
std::string h2d(T** dst, const std::vector<T>& src) {
    hipFree(*dst);
    *dst = nullptr;
    size_t n = src.empty() ? 1 : src.size();
    hipError_t e = hipMalloc(dst, sizeof(T) * n);
    if (e != hipSuccess) return hipGetErrorString(e);
    if (!src.empty()) e = hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice);
    return e == hipSuccess ? std::string() : std::string(hipGetErrorString(e));
}

void free_layout(VbdContext* c) {
    for (int** p : {&c->lay_batch, &c->lay_vert, &c->lay_tri, &c->lay_edge, &c->lay_edge_meta, &c->lay_tet,
                    &c->lay_tet_meta, &c->lay_slow, &c->bat_cinfo, &c->bat_coff, &c->lay_spr, &c->lay_pre,
                    &c->lay_slow_spr_off, &c->lay_slow_spr, &c->batch_list}) {
        hipFree(*p);
        *p = nullptr;
    }
    c->bw_key.clear();
    hipFree(c->spr_rows);
    c->spr_rows = nullptr;
}

// Smallest / largest particle of each packed batch (the batched drivers launch live worlds' batches).
void set_batch_particle_ranges(VbdContext* c, const vbd_layout::PassLayout& L) {
    const int nb = L.fast_color_off.back();
    c->bat_pmin.assign((size_t)nb, 0);
    c->bat_pmax.assign((size_t)nb, 0);
    for (int b = 0; b < nb; ++b) {
        const int32_t* blk = &L.batch_blk[(size_t)b * L.blk_words];
        int lo = INT32_MAX, hi = -1;
        for (int i = 0; i < blk[1]; ++i) {
            const int v = blk[kBatchHdr + 4 * i];
            lo = std::min(lo, v);
            hi = std::max(hi, v);
        }
        c->bat_pmin[b] = hi < 0 ? 0 : lo;
        c->bat_pmax[b] = hi < 0 ? 0 : hi;
    }
    c->bw_key.clear();
}

// Reads the model arrays back once and builds the color-pass layout on the host (vbd_layout.h).
int32_t build_layout(VbdContext* c) {
    const VbdModelDesc* m = &c->model;
    vbd_layout::HostModel hm;
    hm.N = m->particle_count;
    hm.T = m->tri_count;
    hm.E = m->edge_count;
    hm.K = m->tet_count;
    std::string err;
    auto rd = [&](auto& vec, const void* src, size_t n) {
        if (err.empty()) err = d2h(vec, src, n);
    };
    rd(hm.mass, m->particle_mass, (size_t)hm.N);
    rd(hm.flags, m->particle_flags, (size_t)hm.N);
    rd(hm.tri_idx, m->tri_indices, 3 * (size_t)hm.T);
    rd(hm.tri_pose, m->tri_poses, 4 * (size_t)hm.T);
    rd(hm.tri_mat, m->tri_materials, 5 * (size_t)hm.T);
    rd(hm.tri_area, m->tri_areas, (size_t)hm.T);
    rd(hm.edge_idx, m->edge_indices, 4 * (size_t)hm.E);
    rd(hm.edge_rest, m->edge_rest_angle, (size_t)hm.E);
    rd(hm.edge_len, m->edge_rest_length, (size_t)hm.E);
    rd(hm.edge_bend, m->edge_bending_properties, 2 * (size_t)hm.E);
    rd(hm.tet_idx, m->tet_indices, 4 * (size_t)hm.K);
    rd(hm.tet_pose, m->tet_poses, 9 * (size_t)hm.K);
    rd(hm.tet_mat, m->tet_materials, 3 * (size_t)hm.K);
    if (hm.T > 0) rd(hm.adj_tri_off, m->v_adj_tris_offsets, (size_t)hm.N + 1);
    if (hm.E > 0) rd(hm.adj_edge_off, m->v_adj_edges_offsets, (size_t)hm.N + 1);
    if (hm.K > 0) rd(hm.adj_tet_off, m->v_adj_tets_offsets, (size_t)hm.N + 1);
    if (err.empty() && hm.T > 0) rd(hm.adj_tri, m->v_adj_tris, (size_t)hm.adj_tri_off[hm.N]);
    if (err.empty() && hm.E > 0) rd(hm.adj_edge, m->v_adj_edges, (size_t)hm.adj_edge_off[hm.N]);
    if (err.empty() && hm.K > 0) rd(hm.adj_tet, m->v_adj_tets, (size_t)hm.adj_tet_off[hm.N]);
    hm.S = m->spring_count;
    rd(hm.spring_idx, m->spring_indices, 2 * (size_t)hm.S);
    rd(hm.spring_rest, m->spring_rest_length, (size_t)hm.S);
    rd(hm.spring_ke, m->spring_stiffness, (size_t)hm.S);
    rd(hm.spring_kd, m->spring_damping, (size_t)hm.S);
    hm.color_off = c->color_offsets;
    rd(hm.color_particles, m->color_particles, (size_t)c->color_offsets.back());
    if (!err.empty()) return fail(VBD_ERR_HIP, "vbd_create: reading the model: " + err);
    // validate the color groups: every particle exactly once
    {
        std::vector<char> seen((size_t)hm.N, 0);
        if ((int)hm.color_particles.size() != hm.N)
            return fail(VBD_ERR_ARG, "vbd_create: color groups must contain every particle exactly once");
        for (int32_t v : hm.color_particles) {
            if (v < 0 || v >= hm.N || seen[v])
                return fail(VBD_ERR_ARG, "vbd_create: color groups must contain every particle exactly once");
            seen[v] = 1;
        }
    }
    vbd_layout::PassLayout L;
    err = vbd_layout::build_pass_layout(hm, L, c->block, c->block / kComp, c->aligned != 0);
    if (!err.empty()) return fail(VBD_ERR_ARG, "vbd_create: " + err);
    const std::pair<int**, std::vector<int32_t>*> arrays[] = {
        {&c->lay_batch, &L.batch_blk}, {&c->lay_vert, &L.vert_rec}, {&c->lay_tri, &L.tri_rec},
        {&c->lay_edge, &L.edge_rec}, {&c->lay_edge_meta, &L.edge_meta}, {&c->lay_tet, &L.tet_rec},
        {&c->lay_tet_meta, &L.tet_meta}, {&c->lay_slow, &L.slow_pos}, {&c->lay_spr, &L.spr_rec},
        {&c->lay_pre, &L.pre_item}, {&c->lay_slow_spr_off, &L.slow_spr_off}, {&c->lay_slow_spr, &L.slow_spr}};
    std::vector<size_t> sizes;
    for (const auto& pr : arrays) sizes.push_back(pr.second->size());
    // A rebuild whose layout has the same shape (array sizes, per-color batch / generic / pre-row
    // ranges, batch width) -- the case of material or rest-data edits -- rewrites the existing
    // device buffers, so every pointer and launch shape a captured graph holds stays valid.
    const bool in_place = c->lay_batch != nullptr && sizes == c->lay_sizes && L.fast_color_off == c->fast_color_off &&
                          L.slow_color_off == c->slow_color_off && L.pre_color_off == c->pre_color_off &&
                          L.max_v == c->max_v && L.blk_words == c->blk_words;
    if (in_place) {
        for (const auto& pr : arrays) {
            if (pr.second->empty()) continue;
            hipError_t e = hipMemcpy(*pr.first, pr.second->data(), sizeof(int32_t) * pr.second->size(),
                                     hipMemcpyHostToDevice);
            if (e != hipSuccess)
                return fail(VBD_ERR_HIP, std::string("vbd_rebuild: rewriting the layout: ") + hipGetErrorString(e));
        }
        std::vector<int32_t> pos_of((size_t)hm.N);
        for (int k = 0; k < hm.N; ++k) pos_of[hm.color_particles[k]] = k;
        hipError_t e = hipMemcpy(c->pos_of, pos_of.data(), sizeof(int32_t) * (size_t)hm.N, hipMemcpyHostToDevice);
        if (e != hipSuccess)
            return fail(VBD_ERR_HIP, std::string("vbd_rebuild: color positions: ") + hipGetErrorString(e));
        c->fast_items = L.fast_items;
        set_batch_particle_ranges(c, L);
        // the batches may hold other vertices now: recompute their contact ranges of the current substep
        const int nbat = c->fast_color_off.back();
        if (nbat > 0) {
            hipLaunchKernelGGL(k_batch_contacts, dim3(blocks_for(nbat, 256)), dim3(256), 0, 0, nbat, c->lay_batch,
                               c->blk_words, c->max_v, c->pos_coff, c->bat_cinfo, c->bat_coff);
            VBD_LAUNCH_CHECK("k_batch_contacts");
        }
        VBD_HIP_TRY(hipDeviceSynchronize());
        return VBD_OK;
    }
    free_layout(c);
    for (const auto& pr : arrays) {
        err = h2d(pr.first, *pr.second);
        if (!err.empty()) return fail(VBD_ERR_HIP, "vbd_create: uploading the layout: " + err);
    }
    c->lay_sizes = sizes;
    c->fast_color_off = L.fast_color_off;
    c->slow_color_off = L.slow_color_off;
    c->pre_color_off = L.pre_color_off;
    {
        std::vector<int32_t> pos_of((size_t)hm.N);
        for (int k = 0; k < hm.N; ++k) pos_of[hm.color_particles[k]] = k;
        err = h2d(&c->pos_of, pos_of);
        if (!err.empty()) return fail(VBD_ERR_HIP, "vbd_create: color positions: " + err);
        c->sp.pos_of = c->pos_of;
    }
    {
        const size_t npre = (size_t)std::max(L.pre_color_off.back(), 1);
        hipError_t e = hipMalloc(&c->spr_rows, sizeof(float) * 12 * npre);
        if (e != hipSuccess) return fail(VBD_ERR_HIP, std::string("vbd_create: spring rows: ") + hipGetErrorString(e));
    }
    c->fast_items = L.fast_items;
    c->max_v = L.max_v;
    c->blk_words = L.blk_words;
    {
        const int nb = L.fast_color_off.back();
        set_batch_particle_ranges(c, L);
        hipError_t e = hipMalloc(&c->batch_list, sizeof(int) * (size_t)std::max(nb, 1));
        if (e != hipSuccess) return fail(VBD_ERR_HIP, std::string("vbd_create: batch list: ") + hipGetErrorString(e));
    }
    {
        // contact ranges per batch: zero (no contacts) until vbd_contact_prepare fills them
        const size_t nbat = (size_t)std::max(L.fast_color_off.back(), 1);
        hipError_t e;
        if ((e = hipMalloc(&c->bat_cinfo, sizeof(int) * 2 * nbat)) != hipSuccess ||
            (e = hipMalloc(&c->bat_coff, sizeof(int) * (L.max_v + 1) * nbat)) != hipSuccess ||
            (e = hipMemset(c->bat_cinfo, 0, sizeof(int) * 2 * nbat)) != hipSuccess ||
            (e = hipMemset(c->bat_coff, 0, sizeof(int) * (L.max_v + 1) * nbat)) != hipSuccess ||
            (e = hipDeviceSynchronize()) != hipSuccess)
            return fail(VBD_ERR_HIP, std::string("vbd_create: batch contact ranges: ") + hipGetErrorString(e));
    }
    KParams& k = c->kp;
    k.batch_blk = c->lay_batch;
    k.bat_cinfo = c->bat_cinfo;
    k.bat_coff = c->bat_coff;
    k.vert_rec = c->lay_vert;
    k.tri_rec = c->lay_tri;
    k.edge_rec = c->lay_edge;
    k.edge_meta = c->lay_edge_meta;
    k.tet_rec = c->lay_tet;
    k.tet_meta = c->lay_tet_meta;
    k.slow_pos = c->lay_slow;
    k.has_spring = m->spring_count > 0;
    k.spring_idx = m->spring_indices;
    k.spring_rest = m->spring_rest_length;
    k.spring_ke = m->spring_stiffness;
    k.spring_kd = m->spring_damping;
    k.spr_rec = c->lay_spr;
    k.pre_item = c->lay_pre;
    k.spr_rows = c->spr_rows;
    k.slow_spr_off = c->lay_slow_spr_off;
    k.slow_spr = c->lay_slow_spr;
    k.aligned = c->aligned;
    return VBD_OK;
}


// ---------------------------------------------------------------------------
// cloth self-contact: buffers, detection pipeline, truncation (kernels in vbd_self.h)
// ---------------------------------------------------------------------------
template <typename T>
hipError_t sc_malloc(VbdContext* c, T** p, size_t count) {
    hipError_t e = hipMalloc(p, sizeof(T) * (count > 0 ? count : 1));
    if (e == hipSuccess) c->sc_alloc.push_back((void*)*p);
    return e;
}

inline int next_pow2(long n) {
    int p = 1;
    while (p < n && p < (1 << 30)) p <<= 1;
    return p;
}

int32_t setup_self(VbdContext* c, const VbdModelDesc* m, const VbdSolverParams* prm) {
    const int N = m->particle_count, T = m->tri_count, E = m->edge_count;
    if (prm->vertex_contact_buffer_size < 1 || prm->vertex_contact_buffer_size > 64 ||
        prm->edge_contact_buffer_size < 1 || prm->edge_contact_buffer_size > 64)
        return fail(VBD_ERR_ARG, "vbd_create: vertex/edge contact buffer sizes must be in [1, 64]");
    if (!(prm->self_contact_radius > 0.0f) || prm->self_contact_margin < prm->self_contact_radius)
        return fail(VBD_ERR_ARG, "vbd_create: self-contact margin must be >= radius > 0");
    if (T > 0 && (!m->tri_indices)) return fail(VBD_ERR_ARG, "vbd_create: self-contact needs tri_indices");
    if (E > 0 && (!m->edge_indices)) return fail(VBD_ERR_ARG, "vbd_create: self-contact needs edge_indices");
    if ((prm->vt_filter == nullptr) != (prm->vt_filter_offsets == nullptr) ||
        (prm->ee_filter == nullptr) != (prm->ee_filter_offsets == nullptr))
        return fail(VBD_ERR_ARG, "vbd_create: a contact filter needs both its list and its offsets");
    const long vt_cap = prm->vertex_contact_buffer_size, ee_cap = prm->edge_contact_buffer_size;
    const long nslots = (long)E * ee_cap + (long)N * vt_cap;
    const long it_cap = 2L * E * ee_cap + 4L * N * vt_cap;
    if (it_cap >= (1L << 31) - 1)
        return fail(VBD_ERR_ARG, "vbd_create: self-contact buffers exceed 2^31 items; lower the buffer sizes");
    // per-primitive worlds and whether the grid key can include the world
    std::vector<int32_t> tri, edge, pw;
    std::string err = d2h(tri, m->tri_indices, 3 * (size_t)T);
    if (err.empty()) err = d2h(edge, m->edge_indices, 4 * (size_t)E);
    if (err.empty() && prm->particle_world) err = d2h(pw, prm->particle_world, (size_t)N);
    if (!err.empty()) return fail(VBD_ERR_HIP, "vbd_create: reading the mesh for self-contact: " + err);
    if (E > 0) {
        // edge-edge items are listed by walking each particle's edge CSR, which must be ascending
        std::vector<int32_t> off, adj;
        err = d2h(off, m->v_adj_edges_offsets, (size_t)N + 1);
        if (err.empty()) err = d2h(adj, m->v_adj_edges, (size_t)off[N]);
        if (!err.empty()) return fail(VBD_ERR_HIP, "vbd_create: reading the edge adjacency: " + err);
        for (int v = 0; v < N; ++v)
            for (int c2 = off[v] + 2; c2 < off[v + 1]; c2 += 2)
                if (adj[c2] <= adj[c2 - 2])
                    return fail(VBD_ERR_ARG, "vbd_create: self-contact needs each particle's edge adjacency in "
                                             "ascending edge order");
    }
    std::vector<int32_t> tw((size_t)T, 0), ew((size_t)E, 0);
    bool any_global = false;
    int W = 1;
    for (int i = 0; i < (int)pw.size(); ++i) {
        any_global |= pw[i] < 0;
        W = std::max(W, pw[i] + 1);
    }
    for (int t = 0; t < T; ++t) tw[t] = pw.empty() ? 0 : pw[tri[3 * t]];
    for (int e = 0; e < E; ++e) ew[e] = pw.empty() ? 0 : pw[edge[4 * e + 2]];
    if ((err = h2d(&c->sc_tri_world, tw)).size() || (err = h2d(&c->sc_edge_world, ew)).size())
        return fail(VBD_ERR_HIP, "vbd_create: " + err);
    c->sc_alloc.push_back(c->sc_tri_world);
    c->sc_alloc.push_back(c->sc_edge_world);

    SelfParams& S = c->sp;
    S.N = N;
    S.T = T;
    S.E = E;
    S.W = W;
    S.world_hash = (!pw.empty() && !any_global && W > 1) ? 1 : 0;
    S.tri_idx = m->tri_indices;
    S.edge_idx = m->edge_indices;
    S.tri_world = c->sc_tri_world;
    S.edge_world = c->sc_edge_world;
    S.pworld = prm->particle_world;
    S.color_particles = m->color_particles;
    S.radius = prm->self_contact_radius;
    S.margin = prm->self_contact_margin;
    S.relax = prm->conservative_bound_relaxation;
    S.max_disp = prm->max_displacement;
    S.par_eps = prm->edge_parallel_epsilon;
    S.rest_excl = prm->rest_shape_contact_exclusion_radius;
    S.ke = m->soft_contact_ke;
    S.kd = m->soft_contact_kd;
    S.mu = m->soft_contact_mu;
    S.friction_eps = prm->friction_epsilon;
    S.vt_cap = (int)vt_cap;
    S.ee_cap = (int)ee_cap;
    S.q_rest = prm->particle_q_rest;
    S.adj_edge = m->v_adj_edges;
    S.adj_edge_off = m->v_adj_edges_offsets;
    S.vt_filter = prm->vt_filter;
    S.vt_filter_off = prm->vt_filter_offsets;
    S.ee_filter = prm->ee_filter;
    S.ee_filter_off = prm->ee_filter_offsets;
    const long prims = std::max(std::max(T, E), 1);
    S.nbucket = next_pow2(std::max(4L * prims, 1024L));
    S.g_cap = (int)(27L * prims);
    hipError_t e;
    if ((e = sc_malloc(c, &S.vt_list, (size_t)N * vt_cap)) != hipSuccess ||
        (e = sc_malloc(c, &S.vt_count, (size_t)N)) != hipSuccess ||
        (e = sc_malloc(c, &S.vt_min, (size_t)N)) != hipSuccess || (e = sc_malloc(c, &S.tri_min, (size_t)T)) != hipSuccess ||
        (e = sc_malloc(c, &S.ee_list, (size_t)E * ee_cap)) != hipSuccess ||
        (e = sc_malloc(c, &S.ee_count, (size_t)E)) != hipSuccess || (e = sc_malloc(c, &S.ee_min, (size_t)E)) != hipSuccess ||
        (e = sc_malloc(c, &S.ext, 2)) != hipSuccess || (e = sc_malloc(c, &S.g_count, (size_t)S.nbucket + 1)) != hipSuccess ||
        (e = sc_malloc(c, &S.g_off, (size_t)S.nbucket + 1)) != hipSuccess ||
        (e = sc_malloc(c, &S.g_ent, (size_t)S.g_cap)) != hipSuccess ||
        (e = sc_malloc(c, &S.g_box, 2 * (size_t)S.g_cap)) != hipSuccess || (e = sc_malloc(c, &S.g_overflow, 1)) != hipSuccess ||
        (e = sc_malloc(c, &S.it_cnt, (size_t)N + 1)) != hipSuccess ||
        (e = sc_malloc(c, &S.it_off, (size_t)N + 1)) != hipSuccess || (e = sc_malloc(c, &S.it_cur, (size_t)N)) != hipSuccess ||
        (e = sc_malloc(c, &S.ee_n, (size_t)N)) != hipSuccess ||
        (e = sc_malloc(c, &S.items, (size_t)it_cap)) != hipSuccess ||
        (e = sc_malloc(c, &S.rows, (size_t)it_cap * 12)) != hipSuccess ||
        (e = sc_malloc(c, &S.pos_cd, 3 * (size_t)N)) != hipSuccess || (e = sc_malloc(c, &S.trunc, (size_t)N)) != hipSuccess ||
        (e = sc_malloc(c, &S.coll, (size_t)nslots)) != hipSuccess || (e = sc_malloc(c, &S.ncoll, 1)) != hipSuccess)
        return fail(VBD_ERR_HIP, std::string("vbd_create: self-contact buffers: ") + hipGetErrorString(e));
    size_t b1 = 0, b2 = 0;
    if ((e = hipcub::DeviceScan::ExclusiveSum(nullptr, b1, S.g_count, S.g_off, S.nbucket + 1)) != hipSuccess ||
        (e = hipcub::DeviceScan::ExclusiveSum(nullptr, b2, S.it_cnt, S.it_off, N + 1)) != hipSuccess)
        return fail(VBD_ERR_HIP, std::string("vbd_create: scan sizing: ") + hipGetErrorString(e));
    c->sc_cub_bytes = std::max(std::max(b1, b2), (size_t)16);
    if ((e = hipMalloc(&c->sc_cub, c->sc_cub_bytes)) != hipSuccess)
        return fail(VBD_ERR_HIP, std::string("vbd_create: scan scratch: ") + hipGetErrorString(e));
    c->sc_alloc.push_back(c->sc_cub);
    // empty lists until the first detection; truncation factors 1; overflow flag clear
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(N + 1, 256)), dim3(256), 0, 0, S.it_off, 0, N + 1);
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(N, 256)), dim3(256), 0, 0, S.vt_count, 0, N);
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(std::max(E, 1), 256)), dim3(256), 0, 0, S.ee_count, 0, std::max(E, 1));
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(N, 256)), dim3(256), 0, 0, S.trunc, vbd_layout::fbits(1.0f), N);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, 0, S.ncoll, 0, 1);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, 0, S.g_overflow, 0, 1);
    // bucket counts start at zero; each detection's write pass counts them back down to zero
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(S.nbucket + 1, 256)), dim3(256), 0, 0, S.g_count, 0,
                       S.nbucket + 1);
    if ((e = hipDeviceSynchronize()) != hipSuccess)
        return fail(VBD_ERR_HIP, std::string("vbd_create: self-contact init: ") + hipGetErrorString(e));
    c->self_on = true;
    c->interval = prm->collision_detection_interval;
    KParams& k = c->kp;
    k.has_self = 1;
    k.so_off = S.it_off;
    k.so_rows = S.rows;
    k.write_q = 0;
    return VBD_OK;
}

int32_t sc_scan(VbdContext* c, hipStream_t s, const int* in, int* out, int n) {
    size_t bytes = c->sc_cub_bytes;
    VBD_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(c->sc_cub, bytes, in, out, n, s));
    return VBD_OK;
}

// TriMeshCollisionDetector.refit + vertex_triangle / edge_edge_collision_detection
// (solver_vbd.py:3221-3237) at positions q, then the per-color-position item lists
int32_t self_detect(VbdContext* c, hipStream_t s, const float* q) {
    SelfParams& S = c->sp;
    const int N = S.N, T = S.T, E = S.E;
    const long nslots = (long)E * S.ee_cap + (long)N * S.vt_cap;
    const int nb = S.nbucket;
    // one launch clears what this detection accumulates into (the bucket counts are already zero)
    FillList F{};
    auto add_fill = [&F](int* p, int value, int n) {
        F.p[F.count] = p;
        F.value[F.count] = value;
        F.n[F.count] = n;
        ++F.count;
    };
    add_fill(S.ext, 0, 2);
    add_fill(S.tri_min, vbd_layout::fbits(S.margin), T);
    add_fill(S.it_cnt, 0, N + 1);
    add_fill(S.ee_n, 0, N);
    add_fill(S.ncoll, 0, 1);
    hipLaunchKernelGGL(k_fill_many, dim3(blocks_for(std::max(std::max(T, N + 1), 1), 256)), dim3(256), 0, s, F);
    VBD_LAUNCH_CHECK("k_fill_many");
    const int np = std::max(std::max(T, E), 1);
    hipLaunchKernelGGL(k_sc_extents, dim3(blocks_for(np, 256)), dim3(256), 0, s, S, q);
    VBD_LAUNCH_CHECK("k_sc_extents");
    for (int mode = 0; mode < 2; ++mode) {
        const int n = mode == 0 ? T : E;
        if (n > 0) {
            hipLaunchKernelGGL(k_sc_grid_insert, dim3(blocks_for(n, 256)), dim3(256), 0, s, S, q, mode, 0);
            VBD_LAUNCH_CHECK("k_sc_grid_insert");
        }
        int32_t rc = sc_scan(c, s, S.g_count, S.g_off, nb + 1);
        if (rc) return rc;
        if (n > 0) {
            hipLaunchKernelGGL(k_sc_grid_insert, dim3(blocks_for(n, 256)), dim3(256), 0, s, S, q, mode, 1);
            VBD_LAUNCH_CHECK("k_sc_grid_insert");
        }
        if (mode == 0) {
            hipLaunchKernelGGL(k_sc_vt_query, dim3(blocks_for(N, kQPerBlock)), dim3(256), 0, s, S, q);
            VBD_LAUNCH_CHECK("k_sc_vt_query");
        } else if (E > 0) {
            hipLaunchKernelGGL(k_sc_ee_query, dim3(blocks_for(E, kQPerBlock)), dim3(256), 0, s, S, q);
            VBD_LAUNCH_CHECK("k_sc_ee_query");
        }
    }
    // item lists per color position
    hipLaunchKernelGGL(k_sc_items_count_ee, dim3(blocks_for(N, 256)), dim3(256), 0, s, S);
    VBD_LAUNCH_CHECK("k_sc_items_count_ee");
    const int gv = (int)std::max(1L, ((long)N * S.vt_cap + 255) / 256);
    hipLaunchKernelGGL(k_sc_items_count, dim3(gv), dim3(256), 0, s, S);
    VBD_LAUNCH_CHECK("k_sc_items_count");
    int32_t rc = sc_scan(c, s, S.it_cnt, S.it_off, N + 1);
    if (rc) return rc;
    hipLaunchKernelGGL(k_sc_items_ee, dim3(blocks_for(N, 256)), dim3(256), 0, s, S);
    VBD_LAUNCH_CHECK("k_sc_items_ee");
    hipLaunchKernelGGL(k_sc_items_vt, dim3(gv), dim3(256), 0, s, S);
    VBD_LAUNCH_CHECK("k_sc_items_vt");
    hipLaunchKernelGGL(k_sc_items_sort, dim3(blocks_for(N, 256)), dim3(256), 0, s, S);
    VBD_LAUNCH_CHECK("k_sc_items_sort");
    // compact collision list for the truncation
    hipLaunchKernelGGL(k_sc_collect, dim3((int)std::max(1L, (nslots + 256L * kCollectPer - 1) / (256L * kCollectPer))),
                       dim3(256), 0, s, S);
    VBD_LAUNCH_CHECK("k_sc_collect");
    // displacements are measured from the positions of the last detection
    hipLaunchKernelGGL(k_sc_copy3, dim3(blocks_for(3L * N, 256)), dim3(256), 0, s, S.pos_cd, q, N);
    VBD_LAUNCH_CHECK("k_sc_copy3");
    return VBD_OK;
}

// _penetration_free_truncation with self-contact (solver_vbd.py:2025-2063)
int32_t self_truncate(VbdContext* c, hipStream_t s, float* q) {
    SelfParams& S = c->sp;
    hipLaunchKernelGGL(k_sc_trunc_collisions, dim3(kSelfGrid), dim3(256), 0, s, S, c->kp.disp);
    VBD_LAUNCH_CHECK("k_sc_trunc_collisions");
    hipLaunchKernelGGL(k_sc_trunc_apply, dim3(blocks_for(S.N, 256)), dim3(256), 0, s, S, c->kp.disp, q);
    VBD_LAUNCH_CHECK("k_sc_trunc_apply");
    return VBD_OK;
}

}  // namespace

// Device scratch of vbd_collide_generate_noalloc: one int per tile of kColTile pairs (hit counts,
// then their exclusive scan).  Each GPU holds its own copy for the life of the process; every
// call on a GPU reuses it, so calls that overlap on one GPU race.  Nothing in it needs clearing:
// k_collide_tile_count writes every slot the call reads.
static_assert(VBD_NOALLOC_MAX_PAIRS % kColTile == 0, "whole tiles");
constexpr int kNoallocTiles = VBD_NOALLOC_MAX_PAIRS / kColTile;
static __device__ __align__(256) int32_t scratch[kNoallocTiles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int32_t* g_scratch_dev[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per device (the array's
// name in host code is only a host-side handle, never a device address).
std::string noalloc_scratch(int32_t** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return std::string("hipGetDevice: ") + hipGetErrorString(e);
    if (dev < 0 || dev >= kMaxDevices) return "device index " + std::to_string(dev) + " out of range";
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        if (!p) return "hipGetSymbolAddress(scratch) returned null";
        g_scratch_dev[dev] = static_cast<int32_t*>(p);
    }
    *out = g_scratch_dev[dev];
    return std::string();
}

// Soft contacts in pair order from per-tile hit counts (count, scan of the tile counts, write).
int32_t launch_collide_tiles(const ColParams& cp, hipStream_t s, const float* q, float margin, int* tiles_dev,
                             const VbdContactsOut& out) {
    const int tiles = blocks_for((long)cp.P + cp.E + cp.F, kColTile);
    hipLaunchKernelGGL(k_collide_tile_count, dim3(tiles), dim3(kColTile), 0, s, cp, q, margin, tiles_dev);
    VBD_LAUNCH_CHECK("k_collide_tile_count");
    hipLaunchKernelGGL(k_collide_tile_scan, dim3(1), dim3(1024), 0, s, tiles, tiles_dev, out.soft_contact_count);
    VBD_LAUNCH_CHECK("k_collide_tile_scan");
    hipLaunchKernelGGL(k_collide_tile_write, dim3(tiles), dim3(kColTile), 0, s, cp, q, margin, tiles_dev, out);
    VBD_LAUNCH_CHECK("k_collide_tile_write");
    return VBD_OK;
}

}  // namespace

extern "C" {

const char* vbd_last_error(void) { return g_last_error.c_str(); }

int32_t vbd_version(void) { return 2; }

int32_t vbd_create(const VbdModelDesc* m, const VbdSolverParams* prm, VbdContext** out) {
    if (!m || !prm || !out) return fail(VBD_ERR_ARG, "vbd_create: null argument");
    *out = nullptr;
    const int N = m->particle_count;
    if (N <= 0) return fail(VBD_ERR_ARG, "vbd_create: particle_count must be positive");
    if (m->color_count <= 0 || !m->color_offsets_host || !m->color_particles)
        return fail(VBD_ERR_ARG, "vbd_create: model.particle_color_groups is empty (color the model first)");
    if (!m->particle_mass || !m->particle_inv_mass || !m->particle_flags || !m->particle_radius || !m->gravity)
        return fail(VBD_ERR_ARG, "vbd_create: missing particle arrays");
    if (m->tri_count > 0 && (!m->v_adj_tris || !m->v_adj_tris_offsets || !m->tri_indices || !m->tri_poses ||
                             !m->tri_materials || !m->tri_areas))
        return fail(VBD_ERR_ARG, "vbd_create: triangle arrays or adjacency missing");
    if (m->edge_count > 0 && (!m->v_adj_edges || !m->v_adj_edges_offsets || !m->edge_indices ||
                              !m->edge_rest_angle || !m->edge_rest_length || !m->edge_bending_properties))
        return fail(VBD_ERR_ARG, "vbd_create: edge arrays or adjacency missing");
    if (m->tet_count > 0 && (!m->v_adj_tets || !m->v_adj_tets_offsets || !m->tet_indices || !m->tet_poses ||
                             !m->tet_materials))
        return fail(VBD_ERR_ARG, "vbd_create: tet arrays or adjacency missing");
    if (m->spring_count < 0 || (m->spring_count > 0 && (!m->spring_indices || !m->spring_rest_length ||
                                                         !m->spring_stiffness || !m->spring_damping)))
        return fail(VBD_ERR_ARG, "vbd_create: spring arrays missing");
    if (prm->iterations < 0) return fail(VBD_ERR_ARG, "vbd_create: iterations must be >= 0");
    if (prm->soft_contact_capacity < 0) return fail(VBD_ERR_ARG, "vbd_create: negative soft_contact_capacity");
    if (prm->soft_contact_capacity >= (1 << 29))
        return fail(VBD_ERR_ARG, "vbd_create: soft_contact_capacity must be below 2^29");
    if (m->body_count < 0) return fail(VBD_ERR_ARG, "vbd_create: negative body_count");

    VbdContext* c = new VbdContext();
    c->model = *m;
    c->params = *prm;
    c->color_offsets.assign(m->color_offsets_host, m->color_offsets_host + m->color_count + 1);
    if (c->color_offsets[0] != 0 || c->color_offsets.back() != N) {
        delete c;
        return fail(VBD_ERR_ARG, "vbd_create: color_offsets must run from 0 to particle_count");
    }
    c->capacity = prm->soft_contact_capacity > 0 ? prm->soft_contact_capacity : 1;
    auto cleanup = [&](const char* what, hipError_t e) {
        std::string msg = std::string("vbd_create: ") + what + ": " + hipGetErrorString(e);
        vbd_destroy(c);
        return fail(VBD_ERR_HIP, msg);
    };
    hipError_t e;
    const size_t cap = (size_t)c->capacity;
    c->entry_cap = prm->soft_contact_full_surface ? 3 * cap : cap;
    if (m->shape_body && m->shape_count > 0) {
        std::vector<int32_t> body;
        std::string err = d2h(body, m->shape_body, (size_t)m->shape_count);
        if (!err.empty()) {
            vbd_destroy(c);
            return fail(VBD_ERR_HIP, "vbd_create: shape_body: " + err);
        }
        for (int32_t b : body) c->max_body = std::max(c->max_body, b);
        if (c->max_body >= m->body_count) {
            vbd_destroy(c);
            return fail(VBD_ERR_ARG, "vbd_create: shape_body names body " + std::to_string(c->max_body) +
                                         " of a model with body_count " + std::to_string(m->body_count));
        }
    }
    if ((e = hipMalloc(&c->scratch3, sizeof(float) * 9 * (size_t)N)) != hipSuccess) return cleanup("hipMalloc", e);
    if ((e = hipMalloc(&c->c_cnt, sizeof(int) * (size_t)(N + 1))) != hipSuccess) return cleanup("hipMalloc", e);
    if ((e = hipMalloc(&c->pos_coff, sizeof(int) * (size_t)(N + 1))) != hipSuccess) return cleanup("hipMalloc", e);
    if ((e = hipMalloc(&c->c_ids, sizeof(int) * c->entry_cap)) != hipSuccess) return cleanup("hipMalloc", e);
    if ((e = hipMalloc(&c->c_ent_pos, sizeof(int) * c->entry_cap)) != hipSuccess) return cleanup("hipMalloc", e);
    if ((e = hipMalloc(&c->cid_to_pack, sizeof(int) * 3 * cap)) != hipSuccess) return cleanup("hipMalloc", e);
    if ((e = hipMalloc(&c->c_state, sizeof(float) * 4 * cap)) != hipSuccess) return cleanup("hipMalloc", e);
    if ((e = hipMalloc(&c->cpack, sizeof(float) * kContactWords * c->entry_cap)) != hipSuccess)
        return cleanup("hipMalloc", e);
    if (prm->soft_contact_full_surface &&
        ((e = hipMalloc(&c->cpack_ef, sizeof(int) * 8 * c->entry_cap)) != hipSuccess ||
         (e = hipMalloc(&c->crows, sizeof(float) * kComp * c->entry_cap)) != hipSuccess))
        return cleanup("hipMalloc", e);
    if ((e = hipMalloc(&c->g_tmp, sizeof(float) * 3 * (size_t)N)) != hipSuccess) return cleanup("hipMalloc", e);
    if ((e = hipMalloc(&c->r_part, sizeof(double) * (size_t)norm_parts(N))) != hipSuccess)
        return cleanup("hipMalloc", e);
    if ((e = hipMalloc(&c->r_dev, sizeof(double))) != hipSuccess) return cleanup("hipMalloc", e);
    if ((e = hipMalloc(&c->zero_world, sizeof(int) * (size_t)std::max(N, 1))) != hipSuccess ||
        (e = hipMemset(c->zero_world, 0, sizeof(int) * (size_t)std::max(N, 1))) != hipSuccess ||
        (e = hipMalloc(&c->one_state, sizeof(int) * 3)) != hipSuccess)
        return cleanup("hipMalloc", e);
    if ((e = hipHostMalloc(&c->r_host, sizeof(double))) != hipSuccess) return cleanup("hipHostMalloc", e);
    if ((e = hipHostMalloc(&c->w_host, sizeof(int32_t) * kWorldPinned)) != hipSuccess)
        return cleanup("hipHostMalloc", e);
    if ((e = hipcub::DeviceScan::ExclusiveSum(nullptr, c->cub_bytes, c->c_cnt, c->pos_coff, N + 1)) != hipSuccess)
        return cleanup("scan sizing", e);
    if ((e = hipMalloc(&c->cub_tmp, c->cub_bytes > 0 ? c->cub_bytes : 16)) != hipSuccess)
        return cleanup("hipMalloc", e);
    // no contacts until vbd_contact_prepare: every particle's list is empty
    // per-color-position contact counters: zero between substeps (vbd_contact_prepare counts them
    // up and back down to zero)
    if ((e = hipMemset(c->c_cnt, 0, sizeof(int) * (size_t)(N + 1))) != hipSuccess) return cleanup("hipMemset", e);
    if ((e = hipMemset(c->pos_coff, 0, sizeof(int) * (size_t)(N + 1))) != hipSuccess) return cleanup("hipMemset", e);
    if ((e = hipMemset(c->cid_to_pack, 0xff, sizeof(int) * 3 * cap)) != hipSuccess) return cleanup("hipMemset", e);
    if ((e = hipDeviceSynchronize()) != hipSuccess) return cleanup("sync", e);

    KParams& k = c->kp;
    std::memset(&k, 0, sizeof(k));
    k.N = N;
    k.mass = m->particle_mass;
    k.inv_mass = m->particle_inv_mass;
    k.radius = m->particle_radius;
    k.flags = m->particle_flags;
    k.tri_idx = m->tri_indices;
    k.tri_pose = m->tri_poses;
    k.tri_mat = m->tri_materials;
    k.tri_area = m->tri_areas;
    k.edge_idx = m->edge_indices;
    k.edge_rest_angle = m->edge_rest_angle;
    k.edge_rest_len = m->edge_rest_length;
    k.edge_bend = m->edge_bending_properties;
    k.tet_idx = m->tet_indices;
    k.tet_pose = m->tet_poses;
    k.tet_mat = m->tet_materials;
    k.adj_tri = m->v_adj_tris;
    k.adj_tri_off = m->v_adj_tris_offsets;
    k.adj_edge = m->v_adj_edges;
    k.adj_edge_off = m->v_adj_edges_offsets;
    k.adj_tet = m->v_adj_tets;
    k.adj_tet_off = m->v_adj_tets_offsets;
    k.has_tri = m->tri_count > 0;
    k.has_edge = m->edge_count > 0;
    k.has_tet = m->tet_count > 0;
    k.color_particles = m->color_particles;
    k.gravity = m->gravity;
    k.q_prev = c->scratch3;
    k.inertia = c->scratch3 + 3 * (size_t)N;
    k.disp = c->scratch3 + 6 * (size_t)N;
    k.c_pk = c->c_state;
    k.c_kd = c->c_state + 2 * cap;
    k.c_mu = c->c_state + 3 * cap;
    k.pos_coff = c->pos_coff;
    k.cpack = c->cpack;
    k.cpack_ef = c->cpack_ef;
    k.crows = c->crows;
    k.shape_margin = m->shape_margin;
    k.shape_count = m->shape_count;
    k.friction_epsilon = prm->friction_epsilon;
    k.fp64_ill = prm->fp64_ill_conditioned;
    k.write_q = 1;
    int32_t rc = build_layout(c);
    if (rc == VBD_OK && prm->self_contact) rc = setup_self(c, m, prm);
    if (rc) {
        std::string msg = g_last_error;
        vbd_destroy(c);
        return fail(rc, msg);
    }
    *out = c;
    return VBD_OK;
}

int32_t vbd_rebuild(VbdContext* c) {
    if (!c) return fail(VBD_ERR_ARG, "vbd_rebuild: null context");
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return fail(VBD_ERR_HIP, std::string("vbd_rebuild: ") + hipGetErrorString(e));
    return build_layout(c);
}

int32_t vbd_destroy(VbdContext* c) {
    if (!c) return VBD_OK;
    hipFree(c->scratch3);
    hipFree(c->c_cnt);
    hipFree(c->pos_coff);
    hipFree(c->c_ids);
    hipFree(c->c_ent_pos);
    hipFree(c->cid_to_pack);
    hipFree(c->c_state);
    hipFree(c->cpack);
    hipFree(c->cpack_ef);
    hipFree(c->crows);
    hipFree(c->g_tmp);
    hipFree(c->r_part);
    hipFree(c->r_dev);
    hipFree(c->zero_world);
    hipFree(c->one_state);
    if (c->r_host) hipHostFree(c->r_host);
    if (c->w_host) hipHostFree(c->w_host);
    hipFree(c->cub_tmp);
    free_layout(c);
    hipFree(c->pos_of);
    for (void* p : c->sc_alloc) hipFree(p);
    delete c;
    return VBD_OK;
}

int32_t vbd_self_detect(VbdContext* c, void* stream, const float* q) {
    if (!c || !q) return fail(VBD_ERR_ARG, "vbd_self_detect: null argument");
    if (!c->self_on) return fail(VBD_ERR_STATE, "vbd_self_detect: self-contact is off");
    return self_detect(c, (hipStream_t)stream, q);
}

int32_t vbd_self_copy_lists(VbdContext* c, void* stream, int32_t* vt_list, int32_t* vt_count, float* vt_min,
                            float* tri_min, int32_t* ee_list, int32_t* ee_count, float* ee_min) {
    if (!c) return fail(VBD_ERR_ARG, "vbd_self_copy_lists: null context");
    if (!c->self_on) return fail(VBD_ERR_STATE, "vbd_self_copy_lists: self-contact is off");
    hipStream_t s = (hipStream_t)stream;
    const SelfParams& S = c->sp;
    auto cp = [&](void* dst, const void* src, size_t bytes) {
        return (dst && bytes) ? hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s) : hipSuccess;
    };
    VBD_HIP_TRY(cp(vt_list, S.vt_list, sizeof(int) * (size_t)S.N * S.vt_cap));
    VBD_HIP_TRY(cp(vt_count, S.vt_count, sizeof(int) * (size_t)S.N));
    VBD_HIP_TRY(cp(vt_min, S.vt_min, sizeof(float) * (size_t)S.N));
    VBD_HIP_TRY(cp(tri_min, S.tri_min, sizeof(float) * (size_t)S.T));
    VBD_HIP_TRY(cp(ee_list, S.ee_list, sizeof(int) * (size_t)S.E * S.ee_cap));
    VBD_HIP_TRY(cp(ee_count, S.ee_count, sizeof(int) * (size_t)S.E));
    VBD_HIP_TRY(cp(ee_min, S.ee_min, sizeof(float) * (size_t)S.E));
    return VBD_OK;
}

int32_t vbd_self_status(VbdContext* c, int32_t* grid_overflow) {
    if (!c || !grid_overflow) return fail(VBD_ERR_ARG, "vbd_self_status: null argument");
    *grid_overflow = 0;
    if (!c->self_on) return VBD_OK;
    VBD_HIP_TRY(hipDeviceSynchronize());
    VBD_HIP_TRY(hipMemcpy(grid_overflow, c->sp.g_overflow, sizeof(int32_t), hipMemcpyDeviceToHost));
    return VBD_OK;
}

int32_t vbd_copy_scratch(VbdContext* c, void* stream, int32_t which, float* dst) {
    if (!c || !dst || which < 0 || which > 2) return fail(VBD_ERR_ARG, "vbd_copy_scratch: bad argument");
    const size_t n3 = 3 * (size_t)c->model.particle_count;
    // scratch3 holds q_prev | inertia | displacement
    const float* src = which == 0 ? c->kp.inertia : (which == 1 ? c->kp.q_prev : c->kp.disp);
    VBD_HIP_TRY(hipMemcpyAsync(dst, src, sizeof(float) * n3, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return VBD_OK;
}

int32_t vbd_set_displacement(VbdContext* c, void* stream, const float* src) {
    if (!c || !src) return fail(VBD_ERR_ARG, "vbd_set_displacement: null argument");
    const size_t n3 = 3 * (size_t)c->model.particle_count;
    VBD_HIP_TRY(hipMemcpyAsync(c->kp.disp, src, sizeof(float) * n3, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return VBD_OK;
}

int32_t vbd_layout_stats(VbdContext* c, int64_t* out4) {
    if (!c || !out4) return fail(VBD_ERR_ARG, "vbd_layout_stats: null argument");
    out4[0] = c->fast_color_off.back();
    out4[1] = c->slow_color_off.back();
    out4[2] = c->fast_items;
    out4[3] = c->model.color_count;
    return VBD_OK;
}

// Body poses of the current contacts (static shapes when the model has no body-attached shape).
static BodyPoses body_poses(const VbdContext* c) {
    BodyPoses B{};
    if (c->max_body >= 0) {
        B.shape_body = c->model.shape_body;
        B.body_q = c->contacts.body_q;
        B.body_q_prev = c->contacts.body_q_prev;
        B.body_count = c->model.body_count;
    }
    B.shape_count = c->model.shape_count;
    return B;
}

static int32_t launch_batch_contacts(VbdContext* c, hipStream_t s) {
    const int nbat = c->fast_color_off.back();
    if (nbat <= 0) return VBD_OK;
    hipLaunchKernelGGL(k_batch_contacts, dim3(blocks_for(nbat, 256)), dim3(256), 0, s, nbat, c->lay_batch,
                       c->blk_words, c->max_v, c->pos_coff, c->bat_cinfo, c->bat_coff);
    VBD_LAUNCH_CHECK("k_batch_contacts");
    return VBD_OK;
}

int32_t vbd_contact_prepare(VbdContext* c, void* stream, const VbdContactsDesc* cd) {
    if (!c) return fail(VBD_ERR_ARG, "vbd_contact_prepare: null context");
    hipStream_t s = (hipStream_t)stream;
    const int N = c->model.particle_count;
    if (!cd || cd->soft_contact_max <= 0) {
        c->have_contacts = false;
        hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(N + 1, 256)), dim3(256), 0, s, c->pos_coff, 0, N + 1);
        VBD_LAUNCH_CHECK("k_fill_i32");
        return launch_batch_contacts(c, s);
    }
    if (cd->soft_contact_max > c->capacity)
        return fail(VBD_ERR_STATE, "vbd_contact_prepare: soft_contact_max " + std::to_string(cd->soft_contact_max) +
                                       " exceeds the capacity " + std::to_string(c->capacity) +
                                       " given at creation (buffers cannot grow inside a step)");
    if (!cd->soft_contact_count || !cd->soft_contact_indices || !cd->soft_contact_shape || !cd->soft_contact_body_pos ||
        !cd->soft_contact_body_vel || !cd->soft_contact_normal)
        return fail(VBD_ERR_ARG, "vbd_contact_prepare: missing contact arrays");
    if (c->model.shape_count <= 0 || !c->model.shape_material_ke)
        return fail(VBD_ERR_ARG, "vbd_contact_prepare: contacts given but the model has no shapes");
    if (cd->soft_contact_barycentric && !c->params.soft_contact_full_surface)
        return fail(VBD_ERR_STATE, "vbd_contact_prepare: barycentric weights given, but the solver was created "
                                   "without soft_contact_full_surface (no room for edge/face entries)");
    if (c->max_body >= 0) {
        if (!cd->body_q || !cd->body_q_prev)
            return fail(VBD_ERR_ARG, "vbd_contact_prepare: shapes are attached to bodies; body_q and body_q_prev "
                                     "(the external rigid solver's poses at the step's end and start) are required");
        if (!(cd->dt > 0.0f)) return fail(VBD_ERR_ARG, "vbd_contact_prepare: dt must be positive with body poses");
    }
    c->contacts = *cd;
    c->have_contacts = true;
    const int cap = c->capacity;
    const int ef = cd->soft_contact_barycentric != nullptr;
    float* pk = c->c_state;
    float* mke = c->c_state + cap;
    float* mkd = c->c_state + 2 * (size_t)cap;
    float* mmu = c->c_state + 3 * (size_t)cap;
    const int cb = blocks_for(cd->soft_contact_max, 256);
    // counts per color position (c_cnt is all zero between substeps: k_contact_place counts it back down)
    hipLaunchKernelGGL(k_contact_init_pos, dim3(cb), dim3(256), 0, s, cd->soft_contact_count, cd->soft_contact_max,
                       cd->soft_contact_indices, cd->soft_contact_shape, c->model.soft_contact_ke,
                       c->model.soft_contact_kd, c->model.soft_contact_mu, c->model.shape_material_ke,
                       c->model.shape_material_kd, c->model.shape_material_mu, c->params.k_start, pk, mke, mkd, mmu,
                       c->pos_of, c->c_cnt, N, c->model.shape_count, ef);
    VBD_LAUNCH_CHECK("k_contact_init_pos");
    size_t bytes = c->cub_bytes;
    VBD_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(c->cub_tmp, bytes, c->c_cnt, c->pos_coff, N + 1, s));
    hipLaunchKernelGGL(k_contact_place, dim3(cb), dim3(256), 0, s, cd->soft_contact_count, cd->soft_contact_max,
                       cd->soft_contact_indices, cd->soft_contact_shape, c->pos_of, c->pos_coff, c->c_cnt, c->c_ids,
                       N, c->model.shape_count, ef);
    VBD_LAUNCH_CHECK("k_contact_place");
    if (c->params.avbd_beta != 0.0f) {
        hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(3 * (long)cap, 256)), dim3(256), 0, s, c->cid_to_pack, -1,
                           3 * cap);
        VBD_LAUNCH_CHECK("k_fill_i32");
    }
    PackArgs pa;
    pa.pk = pk;
    pa.mkd = mkd;
    pa.mmu = mmu;
    pa.c_indices = cd->soft_contact_indices;
    pa.c_bary = cd->soft_contact_barycentric;
    pa.c_shape = cd->soft_contact_shape;
    pa.body_pos = cd->soft_contact_body_pos;
    pa.body_vel = cd->soft_contact_body_vel;
    pa.normal = cd->soft_contact_normal;
    pa.shape_margin = c->model.shape_margin;
    pa.shape_count = c->model.shape_count;
    pa.radius = c->model.particle_radius;
    pa.bodies = body_poses(c);
    pa.dt = cd->dt;
    hipLaunchKernelGGL(k_contact_sort, dim3(blocks_for(N, 256)), dim3(256), 0, s, N, c->pos_coff, c->c_ids,
                       c->c_ent_pos);
    VBD_LAUNCH_CHECK("k_contact_sort");
    hipLaunchKernelGGL(k_contact_pack, dim3(kPackGrid), dim3(256), 0, s, N, c->model.color_particles, c->lay_vert,
                       c->pos_coff, c->c_ids, c->c_ent_pos, pa, c->cpack, c->cpack_ef,
                       c->params.avbd_beta != 0.0f ? c->cid_to_pack : nullptr);
    VBD_LAUNCH_CHECK("k_contact_pack");
    int32_t rc = launch_batch_contacts(c, s);
    if (rc) return rc;
    KParams& k = c->kp;
    k.c_shape = cd->soft_contact_shape;
    k.c_body_pos = cd->soft_contact_body_pos;
    k.c_body_vel = cd->soft_contact_body_vel;
    k.c_normal = cd->soft_contact_normal;
    return VBD_OK;
}

int32_t vbd_forward(VbdContext* c, void* stream, float* q, const float* qd, const float* f_ext, float dt) {
    if (!c || !q || !qd) return fail(VBD_ERR_ARG, "vbd_forward: null argument");
    hipStream_t s = (hipStream_t)stream;
    // with self-contact: detection before initialization (solver_vbd.py:2074-2075)
    if (c->self_on) {
        int32_t rc = self_detect(c, s, q);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_forward, dim3(blocks_for(c->model.particle_count, 256)), dim3(256), 0, s, c->kp, q, qd, f_ext,
                       dt);
    VBD_LAUNCH_CHECK("k_forward");
    c->iter = 0;
    if (c->self_on) return self_truncate(c, s, q);
    return VBD_OK;
}

namespace {
// mode 0 solves; mode 1 writes the assembled force g: by the force-only instantiations when
// force_only (the residual), else by the solving kernels themselves (the autotuner times those)
int32_t launch_fast(VbdContext* c, hipStream_t s, int b0, int nb, float* q, float dt, int mode, float* g,
                    bool force_only = false) {
    if (mode == 1 && force_only) {
        if (c->block == 128) {
            hipLaunchKernelGGL((k_solve_fast<128, 1, true>), dim3(nb), dim3(128), 0, s, c->kp, q, b0, nb, dt, mode, g);
        } else if (c->block == 192) {
            hipLaunchKernelGGL((k_solve_fast<192, 1, true>), dim3(nb), dim3(192), 0, s, c->kp, q, b0, nb, dt, mode, g);
        } else {
            hipLaunchKernelGGL((k_solve_fast<256, 1, true>), dim3(nb), dim3(256), 0, s, c->kp, q, b0, nb, dt, mode, g);
        }
        VBD_LAUNCH_CHECK("k_solve_fast");
        return VBD_OK;
    }
    if (c->block == 128) {
        hipLaunchKernelGGL((k_solve_fast<128, 1>), dim3(nb), dim3(128), 0, s, c->kp, q, b0, nb, dt, mode, g);
    } else if (c->block == 192 && c->waves == 7) {
        hipLaunchKernelGGL((k_solve_fast<192, 7>), dim3(nb), dim3(192), 0, s, c->kp, q, b0, nb, dt, mode, g);
    } else if (c->block == 192) {
        hipLaunchKernelGGL((k_solve_fast<192, 1>), dim3(nb), dim3(192), 0, s, c->kp, q, b0, nb, dt, mode, g);
    } else if (c->waves == 7) {
        hipLaunchKernelGGL((k_solve_fast<256, 7>), dim3(nb), dim3(256), 0, s, c->kp, q, b0, nb, dt, mode, g);
    } else {
        hipLaunchKernelGGL((k_solve_fast<256, 1>), dim3(nb), dim3(256), 0, s, c->kp, q, b0, nb, dt, mode, g);
    }
    VBD_LAUNCH_CHECK("k_solve_fast");
    return VBD_OK;
}

int32_t launch_spring_rows(VbdContext* c, hipStream_t s, int j0, int j1, const float* q, float dt) {
    if (j1 <= j0) return VBD_OK;
    hipLaunchKernelGGL(k_spring_rows, dim3(blocks_for(j1 - j0, 256)), dim3(256), 0, s, c->kp, q, j0, j1, dt);
    VBD_LAUNCH_CHECK("k_spring_rows");
    return VBD_OK;
}

int32_t launch_contact_rows(VbdContext* c, hipStream_t s, int p_lo, int p_hi, const float* q, float dt) {
    if (!c->crows || !c->have_contacts || p_hi <= p_lo) return VBD_OK;
    hipLaunchKernelGGL(k_contact_rows, dim3(kContactRowsGrid), dim3(256), 0, s, c->kp, q, p_lo, p_hi, dt);
    VBD_LAUNCH_CHECK("k_contact_rows");
    return VBD_OK;
}

int32_t launch_color(VbdContext* c, hipStream_t s, int color, float* q, float dt, int mode, float* g) {
    int32_t rc0 = launch_spring_rows(c, s, c->pre_color_off[color], c->pre_color_off[color + 1], q, dt);
    if (rc0) return rc0;
    rc0 = launch_contact_rows(c, s, c->color_offsets[color], c->color_offsets[color + 1], q, dt);
    if (rc0) return rc0;
    if (c->self_on) {
        hipLaunchKernelGGL(k_sc_rows, dim3(kSelfGrid), dim3(256), 0, s, c->sp, q, c->kp.q_prev,
                           c->color_offsets[color], c->color_offsets[color + 1], dt);
        VBD_LAUNCH_CHECK("k_sc_rows");
    }
    int b0 = c->fast_color_off[color], b1 = c->fast_color_off[color + 1];
    if (c->use_list) {
        b0 = c->live_off[color];
        b1 = c->live_off[color + 1];
    }
    if (b1 > b0) {
        int32_t rc = launch_fast(c, s, b0, b1 - b0, q, dt, mode, g);
        if (rc) return rc;
    }
    int s0 = c->slow_color_off[color], s1 = c->slow_color_off[color + 1];
    if (s1 > s0) {
        hipLaunchKernelGGL(k_solve_generic, dim3(s1 - s0), dim3(kBlock), 0, s, c->kp, q, s0, dt, mode, g);
        VBD_LAUNCH_CHECK("k_solve_generic");
    }
    return VBD_OK;
}
// The force residual assembled by the color-pass kernels' force-only instantiations (one launch
// over every batch) and, for vertices with more than 256 items, the generic kernel.
int32_t residual_by_passes(VbdContext* c, hipStream_t s, const float* q, float dt, float* g) {
    int32_t rc0 = launch_spring_rows(c, s, 0, c->pre_color_off.back(), q, dt);
    if (rc0) return rc0;
    rc0 = launch_contact_rows(c, s, 0, c->model.particle_count, q, dt);
    if (rc0) return rc0;
    if (c->self_on) {
        hipLaunchKernelGGL(k_sc_rows, dim3(kSelfGrid), dim3(256), 0, s, c->sp, q, c->kp.q_prev, 0,
                           c->model.particle_count, dt);
        VBD_LAUNCH_CHECK("k_sc_rows");
    }
    int nb = c->use_list ? c->live_off.back() : c->fast_color_off.back();
    if (nb > 0) {
        int32_t rc = launch_fast(c, s, 0, nb, (float*)q, dt, 1, g, true);
        if (rc) return rc;
    }
    int ns = c->slow_color_off.back();
    if (ns > 0) {
        hipLaunchKernelGGL(k_solve_generic, dim3(ns), dim3(kBlock), 0, s, c->kp, (float*)q, 0, dt, 1, g);
        VBD_LAUNCH_CHECK("k_solve_generic(residual)");
    }
    return VBD_OK;
}
}  // namespace

int32_t vbd_color_pass(VbdContext* c, void* stream, int32_t color, float* q, float dt) {
    if (!c || !q) return fail(VBD_ERR_ARG, "vbd_color_pass: null argument");
    if (color < 0 || color >= c->model.color_count) return fail(VBD_ERR_ARG, "vbd_color_pass: color out of range");
    int32_t rc = launch_color(c, (hipStream_t)stream, color, q, dt, 0, nullptr);
    if (rc == VBD_OK && c->self_on) rc = self_truncate(c, (hipStream_t)stream, q);
    return rc;
}

int32_t vbd_iteration(VbdContext* c, void* stream, float* q, float dt) {
    if (!c || !q) return fail(VBD_ERR_ARG, "vbd_iteration: null argument");
    hipStream_t s = (hipStream_t)stream;
    // self-contact re-detection (solver_vbd.py:2540-2545); displacements restart from the new base
    if (c->self_on && ((c->interval == 0 && c->iter == 0) || (c->interval >= 1 && c->iter % c->interval == 0))) {
        int32_t rc = self_detect(c, s, q);
        if (rc) return rc;
        const int n3 = 3 * c->model.particle_count;
        hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(n3, 256)), dim3(256), 0, s, (int*)c->kp.disp, 0, n3);
        VBD_LAUNCH_CHECK("k_fill_i32");
    }
    if (c->have_contacts && c->params.avbd_beta != 0.0f) {
        const VbdContactsDesc& cd = c->contacts;
        hipLaunchKernelGGL(k_dual_update, dim3(blocks_for(cd.soft_contact_max, 256)), dim3(256), 0, s,
                           cd.soft_contact_count, cd.soft_contact_max, cd.soft_contact_indices,
                           cd.soft_contact_barycentric, cd.soft_contact_shape, cd.soft_contact_body_pos,
                           cd.soft_contact_body_vel, cd.soft_contact_normal, q, c->model.particle_radius,
                           c->model.shape_margin, c->model.shape_count, body_poses(c), cd.dt,
                           c->c_state + c->capacity, c->params.avbd_beta, c->c_state, c->model.particle_count,
                           c->cid_to_pack, c->cpack);
        VBD_LAUNCH_CHECK("k_dual_update");
    }
    for (int color = 0; color < c->model.color_count; ++color) {
        int32_t rc = launch_color(c, s, color, q, dt, 0, nullptr);
        if (rc) return rc;
        if (c->self_on) {
            rc = self_truncate(c, s, q);
            if (rc) return rc;
        }
    }
    ++c->iter;
    return VBD_OK;
}

int32_t vbd_integrate_particles(void* stream, int32_t n, const float* x, const float* v, const float* f,
                                const float* inv_mass, const int32_t* flags, const int32_t* world,
                                const float* gravity, int32_t gravity_count, float dt, float v_max, float* x_new,
                                float* v_new) {
    if (n < 0) return fail(VBD_ERR_ARG, "vbd_integrate_particles: negative particle count");
    if (n == 0) return VBD_OK;
    if (!x || !v || !f || !inv_mass || !flags || !gravity || !x_new || !v_new || gravity_count <= 0)
        return fail(VBD_ERR_ARG, "vbd_integrate_particles: null array or empty gravity table");
    hipLaunchKernelGGL(k_integrate_particles, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, n, x, v,
                       f, inv_mass, flags, world, gravity, gravity_count, dt, v_max, x_new, v_new);
    VBD_LAUNCH_CHECK("k_integrate_particles");
    return VBD_OK;
}

int32_t vbd_finalize(VbdContext* c, void* stream, const float* q, float* q_out, float* qd_out, float dt) {
    if (!c || !q || !q_out || !qd_out) return fail(VBD_ERR_ARG, "vbd_finalize: null argument");
    hipLaunchKernelGGL(k_finalize, dim3(blocks_for(c->model.particle_count, 256)), dim3(256), 0, (hipStream_t)stream,
                       c->kp, q, q_out, qd_out, dt);
    VBD_LAUNCH_CHECK("k_finalize");
    return VBD_OK;
}

int32_t vbd_step(VbdContext* c, void* stream, float* q_in, const float* qd_in, const float* f_ext, float* q_out,
                 float* qd_out, const VbdContactsDesc* contacts, float dt) {
    if (!c) return fail(VBD_ERR_ARG, "vbd_step: null context");
    if (!(dt > 0.0f)) return fail(VBD_ERR_ARG, "vbd_step: dt must be positive");
    VbdContactsDesc cd{};
    if (contacts) {
        cd = *contacts;
        cd.dt = dt;
    }
    int32_t rc = vbd_contact_prepare(c, stream, contacts ? &cd : nullptr);
    if (rc) return rc;
    rc = vbd_forward(c, stream, q_in, qd_in, f_ext, dt);
    if (rc) return rc;
    for (int it = 0; it < c->params.iterations; ++it) {
        rc = vbd_iteration(c, stream, q_in, dt);
        if (rc) return rc;
    }
    return vbd_finalize(c, stream, q_in, q_out, qd_out, dt);
}

int32_t vbd_set_block_shape(VbdContext* c, void* stream, int32_t block) {
    if (!c || (block != 256 && block != 192 && block != 128))
        return fail(VBD_ERR_ARG, "vbd_set_block_shape: block must be 256, 192 or 128");
    if (c->block == block) return VBD_OK;
    c->block = block;
    int32_t rc = build_layout(c);
    if (rc) return rc;
    rc = launch_batch_contacts(c, (hipStream_t)stream);
    if (rc) return rc;
    VBD_HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return VBD_OK;
}

int32_t vbd_set_aligned_items(VbdContext* c, void* stream, int32_t aligned) {
    if (!c || (aligned != 0 && aligned != 1)) return fail(VBD_ERR_ARG, "vbd_set_aligned_items: aligned must be 0 or 1");
    if (c->aligned == aligned) return VBD_OK;
    c->aligned = aligned;
    int32_t rc = build_layout(c);
    if (rc) return rc;
    rc = launch_batch_contacts(c, (hipStream_t)stream);
    if (rc) return rc;
    VBD_HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return VBD_OK;
}

int32_t vbd_set_pass_waves(VbdContext* c, int32_t waves) {
    if (!c || (waves != 1 && waves != 7)) return fail(VBD_ERR_ARG, "vbd_set_pass_waves: waves must be 1 or 7");
    c->waves = waves;
    return VBD_OK;
}

int32_t vbd_set_pass_shape(VbdContext* c, void* stream, int32_t block, int32_t waves, int32_t aligned) {
    if (!c || (block != 256 && block != 192 && block != 128) || (waves != 1 && waves != 7) ||
        (aligned != 0 && aligned != 1))
        return fail(VBD_ERR_ARG, "vbd_set_pass_shape: block must be 256, 192 or 128, waves 1 or 7, aligned 0 or 1");
    c->waves = waves;
    if (c->block == block && c->aligned == aligned) return VBD_OK;
    c->block = block;
    c->aligned = aligned;
    int32_t rc = build_layout(c);
    if (rc) return rc;
    rc = launch_batch_contacts(c, (hipStream_t)stream);
    if (rc) return rc;
    VBD_HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return VBD_OK;
}

int32_t vbd_get_pass_shape(VbdContext* c, int32_t* out3) {
    if (!c || !out3) return fail(VBD_ERR_ARG, "vbd_get_pass_shape: null argument");
    out3[0] = c->block;
    out3[1] = c->waves;
    out3[2] = c->aligned;
    return VBD_OK;
}

int32_t vbd_set_scalars(VbdContext* c, int32_t iterations, float friction_epsilon, float soft_contact_ke,
                        float soft_contact_kd, float soft_contact_mu) {
    if (!c || iterations < 0) return fail(VBD_ERR_ARG, "vbd_set_scalars: null context or negative iterations");
    c->params.iterations = iterations;
    c->params.friction_epsilon = friction_epsilon;
    c->kp.friction_epsilon = friction_epsilon;
    c->model.soft_contact_ke = soft_contact_ke;
    c->model.soft_contact_kd = soft_contact_kd;
    c->model.soft_contact_mu = soft_contact_mu;
    return VBD_OK;
}

int32_t vbd_autotune(VbdContext* c, void* stream, const float* q, float dt, int32_t* block_out, float* ms_out) {
    if (!c || !q || !(dt > 0.0f)) return fail(VBD_ERR_ARG, "vbd_autotune: bad argument");
    hipStream_t s = (hipStream_t)stream;
    constexpr int kCand = 7;
    const int shapes[kCand] = {256, 256, 128, 256, 256, 192, 192};
    const int waves[kCand] = {1, 7, 1, 1, 7, 1, 7};
    const int align[kCand] = {0, 0, 0, 1, 1, 1, 1};
    float ms[kCand] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    hipEvent_t e0, e1;
    VBD_HIP_TRY(hipEventCreate(&e0));
    VBD_HIP_TRY(hipEventCreate(&e1));
    int32_t rc = VBD_OK;
    for (int k = 0; k < kCand && rc == VBD_OK; ++k) {
        c->waves = waves[k];
        if (k == 0 || shapes[k] != shapes[k - 1] || align[k] != align[k - 1]) {
            c->block = shapes[k];
            c->aligned = align[k];
            rc = build_layout(c);
            if (rc) break;
            rc = launch_batch_contacts(c, s);
            if (rc) break;
        }
        // warm, then time three force-assembly passes over every color (q is only read)
        for (int rep = 0; rep < 4 && rc == VBD_OK; ++rep) {
            if (rep == 1 && hipEventRecord(e0, s) != hipSuccess) rc = fail(VBD_ERR_HIP, "vbd_autotune: event");
            for (int color = 0; color < c->model.color_count && rc == VBD_OK; ++color)
                rc = launch_color(c, s, color, (float*)q, dt, 1, c->g_tmp);
        }
        if (rc) break;
        if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms[k], e0, e1) != hipSuccess)
            rc = fail(VBD_ERR_HIP, "vbd_autotune: timing");
        ms[k] /= 3.0f;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) return rc;
    int best = 0;
    for (int k = 1; k < kCand; ++k)
        if (ms[k] < ms[best]) best = k;
    c->waves = waves[best];
    if (c->block != shapes[best] || c->aligned != align[best]) {
        c->block = shapes[best];
        c->aligned = align[best];
        rc = build_layout(c);
        if (rc) return rc;
        rc = launch_batch_contacts(c, s);
        if (rc) return rc;
        VBD_HIP_TRY(hipStreamSynchronize(s));
    }
    if (block_out) {
        block_out[0] = shapes[best];
        block_out[1] = waves[best];
        block_out[2] = align[best];
    }
    if (ms_out)
        for (int k = 0; k < kCand; ++k) ms_out[k] = ms[k];
    return VBD_OK;
}

int32_t vbd_residual(VbdContext* c, void* stream, const float* q, float dt, float* g_out, double* norm_sq_out) {
    if (!c || !q) return fail(VBD_ERR_ARG, "vbd_residual: null argument");
    hipStream_t s = (hipStream_t)stream;
    float* g = g_out ? g_out : c->g_tmp;
    {
        int32_t rc = residual_by_passes(c, s, q, dt, g);
        if (rc) return rc;
    }
    if (norm_sq_out) {
        const int N = c->model.particle_count;
        const int parts = norm_parts(N);
        hipLaunchKernelGGL(k_residual_partial, dim3(parts), dim3(kNormBlock), 0, s, N, g, c->r_part);
        VBD_LAUNCH_CHECK("k_residual_partial");
        hipLaunchKernelGGL(k_residual_final, dim3(1), dim3(kNormBlock), 0, s, parts, c->r_part, norm_sq_out);
        VBD_LAUNCH_CHECK("k_residual_final");
    }
    return VBD_OK;
}



int32_t vbd_set_iterate(VbdContext* c, void* stream, float* q_dst, const float* guess, int32_t commit) {
    if (!c || !q_dst || !guess) return fail(VBD_ERR_ARG, "vbd_set_iterate: null argument");
    if (c->self_on) return fail(VBD_ERR_STATE, "vbd_set_iterate: not available with self-contact");
    hipLaunchKernelGGL(k_set_iterate, dim3(blocks_for(c->model.particle_count, 256)), dim3(256), 0,
                       (hipStream_t)stream, c->kp, q_dst, guess, commit ? 1 : 0, 0, c->model.particle_count);
    VBD_LAUNCH_CHECK("k_set_iterate");
    return VBD_OK;
}

int32_t vbd_set_iterate_range(VbdContext* c, void* stream, float* q_dst, const float* guess, int32_t commit,
                              int32_t p_begin, int32_t p_end) {
    if (!c || !q_dst || !guess) return fail(VBD_ERR_ARG, "vbd_set_iterate_range: null argument");
    if (c->self_on) return fail(VBD_ERR_STATE, "vbd_set_iterate_range: not available with self-contact");
    if (p_begin < 0 || p_end < p_begin || p_end > c->model.particle_count)
        return fail(VBD_ERR_ARG, "vbd_set_iterate_range: need 0 <= p_begin <= p_end <= particle_count");
    if (p_end == p_begin) return VBD_OK;
    hipLaunchKernelGGL(k_set_iterate, dim3(blocks_for(p_end - p_begin, 256)), dim3(256), 0, (hipStream_t)stream,
                       c->kp, q_dst, guess, commit ? 1 : 0, p_begin, p_end);
    VBD_LAUNCH_CHECK("k_set_iterate(range)");
    return VBD_OK;
}

namespace {
// slices per world of the per-world norms: up to 16, within the first-stage scratch (r_part)
int world_split(VbdContext* c, int32_t world_count) {
    const int cap = norm_parts(c->model.particle_count);
    return std::max(1, std::min(16, cap / std::max(world_count, 1)));
}

// the residual and the per-world slice sums (of the live worlds only when `live` is given) in
// c->r_part, world_split slices per world
int32_t residual_world_parts(VbdContext* c, void* stream, const float* q, float dt, int32_t world_count,
                             const int32_t* world_offsets, const int* live) {
    int32_t rc = vbd_residual(c, stream, q, dt, nullptr, nullptr);
    if (rc) return rc;
    const int split = world_split(c, world_count);
    hipLaunchKernelGGL(k_residual_worlds, dim3(world_count * split), dim3(kNormBlock), 0, (hipStream_t)stream,
                       c->model.particle_count, world_offsets, c->g_tmp, split, c->r_part, live);
    VBD_LAUNCH_CHECK("k_residual_worlds");
    return VBD_OK;
}
}  // namespace

int32_t vbd_residual_worlds(VbdContext* c, void* stream, const float* q, float dt, int32_t world_count,
                            const int32_t* world_offsets, double* norm_sq_out) {
    if (!c || !q || !world_offsets || !norm_sq_out) return fail(VBD_ERR_ARG, "vbd_residual_worlds: null argument");
    if (world_count < 1) return fail(VBD_ERR_ARG, "vbd_residual_worlds: world_count must be >= 1");
    int32_t rc = residual_world_parts(c, stream, q, dt, world_count, world_offsets, nullptr);
    if (rc) return rc;
    hipLaunchKernelGGL(k_world_norms, dim3(blocks_for(world_count, 256)), dim3(256), 0, (hipStream_t)stream,
                       world_count, c->r_part, world_split(c, world_count), norm_sq_out);
    VBD_LAUNCH_CHECK("k_world_norms");
    return VBD_OK;
}

namespace {
// Installs a world mask in the context's kernel parameters for the lifetime of one driver call.
struct WorldMask {
    VbdContext* c;
    WorldMask(VbdContext* ctx, const int* p_world, const int* w_active) : c(ctx) {
        c->kp.p_world = p_world;
        c->kp.w_active = w_active;
    }
    ~WorldMask() {
        c->kp.p_world = nullptr;
        c->kp.w_active = nullptr;
        c->kp.batch_list = nullptr;
        c->use_list = false;
    }
};

// Worlds of each packed batch's smallest and largest particle, for the given world offsets.
void batch_worlds(VbdContext* c, const std::vector<int32_t>& off) {
    if (c->bw_key == off) return;
    const size_t nb = c->bat_pmin.size();
    c->bat_wlo.resize(nb);
    c->bat_whi.resize(nb);
    auto world_of = [&](int p) {
        // last w with off[w] <= p (clipped to the world range)
        int w = (int)(std::upper_bound(off.begin(), off.end(), p) - off.begin()) - 1;
        return std::min(std::max(w, 0), (int)off.size() - 2);
    };
    for (size_t b = 0; b < nb; ++b) {
        c->bat_wlo[b] = world_of(c->bat_pmin[b]);
        c->bat_whi[b] = world_of(c->bat_pmax[b]);
    }
    c->bw_key = off;
}

// Rebuilds the live-batch list (batches with a vertex in a live world, per color) and uploads it.
int32_t build_live_list(VbdContext* c, hipStream_t s, const std::vector<int32_t>& live) {
    const int C = c->model.color_count;
    c->live_off.assign((size_t)C + 1, 0);
    c->h_batch_list.clear();
    std::vector<int32_t> pre((size_t)live.size() + 1, 0);  // prefix counts of live worlds
    for (size_t w = 0; w < live.size(); ++w) pre[w + 1] = pre[w] + (live[w] ? 1 : 0);
    for (int col = 0; col < C; ++col) {
        c->live_off[col] = (int32_t)c->h_batch_list.size();
        for (int b = c->fast_color_off[col]; b < c->fast_color_off[col + 1]; ++b)
            if (pre[c->bat_whi[b] + 1] - pre[c->bat_wlo[b]] > 0) c->h_batch_list.push_back(b);
    }
    c->live_off[C] = (int32_t)c->h_batch_list.size();
    if (!c->h_batch_list.empty())
        VBD_HIP_TRY(hipMemcpyAsync(c->batch_list, c->h_batch_list.data(), sizeof(int32_t) * c->h_batch_list.size(),
                                   hipMemcpyHostToDevice, s));
    c->kp.batch_list = c->batch_list;
    c->use_list = true;
    return VBD_OK;
}
}  // namespace

// Adds each live world's norm slices (k_residual_worlds) into norm[w] and freezes the worlds whose
// squared norm reached their target after `done` iterations (one work-group; st = [active W |
// iterations W | live count]).  Frozen worlds stay frozen.  Once at
// most stop_at_live (> 0) worlds are live, the rest are frozen too (the driver stops there).
__global__ __launch_bounds__(1024) void k_world_freeze(int nw, const double* __restrict__ part, int split,
                                                      double* __restrict__ norm,
                                                      const double* __restrict__ target, int* __restrict__ st,
                                                      int done, int stop_at_live) {
    __shared__ int s_live[1024 / 64];
    __shared__ int s_total;
    int live = 0;
    for (int w = threadIdx.x; w < nw; w += blockDim.x) {
        int a = st[w];
        if (a) norm[w] = world_norm(part, split, w);  // a frozen world keeps its norm
        if (a && norm[w] <= target[w]) {
            a = 0;
            st[w] = 0;
            st[nw + w] = done;
        }
        live += a;
    }
    for (int off = 32; off > 0; off >>= 1) live += __shfl_down(live, off, 64);
    if ((threadIdx.x & 63) == 0) s_live[threadIdx.x >> 6] = live;
    __syncthreads();
    if (threadIdx.x == 0) {
        int sum = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) sum += s_live[k];
        s_total = sum;
    }
    __syncthreads();
    const int total = s_total;
    const bool stop_all = stop_at_live > 0 && total <= stop_at_live;
    if (stop_all) {
        for (int w = threadIdx.x; w < nw; w += blockDim.x)
            if (st[w]) {
                st[w] = 0;
                st[nw + w] = done;
            }
    }
    if (threadIdx.x == 0) st[2 * nw] = stop_all ? 0 : total;
}

__global__ void k_world_init(int nw, int* __restrict__ st) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nw) {
        st[i] = 1;
        st[nw + i] = 0;
    }
    if (i == 0) st[2 * nw] = nw;
}

// Iterations between the host's polls of the live-world count; the freezing itself is decided
// on the device after every evaluation, so the poll interval only bounds the idle tail.
constexpr int kWorldPoll = 8;

int32_t vbd_iterate_to_residual_worlds(VbdContext* c, void* stream, float* q, float dt, int32_t world_count,
                                       const int32_t* world_offsets, const int32_t* particle_world,
                                       const double* target_norm_sq, int32_t max_iterations, int32_t check_every,
                                       int32_t stop_at_live, double* world_scratch_dev, int32_t* world_state_dev,
                                       int32_t* iterations_out, int32_t* world_iterations_out, double* norm_sq_out) {
    if (!c || !q || !world_offsets || !particle_world || !target_norm_sq || !world_scratch_dev || !world_state_dev ||
        !norm_sq_out)
        return fail(VBD_ERR_ARG, "vbd_iterate_to_residual_worlds: null argument");
    if (stop_at_live < 0 || stop_at_live >= world_count)
        return fail(VBD_ERR_ARG, "vbd_iterate_to_residual_worlds: need 0 <= stop_at_live < world_count");
    if (world_count < 1 || max_iterations < 0 || check_every < 1)
        return fail(VBD_ERR_ARG,
                    "vbd_iterate_to_residual_worlds: world_count and check_every must be >= 1, max_iterations >= 0");
    if (c->self_on) return fail(VBD_ERR_STATE, "vbd_iterate_to_residual_worlds: not available with self-contact");
    hipStream_t s = (hipStream_t)stream;
    const int W = world_count;
    double* norm = world_scratch_dev;
    double* target = world_scratch_dev + W;
    VBD_HIP_TRY(hipMemcpyAsync(target, target_norm_sq, sizeof(double) * (size_t)W, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_world_init, dim3(blocks_for(W, 256)), dim3(256), 0, s, W, world_state_dev);
    VBD_LAUNCH_CHECK("k_world_init");
    WorldMask mask(c, particle_world, world_state_dev);
    {
        std::vector<int32_t> off((size_t)W + 1);
        VBD_HIP_TRY(hipMemcpyAsync(off.data(), world_offsets, sizeof(int32_t) * (size_t)(W + 1),
                                   hipMemcpyDeviceToHost, s));
        VBD_HIP_TRY(hipStreamSynchronize(s));
        for (int w = 0; w < W; ++w)
            if (off[w] < 0 || off[w + 1] < off[w] || off[w + 1] > c->model.particle_count)
                return fail(VBD_ERR_ARG, "vbd_iterate_to_residual_worlds: world_offsets must ascend within [0, N]");
        batch_worlds(c, off);
    }
    std::vector<int32_t> st_host((size_t)(2 * W + 1), 1), live_prev((size_t)W, 1);
    int32_t done = 0, live = W;
    const int split = world_split(c, W);
    int32_t rc = residual_world_parts(c, stream, q, dt, W, world_offsets, world_state_dev);
    if (rc) return rc;
    hipLaunchKernelGGL(k_world_freeze, dim3(1), dim3(1024), 0, s, W, c->r_part, split, norm, target, world_state_dev,
                       0, stop_at_live);
    VBD_LAUNCH_CHECK("k_world_freeze");
    const int32_t poll_every = kWorldPoll;
    int32_t since_poll = poll_every;  // poll before the first iteration
    for (;;) {
        if (since_poll >= poll_every || done >= max_iterations) {
            // the live count through pinned memory; the flags only when it changed
            VBD_HIP_TRY(hipMemcpyAsync(c->w_host, world_state_dev + 2 * W, sizeof(int32_t), hipMemcpyDeviceToHost, s));
            VBD_HIP_TRY(hipStreamSynchronize(s));
            since_poll = 0;
            const int32_t now_live = c->w_host[0];
            if (now_live <= stop_at_live || done >= max_iterations) {
                live = now_live;
                break;
            }
            if (now_live != live) {
                live = now_live;
                int32_t* flags = W <= kWorldPinned ? c->w_host : st_host.data();
                VBD_HIP_TRY(hipMemcpyAsync(flags, world_state_dev, sizeof(int32_t) * (size_t)W, hipMemcpyDeviceToHost,
                                           s));
                VBD_HIP_TRY(hipStreamSynchronize(s));
                // launch only the batches of worlds still live (frozen ones inside them are masked)
                std::vector<int32_t> now(flags, flags + W);
                live_prev = now;
                rc = build_live_list(c, s, now);
                if (rc) return rc;
            }
        }
        const int32_t n = std::min(check_every, max_iterations - done);
        for (int32_t k = 0; k < n; ++k) {
            rc = vbd_iteration(c, stream, q, dt);
            if (rc) return rc;
        }
        done += n;
        rc = residual_world_parts(c, stream, q, dt, W, world_offsets, world_state_dev);
        if (rc) return rc;
        hipLaunchKernelGGL(k_world_freeze, dim3(1), dim3(1024), 0, s, W, c->r_part, split, norm, target,
                           world_state_dev, done,
                           stop_at_live);
        VBD_LAUNCH_CHECK("k_world_freeze");
        ++since_poll;
    }
    VBD_HIP_TRY(hipMemcpyAsync(norm_sq_out, norm, sizeof(double) * (size_t)W, hipMemcpyDeviceToHost, s));
    std::vector<int32_t> st((size_t)(2 * W));
    VBD_HIP_TRY(hipMemcpyAsync(st.data(), world_state_dev, sizeof(int32_t) * (size_t)(2 * W), hipMemcpyDeviceToHost, s));
    VBD_HIP_TRY(hipStreamSynchronize(s));
    if (world_iterations_out)
        for (int w = 0; w < W; ++w) world_iterations_out[w] = st[w] ? done : st[W + w];
    if (iterations_out) *iterations_out = done;
    return VBD_OK;
}

// Single-world freezing: st = [active, iterations, live]; the substep's passes see it through the
// world mask (every particle in world 0), so once the residual is within the target every later
// pass exits at once.
__global__ void k_single_freeze(const double* __restrict__ norm, double target, int* __restrict__ st, int done) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        if (st[0] && norm[0] <= target) {
            st[0] = 0;
            st[1] = done;
        }
        st[2] = st[0];
    }
}

__global__ void k_single_init(int* __restrict__ st) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        st[0] = 1;
        st[1] = 0;
        st[2] = 1;
    }
}

int32_t vbd_iterate_to_residual(VbdContext* c, void* stream, float* q, float dt, double target_norm_sq,
                                int32_t max_iterations, int32_t check_every, int32_t* iterations_out,
                                double* norm_sq_out) {
    if (!c || !q) return fail(VBD_ERR_ARG, "vbd_iterate_to_residual: null argument");
    if (max_iterations < 0 || check_every < 1)
        return fail(VBD_ERR_ARG, "vbd_iterate_to_residual: max_iterations must be >= 0 and check_every >= 1");
    hipStream_t s = (hipStream_t)stream;
    if (c->self_on) {
        // self-contact truncation rewrites every q after each color pass: no freezing, a host check
        // after every check_every iterations
        int32_t done = 0;
        double r = 0.0;
        for (;;) {
            int32_t rc = vbd_residual(c, stream, q, dt, nullptr, c->r_dev);
            if (rc) return rc;
            VBD_HIP_TRY(hipMemcpyAsync(c->r_host, c->r_dev, sizeof(double), hipMemcpyDeviceToHost, s));
            VBD_HIP_TRY(hipStreamSynchronize(s));
            r = *c->r_host;
            if (r <= target_norm_sq || done >= max_iterations) break;
            int32_t n = std::min(check_every, max_iterations - done);
            for (int32_t k = 0; k < n; ++k) {
                rc = vbd_iteration(c, stream, q, dt);
                if (rc) return rc;
            }
            done += n;
        }
        if (iterations_out) *iterations_out = done;
        if (norm_sq_out) *norm_sq_out = r;
        return VBD_OK;
    }
    // The stop is decided on the device after every evaluation (the substep freezes there: its passes
    // exit at once); the host reads the state every kWorldPoll evaluations, so the iterate and the
    // count are those of a host check after every evaluation, with one round trip per poll.
    hipLaunchKernelGGL(k_single_init, dim3(1), dim3(64), 0, s, c->one_state);
    VBD_LAUNCH_CHECK("k_single_init");
    WorldMask mask(c, c->zero_world, c->one_state);
    int32_t rc = vbd_residual(c, stream, q, dt, nullptr, c->r_dev);
    if (rc) return rc;
    hipLaunchKernelGGL(k_single_freeze, dim3(1), dim3(64), 0, s, c->r_dev, target_norm_sq, c->one_state, 0);
    VBD_LAUNCH_CHECK("k_single_freeze");
    int32_t done = 0, since_poll = kWorldPoll;
    for (;;) {
        if (since_poll >= kWorldPoll || done >= max_iterations) {
            VBD_HIP_TRY(hipMemcpyAsync(c->w_host, c->one_state, sizeof(int32_t) * 3, hipMemcpyDeviceToHost, s));
            VBD_HIP_TRY(hipStreamSynchronize(s));
            since_poll = 0;
            if (c->w_host[2] == 0 || done >= max_iterations) break;
        }
        const int32_t n = std::min(check_every, max_iterations - done);
        for (int32_t k = 0; k < n; ++k) {
            rc = vbd_iteration(c, stream, q, dt);
            if (rc) return rc;
        }
        done += n;
        rc = vbd_residual(c, stream, q, dt, nullptr, c->r_dev);
        if (rc) return rc;
        hipLaunchKernelGGL(k_single_freeze, dim3(1), dim3(64), 0, s, c->r_dev, target_norm_sq, c->one_state, done);
        VBD_LAUNCH_CHECK("k_single_freeze");
        ++since_poll;
    }
    VBD_HIP_TRY(hipMemcpyAsync(c->r_host, c->r_dev, sizeof(double), hipMemcpyDeviceToHost, s));
    VBD_HIP_TRY(hipStreamSynchronize(s));
    if (iterations_out) *iterations_out = c->w_host[0] ? done : c->w_host[1];
    if (norm_sq_out) *norm_sq_out = *c->r_host;
    return VBD_OK;
}

namespace {

// Checks a collision descriptor and fills the kernels' parameters (body poses are set per call).
int32_t collide_params(const VbdCollideDesc* d, const char* who, ColParams& cp) {
    if (d->pair_count < 0 || d->edge_pair_count < 0 || d->face_pair_count < 0)
        return fail(VBD_ERR_ARG, std::string(who) + ": negative pair count");
    const long total = (long)d->pair_count + d->edge_pair_count + d->face_pair_count;
    if (total > INT32_MAX) return fail(VBD_ERR_ARG, std::string(who) + ": more than 2^31 - 1 candidate pairs");
    if (total > 0 && (!d->particle_radius || !d->shape_type || !d->shape_flags || !d->shape_world ||
                      !d->shape_transform || !d->shape_scale || !d->shape_margin))
        return fail(VBD_ERR_ARG, std::string(who) + ": missing arrays");
    if (d->pair_count > 0 && (!d->pairs || !d->particle_flags || !d->particle_world))
        return fail(VBD_ERR_ARG, std::string(who) + ": missing particle-pair arrays");
    if (d->edge_pair_count > 0 && (!d->edge_pairs || !d->edge_indices))
        return fail(VBD_ERR_ARG, std::string(who) + ": edge pairs given without edge_pairs / edge_indices");
    if (d->face_pair_count > 0 && (!d->face_pairs || !d->tri_indices))
        return fail(VBD_ERR_ARG, std::string(who) + ": face pairs given without face_pairs / tri_indices");
    cp = ColParams{};
    cp.P = d->pair_count;
    cp.E = d->edge_pair_count;
    cp.F = d->face_pair_count;
    cp.pairs = d->pairs;
    cp.epairs = d->edge_pairs;
    cp.fpairs = d->face_pairs;
    cp.edge_idx = d->edge_indices;
    cp.tri_idx = d->tri_indices;
    cp.radius = d->particle_radius;
    cp.pflags = d->particle_flags;
    cp.pworld = d->particle_world;
    cp.stype = d->shape_type;
    cp.sflags = d->shape_flags;
    cp.sworld = d->shape_world;
    cp.sbody = d->shape_body;
    cp.sxf = d->shape_transform;
    cp.sscale = d->shape_scale;
    cp.smargin = d->shape_margin;
    return VBD_OK;
}

bool contacts_out_ok(const VbdContactsOut* out) {
    return out->soft_contact_indices && out->soft_contact_shape && out->soft_contact_body_pos &&
           out->soft_contact_body_vel && out->soft_contact_normal;
}

int32_t collide_run(VbdCollide* col, hipStream_t s, const float* q, const float* body_q, int32_t body_count,
                    float margin, const VbdContactsOut* out, const char* who) {
    const long total = (long)col->cp.P + col->cp.E + col->cp.F;
    if (total > 0 && !contacts_out_ok(out)) return fail(VBD_ERR_ARG, std::string(who) + ": missing output arrays");
    if (col->has_bodies && (!body_q || body_count < col->max_body + 1))
        return fail(VBD_ERR_ARG, std::string(who) + ": shapes are attached to bodies; body_q must hold at least " +
                                     std::to_string(col->max_body + 1) + " poses");
    if (total == 0 || out->soft_contact_max <= 0) {
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, out->soft_contact_count, 0, 1);
        VBD_LAUNCH_CHECK("k_fill_i32");
        return VBD_OK;
    }
    ColParams cp = col->cp;
    cp.body_q = body_q;
    cp.body_count = body_count;
    return launch_collide_tiles(cp, s, q, margin, col->tiles, *out);
}

}  // namespace

int32_t vbd_collide_create(const VbdCollideDesc* d, VbdCollide** out) {
    if (!d || !out) return fail(VBD_ERR_ARG, "vbd_collide_create: null argument");
    *out = nullptr;
    ColParams cp;
    int32_t rc = collide_params(d, "vbd_collide_create", cp);
    if (rc) return rc;
    VbdCollide* col = new VbdCollide();
    col->desc = *d;
    col->cp = cp;
    if (d->shape_body && d->shape_count > 0) {
        std::vector<int32_t> body;
        std::string err = d2h(body, d->shape_body, (size_t)d->shape_count);
        if (!err.empty()) {
            vbd_collide_destroy(col);
            return fail(VBD_ERR_HIP, "vbd_collide_create: shape_body: " + err);
        }
        for (int32_t b : body) col->max_body = std::max(col->max_body, b);
        col->has_bodies = col->max_body >= 0;
    }
    const long total = (long)cp.P + cp.E + cp.F;
    const int ntiles = std::max(blocks_for(total, kColTile), 1);
    hipError_t e = hipMalloc(&col->tiles, sizeof(int) * (size_t)ntiles);
    if (e == hipSuccess && cp.E + cp.F > 0) {
        e = hipMalloc(&col->ef_cache, sizeof(float4) * 3 * ((size_t)cp.E + cp.F));
        col->cp.ef_cache = col->ef_cache;
    }
    if (e != hipSuccess) {
        std::string msg = std::string("vbd_collide_create: ") + hipGetErrorString(e);
        vbd_collide_destroy(col);
        return fail(VBD_ERR_HIP, msg);
    }
    *out = col;
    return VBD_OK;
}

int32_t vbd_collide_destroy(VbdCollide* col) {
    if (!col) return VBD_OK;
    hipFree(col->tiles);
    hipFree(col->ef_cache);
    delete col;
    return VBD_OK;
}

int32_t vbd_collide_generate(VbdCollide* col, void* stream, const float* q, float margin, const VbdContactsOut* out) {
    if (!col || !q || !out || !out->soft_contact_count) return fail(VBD_ERR_ARG, "vbd_collide_generate: null argument");
    return collide_run(col, (hipStream_t)stream, q, nullptr, 0, margin, out, "vbd_collide_generate");
}

int32_t vbd_collide_generate_bodies(VbdCollide* col, void* stream, const float* q, const float* body_q,
                                    int32_t body_count, float margin, const VbdContactsOut* out) {
    if (!col || !q || !out || !out->soft_contact_count)
        return fail(VBD_ERR_ARG, "vbd_collide_generate_bodies: null argument");
    if (body_count < 0) return fail(VBD_ERR_ARG, "vbd_collide_generate_bodies: negative body_count");
    return collide_run(col, (hipStream_t)stream, q, body_count > 0 ? body_q : nullptr, body_count, margin, out,
                       "vbd_collide_generate_bodies");
}

int32_t vbd_collide_generate_noalloc(const VbdCollideDesc* d, void* stream, const float* q, float margin,
                                     const VbdContactsOut* out) {
    if (!d || !q || !out || !out->soft_contact_count)
        return fail(VBD_ERR_ARG, "vbd_collide_generate_noalloc: null argument");
    // the size limit first: it is the documented rejection, whatever else the descriptor holds
    const long total = (long)d->pair_count + d->edge_pair_count + d->face_pair_count;
    if (total > VBD_NOALLOC_MAX_PAIRS)
        return fail(VBD_ERR_ARG, "vbd_collide_generate_noalloc: the candidate pair count " + std::to_string(total) +
                                     " exceeds the maximum of " + std::to_string(VBD_NOALLOC_MAX_PAIRS) +
                                     " (use vbd_collide_create / vbd_collide_generate)");
    ColParams cp;
    int32_t rc = collide_params(d, "vbd_collide_generate_noalloc", cp);
    if (rc) return rc;
    if (d->shape_body)
        return fail(VBD_ERR_ARG, "vbd_collide_generate_noalloc: shapes are treated as static here (shape_body must be "
                                 "NULL; use vbd_collide_generate_bodies for body-attached shapes)");
    if (total > 0 && !contacts_out_ok(out))
        return fail(VBD_ERR_ARG, "vbd_collide_generate_noalloc: missing output arrays");
    hipStream_t s = (hipStream_t)stream;
    if (total == 0 || out->soft_contact_max <= 0) {
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, out->soft_contact_count, 0, 1);
        VBD_LAUNCH_CHECK("k_fill_i32");
        return VBD_OK;
    }
    int32_t* tiles_dev = nullptr;
    std::string err = noalloc_scratch(&tiles_dev);
    if (!err.empty()) return fail(VBD_ERR_HIP, "vbd_collide_generate_noalloc: " + err);
    return launch_collide_tiles(cp, s, q, margin, tiles_dev, *out);
}

}  // extern "C"
