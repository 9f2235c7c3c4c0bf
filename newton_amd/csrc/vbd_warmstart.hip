// libvbd_ws.so -- learned warm start of the VBD iteration: 1-ring features and
// a BF16 MLP (48 -> 128 -> 128 -> 128 -> 3, SiLU) with forward, backward and
// Adam, written for gfx950's MFMA units.  C ABI in include/vbd_warmstart.h.
//
// Matrix products use v_mfma_f32_16x16x16_bf16 (16x16 output tile, K = 16,
// FP32 accumulation).  Operand layout per 64-lane wave, with li = lane % 16 and
// lq = lane / 16:
//   A (16 x 16, M x K): lane holds A[li][4*lq + 0..3]
//   B (16 x 16, K x N): lane holds B[4*lq + 0..3][li]
//   D (16 x 16, M x N): lane holds D[4*lq + 0..3][li]
// Each wave owns 16 rows (particles); a workgroup of 4 waves owns 64 rows.
// Hidden activations of a wave live in LDS as BF16 rows, so the next layer's
// A operand is one aligned 8-byte LDS read per lane.  Weights are read from
// global memory (L2-resident, 160 KB) in pre-packed layouts whose per-lane
// operand is one aligned 8-byte load: W^T for the forward pass, W for dH = dZ W^T.
//
// BF16 rounding points (restated by the CPU oracle, oracle/mlp_oracle.py):
// features; H_l = bf16(silu(Z_l)); dZ_l = bf16(dH_l * silu'(Z_l)); dZ_4 =
// bf16(dY).  Z, the output, dH (accumulators), bias sums and all gradients
// are FP32.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <algorithm>
#include <string>

#include "vbd_warmstart.h"

namespace {

constexpr int kIn = VBD_WS_IN;          // 48
constexpr int kHid = VBD_WS_HIDDEN;     // 128
constexpr int kOut = VBD_WS_OUT;        // 3
constexpr int kOutP = 16;               // last layer padded to one MFMA tile
constexpr int kWaveRows = 16;
constexpr int kWaves = 4;
constexpr int kBlock = 64 * kWaves;
constexpr int kRowsWG = kWaveRows * kWaves;  // 64 rows per workgroup
constexpr int kStride = kHid + 8;            // BF16 elements per LDS activation row (bank skew)
constexpr int kStrideIn = kIn + 8;
constexpr int kStrideOut = kOutP + 8;

// FP32 parameter offsets
constexpr int oW1 = 0;
constexpr int ob1 = oW1 + kIn * kHid;
constexpr int oW2 = ob1 + kHid;
constexpr int ob2 = oW2 + kHid * kHid;
constexpr int oW3 = ob2 + kHid;
constexpr int ob3 = oW3 + kHid * kHid;
constexpr int oW4 = ob3 + kHid;
constexpr int ob4 = oW4 + kHid * kOut;
constexpr int kParams = ob4 + kOut;  // 39683

// packed BF16 offsets (elements)
constexpr int pWt1 = 0;                       // [128][48]
constexpr int pW1 = pWt1 + kHid * kIn;        // [48][128]
constexpr int pWt2 = pW1 + kIn * kHid;        // [128][128]
constexpr int pW2 = pWt2 + kHid * kHid;
constexpr int pWt3 = pW2 + kHid * kHid;
constexpr int pW3 = pWt3 + kHid * kHid;
constexpr int pWt4 = pW3 + kHid * kHid;       // [16][128], rows >= 3 zero
constexpr int pW4 = pWt4 + kOutP * kHid;      // [128][16], cols >= 3 zero
constexpr int kPacked = pW4 + kHid * kOutP;   // 81920

constexpr int kFeatBlock = 256;

thread_local std::string g_err;

int32_t fail(int32_t code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define WS_LAUNCH_CHECK(what)                                                                     \
    do {                                                                                          \
        hipError_t e_ = hipGetLastError();                                                        \
        if (e_ != hipSuccess) return fail(-2, std::string(what) + ": " + hipGetErrorString(e_)); \
    } while (0)

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __host__ inline uint16_t f2bf(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);  // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ inline float bf2f(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

__device__ inline float sigmoid(float z) { return 1.0f / (1.0f + expf(-z)); }

__device__ inline floatx4 mfma(bf16x4 a, bf16x4 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ inline bf16x4 ld4(const uint16_t* p) { return *reinterpret_cast<const bf16x4*>(p); }

__device__ inline bf16x4 ld4_strided(const uint16_t* p, int stride) {
    bf16x4 r;
    r[0] = (short)p[0];
    r[1] = (short)p[stride];
    r[2] = (short)p[2 * stride];
    r[3] = (short)p[3 * stride];
    return r;
}

__device__ inline floatx4 zero4() { return floatx4{0.0f, 0.0f, 0.0f, 0.0f}; }

// ---- features ----------------------------------------------------------------
// Columns (a = (y - x0) / s_disp, c = corr / s_corr, over 1-ring neighbours j):
//  0-2 a          3-5 c          6-8 mean(a_j) - a    9-11 mean(c_j) - c
//  12-14 max(a_j) - a            15-17 min(a_j) - a   18-20 max(c_j) - c   21-23 min(c_j) - c
//  24-26 mean(x0_j - x0_i) / s_len   27-29 mean, max, min |x0_j - x0_i| / s_len
//  30-32 acc * dt^2 / s_disp     33 |a|  34 |c|  35 mean |c_j|  36 max |c_j|
//  37 mass / mean mass   38 area / mean area   39 pinned fraction of neighbours   40 degree / 6
//  41-43 mean(y_j - y_i) / s_len   44-46 sum over pinned j of (x0_j - x0_i) / (degree * s_len)   47 1
// A particle without neighbours gets its own values for the neighbour statistics.
__global__ __launch_bounds__(kFeatBlock) void k_features(int n, const int* __restrict__ nbr_off,
                                                        const int* __restrict__ nbr, const float* __restrict__ x0,
                                                        const float* __restrict__ y, const float* __restrict__ corr,
                                                        const float* __restrict__ acc, const float* __restrict__ mass,
                                                        const float* __restrict__ area, float dt,
                                                        const float* __restrict__ scales, uint16_t* __restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float sd = scales[0], sc = scales[1], sl = scales[2], mm = scales[3], ma = scales[4];
    float xi[3], yi[3], a[3], c[3];
    for (int k = 0; k < 3; ++k) {
        xi[k] = x0[3 * i + k];
        yi[k] = y[3 * i + k];
        a[k] = (yi[k] - xi[k]) / sd;
        c[k] = corr ? corr[3 * i + k] / sc : 0.0f;
    }
    float sa[3] = {0, 0, 0}, sc3[3] = {0, 0, 0}, mxa[3], mna[3], mxc[3], mnc[3];
    float sdx[3] = {0, 0, 0}, sdy[3] = {0, 0, 0}, spin[3] = {0, 0, 0};
    float slen = 0.0f, mxl = 0.0f, mnl = 0.0f, sabsc = 0.0f, mxabsc = 0.0f;
    int npin = 0;
    const int j0 = nbr_off[i], j1 = nbr_off[i + 1], deg = j1 - j0;
    for (int k = 0; k < 3; ++k) {
        mxa[k] = mna[k] = a[k];
        mxc[k] = mnc[k] = c[k];
    }
    for (int jj = j0; jj < j1; ++jj) {
        int j = nbr[jj];
        float aj[3], cj[3], dx[3];
        for (int k = 0; k < 3; ++k) {
            float xj = x0[3 * j + k];
            aj[k] = (y[3 * j + k] - xj) / sd;
            cj[k] = corr ? corr[3 * j + k] / sc : 0.0f;
            dx[k] = xj - xi[k];
            sdy[k] += y[3 * j + k] - yi[k];
        }
        float len = sqrtf(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]) / sl;
        float ac = sqrtf(cj[0] * cj[0] + cj[1] * cj[1] + cj[2] * cj[2]);
        if (jj == j0) {
            for (int k = 0; k < 3; ++k) {
                mxa[k] = mna[k] = aj[k];
                mxc[k] = mnc[k] = cj[k];
            }
            mxl = mnl = len;
            mxabsc = ac;
        } else {
            for (int k = 0; k < 3; ++k) {
                mxa[k] = fmaxf(mxa[k], aj[k]);
                mna[k] = fminf(mna[k], aj[k]);
                mxc[k] = fmaxf(mxc[k], cj[k]);
                mnc[k] = fminf(mnc[k], cj[k]);
            }
            mxl = fmaxf(mxl, len);
            mnl = fminf(mnl, len);
            mxabsc = fmaxf(mxabsc, ac);
        }
        for (int k = 0; k < 3; ++k) {
            sa[k] += aj[k];
            sc3[k] += cj[k];
            sdx[k] += dx[k];
        }
        slen += len;
        sabsc += ac;
        if (mass[j] == 0.0f) {
            ++npin;
            for (int k = 0; k < 3; ++k) spin[k] += dx[k];
        }
    }
    float f[kIn];
    const float inv = deg > 0 ? 1.0f / (float)deg : 0.0f;
    const float absa = sqrtf(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    const float absc = sqrtf(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    for (int k = 0; k < 3; ++k) {
        f[0 + k] = a[k];
        f[3 + k] = c[k];
        f[6 + k] = deg > 0 ? sa[k] * inv - a[k] : 0.0f;
        f[9 + k] = deg > 0 ? sc3[k] * inv - c[k] : 0.0f;
        f[12 + k] = mxa[k] - a[k];
        f[15 + k] = mna[k] - a[k];
        f[18 + k] = mxc[k] - c[k];
        f[21 + k] = mnc[k] - c[k];
        f[24 + k] = sdx[k] * inv / sl;
        f[30 + k] = acc ? acc[3 * i + k] * dt * dt / sd : 0.0f;
        f[41 + k] = sdy[k] * inv / sl;
        f[44 + k] = spin[k] * inv / sl;
    }
    f[27] = slen * inv;
    f[28] = mxl;
    f[29] = mnl;
    f[33] = absa;
    f[34] = absc;
    f[35] = deg > 0 ? sabsc * inv : absc;
    f[36] = deg > 0 ? mxabsc : absc;
    f[37] = mass[i] / mm;
    f[38] = area[i] / ma;
    f[39] = (float)npin * inv;
    f[40] = (float)deg / 6.0f;
    f[47] = 1.0f;
    uint16_t* o = out + (size_t)i * kIn;
    for (int k = 0; k < kIn; k += 4) {
        bf16x4 v;
        v[0] = (short)f2bf(f[k]);
        v[1] = (short)f2bf(f[k + 1]);
        v[2] = (short)f2bf(f[k + 2]);
        v[3] = (short)f2bf(f[k + 3]);
        *reinterpret_cast<bf16x4*>(o + k) = v;
    }
}

// ---- weight packing ------------------------------------------------------------
// ---- low-pass of correction fields --------------------------------------------
// One sweep of the 1-ring mean with the particle itself, for `width` floats per particle (e.g. two
// [n,3] fields side by side): out_i = (v_i + sum_j v_j) / (1 + degree_i) over solved particles
// (inactive ones read and write 0).  Repeated sweeps damp rough, neighbour-to-neighbour components
// of a predicted correction fast and its smooth ones slowly.
constexpr int kSmoothMaxWidth = 12;

__global__ __launch_bounds__(256) void k_smooth(int n, int width, const int* __restrict__ nbr_off,
                                                const int* __restrict__ nbr, const float* __restrict__ active,
                                                const float* __restrict__ in, float* __restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float* o = out + (size_t)width * i;
    if (active[i] == 0.0f) {
        for (int c = 0; c < width; ++c) o[c] = 0.0f;
        return;
    }
    float acc[kSmoothMaxWidth];
    const float* vi = in + (size_t)width * i;
    for (int c = 0; c < width; ++c) acc[c] = vi[c];
    const int b = nbr_off[i], e = nbr_off[i + 1];
    for (int k = b; k < e; ++k) {
        const int j = nbr[k];
        const float w = active[j];
        const float* vj = in + (size_t)width * j;
        for (int c = 0; c < width; ++c) acc[c] += w * vj[c];
    }
    const float inv = 1.0f / (float)(1 + e - b);
    for (int c = 0; c < width; ++c) o[c] = acc[c] * inv;
}

__global__ void k_pack(const float* __restrict__ prm, uint16_t* __restrict__ pk) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= kPacked) return;
    float v = 0.0f;
    if (e < pW1) {  // Wt1[j][k] = W1[k][j]
        int j = (e - pWt1) / kIn, k = (e - pWt1) % kIn;
        v = prm[oW1 + k * kHid + j];
    } else if (e < pWt2) {
        v = prm[oW1 + (e - pW1)];
    } else if (e < pW2) {
        int j = (e - pWt2) / kHid, k = (e - pWt2) % kHid;
        v = prm[oW2 + k * kHid + j];
    } else if (e < pWt3) {
        v = prm[oW2 + (e - pW2)];
    } else if (e < pW3) {
        int j = (e - pWt3) / kHid, k = (e - pWt3) % kHid;
        v = prm[oW3 + k * kHid + j];
    } else if (e < pWt4) {
        v = prm[oW3 + (e - pW3)];
    } else if (e < pW4) {  // Wt4[j][k] = W4[k][j], j < 3
        int j = (e - pWt4) / kHid, k = (e - pWt4) % kHid;
        v = j < kOut ? prm[oW4 + k * kOut + j] : 0.0f;
    } else {  // W4p[k][j], j < 3
        int k = (e - pW4) / kOutP, j = (e - pW4) % kOutP;
        v = j < kOut ? prm[oW4 + k * kOut + j] : 0.0f;
    }
    pk[e] = f2bf(v);
}

// ---- forward ---------------------------------------------------------------------
// One hidden layer for the wave's 16 rows: acc = A (16 x K) * Wt^T, A read from
// LDS rows (stride kStride) or, for the first layer, from the feature rows.
template <int K>
__device__ inline void layer_mm(const uint16_t* a_base, int a_stride, const uint16_t* __restrict__ wt, int li, int lq,
                                floatx4 (&acc)[8], bool a_valid) {
#pragma unroll
    for (int nt = 0; nt < 8; ++nt) acc[nt] = zero4();
#pragma unroll
    for (int kt = 0; kt < K / 16; ++kt) {
        bf16x4 a = a_valid ? ld4(a_base + li * a_stride + kt * 16 + lq * 4) : bf16x4{0, 0, 0, 0};
#pragma unroll
        for (int nt = 0; nt < 8; ++nt) {
            bf16x4 b = ld4(wt + (nt * 16 + li) * K + kt * 16 + lq * 4);
            acc[nt] = mfma(a, b, acc[nt]);
        }
    }
}

// bias + SiLU; writes the BF16 activations to the wave's LDS rows and, when
// saving, Z (FP32) and H (BF16) of layer l to global memory.
__device__ inline void layer_epilogue(floatx4 (&acc)[8], const float* __restrict__ bias, uint16_t* h_lds, int li,
                                      int lq, int row0, int n, int l, uint16_t* __restrict__ h_save,
                                      float* __restrict__ z_save) {
#pragma unroll
    for (int nt = 0; nt < 8; ++nt) {
        int j = nt * 16 + li;
        float bj = bias[j];
#pragma unroll
        for (int it = 0; it < 4; ++it) {
            int i = lq * 4 + it;
            float z = acc[nt][it] + bj;
            uint16_t hb = f2bf(z * sigmoid(z));
            h_lds[i * kStride + j] = hb;
            int r = row0 + i;
            if (h_save && r < n) {
                size_t o = ((size_t)r * 3 + l) * kHid + j;
                z_save[o] = z;
                h_save[o] = hb;
            }
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_forward(int n, const uint16_t* __restrict__ feat,
                                                   const uint16_t* __restrict__ pk, const float* __restrict__ prm,
                                                   float* __restrict__ out, uint16_t* __restrict__ h_save,
                                                   float* __restrict__ z_save) {
    __shared__ __align__(16) uint16_t s_h[kWaves][kWaveRows * kStride];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lq = lane >> 4;
    const int row0 = (blockIdx.x * kWaves + wave) * kWaveRows;
    uint16_t* H = s_h[wave];
    floatx4 acc[8];
    const bool valid = row0 + li < n;
    layer_mm<kIn>(feat + (size_t)row0 * kIn, kIn, pk + pWt1, li, lq, acc, valid);
    layer_epilogue(acc, prm + ob1, H, li, lq, row0, n, 0, h_save, z_save);
    __syncthreads();
    layer_mm<kHid>(H, kStride, pk + pWt2, li, lq, acc, true);
    __syncthreads();
    layer_epilogue(acc, prm + ob2, H, li, lq, row0, n, 1, h_save, z_save);
    __syncthreads();
    layer_mm<kHid>(H, kStride, pk + pWt3, li, lq, acc, true);
    __syncthreads();
    layer_epilogue(acc, prm + ob3, H, li, lq, row0, n, 2, h_save, z_save);
    __syncthreads();
    floatx4 o = zero4();
#pragma unroll
    for (int kt = 0; kt < kHid / 16; ++kt) {
        bf16x4 a = ld4(H + li * kStride + kt * 16 + lq * 4);
        bf16x4 b = ld4(pk + pWt4 + li * kHid + kt * 16 + lq * 4);
        o = mfma(a, b, o);
    }
    if (li < kOut) {
        float b4 = prm[ob4 + li];
#pragma unroll
        for (int it = 0; it < 4; ++it) {
            int r = row0 + lq * 4 + it;
            if (r < n) out[(size_t)r * kOut + li] = o[it] + b4;
        }
    }
}

// ---- inference: weights resident in LDS, 16x16x32 BF16 MFMA ------------------------
// k_forward re-reads every weight tile from L2 for each 16 rows (about 6 KB per row), which bounds
// it far below the matrix cores.  The inference kernel keeps all four BF16 weight matrices in LDS
// (92 KB, loaded once per work-group), runs one 512-lane work-group per CU over 128-row groups
// (8 waves x 16 rows), and uses gfx950's v_mfma_f32_16x16x32_bf16 (A: lane = row l % 16, k =
// 8 (l / 16) .. + 7 of the 32-wide k step; B: lane = column l % 16, same k; C: lane holds rows
// 4 (l / 16) .. + 3 of column l % 16).  Each layer is computed transposed, Z^T = W^T H^T: the weight
// tile is the A operand and the activation rows the B operand, so a lane ends with four
// consecutive output columns of one row and writes them to LDS as one 8-byte store.  Layer 1's k
// runs to 64 with the columns past 48 zero.  SiLU uses the hardware exp2 and reciprocal (the
// activations are rounded to BF16 right after).
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kK1 = 64;                 // layer-1 k padded to two 32-wide steps
constexpr int kS1 = kK1 + 8;            // LDS row strides (BF16 elements; 16-byte skew)
constexpr int kSH = kHid + 8;
constexpr int kInfWaves = 8;
constexpr int kInfBlock = 64 * kInfWaves;
constexpr int kInfRows = 16 * kInfWaves;  // rows per work-group iteration
constexpr int kLdsW1 = kHid * kS1;
constexpr int kLdsW = kHid * kSH;
constexpr int kLdsW4 = kOutP * kSH;
constexpr int kLdsH = kInfWaves * 16 * kSH;

__device__ inline floatx4 mfma32(bf16x8 a, bf16x8 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ inline bf16x8 ld8(const uint16_t* p) { return *reinterpret_cast<const bf16x8*>(p); }

__device__ inline float silu_fast(float z) {
    return z * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.4426950408889634f * z));
}

typedef float floatx2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// two floats -> two BF16 (round to nearest even) in one v_cvt_pk_bf16_f32
__device__ inline uint32_t pack_bf16x2(float a, float b) {
    const bf16x2 p = __builtin_convertvector((floatx2){a, b}, bf16x2);
    uint32_t u;
    __builtin_memcpy(&u, &p, 4);
    return u;
}

// bias + SiLU of the wave's tile (lane: row li, columns 16 nt + 4 lq .. + 3) into its LDS rows
__device__ inline void infer_epilogue(const floatx4 (&acc)[8], const float* __restrict__ bias, uint16_t* h, int li,
                                      int lq) {
#pragma unroll
    for (int nt = 0; nt < 8; ++nt) {
        const int j = nt * 16 + lq * 4;
        const uint32_t lo = pack_bf16x2(silu_fast(acc[nt][0] + bias[j]), silu_fast(acc[nt][1] + bias[j + 1]));
        const uint32_t hi = pack_bf16x2(silu_fast(acc[nt][2] + bias[j + 2]), silu_fast(acc[nt][3] + bias[j + 3]));
        *reinterpret_cast<uint2*>(h + li * kSH + j) = make_uint2(lo, hi);
    }
}

template <int KSTEPS, int STRIDE>
__device__ inline void infer_layer(const uint16_t* h_rows, const uint16_t* w, int li, int lq, floatx4 (&acc)[8]) {
#pragma unroll
    for (int nt = 0; nt < 8; ++nt) acc[nt] = zero4();
#pragma unroll
    for (int kt = 0; kt < KSTEPS; ++kt) {
        const bf16x8 hb = ld8(h_rows + li * kSH + kt * 32 + lq * 8);
#pragma unroll
        for (int nt = 0; nt < 8; ++nt)
            acc[nt] = mfma32(ld8(w + (nt * 16 + li) * STRIDE + kt * 32 + lq * 8), hb, acc[nt]);
    }
}

__global__ __launch_bounds__(kInfBlock) void k_infer(int n, const uint16_t* __restrict__ feat,
                                                    const uint16_t* __restrict__ pk, const float* __restrict__ prm,
                                                    float* __restrict__ out) {
    __shared__ __align__(16) uint16_t s_w1[kLdsW1];
    __shared__ __align__(16) uint16_t s_w2[kLdsW];
    __shared__ __align__(16) uint16_t s_w3[kLdsW];
    __shared__ __align__(16) uint16_t s_w4[kLdsW4];
    __shared__ __align__(16) uint16_t s_h[kLdsH];
    for (int e = threadIdx.x; e < kLdsW1; e += kInfBlock) {
        const int r = e / kS1, k = e % kS1;
        s_w1[e] = k < kIn ? pk[pWt1 + r * kIn + k] : (uint16_t)0;
    }
    for (int e = threadIdx.x; e < kLdsW; e += kInfBlock) {
        const int r = e / kSH, k = e % kSH;
        s_w2[e] = k < kHid ? pk[pWt2 + r * kHid + k] : (uint16_t)0;
        s_w3[e] = k < kHid ? pk[pWt3 + r * kHid + k] : (uint16_t)0;
    }
    for (int e = threadIdx.x; e < kLdsW4; e += kInfBlock) {
        const int r = e / kSH, k = e % kSH;
        s_w4[e] = k < kHid ? pk[pWt4 + r * kHid + k] : (uint16_t)0;
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lq = lane >> 4;
    uint16_t* H = s_h + wave * 16 * kSH;
    floatx4 acc[8];
    // every wave runs the same number of iterations (the barriers below are work-group wide)
    for (int base = blockIdx.x * kInfRows; base < n; base += gridDim.x * kInfRows) {
        const int r = base + wave * 16 + li;  // the row this lane carries
        // layer 1: the B operand straight from the feature rows (48 BF16 each); k >= 48 and rows >= n
        // read as 0
#pragma unroll
        for (int nt = 0; nt < 8; ++nt) acc[nt] = zero4();
#pragma unroll
        for (int kt = 0; kt < kK1 / 32; ++kt) {
            const int k0 = kt * 32 + lq * 8;
            bf16x8 fb = {};
            if (r < n && k0 < kIn) fb = ld8(feat + (size_t)r * kIn + k0);
#pragma unroll
            for (int nt = 0; nt < 8; ++nt) acc[nt] = mfma32(ld8(s_w1 + (nt * 16 + li) * kS1 + k0), fb, acc[nt]);
        }
        infer_epilogue(acc, prm + ob1, H, li, lq);
        __syncthreads();
        infer_layer<kHid / 32, kSH>(H, s_w2, li, lq, acc);
        __syncthreads();
        infer_epilogue(acc, prm + ob2, H, li, lq);
        __syncthreads();
        infer_layer<kHid / 32, kSH>(H, s_w3, li, lq, acc);
        __syncthreads();
        infer_epilogue(acc, prm + ob3, H, li, lq);
        __syncthreads();
        floatx4 o = zero4();
#pragma unroll
        for (int kt = 0; kt < kHid / 32; ++kt)
            o = mfma32(ld8(s_w4 + li * kSH + kt * 32 + lq * 8), ld8(H + li * kSH + kt * 32 + lq * 8), o);
        // lane (li, lq = 0) holds outputs 0..3 of row r; the first three are the network's
        if (lq == 0 && r < n) {
#pragma unroll
            for (int c = 0; c < kOut; ++c) out[(size_t)r * kOut + c] = o[c] + prm[ob4 + c];
        }
        __syncthreads();  // H is rewritten by the next group's first epilogue
    }
}

// ---- backward --------------------------------------------------------------------
// dA (16 x Kout) of the wave's rows times W^T: dH[r][j] = sum_k dZ[r][k] W[j][k];
// B operand B[k][j] = W[j][k] read from the natural W layout [in][out] (stride KO).
template <int KO>
__device__ inline void back_mm(const uint16_t* dz_lds, int dz_stride, const uint16_t* __restrict__ w, int li, int lq,
                               floatx4 (&acc)[8]) {
#pragma unroll
    for (int nt = 0; nt < 8; ++nt) acc[nt] = zero4();
#pragma unroll
    for (int kt = 0; kt < KO / 16; ++kt) {
        bf16x4 a = ld4(dz_lds + li * dz_stride + kt * 16 + lq * 4);
#pragma unroll
        for (int nt = 0; nt < 8; ++nt) {
            bf16x4 b = ld4(w + (nt * 16 + li) * KO + kt * 16 + lq * 4);
            acc[nt] = mfma(a, b, acc[nt]);
        }
    }
}

// dZ = bf16(dH * silu'(Z)) for the wave's rows into LDS (rows past n get 0).
__device__ inline void back_epilogue(floatx4 (&acc)[8], const float* __restrict__ z_save, int l, uint16_t* dz_lds,
                                     int li, int lq, int row0, int n) {
#pragma unroll
    for (int nt = 0; nt < 8; ++nt) {
        int j = nt * 16 + li;
#pragma unroll
        for (int it = 0; it < 4; ++it) {
            int i = lq * 4 + it;
            int r = row0 + i;
            float d = 0.0f;
            if (r < n) {
                float z = z_save[((size_t)r * 3 + l) * kHid + j];
                float s = sigmoid(z);
                d = acc[nt][it] * (s * (1.0f + z * (1.0f - s)));
            }
            dz_lds[i * kStride + j] = f2bf(d);
        }
    }
}

// Partial weight gradient dW[i][j] = sum_rows A[r][i] * dZ[r][j] over the
// workgroup's 64 rows, for output tiles (it_i, it_j) assigned to this wave.
// a: [64][a_stride] BF16 rows, dz: [64][dz_stride] BF16 rows.
template <int MI, int NJ>
__device__ inline void weight_grad(const uint16_t* a, int a_stride, const uint16_t* dz, int dz_stride, int wave,
                                   int li, int lq, float* __restrict__ gw, int ld_out, int ncols) {
    constexpr int tiles = (MI / 16) * (NJ / 16);
    for (int t = wave; t < tiles; t += kWaves) {
        int ti = t / (NJ / 16), tj = t % (NJ / 16);
        floatx4 acc = zero4();
#pragma unroll
        for (int kt = 0; kt < kRowsWG / 16; ++kt) {
            int r = kt * 16 + lq * 4;
            bf16x4 av = ld4_strided(a + r * a_stride + ti * 16 + li, a_stride);
            bf16x4 bv = ld4_strided(dz + r * dz_stride + tj * 16 + li, dz_stride);
            acc = mfma(av, bv, acc);
        }
        int j = tj * 16 + li;
        if (j < ncols) {
#pragma unroll
            for (int it = 0; it < 4; ++it) {
                int i = ti * 16 + lq * 4 + it;
                gw[i * ld_out + j] = acc[it];
            }
        }
    }
}

// column sums of dZ (BF16 rows) over the workgroup's rows, in row order
__device__ inline void bias_grad(const uint16_t* dz, int dz_stride, int ncols, float* __restrict__ gb) {
    for (int j = threadIdx.x; j < ncols; j += kBlock) {
        float s = 0.0f;
        for (int r = 0; r < kRowsWG; ++r) s += bf2f(dz[r * dz_stride + j]);
        gb[j] = s;
    }
}

// stage a [64][K] BF16 block of global rows (row stride ld, first column col0) into LDS
template <int K>
__device__ inline void stage_rows(const uint16_t* __restrict__ src, size_t ld, int col0, int row0, int n,
                                  uint16_t* dst, int dst_stride) {
    for (int e = threadIdx.x; e < kRowsWG * (K / 4); e += kBlock) {
        int r = e / (K / 4), c = (e % (K / 4)) * 4;
        bf16x4 v = bf16x4{0, 0, 0, 0};
        if (row0 + r < n) v = ld4(src + (size_t)(row0 + r) * ld + col0 + c);
        *reinterpret_cast<bf16x4*>(dst + r * dst_stride + c) = v;
    }
}

// LDS: dZ4 (64 x 16) and three 64 x 128 BF16 row blocks A, B, C (~55 KB), reused as
//   A: dZ3, later dZ1     B: staged H3, H2, H1, X     C: dZ2
__global__ __launch_bounds__(kBlock) void k_backward(int n, const uint16_t* __restrict__ feat,
                                                    const float* __restrict__ target, const float* __restrict__ roww,
                                                    const uint16_t* __restrict__ pk, const float* __restrict__ out,
                                                    const uint16_t* __restrict__ h_save,
                                                    const float* __restrict__ z_save, float* __restrict__ pgrad,
                                                    float* __restrict__ ploss, float inv_count) {
    __shared__ __align__(16) uint16_t s_dz4[kRowsWG * kStrideOut];
    __shared__ __align__(16) uint16_t s_a[kRowsWG * kStride];
    __shared__ __align__(16) uint16_t s_b[kRowsWG * kStride];
    __shared__ __align__(16) uint16_t s_c[kRowsWG * kStride];
    __shared__ float s_loss[kBlock];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lq = lane >> 4;
    const int wg_row0 = blockIdx.x * kRowsWG;
    const int wr = wave * kWaveRows;
    const int row0 = wg_row0 + wr;
    float* g = pgrad + (size_t)blockIdx.x * kParams;

    // dY = 2 w (out - target) / (3 n); dZ4 = bf16(dY), padded to 16 columns
    float lsum = 0.0f;
    for (int e = threadIdx.x; e < kRowsWG * kOutP; e += kBlock) {
        int r = e / kOutP, j = e % kOutP, gr = wg_row0 + r;
        float d = 0.0f;
        if (j < kOut && gr < n) {
            float w = roww ? roww[gr] : 1.0f;
            float diff = out[(size_t)gr * kOut + j] - target[(size_t)gr * kOut + j];
            lsum += w * diff * diff;
            d = 2.0f * w * diff * inv_count;
        }
        s_dz4[r * kStrideOut + j] = f2bf(d);
    }
    s_loss[threadIdx.x] = lsum;
    __syncthreads();
    if (threadIdx.x == 0) {
        float s = 0.0f;
        for (int t = 0; t < kBlock; ++t) s += s_loss[t];
        ploss[blockIdx.x] = s * inv_count;
    }

    floatx4 acc[8];
    // dZ3 (wave rows) -> A; H3 -> B
    back_mm<kOutP>(s_dz4 + wr * kStrideOut, kStrideOut, pk + pW4, li, lq, acc);
    back_epilogue(acc, z_save, 2, s_a + wr * kStride, li, lq, row0, n);
    stage_rows<kHid>(h_save, (size_t)3 * kHid, 2 * kHid, wg_row0, n, s_b, kStride);
    __syncthreads();
    // dW4 = H3^T dZ4, db4; dZ2 (wave rows) -> C
    weight_grad<kHid, kOutP>(s_b, kStride, s_dz4, kStrideOut, wave, li, lq, g + oW4, kOut, kOut);
    bias_grad(s_dz4, kStrideOut, kOut, g + ob4);
    back_mm<kHid>(s_a + wr * kStride, kStride, pk + pW3, li, lq, acc);
    back_epilogue(acc, z_save, 1, s_c + wr * kStride, li, lq, row0, n);
    __syncthreads();
    // H2 -> B; dW3 = H2^T dZ3, db3
    stage_rows<kHid>(h_save, (size_t)3 * kHid, kHid, wg_row0, n, s_b, kStride);
    __syncthreads();
    weight_grad<kHid, kHid>(s_b, kStride, s_a, kStride, wave, li, lq, g + oW3, kHid, kHid);
    bias_grad(s_a, kStride, kHid, g + ob3);
    __syncthreads();
    // dZ1 (wave rows) -> A; H1 -> B
    back_mm<kHid>(s_c + wr * kStride, kStride, pk + pW2, li, lq, acc);
    back_epilogue(acc, z_save, 0, s_a + wr * kStride, li, lq, row0, n);
    stage_rows<kHid>(h_save, (size_t)3 * kHid, 0, wg_row0, n, s_b, kStride);
    __syncthreads();
    // dW2 = H1^T dZ2, db2
    weight_grad<kHid, kHid>(s_b, kStride, s_c, kStride, wave, li, lq, g + oW2, kHid, kHid);
    bias_grad(s_c, kStride, kHid, g + ob2);
    __syncthreads();
    // X -> B; dW1 = X^T dZ1, db1
    stage_rows<kIn>(feat, (size_t)kIn, 0, wg_row0, n, s_b, kStrideIn);
    __syncthreads();
    weight_grad<kIn, kHid>(s_b, kStrideIn, s_a, kStride, wave, li, lq, g + oW1, kHid, kHid);
    bias_grad(s_a, kStride, kHid, g + ob1);
}

__global__ void k_reduce(int nparts, const float* __restrict__ pgrad, const float* __restrict__ ploss,
                         float* __restrict__ grad, float* __restrict__ loss) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kParams) {
        float s = 0.0f;
        for (int w = 0; w < nparts; ++w) s += pgrad[(size_t)w * kParams + p];
        grad[p] = s;
    }
    if (p == 0 && loss) {
        float s = 0.0f;
        for (int w = 0; w < nparts; ++w) s += ploss[w];
        loss[0] = s;
    }
}

__global__ void k_adam(float* __restrict__ prm, const float* __restrict__ grad, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kParams) return;
    float gp = grad[p];
    float mp = b1 * m[p] + (1.0f - b1) * gp;
    float vp = b2 * v[p] + (1.0f - b2) * gp * gp;
    m[p] = mp;
    v[p] = vp;
    prm[p] -= lr * (mp / c1) / (sqrtf(vp / c2) + eps);
}

inline int blocks_for(long n, int bs) { return (int)((n + bs - 1) / bs); }

}  // namespace

extern "C" {

const char* vbd_ws_last_error(void) { return g_err.c_str(); }

int32_t vbd_ws_sizes(int64_t* out3) {
    if (!out3) return fail(-1, "vbd_ws_sizes: null argument");
    out3[0] = kParams;
    out3[1] = kPacked;
    out3[2] = kRowsWG;
    return 0;
}

int32_t vbd_ws_features(void* stream, int32_t n, const int32_t* nbr_off, const int32_t* nbr, const float* x0,
                        const float* y, const float* corr, const float* acc, const float* mass, const float* area,
                        float dt, const float* scales, uint16_t* feat_out) {
    if (n < 0 || (n > 0 && (!nbr_off || !x0 || !y || !mass || !area || !scales || !feat_out)))
        return fail(-1, "vbd_ws_features: bad argument");
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_features, dim3(blocks_for(n, kFeatBlock)), dim3(kFeatBlock), 0, (hipStream_t)stream, n,
                       nbr_off, nbr, x0, y, corr, acc, mass, area, dt, scales, feat_out);
    WS_LAUNCH_CHECK("k_features");
    return 0;
}

int32_t vbd_ws_smooth(void* stream, int32_t n, int32_t width, const int32_t* nbr_off, const int32_t* nbr,
                      const float* active, int32_t sweeps, float* v, float* tmp) {
    if (n < 0 || sweeps < 0 || width < 1 || width > kSmoothMaxWidth ||
        (n > 0 && (!nbr_off || !active || !v || (sweeps > 0 && !tmp))))
        return fail(-1, "vbd_ws_smooth: bad argument (width 1..12)");
    if (n == 0 || sweeps == 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    float* src = v;
    float* dst = tmp;
    for (int k = 0; k < sweeps; ++k) {
        hipLaunchKernelGGL(k_smooth, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, width, nbr_off, nbr, active, src,
                           dst);
        WS_LAUNCH_CHECK("k_smooth");
        float* t = src;
        src = dst;
        dst = t;
    }
    if (src != v) {
        hipError_t e = hipMemcpyAsync(v, src, sizeof(float) * (size_t)width * (size_t)n, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return fail(-2, std::string("vbd_ws_smooth: ") + hipGetErrorString(e));
    }
    return 0;
}

int32_t vbd_ws_pack(void* stream, const float* params, uint16_t* packed) {
    if (!params || !packed) return fail(-1, "vbd_ws_pack: null argument");
    hipLaunchKernelGGL(k_pack, dim3(blocks_for(kPacked, 256)), dim3(256), 0, (hipStream_t)stream, params, packed);
    WS_LAUNCH_CHECK("k_pack");
    return 0;
}

int32_t vbd_ws_forward(void* stream, int32_t n, const uint16_t* feat, const uint16_t* packed, const float* params,
                       float* out) {
    if (n < 0 || (n > 0 && (!feat || !packed || !params || !out))) return fail(-1, "vbd_ws_forward: bad argument");
    if (n == 0) return 0;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                                                 hipSuccess || cus <= 0)
        cus = 256;
    const int grid = std::min(blocks_for(n, kInfRows), cus);
    hipLaunchKernelGGL(k_infer, dim3(grid), dim3(kInfBlock), 0, (hipStream_t)stream, n, feat, packed, params, out);
    WS_LAUNCH_CHECK("k_infer");
    return 0;
}

int32_t vbd_ws_forward_train(void* stream, int32_t n, const uint16_t* feat, const uint16_t* packed,
                             const float* params, float* out, uint16_t* h_save, float* z_save) {
    if (n < 0 || (n > 0 && (!feat || !packed || !params || !out || !h_save || !z_save)))
        return fail(-1, "vbd_ws_forward_train: bad argument");
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_forward, dim3(blocks_for(n, kRowsWG)), dim3(kBlock), 0, (hipStream_t)stream, n, feat, packed,
                       params, out, h_save, z_save);
    WS_LAUNCH_CHECK("k_forward(train)");
    return 0;
}

int32_t vbd_ws_backward(void* stream, int32_t n, const uint16_t* feat, const float* target, const float* row_weight,
                        const uint16_t* packed, const float* out, const uint16_t* h_save, const float* z_save,
                        float* partial_grad, float* partial_loss) {
    if (n <= 0 || !feat || !target || !packed || !out || !h_save || !z_save || !partial_grad || !partial_loss)
        return fail(-1, "vbd_ws_backward: bad argument");
    float inv_count = 1.0f / (3.0f * (float)n);
    hipLaunchKernelGGL(k_backward, dim3(blocks_for(n, kRowsWG)), dim3(kBlock), 0, (hipStream_t)stream, n, feat, target,
                       row_weight, packed, out, h_save, z_save, partial_grad, partial_loss, inv_count);
    WS_LAUNCH_CHECK("k_backward");
    return 0;
}

int32_t vbd_ws_reduce(void* stream, int32_t n, const float* partial_grad, const float* partial_loss, float* grad,
                      float* loss) {
    if (n <= 0 || !partial_grad || !partial_loss || !grad) return fail(-1, "vbd_ws_reduce: bad argument");
    hipLaunchKernelGGL(k_reduce, dim3(blocks_for(kParams, 256)), dim3(256), 0, (hipStream_t)stream,
                       blocks_for(n, kRowsWG), partial_grad, partial_loss, grad, loss);
    WS_LAUNCH_CHECK("k_reduce");
    return 0;
}

int32_t vbd_ws_adam(void* stream, float* params, const float* grad, float* m, float* v, float lr, float beta1,
                    float beta2, float eps, int32_t step) {
    if (!params || !grad || !m || !v || step < 1) return fail(-1, "vbd_ws_adam: bad argument");
    float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(kParams, 256)), dim3(256), 0, (hipStream_t)stream, params, grad, m, v,
                       lr, beta1, beta2, eps, c1, c2);
    WS_LAUNCH_CHECK("k_adam");
    return 0;
}

}  // extern "C"
