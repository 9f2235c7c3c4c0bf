/*
 * vbd_warmstart.h -- C ABI of libvbd_ws.so: the learned warm start of the VBD
 * iteration on MI355X (gfx950).  No counterpart in the reference (Newton's
 * SolverVBD starts every substep from the inertial guess, forward_step,
 * newton/_src/solvers/vbd/particle_vbd_kernels.py:1350-1375); this library
 * predicts a better starting iterate, which SolverVBD accepts only when its
 * force residual is lower (vbd_set_iterate / vbd_residual in vbd_hip.h).
 *
 * A per-particle MLP 48 -> 128 -> 128 -> 128 -> 3 with SiLU activations maps
 * 1-ring features to the correction dx = x* - y of the converged iterate x*
 * over the inertial guess y (in units of the caller's correction scale).
 * Weights and activations are BF16, every product accumulates in FP32 on the
 * MFMA units (v_mfma_f32_16x16x16_bf16); biases, pre-activations, output,
 * gradients and the Adam state are FP32.
 *
 * Every pointer is a DEVICE pointer, every call is asynchronous on the given
 * stream, nothing allocates: callers pass workspaces sized by the functions
 * below.  BF16 tensors are uint16_t bit patterns (round-to-nearest-even).
 * Returns 0 or a negative code (-1 bad argument, -2 HIP error); the message of
 * the calling thread's last failure is in vbd_ws_last_error().
 *
 * Parameter vector (FP32, this order, row-major, in x out):
 *   W1[48][128] b1[128] W2[128][128] b2[128] W3[128][128] b3[128] W4[128][3] b4[3]
 * Packed BF16 weights (vbd_ws_pack): for each layer both W^T (forward) and W
 * (backward), the last layer zero-padded from 3 to 16 outputs.
 */
#ifndef VBD_WARMSTART_H
#define VBD_WARMSTART_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VBD_WS_IN 48
#define VBD_WS_HIDDEN 128
#define VBD_WS_OUT 3

const char* vbd_ws_last_error(void);

/* Sizes: [0] FP32 parameters, [1] packed BF16 weight elements, [2] rows per
 * training workgroup (partial-gradient granularity). */
int32_t vbd_ws_sizes(int64_t* out3);

/* Per-particle features (BF16 [n,48]).  nbr_off/nbr: 1-ring CSR (int32);
 * x0: start-of-substep positions, y: inertial guess, corr: the previous
 * substep's correction x* - y, acc: external acceleration [n,3] (may be NULL);
 * mass, area: per particle; dt: substep; scales[5] = {displacement, correction,
 * length, mean mass, mean area} (the 48 columns are listed in vbd_warmstart.hip). */
int32_t vbd_ws_features(void* stream, int32_t n, const int32_t* nbr_off, const int32_t* nbr, const float* x0,
                        const float* y, const float* corr, const float* acc, const float* mass, const float* area,
                        float dt, const float* scales, uint16_t* feat_out);

/* Low-pass of correction fields v ([n, width] floats, width 1..12, in place): `sweeps` sweeps of
 * v_i <- (v_i + sum_j v_j) / (1 + degree_i) over the 1-ring CSR (nbr_off [n+1], nbr), with particles
 * whose active[i] == 0 read and written as 0; tmp is [n, width] device scratch. */
int32_t vbd_ws_smooth(void* stream, int32_t n, int32_t width, const int32_t* nbr_off, const int32_t* nbr,
                      const float* active, int32_t sweeps, float* v, float* tmp);

/* FP32 parameters -> packed BF16 weights. */
int32_t vbd_ws_pack(void* stream, const float* params, uint16_t* packed);

/* Inference: out[n,3] = MLP(feat) (FP32). */
int32_t vbd_ws_forward(void* stream, int32_t n, const uint16_t* feat, const uint16_t* packed, const float* params,
                       float* out);

/* Training step pieces for a batch of rows.
 * forward_train additionally saves H (BF16 [n,3,128]) and Z (FP32 [n,3,128]).
 * backward: loss = mean over rows and the 3 outputs of (out - target)^2 * w_row
 * (row_weight may be NULL = 1); writes per-workgroup partial gradients
 * (FP32 [ceil(n / rows_per_wg), n_params]) and the partial loss sums.
 * reduce: grad[p] = sum over partials in workgroup order (deterministic). */
int32_t vbd_ws_forward_train(void* stream, int32_t n, const uint16_t* feat, const uint16_t* packed,
                             const float* params, float* out, uint16_t* h_save, float* z_save);
int32_t vbd_ws_backward(void* stream, int32_t n, const uint16_t* feat, const float* target, const float* row_weight,
                        const uint16_t* packed, const float* out, const uint16_t* h_save, const float* z_save,
                        float* partial_grad, float* partial_loss);
int32_t vbd_ws_reduce(void* stream, int32_t n, const float* partial_grad, const float* partial_loss, float* grad,
                      float* loss);

/* Adam (FP32): m, v, params updated in place; step is the 1-based step count. */
int32_t vbd_ws_adam(void* stream, float* params, const float* grad, float* m, float* v, float lr, float beta1,
                    float beta2, float eps, int32_t step);

#ifdef __cplusplus
}
#endif

#endif
