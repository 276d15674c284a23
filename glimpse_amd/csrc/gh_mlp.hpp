// gh_mlp.hpp -- the learned state selector's network (shared host/device types).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gh_internal.hpp"

namespace gh {

constexpr int MLP_IN = 32;    // features per (target, candidate panel haplotype)
constexpr int MLP_H = 128;    // width of both hidden layers

// Parameter block (FP32, contiguous, in this order):
//   W1 [MLP_IN][MLP_H], b1 [MLP_H], W2 [MLP_H][MLP_H], b2 [MLP_H], W3 [MLP_H], b3 [1]
// y = W3 . relu(W2^T relu(W1^T x + b1) + b2) + b3, W1[i][j] = weight from input i to unit j.
constexpr int MLP_OFF_W1 = 0;
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_IN * MLP_H;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_H;
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_H;
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_H;
constexpr int MLP_NPARAM = MLP_OFF_B3 + 1;          // 20,865
constexpr int MLP_NBF16 = MLP_IN * MLP_H + MLP_H * MLP_H;   // BF16 copies of W1 then W2

// Row selection for a training batch: batch element r is row
// (a * (i0 + r) + b) mod n_total of the data set (a coprime to n_total gives a
// permutation; a = 1, b = 0 is the identity).
struct MlpBatch {
    const float *X;      // [n_total][MLP_IN]
    const float *t;      // [n_total] regression targets
    const float *w;      // [n_total] row weights (may be null: all ones)
    int64_t n_total, i0;
    int32_t n;           // rows in this batch
    uint64_t a, b;
};

size_t mlp_fwd_lds_bytes();
size_t mlp_grad_lds_bytes();
// BF16 copies of W1 and W2 (wb: MLP_NBF16 halves) from the FP32 block w.
void mlp_refresh_bf16(const float *w, uint16_t *wb, hipStream_t s);
// y[r] = net(X[r]) for r < n.
void mlp_forward_launch(const float *w, const uint16_t *wb, const float *X, int64_t n, float *y, hipStream_t s);
// Gradient of the batch loss  sum_r w_r (y_r - t_r)^2 / sum_r w_r  into g
// (MLP_NPARAM floats, overwritten) and the loss into g[MLP_NPARAM].
// part: scratch of mlp_grad_part_floats() floats.
size_t mlp_grad_part_floats();
void mlp_grad_launch(const float *w, const uint16_t *wb, const MlpBatch &B, float inv_wsum, float *part, float *g,
                     hipStream_t s);
// Adam (Kingma & Ba 2015) on the whole block; step counts from 1.
void adam_launch(float *w, const float *g, float *m, float *v, float lr, int step, hipStream_t s);

}  // namespace gh

// A network with its optimiser state, resident on the device that was current
// at creation.  Weights are read-only while a run uses the network.
struct gh_mlp {
    gh::DBuf w, wb, m, v, g, part, X, y, t, wt, acc;
    int step = 0;
    hipStream_t s = nullptr;
};

namespace gh {
void mlp_create(uint32_t seed, gh_mlp **out);
void mlp_destroy(gh_mlp *M);
void mlp_get(gh_mlp *M, float *w);
void mlp_set(gh_mlp *M, const float *w);
void mlp_forward_host(gh_mlp *M, int64_t n, const float *X, float *y);
void mlp_backward_host(gh_mlp *M, int n, const float *X, const float *t, const float *w, float *grad, float *loss);
void mlp_step_host(gh_mlp *M, int n, const float *X, const float *t, const float *w, float lr, float *loss);
// Adam over a device-resident data set in shuffled batches; learning rate decays
// linearly to lr/10 over the run.  loss_per_epoch (host, optional) [epochs].
void mlp_train_device(gh_mlp *M, const float *X, const float *Y, const float *W, int64_t n_total, int epochs,
                      int batch, float lr, uint32_t seed, hipStream_t data_stream, float *loss_per_epoch);
}  // namespace gh
