// mlp.hpp -- launchers of the learned-predictor kernels (mlp.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hps
{
  enum MlpEpilogue
  {
    EPI_F32       = 0, // store the FP32 product (split-K partials)
    EPI_BIAS_TANH = 1, // bf16 tanh(C + bias)
    EPI_BIAS_F32  = 2, // FP32 C + bias
    EPI_DTANH     = 3, // bf16 C * (1 - H^2)
  };

  constexpr int FEAT_STRIDE = 272; // 27 neighbours x 10 + dt + 1

  struct FeatParams
  {
    int64_t n_nodes;
    int     nn[3];
    int     n_comp;
    double  dt, dt_scale, inc_scale, inc_floor;
  };

  void mlp_gemm(hipStream_t, int ta, int tb, int epi, int M, int N, int K, const uint16_t *A,
                int lda, const uint16_t *B, int ldb, void *C, int ldc, const float *bias,
                const uint16_t *H, int nsplit, int64_t split_stride);
  void mlp_sum_splits(hipStream_t, int64_t n, int nsplit, const float *part, float *out);
  void mlp_colsum(hipStream_t, int M, int N, const uint16_t *X, float *out);
  void mlp_loss(hipStream_t, int64_t n, const float *Y, const float *T, uint16_t *dY,
                float *partial, int nblocks);
  void mlp_adam(hipStream_t, int n_out, int n_in, float *W, float *m1, float *m2, const float *g,
                float lr, float b1, float b2, float eps, float c1, float c2, uint16_t *Wb,
                uint16_t *WTb);
  void mlp_features(hipStream_t, const FeatParams &, const double *u_prev, const double *u_cur,
                    uint16_t *X, int *top2, float *lscale);
  void mlp_predict(hipStream_t, int64_t n_nodes, int nc, const double *u_prev,
                   const double *u_cur, const float *Y, const int *top2, const float *out_scale,
                   const float *lscale, double *u_pred);
  void mlp_targets(hipStream_t, int64_t n_nodes, int nc, const double *u_prev,
                   const double *u_cur, const double *u_next, const int *top2,
                   const float *out_scale, const float *lscale, float *T);
} // namespace hps
