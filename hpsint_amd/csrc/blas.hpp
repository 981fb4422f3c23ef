// blas.hpp -- launchers of the block-vector BLAS-1 and Jacobi kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hps
{
  constexpr int RED_BLOCKS = 2048; // fixed reduction grid (deterministic sums)
  constexpr int MAXK       = 32;   // max vectors of one multi-dot

  void blas_dot(hipStream_t, int64_t n, const double *x, const double *y, double *partial,
                double *out);
  void blas_norm2(hipStream_t, int64_t n, const double *x, double *partial, double *out);
  void blas_block_norms(hipStream_t, int64_t n_nodes, int nc, const double *x,
                        double *partial, double *out);
  void blas_multi_dot(hipStream_t, int64_t n, int k, const double *V, const double *w,
                      double *partial, double *out);
  void blas_multi_axpy_norm(hipStream_t, int64_t n, int k, const double *V, const double *h,
                            double *w, double *partial, double *out);
  void blas_axpy(hipStream_t, int64_t n, double a, const double *x, double *y);
  void blas_sadd(hipStream_t, int64_t n, double s, double a, const double *x, double *y);
  void blas_scale(hipStream_t, int64_t n, double a, double *x);
  void blas_fill(hipStream_t, int64_t n, double a, double *x);
  void blas_f64_to_f32(hipStream_t, int64_t n, float *dst, const double *src);
  void blas_f32_to_f64(hipStream_t, int64_t n, double *dst, const float *src);
  void invert_diag(hipStream_t, int64_t n, const double *d, double *inv64, float *inv32);
  void diag_apply64(hipStream_t, int64_t n, const double *inv, const double *src, double *dst);
  void diag_apply32(hipStream_t, int64_t n, const float *inv, const float *src, float *dst);
  bool unit_empty_rows(hipStream_t, int64_t n_nodes, int nc, double *blocks);
  // node-block inverses are stored entry-major with a node stride of
  // inv_stride(n_nodes) (a multiple of 4): entry e of node n at e * stride + n;
  // a buffer of them holds inv_stride(n_nodes) * nc * nc values
  inline int64_t inv_stride(int64_t n_nodes)
  {
    return (n_nodes + 3) & ~(int64_t)3;
  }
  bool invert_blocks(hipStream_t, int64_t n_nodes, int nc, const double *blocks, double *inv64,
                     float *inv32, int *singular);
  bool block_apply64(hipStream_t, int64_t n_nodes, int nc, const double *inv, const double *src,
                     double *dst);
  bool block_apply32(hipStream_t, int64_t n_nodes, int nc, const float *inv, const float *src,
                     float *dst);
  void from_blocks(hipStream_t, int64_t n_nodes, int nc, double *x, const double *const *blocks_dev);
  void to_blocks(hipStream_t, int64_t n_nodes, int nc, double *const *blocks_dev, const double *x);
  void blas_multi_dot_f32(hipStream_t, int64_t n, int k, const float *V, const float *w,
                          double *partial, double *out);
  void blas_multi_axpy_norm_f32(hipStream_t, int64_t n, int k, const float *V, const double *h,
                                float *w, double *partial, double *out);
  void blas_norm2_f32(hipStream_t, int64_t n, const float *x, double *partial, double *out);
  void blas_scale_f32(hipStream_t, int64_t n, double a, float *x);
  void blas_axpy_multi_f64(hipStream_t, int64_t n, int k, const double *a, const double *X,
                           double *y);
  void blas_axpy_multi_f32_f64(hipStream_t, int64_t n, int k, const double *a, const float *X,
                               double *y);
  // constrained nodes (node indices, each listed once)
  void constraint_zero64(hipStream_t, int64_t n, int nc, const int64_t *nodes, double *src,
                         double *save);
  void constraint_zero32(hipStream_t, int64_t n, int nc, const int64_t *nodes, float *src,
                         float *save);
  void constraint_restore64(hipStream_t, int64_t n, int nc, const int64_t *nodes, double *src,
                            double *dst, const double *save);
  void constraint_restore32(hipStream_t, int64_t n, int nc, const int64_t *nodes, float *src,
                            float *dst, const float *save);
  void constraint_set(hipStream_t, int64_t n, int nc, const int64_t *nodes, double *dst,
                      double value);
  void constraint_identity_blocks(hipStream_t, int64_t n, int nc, const int64_t *nodes,
                                  double *blocks);
  // node blocks stored column by column -> row by row (in != out)
  void transpose_blocks(hipStream_t, int64_t n_nodes, int nc, const double *in, double *out);
} // namespace hps
