// blas.hip -- block-vector BLAS-1 and Jacobi preconditioner kernels.
//
// DynamicBlockVector operations (I/lac/dynamic_block_vector.h:346-537) on
// the component-interleaved storage, the inverse-diagonal and inverse
// node-block applies (I/lac/preconditioners.h:160-318) and the layout
// conversions (I/numerics/vector_tools.h:40-100).  Reductions use a fixed
// grid and a fixed second-pass order, so results do not depend on timing.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "blas.hpp"

namespace hps
{
  constexpr int RB = 256; // threads per reduction block

  template <typename T>
  __device__ __forceinline__ T block_reduce(T v, T *sh)
  {
    // wave-level tree with DPP-free shuffles, then across waves in LDS
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
      v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
      sh[w] = v;
    __syncthreads();
    T s = T(0);
    if (threadIdx.x == 0)
      for (int i = 0; i < RB / 64; ++i)
        s += sh[i];
    __syncthreads();
    return s;
  }

  __global__ void __launch_bounds__(RB)
    k_dot(int64_t n, const double *__restrict__ x, const double *__restrict__ y,
          double *__restrict__ partial)
  {
    __shared__ double sh[RB / 64];
    double            s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * RB)
      s += x[i] * y[i];
    s = block_reduce(s, sh);
    if (threadIdx.x == 0)
      partial[blockIdx.x] = s;
  }

  // sums nb partial rows of k values each: out[j] = sum_b partial[b*k + j]
  __global__ void __launch_bounds__(RB)
    k_finish(int nb, int k, const double *__restrict__ partial, double *__restrict__ out,
             int do_sqrt)
  {
    __shared__ double sh[RB / 64];
    for (int j = 0; j < k; ++j)
      {
        double s = 0.0;
        for (int b = threadIdx.x; b < nb; b += RB)
          s += partial[(int64_t)b * k + j];
        s = block_reduce(s, sh);
        if (threadIdx.x == 0)
          out[j] = do_sqrt ? sqrt(s) : s;
      }
  }

  // h[j] = <V_j, w>, j < k (k <= MAXK)
  template <int KK>
  __global__ void __launch_bounds__(RB)
    k_multi_dot(int64_t n, int k, const double *__restrict__ V, const double *__restrict__ w,
                double *__restrict__ partial)
  {
    __shared__ double sh[RB / 64];
    double            acc[KK];
#pragma unroll
    for (int j = 0; j < KK; ++j)
      acc[j] = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * RB)
      {
        const double wi = w[i];
#pragma unroll
        for (int j = 0; j < KK; ++j)
          if (j < k)
            acc[j] += V[(int64_t)j * n + i] * wi;
      }
#pragma unroll
    for (int j = 0; j < KK; ++j)
      if (j < k)
        {
          const double s = block_reduce(acc[j], sh);
          if (threadIdx.x == 0)
            partial[(int64_t)blockIdx.x * k + j] = s;
        }
  }

  // w -= sum_j h[j] V_j ; partial norms of the updated w
  template <int KK>
  __global__ void __launch_bounds__(RB)
    k_multi_axpy_norm(int64_t n, int k, const double *__restrict__ V,
                      const double *__restrict__ h, double *__restrict__ w,
                      double *__restrict__ partial)
  {
    __shared__ double sh[RB / 64];
    double            hk[KK];
#pragma unroll
    for (int j = 0; j < KK; ++j)
      hk[j] = j < k ? h[j] : 0.0;
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * RB)
      {
        double wi = w[i];
#pragma unroll
        for (int j = 0; j < KK; ++j)
          if (j < k)
            wi -= hk[j] * V[(int64_t)j * n + i];
        w[i] = wi;
        s += wi * wi;
      }
    s = block_reduce(s, sh);
    if (threadIdx.x == 0)
      partial[blockIdx.x] = s;
  }

  // per-component sums of squares over an interleaved vector
  __global__ void __launch_bounds__(RB)
    k_block_sq(int64_t n_nodes, int nc, const double *__restrict__ x,
               double *__restrict__ partial)
  {
    __shared__ double sh[RB / 64];
    for (int c = 0; c < nc; ++c)
      {
        double s = 0.0;
        for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n_nodes;
             i += (int64_t)gridDim.x * RB)
          {
            const double v = x[i * nc + c];
            s += v * v;
          }
        s = block_reduce(s, sh);
        if (threadIdx.x == 0)
          partial[(int64_t)blockIdx.x * nc + c] = s;
      }
  }

  __global__ void k_axpy(int64_t n, double a, const double *__restrict__ x,
                         double *__restrict__ y)
  {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
      y[i] += a * x[i];
  }

  __global__ void k_sadd(int64_t n, double s, double a, const double *__restrict__ x,
                         double *__restrict__ y)
  {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
      y[i] = s * y[i] + a * x[i];
  }

  __global__ void k_scale(int64_t n, double a, double *__restrict__ x)
  {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
      x[i] *= a;
  }

  __global__ void k_fill(int64_t n, double a, double *__restrict__ x)
  {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
      x[i] = a;
  }

  template <typename A, typename B>
  __global__ void k_convert(int64_t n, B *__restrict__ dst, const A *__restrict__ src)
  {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
      dst[i] = (B)src[i];
  }

  // OperatorBase::compute_inverse_diagonal, S/operator_base.h:385-388
  __global__ void k_invert_diag(int64_t n, const double *__restrict__ d,
                                double *__restrict__ inv64, float *__restrict__ inv32)
  {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
      {
        const double v = d[i];
        const double r = (fabs(v) > 1.0e-10) ? (1.0 / v) : 1.0;
        inv64[i]       = r;
        inv32[i]       = (float)r;
      }
  }

  template <typename T>
  __global__ void k_diag_apply(int64_t n, const T *__restrict__ inv, const T *__restrict__ src,
                               T *__restrict__ dst)
  {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
      dst[i] = inv[i] * src[i];
  }

  // Inverse of every node's n_c x n_c block by Gauss-Jordan elimination with
  // partial pivoting on [A | I], one node per thread (the result of
  // FullMatrix::gauss_jordan in InverseComponentBlockDiagonalMatrix::
  // do_update, I/lac/preconditioners.h:288-308)
  template <int NC>
  __global__ void k_invert_blocks(int64_t n_nodes, const double *__restrict__ blocks,
                                  double *__restrict__ inv64, float *__restrict__ inv32,
                                  int *__restrict__ singular)
  {
    const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= n_nodes)
      return;
    double a[NC][NC], b[NC][NC];
    for (int i = 0; i < NC; ++i)
      for (int j = 0; j < NC; ++j)
        {
          a[i][j] = blocks[(node * NC + i) * NC + j];
          b[i][j] = i == j ? 1.0 : 0.0;
        }
    for (int j = 0; j < NC; ++j)
      {
        int    r   = j;
        double max = fabs(a[j][j]);
        for (int i = j + 1; i < NC; ++i)
          if (fabs(a[i][j]) > max)
            {
              max = fabs(a[i][j]);
              r   = i;
            }
        if (max == 0.0)
          {
            atomicAdd(singular, 1);
            return;
          }
        if (r != j)
          for (int k = 0; k < NC; ++k)
            {
              double t = a[j][k];
              a[j][k]  = a[r][k];
              a[r][k]  = t;
              t        = b[j][k];
              b[j][k]  = b[r][k];
              b[r][k]  = t;
            }
        const double hr = 1.0 / a[j][j];
        for (int k = 0; k < NC; ++k)
          {
            a[j][k] *= hr;
            b[j][k] *= hr;
          }
        for (int i = 0; i < NC; ++i)
          {
            if (i == j)
              continue;
            const double f = a[i][j];
            for (int k = 0; k < NC; ++k)
              {
                a[i][k] -= f * a[j][k];
                b[i][k] -= f * b[j][k];
              }
          }
      }
    for (int i = 0; i < NC; ++i)
      for (int j = 0; j < NC; ++j)
        {
          inv64[(node * NC + i) * NC + j] = b[i][j];
          inv32[(node * NC + i) * NC + j] = (float)b[i][j];
        }
  }

  template <int NC, typename T>
  __global__ void k_block_apply(int64_t n_nodes, const T *__restrict__ inv,
                                const T *__restrict__ src, T *__restrict__ dst)
  {
    for (int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; node < n_nodes;
         node += (int64_t)gridDim.x * blockDim.x)
      {
        T s[NC];
#pragma unroll
        for (int j = 0; j < NC; ++j)
          s[j] = src[node * NC + j];
#pragma unroll
        for (int i = 0; i < NC; ++i)
          {
            T acc = T(0);
#pragma unroll
            for (int j = 0; j < NC; ++j)
              acc += inv[(node * NC + i) * NC + j] * s[j];
            dst[node * NC + i] = acc;
          }
      }
  }

  __global__ void k_from_blocks(int64_t n_nodes, int nc, double *__restrict__ x,
                                const double *const *__restrict__ blocks)
  {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_nodes * nc;
         i += (int64_t)gridDim.x * blockDim.x)
      x[i] = blocks[i % nc][i / nc];
  }

  __global__ void k_to_blocks(int64_t n_nodes, int nc, double *const *__restrict__ blocks,
                              const double *__restrict__ x)
  {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_nodes * nc;
         i += (int64_t)gridDim.x * blockDim.x)
      blocks[i % nc][i / nc] = x[i];
  }

  // ----------------------------------------------------------- launchers
  static int grid_for(int64_t n, int per_block, int cap = 8192)
  {
    int64_t g = (n + per_block - 1) / per_block;
    if (g < 1)
      g = 1;
    if (g > cap)
      g = cap;
    return (int)g;
  }

  void blas_dot(hipStream_t s, int64_t n, const double *x, const double *y,
                double *partial, double *out)
  {
    k_dot<<<RED_BLOCKS, RB, 0, s>>>(n, x, y, partial);
    k_finish<<<1, RB, 0, s>>>(RED_BLOCKS, 1, partial, out, 0);
  }

  void blas_norm2(hipStream_t s, int64_t n, const double *x, double *partial, double *out)
  {
    k_dot<<<RED_BLOCKS, RB, 0, s>>>(n, x, x, partial);
    k_finish<<<1, RB, 0, s>>>(RED_BLOCKS, 1, partial, out, 1);
  }

  void blas_block_norms(hipStream_t s, int64_t n_nodes, int nc, const double *x,
                        double *partial, double *out)
  {
    k_block_sq<<<RED_BLOCKS, RB, 0, s>>>(n_nodes, nc, x, partial);
    k_finish<<<1, RB, 0, s>>>(RED_BLOCKS, nc, partial, out, 1);
  }

  void blas_multi_dot(hipStream_t s, int64_t n, int k, const double *V, const double *w,
                      double *partial, double *out)
  {
    if (k <= 8)
      k_multi_dot<8><<<RED_BLOCKS, RB, 0, s>>>(n, k, V, w, partial);
    else if (k <= 16)
      k_multi_dot<16><<<RED_BLOCKS, RB, 0, s>>>(n, k, V, w, partial);
    else
      k_multi_dot<MAXK><<<RED_BLOCKS, RB, 0, s>>>(n, k, V, w, partial);
    k_finish<<<1, RB, 0, s>>>(RED_BLOCKS, k, partial, out, 0);
  }

  void blas_multi_axpy_norm(hipStream_t s, int64_t n, int k, const double *V,
                            const double *h, double *w, double *partial, double *out)
  {
    if (k <= 8)
      k_multi_axpy_norm<8><<<RED_BLOCKS, RB, 0, s>>>(n, k, V, h, w, partial);
    else if (k <= 16)
      k_multi_axpy_norm<16><<<RED_BLOCKS, RB, 0, s>>>(n, k, V, h, w, partial);
    else
      k_multi_axpy_norm<MAXK><<<RED_BLOCKS, RB, 0, s>>>(n, k, V, h, w, partial);
    k_finish<<<1, RB, 0, s>>>(RED_BLOCKS, 1, partial, out, 1);
  }

  void blas_axpy(hipStream_t s, int64_t n, double a, const double *x, double *y)
  {
    k_axpy<<<grid_for(n, 256 * 4), 256, 0, s>>>(n, a, x, y);
  }
  void blas_sadd(hipStream_t s, int64_t n, double sc, double a, const double *x, double *y)
  {
    k_sadd<<<grid_for(n, 256 * 4), 256, 0, s>>>(n, sc, a, x, y);
  }
  void blas_scale(hipStream_t s, int64_t n, double a, double *x)
  {
    k_scale<<<grid_for(n, 256 * 4), 256, 0, s>>>(n, a, x);
  }
  void blas_fill(hipStream_t s, int64_t n, double a, double *x)
  {
    k_fill<<<grid_for(n, 256 * 4), 256, 0, s>>>(n, a, x);
  }
  void blas_f64_to_f32(hipStream_t s, int64_t n, float *dst, const double *src)
  {
    k_convert<double, float><<<grid_for(n, 256 * 4), 256, 0, s>>>(n, dst, src);
  }
  void blas_f32_to_f64(hipStream_t s, int64_t n, double *dst, const float *src)
  {
    k_convert<float, double><<<grid_for(n, 256 * 4), 256, 0, s>>>(n, dst, src);
  }
  void invert_diag(hipStream_t s, int64_t n, const double *d, double *inv64, float *inv32)
  {
    k_invert_diag<<<grid_for(n, 256 * 4), 256, 0, s>>>(n, d, inv64, inv32);
  }
  void diag_apply64(hipStream_t s, int64_t n, const double *inv, const double *src,
                    double *dst)
  {
    k_diag_apply<double><<<grid_for(n, 256 * 4), 256, 0, s>>>(n, inv, src, dst);
  }
  void diag_apply32(hipStream_t s, int64_t n, const float *inv, const float *src, float *dst)
  {
    k_diag_apply<float><<<grid_for(n, 256 * 4), 256, 0, s>>>(n, inv, src, dst);
  }

#define HPS_NC_SWITCH(nc, CALL) \
  switch (nc)                   \
    {                           \
      case 3:                   \
        CALL(3);                \
        break;                  \
      case 4:                   \
        CALL(4);                \
        break;                  \
      case 5:                   \
        CALL(5);                \
        break;                  \
      case 6:                   \
        CALL(6);                \
        break;                  \
      case 7:                   \
        CALL(7);                \
        break;                  \
      case 8:                   \
        CALL(8);                \
        break;                  \
      case 10:                  \
        CALL(10);               \
        break;                  \
      case 12:                  \
        CALL(12);               \
        break;                  \
      case 14:                  \
        CALL(14);               \
        break;                  \
      default:                  \
        return false;           \
    }

  bool invert_blocks(hipStream_t s, int64_t n_nodes, int nc, const double *blocks,
                     double *inv64, float *inv32, int *singular)
  {
    const int g = (int)((n_nodes + 127) / 128);
#define HPS_CALL(N) k_invert_blocks<N><<<g, 128, 0, s>>>(n_nodes, blocks, inv64, inv32, singular)
    HPS_NC_SWITCH(nc, HPS_CALL)
#undef HPS_CALL
    return true;
  }

  bool block_apply64(hipStream_t s, int64_t n_nodes, int nc, const double *inv,
                     const double *src, double *dst)
  {
    const int g = grid_for(n_nodes, 256);
#define HPS_CALL(N) k_block_apply<N, double><<<g, 256, 0, s>>>(n_nodes, inv, src, dst)
    HPS_NC_SWITCH(nc, HPS_CALL)
#undef HPS_CALL
    return true;
  }

  bool block_apply32(hipStream_t s, int64_t n_nodes, int nc, const float *inv,
                     const float *src, float *dst)
  {
    const int g = grid_for(n_nodes, 256);
#define HPS_CALL(N) k_block_apply<N, float><<<g, 256, 0, s>>>(n_nodes, inv, src, dst)
    HPS_NC_SWITCH(nc, HPS_CALL)
#undef HPS_CALL
    return true;
  }

  void from_blocks(hipStream_t s, int64_t n_nodes, int nc, double *x,
                   const double *const *blocks_dev)
  {
    k_from_blocks<<<grid_for(n_nodes * nc, 256 * 4), 256, 0, s>>>(n_nodes, nc, x, blocks_dev);
  }
  void to_blocks(hipStream_t s, int64_t n_nodes, int nc, double *const *blocks_dev,
                 const double *x)
  {
    k_to_blocks<<<grid_for(n_nodes * nc, 256 * 4), 256, 0, s>>>(n_nodes, nc, blocks_dev, x);
  }
  // ---- FP32 storage, FP64 accumulation (mixed-precision Krylov basis)
  template <int KK>
  __global__ void __launch_bounds__(RB)
    k_multi_dot_f32(int64_t n, int k, const float *__restrict__ V, const float *__restrict__ w,
                    double *__restrict__ partial)
  {
    __shared__ double sh[RB / 64];
    double            acc[KK];
#pragma unroll
    for (int j = 0; j < KK; ++j)
      acc[j] = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * RB)
      {
        const double wi = w[i];
#pragma unroll
        for (int j = 0; j < KK; ++j)
          if (j < k)
            acc[j] += (double)V[(int64_t)j * n + i] * wi;
      }
#pragma unroll
    for (int j = 0; j < KK; ++j)
      if (j < k)
        {
          const double s = block_reduce(acc[j], sh);
          if (threadIdx.x == 0)
            partial[(int64_t)blockIdx.x * k + j] = s;
        }
  }

  template <int KK>
  __global__ void __launch_bounds__(RB)
    k_multi_axpy_norm_f32(int64_t n, int k, const float *__restrict__ V,
                          const double *__restrict__ h, float *__restrict__ w,
                          double *__restrict__ partial)
  {
    __shared__ double sh[RB / 64];
    double            hk[KK];
#pragma unroll
    for (int j = 0; j < KK; ++j)
      hk[j] = j < k ? h[j] : 0.0;
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * RB)
      {
        double wi = w[i];
#pragma unroll
        for (int j = 0; j < KK; ++j)
          if (j < k)
            wi -= hk[j] * (double)V[(int64_t)j * n + i];
        const float wf = (float)wi;
        w[i]           = wf;
        s += (double)wf * (double)wf;
      }
    s = block_reduce(s, sh);
    if (threadIdx.x == 0)
      partial[blockIdx.x] = s;
  }

  __global__ void __launch_bounds__(RB)
    k_norm_f32(int64_t n, const float *__restrict__ x, double *__restrict__ partial)
  {
    __shared__ double sh[RB / 64];
    double            s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * RB)
      s += (double)x[i] * (double)x[i];
    s = block_reduce(s, sh);
    if (threadIdx.x == 0)
      partial[blockIdx.x] = s;
  }

  __global__ void k_scale_f32(int64_t n, double a, float *__restrict__ x)
  {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
      x[i] = (float)(a * (double)x[i]);
  }

  // y (FP64) += sum_j a_j x_j (FP32), j < k
  __global__ void k_axpy_multi_f32_f64(int64_t n, int k, const double *__restrict__ a,
                                       const float *__restrict__ X, double *__restrict__ y)
  {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
      {
        double s = y[i];
        for (int j = 0; j < k; ++j)
          s += a[j] * (double)X[(int64_t)j * n + i];
        y[i] = s;
      }
  }

  void blas_multi_dot_f32(hipStream_t s, int64_t n, int k, const float *V, const float *w,
                          double *partial, double *out)
  {
    if (k <= 8)
      k_multi_dot_f32<8><<<RED_BLOCKS, RB, 0, s>>>(n, k, V, w, partial);
    else if (k <= 16)
      k_multi_dot_f32<16><<<RED_BLOCKS, RB, 0, s>>>(n, k, V, w, partial);
    else
      k_multi_dot_f32<MAXK><<<RED_BLOCKS, RB, 0, s>>>(n, k, V, w, partial);
    k_finish<<<1, RB, 0, s>>>(RED_BLOCKS, k, partial, out, 0);
  }

  void blas_multi_axpy_norm_f32(hipStream_t s, int64_t n, int k, const float *V,
                                const double *h, float *w, double *partial, double *out)
  {
    if (k <= 8)
      k_multi_axpy_norm_f32<8><<<RED_BLOCKS, RB, 0, s>>>(n, k, V, h, w, partial);
    else if (k <= 16)
      k_multi_axpy_norm_f32<16><<<RED_BLOCKS, RB, 0, s>>>(n, k, V, h, w, partial);
    else
      k_multi_axpy_norm_f32<MAXK><<<RED_BLOCKS, RB, 0, s>>>(n, k, V, h, w, partial);
    k_finish<<<1, RB, 0, s>>>(RED_BLOCKS, 1, partial, out, 1);
  }

  void blas_norm2_f32(hipStream_t s, int64_t n, const float *x, double *partial, double *out)
  {
    k_norm_f32<<<RED_BLOCKS, RB, 0, s>>>(n, x, partial);
    k_finish<<<1, RB, 0, s>>>(RED_BLOCKS, 1, partial, out, 1);
  }

  void blas_scale_f32(hipStream_t s, int64_t n, double a, float *x)
  {
    k_scale_f32<<<grid_for(n, 256 * 4), 256, 0, s>>>(n, a, x);
  }

  void blas_axpy_multi_f32_f64(hipStream_t s, int64_t n, int k, const double *a, const float *X,
                               double *y)
  {
    k_axpy_multi_f32_f64<<<grid_for(n, 256 * 4), 256, 0, s>>>(n, k, a, X, y);
  }
} // namespace hps
