// blas.hip -- block-vector BLAS-1 and Jacobi preconditioner kernels.
//
// DynamicBlockVector operations (I/lac/dynamic_block_vector.h:346-537) on
// the component-interleaved storage, the inverse-diagonal and inverse
// node-block applies (I/lac/preconditioners.h:160-318) and the layout
// conversions (I/numerics/vector_tools.h:40-100).  Reductions use a fixed
// grid and a fixed second-pass order, so results do not depend on timing.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

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

  // sums the RED_BLOCKS partial rows of k values each: out[j] = sum_b
  // partial[b*k + j].  Block j owns column j; thread t loads rows t, t + RB,
  // ... all at once and adds them in row order, then the wave trees and the
  // waves in wave order (fixed order: deterministic).
  static_assert(RED_BLOCKS % RB == 0, "finish rows per thread");
  __global__ void __launch_bounds__(RB)
    k_finish(int k, const double *__restrict__ partial, double *__restrict__ out, int do_sqrt)
  {
    constexpr int     PER = RED_BLOCKS / RB;
    __shared__ double sh[RB / 64];
    const int         j = blockIdx.x;
    double            v[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i)
      v[i] = partial[(int64_t)(threadIdx.x + i * RB) * k + j];
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < PER; ++i)
      s += v[i];
    s = block_reduce(s, sh);
    if (threadIdx.x == 0)
      out[j] = do_sqrt ? sqrt(s) : s;
  }

  static void finish(hipStream_t s, int k, const double *partial, double *out, int do_sqrt)
  {
    k_finish<<<k, RB, 0, s>>>(k, partial, out, do_sqrt);
  }

  // K values reduced over the block in one pass: per-wave shuffle trees,
  // then thread j < K sums the waves' partials in wave order
  template <int K>
  __device__ __forceinline__ void block_reduce_k(double *acc, double (*sh)[K],
                                                 double *__restrict__ out)
  {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < K; ++j)
      {
        double v = acc[j];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1)
          v += __shfl_down(v, o, 64);
        if (lane == 0)
          sh[w][j] = v;
      }
    __syncthreads();
    if (threadIdx.x < K)
      {
        double t = 0.0;
#pragma unroll
        for (int i = 0; i < RB / 64; ++i)
          t += sh[i][threadIdx.x];
        out[threadIdx.x] = t;
      }
  }

  // grid-stride steps unrolled so that a thread has >= 8 loads in flight
  template <int K>
  constexpr int unroll_for()
  {
    return K >= 8 ? 1 : (K >= 4 ? 2 : 4);
  }

  // h[j] = <V_j, w>, j < K; rows of V have length n; storage T (FP64 or
  // FP32), FP64 accumulation.  K is exact, so every row load of a step is
  // issued before the first is consumed.
  // (partial row of block b: partial[b * pstride + pcol + j])
  template <int K, typename T>
  __global__ void __launch_bounds__(RB)
    k_mdot(int64_t n, const T *__restrict__ V, const T *__restrict__ w,
           double *__restrict__ partial, int pstride, int pcol)
  {
    constexpr int     U = unroll_for<K>();
    __shared__ double sh[RB / 64][K];
    double            acc[K];
#pragma unroll
    for (int j = 0; j < K; ++j)
      acc[j] = 0.0;
    const int64_t st = (int64_t)gridDim.x * RB;
    int64_t       i  = (int64_t)blockIdx.x * RB + threadIdx.x;
    for (; i + (U - 1) * st < n; i += U * st)
      {
        T wv[U], vv[U][K];
#pragma unroll
        for (int u = 0; u < U; ++u)
          {
            wv[u] = w[i + u * st];
#pragma unroll
            for (int j = 0; j < K; ++j)
              vv[u][j] = V[(int64_t)j * n + i + u * st];
          }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int j = 0; j < K; ++j)
            acc[j] += (double)vv[u][j] * (double)wv[u];
      }
    for (; i < n; i += st)
      {
        const double wi = w[i];
#pragma unroll
        for (int j = 0; j < K; ++j)
          acc[j] += (double)V[(int64_t)j * n + i] * wi;
      }
    block_reduce_k<K>(acc, sh, partial + (int64_t)blockIdx.x * pstride + pcol);
  }

  // w -= sum_j h[j] V_j (FP64 arithmetic, stored as T); partial sums of
  // squares of the stored w
  template <int K, typename T>
  __global__ void __launch_bounds__(RB)
    k_maxpy_norm(int64_t n, const T *__restrict__ V, const double *__restrict__ h,
                 T *__restrict__ w, double *__restrict__ partial)
  {
    constexpr int     U = unroll_for<K>();
    __shared__ double sh[RB / 64][1];
    double            hk[K];
#pragma unroll
    for (int j = 0; j < K; ++j)
      hk[j] = h[j];
    double        s  = 0.0;
    const int64_t st = (int64_t)gridDim.x * RB;
    int64_t       i  = (int64_t)blockIdx.x * RB + threadIdx.x;
    for (; i + (U - 1) * st < n; i += U * st)
      {
        T wv[U], vv[U][K];
#pragma unroll
        for (int u = 0; u < U; ++u)
          {
            wv[u] = w[i + u * st];
#pragma unroll
            for (int j = 0; j < K; ++j)
              vv[u][j] = V[(int64_t)j * n + i + u * st];
          }
#pragma unroll
        for (int u = 0; u < U; ++u)
          {
            double wi = wv[u];
#pragma unroll
            for (int j = 0; j < K; ++j)
              wi -= hk[j] * (double)vv[u][j];
            const T wt      = (T)wi;
            w[i + u * st]   = wt;
            s += (double)wt * (double)wt;
          }
      }
    for (; i < n; i += st)
      {
        double wi = w[i];
#pragma unroll
        for (int j = 0; j < K; ++j)
          wi -= hk[j] * (double)V[(int64_t)j * n + i];
        const T wt = (T)wi;
        w[i]       = wt;
        s += (double)wt * (double)wt;
      }
    double acc[1] = {s};
    block_reduce_k<1>(acc, sh, partial + blockIdx.x);
  }

  // y (FP64) += sum_j a_j X_j (storage T), j < K
  template <int K, typename T>
  __global__ void k_axpy_multi(int64_t n, const double *__restrict__ a, const T *__restrict__ X,
                               double *__restrict__ y)
  {
    double ak[K];
#pragma unroll
    for (int j = 0; j < K; ++j)
      ak[j] = a[j];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
      {
        T xv[K];
#pragma unroll
        for (int j = 0; j < K; ++j)
          xv[j] = X[(int64_t)j * n + i];
        double s = y[i];
#pragma unroll
        for (int j = 0; j < K; ++j)
          s += ak[j] * (double)xv[j];
        y[i] = s;
      }
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
  __global__ void __launch_bounds__(64)
    k_invert_blocks(int64_t n_nodes, int64_t ist, const double *__restrict__ blocks,
                                  double *__restrict__ inv64, float *__restrict__ inv32,
                                  int *__restrict__ singular)
  {
    const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= n_nodes)
      return;
    // fully unrolled; the pivot row swap is done by selects, so both
    // matrices stay in registers
    double a[NC][NC], b[NC][NC];
#pragma unroll
    for (int i = 0; i < NC; ++i)
#pragma unroll
      for (int j = 0; j < NC; ++j)
        {
          a[i][j] = blocks[(node * NC + j) * NC + i]; // stored column by column
          b[i][j] = i == j ? 1.0 : 0.0;
        }
#pragma unroll
    for (int j = 0; j < NC; ++j)
      {
        int    r   = j;
        double max = fabs(a[j][j]);
#pragma unroll
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
#pragma unroll
        for (int i = j + 1; i < NC; ++i)
          {
            const bool sw = i == r;
#pragma unroll
            for (int k = 0; k < NC; ++k)
              {
                const double aj = a[j][k], ai = a[i][k], bj = b[j][k], bi = b[i][k];
                a[j][k]         = sw ? ai : aj;
                a[i][k]         = sw ? aj : ai;
                b[j][k]         = sw ? bi : bj;
                b[i][k]         = sw ? bj : bi;
              }
          }
        const double hr = 1.0 / a[j][j];
#pragma unroll
        for (int k = 0; k < NC; ++k)
          {
            a[j][k] *= hr;
            b[j][k] *= hr;
          }
#pragma unroll
        for (int i = 0; i < NC; ++i)
          {
            if (i == j)
              continue;
            const double f = a[i][j];
#pragma unroll
            for (int k = 0; k < NC; ++k)
              {
                a[i][k] -= f * a[j][k];
                b[i][k] -= f * b[j][k];
              }
          }
      }
    // inverses entry-major ([row * NC + column][node]): the apply reads
    // each entry for consecutive nodes in one coalesced access
    for (int i = 0; i < NC; ++i)
      for (int j = 0; j < NC; ++j)
        {
          inv64[(int64_t)(i * NC + j) * ist + node] = b[i][j];
          inv32[(int64_t)(i * NC + j) * ist + node] = (float)b[i][j];
        }
  }

  // the same inversion for larger blocks, whose two matrices do not fit in
  // one lane's registers: a group of 16 lanes per node block, lane i holds
  // row i of the block and of the inverse.  Partial pivoting by the largest
  // magnitude among the rows not yet used (lowest row on ties); pivot rows
  // are not moved, so the lane that pivots column j ends with row j of the
  // inverse.  The normalised pivot row reaches the group through LDS.
  // (16 groups per workgroup: the entry-major writes of the inverses cover
  // 16 consecutive nodes)
  template <int NC>
  __global__ void __launch_bounds__(256)
    k_invert_blocks_rows(int64_t n_nodes, int64_t ist, const double *__restrict__ blocks,
                         double *__restrict__ inv64, float *__restrict__ inv32,
                         int *__restrict__ singular)
  {
    static_assert(NC <= 16, "one row per lane of a 16-lane group");
    __shared__ double piv[16][2 * NC];
    const int     grp  = threadIdx.x >> 4, i = threadIdx.x & 15;
    const int64_t node = (int64_t)blockIdx.x * 16 + grp;
    const bool    live = node < n_nodes && i < NC;
    double        a[NC], b[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k)
      {
        a[k] = live ? blocks[(node * NC + k) * NC + i] : 0.0; // stored column by column
        b[k] = k == i ? 1.0 : 0.0;
      }
    bool used = !live; // rows of absent lanes never pivot
    bool sing = false;
    int  col  = 0;     // the column this lane pivoted
#pragma unroll
    for (int j = 0; j < NC; ++j)
      {
        double v = used ? -1.0 : fabs(a[j]);
        int    r = i;
#pragma unroll
        for (int m = 1; m < 16; m <<= 1)
          {
            const double ov = __shfl_xor(v, m, 16);
            const int    orr = __shfl_xor(r, m, 16);
            if (ov > v || (ov == v && orr < r))
              {
                v = ov;
                r = orr;
              }
          }
        sing         = sing || !(v > 0.0);
        const bool me = i == r;
        if (me)
          {
            const double hr = sing ? 0.0 : 1.0 / a[j];
#pragma unroll
            for (int k = 0; k < NC; ++k)
              {
                a[k] *= hr;
                b[k] *= hr;
                piv[grp][k]      = a[k];
                piv[grp][NC + k] = b[k];
              }
            used = true;
            col  = j;
          }
        __syncthreads();
        if (!me && live)
          {
            const double f = a[j];
#pragma unroll
            for (int k = 0; k < NC; ++k)
              {
                a[k] -= f * piv[grp][k];
                b[k] -= f * piv[grp][NC + k];
              }
          }
        __syncthreads();
      }
    if (sing)
      {
        if (i == 0 && node < n_nodes)
          atomicAdd(singular, 1);
        return;
      }
    if (live)
#pragma unroll
      for (int k = 0; k < NC; ++k)
        {
          inv64[(int64_t)(col * NC + k) * ist + node] = b[k];
          inv32[(int64_t)(col * NC + k) * ist + node] = (float)b[k];
        }
  }

  // dst_n = B_n^-1 src_n for the nodes of a chunk of 256 NPT, thread t
  // taking the NPT adjacent nodes NPT t ...  The chunk's src and dst rows
  // (256 NPT NC consecutive values) move between HBM and LDS as 16-byte
  // vectors, so the strided per-node rows are never read or written lane by
  // lane from HBM; the inverse blocks are read entry-major (entry e of node
  // n at e ist + n, ist a multiple of 4), the NPT nodes of a lane in one
  // vector access, so that one access of a wave covers 256 NPT consecutive
  // values of an entry.  The chunk's src is read completely before any of
  // its dst is written, so dst may be src.
  template <int NC, typename T, int NPT>
  __global__ void __launch_bounds__(256)
    k_block_apply(int64_t n_nodes, int64_t ist, const T *__restrict__ inv, const T *src, T *dst)
  {
    static_assert(NPT == 1 || NPT == 2, "one or two nodes per thread");
    constexpr int NB = 256 * NPT;
    constexpr int W  = 16 / (int)sizeof(T);
    using V          = typename std::conditional<sizeof(T) == 8, double2, float4>::type;
    using P          = typename std::conditional<sizeof(T) == 8, double2, float2>::type;
    __shared__ T  buf[NB * NC];
    const int64_t n0  = (int64_t)blockIdx.x * NB;
    const int     nb  = n_nodes - n0 < NB ? (int)(n_nodes - n0) : NB;
    const int     len = nb * NC;
    const int     t   = (int)threadIdx.x;
    const T      *s0  = src + n0 * NC;
    T            *d0  = dst + n0 * NC;
    const bool    vec = ((reinterpret_cast<uintptr_t>(s0) | reinterpret_cast<uintptr_t>(d0)) & 15) == 0;
    const int     nv  = vec ? len / W : 0; // whole vectors; the rest by value
    for (int v = t; v < nv; v += 256)
      {
        const V x = reinterpret_cast<const V *>(s0)[v];
        __builtin_memcpy(buf + v * W, &x, 16);
      }
    for (int e = nv * W + t; e < len; e += 256)
      buf[e] = s0[e];
    __syncthreads();
    const int tn = NPT * t; // first node of the thread in the chunk
    T         out[NPT][NC];
    if (tn < nb)
      {
        T x[NPT][NC];
#pragma unroll
        for (int p = 0; p < NPT; ++p)
#pragma unroll
          for (int j = 0; j < NC; ++j)
            x[p][j] = buf[(tn + p) * NC + j]; // the chunk's rows past nb hold junk: unused
        const T *ib = inv + n0 + tn;
#pragma unroll
        for (int i = 0; i < NC; ++i)
          {
            T acc[NPT];
#pragma unroll
            for (int p = 0; p < NPT; ++p)
              acc[p] = T(0);
#pragma unroll
            for (int j = 0; j < NC; ++j)
              {
                const T *e = ib + (int64_t)(i * NC + j) * ist;
                if constexpr (NPT == 1)
                  acc[0] += e[0] * x[0][j];
                else
                  {
                    const P y = *reinterpret_cast<const P *>(e);
                    acc[0] += y.x * x[0][j];
                    acc[1] += y.y * x[1][j];
                  }
              }
#pragma unroll
            for (int p = 0; p < NPT; ++p)
              out[p][i] = acc[p];
          }
      }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < NPT; ++p)
      if (tn + p < nb)
#pragma unroll
        for (int i = 0; i < NC; ++i)
          buf[(tn + p) * NC + i] = out[p][i];
    __syncthreads();
    for (int v = t; v < nv; v += 256)
      {
        V x;
        __builtin_memcpy(&x, buf + v * W, 16);
        reinterpret_cast<V *>(d0)[v] = x;
      }
    for (int e = nv * W + t; e < len; e += 256)
      d0[e] = buf[e];
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
    finish(s, 1, partial, out, 0);
  }

  void blas_norm2(hipStream_t s, int64_t n, const double *x, double *partial, double *out)
  {
    k_dot<<<RED_BLOCKS, RB, 0, s>>>(n, x, x, partial);
    finish(s, 1, partial, out, 1);
  }

  void blas_block_norms(hipStream_t s, int64_t n_nodes, int nc, const double *x,
                        double *partial, double *out)
  {
    k_block_sq<<<RED_BLOCKS, RB, 0, s>>>(n_nodes, nc, x, partial);
    finish(s, nc, partial, out, 1);
  }

#define HPS_K_SWITCH(k, CALL)                                                  \
  switch (k)                                                                   \
    {                                                                          \
      case 1: CALL(1); break;   case 2: CALL(2); break;   case 3: CALL(3); break;   \
      case 4: CALL(4); break;   case 5: CALL(5); break;   case 6: CALL(6); break;   \
      case 7: CALL(7); break;   case 8: CALL(8); break;   case 9: CALL(9); break;   \
      case 10: CALL(10); break; case 11: CALL(11); break; case 12: CALL(12); break; \
      case 13: CALL(13); break; case 14: CALL(14); break; case 15: CALL(15); break; \
      case 16: CALL(16); break; case 17: CALL(17); break; case 18: CALL(18); break; \
      case 19: CALL(19); break; case 20: CALL(20); break; case 21: CALL(21); break; \
      case 22: CALL(22); break; case 23: CALL(23); break; case 24: CALL(24); break; \
      case 25: CALL(25); break; case 26: CALL(26); break; case 27: CALL(27); break; \
      case 28: CALL(28); break; case 29: CALL(29); break; case 30: CALL(30); break; \
      case 31: CALL(31); break; case 32: CALL(32); break;                       \
      default: break;                                                          \
    }

  // rows in passes of at most 16 (beyond that the compiler serialises the
  // row loads of one step and the kernel drops to half bandwidth)
  template <typename T>
  static void multi_dot_any(hipStream_t s, int64_t n, int k, const T *V, const T *w,
                            double *partial, double *out)
  {
    for (int j0 = 0; j0 < k; j0 += 16)
      {
        const int kk = k - j0 < 16 ? k - j0 : 16;
        const T  *Vj = V + (int64_t)j0 * n;
#define HPS_CALL(K) k_mdot<K, T><<<RED_BLOCKS, RB, 0, s>>>(n, Vj, w, partial, k, j0)
        HPS_K_SWITCH(kk, HPS_CALL)
#undef HPS_CALL
      }
    finish(s, k, partial, out, 0);
  }

  template <typename T>
  static void multi_axpy_norm_any(hipStream_t s, int64_t n, int k, const T *V,
                                  const double *h, T *w, double *partial, double *out)
  {
#define HPS_CALL(K) k_maxpy_norm<K, T><<<RED_BLOCKS, RB, 0, s>>>(n, V, h, w, partial)
    HPS_K_SWITCH(k, HPS_CALL)
#undef HPS_CALL
    finish(s, 1, partial, out, 1);
  }

  void blas_multi_dot(hipStream_t s, int64_t n, int k, const double *V, const double *w,
                      double *partial, double *out)
  {
    multi_dot_any<double>(s, n, k, V, w, partial, out);
  }

  void blas_multi_axpy_norm(hipStream_t s, int64_t n, int k, const double *V,
                            const double *h, double *w, double *partial, double *out)
  {
    multi_axpy_norm_any<double>(s, n, k, V, h, w, partial, out);
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
      case 9:                   \
        CALL(9);                \
        break;                  \
      case 10:                  \
        CALL(10);               \
        break;                  \
      case 11:                  \
        CALL(11);               \
        break;                  \
      case 12:                  \
        CALL(12);               \
        break;                  \
      case 13:                  \
        CALL(13);               \
        break;                  \
      case 14:                  \
        CALL(14);               \
        break;                  \
      default:                  \
        return false;           \
    }

  // grain cut-off: a component with no entry in its node block's row and
  // column (an order parameter dropped in every cell around the node) gets a
  // unit diagonal, the rule of the point-Jacobi diagonal (|d| <= 1e-10 -> 1,
  // S/operator_base.h:364-389); blocks stored column by column per node.
  // One wave per node: lane l reads entries l, l + 64, ... of the node's
  // block (coalesced), marks the row and the column of each nonzero entry
  // in a bit mask, the masks are OR-reduced over the wave, and lane a sets
  // the diagonal of an order parameter a whose row and column are empty.
  template <int NC>
  __global__ void __launch_bounds__(256) k_unit_empty_rows(int64_t n_nodes, double *blocks)
  {
    static_assert(NC <= 32, "row/column bits in one mask");
    constexpr int NE   = NC * NC;
    constexpr int PER  = (NE + 63) / 64;
    const int     lane = (int)threadIdx.x & 63;
    const int64_t nw   = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t node = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; node < n_nodes;
         node += nw)
      {
        double  *b = blocks + node * NE;
        unsigned m = 0;
#pragma unroll
        for (int k = 0; k < PER; ++k)
          {
            const int e = lane + 64 * k;
            if (e < NE && b[e] != 0.0)
              {
                const int j = e / NC, i = e - j * NC; // column j, row i
                m |= (1u << i) | (1u << j);
              }
          }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1)
          m |= (unsigned)__shfl_xor((int)m, o, 64);
        if (lane >= 2 && lane < NC && ((m >> lane) & 1u) == 0)
          b[lane * NC + lane] = 1.0;
      }
  }

  bool unit_empty_rows(hipStream_t s, int64_t n_nodes, int nc, double *blocks)
  {
    const int g = grid_for(n_nodes * 64, 256);
#define HPS_CALL(N) k_unit_empty_rows<N><<<g, 256, 0, s>>>(n_nodes, blocks)
    HPS_NC_SWITCH(nc, HPS_CALL)
#undef HPS_CALL
    return true;
  }

  bool invert_blocks(hipStream_t s, int64_t n_nodes, int nc, const double *blocks,
                     double *inv64, float *inv32, int *singular)
  {
    // up to 6 components both matrices stay in one lane's registers
    const int g  = (int)((n_nodes + 63) / 64);
    const int g16 = (int)((n_nodes + 15) / 16);
    const int64_t ist = inv_stride(n_nodes);
#define HPS_CALL(N)                                                                          \
  if constexpr (N <= 6)                                                                      \
    k_invert_blocks<N><<<g, 64, 0, s>>>(n_nodes, ist, blocks, inv64, inv32, singular);       \
  else                                                                                       \
    k_invert_blocks_rows<N><<<g16, 256, 0, s>>>(n_nodes, ist, blocks, inv64, inv32, singular)
    HPS_NC_SWITCH(nc, HPS_CALL)
#undef HPS_CALL
    return true;
  }

  bool block_apply64(hipStream_t s, int64_t n_nodes, int nc, const double *inv,
                     const double *src, double *dst)
  {
    if (n_nodes <= 0)
      return true;
    const unsigned g = (unsigned)((n_nodes + 255) / 256);
#define HPS_CALL(N) \
  k_block_apply<N, double, 1><<<g, 256, 0, s>>>(n_nodes, inv_stride(n_nodes), inv, src, dst)
    HPS_NC_SWITCH(nc, HPS_CALL)
#undef HPS_CALL
    return true;
  }

  bool block_apply32(hipStream_t s, int64_t n_nodes, int nc, const float *inv,
                     const float *src, float *dst)
  {
    if (n_nodes <= 0)
      return true;
    // two nodes per thread from 9 components on (one access of a wave then
    // covers 512 consecutive values of an entry); below, the 512-node
    // chunks leave too few workgroups and one node per thread is faster
#define HPS_CALL(N)                                                                       \
  if constexpr (N >= 9)                                                                   \
    k_block_apply<N, float, 2><<<(unsigned)((n_nodes + 511) / 512), 256, 0, s>>>(         \
      n_nodes, inv_stride(n_nodes), inv, src, dst);                                       \
  else                                                                                    \
    k_block_apply<N, float, 1><<<(unsigned)((n_nodes + 255) / 256), 256, 0, s>>>(         \
      n_nodes, inv_stride(n_nodes), inv, src, dst)
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

  void blas_multi_dot_f32(hipStream_t s, int64_t n, int k, const float *V, const float *w,
                          double *partial, double *out)
  {
    multi_dot_any<float>(s, n, k, V, w, partial, out);
  }

  void blas_multi_axpy_norm_f32(hipStream_t s, int64_t n, int k, const float *V,
                                const double *h, float *w, double *partial, double *out)
  {
    multi_axpy_norm_any<float>(s, n, k, V, h, w, partial, out);
  }

  void blas_norm2_f32(hipStream_t s, int64_t n, const float *x, double *partial, double *out)
  {
    k_norm_f32<<<RED_BLOCKS, RB, 0, s>>>(n, x, partial);
    finish(s, 1, partial, out, 1);
  }

  void blas_scale_f32(hipStream_t s, int64_t n, double a, float *x)
  {
    k_scale_f32<<<grid_for(n, 256 * 4), 256, 0, s>>>(n, a, x);
  }

  void blas_axpy_multi_f64(hipStream_t s, int64_t n, int k, const double *a, const double *X,
                           double *y)
  {
#define HPS_CALL(K) k_axpy_multi<K, double><<<grid_for(n, 256 * 4), 256, 0, s>>>(n, a, X, y)
    HPS_K_SWITCH(k, HPS_CALL)
#undef HPS_CALL
  }

  void blas_axpy_multi_f32_f64(hipStream_t s, int64_t n, int k, const double *a, const float *X,
                               double *y)
  {
#define HPS_CALL(K) k_axpy_multi<K, float><<<grid_for(n, 256 * 4), 256, 0, s>>>(n, a, X, y)
    HPS_K_SWITCH(k, HPS_CALL)
#undef HPS_CALL
  }

  // ------------------------------------------------ constrained nodes
  // OperatorBase::vmult (S/operator_base.h:294-353): before the cell
  // operator the constrained entries of src are saved and zeroed; after it
  // src is restored and dst equals src on them (identity rows).  One item
  // per (constrained node, component).
  template <typename T>
  __global__ void k_constraint_zero(int64_t n, int nc, const int64_t *nodes, T *src, T *save)
  {
    const int64_t total = n * nc;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x)
      {
        const int64_t k = i / nc;
        T *p            = src + nodes[k] * nc + (i - k * nc);
        save[i]         = *p;
        *p              = T(0);
      }
  }

  template <typename T>
  __global__ void k_constraint_restore(int64_t n, int nc, const int64_t *nodes, T *src, T *dst,
                                       const T *save)
  {
    const int64_t total = n * nc;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x)
      {
        const int64_t k   = i / nc;
        const int64_t idx = nodes[k] * nc + (i - k * nc);
        src[idx]          = save[i];
        if (dst)
          dst[idx] = save[i];
      }
  }

  // constrained rows of dst set to value (residual rows: 0; point diagonal: 1)
  __global__ void k_constraint_set(int64_t n, int nc, const int64_t *nodes, double *dst,
                                   double value)
  {
    const int64_t total = n * nc;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x)
      {
        const int64_t k       = i / nc;
        dst[nodes[k] * nc + (i - k * nc)] = value;
      }
  }

  // node blocks of constrained nodes set to the identity
  __global__ void k_constraint_identity_blocks(int64_t n, int nc, const int64_t *nodes,
                                               double *blocks)
  {
    const int64_t per   = (int64_t)nc * nc;
    const int64_t total = n * per;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x)
      {
        const int64_t k = i / per;
        const int     e = (int)(i - k * per);
        blocks[nodes[k] * per + e] = (e / nc == e % nc) ? 1.0 : 0.0;
      }
  }

  void constraint_zero64(hipStream_t s, int64_t n, int nc, const int64_t *nodes, double *src,
                         double *save)
  {
    k_constraint_zero<double><<<grid_for(n * nc, 256), 256, 0, s>>>(n, nc, nodes, src, save);
  }
  void constraint_zero32(hipStream_t s, int64_t n, int nc, const int64_t *nodes, float *src,
                         float *save)
  {
    k_constraint_zero<float><<<grid_for(n * nc, 256), 256, 0, s>>>(n, nc, nodes, src, save);
  }
  void constraint_restore64(hipStream_t s, int64_t n, int nc, const int64_t *nodes, double *src,
                            double *dst, const double *save)
  {
    k_constraint_restore<double>
      <<<grid_for(n * nc, 256), 256, 0, s>>>(n, nc, nodes, src, dst, save);
  }
  void constraint_restore32(hipStream_t s, int64_t n, int nc, const int64_t *nodes, float *src,
                            float *dst, const float *save)
  {
    k_constraint_restore<float>
      <<<grid_for(n * nc, 256), 256, 0, s>>>(n, nc, nodes, src, dst, save);
  }
  void constraint_set(hipStream_t s, int64_t n, int nc, const int64_t *nodes, double *dst,
                      double value)
  {
    k_constraint_set<<<grid_for(n * nc, 256), 256, 0, s>>>(n, nc, nodes, dst, value);
  }
  void constraint_identity_blocks(hipStream_t s, int64_t n, int nc, const int64_t *nodes,
                                  double *blocks)
  {
    k_constraint_identity_blocks<<<grid_for(n * nc * nc, 256), 256, 0, s>>>(n, nc, nodes,
                                                                             blocks);
  }

  // node blocks stored column by column -> row by row
  __global__ void k_transpose_blocks(int64_t n_nodes, int nc, const double *__restrict__ in,
                                     double *__restrict__ out)
  {
    const int64_t per   = (int64_t)nc * nc;
    const int64_t total = n_nodes * per;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x)
      {
        const int64_t node = e / per;
        const int     r    = (int)(e - node * per);
        const int     a = r / nc, b = r - (r / nc) * nc;
        out[e]             = in[node * per + (int64_t)b * nc + a];
      }
  }
  void transpose_blocks(hipStream_t s, int64_t n_nodes, int nc, const double *in, double *out)
  {
    k_transpose_blocks<<<grid_for(n_nodes * nc * nc, 256), 256, 0, s>>>(n_nodes, nc, in, out);
  }
} // namespace hps
