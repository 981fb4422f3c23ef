// mlp.hip -- BF16-MFMA kernels of the learned Newton predictor.
//
// No counterpart in the reference: the predictor plugs into the driver's
// predictor slot (S/driver.h:2002-2050, S/ =
// applications/sintering/include/pf-applications/sintering/), where the
// reference offers Euler/Midpoint/Linear extrapolation of the initial guess.
//
// A small MLP (tanh hidden layers) maps per-node stencil features to a
// correction of the linear extrapolation 2 u_n - u_{n-1}.  Training and
// inference run on these kernels:
//   k_gemm      C = op(A) op(B) on v_mfma_f32_32x32x16_bf16 (BF16 inputs,
//               FP32 accumulation), LDS-staged 64x64 tiles, optional split-K,
//               fused epilogues (bias + tanh, bias, (1 - H^2) scaling)
//   k_features  27-node stencil features of (c, mu, sum eta^2, top-2 eta)
//   k_loss      MSE and its gradient
//   k_adam      Adam step on FP32 master weights + BF16 copies
//   k_predict   applies the corrections to the extrapolated state
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mlp.hpp"

namespace hps
{
  typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
  typedef float  f32x16 __attribute__((ext_vector_type(16)));

  __device__ __forceinline__ uint16_t f2bf(float f)
  {
    // round to nearest even
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
  }
  __device__ __forceinline__ float bf2f(uint16_t b)
  {
    return __uint_as_float(((uint32_t)b) << 16);
  }

  constexpr int GBM = 64, GBN = 64, GBK = 32, GPAD = 8; // tile and LDS padding

  // C[m][n] = sum_k A(m,k) B(k,n);  A(m,k) = TA ? A[k*lda+m] : A[m*lda+k],
  // B(k,n) = TB ? B[n*ldb+k] : B[k*ldb+n].  blockIdx.z selects a K-slice of
  // length kslice (split-K: output offset z*split_stride).
  template <bool TA, bool TB, int EPI>
  __global__ void __launch_bounds__(256)
    k_gemm(int M, int N, int K, const uint16_t *__restrict__ A, int lda,
           const uint16_t *__restrict__ B, int ldb, void *__restrict__ Cout, int ldc,
           const float *__restrict__ bias, const uint16_t *__restrict__ H, int kslice,
           int64_t split_stride)
  {
    __shared__ __align__(16) uint16_t As[GBM][GBK + GPAD];
    __shared__ __align__(16) uint16_t Bs[GBN][GBK + GPAD];
    // 16-byte row loads when every row start is 16-byte aligned
    const bool vec = ((lda | ldb) & 7) == 0 && (((uintptr_t)A | (uintptr_t)B) & 15) == 0;
    const int tid  = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int m0 = blockIdx.x * GBM, n0 = blockIdx.y * GBN;
    const int kbeg = blockIdx.z * kslice;
    const int kend = min(K, kbeg + kslice);

    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r)
      acc[r] = 0.f;

    for (int k0 = kbeg; k0 < kend; k0 += GBK)
      {
        // ---- stage A tile (GBM x GBK) as As[m][k]
        if (!TA)
          {
            const int m = tid >> 2, kc = (tid & 3) * 8;
            const int gm = m0 + m;
            if (vec && gm < M && k0 + kc + 8 <= kend)
              *(uint4 *)&As[m][kc] = *(const uint4 *)&A[(int64_t)gm * lda + k0 + kc];
            else
#pragma unroll
              for (int j = 0; j < 8; ++j)
                {
                  const int gk = k0 + kc + j;
                  As[m][kc + j] = (gm < M && gk < kend) ? A[(int64_t)gm * lda + gk] : 0;
                }
          }
        else
          {
            const int k = tid >> 3, mc = (tid & 7) * 8;
#pragma unroll
            for (int j = 0; j < 8; ++j)
              {
                const int gm = m0 + mc + j, gk = k0 + k;
                As[mc + j][k] = (gm < M && gk < kend) ? A[(int64_t)gk * lda + gm] : 0;
              }
          }
        // ---- stage B tile (GBK x GBN) as Bs[n][k]
        if (TB)
          {
            const int n = tid >> 2, kc = (tid & 3) * 8;
            const int gn = n0 + n;
            if (vec && gn < N && k0 + kc + 8 <= kend)
              *(uint4 *)&Bs[n][kc] = *(const uint4 *)&B[(int64_t)gn * ldb + k0 + kc];
            else
#pragma unroll
              for (int j = 0; j < 8; ++j)
                {
                  const int gk = k0 + kc + j;
                  Bs[n][kc + j] = (gn < N && gk < kend) ? B[(int64_t)gn * ldb + gk] : 0;
                }
          }
        else
          {
            const int k = tid >> 3, nc = (tid & 7) * 8;
#pragma unroll
            for (int j = 0; j < 8; ++j)
              {
                const int gn = n0 + nc + j, gk = k0 + k;
                Bs[nc + j][k] = (gn < N && gk < kend) ? B[(int64_t)gk * ldb + gn] : 0;
              }
          }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GBK; kk += 16)
          {
            const uint4 ra = *(const uint4 *)&As[wm * 32 + (lane & 31)][kk + (lane >> 5) * 8];
            const uint4 rb = *(const uint4 *)&Bs[wn * 32 + (lane & 31)][kk + (lane >> 5) * 8];
            const bf16x8 a = __builtin_bit_cast(bf16x8, ra);
            const bf16x8 b = __builtin_bit_cast(bf16x8, rb);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
          }
        __syncthreads();
      }

    // ---- epilogue: acc[r] holds C[row][col] with
    //      row = 8*(r/4) + 4*(lane/32) + r%4, col = lane%32 (within the 32x32 tile)
    const int col = n0 + wn * 32 + (lane & 31);
#pragma unroll
    for (int r = 0; r < 16; ++r)
      {
        const int row = m0 + wm * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        if (row >= M || col >= N)
          continue;
        float v = acc[r];
        const int64_t o = (int64_t)row * ldc + col;
        if constexpr (EPI == EPI_F32)
          ((float *)Cout)[blockIdx.z * split_stride + o] = v;
        else if constexpr (EPI == EPI_BIAS_TANH)
          ((uint16_t *)Cout)[o] = f2bf(tanhf(v + bias[col]));
        else if constexpr (EPI == EPI_BIAS_F32)
          ((float *)Cout)[o] = v + bias[col];
        else if constexpr (EPI == EPI_DTANH)
          {
            const float h = bf2f(H[o]);
            ((uint16_t *)Cout)[o] = f2bf(v * (1.f - h * h));
          }
      }
  }

  // sum of split-K partials: out[i] = sum_z part[z*n + i]
  __global__ void k_sum_splits(int64_t n, int nsplit, const float *__restrict__ part,
                               float *__restrict__ out)
  {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
      {
        float s = 0.f;
        for (int z = 0; z < nsplit; ++z)
          s += part[(int64_t)z * n + i];
        out[i] = s;
      }
  }

  // column sums of a bf16 matrix (bias gradients): out[n] = sum_m X[m][n]
  __global__ void __launch_bounds__(256)
    k_colsum(int M, int N, const uint16_t *__restrict__ X, float *__restrict__ out)
  {
    const int n = blockIdx.x;
    __shared__ float sh[256];
    float s = 0.f;
    for (int m = threadIdx.x; m < M; m += 256)
      s += bf2f(X[(int64_t)m * N + n]);
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1)
      {
        if (threadIdx.x < o)
          sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
      }
    if (threadIdx.x == 0)
      out[n] = sh[0];
  }

  // MSE loss L = mean((Y - T)^2) over M x N; dY = 2 (Y - T) / (M N) as bf16;
  // per-block partial losses
  __global__ void __launch_bounds__(256)
    k_loss(int64_t n, const float *__restrict__ Y, const float *__restrict__ T,
           uint16_t *__restrict__ dY, float *__restrict__ partial)
  {
    __shared__ float sh[256];
    float s = 0.f;
    const float inv = 1.f / (float)n;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
      {
        const float d = Y[i] - T[i];
        s += d * d;
        dY[i] = f2bf(2.f * d * inv);
      }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1)
      {
        if (threadIdx.x < o)
          sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
      }
    if (threadIdx.x == 0)
      partial[blockIdx.x] = sh[0] * inv;
  }

  // Adam on FP32 master weights W (out x in), refreshes the bf16 copies W
  // (out x in) and W^T (in x out)
  __global__ void k_adam(int n_out, int n_in, float *__restrict__ W, float *__restrict__ m1,
                         float *__restrict__ m2, const float *__restrict__ g, float lr,
                         float b1, float b2, float eps, float c1, float c2,
                         uint16_t *__restrict__ Wb, uint16_t *__restrict__ WTb)
  {
    const int n = n_out * n_in;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
      {
        const float gi = g[i];
        const float a  = b1 * m1[i] + (1.f - b1) * gi;
        const float b  = b2 * m2[i] + (1.f - b2) * gi * gi;
        m1[i]          = a;
        m2[i]          = b;
        const float w  = W[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
        W[i]           = w;
        if (Wb)
          Wb[i] = f2bf(w);
        if (WTb)
          {
            const int o = i / n_in, k = i - o * n_in;
            WTb[(int64_t)k * n_out + o] = f2bf(w);
          }
      }
  }

  // stencil features of every node (3D: 27 neighbours, clamped at the
  // boundary; 2D: 9, the z offsets repeat the plane).  For the fields
  // f = c, mu, sum eta^2, eta_{i1}, eta_{i2} of neighbour s:
  //   x[s*10 + f]     = f(t_n)
  //   x[s*10 + 5 + f] = (f(t_n) - f(t_{n-1})) / ls
  // with (i1, i2) the two largest order parameters at the node at t_n and
  // ls the node's local increment scale, the RMS of the 135 increments plus
  // a floor; x[270] = dt * dt_scale, x[271] = log10(ls * inc_scale) (the
  // local scale relative to the global one, 1 / inc_scale).
  // One wave per 16 consecutive nodes, 4 lanes per node: each lane handles
  // a quarter of the stencil (the increment scale is summed over the 4
  // lanes), the lanes build the node's feature row in LDS, then the wave
  // writes the 16 rows as contiguous 16-byte stores.
  constexpr int FEAT_LANES = 4;
  constexpr int FEAT_NODES = 64 / FEAT_LANES;

  __device__ __forceinline__ void stencil_node(const FeatParams &fp, int gx, int gy, int gz,
                                               int dx, int dy, int dz, int64_t &nb)
  {
    const int xx = min(max(gx + dx, 0), fp.nn[0] - 1);
    const int yy = min(max(gy + dy, 0), fp.nn[1] - 1);
    const int zz = min(max(gz + dz, 0), fp.nn[2] - 1);
    nb           = ((int64_t)zz * fp.nn[1] + yy) * fp.nn[0] + xx;
  }

  __device__ __forceinline__ void node_fields(const double *__restrict__ a, int G, int i1,
                                              int i2, double *f)
  {
    double sq = 0.0;
    for (int i = 0; i < G; ++i)
      sq += a[2 + i] * a[2 + i];
    f[0] = a[0];
    f[1] = a[1];
    f[2] = sq;
    f[3] = a[2 + i1];
    f[4] = a[2 + i2];
  }

  __global__ void __launch_bounds__(64)
    k_features(FeatParams fp, const double *__restrict__ u_prev,
               const double *__restrict__ u_cur, uint16_t *__restrict__ X,
               int *__restrict__ top2, float *__restrict__ lscale)
  {
    __shared__ __align__(16) uint16_t rows[FEAT_NODES * FEAT_STRIDE];
    const int     sub  = threadIdx.x % FEAT_LANES; // quarter of the stencil
    const int     slot = threadIdx.x / FEAT_LANES;
    const int64_t base = (int64_t)blockIdx.x * FEAT_NODES;
    const int64_t node = base + slot;
    const bool    live = node < fp.n_nodes;
    const int     s0 = (27 * sub) / FEAT_LANES, s1 = (27 * (sub + 1)) / FEAT_LANES;
    const int     nc = fp.n_comp, G = nc - 2;
    int           gx = 0, gy = 0, gz = 0, i1 = 0, i2 = 0;
    double        ss = 0.0;
    if (live)
      {
        gx = (int)(node % fp.nn[0]);
        gy = (int)((node / fp.nn[0]) % fp.nn[1]);
        gz = (int)(node / ((int64_t)fp.nn[0] * fp.nn[1]));
        // top-2 order parameters at the node at t_n
        i2        = G > 1 ? 1 : 0;
        double e1 = -1e300, e2 = -1e300;
        for (int i = 0; i < G; ++i)
          {
            const double e = u_cur[node * nc + 2 + i];
            if (e > e1)
              {
                e2 = e1;
                i2 = i1;
                e1 = e;
                i1 = i;
              }
            else if (e > e2)
              {
                e2 = e;
                i2 = i;
              }
          }
        if (G == 1)
          i2 = 0;
        // pass 1: this lane's part of the sum of squares of the increments
        for (int s = s0; s < s1; ++s)
          {
            int64_t nb;
            stencil_node(fp, gx, gy, gz, s % 3 - 1, (s / 3) % 3 - 1, s / 9 - 1, nb);
            double fa[5], fb[5];
            node_fields(u_cur + nb * nc, G, i1, i2, fa);
            node_fields(u_prev + nb * nc, G, i1, i2, fb);
#pragma unroll
            for (int f = 0; f < 5; ++f)
              {
                const double d = fa[f] - fb[f];
                ss += d * d;
              }
          }
      }
    // sum over the node's lanes in a fixed order: (q0 + q1) + (q2 + q3)
    ss += __shfl_xor(ss, 1, 64);
    ss += __shfl_xor(ss, 2, 64);
    if (live)
      {
        const double ls = sqrt(ss / 135.0) + fp.inc_floor;
        // pass 2: values and scaled increments of this lane's stencil part
        uint16_t *x = rows + slot * FEAT_STRIDE;
        for (int s = s0; s < s1; ++s)
          {
            int64_t nb;
            stencil_node(fp, gx, gy, gz, s % 3 - 1, (s / 3) % 3 - 1, s / 9 - 1, nb);
            double fa[5], fb[5];
            node_fields(u_cur + nb * nc, G, i1, i2, fa);
            node_fields(u_prev + nb * nc, G, i1, i2, fb);
#pragma unroll
            for (int f = 0; f < 5; ++f)
              {
                x[s * 10 + f]     = f2bf((float)fa[f]);
                x[s * 10 + 5 + f] = f2bf((float)((fa[f] - fb[f]) / ls));
              }
          }
        if (sub == 0)
          {
            x[270]             = f2bf((float)(fp.dt * fp.dt_scale));
            x[271]             = f2bf((float)log10(ls * fp.inc_scale));
            lscale[node]       = (float)ls;
            top2[node * 2]     = i1;
            top2[node * 2 + 1] = i2;
          }
      }
    __syncthreads();
    // write the block's rows: contiguous, 16 bytes per lane and store
    const int64_t n_rows = min((int64_t)FEAT_NODES, fp.n_nodes - base);
    const int     n_vec  = (int)(n_rows * FEAT_STRIDE / 8);
    const uint4  *src    = (const uint4 *)rows;
    uint4        *dst    = (uint4 *)(X + base * FEAT_STRIDE);
    for (int i = threadIdx.x; i < n_vec; i += 64)
      dst[i] = src[i];
  }

  // u_pred = 2 u_n - u_{n-1} + ls * out_scale[k] * y[k] on (c, mu, eta_i1, eta_i2)
  __global__ void k_predict(int64_t n_nodes, int nc, const double *__restrict__ u_prev,
                            const double *__restrict__ u_cur, const float *__restrict__ Y,
                            const int *__restrict__ top2, const float *__restrict__ out_scale,
                            const float *__restrict__ lscale, double *__restrict__ u_pred)
  {
    const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= n_nodes)
      return;
    for (int c = 0; c < nc; ++c)
      u_pred[node * nc + c] = 2.0 * u_cur[node * nc + c] - u_prev[node * nc + c];
    if (!Y)
      return;
    const int comp[4] = {0, 1, 2 + top2[node * 2], 2 + top2[node * 2 + 1]};
    for (int k = 0; k < 4; ++k)
      {
        if (k == 3 && comp[3] == comp[2])
          break;
        u_pred[node * nc + comp[k]] += (double)lscale[node] * (double)(out_scale[k] * Y[node * 4 + k]);
      }
  }

  // per-row training targets: (u_{n+1} - (2 u_n - u_{n-1}))[comp_k] / (ls out_scale[k])
  __global__ void k_targets(int64_t n_nodes, int nc, const double *__restrict__ u_prev,
                            const double *__restrict__ u_cur, const double *__restrict__ u_next,
                            const int *__restrict__ top2, const float *__restrict__ out_scale,
                            const float *__restrict__ lscale, float *__restrict__ T)
  {
    const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= n_nodes)
      return;
    const int comp[4] = {0, 1, 2 + top2[node * 2], 2 + top2[node * 2 + 1]};
    for (int k = 0; k < 4; ++k)
      {
        const int64_t i = node * nc + comp[k];
        const double  r = u_next[i] - (2.0 * u_cur[i] - u_prev[i]);
        T[node * 4 + k] = (k == 3 && comp[3] == comp[2])
                            ? 0.f
                            : (float)(r / ((double)lscale[node] * (double)out_scale[k]));
      }
  }

  // ------------------------------------------------------------ launchers
  template <bool TA, bool TB, int EPI>
  static void gemm(hipStream_t s, int M, int N, int K, const uint16_t *A, int lda,
                   const uint16_t *B, int ldb, void *C, int ldc, const float *bias,
                   const uint16_t *H, int nsplit, int64_t split_stride)
  {
    const int kslice = ((K + nsplit - 1) / nsplit + GBK - 1) / GBK * GBK;
    dim3 grid((M + GBM - 1) / GBM, (N + GBN - 1) / GBN, nsplit);
    k_gemm<TA, TB, EPI><<<grid, 256, 0, s>>>(M, N, K, A, lda, B, ldb, C, ldc, bias, H, kslice,
                                             split_stride);
  }

  void mlp_gemm(hipStream_t s, int ta, int tb, int epi, int M, int N, int K, const uint16_t *A,
                int lda, const uint16_t *B, int ldb, void *C, int ldc, const float *bias,
                const uint16_t *H, int nsplit, int64_t split_stride)
  {
#define HPS_G(TA, TB, E)                                                                   \
  if (ta == TA && tb == TB && epi == E)                                                    \
    return gemm<TA, TB, E>(s, M, N, K, A, lda, B, ldb, C, ldc, bias, H, nsplit, split_stride);
    HPS_G(0, 1, EPI_BIAS_TANH)
    HPS_G(0, 1, EPI_BIAS_F32)
    HPS_G(0, 1, EPI_F32)
    HPS_G(0, 0, EPI_DTANH)
    HPS_G(0, 0, EPI_F32)
    HPS_G(1, 0, EPI_F32)
    HPS_G(1, 1, EPI_F32)
#undef HPS_G
  }

  void mlp_sum_splits(hipStream_t s, int64_t n, int nsplit, const float *part, float *out)
  {
    k_sum_splits<<<(int)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096), 256, 0, s>>>(
      n, nsplit, part, out);
  }
  void mlp_colsum(hipStream_t s, int M, int N, const uint16_t *X, float *out)
  {
    k_colsum<<<N, 256, 0, s>>>(M, N, X, out);
  }
  void mlp_loss(hipStream_t s, int64_t n, const float *Y, const float *T, uint16_t *dY,
                float *partial, int nblocks)
  {
    k_loss<<<nblocks, 256, 0, s>>>(n, Y, T, dY, partial);
  }
  void mlp_adam(hipStream_t s, int n_out, int n_in, float *W, float *m1, float *m2,
                const float *g, float lr, float b1, float b2, float eps, float c1, float c2,
                uint16_t *Wb, uint16_t *WTb)
  {
    const int n = n_out * n_in;
    k_adam<<<(n + 255) / 256, 256, 0, s>>>(n_out, n_in, W, m1, m2, g, lr, b1, b2, eps, c1, c2,
                                           Wb, WTb);
  }
  void mlp_features(hipStream_t s, const FeatParams &fp, const double *u_prev,
                    const double *u_cur, uint16_t *X, int *top2, float *lscale)
  {
    k_features<<<(int)((fp.n_nodes + FEAT_NODES - 1) / FEAT_NODES), 64, 0, s>>>(
      fp, u_prev, u_cur, X, top2, lscale);
  }
  void mlp_predict(hipStream_t s, int64_t n_nodes, int nc, const double *u_prev,
                   const double *u_cur, const float *Y, const int *top2,
                   const float *out_scale, const float *lscale, double *u_pred)
  {
    k_predict<<<(int)((n_nodes + 255) / 256), 256, 0, s>>>(n_nodes, nc, u_prev, u_cur, Y, top2,
                                                          out_scale, lscale, u_pred);
  }
  void mlp_targets(hipStream_t s, int64_t n_nodes, int nc, const double *u_prev,
                   const double *u_cur, const double *u_next, const int *top2,
                   const float *out_scale, const float *lscale, float *T)
  {
    k_targets<<<(int)((n_nodes + 255) / 256), 256, 0, s>>>(n_nodes, nc, u_prev, u_cur, u_next,
                                                          top2, out_scale, lscale, T);
  }
} // namespace hps
