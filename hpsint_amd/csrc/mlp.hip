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

  // tanh for the hidden layers (the result is rounded to bf16, relative
  // spacing 2^-8): the odd Taylor polynomial through x^9 for |x| < 0.5
  // (relative error below 1e-5, largest at |x| = 0.5), and
  // 1 - 2 / (exp(2|x|) + 1) with the hardware exponential elsewhere
  // (relative error below 1e-6)
  __device__ __forceinline__ float tanh_fast(float x)
  {
    const float a = fabsf(x);
    float       r;
    if (a < 0.5f)
      {
        const float x2 = x * x;
        // x - x^3/3 + 2x^5/15 - 17x^7/315 + 62x^9/2835
        r = x * (1.0f + x2 * (-0.333333333f +
                               x2 * (0.133333333f + x2 * (-0.0539682540f + x2 * 0.0218694885f))));
      }
    else
      {
        const float e = __expf(2.0f * a);
        r             = copysignf(1.0f - 2.0f / (e + 1.0f), x);
      }
    return r;
  }

  // tile: 64 rows x 128 columns (a 128-wide layer reads its input once),
  // 4 waves as 2 x 2, each 32 rows x 64 columns = two 32x32 MFMA tiles
  constexpr int GBM = 64, GBN = 128, GBK = 32, GPAD = 8;

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

    f32x16 acc[2];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int r = 0; r < 16; ++r)
        acc[t][r] = 0.f;

    // 8 consecutive bf16 from p, the first `valid` of them (the rest zero)
    auto load8 = [](const uint16_t *p, int valid) -> uint4 {
      if (valid >= 8)
        return *(const uint4 *)p;
      uint16_t e[8];
#pragma unroll
      for (int j = 0; j < 8; ++j)
        e[j] = j < valid ? p[j] : 0;
      uint4 r;
      r.x = e[0] | ((uint32_t)e[1] << 16);
      r.y = e[2] | ((uint32_t)e[3] << 16);
      r.z = e[4] | ((uint32_t)e[5] << 16);
      r.w = e[6] | ((uint32_t)e[7] << 16);
      return r;
    };
    if (!TA && TB && vec)
      {
        // row-major A and row-major B^T (the inference layout): the next
        // K-tile is fetched into registers while the current one is on MFMA
        const int am = tid >> 2, kc = (tid & 3) * 8;
        uint4     pa, pb[2];
        auto      fetch = [&](int k0) {
          const int valid = min(8, max(0, kend - (k0 + kc)));
          const int gm    = m0 + am;
          pa = gm < M ? load8(&A[(int64_t)gm * lda + k0 + kc], valid) : make_uint4(0, 0, 0, 0);
#pragma unroll
          for (int j = 0; j < 2; ++j)
            {
              const int gn = n0 + (tid >> 2) + j * 64;
              pb[j] = gn < N ? load8(&B[(int64_t)gn * ldb + k0 + kc], valid)
                             : make_uint4(0, 0, 0, 0);
            }
        };
        fetch(kbeg);
        for (int k0 = kbeg; k0 < kend; k0 += GBK)
          {
            *(uint4 *)&As[am][kc] = pa;
#pragma unroll
            for (int j = 0; j < 2; ++j)
              *(uint4 *)&Bs[(tid >> 2) + j * 64][kc] = pb[j];
            __syncthreads();
            if (k0 + GBK < kend)
              fetch(k0 + GBK);
#pragma unroll
            for (int kk = 0; kk < GBK; kk += 16)
              {
                const uint4 ra =
                  *(const uint4 *)&As[wm * 32 + (lane & 31)][kk + (lane >> 5) * 8];
                const bf16x8 a = __builtin_bit_cast(bf16x8, ra);
#pragma unroll
                for (int t = 0; t < 2; ++t)
                  {
                    const uint4 rb =
                      *(const uint4 *)&Bs[wn * 64 + t * 32 + (lane & 31)][kk + (lane >> 5) * 8];
                    const bf16x8 b = __builtin_bit_cast(bf16x8, rb);
                    acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[t], 0, 0, 0);
                  }
              }
            __syncthreads();
          }
      }
    else
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
          // ---- stage B tile (GBK x GBN) as Bs[n][k], 8 consecutive k or n per item
          if (TB)
            {
              for (int ch = tid; ch < GBN * GBK / 8; ch += 256)
                {
                  const int n = ch >> 2, kc = (ch & 3) * 8;
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
            }
          else
            {
              for (int ch = tid; ch < GBN * GBK / 8; ch += 256)
                {
                  const int k = ch / (GBN / 8), nc = (ch % (GBN / 8)) * 8;
#pragma unroll
                  for (int j = 0; j < 8; ++j)
                    {
                      const int gn = n0 + nc + j, gk = k0 + k;
                      Bs[nc + j][k] = (gn < N && gk < kend) ? B[(int64_t)gk * ldb + gn] : 0;
                    }
                }
            }
          __syncthreads();
#pragma unroll
          for (int kk = 0; kk < GBK; kk += 16)
            {
              const uint4  ra = *(const uint4 *)&As[wm * 32 + (lane & 31)][kk + (lane >> 5) * 8];
              const bf16x8 a  = __builtin_bit_cast(bf16x8, ra);
#pragma unroll
              for (int t = 0; t < 2; ++t)
                {
                  const uint4 rb =
                    *(const uint4 *)&Bs[wn * 64 + t * 32 + (lane & 31)][kk + (lane >> 5) * 8];
                  const bf16x8 b = __builtin_bit_cast(bf16x8, rb);
                  acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[t], 0, 0, 0);
                }
            }
          __syncthreads();
        }

  
    // ---- epilogue: acc[r] holds C[row][col] with
    //      row = 8*(r/4) + 4*(lane/32) + r%4, col = lane%32 (within the 32x32 tile)
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int r = 0; r < 16; ++r)
      {
        const int col = n0 + wn * 64 + t * 32 + (lane & 31);
        const int row = m0 + wm * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        if (row >= M || col >= N)
          continue;
        float v = acc[t][r];
        const int64_t o = (int64_t)row * ldc + col;
        if constexpr (EPI == EPI_F32)
          ((float *)Cout)[blockIdx.z * split_stride + o] = v;
        else if constexpr (EPI == EPI_BIAS_TANH)
          ((uint16_t *)Cout)[o] = f2bf(tanh_fast(v + bias[col]));
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
  // A workgroup of 256 lanes handles a tile of 16 x 4 nodes (x, y) of one
  // z-plane.  The tile's stencil halo (18 x 6 x 3 nodes, both time levels,
  // clamped at the domain boundary like the stencil) is loaded into LDS
  // once.  Phase 1, 4 lanes per node: top-2 order parameters and the local
  // increment scale (sum over the node's 4 lanes in a fixed order).
  // Phase 2, one lane per (node, stencil point): its 10 features into the
  // node's row in LDS; the rows then go out as contiguous 16-byte stores.
  constexpr int FEAT_TX    = 16, FEAT_TY = 4, FEAT_TN = FEAT_TX * FEAT_TY;
  constexpr int FEAT_HX    = FEAT_TX + 2, FEAT_HY = FEAT_TY + 2;
  constexpr int FEAT_HALO  = FEAT_HX * FEAT_HY * 3; // halo nodes per time level
  constexpr int FEAT_CHUNK = FEAT_STRIDE / 8;

  __device__ __forceinline__ void node_fields(const double *a, int G, int i1, int i2,
                                              double *f)
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

  __global__ void __launch_bounds__(256)
    k_features(FeatParams fp, const double *__restrict__ u_prev,
               const double *__restrict__ u_cur, uint16_t *__restrict__ X,
               int *__restrict__ top2, float *__restrict__ lscale)
  {
    extern __shared__ __align__(16) unsigned char feat_smem[];
    uint16_t *rows  = (uint16_t *)feat_smem;                  // [FEAT_TN][FEAT_STRIDE]
    double   *ls_sh = (double *)(rows + FEAT_TN * FEAT_STRIDE); // [FEAT_TN]
    int      *i_sh  = (int *)(ls_sh + FEAT_TN);               // [FEAT_TN][2]
    double   *halo  = (double *)(i_sh + 2 * FEAT_TN);
    const int nc = fp.n_comp, G = nc - 2;
    const int nsx = (fp.nn[0] + FEAT_TX - 1) / FEAT_TX;
    const int nsy = (fp.nn[1] + FEAT_TY - 1) / FEAT_TY;
    const int bx  = (int)(blockIdx.x % nsx);
    const int by  = (int)((blockIdx.x / nsx) % nsy);
    const int gz  = (int)(blockIdx.x / (nsx * nsy));
    const int x0 = bx * FEAT_TX, y0 = by * FEAT_TY;

    // halo[t][(hz * FEAT_HY + hy) * FEAT_HX + hx][c], t = 0: u_cur, 1: u_prev
    // (batches of 8 loads per lane in flight before the LDS stores)
    const int n_load = 2 * FEAT_HALO * nc;
    for (int i0 = threadIdx.x; i0 < n_load; i0 += 8 * 256)
      {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
          {
            const int i = i0 + u * 256;
            v[u]        = 0.0;
            if (i < n_load)
              {
                const int c  = i % nc;
                const int hn = (i / nc) % FEAT_HALO;
                const int t  = i / (nc * FEAT_HALO);
                const int hx = hn % FEAT_HX, hy = (hn / FEAT_HX) % FEAT_HY;
                const int hz = hn / (FEAT_HX * FEAT_HY);
                const int xx = min(max(x0 + hx - 1, 0), fp.nn[0] - 1);
                const int yy = min(max(y0 + hy - 1, 0), fp.nn[1] - 1);
                const int zz = min(max(gz + hz - 1, 0), fp.nn[2] - 1);
                const int64_t nb = ((int64_t)zz * fp.nn[1] + yy) * fp.nn[0] + xx;
                v[u]             = (t == 0 ? u_cur : u_prev)[nb * nc + c];
              }
          }
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (i0 + u * 256 < n_load)
            halo[i0 + u * 256] = v[u];
      }
    __syncthreads();
    const double *hc = halo;                  // u_cur
    const double *hp = halo + FEAT_HALO * nc; // u_prev
    // halo offset of stencil point s of tile node n
    auto at = [&](int n, int s) {
      const int dx = s % 3 - 1, dy = (s / 3) % 3 - 1, dz = s / 9 - 1;
      return (((1 + dz) * FEAT_HY + (1 + n / FEAT_TX + dy)) * FEAT_HX + 1 + n % FEAT_TX + dx) *
             nc;
    };

    // ---- phase 1: 4 lanes per node
    {
      const int  sub  = threadIdx.x % 4;
      const int  n    = threadIdx.x / 4;
      const int  gx = x0 + n % FEAT_TX, gy = y0 + n / FEAT_TX;
      const bool live = gx < fp.nn[0] && gy < fp.nn[1];
      const int  s0 = (27 * sub) / 4, s1 = (27 * (sub + 1)) / 4;
      int        i1 = 0, i2 = G > 1 ? 1 : 0;
      double     ss = 0.0;
      if (live)
        {
          const double *ctr = hc + at(n, 13);
          double        e1 = -1e300, e2 = -1e300;
          for (int i = 0; i < G; ++i)
            {
              const double e = ctr[2 + i];
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
          for (int s = s0; s < s1; ++s)
            {
              double fa[5], fb[5];
              node_fields(hc + at(n, s), G, i1, i2, fa);
              node_fields(hp + at(n, s), G, i1, i2, fb);
#pragma unroll
              for (int f = 0; f < 5; ++f)
                {
                  const double d = fa[f] - fb[f];
                  ss += d * d;
                }
            }
        }
      // fixed order: (q0 + q1) + (q2 + q3)
      ss += __shfl_xor(ss, 1, 64);
      ss += __shfl_xor(ss, 2, 64);
      if (live && sub == 0)
        {
          const double  ls   = sqrt(ss / 135.0) + fp.inc_floor;
          const int64_t node = ((int64_t)gz * fp.nn[1] + gy) * fp.nn[0] + gx;
          ls_sh[n]           = ls;
          i_sh[2 * n]        = i1;
          i_sh[2 * n + 1]    = i2;
          lscale[node]       = (float)ls;
          top2[node * 2]     = i1;
          top2[node * 2 + 1] = i2;
        }
    }
    __syncthreads();

    // ---- phase 2: one lane per (node, stencil point): the point's 10
    // features into the node's row in LDS
    for (int it = threadIdx.x; it < FEAT_TN * 27; it += 256)
      {
        const int n = it / 27, sp = it % 27;
        const int gx = x0 + n % FEAT_TX, gy = y0 + n / FEAT_TX;
        if (gx >= fp.nn[0] || gy >= fp.nn[1])
          continue;
        const double inv_ls = 1.0 / ls_sh[n];
        double       fa[5], fb[5];
        node_fields(hc + at(n, sp), G, i_sh[2 * n], i_sh[2 * n + 1], fa);
        node_fields(hp + at(n, sp), G, i_sh[2 * n], i_sh[2 * n + 1], fb);
        uint16_t *x = rows + n * FEAT_STRIDE + sp * 10;
#pragma unroll
        for (int f = 0; f < 5; ++f)
          {
            x[f]     = f2bf((float)fa[f]);
            x[5 + f] = f2bf((float)((fa[f] - fb[f]) * inv_ls));
          }
      }
    if (threadIdx.x < FEAT_TN)
      {
        const int n = threadIdx.x;
        if (x0 + n % FEAT_TX < fp.nn[0] && y0 + n / FEAT_TX < fp.nn[1])
          {
            rows[n * FEAT_STRIDE + 270] = f2bf((float)(fp.dt * fp.dt_scale));
            rows[n * FEAT_STRIDE + 271] = f2bf((float)log10(ls_sh[n] * fp.inc_scale));
          }
      }
    __syncthreads();
    // ---- write the rows: per tile row of up to 16 nodes, contiguous
    //      16-byte stores
    for (int ty = 0; ty < FEAT_TY; ++ty)
      {
        const int gy = y0 + ty;
        if (gy >= fp.nn[1])
          break;
        const int     n_rows = min(FEAT_TX, fp.nn[0] - x0);
        const int     n_vec  = n_rows * FEAT_CHUNK;
        const int64_t first  = ((int64_t)gz * fp.nn[1] + gy) * fp.nn[0] + x0;
        const uint4  *src    = (const uint4 *)(rows + ty * FEAT_TX * FEAT_STRIDE);
        uint4        *dst    = (uint4 *)(X + first * FEAT_STRIDE);
        for (int i = threadIdx.x; i < n_vec; i += 256)
          dst[i] = src[i];
      }
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
    const int    nsx    = (fp.nn[0] + FEAT_TX - 1) / FEAT_TX;
    const int    nsy    = (fp.nn[1] + FEAT_TY - 1) / FEAT_TY;
    const size_t shared = FEAT_TN * (FEAT_STRIDE * sizeof(uint16_t) + sizeof(double) +
                                     2 * sizeof(int)) +
                          2 * FEAT_HALO * fp.n_comp * sizeof(double);
    k_features<<<nsx * nsy * fp.nn[2], 256, shared, s>>>(fp, u_prev, u_cur, X, top2, lscale);
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
