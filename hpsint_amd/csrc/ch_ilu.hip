// ch_ilu.hip -- the block preconditioner of the reference's default setup
// on the GPU: ILU(0) of the Cahn-Hilliard block and the inverse diagonal of
// the Allen-Cahn block (BlockPreconditioner2, S/preconditioners.h:983-1282,
// with BlockPreconditioner2Data's defaults "ILU" / "InverseDiagonalMatrix",
// S/preconditioner_parameters.h:57-64; S/ = applications/sintering/include/
// pf-applications/sintering/).
//
// The Cahn-Hilliard block is the tangent restricted to the rows and columns
// of c and mu.  Q1 couples a node with its 3^DIM neighbours, so a node's row
// is 3^DIM stencil entries of 2x2 (row c/mu, column c/mu), s = (dx+1) +
// 3 (dy+1) [+ 9 (dz+1)]; the rows are stored colour by colour and entry by
// entry (ChParams::off, cnt), so that the threads of one colour stage read
// consecutive addresses.
//
// Ordering: nodes in 2^DIM colours by the parity of their coordinates
// (colour = x&1 | (y&1)<<1 | (z&1)<<2), nodes of one colour in index order,
// c before mu within a node.  No two neighbours share a colour, so every
// stage of the factorisation and of the triangular solves (one colour) is a
// parallel loop over the colour's nodes: 2^DIM launches per factorisation
// and per solve.  ILU(0) keeps the pattern of the block; the rows are
// eliminated in the IKJ order (lower entries in increasing column order).
//
// The assembly forms the Cahn-Hilliard part of every cell's element matrix
// from the linearisation cache (scalar mobility):
//   (c, c):   JxW (w0 phi_k phi_m + dM/dc phi_m (grad mu . grad phi_k))
//   (c, mu):  JxW M (grad phi_m . grad phi_k)
//   (mu, c):  JxW (-d2f/dc2 phi_m phi_k - kappa_c grad phi_m . grad phi_k)
//   (mu, mu): JxW w_alg phi_m phi_k
// (S/operator_sintering_generic.h:104-190 with MobilityScalar,
// S/mobility.h:441-534), in 2^DIM cell-parity passes so that no two threads
// add to one node's row.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "ch_ilu.hpp"
#include "physics.hpp"

namespace hps
{
  namespace
  {
    template <int DIM>
    struct Sten
    {
      static constexpr int S = DIM == 3 ? 27 : 9;
      __device__ __forceinline__ static int index(int dx, int dy, int dz)
      {
        return DIM == 3 ? (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1) : (dx + 1) + 3 * (dy + 1);
      }
    };

    __device__ __forceinline__ void sten_offset(int DIMv, int s, int &dx, int &dy, int &dz)
    {
      dx = s % 3 - 1;
      dy = (s / 3) % 3 - 1;
      dz = DIMv == 3 ? s / 9 - 1 : 0;
    }

    // node of colour `col` number i (x fastest among the colour's nodes)
    __device__ __forceinline__ bool colour_node(const ChParams &cp, int col, int64_t i, int &x,
                                                int &y, int &z)
    {
      const int px = col & 1, py = (col >> 1) & 1, pz = (col >> 2) & 1;
      const int mx = (cp.nn[0] - px + 1) / 2, my = (cp.nn[1] - py + 1) / 2;
      const int mz = cp.dim == 3 ? (cp.nn[2] - pz + 1) / 2 : 1;
      if (i >= (int64_t)mx * my * mz)
        return false;
      x = px + 2 * (int)(i % mx);
      y = py + 2 * (int)((i / mx) % my);
      z = cp.dim == 3 ? pz + 2 * (int)(i / ((int64_t)mx * my)) : 0;
      return true;
    }

    __device__ __forceinline__ bool inside(const ChParams &cp, int x, int y, int z)
    {
      return x >= 0 && x < cp.nn[0] && y >= 0 && y < cp.nn[1] && z >= 0 &&
             z < (cp.dim == 3 ? cp.nn[2] : 1);
    }

    __device__ __forceinline__ int64_t node_id(const ChParams &cp, int x, int y, int z)
    {
      return ((int64_t)z * cp.nn[1] + y) * cp.nn[0] + x;
    }

    // the stored row of a node: entry k = 4 s + 2 a + b at base[k * stride]
    template <typename T>
    struct RowRef
    {
      T      *base;
      int64_t stride;
      __device__ __forceinline__ T &operator[](int k) const { return base[k * stride]; }
    };

    template <typename T>
    __device__ __forceinline__ RowRef<T> row_of(const ChParams &cp, T *F, int x, int y, int z)
    {
      const int     col = (x & 1) | ((y & 1) << 1) | ((z & 1) << 2);
      const int64_t mx = (cp.nn[0] - (col & 1) + 1) / 2, my = (cp.nn[1] - ((col >> 1) & 1) + 1) / 2;
      const int64_t ic = (x >> 1) + mx * ((y >> 1) + my * (z >> 1));
      return RowRef<T>{F + cp.off[col] + ic, cp.cnt[col]};
    }
  } // namespace

  // ---------------------------------------------------------------- assembly
  // one thread per cell of parity colour `par`; the cell's 2^DIM x 2^DIM
  // element matrix (2x2 per node pair) is added to the rows of its nodes
  template <int DIM>
  __global__ void __launch_bounds__(256)
    k_ch_assemble(const ChParams cp, int par, const double *__restrict__ cache,
                  double *__restrict__ A)
  {
    constexpr int NV = 1 << DIM;
    constexpr int S  = Sten<DIM>::S;
    const int     px = par & 1, py = (par >> 1) & 1, pz = (par >> 2) & 1;
    const int     mx = (cp.nc[0] - px + 1) / 2, my = (cp.nc[1] - py + 1) / 2;
    const int     mz = DIM == 3 ? (cp.nc[2] - pz + 1) / 2 : 1;
    const int64_t i  = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)mx * my * mz)
      return;
    const int cx = px + 2 * (int)(i % mx);
    const int cy = py + 2 * (int)((i / mx) % my);
    const int cz = DIM == 3 ? pz + 2 * (int)(i / ((int64_t)mx * my)) : 0;

    // the cell's place in the cache (lane order of the Q1 kernels: tiles of
    // tx x ty cells, tz layers along the last direction)
    const int cm   = DIM == 3 ? cz : cy;  // march direction
    const int cyy  = DIM == 3 ? cy : 0;
    const int64_t tile =
      (int64_t)(cx / cp.tx) + (int64_t)cp.nt[0] * ((cyy / cp.ty) + (int64_t)cp.nt[1] * (cm / cp.tz));
    const int     layer = cm % cp.tz;
    const int     lane0 = (((cyy % cp.ty) * cp.tx) + (cx % cp.tx)) << DIM;
    const double *cl    = cache + ((tile * cp.tz + layer) * cp.nf) * (int64_t)cp.nt_lanes;

    const double A_ = cp.pp.A, B_ = cp.pp.B;
    // quadrature-point coefficients of the Cahn-Hilliard block
    double Mq[NV], dMq[NV], d2q[NV], lmu[NV][3];
#pragma unroll
    for (int q = 0; q < NV; ++q)
      {
        const double c       = cl[lane0 + q];
        double       eta_sum = 0.0, eta_ij = 0.0;
        for (int g = 0; g < cp.G; ++g)
          {
            const double e = cl[(1 + g) * (int64_t)cp.nt_lanes + lane0 + q];
            eta_ij += e * eta_sum;
            eta_sum += e;
          }
#pragma unroll
        for (int d = 0; d < 3; ++d)
          lmu[q][d] = d < DIM ? cl[(1 + cp.G + d) * (int64_t)cp.nt_lanes + lane0 + q] : 0.0;
        const double phi    = dmin(1.0, dmax(0.0, c * c * c * (10.0 - 15.0 * c + 6.0 * c * c)));
        Mq[q]               = cp.pp.Mvol * phi + cp.pp.Mvap * (1.0 - phi) +
                cp.pp.Msurf * 4.0 * c * c * (1.0 - c) * (1.0 - c) + (2.0 * cp.pp.Mgb) * eta_ij;
        const double dphidc = 30.0 * c * c * (1.0 - c) * (1.0 - c);
        dMq[q]              = cp.pp.Mvol * dphidc - cp.pp.Mvap * dphidc +
                 cp.pp.Msurf * 8.0 * c * (1.0 - 3.0 * c + 2.0 * c * c);
        d2q[q] = 2.0 * A_ * (c * c) + 4.0 * A_ * c * (2.0 * c - 2.0) +
                 2.0 * A_ * ((1.0 - c) * (1.0 - c)) + 2.0 * B_;
      }
    // shape value of corner k at Gauss point q along d: sa on the corner's
    // side, sb on the other; derivative +-1/h
    constexpr double sa = 0.78867513459481288225457439025098;
    constexpr double sb = 0.21132486540518711774542560974902;
    const double     jxw = cp.jxw;
#pragma unroll 1
    for (int k = 0; k < NV; ++k)
      {
        const int     kx = cx + (k & 1), ky = cy + ((k >> 1) & 1);
        const int     kz = DIM == 3 ? cz + ((k >> 2) & 1) : 0;
        const RowRef<double> ra = row_of(cp, A, kx, ky, kz);
#pragma unroll
        for (int m = 0; m < NV; ++m)
          {
            double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
#pragma unroll
            for (int q = 0; q < NV; ++q)
              {
                double vk = 1.0, vm = 1.0, gk[3], gm[3];
#pragma unroll
                for (int d = 0; d < DIM; ++d)
                  {
                    vk *= (((k >> d) & 1) == ((q >> d) & 1)) ? sa : sb;
                    vm *= (((m >> d) & 1) == ((q >> d) & 1)) ? sa : sb;
                  }
#pragma unroll
                for (int d = 0; d < DIM; ++d)
                  {
                    double fk = ((k >> d) & 1) ? 1.0 : -1.0, fm = ((m >> d) & 1) ? 1.0 : -1.0;
#pragma unroll
                    for (int e = 0; e < DIM; ++e)
                      if (e != d)
                        {
                          fk *= (((k >> e) & 1) == ((q >> e) & 1)) ? sa : sb;
                          fm *= (((m >> e) & 1) == ((q >> e) & 1)) ? sa : sb;
                        }
                    gk[d] = fk * cp.inv_h[d];
                    gm[d] = fm * cp.inv_h[d];
                  }
                double gg = 0.0, mug = 0.0;
#pragma unroll
                for (int d = 0; d < DIM; ++d)
                  {
                    gg += gm[d] * gk[d];
                    mug += lmu[q][d] * gk[d];
                  }
                const double vv = vm * vk;
                a0 += cp.pp.w0 * vv + dMq[q] * vm * mug;
                a1 += Mq[q] * gg;
                a2 += -d2q[q] * vv - cp.pp.kappa_c * gg;
                a3 += cp.pp.w_alg * vv;
              }
            const int dx = (m & 1) - (k & 1), dy = ((m >> 1) & 1) - ((k >> 1) & 1);
            const int dz = DIM == 3 ? ((m >> 2) & 1) - ((k >> 2) & 1) : 0;
            const int s4 = 4 * Sten<DIM>::index(dx, dy, dz);
            ra[s4 + 0] += jxw * a0;
            ra[s4 + 1] += jxw * a1;
            ra[s4 + 2] += jxw * a2;
            ra[s4 + 3] += jxw * a3;
          }
      }
  }

  // ------------------------------------------------------------ factorisation
  // rows (i, c) then (i, mu) of the nodes of colour `col`, IKJ order: the
  // lower entries (j, b) of a row in increasing column order -- the
  // neighbours j of lower colours by (colour, index), c before mu, and for the
  // mu row finally (i, c) -- each eliminated with the already final row
  // (j, b): l = a / u_(j,b),(j,b); a_(i,a),(k,e) -= l u_(j,b),(k,e) for the
  // (k, e) after (j, b) in row (j, b) that are in the pattern of row (i, a)
  template <int DIM>
  __global__ void __launch_bounds__(256)
    k_ch_factor(const ChParams cp, const ChOrder ord, int col, double *__restrict__ A)
  {
    constexpr int S = Sten<DIM>::S;
    int           x, y, z;
    if (!colour_node(cp, col, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, x, y, z))
      return;
    const RowRef<double> Ai     = row_of(cp, A, x, y, z);
    const int            center = (S - 1) / 2;
    for (int a = 0; a < 2; ++a)
      {
        for (int t = 0; t < ord.n_lower[col]; ++t)
          {
            const int s = ord.order[col][t];
            int       dx, dy, dz;
            sten_offset(DIM, s, dx, dy, dz);
            if (!inside(cp, x + dx, y + dy, z + dz))
              continue;
            const RowRef<double> Aj = row_of(cp, A, x + dx, y + dy, z + dz);
            const int            jc = col ^ ((dx & 1) | ((dy & 1) << 1) | ((dz & 1) << 2));
            for (int b = 0; b < 2; ++b)
              {
                const double l = Ai[s * 4 + 2 * a + b] / Aj[center * 4 + 3 * b];
                Ai[s * 4 + 2 * a + b] = l;
                // row (j, b) after (j, b): (j, mu) when b = c, then the
                // neighbours of j of higher colour than jc; those that are
                // in row (i, a)'s pattern: i itself and i's neighbours
                if (b == 0)
                  Ai[s * 4 + 2 * a + 1] -= l * Aj[center * 4 + 1];
                for (int u = 0; u < S; ++u)
                  {
                    int ex, ey, ez;
                    sten_offset(DIM, u, ex, ey, ez);
                    if (u == center)
                      continue;
                    const int kc = jc ^ ((ex & 1) | ((ey & 1) << 1) | ((ez & 1) << 2));
                    if (kc <= jc)
                      continue;
                    // k = j + e relative to i: (dx + ex, ...)
                    const int fx = dx + ex, fy = dy + ey, fz = dz + ez;
                    if (fx < -1 || fx > 1 || fy < -1 || fy > 1 || fz < -1 || fz > 1)
                      continue;
                    if (!inside(cp, x + fx, y + fy, z + fz))
                      continue;
                    const int v = DIM == 3 ? (fx + 1) + 3 * (fy + 1) + 9 * (fz + 1)
                                           : (fx + 1) + 3 * (fy + 1);
                    Ai[v * 4 + 2 * a + 0] -= l * Aj[u * 4 + 2 * b + 0];
                    Ai[v * 4 + 2 * a + 1] -= l * Aj[u * 4 + 2 * b + 1];
                  }
              }
          }
        if (a == 1)
          {
            // the lower entry (i, c) of row (i, mu), with the final row (i, c)
            const double l = Ai[center * 4 + 2] / Ai[center * 4 + 0];
            Ai[center * 4 + 2] = l;
            Ai[center * 4 + 3] -= l * Ai[center * 4 + 1];
            for (int u = 0; u < S; ++u)
              {
                int ex, ey, ez;
                sten_offset(DIM, u, ex, ey, ez);
                if (u == center || !inside(cp, x + ex, y + ey, z + ez))
                  continue;
                const int kc = col ^ ((ex & 1) | ((ey & 1) << 1) | ((ez & 1) << 2));
                if (kc <= col)
                  continue;
                Ai[u * 4 + 2] -= l * Ai[u * 4 + 0];
                Ai[u * 4 + 3] -= l * Ai[u * 4 + 1];
              }
          }
      }
  }

  template <int DIM>
  __global__ void k_ch_to_f32(int64_t n, const double *__restrict__ a, float *__restrict__ b)
  {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
      b[i] = (float)a[i];
  }

  // ------------------------------------------------------------------ solves
  // forward: y_(i,a) = r_(i,a) - sum_lower L y; the (c, mu) components of
  // the interleaved vectors (n_comp per node)
  template <int DIM, typename T>
  __global__ void __launch_bounds__(256)
    k_ch_forward(const ChParams cp, int col, const T *__restrict__ F, const T *__restrict__ r,
                 T *__restrict__ y)
  {
    constexpr int S = Sten<DIM>::S;
    int           x, y0, z;
    if (!colour_node(cp, col, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, x, y0, z))
      return;
    const int64_t   i  = node_id(cp, x, y0, z);
    const RowRef<const T> Fi = row_of(cp, F, x, y0, z);
    const int       nc = cp.n_comp;
    T             s0 = r[i * nc + 0], s1 = r[i * nc + 1];
    for (int s = 0; s < S; ++s)
      {
        int dx, dy, dz;
        sten_offset(DIM, s, dx, dy, dz);
        const int jc = col ^ ((dx & 1) | ((dy & 1) << 1) | ((dz & 1) << 2));
        if (jc >= col || !inside(cp, x + dx, y0 + dy, z + dz))
          continue;
        const int64_t j  = node_id(cp, x + dx, y0 + dy, z + dz);
        const T       yc = y[j * nc + 0], ym = y[j * nc + 1];
        s0 -= Fi[s * 4 + 0] * yc + Fi[s * 4 + 1] * ym;
        s1 -= Fi[s * 4 + 2] * yc + Fi[s * 4 + 3] * ym;
      }
    const int center = (S - 1) / 2;
    s1 -= Fi[center * 4 + 2] * s0;
    y[i * nc + 0] = s0;
    y[i * nc + 1] = s1;
  }

  // backward: x_(i,a) = (y_(i,a) - sum_upper U x) / u_(i,a),(i,a), mu first;
  // the Allen-Cahn components of the node get the inverse diagonal (in the
  // colour pass of their node)
  template <int DIM, typename T>
  __global__ void __launch_bounds__(256)
    k_ch_backward(const ChParams cp, int col, const T *__restrict__ F, const T *__restrict__ inv_diag,
                  const T *__restrict__ src, T *__restrict__ y)
  {
    constexpr int S = Sten<DIM>::S;
    int           x, y0, z;
    if (!colour_node(cp, col, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, x, y0, z))
      return;
    const int64_t   i  = node_id(cp, x, y0, z);
    const RowRef<const T> Fi = row_of(cp, F, x, y0, z);
    const int       nc = cp.n_comp;
    T             s0 = y[i * nc + 0], s1 = y[i * nc + 1];
    for (int s = 0; s < S; ++s)
      {
        int dx, dy, dz;
        sten_offset(DIM, s, dx, dy, dz);
        const int jc = col ^ ((dx & 1) | ((dy & 1) << 1) | ((dz & 1) << 2));
        if (jc <= col || !inside(cp, x + dx, y0 + dy, z + dz))
          continue;
        const int64_t j  = node_id(cp, x + dx, y0 + dy, z + dz);
        const T       xc = y[j * nc + 0], xm = y[j * nc + 1];
        s0 -= Fi[s * 4 + 0] * xc + Fi[s * 4 + 1] * xm;
        s1 -= Fi[s * 4 + 2] * xc + Fi[s * 4 + 3] * xm;
      }
    const int center = (S - 1) / 2;
    const T   xm     = s1 / Fi[center * 4 + 3];
    const T   xc     = (s0 - Fi[center * 4 + 1] * xm) / Fi[center * 4 + 0];
    y[i * nc + 0]    = xc;
    y[i * nc + 1]    = xm;
    for (int c = 2; c < nc; ++c)
      y[i * nc + c] = inv_diag[i * nc + c] * src[i * nc + c];
  }

  // ---------------------------------------------------------------- launchers
  static unsigned blocks_for(int64_t n)
  {
    return (unsigned)((n + 255) / 256);
  }

  static int64_t colour_count(const ChParams &cp, int col)
  {
    const int px = col & 1, py = (col >> 1) & 1, pz = (col >> 2) & 1;
    const int64_t mx = (cp.nn[0] - px + 1) / 2, my = (cp.nn[1] - py + 1) / 2;
    const int64_t mz = cp.dim == 3 ? (cp.nn[2] - pz + 1) / 2 : 1;
    return mx * my * mz;
  }

  void ch_layout(ChParams &cp)
  {
    const int S = cp.dim == 3 ? 27 : 9;
    cp.off[0]   = 0;
    for (int col = 0; col < 8; ++col)
      {
        cp.cnt[col]     = col < (1 << cp.dim) ? colour_count(cp, col) : 0;
        cp.off[col + 1] = cp.off[col] + cp.cnt[col] * S * 4;
      }
  }

  ChOrder ch_order(int dim)
  {
    ChOrder o{};
    const int S = dim == 3 ? 27 : 9, center = (S - 1) / 2, ncol = 1 << dim;
    for (int col = 0; col < ncol; ++col)
      {
        // lower neighbours (colour < col) by (colour, linear offset): the
        // linear offset orders as (dz, dy, dx) lexicographically, which is
        // the stencil index order
        int n = 0;
        for (int jc = 0; jc < col; ++jc)
          for (int s = 0; s < S; ++s)
            {
              if (s == center)
                continue;
              const int dx = s % 3 - 1, dy = (s / 3) % 3 - 1, dz = dim == 3 ? s / 9 - 1 : 0;
              const int m  = (dx & 1) | ((dy & 1) << 1) | ((dz & 1) << 2);
              if ((col ^ m) == jc)
                o.order[col][n++] = (signed char)s;
            }
        o.n_lower[col] = n;
      }
    return o;
  }

  void ch_assemble(const ChParams &cp, const double *cache, double *A, hipStream_t s)
  {
    const int64_t nA = cp.n_nodes * (cp.dim == 3 ? 27 : 9) * 4;
    (void)hipMemsetAsync(A, 0, nA * sizeof(double), s);
    for (int par = 0; par < (1 << cp.dim); ++par)
      {
        const int px = par & 1, py = (par >> 1) & 1, pz = (par >> 2) & 1;
        const int64_t mx = (cp.nc[0] - px + 1) / 2, my = (cp.nc[1] - py + 1) / 2;
        const int64_t mz = cp.dim == 3 ? (cp.nc[2] - pz + 1) / 2 : 1;
        const int64_t n  = mx * my * mz;
        if (n == 0)
          continue;
        if (cp.dim == 3)
          k_ch_assemble<3><<<blocks_for(n), 256, 0, s>>>(cp, par, cache, A);
        else
          k_ch_assemble<2><<<blocks_for(n), 256, 0, s>>>(cp, par, cache, A);
      }
  }

  void ch_factor(const ChParams &cp, double *A, float *A32, hipStream_t s)
  {
    const ChOrder ord = ch_order(cp.dim);
    for (int col = 0; col < (1 << cp.dim); ++col)
      {
        const int64_t n = colour_count(cp, col);
        if (n == 0)
          continue;
        if (cp.dim == 3)
          k_ch_factor<3><<<blocks_for(n), 256, 0, s>>>(cp, ord, col, A);
        else
          k_ch_factor<2><<<blocks_for(n), 256, 0, s>>>(cp, ord, col, A);
      }
    const int64_t nA = cp.n_nodes * (cp.dim == 3 ? 27 : 9) * 4;
    k_ch_to_f32<3><<<(unsigned)std::min<int64_t>((nA + 255) / 256, 8192), 256, 0, s>>>(nA, A, A32);
  }

  template <typename T>
  void ch_apply(const ChParams &cp, const T *F, const T *inv_diag, const T *src, T *dst,
                hipStream_t s)
  {
    const int ncol = 1 << cp.dim;
    for (int col = 0; col < ncol; ++col)
      {
        const int64_t n = colour_count(cp, col);
        if (n == 0)
          continue;
        if (cp.dim == 3)
          k_ch_forward<3, T><<<blocks_for(n), 256, 0, s>>>(cp, col, F, src, dst);
        else
          k_ch_forward<2, T><<<blocks_for(n), 256, 0, s>>>(cp, col, F, src, dst);
      }
    for (int col = ncol - 1; col >= 0; --col)
      {
        const int64_t n = colour_count(cp, col);
        if (n == 0)
          continue;
        if (cp.dim == 3)
          k_ch_backward<3, T><<<blocks_for(n), 256, 0, s>>>(cp, col, F, inv_diag, src, dst);
        else
          k_ch_backward<2, T><<<blocks_for(n), 256, 0, s>>>(cp, col, F, inv_diag, src, dst);
      }
  }
  template void ch_apply<double>(const ChParams &, const double *, const double *,
                                 const double *, double *, hipStream_t);
  template void ch_apply<float>(const ChParams &, const float *, const float *, const float *,
                                float *, hipStream_t);
} // namespace hps
