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
// 3 (dy+1) [+ 9 (dz+1)]; the rows are stored colour by colour and stencil
// entry by stencil entry (ChParams::off, cnt), so that the threads of one
// colour stage read consecutive 4-vectors.
//
// Ordering: nodes in 2^DIM colours by the parity of their coordinates
// (colour = x&1 | (y&1)<<1 | (z&1)<<2), nodes of one colour in index order,
// c before mu within a node.  No two neighbours share a colour, so every
// stage of the factorisation and of the triangular solves (one colour) is a
// parallel loop over the colour's nodes: 2^DIM launches per factorisation
// and 2^(DIM+1) - 1 per solve (the last forward stage is also the first
// backward stage).  ILU(0) keeps the pattern of the block; the rows are
// eliminated in the IKJ order (lower entries in increasing column order).
// The solves keep (c, mu) in a work vector in the same colour-major node
// order, so the neighbours that a thread reads are consecutive for
// consecutive threads, as its rows are; one more launch writes dst in node
// order, one thread per entry (coalesced), with the Allen-Cahn components.
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
#include <utility>

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

    // the stored row of a node: entry k = 4 s + 2 a + b at
    // base[(k / 4) stride + k % 4]
    template <typename T>
    struct RowRef
    {
      T      *base;
      int64_t stride;
      __device__ __forceinline__ T &operator[](int k) const
      {
        return base[(k >> 2) * stride + (k & 3)];
      }
    };

    // the row of node (x, y, z) of colour col (uniform across the lanes
    // where it is called, so the parameter arrays are read by scalar loads)
    template <typename T>
    __device__ __forceinline__ RowRef<T> row_of(const ChParams &cp, T *F, int col, int x, int y,
                                                int z)
    {
      const int64_t ic = (x >> 1) + (int64_t)cp.mx[col] * ((y >> 1) + (int64_t)cp.my[col] * (z >> 1));
      return RowRef<T>{F + cp.off[col] + 4 * ic, 4 * cp.cnt[col]};
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
        // the colour of corner k: the cell parity flipped by the corner's
        // bits (uniform)
        const RowRef<double> ra = row_of(cp, A, par ^ k, kx, ky, kz);
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

  // Tensorial mobility (S/mobility.h:758-916): the c row's flux is linear
  // in (delta c, grad delta c, grad delta mu) with Gauss-point coefficients
  // b (of delta c), Dc (of grad delta c: the projector derivative) and Mt
  // (of grad delta mu: the mobility tensor).  One thread per (cell, corner
  // k): thread k finds the coefficients of Gauss point k by evaluating the
  // tangent's c row (tangent_qp_rows) at unit inputs, shares them through
  // LDS, and adds the element-matrix row of its corner:
  //   (c, c):   JxW (w0 phi_k phi_m + (b phi_m + Dc grad phi_m) . grad phi_k)
  //   (c, mu):  JxW (Mt grad phi_m) . grad phi_k
  //   (mu, c), (mu, mu) as with scalar mobility.
  template <int DIM, int G>
  __global__ void __launch_bounds__(256)
    k_ch_assemble_tens(const ChParams cp, int par, const double *__restrict__ cache,
                       double *__restrict__ A)
  {
    using CL            = CacheLayout<DIM, G, true>;
    constexpr int NV    = 1 << DIM;
    constexpr int NF    = CL::n_fields;
    constexpr int CPB   = 256 / NV; // cells per workgroup
    constexpr int NCOEF = DIM + 2 * DIM * DIM + 1;
    __shared__ double coef[CPB][NV][NCOEF]; // b, Dc, Mt (row-major), d2f/dc2
    const int     px = par & 1, py = (par >> 1) & 1, pz = (par >> 2) & 1;
    const int     mx = (cp.nc[0] - px + 1) / 2, my = (cp.nc[1] - py + 1) / 2;
    const int     mz = DIM == 3 ? (cp.nc[2] - pz + 1) / 2 : 1;
    const int     lc = (int)threadIdx.x / NV, k = (int)threadIdx.x % NV;
    const int64_t i  = (int64_t)blockIdx.x * CPB + lc;
    const bool    valid = i < (int64_t)mx * my * mz;
    const int     cx = px + 2 * (int)(i % mx);
    const int     cy = py + 2 * (int)((i / mx) % my);
    const int     cz = DIM == 3 ? pz + 2 * (int)(i / ((int64_t)mx * my)) : 0;
    if (valid)
      {
        // the cell's place in the cache (as k_ch_assemble)
        const int     cm    = DIM == 3 ? cz : cy;
        const int     cyy   = DIM == 3 ? cy : 0;
        const int64_t tile  = (int64_t)(cx / cp.tx) +
                             (int64_t)cp.nt[0] * ((cyy / cp.ty) + (int64_t)cp.nt[1] * (cm / cp.tz));
        const int     layer = cm % cp.tz;
        const int     lane  = ((((cyy % cp.ty) * cp.tx) + (cx % cp.tx)) << DIM) + k;
        const double *cl    = cache + ((tile * cp.tz + layer) * cp.nf) * (int64_t)cp.nt_lanes;
        double        lin[NF];
#pragma unroll
        for (int f = 0; f < NF; ++f)
          lin[f] = cl[f * (int64_t)cp.nt_lanes + lane];
        double v[G + 2], g[G + 2][DIM], vr[G + 2], gr[G + 2][DIM];
#pragma unroll
        for (int c = 0; c < G + 2; ++c)
          {
            v[c] = 0.0;
#pragma unroll
            for (int d = 0; d < DIM; ++d)
              g[c][d] = 0.0;
          }
        double *out = coef[lc][k];
        v[0]        = 1.0;
        tangent_qp_rows<DIM, G, true, false, true>(cp.pp, lin, v, g, vr, gr);
        v[0] = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d)
          out[d] = gr[0][d];
#pragma unroll
        for (int e = 0; e < DIM; ++e)
          {
            g[0][e] = 1.0;
            tangent_qp_rows<DIM, G, true, false, true>(cp.pp, lin, v, g, vr, gr);
            g[0][e] = 0.0;
#pragma unroll
            for (int d = 0; d < DIM; ++d)
              out[DIM + d * DIM + e] = gr[0][d];
            g[1][e] = 1.0;
            tangent_qp_rows<DIM, G, true, false, true>(cp.pp, lin, v, g, vr, gr);
            g[1][e] = 0.0;
#pragma unroll
            for (int d = 0; d < DIM; ++d)
              out[DIM + DIM * DIM + d * DIM + e] = gr[0][d];
          }
        const double c = lin[CL::c], A_ = cp.pp.A, B_ = cp.pp.B;
        out[DIM + 2 * DIM * DIM] = 2.0 * A_ * (c * c) + 4.0 * A_ * c * (2.0 * c - 2.0) +
                                   2.0 * A_ * ((1.0 - c) * (1.0 - c)) + 2.0 * B_;
      }
    __syncthreads();
    if (!valid)
      return;
    constexpr double sa  = 0.78867513459481288225457439025098;
    constexpr double sb  = 0.21132486540518711774542560974902;
    const double     jxw = cp.jxw;
    const int        kx = cx + (k & 1), ky = cy + ((k >> 1) & 1);
    const int        kz = DIM == 3 ? cz + ((k >> 2) & 1) : 0;
    // the colour of corner k: the cell parity flipped by the corner's bits
    // (uniform only per corner, so the row is addressed per lane here)
    const int            col = par ^ k;
    const int64_t        ic  = (kx >> 1) + (int64_t)cp.mx[col] * ((ky >> 1) + (int64_t)cp.my[col] * (kz >> 1));
    double              *rb  = A + cp.off[col] + 4 * ic;
    const int64_t        rs  = 4 * cp.cnt[col];
#pragma unroll 1
    for (int m = 0; m < NV; ++m)
      {
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
#pragma unroll
        for (int q = 0; q < NV; ++q)
          {
            const double *cf = coef[lc][q];
            double        vk = 1.0, vm = 1.0, gk[DIM], gm[DIM];
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
            double gg = 0.0, cc = 0.0, cm = 0.0;
#pragma unroll
            for (int d = 0; d < DIM; ++d)
              {
                gg += gm[d] * gk[d];
                double fc = cf[d] * vm, fmu = 0.0;
#pragma unroll
                for (int e = 0; e < DIM; ++e)
                  {
                    fc += cf[DIM + d * DIM + e] * gm[e];
                    fmu += cf[DIM + DIM * DIM + d * DIM + e] * gm[e];
                  }
                cc += fc * gk[d];
                cm += fmu * gk[d];
              }
            const double vv = vm * vk;
            a0 += cp.pp.w0 * vv + cc;
            a1 += cm;
            a2 += -cf[DIM + 2 * DIM * DIM] * vv - cp.pp.kappa_c * gg;
            a3 += cp.pp.w_alg * vv;
          }
        const int dx = (m & 1) - (k & 1), dy = ((m >> 1) & 1) - ((k >> 1) & 1);
        const int dz = DIM == 3 ? ((m >> 2) & 1) - ((k >> 2) & 1) : 0;
        const int s  = Sten<DIM>::index(dx, dy, dz);
        double   *e4 = rb + s * rs;
        e4[0] += jxw * a0;
        e4[1] += jxw * a1;
        e4[2] += jxw * a2;
        e4[3] += jxw * a3;
      }
  }

  // constrained nodes (OperatorBase::vmult, S/operator_base.h:294-353: the
  // constrained src entries are zeroed and the constrained rows are
  // identity rows): the Cahn-Hilliard block of the constrained operator has
  // identity rows and columns at those nodes.  One thread per constrained
  // node: its row becomes the identity, and the entry of each neighbour's
  // row that couples to it becomes zero.  Two neighbouring constrained
  // nodes may zero the same entries; every write is a zero or the node's
  // own identity block, so the order does not matter.
  template <int DIM>
  __global__ void __launch_bounds__(256)
    k_ch_constrain(const ChParams cp, const int64_t *__restrict__ nodes, int64_t n,
                   double *__restrict__ A)
  {
    constexpr int S = Sten<DIM>::S;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n)
      return;
    const int64_t node = nodes[t];
    const int     x = (int)(node % cp.nn[0]);
    const int     y = (int)((node / cp.nn[0]) % cp.nn[1]);
    const int     z = DIM == 3 ? (int)(node / ((int64_t)cp.nn[0] * cp.nn[1])) : 0;
    auto row = [&](int xx, int yy, int zz, int64_t &base, int64_t &stride) {
      const int col = (xx & 1) | ((yy & 1) << 1) | ((zz & 1) << 2);
      base = cp.off[col] +
             4 * ((xx >> 1) + (int64_t)cp.mx[col] * ((yy >> 1) + (int64_t)cp.my[col] * (zz >> 1)));
      stride = 4 * cp.cnt[col];
    };
    int64_t b, st;
    row(x, y, z, b, st);
    const int centre = Sten<DIM>::index(0, 0, 0);
    for (int s = 0; s < S; ++s)
      {
        double *e = A + b + s * st;
        e[0]      = s == centre ? 1.0 : 0.0;
        e[1]      = 0.0;
        e[2]      = 0.0;
        e[3]      = s == centre ? 1.0 : 0.0;
      }
    for (int dz = DIM == 3 ? -1 : 0; dz <= (DIM == 3 ? 1 : 0); ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx)
          {
            const int xx = x + dx, yy = y + dy, zz = z + dz;
            if ((dx == 0 && dy == 0 && dz == 0) || xx < 0 || yy < 0 || zz < 0 ||
                xx >= cp.nn[0] || yy >= cp.nn[1] || zz >= cp.nn[2])
              continue;
            int64_t nb, nst;
            row(xx, yy, zz, nb, nst);
            double *e = A + nb + Sten<DIM>::index(-dx, -dy, -dz) * nst;
            e[0]      = 0.0;
            e[1]      = 0.0;
            e[2]      = 0.0;
            e[3]      = 0.0;
          }
  }

  // a node of colour COL with in-colour index ic: its place in the colour's
  // grid (ix, iy, iz), its coordinates and its node number
  template <int DIM, int COL>
  struct ColourNode
  {
    int64_t ic, node;
    int     ix, iy, iz, x, y, z;
    __device__ __forceinline__ ColourNode(const ChParams &cp, int64_t ic_)
      : ic(ic_)
    {
      const int mx = cp.mx[COL], my = cp.my[COL];
      ix   = (int)(ic % mx);
      iy   = (int)((ic / mx) % my);
      iz   = DIM == 3 ? (int)(ic / ((int64_t)mx * my)) : 0;
      x    = (COL & 1) + 2 * ix;
      y    = ((COL >> 1) & 1) + 2 * iy;
      z    = DIM == 3 ? ((COL >> 2) & 1) + 2 * iz : 0;
      node = ((int64_t)z * cp.nn[1] + y) * cp.nn[0] + x;
    }
  };

  // ------------------------------------------------------------ factorisation
  // stencil index s <-> offset (dx, dy, dz); parity mask of the offset
  __host__ __device__ constexpr int st_dx(int s) { return s % 3 - 1; }
  __host__ __device__ constexpr int st_dy(int s) { return (s / 3) % 3 - 1; }
  __host__ __device__ constexpr int st_dz(int DIM, int s) { return DIM == 3 ? s / 9 - 1 : 0; }
  __host__ __device__ constexpr int st_mask(int DIM, int s)
  {
    return (st_dx(s) & 1) | ((st_dy(s) & 1) << 1) | ((st_dz(DIM, s) & 1) << 2);
  }

  // the neighbours of a node of colour COL with a lower colour, in the
  // order of their columns: by colour, then by stencil index (the node
  // index order within a colour); and those with a higher colour
  template <int DIM, int COL>
  struct FactorLists
  {
    int n_lower = 0, n_upper = 0;
    int lower[26] = {}, upper[26] = {};
    constexpr FactorLists()
    {
      constexpr int S = DIM == 3 ? 27 : 9;
      for (int jc = 0; jc < COL; ++jc)
        for (int t = 0; t < S; ++t)
          if ((COL ^ st_mask(DIM, t)) == jc)
            lower[n_lower++] = t;
      for (int t = 0; t < S; ++t)
        if ((COL ^ st_mask(DIM, t)) > COL)
          upper[n_upper++] = t;
    }
  };

  // row (i, AR) (AR = 0: c, 1: mu) of the nodes of colour COL, held in
  // registers: R[s] = the entries (AR, c) and (AR, mu) of stencil entry s.
  // IKJ order: the lower entries (j, b) in increasing column order -- the
  // neighbours j of lower colours by (colour, index), c before mu, and for
  // the mu row finally (i, c) -- each eliminated with the already final row
  // (j, b): l = a / u_(j,b),(j,b); a_(i,AR),(k,e) -= l u_(j,b),(k,e) for
  // the (k, e) after (j, b) in row (j, b) that are in the pattern of row
  // (i, AR): (j, mu) when b = c, and the neighbours k of j of a higher
  // colour than j's that are neighbours of i (or i itself).  Entries of
  // neighbours outside the domain are zero in the assembled block and stay
  // zero.  The final row is written in FP64 and FP32.
  template <int DIM, int COL, int AR>
  __device__ __forceinline__ void ch_factor_row(const ChParams &cp, const ColourNode<DIM, COL> &n,
                                                double *__restrict__ A, float *__restrict__ A32)
  {
    constexpr int                   S      = DIM == 3 ? 27 : 9;
    constexpr int                   center = (S - 1) / 2;
    constexpr FactorLists<DIM, COL> fl{};
    constexpr int px = COL & 1, py = (COL >> 1) & 1, pz = (COL >> 2) & 1;
    double2      *Ai = reinterpret_cast<double2 *>(A + cp.off[COL]) + 2 * n.ic + AR;
    const int64_t st = 2 * cp.cnt[COL];  // double2 stride between stencil entries
    double        R[S][2];
#pragma unroll
    for (int s = 0; s < S; ++s)
      {
        const double2 v = Ai[s * st];
        R[s][0]         = v.x;
        R[s][1]         = v.y;
      }
#pragma unroll
    for (int t = 0; t < fl.n_lower; ++t)
      {
        const int  tt = fl.lower[t];
        const int  dx = st_dx(tt), dy = st_dy(tt), dz = st_dz(DIM, tt);
        const int  jc = COL ^ st_mask(DIM, tt);
        const bool ok = (dx >= 0 || n.x > 0) && (dx <= 0 || n.x + 1 < cp.nn[0]) &&
                        (dy >= 0 || n.y > 0) && (dy <= 0 || n.y + 1 < cp.nn[1]) &&
                        (dz >= 0 || n.z > 0) && (dz <= 0 || n.z + 1 < cp.nn[2]);
        if (!ok)
          continue;
        const int      jx = n.ix + ((px + dx) >> 1), jy = n.iy + ((py + dy) >> 1);
        const int      jz = n.iz + ((pz + dz) >> 1);
        const int64_t  j  = jx + (int64_t)cp.mx[jc] * (jy + (int64_t)cp.my[jc] * jz);
        const double4 *Aj = reinterpret_cast<const double4 *>(A + cp.off[jc]) + j;
        const double4  uc = Aj[center * cp.cnt[jc]];
        // (j, c), then (j, mu) with the entry that (j, c) updated
        const double l0 = R[tt][0] / uc.x;
        R[tt][0]        = l0;
        R[tt][1] -= l0 * uc.y;
        const double l1 = R[tt][1] / uc.w;
        R[tt][1]        = l1;
#pragma unroll
        for (int u = 0; u < S; ++u)
          {
            if (u == center || (jc ^ st_mask(DIM, u)) <= jc)
              continue;
            const int fx = dx + st_dx(u), fy = dy + st_dy(u), fz = dz + st_dz(DIM, u);
            if (fx < -1 || fx > 1 || fy < -1 || fy > 1 || fz < -1 || fz > 1)
              continue;
            const int     v  = DIM == 3 ? (fx + 1) + 3 * (fy + 1) + 9 * (fz + 1)
                                        : (fx + 1) + 3 * (fy + 1);
            const double4 uu = Aj[u * cp.cnt[jc]];
            R[v][0] -= l0 * uu.x;
            R[v][1] -= l0 * uu.y;
            R[v][0] -= l1 * uu.z;
            R[v][1] -= l1 * uu.w;
          }
      }
    if (AR == 1)
      {
        // the lower entry (i, c) of row (i, mu), with the final row (i, c)
        const double2 *Ac = Ai - 1;
        const double2  cc = Ac[center * st];
        const double   l  = R[center][0] / cc.x;
        R[center][0]      = l;
        R[center][1] -= l * cc.y;
#pragma unroll
        for (int t = 0; t < fl.n_upper; ++t)
          {
            const double2 cu = Ac[fl.upper[t] * st];
            R[fl.upper[t]][0] -= l * cu.x;
            R[fl.upper[t]][1] -= l * cu.y;
          }
      }
    float2 *Ai32 = reinterpret_cast<float2 *>(A32 + cp.off[COL]) + 2 * n.ic + AR;
#pragma unroll
    for (int s = 0; s < S; ++s)
      {
        Ai[s * st]   = double2{R[s][0], R[s][1]};
        Ai32[s * st] = float2{(float)R[s][0], (float)R[s][1]};
      }
  }

  // Workgroups are dispatched to the 8 XCDs round-robin (workgroup b runs
  // on XCD b % 8).  The logical block returned here gives each XCD one
  // contiguous range of the colour grid, so that the neighbour rows a stage
  // gathers from the other colours are shared in that XCD's L2.  The map is
  // a bijection of [0, gridDim.x): XCD x < r takes q + 1 blocks, the others
  // q, with nb = 8 q + r.
  __device__ __forceinline__ int xcd_block()
  {
    const int nb = (int)gridDim.x, b = (int)blockIdx.x;
    const int q = nb >> 3, r = nb & 7, x = b & 7, k = b >> 3;
    return x * q + (x < r ? x : r) + k;
  }

  template <int DIM, int COL>
  __global__ void __launch_bounds__(256)
    k_ch_factor(const ChParams cp, double *__restrict__ A, float *__restrict__ A32)
  {
    const int64_t ic = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
    if (ic >= cp.cnt[COL])
      return;
    const ColourNode<DIM, COL> n(cp, ic);
    ch_factor_row<DIM, COL, 0>(cp, n, A, A32);
    ch_factor_row<DIM, COL, 1>(cp, n, A, A32);
  }

  // ------------------------------------------------------------------ solves
  template <typename T>
  struct Vec;
  template <>
  struct Vec<float>
  {
    using V2 = float2;
    using V4 = float4;
  };
  template <>
  struct Vec<double>
  {
    using V2 = double2;
    using V4 = double4;
  };

  // the stencil entries of a node of colour COL whose neighbour has a lower
  // (LOWER) or a higher colour, in stencil order
  template <int DIM, int COL, bool LOWER>
  struct Neighbours
  {
    int n     = 0;
    int s[26] = {};
    constexpr Neighbours()
    {
      for (int t = 0; t < (DIM == 3 ? 27 : 9); ++t)
        {
          const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = DIM == 3 ? t / 9 - 1 : 0;
          const int jc = COL ^ ((dx & 1) | ((dy & 1) << 1) | ((dz & 1) << 2));
          if (LOWER ? jc < COL : jc > COL)
            s[n++] = t;
        }
    }
  };

  // (s0, s1) -= sum over the lower (LOWER) or upper neighbours j of the
  // node's 2x2 entry times the work vector's (c, mu) at j
  template <int DIM, int COL, bool LOWER, typename T>
  __device__ __forceinline__ void ch_gather(const ChParams &cp, const T *__restrict__ F,
                                            const T *__restrict__ W, const ColourNode<DIM, COL> &n,
                                            T &s0, T &s1)
  {
    using V2 = typename Vec<T>::V2;
    using V4 = typename Vec<T>::V4;
    constexpr Neighbours<DIM, COL, LOWER> nb{};
    constexpr int px = COL & 1, py = (COL >> 1) & 1, pz = (COL >> 2) & 1;
    const V4     *Fr = reinterpret_cast<const V4 *>(F + cp.off[COL]) + n.ic;
    const V2     *W2 = reinterpret_cast<const V2 *>(W);
#pragma unroll
    for (int t = 0; t < nb.n; ++t)
      {
        const int sx = nb.s[t];
        const int dx = sx % 3 - 1, dy = (sx / 3) % 3 - 1, dz = DIM == 3 ? sx / 9 - 1 : 0;
        const int jc = COL ^ ((dx & 1) | ((dy & 1) << 1) | ((dz & 1) << 2));
        const bool ok = (dx >= 0 || n.x > 0) && (dx <= 0 || n.x + 1 < cp.nn[0]) &&
                        (dy >= 0 || n.y > 0) && (dy <= 0 || n.y + 1 < cp.nn[1]) &&
                        (dz >= 0 || n.z > 0) && (dz <= 0 || n.z + 1 < cp.nn[2]);
        // the neighbour's place in its colour's grid: (x + dx) / 2 etc.;
        // outside the domain the loads read the colour's first node and the
        // select drops the term (the loads of all neighbours stay
        // unconditional, so they issue together)
        const int     jx = n.ix + ((px + dx) >> 1), jy = n.iy + ((py + dy) >> 1);
        const int     jz = n.iz + ((pz + dz) >> 1);
        const int64_t j  = ok ? jx + (int64_t)cp.mx[jc] * (jy + (int64_t)cp.my[jc] * jz) : 0;
        const V2      w  = W2[cp.noff[jc] + j];
        const V4      f  = Fr[sx * cp.cnt[COL]];
        const T       t0 = f.x * w.x + f.y * w.y, t1 = f.z * w.x + f.w * w.y;
        s0 -= ok ? t0 : T(0);
        s1 -= ok ? t1 : T(0);
      }
  }

  // forward stage of colour COL: y_(i,c) = r_(i,c) - sum_lower L y,
  // y_(i,mu) = r_(i,mu) - sum_lower L y - l_(mu,c) y_(i,c), into the work
  // vector.  The last colour has no upper neighbours, so its stage also
  // finishes the backward solve of its nodes:
  // x_mu = y_mu / u_(mu,mu), x_c = (y_c - u_(c,mu) x_mu) / u_(c,c)
  template <int DIM, int COL, typename T>
  __global__ void __launch_bounds__(256)
    k_ch_forward(const ChParams cp, const T *__restrict__ F, const T *__restrict__ src,
                 T *__restrict__ W)
  {
    using V2            = typename Vec<T>::V2;
    using V4            = typename Vec<T>::V4;
    constexpr bool LAST = COL == (1 << DIM) - 1;
    const int64_t  ic   = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
    if (ic >= cp.cnt[COL])
      return;
    const ColourNode<DIM, COL> n(cp, ic);
    T                          s0 = src[n.node * cp.n_comp + 0];
    T                          s1 = src[n.node * cp.n_comp + 1];
    ch_gather<DIM, COL, true>(cp, F, W, n, s0, s1);
    constexpr int center = DIM == 3 ? 13 : 4;
    const V4      u      = reinterpret_cast<const V4 *>(F + cp.off[COL])[center * cp.cnt[COL] + ic];
    s1 -= u.z * s0;
    if (LAST)
      {
        s1 = s1 / u.w;
        s0 = (s0 - u.y * s1) / u.x;
      }
    reinterpret_cast<V2 *>(W)[cp.noff[COL] + ic] = V2{s0, s1};
  }

  // backward stage of colour COL (all but the last): x_(i,a) from the work
  // vector's y_(i,a) minus the upper entries times the final x of the
  // higher colours, x_mu first, into the work vector
  template <int DIM, int COL, typename T>
  __global__ void __launch_bounds__(256)
    k_ch_backward(const ChParams cp, const T *__restrict__ F, T *__restrict__ W)
  {
    using V2         = typename Vec<T>::V2;
    using V4         = typename Vec<T>::V4;
    const int64_t ic = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
    if (ic >= cp.cnt[COL])
      return;
    const ColourNode<DIM, COL> n(cp, ic);
    V2                        *W2 = reinterpret_cast<V2 *>(W);
    const V2                   yv = W2[cp.noff[COL] + ic];
    T                          s0 = yv.x, s1 = yv.y;
    ch_gather<DIM, COL, false>(cp, F, W, n, s0, s1);
    constexpr int center = DIM == 3 ? 13 : 4;
    const V4      u      = reinterpret_cast<const V4 *>(F + cp.off[COL])[center * cp.cnt[COL] + ic];
    const T       xm     = s1 / u.w;
    const T       xc     = (s0 - u.y * xm) / u.x;
    W2[cp.noff[COL] + ic] = V2{xc, xm};
  }

  // dst in node order, one workgroup per node row (y, z), its threads
  // striding over the row's entries: (c, mu) from the work vector (the
  // node's colour and place in its colour's grid), the Allen-Cahn components
  // inv_diag * src.  Each thread has several entries in flight, and the
  // index arithmetic stays in 32 bits.
  template <int DIM, typename T>
  __global__ void __launch_bounds__(256)
    k_ch_scatter(const ChParams cp, const T *__restrict__ W, const T *__restrict__ inv_diag,
                 const T *__restrict__ src, T *__restrict__ dst)
  {
    const int nc  = cp.n_comp;
    const int len = cp.nn[0] * nc;
    const int row = (int)blockIdx.x;
    // the row's two colours (x even, x odd): uniform values, selected per
    // lane (a per-lane index into the parameter arrays would load them from
    // memory lane by lane)
    const int     y   = row % cp.nn[1];
    const int     z   = row / cp.nn[1];
    const int     c0  = ((y & 1) << 1) | ((z & 1) << 2);
    const int64_t yz0 = (y >> 1) + (int64_t)cp.my[c0] * (z >> 1);
    const int64_t yz1 = (y >> 1) + (int64_t)cp.my[c0 | 1] * (z >> 1);
    const int64_t b0  = cp.noff[c0] + (int64_t)cp.mx[c0] * yz0;
    const int64_t b1  = cp.noff[c0 | 1] + (int64_t)cp.mx[c0 | 1] * yz1;
    const int64_t r0  = (int64_t)row * len;
#pragma unroll 4
    for (int t = (int)threadIdx.x; t < len; t += 256)
      {
        const int64_t i = r0 + t;
        const int     x = t / nc;
        const int     c = t - x * nc;
        T             v;
        if (c >= 2)
          v = inv_diag[i] * src[i];
        else
          v = W[2 * (((x & 1) ? b1 : b0) + (x >> 1)) + c];
        dst[i] = v;
      }
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
    cp.noff[0]  = 0;
    for (int col = 0; col < 8; ++col)
      {
        const int px = col & 1, py = (col >> 1) & 1;
        cp.mx[col]       = (cp.nn[0] - px + 1) / 2;
        cp.my[col]       = (cp.nn[1] - py + 1) / 2;
        cp.cnt[col]      = col < (1 << cp.dim) ? colour_count(cp, col) : 0;
        cp.off[col + 1]  = cp.off[col] + cp.cnt[col] * S * 4;
        cp.noff[col + 1] = cp.noff[col] + cp.cnt[col];
      }
  }

  template <int DIM>
  static void ch_assemble_tens_dim(const ChParams &cp, int par, int64_t n, const double *cache,
                                   double *A, hipStream_t s)
  {
    const unsigned b = (unsigned)((n + (256 >> DIM) - 1) / (256 >> DIM));
    switch (cp.G)
      {
#define HPS_CASE(G)                                                                \
  case G:                                                                          \
    k_ch_assemble_tens<DIM, G><<<b, 256, 0, s>>>(cp, par, cache, A);               \
    break;
        HPS_CASE(1)
        HPS_CASE(2)
        HPS_CASE(3)
        HPS_CASE(4)
        HPS_CASE(5)
        HPS_CASE(6)
        HPS_CASE(7)
        HPS_CASE(8)
        HPS_CASE(9)
        HPS_CASE(10)
        HPS_CASE(11)
        HPS_CASE(12)
#undef HPS_CASE
        default:
          break;
      }
  }

  static void ch_assemble_tens(const ChParams &cp, int par, int64_t n, const double *cache,
                               double *A, hipStream_t s)
  {
    if (cp.dim == 3)
      ch_assemble_tens_dim<3>(cp, par, n, cache, A, s);
    else
      ch_assemble_tens_dim<2>(cp, par, n, cache, A, s);
  }

  void ch_constrain(const ChParams &cp, const int64_t *nodes, int64_t n, double *A,
                    hipStream_t s)
  {
    if (n <= 0)
      return;
    if (cp.dim == 3)
      k_ch_constrain<3><<<blocks_for(n), 256, 0, s>>>(cp, nodes, n, A);
    else
      k_ch_constrain<2><<<blocks_for(n), 256, 0, s>>>(cp, nodes, n, A);
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
        if (cp.tens)
          ch_assemble_tens(cp, par, n, cache, A, s);
        else if (cp.dim == 3)
          k_ch_assemble<3><<<blocks_for(n), 256, 0, s>>>(cp, par, cache, A);
        else
          k_ch_assemble<2><<<blocks_for(n), 256, 0, s>>>(cp, par, cache, A);
      }
  }

  template <int DIM, int... C>
  static void ch_factor_stages(std::integer_sequence<int, C...>, const ChParams &cp, double *A,
                               float *A32, hipStream_t s)
  {
    ((cp.cnt[C] ? k_ch_factor<DIM, C><<<blocks_for(cp.cnt[C]), 256, 0, s>>>(cp, A, A32) : void()),
     ...);
  }

  void ch_factor(const ChParams &cp, double *A, float *A32, hipStream_t s)
  {
    if (cp.dim == 3)
      ch_factor_stages<3>(std::make_integer_sequence<int, 8>{}, cp, A, A32, s);
    else
      ch_factor_stages<2>(std::make_integer_sequence<int, 4>{}, cp, A, A32, s);
  }

  // one solve stage: forward (FWD) or backward of colour COL; the backward
  // stage of the last colour is part of its forward stage
  template <int DIM, int COL, bool FWD, typename T>
  static void ch_stage(const ChParams &cp, const T *F, const T *inv_diag, const T *src, T *dst,
                       hipStream_t s)
  {
    if (cp.cnt[COL] == 0)
      return;
    T *W = const_cast<T *>(F) + cp.off[8];
    if constexpr (FWD)
      k_ch_forward<DIM, COL, T><<<blocks_for(cp.cnt[COL]), 256, 0, s>>>(cp, F, src, W);
    else if constexpr (COL < (1 << DIM) - 1)
      k_ch_backward<DIM, COL, T><<<blocks_for(cp.cnt[COL]), 256, 0, s>>>(cp, F, W);
  }

  template <int DIM, typename T, int... C>
  static void ch_solves(std::integer_sequence<int, C...>, const ChParams &cp, const T *F,
                        const T *inv_diag, const T *src, T *dst, hipStream_t s)
  {
    // forward in increasing colour order, then backward in decreasing order
    (ch_stage<DIM, C, true, T>(cp, F, inv_diag, src, dst, s), ...);
    (ch_stage<DIM, (int)sizeof...(C) - 1 - C, false, T>(cp, F, inv_diag, src, dst, s), ...);
    k_ch_scatter<DIM, T>
      <<<(unsigned)(cp.nn[1] * cp.nn[2]), 256, 0, s>>>(cp, F + cp.off[8], inv_diag, src, dst);
  }

  template <typename T>
  void ch_apply(const ChParams &cp, const T *F, const T *inv_diag, const T *src, T *dst,
                hipStream_t s)
  {
    if (cp.dim == 3)
      ch_solves<3, T>(std::make_integer_sequence<int, 8>{}, cp, F, inv_diag, src, dst, s);
    else
      ch_solves<2, T>(std::make_integer_sequence<int, 4>{}, cp, F, inv_diag, src, dst, s);
  }
  template void ch_apply<double>(const ChParams &, const double *, const double *,
                                 const double *, double *, hipStream_t);
  template void ch_apply<float>(const ChParams &, const float *, const float *, const float *,
                                float *, hipStream_t);
} // namespace hps
