// fe_q1_cut.hpp -- grain cut-off: per-cell masks and the fast tangent vmult
// of the 3D scalar-mobility Q1 operator.
//
// The reference's cut-off (SinteringOperatorData::set_component_mask,
// S/operator_sintering_data.h:150-298, with do_vmult_range_no_template and
// do_evaluate_nonlinear_residual_nt, S/operator_sintering_generic.h:456-537,
// 1008-1081) evaluates each cell batch with only the order parameters whose
// nodal values exceed the tolerance somewhere in the batch: the others are
// dropped from the batch's physics and receive no contribution from it.
// Here the batch is one cell (the reference with a batch width of one):
// k_cell_masks sets bit g of a cell's mask when max |eta_g| over the cell's
// nodes > tol.  The masked kernels (CUT = true in fe_q1.hpp) apply the
// masks to all G order parameters.  The fast tangent vmult below runs a tile
// (the TX x TY x tz cells of one workgroup) with GA + 2 components: c, mu and
// the GA order parameters active in some cell of the tile (the union of its
// cell masks, k_tile_masks); inside the tile every cell still drops those
// of its own mask's zeros.  It writes zeros to the other components of the
// tile's nodes, so dst and the shared-node partials are complete for the
// usual fix-up.
#pragma once

#include "fe_q1.hpp"

namespace hps
{
  // bit g of masks[cell]: max |eta_g| over the cell's nodes > tol
  // (SinteringOperatorData::set_component_mask's l-infinity rule); one
  // thread per cell, cells x fastest
  __global__ void __launch_bounds__(256)
    k_cell_masks(int dim, int64_t n_cells, int ncx, int ncy, int nnx, int nny, int G,
                 double tol, const double *__restrict__ u, unsigned *__restrict__ masks)
  {
    const int64_t cell = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (cell >= n_cells)
      return;
    const int     cx = (int)(cell % ncx);
    const int64_t r  = cell / ncx;
    const int     cy = (int)(r % ncy);
    const int     cz = (int)(r / ncy);
    const int     NC = G + 2;
    const int     nv = dim == 3 ? 8 : 4;
    unsigned      m  = 0u;
    for (int v = 0; v < nv; ++v)
      {
        const int64_t node = ((int64_t)(cz + (v >> 2)) * nny + (cy + ((v >> 1) & 1))) * nnx +
                             (cx + (v & 1));
        const double *x = u + node * NC + 2;
        for (int g = 0; g < G; ++g)
          if (fabs(x[g]) > tol)
            m |= 1u << g;
      }
    masks[cell] = m;
  }

  // union of the cell masks of each tile
  template <int DIM, int TX, int TY>
  __global__ void __launch_bounds__(TX *TY *(1 << DIM))
    k_tile_masks(const KParams kp, unsigned *__restrict__ masks)
  {
    using TL = Tile<DIM, TX, TY>;
    const TL        t(kp);
    __shared__ unsigned m_sh;
    if (t.tid == 0)
      m_sh = 0u;
    __syncthreads();
    unsigned m = 0u;
    if (t.corner == 0 && t.cell_valid)
      for (int k = 0; k < t.tm_eff; ++k)
        m |= kp.cmask[t.cell_index(kp, k)];
    atomicOr(&m_sh, m);
    __syncthreads();
    if (t.tid == 0)
      masks[t.tile] = m_sh;
  }

  // global component of local component lc (0: c, 1: mu, 2 + s: the s-th
  // active order parameter)
  template <int NCA>
  __device__ __forceinline__ int cut_component(const int (&gc)[NCA], int lc)
  {
    int r = 0;
#pragma unroll
    for (int s = 0; s < NCA; ++s)
      r = lc == s ? gc[s] : r;
    return r;
  }

  // tangent vmult of the tiles tiles[blockIdx.x], each with exactly GA
  // active order parameters (scalar mobility, the folded form of q1_apply
  // with the rescaled butterflies).  FP64 with one active order parameter,
  // the most common tile, is held to 96 registers (5 waves/SIMD; its
  // natural allocation is 99).
  template <int DIM, int TX, int TY, int GA, typename T>
  __global__ void __launch_bounds__(TX *TY *(1 << DIM))
    __attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 && GA == 1 ? 5 : 1)))
    q1_vmult_cut(const KParams kp, int G, const int *__restrict__ tiles,
                 const unsigned *__restrict__ masks, const T *__restrict__ src,
                 const T *__restrict__ cache, T *__restrict__ dst, T *__restrict__ part)
  {
    using TL            = Tile<DIM, TX, TY>;
    constexpr int NCA   = GA + 2;
    constexpr int NT    = TL::NT;
    constexpr int NV    = TL::NV;
    constexpr int PN    = TL::PN;
    constexpr int ITEMS = (PN * NCA + NT - 1) / NT;
    constexpr int NFA   = 1 + GA + DIM; // local cached fields
    const int     NC    = G + 2;
    const int     NF    = 1 + G + DIM; // cached fields of the full state

    __shared__ T ring[2][PN * NCA];
    __shared__ T slots[NV][PN * NCA];

    const TL       t(kp, tiles[blockIdx.x]);
    const int      tid  = t.tid;
    const unsigned mask = masks[t.tile];
    int            gc[NCA];
    gc[0] = 0;
    gc[1] = 1;
    {
      unsigned m = mask;
#pragma unroll
      for (int s = 0; s < GA; ++s)
        {
          gc[2 + s] = 2 + (m ? __builtin_ctz(m) : 0);
          m &= m - 1u;
        }
    }

    FoldScales<T> fs;
    T             jxw_hat;
    fold_scales_hat<DIM>(kp, t.corner, fs, jxw_hat);

    T    regs[ITEMS];
    auto load_plane = [&](int k) {
#pragma unroll
      for (int r = 0; r < ITEMS; ++r)
        {
          const int i = tid + r * NT;
          if (i < PN * NCA)
            {
              const int pn = i / NCA, lc = i - pn * NCA;
              const int nx = pn % TL::PX, ny = pn / TL::PX;
              T         v  = T(0);
              if (t.node_in_tile(nx, ny))
                v = src[t.gnode(kp, nx, ny, k) * NC + cut_component<NCA>(gc, lc)];
              regs[r] = v;
            }
        }
    };
    auto store_plane = [&](int slot) {
#pragma unroll
      for (int r = 0; r < ITEMS; ++r)
        {
          const int i = tid + r * NT;
          if (i < PN * NCA)
            ring[slot][i] = regs[r];
        }
    };
    // zeros for the inactive components of plane k's nodes
    auto put_zeros = [&](int k) {
      for (int i = tid; i < PN * NC; i += NT)
        {
          const int pn = i / NC, c = i - pn * NC;
          if (c >= 2 && !((mask >> (c - 2)) & 1u))
            put(kp, t, NC, k, pn, c, T(0), dst, part, NC, 1, 0);
        }
    };

    T carry[ITEMS];
#pragma unroll
    for (int r = 0; r < ITEMS; ++r)
      carry[r] = T(0);
    load_plane(0);
    store_plane(0);
    load_plane(1);
    store_plane(1);
    __syncthreads();

    // the cache fields of the active components of layer k: c, eta_g,
    // grad mu.  As in q1_apply, the next layer's fields and the plane after
    // next are requested together at the end of a layer.
    const T *cache_tile = cache + (int64_t)t.tile * kp.tz * NF * NT;
    T        lin_next[NFA];
    auto     load_lin = [&](int k) {
      const T *cl = cache_tile + k * (NF * NT) + tid;
      lin_next[0] = cl[0];
#pragma unroll
      for (int s = 0; s < GA; ++s)
        lin_next[1 + s] = cl[(gc[2 + s] - 1) * NT];
#pragma unroll
      for (int d = 0; d < DIM; ++d)
        lin_next[1 + GA + d] = cl[(1 + G + d) * NT];
    };
    if (2 <= t.tm_eff)
      load_plane(2);
    load_lin(0);

    for (int k = 0; k < t.tm_eff; ++k)
      {
        const bool pre = k + 2 <= t.tm_eff;

        const T *pl = &ring[(k + t.bm) & 1][t.pn_lane * NCA];
        T        val[NCA], grad[NCA][DIM];
#pragma unroll
        for (int c = 0; c < NCA; ++c)
          interp_vg_hat<DIM>(pl[c], val[c], grad[c]);
        // the tile's order parameters that the lane's cell drops
        const unsigned cm = t.template cut_mask<true>(kp, k);
        unsigned       on = 0u; // bit s: local order parameter s active here
#pragma unroll
        for (int s = 0; s < GA; ++s)
          {
            const bool a = (cm >> (gc[2 + s] - 2)) & 1u;
            on |= (a ? 1u : 0u) << s;
            val[2 + s] = a ? val[2 + s] : T(0);
#pragma unroll
            for (int d = 0; d < DIM; ++d)
              grad[2 + s][d] = a ? grad[2 + s][d] : T(0);
          }
        __builtin_amdgcn_sched_barrier(0); // the physics waits for lin_next

        T *sl = &slots[t.corner][t.pn_lane * NCA];
        {
          T lin[NFA];
#pragma unroll
          for (int f = 0; f < NFA; ++f)
            lin[f] = lin_next[f];
          using CL = CacheLayout<DIM, GA, false>;
          T vr[NCA], gr[NCA][DIM];
          tangent_qp_rows<DIM, GA, false, true, false>(kp.pp, lin, val, grad, vr, gr);
          T M, sm;
          scalar_mobility_terms<GA>(kp.pp, lin[CL::c], lin + CL::eta, val, M, sm);
          T lmu[DIM];
#pragma unroll
          for (int d = 0; d < DIM; ++d)
            lmu[d] = lin[CL::grad_mu + d];
#pragma unroll
          for (int c = 0; c < NCA; ++c)
            {
              T gs[DIM];
#pragma unroll
              for (int d = 0; d < DIM; ++d)
                gs[d] = c == 0 ? (M * fs.kj[d]) * grad[1][d] + sm * (lmu[d] * fs.gsj[d])
                               : (c == 1 ? fs.kmu[d] * grad[0][d] : fs.keta[d] * grad[c][d]);
              sl[c] = integrate_vg_hat<DIM>(vr[c] * jxw_hat, gs);
              if (c >= 2 && !((on >> (c - 2)) & 1u))
                sl[c] = T(0);
            }
        }
        __syncthreads();
        if (pre)
          store_plane(k & 1);
        if (k + 3 <= t.tm_eff)
          load_plane(k + 3);
        if (k + 1 < t.tm_eff)
          load_lin(k + 1);

#pragma unroll
        for (int r = 0; r < ITEMS; ++r)
          {
            const int i = tid + r * NT;
            if (i < PN * NCA)
              {
                const int pn = i / NCA, lc = i - pn * NCA;
                const int nx = pn % TL::PX, ny = pn / TL::PX;
                T         fin = carry[r], nxt = T(0);
#pragma unroll
                for (int q = 0; q < NV; ++q)
                  {
                    const int qx = q & 1, qy = DIM == 3 ? (q >> 1) & 1 : 0;
                    const int qm = q >> (DIM - 1);
                    const int ccx = nx - qx, ccy = ny - qy;
                    if (ccx >= 0 && ccx < t.tx_eff && ccy >= 0 && ccy < t.ty_eff)
                      {
                        const T v = slots[q][i];
                        if (qm == 0)
                          fin += v;
                        else
                          nxt += v;
                      }
                  }
                carry[r] = nxt;
                put(kp, t, NC, k, pn, cut_component<NCA>(gc, lc), fin, dst, part, NC, 1, 0);
              }
          }
        put_zeros(k);
        __syncthreads();
      }
#pragma unroll
    for (int r = 0; r < ITEMS; ++r)
      {
        const int i = tid + r * NT;
        if (i < PN * NCA)
          {
            const int pn = i / NCA, lc = i - pn * NCA;
            put(kp, t, NC, t.tm_eff, pn, cut_component<NCA>(gc, lc), carry[r], dst, part, NC,
                1, 0);
          }
      }
    put_zeros(t.tm_eff);
  }
} // namespace hps
