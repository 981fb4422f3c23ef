// fe_q1.hpp -- Q1 cell kernels of the sintering operator for gfx950.
//
// What the reference computes with deal.II FEEvaluation on the CPU
// (gather, sum-factorised interpolation, quadrature-point physics,
// integration, scatter-add; S/operator_sintering_generic.h:539-577 for the
// tangent, :877-1006 for the residual, S/operator_sintering_data.h:301-409
// for the linearisation cache) is done here by one kernel per operation.
//
// Lane mapping.  One lane per (cell, corner): the 2^DIM lanes of a cell hold
// the cell's nodal values on entry, the quadrature-point values after
// interpolation and the nodal contributions after integration.  Lane bit d
// is the node/q-point index along direction d.  The 2-point Gauss
// interpolation along d is a butterfly between lanes that differ in bit d
// (DPP quad_perm for bits 0 and 1, ds_swizzle for bit 2):
//   value: sa*own + sb*partner            derivative: sg*(own - partner)/h
// and integration is its transpose:
//   value: sa*own + sb*partner            derivative: sg*(own + partner)
// with sa = (1+1/sqrt3)/2, sb = (1-1/sqrt3)/2, sg = +1 for bit 1, -1 for 0.
//
// Tiling.  A workgroup owns a column of TX x TY cells (x, y) and marches
// through a chunk of `tz` cell layers along the last direction (z in 3D, y
// in 2D).  Two planes of source nodes live in LDS; the planes one layer
// ahead are fetched into registers while the current layer computes.  Every
// lane writes its nodal contribution to an LDS slot indexed by (corner,
// node), so no two lanes write the same slot; after a barrier each node sums
// its slots in a fixed order (deterministic, no atomics).  The bottom corners
// complete node plane k, the top corners start plane k+1 (kept in registers
// as the carry).  Nodes shared with a neighbouring workgroup are written to
// a per-tile partial buffer and summed by q1_fixup in a fixed order.
//
// The linearisation cache is stored in lane order, cache[tile][layer]
// [field][lane], so each field of a layer is one coalesced 2 KiB read.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "physics.hpp"

namespace hps
{
  struct KParams
  {
    int    nc[3];     // cells along x, y (1 in 2D), march direction
    int    nn[3];     // nodes along x, y (1 in 2D), march direction
    int    nt[3];     // tiles along x, y, march
    int    tz;        // cells per tile along the march direction
    int    ctz;       // cells per cache tile along the march (0: tz); the
                      // z-column kernels march tz = a multiple of ctz
    int    order;     // number of history vectors (BDF order)
    double inv_h[3];  // 1/h along the physical directions x, y, z
    double jxw;       // prod(h) * (1/2)^DIM
    double w_hist[3]; // BDF weights[1..order]
    PhysParams pp;
    // uniform scales of the folded scalar tangent for the rescaled
    // butterflies (FoldScales, set by set_fold_scales; sa = (1+1/sqrt3)/2):
    // JxW/h^2 sa^(2 DIM - 2), times -kappa_c, times L kappa_p;
    // JxW/h sa^(2 DIM - 1) (signed per lane by the kernels); JxW sa^(2 DIM)
    double fold_kj[3], fold_kmu[3], fold_keta[3], fold_gs[3], fold_jxw;
    // grain cut-off: active order parameters per cell (bit g: eta_g), cells
    // x fastest; read only by the CUT kernels
    const unsigned *cmask;
  };

  // fills the uniform folded-tangent scales of kp from jxw, inv_h and pp
  inline void set_fold_scales(KParams &kp, int dim)
  {
    const double sa = 0.78867513459481288225457439025098;
    double       eg = 1.0;
    for (int i = 0; i < 2 * dim - 2; ++i)
      eg *= sa;
    for (int d = 0; d < 3; ++d)
      {
        kp.fold_kj[d]   = kp.jxw * kp.inv_h[d] * kp.inv_h[d] * eg;
        kp.fold_kmu[d]  = -kp.pp.kappa_c * kp.fold_kj[d];
        kp.fold_keta[d] = kp.pp.L * kp.pp.kappa_p * kp.fold_kj[d];
        kp.fold_gs[d]   = kp.jxw * kp.inv_h[d] * (eg * sa);
      }
    kp.fold_jxw = kp.jxw * (eg * sa * sa);
  }

  enum Op
  {
    OP_VMULT = 0,
    OP_RES0  = 1, // residual, no time terms
    OP_RES1  = 2, // residual, primary weight only
    OP_RES2  = 3, // residual with BDF history
  };

  // ---------------------------------------------------------------- lanes
  template <int MASK>
  __device__ __forceinline__ int xor_lane(int v);
  template <>
  __device__ __forceinline__ int xor_lane<1>(int v)
  {
    return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false); // quad_perm [1,0,3,2]
  }
  template <>
  __device__ __forceinline__ int xor_lane<2>(int v)
  {
    return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false); // quad_perm [2,3,0,1]
  }
  template <>
  __device__ __forceinline__ int xor_lane<4>(int v)
  {
    return __builtin_amdgcn_ds_swizzle(v, 0x101F); // bitmask mode, xor 4
  }

  template <int MASK>
  __device__ __forceinline__ double xshfl(double v)
  {
    const int lo = __double2loint(v);
    const int hi = __double2hiint(v);
    return __hiloint2double(xor_lane<MASK>(hi), xor_lane<MASK>(lo));
  }
  template <int MASK>
  __device__ __forceinline__ float xshfl(float v)
  {
    return __int_as_float(xor_lane<MASK>(__float_as_int(v)));
  }

  template <typename T>
  struct Q1Coef
  {
    static constexpr T sa = T(0.78867513459481288225457439025098);
    static constexpr T sb = T(0.21132486540518711774542560974902);
  };

  // signed scale of the lane's derivative along d: +1/h_d if the lane's
  // corner bit d is set, -1/h_d otherwise (times an optional factor)
  template <int DIM, typename T>
  __device__ __forceinline__ void signed_scales(int corner, const T *inv_h, T f, T *out)
  {
#pragma unroll
    for (int d = 0; d < DIM; ++d)
      out[d] = ((corner >> d) & 1) ? f * inv_h[d] : -(f * inv_h[d]);
  }

  // values and gradients at the lane's q-point from the lane's nodal value;
  // gsc: signed_scales(corner, inv_h, 1).  The derivative's sign is applied
  // with the scale at the end (exact: partners along the other directions
  // share it).
  template <int DIM, typename T>
  __device__ __forceinline__ void interp_vg(T u, const T *gsc, T &val, T *grad)
  {
    const T sa = Q1Coef<T>::sa, sb = Q1Coef<T>::sb;
    if constexpr (DIM == 2)
      {
        const T p  = xshfl<1>(u);
        const T A  = sa * u + sb * p;
        const T Dx = u - p;
        const T pA = xshfl<2>(A);
        const T pD = xshfl<2>(Dx);
        val     = sa * A + sb * pA;
        grad[1] = (A - pA) * gsc[1];
        grad[0] = (sa * Dx + sb * pD) * gsc[0];
      }
    else
      {
        const T p    = xshfl<4>(u);
        const T A    = sa * u + sb * p;
        const T Dz   = u - p;
        const T pA   = xshfl<2>(A);
        const T pDz  = xshfl<2>(Dz);
        const T AA   = sa * A + sb * pA;
        const T Ay   = A - pA;
        const T DzA  = sa * Dz + sb * pDz;
        const T pAA  = xshfl<1>(AA);
        const T pAy  = xshfl<1>(Ay);
        const T pDzA = xshfl<1>(DzA);
        val          = sa * AA + sb * pAA;
        grad[0]      = (AA - pAA) * gsc[0];
        grad[1]      = (sa * Ay + sb * pAy) * gsc[1];
        grad[2]      = (sa * DzA + sb * pDzA) * gsc[2];
      }
  }

  // interp_vg without the per-lane signed 1/h of the gradients
  template <int DIM, typename T>
  __device__ __forceinline__ void interp_vg_pre(T u, T &val, T *grad)
  {
    const T one[3] = {T(1), T(1), T(1)};
    interp_vg<DIM>(u, one, val, grad);
  }

  template <int DIM, typename T>
  __device__ __forceinline__ T interp_v(T u)
  {
    const T sa = Q1Coef<T>::sa, sb = Q1Coef<T>::sb;
    T       a  = sa * u + sb * xshfl<1>(u);
    a          = sa * a + sb * xshfl<2>(a);
    if constexpr (DIM == 3)
      a = sa * a + sb * xshfl<4>(a);
    return a;
  }

  // transpose of interp_vg: nodal contribution of (v, g) given at q-points,
  // g already multiplied by signed_scales(corner, inv_h, JxW): the partner
  // across direction d holds the opposite sign, so s_d (g + g_partner) is
  // the difference of the signed values.  Where a level combines a value
  // row and a derivative row, the partner's whole contribution
  // (sb * value - derivative) is exchanged as one quantity.
  template <int DIM, typename T>
  __device__ __forceinline__ T integrate_vg(T v, const T *g)
  {
    const T sa = Q1Coef<T>::sa, sb = Q1Coef<T>::sb;
    if constexpr (DIM == 2)
      {
        const T U1 = (sa * v + g[1]) + xshfl<2>(sb * v - g[1]);
        const T Ux = sa * g[0] + sb * xshfl<2>(g[0]);
        return (sa * U1 + Ux) + xshfl<1>(sb * U1 - Ux);
      }
    else
      {
        const T T1 = (sa * v + g[0]) + xshfl<1>(sb * v - g[0]);
        const T Ty = sa * g[1] + sb * xshfl<1>(g[1]);
        const T Tz = sa * g[2] + sb * xshfl<1>(g[2]);
        const T U1 = (sa * T1 + Ty) + xshfl<2>(sb * T1 - Ty);
        const T Uz = sa * Tz + sb * xshfl<2>(Tz);
        return (sa * U1 + Uz) + xshfl<4>(sb * U1 - Uz);
      }
  }

  // ----- rescaled butterflies (one FMA per value level)
  // With r = sb / sa = 2 - sqrt(3), a value level sa x + sb x_partner is
  // sa (x + r x_partner); the factor sa of each level is left to constants
  // the caller folds.
  //   interp_vg_hat: val = sa^DIM val_hat, grad_d = sa^(DIM-1) grad_hat_d
  //     (unsigned, unscaled derivatives as interp_vg_pre);
  //   integrate_vg_hat(sa^DIM v, sa^(DIM-1) g) = integrate_vg(v, g).
  template <typename T>
  struct Q1Hat
  {
    static constexpr T r = T(0.26794919243112270647255365849413);
  };

  template <int DIM, typename T>
  __device__ __forceinline__ void interp_vg_hat(T u, T &val, T *grad)
  {
    const T r = Q1Hat<T>::r;
    if constexpr (DIM == 2)
      {
        const T p  = xshfl<1>(u);
        const T A  = u + r * p;
        const T Dx = u - p;
        const T pA = xshfl<2>(A);
        const T pD = xshfl<2>(Dx);
        val        = A + r * pA;
        grad[1]    = A - pA;
        grad[0]    = Dx + r * pD;
      }
    else
      {
        const T p    = xshfl<4>(u);
        const T A    = u + r * p;
        const T Dz   = u - p;
        const T pA   = xshfl<2>(A);
        const T pDz  = xshfl<2>(Dz);
        const T AA   = A + r * pA;
        const T Ay   = A - pA;
        const T DzA  = Dz + r * pDz;
        const T pAA  = xshfl<1>(AA);
        const T pAy  = xshfl<1>(Ay);
        const T pDzA = xshfl<1>(DzA);
        val          = AA + r * pAA;
        grad[0]      = AA - pAA;
        grad[1]      = Ay + r * pAy;
        grad[2]      = DzA + r * pDzA;
      }
  }

  template <int DIM, typename T>
  __device__ __forceinline__ T integrate_vg_hat(T v, const T *g)
  {
    const T r = Q1Hat<T>::r;
    if constexpr (DIM == 2)
      {
        const T U1 = (v + g[1]) + xshfl<2>(r * v - g[1]);
        const T Ux = g[0] + r * xshfl<2>(g[0]);
        return (U1 + Ux) + xshfl<1>(r * U1 - Ux);
      }
    else
      {
        const T T1 = (v + g[0]) + xshfl<1>(r * v - g[0]);
        const T Ty = g[1] + r * xshfl<1>(g[1]);
        const T Tz = g[2] + r * xshfl<1>(g[2]);
        const T U1 = (T1 + Ty) + xshfl<2>(r * T1 - Ty);
        const T Uz = Tz + r * xshfl<2>(Tz);
        return (U1 + Uz) + xshfl<4>(r * U1 - Uz);
      }
  }

  // the value of interp_vg alone (same levels in the same order: the same
  // rounding)
  template <int DIM, typename T>
  __device__ __forceinline__ T interp_v_vg(T u)
  {
    const T sa = Q1Coef<T>::sa, sb = Q1Coef<T>::sb;
    if constexpr (DIM == 2)
      {
        const T A = sa * u + sb * xshfl<1>(u);
        return sa * A + sb * xshfl<2>(A);
      }
    else
      {
        const T A  = sa * u + sb * xshfl<4>(u);
        const T AA = sa * A + sb * xshfl<2>(A);
        return sa * AA + sb * xshfl<1>(AA);
      }
  }

  // the value of interp_vg_hat alone (same levels in the same order: the
  // same rounding)
  template <int DIM, typename T>
  __device__ __forceinline__ T interp_v_hat(T u)
  {
    const T r = Q1Hat<T>::r;
    if constexpr (DIM == 2)
      {
        const T A = u + r * xshfl<1>(u);
        return A + r * xshfl<2>(A);
      }
    else
      {
        const T A  = u + r * xshfl<4>(u);
        const T AA = A + r * xshfl<2>(A);
        return AA + r * xshfl<1>(AA);
      }
  }

  // FoldScales of the scalar-mobility tangent for the rescaled butterflies
  // (the scales of KParams, the mobility-derivative one signed by the lane's
  // corner) and the value-row factor jxw_hat
  template <int DIM, typename T>
  __device__ __forceinline__ void fold_scales_hat(const KParams &kp, int corner,
                                                  FoldScales<T> &fs, T &jxw_hat)
  {
#pragma unroll
    for (int d = 0; d < DIM; ++d)
      {
        fs.kj[d]   = T(kp.fold_kj[d]);
        fs.kmu[d]  = T(kp.fold_kmu[d]);
        fs.keta[d] = T(kp.fold_keta[d]);
        fs.gsj[d]  = ((corner >> d) & 1) ? T(kp.fold_gs[d]) : -T(kp.fold_gs[d]);
      }
    jxw_hat = T(kp.fold_jxw);
  }

  // sum over the 2^DIM lanes of a cell
  template <int DIM, typename T>
  __device__ __forceinline__ T cell_sum(T v)
  {
    v += xshfl<1>(v);
    v += xshfl<2>(v);
    if constexpr (DIM == 3)
      v += xshfl<4>(v);
    return v;
  }

  // ------------------------------------------------------------ the tile
  template <int DIM, int TX, int TY>
  struct Tile
  {
    static constexpr int NV = 1 << DIM;
    static constexpr int NT = TX * TY * NV;
    static constexpr int PX = TX + 1;
    static constexpr int PY = DIM == 3 ? TY + 1 : 1;
    static constexpr int PN = PX * PY;

    int  tid, corner, cx, cy, bx, by, bm;
    int  x0, y0, m0, tx_eff, ty_eff, tm_eff;
    int  pn_lane; // plane-node index of the lane's node
    bool cell_valid;

    int tile; // index of the tile (workgroup of the full launch)

    __device__ __forceinline__ Tile(const KParams &kp) : Tile(kp, (int)blockIdx.x) {}

    __device__ __forceinline__ Tile(const KParams &kp, int tile_index)
    {
      tile       = tile_index;
      tid        = threadIdx.x;
      corner     = tid & (NV - 1);
      const int cell = tid >> DIM;
      cx         = cell % TX;
      cy         = cell / TX;
      bx         = corner & 1;
      by         = DIM == 3 ? (corner >> 1) & 1 : 0;
      bm         = corner >> (DIM - 1);
      int b      = tile_index;
      const int tix = b % kp.nt[0];
      b /= kp.nt[0];
      const int tiy = b % kp.nt[1];
      const int tim = b / kp.nt[1];
      x0         = tix * TX;
      y0         = tiy * TY;
      m0         = tim * kp.tz;
      tx_eff     = min(TX, kp.nc[0] - x0);
      ty_eff     = DIM == 3 ? min(TY, kp.nc[1] - y0) : 1;
      tm_eff     = min(kp.tz, kp.nc[2] - m0);
      cell_valid = cx < tx_eff && cy < ty_eff;
      pn_lane    = (cx + bx) + PX * (cy + by);
    }

    // the lane's cell in layer k (cells x fastest, as the cut-off masks)
    __device__ __forceinline__ int64_t cell_index(const KParams &kp, int k) const
    {
      return ((int64_t)(m0 + k) * kp.nc[1] + (y0 + cy)) * kp.nc[0] + (x0 + cx);
    }

    // the cut-off mask of the lane's cell in layer k (all ones without CUT)
    template <bool CUT>
    __device__ __forceinline__ unsigned cut_mask(const KParams &kp, int k) const
    {
      if constexpr (CUT)
        return cell_valid ? kp.cmask[cell_index(kp, k)] : 0u;
      else
        return ~0u;
    }

    __device__ __forceinline__ bool node_in_tile(int nx, int ny) const
    {
      return nx <= tx_eff && (DIM == 2 || ny <= ty_eff);
    }

    __device__ __forceinline__ int64_t gnode(const KParams &kp, int nx, int ny, int k) const
    {
      return ((int64_t)(m0 + k) * kp.nn[1] + (y0 + ny)) * kp.nn[0] + (x0 + nx);
    }

    // node shared with another tile (its sum is finished by q1_fixup)
    __device__ __forceinline__ bool shared(const KParams &kp, int nx, int ny, int k) const
    {
      bool s = (nx == 0 && x0 > 0) || (nx == tx_eff && x0 + tx_eff < kp.nc[0]) ||
               (k == 0 && m0 > 0) || (k == tm_eff && m0 + tm_eff < kp.nc[2]);
      if constexpr (DIM == 3)
        s = s || (ny == 0 && y0 > 0) || (ny == ty_eff && y0 + ty_eff < kp.nc[1]);
      return s;
    }
  };

  // output of one node/component of plane k: the node's own entry of dst or
  // this tile's partial
  template <int DIM, int TX, int TY, typename T>
  __device__ __forceinline__ void put(const KParams &kp, const Tile<DIM, TX, TY> &t, int nc,
                                      int k, int pn, int c, T v, T *dst, T *part,
                                      int ostride, int ocstride, int ooffset)
  {
    using TL     = Tile<DIM, TX, TY>;
    const int nx = pn % TL::PX, ny = pn / TL::PX;
    if (!t.node_in_tile(nx, ny))
      return;
    if (t.shared(kp, nx, ny, k))
      part[(((int64_t)t.tile * (kp.tz + 1) + k) * TL::PN + pn) * nc + c] = v;
    else
      dst[t.gnode(kp, nx, ny, k) * ostride + (int64_t)c * ocstride + ooffset] = v;
  }

  // ------------------------------------------- vmult / residual kernel
  // waves per SIMD the register allocation of q1_apply is asked to allow:
  // 3 (at most 168 registers) for the FP64 scalar tangent with up to 4
  // order parameters, whose allocation is close to that limit
  template <int G, bool TENS, int OP, typename T>
  constexpr int q1_apply_waves()
  {
    return (sizeof(T) == 8 && !TENS && OP == OP_VMULT && G <= 4) ? 3 : 1;
  }

  // CUT: grain cut-off (do_vmult_range_no_template and
  // do_evaluate_nonlinear_residual_nt, S/operator_sintering_generic.h:
  // 456-537, 1008-1081): in every cell the order parameters outside the
  // cell's mask are dropped -- their values and gradients do not enter the
  // physics (the cut cache holds zeros for their linearisation values) and
  // their rows get nothing from the cell
  template <int DIM, int TX, int TY, int G, bool TENS, int OP, typename T, bool CUT = false>
  __global__ void __launch_bounds__(TX *TY *(1 << DIM))
    __attribute__((amdgpu_waves_per_eu(q1_apply_waves<G, TENS, OP, T>())))
    q1_apply(const KParams kp, const T *__restrict__ src, const T *__restrict__ old0,
             const T *__restrict__ old1, const T *__restrict__ old2,
             const T *__restrict__ cache, T *__restrict__ dst, T *__restrict__ part)
  {
    using TL             = Tile<DIM, TX, TY>;
    using CL             = CacheLayout<DIM, G, TENS>;
    constexpr int  NC    = G + 2;
    constexpr int  NT    = TL::NT;
    constexpr int  NV    = TL::NV;
    constexpr int  PN    = TL::PN;
    constexpr bool HIST  = OP == OP_RES2;
    constexpr int  NPL   = HIST ? 2 : 1;
    constexpr int  ITEMS = (PN * NC + NT - 1) / NT;
    constexpr int  NF    = CL::n_fields;

    __shared__ T ring[2][NPL][PN * NC];
    __shared__ T slots[NV][PN * NC];

    const TL t(kp);
    const int tid = t.tid;
    // item loops of the 3D scalar tangent with up to 4 order parameters
    // (plane loads and stores, node sums): thread ids rotated by a wave per
    // workgroup, so the wave that takes the items beyond NT differs from
    // one workgroup to the next (elsewhere the plain order: there the
    // rotation costs an occupancy step)
    const int itid = (DIM == 3 && !TENS && OP == OP_VMULT && G <= 4)
                       ? (tid + 64 * (int)(blockIdx.x & 3)) & (NT - 1)
                       : tid;

    const T inv_h[3] = {T(kp.inv_h[0]), T(kp.inv_h[1]), T(kp.inv_h[2])};
    T       gsc[DIM], gsj[DIM];
    signed_scales<DIM>(t.corner, inv_h, T(1), gsc);
    signed_scales<DIM>(t.corner, inv_h, T(kp.jxw), gsj);
    const T jxw = T(kp.jxw);
    // scalar-mobility tangent: constant gradient rows folded into uniform
    // scales (FoldScales)
    // (rescaled butterflies: interp_vg_hat / integrate_vg_hat)
    constexpr bool FOLD = !TENS && OP == OP_VMULT;
    FoldScales<T>  fs;
    T              jxw_hat = T(0);
    if constexpr (FOLD)
      fold_scales_hat<DIM>(kp, t.corner, fs, jxw_hat);

    T regs[NPL][ITEMS];
    auto load_plane = [&](int k) {
#pragma unroll
      for (int r = 0; r < ITEMS; ++r)
        {
          const int i = itid + r * NT;
          if (i < PN * NC)
            {
              const int pn = i / NC, c = i - pn * NC;
              const int nx = pn % TL::PX, ny = pn / TL::PX;
              T         s = T(0), h = T(0);
              if (t.node_in_tile(nx, ny))
                {
                  const int64_t gi = t.gnode(kp, nx, ny, k) * NC + c;
                  s                = src[gi];
                  if constexpr (HIST)
                    {
                      // do_evalute_history, S/operator_sintering_generic.h:950-960
                      h = old0[gi] * T(kp.w_hist[0]);
                      if (kp.order >= 2)
                        h += old1[gi] * T(kp.w_hist[1]);
                      if (kp.order >= 3)
                        h += old2[gi] * T(kp.w_hist[2]);
                    }
                }
              regs[0][r] = s;
              if constexpr (HIST)
                regs[1][r] = h;
            }
        }
    };
    auto store_plane = [&](int slot) {
#pragma unroll
      for (int r = 0; r < ITEMS; ++r)
        {
          const int i = itid + r * NT;
          if (i < PN * NC)
            {
#pragma unroll
              for (int a = 0; a < NPL; ++a)
                ring[slot][a][i] = regs[a][r];
            }
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

    // uniform base of this tile's cache + a 32-bit per-lane offset
    const T *cache_tile = cache + (int64_t)blockIdx.x * kp.tz * NF * NT;
    // PF (3D, up to 4 order parameters: where the registers it holds do
    // not lower the occupancy): the next layer's cache fields and the plane
    // after next are requested together at the end of a layer, so that the
    // layer's waits do not cover them and both arrive during the node sums
    // and the next interpolation
    constexpr bool PF = FOLD && DIM == 3 && G <= 4;
    // TWO (FP64, 3D, 5 or more order parameters): values first, each
    // gradient interpolated again where its row is integrated, so the
    // gradients of all components are never live together (8 order
    // parameters: 246 -> 167 VGPRs, 2 -> 3 waves/SIMD; the FP32 kernels and
    // those with fewer order parameters are faster without it)
    constexpr bool TWO = FOLD && DIM == 3 && G >= 5 && sizeof(T) == 8;
    // the same for the scalar-mobility residual (its gradient rows are
    // M grad mu, -kappa_c grad c and L kappa_p grad eta_i)
    constexpr bool TWO_R = !TENS && OP != OP_VMULT && DIM == 3 && G >= 5 && sizeof(T) == 8;
    T              lin_next[PF ? NF : 1];
    if constexpr (PF)
      {
        if (2 <= t.tm_eff)
          load_plane(2);
#pragma unroll
        for (int f = 0; f < NF; ++f)
          lin_next[f] = cache_tile[f * NT + tid];
      }

    for (int k = 0; k < t.tm_eff; ++k)
      {
        const bool pre = k + 2 <= t.tm_eff;
        if (!PF && pre)
          load_plane(k + 2);

        // ---- cell layer k: interpolate, physics, integrate
        // TWO without the cut-off: the layer's cache fields are requested
        // before the interpolation (the barrier below keeps the physics
        // after it; the cut-off kernel with 8 order parameters would lose
        // its third wave per SIMD)
        constexpr bool TOP = TWO && !CUT;
        T              lin_top[TOP ? NF : 1];
        if constexpr (TOP)
#pragma unroll
          for (int f = 0; f < NF; ++f)
            lin_top[f] = cache_tile[k * (NF * NT) + f * NT + tid];
        const T *pl = &ring[(k + t.bm) & 1][0][t.pn_lane * NC];
        T        val[NC], grad[NC][DIM];
#pragma unroll
        for (int c = 0; c < NC; ++c)
          if constexpr (TWO)
            val[c] = interp_v_hat<DIM>(pl[c]);
          else if constexpr (TWO_R)
            val[c] = interp_v_vg<DIM>(pl[c]);
          else if constexpr (FOLD)
            interp_vg_hat<DIM>(pl[c], val[c], grad[c]);
          else
            interp_vg<DIM>(pl[c], gsc, val[c], grad[c]);
        // cut-off: the dropped order parameters of the lane's cell enter
        // the physics as absent (zero)
        const unsigned cm = t.template cut_mask<CUT>(kp, k);
        if constexpr (CUT)
#pragma unroll
          for (int c = 2; c < NC; ++c)
            {
              const bool on = (cm >> (c - 2)) & 1u;
              val[c]        = on ? val[c] : T(0);
              if constexpr (!TWO && !TWO_R)
#pragma unroll
                for (int d = 0; d < DIM; ++d)
                  grad[c][d] = on ? grad[c][d] : T(0);
            }
        // the physics stays below the interpolation (with PF it waits for
        // lin_next; in 3D it also shortens the register live ranges)
        if constexpr (FOLD && DIM == 3)
          __builtin_amdgcn_sched_barrier(0);

        T vr[NC], gr[NC][DIM];
        T *sl = &slots[t.corner][t.pn_lane * NC];
        if constexpr (FOLD)
          {
            // value rows first, then each component's gradient row is
            // formed and integrated at once (shorter live ranges)
            T lin[NF];
            if constexpr (PF)
#pragma unroll
              for (int f = 0; f < NF; ++f)
                lin[f] = lin_next[f];
            else if constexpr (TOP)
#pragma unroll
              for (int f = 0; f < NF; ++f)
                lin[f] = lin_top[f];
            else
#pragma unroll
              for (int f = 0; f < NF; ++f)
                lin[f] = cache_tile[k * (NF * NT) + f * NT + tid];
            tangent_qp_rows<DIM, G, false, true, false>(kp.pp, lin, val, grad, vr, gr);
            T M, s;
            scalar_mobility_terms<G>(kp.pp, lin[CL::c], lin + CL::eta, val, M, s);
            T lmu[DIM];
#pragma unroll
            for (int d = 0; d < DIM; ++d)
              lmu[d] = lin[CL::grad_mu + d];
#pragma unroll
            for (int c = 0; c < NC; ++c)
              {
                // TWO: the gradient this row needs is interpolated here
                // (grad mu for the c row, grad c for the mu row)
                T g[DIM];
                if constexpr (TWO)
                  {
                    T v_unused;
                    interp_vg_hat<DIM>(pl[c == 0 ? 1 : (c == 1 ? 0 : c)], v_unused, g);
                  }
                else
#pragma unroll
                  for (int d = 0; d < DIM; ++d)
                    g[d] = grad[c == 0 ? 1 : (c == 1 ? 0 : c)][d];
                T gs[DIM];
#pragma unroll
                for (int d = 0; d < DIM; ++d)
                  gs[d] = c == 0 ? (M * fs.kj[d]) * g[d] + s * (lmu[d] * fs.gsj[d])
                                 : (c == 1 ? fs.kmu[d] * g[d] : fs.keta[d] * g[d]);
                sl[c] = integrate_vg_hat<DIM>(vr[c] * jxw_hat, gs);
                if constexpr (CUT)
                  if (c >= 2 && !((cm >> (c - 2)) & 1u))
                    sl[c] = T(0);
              }
          }
        else if constexpr (OP == OP_VMULT)
          {
            T lin[NF];
            const T *cl = cache_tile + k * (NF * NT);
#pragma unroll
            for (int f = 0; f < NF; ++f)
              lin[f] = cl[f * NT + tid];
            tangent_qp<DIM, G, TENS>(kp.pp, lin, val, grad, vr, gr);
          }
        else
          {
            T hist[NC];
            if constexpr (HIST)
              {
                const T *ph = &ring[(k + t.bm) & 1][1][t.pn_lane * NC];
#pragma unroll
                for (int c = 0; c < NC; ++c)
                  hist[c] = interp_v<DIM>(ph[c]);
              }
            if constexpr (TWO_R)
              residual_values<DIM, G, TENS, OP - OP_RES0>(kp.pp, val, hist, vr);
            else
              residual_qp<DIM, G, TENS, OP - OP_RES0>(kp.pp, val, grad, hist, vr, gr);
          }

        if constexpr (!FOLD)
#pragma unroll
          for (int c = 0; c < NC; ++c)
            {
              T gs[DIM];
              if constexpr (TWO_R)
                {
                  // this row's gradient input, interpolated here
                  T v_unused, g[DIM], row[DIM];
                  interp_vg<DIM>(pl[c == 0 ? 1 : (c == 1 ? 0 : c)], gsc, v_unused, g);
                  if (c == 0)
                    residual_grad_c_scalar<DIM, G>(kp.pp, val, g, row);
                  else
                    {
                      const T f = c == 1 ? -T(kp.pp.kappa_c) : T(kp.pp.L) * T(kp.pp.kappa_p);
#pragma unroll
                      for (int d = 0; d < DIM; ++d)
                        row[d] = f * g[d];
                    }
#pragma unroll
                  for (int d = 0; d < DIM; ++d)
                    gs[d] = row[d] * gsj[d];
                }
              else
#pragma unroll
                for (int d = 0; d < DIM; ++d)
                  gs[d] = gr[c][d] * gsj[d];
              sl[c] = integrate_vg<DIM>(vr[c] * jxw, gs);
              if constexpr (CUT)
                if (c >= 2 && !((cm >> (c - 2)) & 1u))
                  sl[c] = T(0);
            }
        __syncthreads();

        if (pre)
          store_plane(k & 1);

        if constexpr (PF)
          {
            if (k + 3 <= t.tm_eff)
              load_plane(k + 3);
            if (k + 1 < t.tm_eff)
              {
                const T *cn = cache_tile + (k + 1) * (NF * NT);
#pragma unroll
                for (int f = 0; f < NF; ++f)
                  lin_next[f] = cn[f * NT + tid];
              }
          }

        // ---- node sums: plane k is complete, plane k+1 begins
#pragma unroll
        for (int r = 0; r < ITEMS; ++r)
          {
            const int i = itid + r * NT;
            if (i < PN * NC)
              {
                const int pn = i / NC, c = i - pn * NC;
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
                put(kp, t, NC, k, pn, c, fin, dst, part, NC, 1, 0);
              }
          }
        __syncthreads();
      }

#pragma unroll
    for (int r = 0; r < ITEMS; ++r)
      {
        const int i = itid + r * NT;
        if (i < PN * NC)
          {
            const int pn = i / NC, c = i - pn * NC;
            put(kp, t, NC, t.tm_eff, pn, c, carry[r], dst, part, NC, 1, 0);
          }
      }
  }

  // ------------------------------------------- linearisation cache fill
  // SinteringOperatorData::fill_quadrature_point_values,
  // S/operator_sintering_data.h:301-409 (only the fields the tangent reads)
  // CUT: the linearisation values (and, tensorial, gradients) of the order
  // parameters dropped in a cell are stored as zeros, as the compacted cache
  // of the reference leaves them out (S/operator_sintering_data.h:360-395)
  template <int DIM, int TX, int TY, int G, bool TENS, bool CUT = false>
  __global__ void __launch_bounds__(TX *TY *(1 << DIM))
    q1_fill(const KParams kp, const double *__restrict__ lin, double *__restrict__ cache64,
            float *__restrict__ cache32)
  {
    using TL            = Tile<DIM, TX, TY>;
    using CL            = CacheLayout<DIM, G, TENS>;
    constexpr int NC    = G + 2;
    constexpr int NT    = TL::NT;
    constexpr int PN    = TL::PN;
    constexpr int ITEMS = (PN * NC + NT - 1) / NT;
    constexpr int NF    = CL::n_fields;

    __shared__ double ring[2][PN * NC];
    const TL          t(kp);
    const int         tid      = t.tid;
    const double      inv_h[3] = {kp.inv_h[0], kp.inv_h[1], kp.inv_h[2]};
    double            gsc[DIM];
    signed_scales<DIM>(t.corner, inv_h, 1.0, gsc);

    double regs[ITEMS];
    auto   load_plane = [&](int k) {
#pragma unroll
      for (int r = 0; r < ITEMS; ++r)
        {
          const int i = tid + r * NT;
          if (i < PN * NC)
            {
              const int pn = i / NC, c = i - pn * NC;
              const int nx = pn % TL::PX, ny = pn / TL::PX;
              regs[r] = t.node_in_tile(nx, ny) ? lin[t.gnode(kp, nx, ny, k) * NC + c] : 0.0;
            }
        }
    };
    auto store_plane = [&](int slot) {
#pragma unroll
      for (int r = 0; r < ITEMS; ++r)
        {
          const int i = tid + r * NT;
          if (i < PN * NC)
            ring[slot][i] = regs[r];
        }
    };

    load_plane(0);
    store_plane(0);
    load_plane(1);
    store_plane(1);
    __syncthreads();

    const int64_t tile_off = (int64_t)blockIdx.x * kp.tz * NF * NT;
    for (int k = 0; k < t.tm_eff; ++k)
      {
        const bool pre = k + 2 <= t.tm_eff;
        if (pre)
          load_plane(k + 2);

        const double  *pl = &ring[(k + t.bm) & 1][t.pn_lane * NC];
        const unsigned cm = t.template cut_mask<CUT>(kp, k);
        double         f[NF];
        if constexpr (!TENS)
          {
            f[CL::c] = interp_v<DIM>(pl[0]);
#pragma unroll
            for (int i = 0; i < G; ++i)
              {
                f[CL::eta + i] = interp_v<DIM>(pl[2 + i]);
                if constexpr (CUT)
                  f[CL::eta + i] = ((cm >> i) & 1u) ? f[CL::eta + i] : 0.0;
              }
            double v, g[DIM];
            interp_vg<DIM>(pl[1], gsc, v, g);
#pragma unroll
            for (int d = 0; d < DIM; ++d)
              f[CL::grad_mu + d] = g[d];
          }
        else
          {
            double v, g[DIM];
            interp_vg<DIM>(pl[0], gsc, v, g);
            f[CL::c] = v;
#pragma unroll
            for (int d = 0; d < DIM; ++d)
              f[CL::grad_c + d] = g[d];
            interp_vg<DIM>(pl[1], gsc, v, g);
#pragma unroll
            for (int d = 0; d < DIM; ++d)
              f[CL::grad_mu + d] = g[d];
            double leg[G][DIM];
#pragma unroll
            for (int i = 0; i < G; ++i)
              {
                interp_vg<DIM>(pl[2 + i], gsc, v, leg[i]);
                f[CL::eta + i] = v;
                if constexpr (CUT)
                  if (!((cm >> i) & 1u))
                    {
                      f[CL::eta + i] = 0.0;
#pragma unroll
                      for (int d = 0; d < DIM; ++d)
                        leg[i][d] = 0.0;
                    }
              }
            gb_fields<DIM, G>(f + CL::eta, leg, f + CL::grad_mu, f[CL::gb_s], f + CL::gb_p,
                              f + CL::gb_q);
          }
        const int64_t o = tile_off + (int64_t)k * NF * NT + tid;
#pragma unroll
        for (int q = 0; q < NF; ++q)
          {
            cache64[o + (int64_t)q * NT] = f[q];
            if (cache32)
              cache32[o + (int64_t)q * NT] = (float)f[q];
          }
        __syncthreads();
        if (pre)
          store_plane(k & 1);
        __syncthreads();
      }
  }

  // --------------------------------- diagonal / block column by probing
  // Matrix-free assembly of the tangent's diagonal (column < 0) or of column
  // `column` of every node's n_c x n_c block, by applying the quadrature
  // operator to unit vectors of the cell's local dofs -- the recipe of
  // MatrixFreeTools::compute_diagonal (used by OperatorBase::
  // compute_inverse_diagonal, S/operator_base.h:364-389).
  // T = float assembles from the FP32 cache in FP32 arithmetic (the
  // preconditioner of the FP32 inner solve)
  // Node-block column (or diagonal) entries of this lane's corner node in
  // its cell: B[c][a][b] = sum_q JxW psi_c,m'(q) M_{a m', b m}(q) psi_c,m(q)
  // with M the quadrature-point tangent from input (component b, value or
  // gradient m) to output (component a, value or gradient m').  M is
  // sparse: value rows read values only; the c gradient row reads the
  // values and the gradients of c and mu; the mu and eta_i gradient rows
  // are -kappa_c grad c and L kappa_p grad eta_i.  Each nonzero term is a
  // tensor product of 1D weights (s s, s s', s' s' of the node's basis at
  // the Gauss points), so the sum over the cell's q-points is one
  // butterfly per direction over the fields
  //   r < n_c      mass        M_{r 0, b 0}                   -> row r
  //   n_c + d      s' s  (d)   M_{0 d+1, b 0}                 -> row 0
  //   n_c + 3 + d  s' s' (d)   M_{0 d+1, b d+1}               -> row 0
  //   n_c + 6 + p  s' s' (d,e) M_{0 d+1, b e+1} + M_{0 e+1, b d+1} -> row 0
  //   n_c + 9 + d  s' s' (d)   -kappa_c (b = 0, row 1), L kappa_p (b >= 2)
  // (diagonal, column < 0: mass row b from input b, the c-row fields from
  // input c, the stiffness fields 1, scaled by L kappa_p for b >= 2)
  template <int DIM, int G, bool TENS, typename T>
  __device__ __forceinline__ void q1_probe_cell(const KParams &kp, int corner, const T *lin,
                                                int column, const T *inv_h, T *out)
  {
    constexpr int NC = G + 2;
    constexpr int NR = NC + 12;
    const T       sa = Q1Coef<T>::sa, sb = Q1Coef<T>::sb;
    const T       kc = -T(kp.pp.kappa_c), lkp = T(kp.pp.L) * T(kp.pp.kappa_p);
    const int     b  = column;
    T             F[NR];
    const T       zero[3] = {T(0), T(0), T(0)};
    {
      // c-row gradient outputs for the inputs of component cb
      const int cb = b >= 0 ? b : 0;
      T         v[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c)
        v[c] = c == cb ? T(1) : T(0);
      T gm[DIM + 1][DIM];
      tangent_grad_c<DIM, G, TENS>(kp.pp, lin, v, zero, zero, gm[0]);
#pragma unroll
      for (int c = 0; c < NC; ++c)
        v[c] = T(0);
#pragma unroll
      for (int d = 0; d < DIM; ++d)
        {
          T u[3] = {T(0), T(0), T(0)};
          u[d]   = T(1);
          if (cb <= 1)
            tangent_grad_c<DIM, G, TENS>(kp.pp, lin, v, cb == 0 ? u : zero, cb == 1 ? u : zero,
                                         gm[d + 1]);
          else
#pragma unroll
            for (int e = 0; e < DIM; ++e)
              gm[d + 1][e] = T(0);
        }
#pragma unroll
      for (int d = 0; d < 3; ++d)
        {
          F[NC + d]     = d < DIM ? gm[0][d] : T(0);
          F[NC + 3 + d] = d < DIM ? gm[d + 1][d] : T(0);
        }
      F[NC + 6] = gm[2][0] + gm[1][1];
      F[NC + 7] = DIM == 3 ? gm[3 % (DIM + 1)][0] + gm[1][2 % DIM] : T(0);
      F[NC + 8] = DIM == 3 ? gm[3 % (DIM + 1)][1] + gm[2][2 % DIM] : T(0);
    }
    if (b >= 0)
      {
        T v[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c)
          v[c] = c == b ? T(1) : T(0);
        tangent_values<DIM, G, TENS>(kp.pp, lin, v, F);
        const T st = b == 0 ? kc : (b >= 2 ? lkp : T(0));
#pragma unroll
        for (int d = 0; d < 3; ++d)
          F[NC + 9 + d] = d < DIM ? st : T(0);
      }
    else
      {
#pragma unroll
        for (int cb = 0; cb < NC; ++cb)
          {
            T v[NC], vr[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c)
              v[c] = c == cb ? T(1) : T(0);
            tangent_values<DIM, G, TENS>(kp.pp, lin, v, vr);
            F[cb] = vr[cb];
          }
#pragma unroll
        for (int d = 0; d < 3; ++d)
          F[NC + 9 + d] = d < DIM ? T(1) : T(0);
      }
    // q -> node: one butterfly per direction; the lane's bit d stands for
    // the Gauss point before and for the node after level d
#pragma unroll
    for (int d = 0; d < DIM; ++d)
      {
        const bool hi = (corner >> d) & 1;
        const T    e  = hi ? inv_h[d] : -inv_h[d]; // s' of the node's basis
#pragma unroll
        for (int r = 0; r < NR; ++r)
          {
            // weight type of field r along d: 0 s s, 1 s' s, 2 s' s'
            int ty = 0;
            if (r >= NC)
              {
                const int k = r - NC;
                if (k < 3)
                  ty = k == d ? 1 : 0;
                else if (k < 6 || k >= 9)
                  ty = (k % 3) == d ? 2 : 0;
                else
                  {
                    const int p = k - 6, e0 = p == 2 ? 1 : 0, e1 = p == 0 ? 1 : 2;
                    ty = (d == e0 || d == e1) ? 1 : 0;
                  }
              }
            const T own = F[r];
            const T par = d == 0 ? xshfl<1>(own) : (d == 1 ? xshfl<2>(own) : xshfl<4>(own));
            if (ty == 0)
              F[r] = (sa * sa) * own + (sb * sb) * par;
            else if (ty == 1)
              F[r] = e * (sa * own + sb * par);
            else
              F[r] = (e * e) * (own + par);
          }
      }
    const T jxw = T(kp.jxw);
    T       c_sum = T(0), st_sum = T(0);
#pragma unroll
    for (int r = NC; r < NC + 9; ++r)
      c_sum += F[r];
#pragma unroll
    for (int r = NC + 9; r < NR; ++r)
      st_sum += F[r];
#pragma unroll
    for (int a = 0; a < NC; ++a)
      {
        T x = F[a];
        if (a == 0)
          x += c_sum;
        if (b >= 0)
          {
            if (b != 1 && a == (b == 0 ? 1 : b))
              x += st_sum;
          }
        else if (a >= 2)
          x += lkp * st_sum;
        out[a] = jxw * x;
      }
  }

  // CUT: entries of rows or columns of order parameters dropped in a cell
  // get nothing from that cell
  template <int DIM, int TX, int TY, int G, bool TENS, typename T, bool CUT = false>
  __global__ void __launch_bounds__(TX *TY *(1 << DIM))
    q1_probe(const KParams kp, const T *__restrict__ cache, T *__restrict__ dst,
             T *__restrict__ part, int column, int ostride, int ocstride, int ooffset)
  {
    using TL            = Tile<DIM, TX, TY>;
    using CL            = CacheLayout<DIM, G, TENS>;
    constexpr int NC    = G + 2;
    constexpr int NT    = TL::NT;
    constexpr int NV    = TL::NV;
    constexpr int PN    = TL::PN;
    constexpr int ITEMS = (PN * NC + NT - 1) / NT;
    constexpr int NF    = CL::n_fields;

    __shared__ T slots[NV][PN * NC];
    const TL     t(kp);
    const int    tid = t.tid;
    const T      inv_h[3] = {T(kp.inv_h[0]), T(kp.inv_h[1]), T(kp.inv_h[2])};

    T carry[ITEMS];
#pragma unroll
    for (int r = 0; r < ITEMS; ++r)
      carry[r] = T(0);

    const T *cache_tile = cache + (int64_t)blockIdx.x * kp.tz * NF * NT;
    for (int k = 0; k < t.tm_eff; ++k)
      {
        T lin[NF];
        const T *cl = cache_tile + (int64_t)k * NF * NT + tid;
#pragma unroll
        for (int f = 0; f < NF; ++f)
          lin[f] = cl[f * NT];

        T mine[NC];
        q1_probe_cell<DIM, G, TENS>(kp, t.corner, lin, column, inv_h, mine);
        if constexpr (CUT)
          {
            const unsigned cm   = t.template cut_mask<CUT>(kp, k);
            const bool     coff = column >= 2 && !((cm >> (column - 2)) & 1u);
#pragma unroll
            for (int c = 2; c < NC; ++c)
              if (coff || !((cm >> (c - 2)) & 1u))
                mine[c] = T(0);
            if (coff)
              mine[0] = mine[1] = T(0);
          }
        T *sl = &slots[t.corner][t.pn_lane * NC];
#pragma unroll
        for (int c = 0; c < NC; ++c)
          sl[c] = mine[c];
        __syncthreads();

#pragma unroll
        for (int r = 0; r < ITEMS; ++r)
          {
            const int i = tid + r * NT;
            if (i < PN * NC)
              {
                const int pn = i / NC, c = i - pn * NC;
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
                        const T val = slots[q][i];
                        if (qm == 0)
                          fin += val;
                        else
                          nxt += val;
                      }
                  }
                carry[r] = nxt;
                put(kp, t, NC, k, pn, c, fin, dst, part, ostride, ocstride, ooffset);
              }
          }
        __syncthreads();
      }
#pragma unroll
    for (int r = 0; r < ITEMS; ++r)
      {
        const int i = tid + r * NT;
        if (i < PN * NC)
          {
            const int pn = i / NC, c = i - pn * NC;
            put(kp, t, NC, t.tm_eff, pn, c, carry[r], dst, part, ostride, ocstride, ooffset);
          }
      }
  }

  // ------------------------------------------- shared-node fix-up
  struct FixParams
  {
    int nc[3], nn[3], nt[3], tsz[3]; // tile sizes: TX, TY, tz
    int n_comp;
    int px, pn;                      // plane-node geometry of the tile
    int nb[3];                       // interior tile boundaries per direction
    int64_t n_m, n_y, n_x;           // thread counts of the three classes
    int ostride, ocstride, ooffset;
    int vdst;                        // dst node records written as vectors
  };

  // the widest load (16 or 8 bytes) that tiles a record of NC values of T
  template <typename T, int NC>
  struct FixVec
  {
    static constexpr int bytes = (NC * sizeof(T)) % 16 == 0 ? 16
                                 : (NC * sizeof(T)) % 8 == 0 ? 8
                                                              : (int)sizeof(T);
    static constexpr int w     = bytes / (int)sizeof(T);
    struct alignas(bytes) V
    {
      T x[w];
    };
  };

  __device__ __forceinline__ bool is_bnd(int g, int tsz, int nc)
  {
    return g > 0 && g < nc && (g % tsz) == 0;
  }

  // one thread per shared node (all NC components: the partial records of
  // a node are NC contiguous values, read together); the three classes of
  // shared nodes (on a march-direction tile boundary; on a y boundary only;
  // on an x boundary only) are enumerated back to back, x fastest
  template <typename T, int NC>
  __global__ void __launch_bounds__(256)
    q1_fixup(const FixParams fp, const T *__restrict__ part, T *__restrict__ dst)
  {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    int g[3];
    if (idx < fp.n_m)
      {
        g[0] = idx % fp.nn[0];
        idx /= fp.nn[0];
        g[1] = idx % fp.nn[1];
        g[2] = (idx / fp.nn[1] + 1) * fp.tsz[2];
      }
    else if ((idx -= (int)fp.n_m) < fp.n_y)
      {
        g[0] = idx % fp.nn[0];
        idx /= fp.nn[0];
        g[2] = idx % fp.nn[2];
        g[1] = (idx / fp.nn[2] + 1) * fp.tsz[1];
        if (is_bnd(g[2], fp.tsz[2], fp.nc[2]))
          return;
      }
    else if ((idx -= (int)fp.n_y) < fp.n_x)
      {
        const int per = fp.nb[0];
        g[0]          = (idx % per + 1) * fp.tsz[0];
        idx /= per;
        g[1] = idx % fp.nn[1];
        g[2] = idx / fp.nn[1];
        if (is_bnd(g[2], fp.tsz[2], fp.nc[2]) || is_bnd(g[1], fp.tsz[1], fp.nc[1]))
          return;
      }
    else
      return;

    // sharing tiles and local indices along each direction
    int tl[3][2], lo[3][2], ns[3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
      {
        if (is_bnd(g[d], fp.tsz[d], fp.nc[d]))
          {
            ns[d]    = 2;
            tl[d][0] = g[d] / fp.tsz[d] - 1;
            lo[d][0] = fp.tsz[d];
            tl[d][1] = g[d] / fp.tsz[d];
            lo[d][1] = 0;
          }
        else
          {
            ns[d]    = 1;
            tl[d][0] = min(g[d] / fp.tsz[d], fp.nt[d] - 1);
            lo[d][0] = g[d] - tl[d][0] * fp.tsz[d];
            tl[d][1] = tl[d][0];
            lo[d][1] = lo[d][0];
          }
      }
    using FV           = FixVec<T, NC>;
    const int64_t node = ((int64_t)g[2] * fp.nn[1] + g[1]) * fp.nn[0] + g[0];
    T             s[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c)
      s[c] = T(0);
    for (int a = 0; a < ns[2]; ++a)
      for (int b = 0; b < ns[1]; ++b)
        for (int e = 0; e < ns[0]; ++e)
          {
            const int64_t tile = ((int64_t)tl[2][a] * fp.nt[1] + tl[1][b]) * fp.nt[0] + tl[0][e];
            const int     pn   = lo[0][e] + fp.px * lo[1][b];
            const T      *p    = part + ((tile * (fp.tsz[2] + 1) + lo[2][a]) * fp.pn + pn) * NC;
            const auto   *pv   = reinterpret_cast<const typename FV::V *>(p);
#pragma unroll
            for (int v = 0; v < NC / FV::w; ++v)
              {
                const typename FV::V x = pv[v];
#pragma unroll
                for (int j = 0; j < FV::w; ++j)
                  s[v * FV::w + j] += x.x[j];
              }
          }
    if (fp.vdst)
      {
        auto *dv = reinterpret_cast<typename FV::V *>(dst + node * NC);
#pragma unroll
        for (int v = 0; v < NC / FV::w; ++v)
          {
            typename FV::V x;
#pragma unroll
            for (int j = 0; j < FV::w; ++j)
              x.x[j] = s[v * FV::w + j];
            dv[v] = x;
          }
        return;
      }
    T *d = dst + node * fp.ostride + fp.ooffset;
#pragma unroll
    for (int c = 0; c < NC; ++c)
      d[(int64_t)c * fp.ocstride] = s[c];
  }
} // namespace hps
