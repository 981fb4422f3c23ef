// fe_q1_zcol.hpp -- the 3D scalar-mobility tangent vmult and nonlinear
// residual with z-column lanes (the Q1 operator of S/operator_sintering_
// generic.h:539-577, 104-190 and 877-1006, 235-321 with MobilityScalar,
// S/mobility.h:441-534).
//
// Lane mapping.  One lane per (cell, x/y corner): lane = 4 cell + (bx | by
// << 1), cell = cx + 8 cy, so a workgroup of 256 lanes holds 8 x 8 cells of
// one layer.  A lane holds both z-levels of its corner: the two nodal
// values, then the two Gauss points (lo, hi) of its (x, y) corner, then the
// two nodal contributions.  The x and y levels of the sum factorisation are
// butterflies with the lane that differs in bit 0 or 1 (DPP quad_perm); the
// z level is arithmetic on the lane's own registers.  Per component and
// layer that is 6 lane exchanges in the interpolation and 6 in the
// integration for two Gauss points -- half of what one lane per
// (cell, corner) needs -- and no ds_swizzle.  The FP32 tangent with up to
// 4 order parameters reads the interpolation's partner values (x, y and
// diagonal corner of the cell face) from the LDS node planes instead of
// exchanging them.
//
// March.  The workgroup marches along z through `tz` cell layers with two
// node planes of 9 x 9 nodes in LDS (planes k and k+1 are a cell layer's
// bottom and top); the plane two layers ahead and the next layer's cache
// fields are fetched into registers during the current layer.  The top contribution of a lane (plane k+1) stays in
// its registers (`carry`) and is added to its bottom contribution of the
// next layer, the same node: a node's z sum never leaves the lane.  The x/y
// sum of the four cells around a node goes through LDS slots indexed by
// (x/y corner, node), each written by one lane, summed in a fixed order
// after a barrier: deterministic, no atomics.  Nodes on a tile face shared
// with another workgroup go to the partial buffer (ZSlots) and are summed by
// q1_fixup_slots in a fixed order.
//
// The linearisation cache is the one the other Q1 kernels read (tiles of
// 8 x 4 cells, lane = 8 cell + corner): the workgroup's 8 x 8 cells are two
// of those tiles, and a lane reads its lo and hi Gauss points' fields.
#pragma once

#include "fe_q1.hpp"

namespace hps
{
  namespace zcol
  {
    constexpr int TX = 8, TY = 8, NT = 256, PX = TX + 1, PN = (TX + 1) * (TY + 1);
  }

  // Partials of the nodes shared by tiles, slot-major per class: nodes on an
  // interior z tile boundary (class M, 8 slots), else on an interior y
  // boundary (class Y, 4 slots), else on an interior x boundary (class X, 2
  // slots); slot bit d is set when the writing tile is the upper one along
  // d.  Within a slot the records (n_comp values) are ordered x fastest (M,
  // Y) or y fastest (X), so that a tile writes runs of consecutive records
  // and the fix-up reads every slot of consecutive nodes contiguously.
  struct ZSlots
  {
    int     nc[3], nn[3], tz, n_comp;
    int64_t n_m, n_y, n_x;
  };

  inline ZSlots zcol_slots(const KParams &kp, int ntz1, int n_comp)
  {
    ZSlots z;
    for (int d = 0; d < 3; ++d)
      {
        z.nc[d] = kp.nc[d];
        z.nn[d] = kp.nn[d];
      }
    z.tz     = kp.tz;
    z.n_comp = n_comp;
    z.n_m    = (int64_t)(kp.nt[2] - 1) * kp.nn[1] * kp.nn[0];
    z.n_y    = (int64_t)(ntz1 - 1) * kp.nn[2] * kp.nn[0];
    z.n_x    = (int64_t)(kp.nt[0] - 1) * kp.nn[2] * kp.nn[1];
    return z;
  }

  // records of the three classes
  inline int64_t zcol_slot_records(const ZSlots &z)
  {
    return 8 * z.n_m + 4 * z.n_y + 2 * z.n_x;
  }

  // record of node (gx, gy, gz) written by the tile on side (sx, sy, sz);
  // bxn, byn, bzk: the node lies on an interior tile boundary along x, y, z
  __device__ __forceinline__ int64_t zcol_record(const ZSlots &z, int gx, int gy, int gz,
                                                 int sx, int sy, int sz, bool byn, bool bzk)
  {
    if (bzk)
      return (int64_t)(sx | sy << 1 | sz << 2) * z.n_m +
             ((int64_t)(gz / z.tz - 1) * z.nn[1] + gy) * z.nn[0] + gx;
    if (byn)
      return 8 * z.n_m + (int64_t)(sx | sy << 1) * z.n_y +
             ((int64_t)(gy / 8 - 1) * z.nn[2] + gz) * z.nn[0] + gx;
    return 8 * z.n_m + 4 * z.n_y + (int64_t)sx * z.n_x +
           ((int64_t)(gx / 8 - 1) * z.nn[2] + gz) * z.nn[1] + gy;
  }

  __device__ __forceinline__ bool zcol_bnd(int g, int t, int n)
  {
    return g > 0 && g < n && (g % t) == 0;
  }

  // one thread per shared node (class M, then Y, then X): the sum of its
  // slots in slot order, all n_comp values, into dst
  template <typename T, int NC>
  __global__ void __launch_bounds__(256)
    q1_fixup_slots(const ZSlots z, const T *__restrict__ part, T *__restrict__ dst)
  {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int     g[3];
    int     slot[8], ns = 0;
    const T *base;
    int64_t  stride;
    if (idx < z.n_m)
      {
        g[0] = (int)(idx % z.nn[0]);
        g[1] = (int)((idx / z.nn[0]) % z.nn[1]);
        g[2] = (int)(idx / ((int64_t)z.nn[0] * z.nn[1]) + 1) * z.tz;
        const bool bx = zcol_bnd(g[0], 8, z.nc[0]), by = zcol_bnd(g[1], 8, z.nc[1]);
        for (int sz = 0; sz < 2; ++sz)
          for (int sy = 0; sy <= (int)by; ++sy)
            for (int sx = 0; sx <= (int)bx; ++sx)
              slot[ns++] = sx | sy << 1 | sz << 2;
        base   = part + idx * NC;
        stride = z.n_m * NC;
      }
    else if ((idx -= z.n_m) < z.n_y)
      {
        g[0] = (int)(idx % z.nn[0]);
        g[2] = (int)((idx / z.nn[0]) % z.nn[2]);
        g[1] = (int)(idx / ((int64_t)z.nn[0] * z.nn[2]) + 1) * 8;
        if (zcol_bnd(g[2], z.tz, z.nc[2]))
          return;
        const bool bx = zcol_bnd(g[0], 8, z.nc[0]);
        for (int sy = 0; sy < 2; ++sy)
          for (int sx = 0; sx <= (int)bx; ++sx)
            slot[ns++] = sx | sy << 1;
        base   = part + (8 * z.n_m + idx) * NC;
        stride = z.n_y * NC;
      }
    else if ((idx -= z.n_y) < z.n_x)
      {
        g[1] = (int)(idx % z.nn[1]);
        g[2] = (int)((idx / z.nn[1]) % z.nn[2]);
        g[0] = (int)(idx / ((int64_t)z.nn[1] * z.nn[2]) + 1) * 8;
        if (zcol_bnd(g[2], z.tz, z.nc[2]) || zcol_bnd(g[1], 8, z.nc[1]))
          return;
        slot[ns++] = 0;
        slot[ns++] = 1;
        base       = part + (8 * z.n_m + 4 * z.n_y + idx) * NC;
        stride     = z.n_x * NC;
      }
    else
      return;
    using FV = FixVec<T, NC>;
    T s[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c)
      s[c] = T(0);
    for (int i = 0; i < ns; ++i)
      {
        const auto *pv = reinterpret_cast<const typename FV::V *>(base + slot[i] * stride);
#pragma unroll
        for (int v = 0; v < NC / FV::w; ++v)
          {
            const typename FV::V x = pv[v];
#pragma unroll
            for (int j = 0; j < FV::w; ++j)
              s[v * FV::w + j] += x.x[j];
          }
      }
    T *d = dst + (((int64_t)g[2] * z.nn[1] + g[1]) * z.nn[0] + g[0]) * NC;
#pragma unroll
    for (int c = 0; c < NC; ++c)
      d[c] = s[c];
  }

  // values and (rescaled, unsigned: own minus partner) gradients at the lane's
  // two Gauss points from its two nodal values, rescaled butterflies as
  // interp_vg_hat (val = sa^3 val_hat, grad_d = sa^2 grad_hat_d); the z
  // derivative of the hi point is -gz_lo
  template <typename T>
  __device__ __forceinline__ void zcol_interp(T ul, T uh, T &vl, T &vh, T &gxl, T &gxh, T &gyl,
                                              T &gyh, T &gzl)
  {
    const T r   = Q1Hat<T>::r;
    const T pl  = xshfl<1>(ul);
    const T ph  = xshfl<1>(uh);
    const T Al  = ul + r * pl;
    const T Ah  = uh + r * ph;
    const T Dl  = ul - pl;
    const T Dh  = uh - ph;
    const T pAl = xshfl<2>(Al);
    const T pAh = xshfl<2>(Ah);
    const T pDl = xshfl<2>(Dl);
    const T pDh = xshfl<2>(Dh);
    const T AAl = Al + r * pAl;
    const T AAh = Ah + r * pAh;
    const T Ayl = Al - pAl;
    const T Ayh = Ah - pAh;
    const T DAl = Dl + r * pDl;
    const T DAh = Dh + r * pDh;
    vl          = AAl + r * AAh;
    vh          = AAh + r * AAl;
    gzl         = AAl - AAh;
    gxl         = DAl + r * DAh;
    gxh         = DAh + r * DAl;
    gyl         = Ayl + r * Ayh;
    gyh         = Ayh + r * Ayl;
  }

  // the four nodes of the lane's cell face in a node plane of the LDS ring:
  // its own corner, the x partner, the y partner and the diagonal one
  template <typename T>
  struct ZQuad
  {
    const T *o, *x, *y, *d;
  };

  // zcol_interp with the partners' values read from the ring instead of
  // exchanged between lanes: the y partner's x-level sums are formed here
  // from its two nodes, with the operations the partner lane would use, so
  // the results are the same bit for bit
  template <typename T>
  __device__ __forceinline__ void zcol_interp_q(const ZQuad<T> &lo, const ZQuad<T> &hi, int c,
                                                T &vl, T &vh, T &gxl, T &gxh, T &gyl, T &gyh,
                                                T &gzl)
  {
    const T r   = Q1Hat<T>::r;
    const T ul  = lo.o[c], uh = hi.o[c];
    const T pl  = lo.x[c], ph = hi.x[c];
    const T ql  = lo.y[c], qh = hi.y[c];
    const T dl  = lo.d[c], dh = hi.d[c];
    const T Al  = ul + r * pl;
    const T Ah  = uh + r * ph;
    const T Dl  = ul - pl;
    const T Dh  = uh - ph;
    const T pAl = ql + r * dl;
    const T pAh = qh + r * dh;
    const T pDl = ql - dl;
    const T pDh = qh - dh;
    const T AAl = Al + r * pAl;
    const T AAh = Ah + r * pAh;
    const T Ayl = Al - pAl;
    const T Ayh = Ah - pAh;
    const T DAl = Dl + r * pDl;
    const T DAh = Dh + r * pDh;
    vl          = AAl + r * AAh;
    vh          = AAh + r * AAl;
    gzl         = AAl - AAh;
    gxl         = DAl + r * DAh;
    gxh         = DAh + r * DAl;
    gyl         = Ayl + r * Ayh;
    gyh         = Ayh + r * Ayl;
  }

  // values only (same levels, same rounding as zcol_interp)
  template <typename T>
  __device__ __forceinline__ void zcol_interp_values(T ul, T uh, T &vl, T &vh)
  {
    const T r   = Q1Hat<T>::r;
    const T Al  = ul + r * xshfl<1>(ul);
    const T Ah  = uh + r * xshfl<1>(uh);
    const T AAl = Al + r * xshfl<2>(Al);
    const T AAh = Ah + r * xshfl<2>(Ah);
    vl          = AAl + r * AAh;
    vh          = AAh + r * AAl;
  }

  // transpose of zcol_interp: the nodal contributions (lo, hi) of a row given
  // at the two Gauss points as (value, gradient) in the scaling of
  // integrate_vg_hat; levels z (registers), y, x
  template <typename T>
  __device__ __forceinline__ void zcol_integrate(T vl, T vh, const T *gl, const T *gh, T &ol,
                                                 T &oh)
  {
    const T r   = Q1Hat<T>::r;
    const T Zl  = (vl + gl[2]) + (r * vh - gh[2]);
    const T Zh  = (vh + gh[2]) + (r * vl - gl[2]);
    const T Xl  = gl[0] + r * gh[0];
    const T Xh  = gh[0] + r * gl[0];
    const T Yl  = gl[1] + r * gh[1];
    const T Yh  = gh[1] + r * gl[1];
    const T Ul  = (Zl + Yl) + xshfl<2>(r * Zl - Yl);
    const T Uh  = (Zh + Yh) + xshfl<2>(r * Zh - Yh);
    const T Xyl = Xl + r * xshfl<2>(Xl);
    const T Xyh = Xh + r * xshfl<2>(Xh);
    ol          = (Ul + Xyl) + xshfl<1>(r * Ul - Xyl);
    oh          = (Uh + Xyh) + xshfl<1>(r * Uh - Xyh);
  }

  // the same butterflies unscaled (value levels sa own + sb partner), for
  // the residual, whose physics needs the true values; derivatives
  // unsigned (own minus partner)
  template <typename T>
  __device__ __forceinline__ void zcol_interp_plain(T ul, T uh, T &vl, T &vh, T &gxl, T &gxh,
                                                    T &gyl, T &gyh, T &gzl)
  {
    const T sa  = Q1Coef<T>::sa, sb = Q1Coef<T>::sb;
    const T pl  = xshfl<1>(ul);
    const T ph  = xshfl<1>(uh);
    const T Al  = sa * ul + sb * pl;
    const T Ah  = sa * uh + sb * ph;
    const T Dl  = ul - pl;
    const T Dh  = uh - ph;
    const T pAl = xshfl<2>(Al);
    const T pAh = xshfl<2>(Ah);
    const T pDl = xshfl<2>(Dl);
    const T pDh = xshfl<2>(Dh);
    const T AAl = sa * Al + sb * pAl;
    const T AAh = sa * Ah + sb * pAh;
    const T Ayl = Al - pAl;
    const T Ayh = Ah - pAh;
    const T DAl = sa * Dl + sb * pDl;
    const T DAh = sa * Dh + sb * pDh;
    vl          = sa * AAl + sb * AAh;
    vh          = sa * AAh + sb * AAl;
    gzl         = AAl - AAh;
    gxl         = sa * DAl + sb * DAh;
    gxh         = sa * DAh + sb * DAl;
    gyl         = sa * Ayl + sb * Ayh;
    gyh         = sa * Ayh + sb * Ayl;
  }

  // values only (same levels, same rounding as zcol_interp_plain)
  template <typename T>
  __device__ __forceinline__ void zcol_interp_values_plain(T ul, T uh, T &vl, T &vh)
  {
    const T sa  = Q1Coef<T>::sa, sb = Q1Coef<T>::sb;
    const T Al  = sa * ul + sb * xshfl<1>(ul);
    const T Ah  = sa * uh + sb * xshfl<1>(uh);
    const T AAl = sa * Al + sb * xshfl<2>(Al);
    const T AAh = sa * Ah + sb * xshfl<2>(Ah);
    vl          = sa * AAl + sb * AAh;
    vh          = sa * AAh + sb * AAl;
  }

  // transpose of zcol_interp_plain; the gradients given as the gradient row
  // times the signed JxW / h_d of the Gauss point's corner (integrate_vg)
  template <typename T>
  __device__ __forceinline__ void zcol_integrate_plain(T vl, T vh, const T *gl, const T *gh,
                                                       T &ol, T &oh)
  {
    const T sa  = Q1Coef<T>::sa, sb = Q1Coef<T>::sb;
    const T Zl  = (sa * vl + gl[2]) + (sb * vh - gh[2]);
    const T Zh  = (sa * vh + gh[2]) + (sb * vl - gl[2]);
    const T Xl  = sa * gl[0] + sb * gh[0];
    const T Xh  = sa * gh[0] + sb * gl[0];
    const T Yl  = sa * gl[1] + sb * gh[1];
    const T Yh  = sa * gh[1] + sb * gl[1];
    const T Ul  = (sa * Zl + Yl) + xshfl<2>(sb * Zl - Yl);
    const T Uh  = (sa * Zh + Yh) + xshfl<2>(sb * Zh - Yh);
    const T Xyl = sa * Xl + sb * xshfl<2>(Xl);
    const T Xyh = sa * Xh + sb * xshfl<2>(Xh);
    ol          = (sa * Ul + Xyl) + xshfl<1>(sb * Ul - Xyl);
    oh          = (sa * Uh + Xyh) + xshfl<1>(sb * Uh - Xyh);
  }

  // ntz1: tiles along y (8 cells each); the workgroups are the tiles
  // (x fastest, then y, then the march direction).  W: waves per SIMD asked
  // of the register allocation.  OP: the tangent (OP_VMULT, reads the
  // cache) or the residual (OP_RES0/1/2, S/operator_sintering_generic.h:
  // 877-1006 with MobilityScalar; OP_RES2 reads the old solutions, FP64).
  template <int G, typename T, int W, int OP = OP_VMULT>
  __global__ void __launch_bounds__(zcol::NT) __attribute__((amdgpu_waves_per_eu(W)))
    q1_vmult_zcol(const KParams kp, const ZSlots zs, int ntz1, const T *__restrict__ src,
                  const T *__restrict__ cache, T *__restrict__ dst, T *__restrict__ part,
                  const T *__restrict__ old0 = nullptr, const T *__restrict__ old1 = nullptr,
                  const T *__restrict__ old2 = nullptr)
  {
    using CL            = CacheLayout<3, G, false>;
    constexpr bool VM   = OP == OP_VMULT;
    constexpr bool HIST = OP == OP_RES2;
    constexpr int  NPL  = HIST ? 2 : 1; // planes per ring slot: src, history
    constexpr int NC    = G + 2;
    constexpr int NF    = CL::n_fields;
    constexpr int NT    = zcol::NT;
    constexpr int PX    = zcol::PX;
    constexpr int PN    = zcol::PN;
    constexpr int ITEMS = (PN * NC + NT - 1) / NT;
    constexpr int NTC   = 256; // lanes of a cache tile (8 x 4 cells x 8 corners)

    __shared__ T ring[2][NPL][PN * NC];
    __shared__ T slots[4][PN * NC];

    const int tid  = threadIdx.x;
    const int xy   = tid & 3;
    const int cell = tid >> 2;
    const int cx = cell & 7, cy = cell >> 3;
    const int bx = xy & 1, by = xy >> 1;
    int       b  = blockIdx.x;
    const int tix = b % kp.nt[0];
    b /= kp.nt[0];
    const int tiy = b % ntz1;
    const int tim = b / ntz1;
    const int x0 = tix * zcol::TX, y0 = tiy * zcol::TY, m0 = tim * kp.tz;
    const int tx_eff  = min(zcol::TX, kp.nc[0] - x0);
    const int ty_eff  = min(zcol::TY, kp.nc[1] - y0);
    const int tm_eff  = min(kp.tz, kp.nc[2] - m0);
    const int pn_lane = (cx + bx) + PX * (cy + by);

    // the cache tile of the lane's cell (cells beyond the mesh read a valid
    // tile; their contributions are never summed)
    int tyc = 2 * tiy + (cy >> 2);
    if (tyc >= kp.nt[1])
      tyc = 2 * tiy;
    // cache tiles are ctz layers deep; the march of tz layers crosses
    // tz / ctz of them
    // (FP64, MW: the lane's x/y tile in a pointer; the cache tile along z
    // and the layer in it are counters that load_lin advances by one layer
    // per call, the layers being requested in order from the march's
    // first.  FP32 marches one cache tile: kp.ctz = kp.tz, the tile's
    // layers at a fixed stride from cache_t; the counters cost it 3 %)
    constexpr bool MW      = sizeof(T) == 8;
    const int      ctz     = MW && kp.ctz > 0 ? kp.ctz : kp.tz;
    const int64_t  ctile   = tix + (int64_t)kp.nt[0] * (tyc + (MW ? 0 : (int64_t)kp.nt[1] * tim));
    const T       *cache_t = cache + ctile * ctz * NF * NTC;
    const int64_t  cslab   = (int64_t)kp.nt[0] * kp.nt[1] * ctz;
    int            lin_ct = MW ? tim * (kp.tz / ctz) : 0, lin_layer = 0;
    const int     lane_lo = ((((cy & 3) * 8) + cx) << 3) | xy; // corner with z bit 0
    const int     lane_hi = lane_lo | 4;

    // the plane items of this thread (item i = node pn, component c), fixed
    // for the whole march: whether the node is in the tile, its index in
    // src/dst at plane 0 (planes are pstride apart), the x/y corner slots
    // that hold a valid cell, and the partial record of a node on an x or y
    // tile boundary at plane 0 (records along z are xystride apart); the
    // records of an interior z boundary plane are formed where needed
    const int64_t pstride = (int64_t)kp.nn[0] * kp.nn[1] * NC;
    bool          it_ok[ITEMS], it_in[ITEMS], it_xy[ITEMS];
    int64_t       it_g0[ITEMS], it_r0[ITEMS];
    int           it_rs[ITEMS];
    unsigned      it_qm[ITEMS];
#pragma unroll
    for (int r = 0; r < ITEMS; ++r)
      {
        const int i = tid + r * NT;
        it_ok[r]    = i < PN * NC;
        const int pn = i / NC, c = i - pn * NC;
        const int nx = pn % PX, ny = pn / PX;
        const int gx = x0 + nx, gy = y0 + ny;
        it_in[r]     = it_ok[r] && nx <= tx_eff && ny <= ty_eff;
        it_g0[r]     = (((int64_t)m0 * kp.nn[1] + gy) * kp.nn[0] + gx) * NC + c;
        unsigned qm  = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q)
          {
            const int ccx = nx - (q & 1), ccy = ny - (q >> 1);
            if (ccx >= 0 && ccx < tx_eff && ccy >= 0 && ccy < ty_eff)
              qm |= 1u << q;
          }
        it_qm[r]       = qm;
        const bool bxn = (nx == 0 && x0 > 0) || (nx == tx_eff && x0 + tx_eff < kp.nc[0]);
        const bool byn = (ny == 0 && y0 > 0) || (ny == ty_eff && y0 + ty_eff < kp.nc[1]);
        it_xy[r]       = bxn || byn;
        it_r0[r] = zcol_record(zs, gx, gy, 0, bxn && nx == 0, byn && ny == 0, 0, byn, false) * NC + c;
        it_rs[r] = (byn ? kp.nn[0] : kp.nn[1]) * NC;
      }
    // record of item r on the interior z boundary plane k (0 or tm_eff)
    auto m_record = [&](int r, int k) {
      const int i  = tid + r * NT;
      const int pn = i / NC, c = i - pn * NC;
      const int nx = pn % PX, ny = pn / PX;
      const bool bxn = (nx == 0 && x0 > 0) || (nx == tx_eff && x0 + tx_eff < kp.nc[0]);
      const bool byn = (ny == 0 && y0 > 0) || (ny == ty_eff && y0 + ty_eff < kp.nc[1]);
      return zcol_record(zs, x0 + nx, y0 + ny, m0 + k, bxn && nx == 0, byn && ny == 0, k == 0,
                         byn, true) *
               NC +
             c;
    };

    // src (and, with the history, the combination of the old solutions
    // with their BDF weights, do_evalute_history, S/operator_sintering_
    // generic.h:950-960) of node plane k
    T    regs[NPL][ITEMS];
    auto load_plane = [&](int k) {
#pragma unroll
      for (int r = 0; r < ITEMS; ++r)
        if (it_ok[r])
          {
            const int64_t gi = it_g0[r] + k * pstride;
            regs[0][r]       = it_in[r] ? src[gi] : T(0);
            if constexpr (HIST)
              {
                T h = T(0);
                if (it_in[r])
                  {
                    h = old0[gi] * T(kp.w_hist[0]);
                    if (kp.order >= 2)
                      h += old1[gi] * T(kp.w_hist[1]);
                    if (kp.order >= 3)
                      h += old2[gi] * T(kp.w_hist[2]);
                  }
                regs[1][r] = h;
              }
          }
    };
    auto store_plane = [&](int slot) {
#pragma unroll
      for (int r = 0; r < ITEMS; ++r)
        if (it_ok[r])
#pragma unroll
          for (int a = 0; a < NPL; ++a)
            ring[slot][a][tid + r * NT] = regs[a][r];
    };
    // plane k: the sum over the x/y corner slots of the valid cells, to dst
    // or to the node's partial record
    auto node_sums = [&](int k) {
      const bool bzk = (k == 0 && m0 > 0) || (k == tm_eff && m0 + tm_eff < kp.nc[2]);
#pragma unroll
      for (int r = 0; r < ITEMS; ++r)
        if (it_in[r])
          {
            const int i   = tid + r * NT;
            T         fin = T(0);
#pragma unroll
            for (int q = 0; q < 4; ++q)
              if ((it_qm[r] >> q) & 1u)
                fin += slots[q][i];
            T *out;
            if (bzk)
              out = part + m_record(r, k);
            else
              out = it_xy[r] ? part + (it_r0[r] + (int64_t)(m0 + k) * it_rs[r])
                             : dst + (it_g0[r] + k * pstride);
            *out = fin;
          }
    };

    // uniform folded-tangent scales (set_fold_scales); the signed ones per
    // Gauss point (lo: z bit 0, hi: z bit 1)
    FoldScales<T> fs;
    T             jxw_hat;
    fold_scales_hat<3>(kp, xy, fs, jxw_hat);
    const T gsz = T(kp.fold_gs[2]);

    T carry[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c)
      carry[c] = T(0);

    // the cache fields of layer k's two Gauss points (tangent).  FP64: the
    // next layer's fields are requested during the current layer (PF); FP32:
    // at the top of their layer, which leaves the registers of the prefetch
    // free (3 % faster on the 128^3 workload; FP64 is 1 % faster with it)
    // more than 4 order parameters (tangent): values first, the gradients of
    // c and mu with them, an eta gradient interpolated again where its row
    // is integrated (its registers are not held across the physics)
    constexpr bool TWO = VM && G > 4;
    constexpr int  NGV = TWO ? 2 : NC; // gradients interpolated with the values
    constexpr bool PF  = VM && sizeof(T) == 8 && !TWO;
    // FP32 tangent with up to 4 order parameters: the interpolation's
    // partner values from the ring (zcol_interp_q; 2-3 % faster on the
    // 128^3 workload with 4; FP64 gains nothing and its allocation spills,
    // FP32 with 8 or 12 loses 0.3-1.7 %)
    constexpr bool LDSQ = VM && sizeof(T) == 4 && G <= 4;
    const int      q_dx = (bx ? -1 : 1) * NC, q_dy = (by ? -PX : PX) * NC;
    T    lin_next[VM ? 2 : 1][VM ? NF : 1];
    auto load_lin = [&](int k) { // (called for k = 0, 1, ... in order)
      if constexpr (VM)
        {
          const T *cl;
          if constexpr (MW)
            {
              cl = cache_t + (cslab * lin_ct + lin_layer) * (NF * NTC);
              if (++lin_layer == ctz)
                {
                  lin_layer = 0;
                  ++lin_ct;
                }
            }
          else
            cl = cache_t + (int64_t)k * (NF * NTC);
#pragma unroll
          for (int f = 0; f < NF; ++f)
            {
              lin_next[0][f] = cl[f * NTC + lane_lo];
              lin_next[1][f] = cl[f * NTC + lane_hi];
            }
        }
    };
    // residual: the gradient rows' uniform scales JxW / h_d^2 (the sign of
    // a Gauss point's derivative and of its test function cancel) and the
    // value rows' JxW
    T k2[3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
      k2[d] = T(kp.jxw * kp.inv_h[d] * kp.inv_h[d]);
    const T jxw = T(kp.jxw);

    load_plane(0);
    store_plane(0);
    load_plane(1);
    store_plane(1);
    __syncthreads();
    // the next layer's cache fields and the plane after next are requested
    // together at the end of a layer (after the wait for the previous
    // plane), so that they arrive during the node sums and the next
    // interpolation
    if (2 <= tm_eff)
      load_plane(2);
    if (PF)
      load_lin(0);

    for (int k = 0; k < tm_eff; ++k)
      {
        if (!PF)
          load_lin(k);
        const T *plo = &ring[k & 1][0][pn_lane * NC];
        const T *phi = &ring[(k + 1) & 1][0][pn_lane * NC];
        const ZQuad<T> qlo{plo, plo + q_dx, plo + q_dy, plo + q_dx + q_dy};
        const ZQuad<T> qhi{phi, phi + q_dx, phi + q_dy, phi + q_dx + q_dy};
        if constexpr (!VM)
          {
            T val[2][NC], gx[2][NC], gy[2][NC], gz[NC], hist[2][NC];
#pragma unroll
            for (int c = 0; c < NC; ++c)
              zcol_interp_plain(plo[c], phi[c], val[0][c], val[1][c], gx[0][c], gx[1][c],
                                gy[0][c], gy[1][c], gz[c]);
            if constexpr (HIST)
              {
                const T *hlo = &ring[k & 1][1][pn_lane * NC];
                const T *hhi = &ring[(k + 1) & 1][1][pn_lane * NC];
#pragma unroll
                for (int c = 0; c < NC; ++c)
                  zcol_interp_values_plain(hlo[c], hhi[c], hist[0][c], hist[1][c]);
              }
            __builtin_amdgcn_sched_barrier(0);
            // value rows and the mobility of the c row per Gauss point
            T vr[2][NC], M[2];
#pragma unroll
            for (int n = 0; n < 2; ++n)
              {
                residual_values<3, G, false, OP - OP_RES0>(kp.pp, val[n], hist[n], vr[n]);
                const T one[3] = {T(1), T(1), T(1)};
                T       m3[3];
                residual_grad_c_scalar<3, G>(kp.pp, val[n], one, m3);
                M[n] = m3[0];
              }
            // gradient rows M grad mu, -kappa_c grad c, L kappa_p grad eta_i
            const T kc = -T(kp.pp.kappa_c), lkp = T(kp.pp.L) * T(kp.pp.kappa_p);
#pragma unroll
            for (int c = 0; c < NC; ++c)
              {
                const int src_c = c == 0 ? 1 : (c == 1 ? 0 : c);
                T         gs[2][3];
#pragma unroll
                for (int n = 0; n < 2; ++n)
                  {
                    const T f     = c == 0 ? M[n] : (c == 1 ? kc : lkp);
                    const T gg[3] = {gx[n][src_c], gy[n][src_c], n ? -gz[src_c] : gz[src_c]};
#pragma unroll
                    for (int d = 0; d < 3; ++d)
                      gs[n][d] = (f * gg[d]) * k2[d];
                  }
                T ol, oh;
                zcol_integrate_plain(vr[0][c] * jxw, vr[1][c] * jxw, gs[0], gs[1], ol, oh);
                slots[xy][pn_lane * NC + c] = ol + carry[c];
                carry[c]                    = oh;
              }
          }
        else
          {
            T lin[2][NF];
#pragma unroll
            for (int f = 0; f < NF; ++f)
              {
                lin[0][f] = lin_next[0][f];
                lin[1][f] = lin_next[1][f];
              }
            T        val[2][NC], gx[2][NGV], gy[2][NGV], gz[NGV];
#pragma unroll
            for (int c = 0; c < NC; ++c)
              if (c < NGV)
                {
                  if constexpr (LDSQ)
                    zcol_interp_q(qlo, qhi, c, val[0][c], val[1][c], gx[0][c], gx[1][c],
                                  gy[0][c], gy[1][c], gz[c]);
                  else
                    zcol_interp(plo[c], phi[c], val[0][c], val[1][c], gx[0][c], gx[1][c],
                                gy[0][c], gy[1][c], gz[c]);
                }
              else
                zcol_interp_values(plo[c], phi[c], val[0][c], val[1][c]);
            __builtin_amdgcn_sched_barrier(0);

            // value rows and the c-row mobility terms per Gauss point
            T vr[2][NC], M[2], s[2];
#pragma unroll
            for (int n = 0; n < 2; ++n)
              {
                // (the value rows read no gradient)
                T g[NC][3], gr_unused[NC][3];
#pragma unroll
                for (int c = 0; c < NC; ++c)
#pragma unroll
                  for (int d = 0; d < 3; ++d)
                    g[c][d] = T(0);
                tangent_qp_rows<3, G, false, true, false>(kp.pp, lin[n], val[n], g, vr[n],
                                                          gr_unused);
                scalar_mobility_terms<G>(kp.pp, lin[n][CL::c], lin[n] + CL::eta, val[n], M[n],
                                         s[n]);
              }
            // gradient rows, one component at a time, and the integration
#pragma unroll
            for (int c = 0; c < NC; ++c)
              {
                // the gradient a row needs: grad mu for the c row, grad c for
                // the mu row, its own for an eta row
                const int src_c = c == 0 ? 1 : (c == 1 ? 0 : c);
                T         ex[2], ey[2], ez;
                if (src_c < NGV)
                  {
                    ex[0] = gx[0][src_c];
                    ex[1] = gx[1][src_c];
                    ey[0] = gy[0][src_c];
                    ey[1] = gy[1][src_c];
                    ez    = gz[src_c];
                  }
                else
                  {
                    T vl_unused, vh_unused;
                    zcol_interp(plo[c], phi[c], vl_unused, vh_unused, ex[0], ex[1], ey[0], ey[1],
                                ez);
                  }
                T gs[2][3];
#pragma unroll
                for (int n = 0; n < 2; ++n)
                  {
                    const T gg[3]  = {ex[n], ey[n], n ? -ez : ez};
                    const T gsj[3] = {fs.gsj[0], fs.gsj[1], n ? gsz : -gsz};
#pragma unroll
                    for (int d = 0; d < 3; ++d)
                      gs[n][d] = c == 0 ? (M[n] * fs.kj[d]) * gg[d] +
                                            s[n] * (lin[n][CL::grad_mu + d] * gsj[d])
                                        : (c == 1 ? fs.kmu[d] * gg[d] : fs.keta[d] * gg[d]);
                  }
                T ol, oh;
                zcol_integrate(vr[0][c] * jxw_hat, vr[1][c] * jxw_hat, gs[0], gs[1], ol, oh);
                slots[xy][pn_lane * NC + c] = ol + carry[c];
                carry[c]                    = oh;
              }
          }
        __syncthreads();
        if (k + 2 <= tm_eff)
          store_plane(k & 1);
        if (k + 3 <= tm_eff)
          load_plane(k + 3);
        if (PF && k + 1 < tm_eff)
          load_lin(k + 1);
        node_sums(k);
        __syncthreads();
      }
    // the top plane: the last layer's top contributions
#pragma unroll
    for (int c = 0; c < NC; ++c)
      slots[xy][pn_lane * NC + c] = carry[c];
    __syncthreads();
    node_sums(tm_eff);
  }
} // namespace hps
