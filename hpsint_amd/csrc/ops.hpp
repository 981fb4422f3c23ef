// ops.hpp -- launch tables of the Q1 and Q2 cell kernels, one entry per
// (dim, n_grains, mobility); filled by the inst_*.hip translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "fe_q1.hpp"
#include "fe_q2.hpp"

namespace hps
{
  struct OpTable
  {
    int nt;       // threads per workgroup
    int tx, ty;   // tile extent (cells) along x and y
    int n_fields; // cached linearisation fields per q-point
    void (*vmult64)(const KParams &, int n_tiles, const double *src, double *dst,
                    double *part, const double *cache, hipStream_t);
    void (*vmult32)(const KParams &, int n_tiles, const float *src, float *dst,
                    float *part, const float *cache, hipStream_t);
    void (*residual)(const KParams &, int n_tiles, int wtd, const double *src,
                     const double *o0, const double *o1, const double *o2,
                     double *dst, double *part, hipStream_t);
    void (*fill)(const KParams &, int n_tiles, const double *lin, double *c64,
                 float *c32, hipStream_t);
    void (*probe)(const KParams &, int n_tiles, const double *cache, double *dst,
                  double *part, int column, int ostride, int ocstride, int ooffset,
                  hipStream_t);
    void (*probe32)(const KParams &, int n_tiles, const float *cache, float *dst, float *part,
                    int column, int ostride, int ocstride, int ooffset, hipStream_t);
    // the same operations with the grain cut-off (per-cell masks kp.cmask)
    void (*vmult64_cut)(const KParams &, int n_tiles, const double *src, double *dst,
                        double *part, const double *cache, hipStream_t);
    void (*vmult32_cut)(const KParams &, int n_tiles, const float *src, float *dst,
                        float *part, const float *cache, hipStream_t);
    void (*residual_cut)(const KParams &, int n_tiles, int wtd, const double *src,
                         const double *o0, const double *o1, const double *o2,
                         double *dst, double *part, hipStream_t);
    void (*fill_cut)(const KParams &, int n_tiles, const double *lin, double *c64,
                     float *c32, hipStream_t);
    void (*probe_cut)(const KParams &, int n_tiles, const double *cache, double *dst,
                      double *part, int column, int ostride, int ocstride, int ooffset,
                      hipStream_t);
    void (*probe32_cut)(const KParams &, int n_tiles, const float *cache, float *dst,
                        float *part, int column, int ostride, int ocstride, int ooffset,
                        hipStream_t);
  };

  template <int DIM, int TX, int TY, int G, bool TENS>
  struct Q1Launch
  {
    static constexpr int NT = TX * TY * (1 << DIM);

    template <bool CUT>
    static void vmult64(const KParams &kp, int n_tiles, const double *src, double *dst,
                        double *part, const double *cache, hipStream_t s)
    {
      q1_apply<DIM, TX, TY, G, TENS, OP_VMULT, double, CUT>
        <<<n_tiles, NT, 0, s>>>(kp, src, nullptr, nullptr, nullptr, cache, dst, part);
    }
    template <bool CUT>
    static void vmult32(const KParams &kp, int n_tiles, const float *src, float *dst,
                        float *part, const float *cache, hipStream_t s)
    {
      q1_apply<DIM, TX, TY, G, TENS, OP_VMULT, float, CUT>
        <<<n_tiles, NT, 0, s>>>(kp, src, nullptr, nullptr, nullptr, cache, dst, part);
    }
    template <bool CUT>
    static void residual(const KParams &kp, int n_tiles, int wtd, const double *src,
                         const double *o0, const double *o1, const double *o2,
                         double *dst, double *part, hipStream_t s)
    {
      if (wtd == 0)
        q1_apply<DIM, TX, TY, G, TENS, OP_RES0, double, CUT>
          <<<n_tiles, NT, 0, s>>>(kp, src, o0, o1, o2, nullptr, dst, part);
      else if (wtd == 1)
        q1_apply<DIM, TX, TY, G, TENS, OP_RES1, double, CUT>
          <<<n_tiles, NT, 0, s>>>(kp, src, o0, o1, o2, nullptr, dst, part);
      else
        q1_apply<DIM, TX, TY, G, TENS, OP_RES2, double, CUT>
          <<<n_tiles, NT, 0, s>>>(kp, src, o0, o1, o2, nullptr, dst, part);
    }
    template <bool CUT>
    static void fill(const KParams &kp, int n_tiles, const double *lin, double *c64,
                     float *c32, hipStream_t s)
    {
      q1_fill<DIM, TX, TY, G, TENS, CUT><<<n_tiles, NT, 0, s>>>(kp, lin, c64, c32);
    }
    template <bool CUT>
    static void probe(const KParams &kp, int n_tiles, const double *cache, double *dst,
                      double *part, int column, int os, int ocs, int oo, hipStream_t s)
    {
      q1_probe<DIM, TX, TY, G, TENS, double, CUT>
        <<<n_tiles, NT, 0, s>>>(kp, cache, dst, part, column, os, ocs, oo);
    }
    template <bool CUT>
    static void probe32(const KParams &kp, int n_tiles, const float *cache, float *dst,
                        float *part, int column, int os, int ocs, int oo, hipStream_t s)
    {
      q1_probe<DIM, TX, TY, G, TENS, float, CUT>
        <<<n_tiles, NT, 0, s>>>(kp, cache, dst, part, column, os, ocs, oo);
    }
    static OpTable table()
    {
      return OpTable{NT,
                     TX,
                     TY,
                     CacheLayout<DIM, G, TENS>::n_fields,
                     &vmult64<false>,
                     &vmult32<false>,
                     &residual<false>,
                     &fill<false>,
                     &probe<false>,
                     &probe32<false>,
                     &vmult64<true>,
                     &vmult32<true>,
                     &residual<true>,
                     &fill<true>,
                     &probe<true>,
                     &probe32<true>};
    }
  };

  // ---------------------------------------------------------------- Q2
  // Every entry runs the 2^dim colour passes in order on one stream.
  struct Q2Table
  {
    int n_fields; // cached linearisation fields per q-point
    int cpw;      // cells per workgroup
    void (*vmult64)(const Q2Params &, const double *src, double *dst, const double *cache,
                    hipStream_t);
    void (*vmult32)(const Q2Params &, const float *src, float *dst, const float *cache,
                    hipStream_t);
    void (*residual)(const Q2Params &, int wtd, const double *src, const double *o0,
                     const double *o1, const double *o2, double *dst, hipStream_t);
    void (*fill)(const Q2Params &, const double *lin, double *c64, float *c32, hipStream_t);
    void (*probe)(const Q2Params &, const double *cache, double *dst, int column, int ostride,
                  int ocstride, int ooffset, hipStream_t);
    void (*probe32)(const Q2Params &, const float *cache, float *dst, int column, int ostride,
                    int ocstride, int ooffset, hipStream_t);
  };

  // parameters of colour `col` (bit d = parity along d); wg: running count
  // of workgroups of the earlier colours (their cache offset)
  inline Q2Params q2_colour(Q2Params kp, int dim, int col, int cpw, int64_t &wg)
  {
    kp.n_col = 1;
    for (int d = 0; d < 3; ++d)
      {
        kp.par[d]  = d < dim ? (col >> d) & 1 : 0;
        kp.ncol[d] = d < dim ? (kp.nc[d] - kp.par[d] + 1) / 2 : 1;
        kp.n_col *= kp.ncol[d];
      }
    kp.wg_base = wg;
    wg += (kp.n_col + cpw - 1) / cpw;
    return kp;
  }

  // workgroups of all colours (the cache holds n_fields * cpw * (p+1)^dim
  // entries each)
  inline int64_t q2_workgroups(const Q2Params &kp, int dim, int cpw)
  {
    int64_t wg = 0;
    for (int col = 0; col < (1 << dim); ++col)
      (void)q2_colour(kp, dim, col, cpw, wg);
    return wg;
  }

  template <int DIM, int G, bool TENS>
  struct Q2Launch
  {
    static constexpr int CPW = Q2Lane<DIM>::CPW;

    template <typename F>
    static void colours(const Q2Params &kp, F &&launch)
    {
      int64_t wg = 0;
      for (int col = 0; col < (1 << DIM); ++col)
        {
          const Q2Params k = q2_colour(kp, DIM, col, CPW, wg);
          if (k.n_col > 0)
            launch(k, (unsigned)((k.n_col + CPW - 1) / CPW));
        }
    }
    // 3D tile march (q2_vmult_tile) with tiles of one cell row (NY = 1, two
    // waves per workgroup): at 96^3 cells, 10 order parameters, it ran the
    // FP64 vmult 10 % faster than tiles of two rows (four waves per
    // workgroup, fewer bytes but more waves held at each barrier)
    template <typename T>
    static void tiles(const Q2Params &kp, const T *src, T *dst, const T *cache, hipStream_t s)
    {
      constexpr int NY = 1;
      Q2TileParams  tp{};
      int64_t       wg = 0;
      for (int col = 0; col < 8; ++col)
        tp.wgb[col] = q2_colour(kp, 3, col, CPW, wg).wg_base;
      const int ext[3] = {4, NY, kp.tile_zc};
      int       nt[3];
      for (int d = 0; d < 3; ++d)
        nt[d] = (kp.nc[d] + ext[d] - 1) / ext[d];
      for (int col = 0; col < 8; ++col)
        {
          for (int d = 0; d < 3; ++d)
            {
              tp.tpar[d] = (col >> d) & 1;
              tp.ntc[d]  = (nt[d] - tp.tpar[d] + 1) / 2;
            }
          const int64_t nb = (int64_t)tp.ntc[0] * tp.ntc[1] * tp.ntc[2];
          if (nb > 0)
            q2_vmult_tile<G, TENS, T, NY, sizeof(T) == 8 ? 2 : 3>
              <<<(unsigned)nb, 128 * NY, 0, s>>>(kp, tp, src, cache, dst);
        }
    }
    static void vmult64(const Q2Params &kp, const double *src, double *dst, const double *cache,
                        hipStream_t s)
    {
      if constexpr (DIM == 3)
        if (kp.tile_zc > 0)
          return tiles<double>(kp, src, dst, cache, s);
      colours(kp, [&](const Q2Params &k, unsigned nb) {
        q2_apply<DIM, G, TENS, OP_VMULT, double>
          <<<nb, 64, 0, s>>>(k, src, nullptr, nullptr, nullptr, cache, dst);
      });
    }
    static void vmult32(const Q2Params &kp, const float *src, float *dst, const float *cache,
                        hipStream_t s)
    {
      if constexpr (DIM == 3)
        if (kp.tile_zc > 0)
          return tiles<float>(kp, src, dst, cache, s);
      colours(kp, [&](const Q2Params &k, unsigned nb) {
        q2_apply<DIM, G, TENS, OP_VMULT, float>
          <<<nb, 64, 0, s>>>(k, src, nullptr, nullptr, nullptr, cache, dst);
      });
    }
    static void residual(const Q2Params &kp, int wtd, const double *src, const double *o0,
                         const double *o1, const double *o2, double *dst, hipStream_t s)
    {
      colours(kp, [&](const Q2Params &k, unsigned nb) {
        if (wtd == 0)
          q2_apply<DIM, G, TENS, OP_RES0, double><<<nb, 64, 0, s>>>(k, src, o0, o1, o2, nullptr, dst);
        else if (wtd == 1)
          q2_apply<DIM, G, TENS, OP_RES1, double><<<nb, 64, 0, s>>>(k, src, o0, o1, o2, nullptr, dst);
        else
          q2_apply<DIM, G, TENS, OP_RES2, double><<<nb, 64, 0, s>>>(k, src, o0, o1, o2, nullptr, dst);
      });
    }
    static void fill(const Q2Params &kp, const double *lin, double *c64, float *c32,
                     hipStream_t s)
    {
      colours(kp, [&](const Q2Params &k, unsigned nb) {
        q2_fill<DIM, G, TENS><<<nb, 64, 0, s>>>(k, lin, c64, c32);
      });
    }
    static void probe(const Q2Params &kp, const double *cache, double *dst, int column, int os,
                      int ocs, int oo, hipStream_t s)
    {
      colours(kp, [&](const Q2Params &k, unsigned nb) {
        q2_probe<DIM, G, TENS, double><<<nb, 64, 0, s>>>(k, cache, dst, column, os, ocs, oo);
      });
    }
    static void probe32(const Q2Params &kp, const float *cache, float *dst, int column, int os,
                        int ocs, int oo, hipStream_t s)
    {
      colours(kp, [&](const Q2Params &k, unsigned nb) {
        q2_probe<DIM, G, TENS, float><<<nb, 64, 0, s>>>(k, cache, dst, column, os, ocs, oo);
      });
    }
    static Q2Table table()
    {
      return Q2Table{CacheLayout<DIM, G, TENS>::n_fields, CPW, &vmult64, &vmult32,
                     &residual, &fill, &probe, &probe32};
    }
  };

  // nullptr when (dim, n_grains, mobility) is not built for Q2
  const Q2Table *q2_table(int dim, int n_grains, bool tensorial);

  // tile extents: 3D 8x4 cells (256 lanes), 2D 64 cells (256 lanes)
  constexpr int TX3 = 8, TY3 = 4;
  constexpr int TX2 = 64, TY2 = 1;

  // nullptr when (n_grains, mobility) is not built for that dimension
  // ty: tile extent along y (4, or 8 for scalar mobility)
  const OpTable *q1_table_d3(int n_grains, bool tensorial, int ty);
  const OpTable *q1_table_d2(int n_grains, bool tensorial);

  template <typename T>
  void launch_fixup(const FixParams &fp, const T *part, T *dst, hipStream_t s);

  // grain cut-off: per-cell masks of the active order parameters (Q1, 2D
  // or 3D); the union of the cell masks of every tile (3D, TX3 x TY3
  // tiles); the fast tangent vmult of `count` tiles (tiles[]) that all have
  // ga active order parameters (3D scalar mobility)
  constexpr int CUT_MAX_GA = 8;
  void cell_masks(const KParams &kp, int dim, int n_grains, double tol, const double *u,
                  unsigned *masks, hipStream_t s);
  void tile_masks_d3(const KParams &kp, int n_tiles, unsigned *masks, hipStream_t s);
  bool vmult_cut_d3(int ga, const KParams &kp, int n_grains, const int *tiles, int count,
                    const unsigned *masks, const double *src, const double *cache, double *dst,
                    double *part, hipStream_t s);
  bool vmult_cut_d3(int ga, const KParams &kp, int n_grains, const int *tiles, int count,
                    const unsigned *masks, const float *src, const float *cache, float *dst,
                    float *part, hipStream_t s);

  // the 3D scalar-mobility tangent with z-column lanes (fe_q1_zcol.hpp):
  // 8 x 8 x tz tiles, ntz1 of them along y; reads the cache of the TX3 x
  // TY3 tiles.  False when n_grains is not built for it.
  constexpr int ZCOL_MAX_G   = 4;  // FP64 tangent and the residual
  constexpr int ZCOL32_MAX_G = 12; // FP32 tangent
  bool vmult_zcol_d3(int n_grains, const KParams &kp, int ntz1, int n_tiles, const double *src,
                     const double *cache, double *dst, double *part, hipStream_t s);
  bool vmult_zcol_d3(int n_grains, const KParams &kp, int ntz1, int n_tiles, const float *src,
                     const float *cache, float *dst, float *part, hipStream_t s);
  // its shared-node fix-up (partials slot-major per boundary class) and the
  // size of its partial buffer (values)
  void    fixup_zcol_d3(const KParams &kp, int ntz1, int n_comp, const double *part,
                        double *dst, hipStream_t s);
  void    fixup_zcol_d3(const KParams &kp, int ntz1, int n_comp, const float *part, float *dst,
                        hipStream_t s);
  int64_t zcol_part_elems(const KParams &kp, int ntz1, int n_comp);
  // the residual (FP64, wtd 0, 1, 2) with z-column lanes; its fix-up is
  // fixup_zcol_d3
  bool residual_zcol_d3(int n_grains, const KParams &kp, int ntz1, int n_tiles, int wtd,
                        const double *src, const double *o0, const double *o1, const double *o2,
                        double *dst, double *part, hipStream_t s);
} // namespace hps
