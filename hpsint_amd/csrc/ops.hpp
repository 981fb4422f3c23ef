// ops.hpp -- launch table of the Q1 cell kernels, one entry per
// (dim, n_grains, mobility); filled by the inst_*.hip translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "fe_q1.hpp"

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
  };

  template <int DIM, int TX, int TY, int G, bool TENS>
  struct Q1Launch
  {
    static constexpr int NT = TX * TY * (1 << DIM);

    static void vmult64(const KParams &kp, int n_tiles, const double *src, double *dst,
                        double *part, const double *cache, hipStream_t s)
    {
      q1_apply<DIM, TX, TY, G, TENS, OP_VMULT, double>
        <<<n_tiles, NT, 0, s>>>(kp, src, nullptr, nullptr, nullptr, cache, dst, part);
    }
    static void vmult32(const KParams &kp, int n_tiles, const float *src, float *dst,
                        float *part, const float *cache, hipStream_t s)
    {
      q1_apply<DIM, TX, TY, G, TENS, OP_VMULT, float>
        <<<n_tiles, NT, 0, s>>>(kp, src, nullptr, nullptr, nullptr, cache, dst, part);
    }
    static void residual(const KParams &kp, int n_tiles, int wtd, const double *src,
                         const double *o0, const double *o1, const double *o2,
                         double *dst, double *part, hipStream_t s)
    {
      if (wtd == 0)
        q1_apply<DIM, TX, TY, G, TENS, OP_RES0, double>
          <<<n_tiles, NT, 0, s>>>(kp, src, o0, o1, o2, nullptr, dst, part);
      else if (wtd == 1)
        q1_apply<DIM, TX, TY, G, TENS, OP_RES1, double>
          <<<n_tiles, NT, 0, s>>>(kp, src, o0, o1, o2, nullptr, dst, part);
      else
        q1_apply<DIM, TX, TY, G, TENS, OP_RES2, double>
          <<<n_tiles, NT, 0, s>>>(kp, src, o0, o1, o2, nullptr, dst, part);
    }
    static void fill(const KParams &kp, int n_tiles, const double *lin, double *c64,
                     float *c32, hipStream_t s)
    {
      q1_fill<DIM, TX, TY, G, TENS><<<n_tiles, NT, 0, s>>>(kp, lin, c64, c32);
    }
    static void probe(const KParams &kp, int n_tiles, const double *cache, double *dst,
                      double *part, int column, int os, int ocs, int oo, hipStream_t s)
    {
      q1_probe<DIM, TX, TY, G, TENS>
        <<<n_tiles, NT, 0, s>>>(kp, cache, dst, part, column, os, ocs, oo);
    }
    static OpTable table()
    {
      return OpTable{NT,       TX,       TY,   CacheLayout<DIM, G, TENS>::n_fields,
                     &vmult64, &vmult32, &residual, &fill, &probe};
    }
  };

  // tile extents: 3D 8x4 cells (256 lanes), 2D 64 cells (256 lanes)
  constexpr int TX3 = 8, TY3 = 4;
  constexpr int TX2 = 64, TY2 = 1;

  // nullptr when (n_grains, mobility) is not built for that dimension
  // ty: tile extent along y (4, or 8 for scalar mobility)
  const OpTable *q1_table_d3(int n_grains, bool tensorial, int ty);
  const OpTable *q1_table_d2(int n_grains, bool tensorial);

  template <typename T>
  void launch_fixup(const FixParams &fp, const T *part, T *dst, hipStream_t s);
} // namespace hps
