// Grain cut-off kernels: per-cell masks (Q1, 2D and 3D); the fast tangent
// vmult (3D Q1, scalar mobility, tiles of TX3 x TY3 cells, up to CUT_MAX_GA
// active order parameters per tile).
#include "fe_q1_cut.hpp"
#include "ops.hpp"

namespace hps
{
  void cell_masks(const KParams &kp, int dim, int n_grains, double tol, const double *u,
                  unsigned *masks, hipStream_t s)
  {
    // KParams: nc = {x, y (1 in 2D), march direction}
    const int     ncx = kp.nc[0];
    const int     ncy = dim == 3 ? kp.nc[1] : kp.nc[2];
    const int64_t n   = (int64_t)kp.nc[0] * kp.nc[1] * kp.nc[2];
    const int     nnx = kp.nn[0];
    const int     nny = dim == 3 ? kp.nn[1] : kp.nn[2];
    k_cell_masks<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(dim, n, ncx, ncy, nnx, nny,
                                                             n_grains, tol, u, masks);
  }

  void tile_masks_d3(const KParams &kp, int n_tiles, unsigned *masks, hipStream_t s)
  {
    k_tile_masks<3, TX3, TY3><<<n_tiles, TX3 * TY3 * 8, 0, s>>>(kp, masks);
  }

  template <typename T>
  static bool vmult_cut(int ga, const KParams &kp, int n_grains, const int *tiles, int count,
                        const unsigned *masks, const T *src, const T *cache, T *dst, T *part,
                        hipStream_t s)
  {
    constexpr int NT = TX3 * TY3 * 8;
#define HPS_CASE(GA)                                                                       \
  case GA:                                                                                 \
    q1_vmult_cut<3, TX3, TY3, GA, T>                                                       \
      <<<count, NT, 0, s>>>(kp, n_grains, tiles, masks, src, cache, dst, part);            \
    return true;
    switch (ga)
      {
        HPS_CASE(0)
        HPS_CASE(1)
        HPS_CASE(2)
        HPS_CASE(3)
        HPS_CASE(4)
        HPS_CASE(5)
        HPS_CASE(6)
        HPS_CASE(7)
        HPS_CASE(8)
        default:
          return false;
      }
#undef HPS_CASE
  }

  bool vmult_cut_d3(int ga, const KParams &kp, int n_grains, const int *tiles, int count,
                    const unsigned *masks, const double *src, const double *cache, double *dst,
                    double *part, hipStream_t s)
  {
    return vmult_cut<double>(ga, kp, n_grains, tiles, count, masks, src, cache, dst, part, s);
  }
  bool vmult_cut_d3(int ga, const KParams &kp, int n_grains, const int *tiles, int count,
                    const unsigned *masks, const float *src, const float *cache, float *dst,
                    float *part, hipStream_t s)
  {
    return vmult_cut<float>(ga, kp, n_grains, tiles, count, masks, src, cache, dst, part, s);
  }
} // namespace hps
