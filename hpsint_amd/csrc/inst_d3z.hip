// Instantiations: 3D Q1 scalar-mobility tangent with z-column lanes, FP64
// with 1-4 order parameters, FP32 with 1-12; the residual with 1-4.
#include <cstdlib>

#include "fe_q1_zcol.hpp"
#include "ops.hpp"

namespace hps
{
  // waves per SIMD of the register allocation: HPSINT_AMD_ZCOL_WAVES (2, 3
  // or 4) or the default (FP64 2; FP32 3, or 2 with more than 4 order
  // parameters, whose allocation spills at 3)
  static int zcol_waves(bool fp64, int n_grains)
  {
    const char *e = std::getenv("HPSINT_AMD_ZCOL_WAVES");
    const int   w = e ? std::atoi(e) : 0;
    if (w >= 2 && w <= 4)
      return w;
    return fp64 || n_grains > 4 ? 2 : 3;
  }

  template <int G, typename T>
  static void zcol_kernel(int w, const KParams &kp, int ntz1, int n_tiles, const T *src,
                          const T *cache, T *dst, T *part, hipStream_t s)
  {
    const ZSlots zs = zcol_slots(kp, ntz1, G + 2);
    if (w == 2)
      q1_vmult_zcol<G, T, 2><<<n_tiles, zcol::NT, 0, s>>>(kp, zs, ntz1, src, cache, dst, part);
    else if (w == 3)
      q1_vmult_zcol<G, T, 3><<<n_tiles, zcol::NT, 0, s>>>(kp, zs, ntz1, src, cache, dst, part);
    else
      q1_vmult_zcol<G, T, 4><<<n_tiles, zcol::NT, 0, s>>>(kp, zs, ntz1, src, cache, dst, part);
  }

  // the residual (FP64) with z-column lanes, wtd 0, 1 or 2
  template <int G>
  static void zcol_residual(int wtd, const KParams &kp, int ntz1, int n_tiles, const double *src,
                            const double *o0, const double *o1, const double *o2, double *dst,
                            double *part, hipStream_t s)
  {
    const ZSlots zs = zcol_slots(kp, ntz1, G + 2);
    if (wtd == 0)
      q1_vmult_zcol<G, double, 2, OP_RES0>
        <<<n_tiles, zcol::NT, 0, s>>>(kp, zs, ntz1, src, nullptr, dst, part, o0, o1, o2);
    else if (wtd == 1)
      q1_vmult_zcol<G, double, 2, OP_RES1>
        <<<n_tiles, zcol::NT, 0, s>>>(kp, zs, ntz1, src, nullptr, dst, part, o0, o1, o2);
    else
      q1_vmult_zcol<G, double, 2, OP_RES2>
        <<<n_tiles, zcol::NT, 0, s>>>(kp, zs, ntz1, src, nullptr, dst, part, o0, o1, o2);
  }

  bool residual_zcol_d3(int n_grains, const KParams &kp, int ntz1, int n_tiles, int wtd,
                        const double *src, const double *o0, const double *o1, const double *o2,
                        double *dst, double *part, hipStream_t s)
  {
    switch (n_grains)
      {
#define HPS_CASE(G)                                                                  \
  case G:                                                                            \
    zcol_residual<G>(wtd, kp, ntz1, n_tiles, src, o0, o1, o2, dst, part, s);         \
    return true;
        HPS_CASE(1)
        HPS_CASE(2)
        HPS_CASE(3)
        HPS_CASE(4)
#undef HPS_CASE
        default:
          return false;
      }
  }

  // the kernel of G order parameters, if built for T
  template <int G, typename T>
  static bool zcol_try(int w, const KParams &kp, int ntz1, int n_tiles, const T *src,
                       const T *cache, T *dst, T *part, hipStream_t s)
  {
    if constexpr (G > (sizeof(T) == 8 ? ZCOL_MAX_G : ZCOL32_MAX_G))
      return false;
    else
      {
        zcol_kernel<G, T>(w, kp, ntz1, n_tiles, src, cache, dst, part, s);
        return true;
      }
  }

  template <typename T>
  static bool zcol_launch(int n_grains, const KParams &kp, int ntz1, int n_tiles, const T *src,
                          const T *cache, T *dst, T *part, hipStream_t s)
  {
    const int w = zcol_waves(sizeof(T) == 8, n_grains);
    switch (n_grains)
      {
#define HPS_CASE(G)                                                           \
  case G:                                                                     \
    return zcol_try<G, T>(w, kp, ntz1, n_tiles, src, cache, dst, part, s);
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
          return false;
      }
  }

  template <typename T>
  static void fixup_slots(const KParams &kp, int ntz1, int n_comp, const T *part, T *dst,
                          hipStream_t s)
  {
    const ZSlots  zs = zcol_slots(kp, ntz1, n_comp);
    const int64_t n  = zs.n_m + zs.n_y + zs.n_x; // one thread per shared node
    if (n == 0)
      return;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    switch (n_comp)
      {
#define HPS_CASE(NC)                                                     \
  case NC:                                                               \
    q1_fixup_slots<T, NC><<<blocks, 256, 0, s>>>(zs, part, dst);         \
    break;
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
        HPS_CASE(13)
        HPS_CASE(14)
#undef HPS_CASE
        default:
          break;
      }
  }

  int64_t zcol_part_elems(const KParams &kp, int ntz1, int n_comp)
  {
    return zcol_slot_records(zcol_slots(kp, ntz1, n_comp)) * n_comp;
  }

  void fixup_zcol_d3(const KParams &kp, int ntz1, int n_comp, const double *part, double *dst,
                     hipStream_t s)
  {
    fixup_slots<double>(kp, ntz1, n_comp, part, dst, s);
  }

  void fixup_zcol_d3(const KParams &kp, int ntz1, int n_comp, const float *part, float *dst,
                     hipStream_t s)
  {
    fixup_slots<float>(kp, ntz1, n_comp, part, dst, s);
  }

  bool vmult_zcol_d3(int n_grains, const KParams &kp, int ntz1, int n_tiles, const double *src,
                     const double *cache, double *dst, double *part, hipStream_t s)
  {
    return zcol_launch<double>(n_grains, kp, ntz1, n_tiles, src, cache, dst, part, s);
  }

  bool vmult_zcol_d3(int n_grains, const KParams &kp, int ntz1, int n_tiles, const float *src,
                     const float *cache, float *dst, float *part, hipStream_t s)
  {
    return zcol_launch<float>(n_grains, kp, ntz1, n_tiles, src, cache, dst, part, s);
  }
} // namespace hps
