// Instantiations: 2D Q1, scalar and tensorial mobility; the fix-up kernel.
#include "ops.hpp"

namespace hps
{
  const OpTable *q1_table_d3_scalar(int n_grains);
  const OpTable *q1_table_d3_scalar_ty8(int n_grains);
  const OpTable *q1_table_d3_tensorial(int n_grains);

  const OpTable *q1_table_d3(int n_grains, bool tensorial, int ty)
  {
    if (tensorial)
      return q1_table_d3_tensorial(n_grains);
    return ty == 8 ? q1_table_d3_scalar_ty8(n_grains) : q1_table_d3_scalar(n_grains);
  }

  const OpTable *q1_table_d2(int n_grains, bool tensorial)
  {
#define HPS_CASE(G, TENS)                                                  \
  case G:                                                                  \
    {                                                                      \
      static const OpTable t = Q1Launch<2, TX2, TY2, G, TENS>::table();    \
      return &t;                                                           \
    }
    if (!tensorial)
      switch (n_grains)
        {
          HPS_CASE(1, false)
          HPS_CASE(2, false)
          HPS_CASE(3, false)
          HPS_CASE(4, false)
          default:
            return nullptr;
        }
    switch (n_grains)
      {
        HPS_CASE(1, true)
        HPS_CASE(2, true)
        HPS_CASE(3, true)
        default:
          return nullptr;
      }
#undef HPS_CASE
  }

  template <typename T>
  void launch_fixup(const FixParams &fp, const T *part, T *dst, hipStream_t s)
  {
    const int64_t n = fp.n_m + fp.n_y + fp.n_x; // one thread per shared node
    if (n == 0)
      return;
    const int blocks = (int)((n + 255) / 256);
    // node-interleaved dst (the vmult and residual layout) of suitable
    // alignment: records stored as vectors
    auto with_vdst = [&](int bytes) {
      FixParams f = fp;
      f.vdst      = fp.ostride == fp.n_comp && fp.ocstride == 1 && fp.ooffset == 0 &&
               reinterpret_cast<uintptr_t>(dst) % bytes == 0;
      return f;
    };
    switch (fp.n_comp)
      {
#define HPS_CASE(NC)                                                               \
  case NC:                                                                         \
    q1_fixup<T, NC><<<blocks, 256, 0, s>>>(with_vdst(FixVec<T, NC>::bytes), part, dst); \
    break;
        HPS_CASE(3)
        HPS_CASE(4)
        HPS_CASE(5)
        HPS_CASE(6)
        HPS_CASE(7)
        HPS_CASE(8)
        HPS_CASE(10)
        HPS_CASE(12)
        HPS_CASE(14)
#undef HPS_CASE
        default:
          break;
      }
  }
  template void launch_fixup<double>(const FixParams &, const double *, double *, hipStream_t);
  template void launch_fixup<float>(const FixParams &, const float *, float *, hipStream_t);
} // namespace hps
