// The Q1 dispatcher and the fix-up kernel.
#include "ops.hpp"

namespace hps
{
  const OpTable *q1_d3s_a(int n_grains);
  const OpTable *q1_d3s_b(int n_grains);
  const OpTable *q1_d3s_c(int n_grains);
  const OpTable *q1_d3t_a(int n_grains);
  const OpTable *q1_d3t_b(int n_grains);
  const OpTable *q1_d3t_c(int n_grains);
  const OpTable *q1_d2s(int n_grains);
  const OpTable *q1_d2t(int n_grains);
  const OpTable *q1_table_d3_scalar_ty8(int n_grains);

  // every grain count 1..12 (EXPAND_OPERATIONS, S/instantiation.h:97-152)
  const OpTable *q1_table_d3(int n_grains, bool tensorial, int ty)
  {
    if (!tensorial && ty == 8)
      return q1_table_d3_scalar_ty8(n_grains);
    const OpTable *t = nullptr;
    if (tensorial)
      {
        if (!(t = q1_d3t_a(n_grains)) && !(t = q1_d3t_b(n_grains)))
          t = q1_d3t_c(n_grains);
      }
    else if (!(t = q1_d3s_a(n_grains)) && !(t = q1_d3s_b(n_grains)))
      t = q1_d3s_c(n_grains);
    return t;
  }

  const OpTable *q1_table_d2(int n_grains, bool tensorial)
  {
    return tensorial ? q1_d2t(n_grains) : q1_d2s(n_grains);
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
  template void launch_fixup<double>(const FixParams &, const double *, double *, hipStream_t);
  template void launch_fixup<float>(const FixParams &, const float *, float *, hipStream_t);
} // namespace hps
