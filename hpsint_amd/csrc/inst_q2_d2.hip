// The Q2 dispatcher.
#include "ops.hpp"

namespace hps
{
  const Q2Table *q2_d3s_a(int n_grains);
  const Q2Table *q2_d3s_b(int n_grains);
  const Q2Table *q2_d3s_c(int n_grains);
  const Q2Table *q2_d3t_a(int n_grains);
  const Q2Table *q2_d3t_b(int n_grains);
  const Q2Table *q2_d3t_c(int n_grains);
  const Q2Table *q2_d3t_d(int n_grains);
  const Q2Table *q2_d2s(int n_grains);
  const Q2Table *q2_d2t(int n_grains);

  // every grain count 1..12 (EXPAND_OPERATIONS, S/instantiation.h:97-152)
  const Q2Table *q2_table(int dim, int n_grains, bool tensorial)
  {
    if (dim == 2)
      return tensorial ? q2_d2t(n_grains) : q2_d2s(n_grains);
    const Q2Table *t = nullptr;
    if (tensorial)
      {
        if (!(t = q2_d3t_a(n_grains)) && !(t = q2_d3t_b(n_grains)) && !(t = q2_d3t_c(n_grains)))
          t = q2_d3t_d(n_grains);
      }
    else if (!(t = q2_d3s_a(n_grains)) && !(t = q2_d3s_b(n_grains)))
      t = q2_d3s_c(n_grains);
    return t;
  }
} // namespace hps
