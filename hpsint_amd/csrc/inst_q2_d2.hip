// Instantiations: 2D Q2, scalar and tensorial mobility; the Q2 dispatcher.
#include "ops.hpp"

namespace hps
{
  const Q2Table *q2_table_d3_scalar(int n_grains);
  const Q2Table *q2_table_d3_tensorial(int n_grains);

  static const Q2Table *q2_table_d2(int n_grains, bool tensorial)
  {
#define HPS_CASE(G, TENS)                                           \
  case G:                                                           \
    {                                                               \
      static const Q2Table t = Q2Launch<2, G, TENS>::table();       \
      return &t;                                                    \
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

  const Q2Table *q2_table(int dim, int n_grains, bool tensorial)
  {
    if (dim == 2)
      return q2_table_d2(n_grains, tensorial);
    return tensorial ? q2_table_d3_tensorial(n_grains) : q2_table_d3_scalar(n_grains);
  }
} // namespace hps
