// Instantiations: 3D Q2, scalar mobility.
#include "ops.hpp"

namespace hps
{
  const Q2Table *q2_table_d3_scalar(int n_grains)
  {
#define HPS_CASE(G)                                                 \
  case G:                                                           \
    {                                                               \
      static const Q2Table t = Q2Launch<3, G, false>::table();      \
      return &t;                                                    \
    }
    switch (n_grains)
      {
        HPS_CASE(1)
        HPS_CASE(2)
        HPS_CASE(3)
        HPS_CASE(4)
        default:
          return nullptr;
      }
#undef HPS_CASE
  }
} // namespace hps
