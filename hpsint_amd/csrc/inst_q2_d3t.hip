// Instantiations: 3D Q2, tensorial mobility (10 order parameters: the
// 12-component Q2 workload).
#include "ops.hpp"

namespace hps
{
  const Q2Table *q2_table_d3_tensorial(int n_grains)
  {
#define HPS_CASE(G)                                                 \
  case G:                                                           \
    {                                                               \
      static const Q2Table t = Q2Launch<3, G, true>::table();       \
      return &t;                                                    \
    }
    switch (n_grains)
      {
        HPS_CASE(1)
        HPS_CASE(2)
        HPS_CASE(3)
        HPS_CASE(4)
        HPS_CASE(10)
        default:
          return nullptr;
      }
#undef HPS_CASE
  }
} // namespace hps
