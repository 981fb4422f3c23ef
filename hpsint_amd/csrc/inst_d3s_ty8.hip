// Instantiations: 3D Q1, scalar mobility, 8x8-cell tiles.
#include "ops.hpp"

namespace hps
{
  const OpTable *q1_table_d3_scalar_ty8(int n_grains)
  {
#define HPS_CASE(G)                                                   \
  case G:                                                             \
    {                                                                 \
      static const OpTable t = Q1Launch<3, TX3, 8, G, false>::table(); \
      return &t;                                                      \
    }
    switch (n_grains)
      {
        HPS_CASE(1)
        HPS_CASE(2)
        HPS_CASE(3)
        HPS_CASE(4)
        HPS_CASE(5)
        HPS_CASE(6)
        HPS_CASE(8)
        HPS_CASE(10)
        HPS_CASE(12)
        default:
          return nullptr;
      }
#undef HPS_CASE
  }
} // namespace hps
