// Instantiations: 3D Q1, scalar mobility, 9-12 order parameters.
#include "ops.hpp"

namespace hps
{
  // nullptr for grain counts outside this unit's range
  const OpTable *q1_d3s_c(int n_grains)
  {
#define HPS_CASE(G)                                                     \
  case G:                                                               \
    {                                                                   \
      static const OpTable t = Q1Launch<3, TX3, TY3, G, false>::table(); \
      return &t;                                                        \
    }
    switch (n_grains)
      {
        HPS_CASE(9)
        HPS_CASE(10)
        HPS_CASE(11)
        HPS_CASE(12)
        default:
          return nullptr;
      }
#undef HPS_CASE
  }
} // namespace hps
