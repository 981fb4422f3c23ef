// Instantiations: 3D Q1, tensorial mobility, 5-8 order parameters.
#include "ops.hpp"

namespace hps
{
  // nullptr for grain counts outside this unit's range
  const OpTable *q1_d3t_b(int n_grains)
  {
#define HPS_CASE(G)                                                     \
  case G:                                                               \
    {                                                                   \
      static const OpTable t = Q1Launch<3, TX3, TY3, G, true>::table(); \
      return &t;                                                        \
    }
    switch (n_grains)
      {
        HPS_CASE(5)
        HPS_CASE(6)
        HPS_CASE(7)
        HPS_CASE(8)
        default:
          return nullptr;
      }
#undef HPS_CASE
  }
} // namespace hps
