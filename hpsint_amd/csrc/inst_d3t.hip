// Instantiations: 3D Q1, tensorial mobility, 1-4 order parameters.
#include "ops.hpp"

namespace hps
{
  // nullptr for grain counts outside this unit's range
  const OpTable *q1_d3t_a(int n_grains)
  {
#define HPS_CASE(G)                                                     \
  case G:                                                               \
    {                                                                   \
      static const OpTable t = Q1Launch<3, TX3, TY3, G, true>::table(); \
      return &t;                                                        \
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
