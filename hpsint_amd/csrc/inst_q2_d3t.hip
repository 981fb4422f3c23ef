// Instantiations: 3D Q2, tensorial mobility, 1-4 order parameters.
#include "ops.hpp"

namespace hps
{
  // nullptr for grain counts outside this unit's range
  const Q2Table *q2_d3t_a(int n_grains)
  {
#define HPS_CASE(G)                                                     \
  case G:                                                               \
    {                                                                   \
      static const Q2Table t = Q2Launch<3, G, true>::table(); \
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
