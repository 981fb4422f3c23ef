// Instantiations: 2D Q1, tensorial mobility, 1-12 order parameters.
#include "ops.hpp"

namespace hps
{
  // nullptr for grain counts outside this unit's range
  const OpTable *q1_d2t(int n_grains)
  {
#define HPS_CASE(G)                                                     \
  case G:                                                               \
    {                                                                   \
      static const OpTable t = Q1Launch<2, TX2, TY2, G, true>::table(); \
      return &t;                                                        \
    }
    switch (n_grains)
      {
        HPS_CASE(1)
        HPS_CASE(2)
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
        default:
          return nullptr;
      }
#undef HPS_CASE
  }
} // namespace hps
