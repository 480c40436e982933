// Instantiates the rule-evaluation kernels of the 6-D table-driven family.
#include "eval_kernels.hpp"

namespace cuhre {
EvalLauncher launcher_table6(int dim, int lanes)
{
  return dim == 6 ? pick_lanes<CUHRE_TABLE6, 6>(lanes) : nullptr;
}
}  // namespace cuhre
