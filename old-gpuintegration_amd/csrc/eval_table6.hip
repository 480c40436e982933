// Instantiates the rule-evaluation kernels of the 6-D table-driven family.
#include "eval_kernels.hpp"

namespace cuhre {
EvalLauncher launcher_table6(int dim) { return dim == 6 ? &launch_eval<6, CUHRE_TABLE6> : nullptr; }
}  // namespace cuhre
