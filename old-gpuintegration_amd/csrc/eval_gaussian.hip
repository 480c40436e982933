// Instantiates the rule-evaluation kernels of the gaussian family for d = 2..10.
#include "eval_kernels.hpp"

namespace cuhre {
EvalLauncher launcher_gaussian(int dim) { return family_launcher<CUHRE_GAUSSIAN>(dim); }
}  // namespace cuhre
