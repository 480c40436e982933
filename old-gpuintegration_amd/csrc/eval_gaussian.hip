// Instantiates the rule-evaluation kernels of the gaussian family for d = 2..10, for the
// degree-7/5 rule and the degree-9 rule.
#include "eval_kernels.hpp"

namespace cuhre {
EvalLauncher launcher_gaussian(int dim, int key) { return family_launcher<CUHRE_GAUSSIAN>(dim, key); }
}  // namespace cuhre
