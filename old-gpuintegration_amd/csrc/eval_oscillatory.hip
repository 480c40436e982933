// Instantiates the rule-evaluation kernels of the oscillatory family for d = 2..10, for the
// degree-7/5 rule and the degree-9 rule.
#include "eval_kernels.hpp"

namespace cuhre {
EvalLauncher launcher_oscillatory(int dim, int key) { return family_launcher<CUHRE_OSCILLATORY>(dim, key); }
}  // namespace cuhre
