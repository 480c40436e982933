// Instantiates the rule-evaluation kernels of the oscillatory family for d = 2..10.
#include "eval_kernels.hpp"

namespace cuhre {
EvalLauncher launcher_oscillatory(int dim) { return family_launcher<CUHRE_OSCILLATORY>(dim); }
}  // namespace cuhre
