// Instantiates the rule-evaluation kernels of the continuous family for d = 2..10.
#include "eval_kernels.hpp"

namespace cuhre {
EvalLauncher launcher_continuous(int dim) { return family_launcher<CUHRE_CONTINUOUS>(dim); }
}  // namespace cuhre
