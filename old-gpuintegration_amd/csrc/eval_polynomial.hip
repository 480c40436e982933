// Instantiates the rule-evaluation kernels of the polynomial family for d = 2..10.
#include "eval_kernels.hpp"

namespace cuhre {
EvalLauncher launcher_polynomial(int dim) { return family_launcher<CUHRE_POLYNOMIAL>(dim); }
}  // namespace cuhre
