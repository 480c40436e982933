// Instantiates the rule-evaluation kernels of the cosine_sum family for d = 2..10.
#include "eval_kernels.hpp"

namespace cuhre {
EvalLauncher launcher_cosine_sum(int dim) { return family_launcher<CUHRE_COSINE_SUM>(dim); }
}  // namespace cuhre
