// Instantiates the rule-evaluation kernels of the corner_peak family for d = 2..10.
#include "eval_kernels.hpp"

namespace cuhre {
EvalLauncher launcher_corner_peak(int dim) { return family_launcher<CUHRE_CORNER_PEAK>(dim); }
}  // namespace cuhre
