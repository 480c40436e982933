// Instantiates the rule-evaluation kernels of the product_peak family for d = 2..10.
#include "eval_kernels.hpp"

namespace cuhre {
EvalLauncher launcher_product_peak(int dim) { return family_launcher<CUHRE_PRODUCT_PEAK>(dim); }
}  // namespace cuhre
