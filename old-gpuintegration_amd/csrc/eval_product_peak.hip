// Instantiates the rule-evaluation kernels of the product_peak family for d = 2..10, for the
// degree-7/5 rule and the degree-9 rule.
#include "eval_kernels.hpp"

namespace cuhre {
EvalLauncher launcher_product_peak(int dim, int key) { return family_launcher<CUHRE_PRODUCT_PEAK>(dim, key); }
}  // namespace cuhre
