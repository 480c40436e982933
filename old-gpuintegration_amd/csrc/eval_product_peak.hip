// Instantiates the rule-evaluation kernels of the product_peak family for d = 2..10, for the
// degree-7/5 rule and the degree-9 rule, and its Phase-2 kernels.
#include "eval_kernels.hpp"
#include "phase2.hpp"

namespace cuhre {
EvalLauncher launcher_product_peak(int dim, int key) { return family_launcher<CUHRE_PRODUCT_PEAK>(dim, key); }
FusedLauncher fused_product_peak(int dim) { return family_fused_launcher<CUHRE_PRODUCT_PEAK>(dim); }
Phase2Launcher p2launcher_product_peak(int dim, int key) { return phase2_launcher<CUHRE_PRODUCT_PEAK>(dim, key); }
}  // namespace cuhre
