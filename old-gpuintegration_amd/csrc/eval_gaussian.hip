// Instantiates the rule-evaluation kernels of the gaussian family for d = 2..10, for the
// degree-7/5 rule and the degree-9 rule, and its Phase-2 kernels.
#include "eval_kernels.hpp"
#include "phase2.hpp"

namespace cuhre {
EvalLauncher launcher_gaussian(int dim, int key) { return family_launcher<CUHRE_GAUSSIAN>(dim, key); }
FusedLauncher fused_gaussian(int dim) { return family_fused_launcher<CUHRE_GAUSSIAN>(dim); }
Phase2Launcher p2launcher_gaussian(int dim, int key) { return phase2_launcher<CUHRE_GAUSSIAN>(dim, key); }
}  // namespace cuhre
