// Instantiates the rule-evaluation kernels of the polynomial family for d = 2..10, for the
// degree-7/5 rule and the degree-9 rule, and its Phase-2 kernels.
#include "eval_kernels.hpp"
#include "phase2.hpp"

namespace cuhre {
EvalLauncher launcher_polynomial(int dim, int key) { return family_launcher<CUHRE_POLYNOMIAL>(dim, key); }
FusedLauncher fused_polynomial(int dim) { return family_fused_launcher<CUHRE_POLYNOMIAL>(dim); }
Phase2Launcher p2launcher_polynomial(int dim, int key) { return phase2_launcher<CUHRE_POLYNOMIAL>(dim, key); }
}  // namespace cuhre
