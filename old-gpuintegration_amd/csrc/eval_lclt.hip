// Instantiates the rule-evaluation kernels of the reference's LC_LT_t
// integrand (three dimensions: lt, zt, lc).
#include "eval_kernels.hpp"
#include "phase2.hpp"

namespace cuhre {
EvalLauncher launcher_lclt(int dim, int key)
{
  if (dim != 3) return nullptr;
  return key == 9 ? &launch_eval9<3, CUHRE_LC_LT_T> : &launch_eval<3, CUHRE_LC_LT_T>;
}
Phase2Launcher p2launcher_lclt(int dim, int key)
{
  if (dim != 3) return nullptr;
  return key == 9 ? &launch_phase2<3, CUHRE_LC_LT_T, 9> : &launch_phase2<3, CUHRE_LC_LT_T, 7>;
}
}  // namespace cuhre
