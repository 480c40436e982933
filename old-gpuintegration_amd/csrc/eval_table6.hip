// Instantiates the rule-evaluation kernels of the 6-D table-driven family.
#include "eval_kernels.hpp"
#include "phase2.hpp"

namespace cuhre {
EvalLauncher launcher_table6(int dim, int key)
{
  if (dim != 6) return nullptr;
  return key == 9 ? &launch_eval9<6, CUHRE_TABLE6> : &launch_eval<6, CUHRE_TABLE6>;
}
Phase2Launcher p2launcher_table6(int dim, int key)
{
  if (dim != 6) return nullptr;
  return key == 9 ? &launch_phase2<6, CUHRE_TABLE6, 9> : &launch_phase2<6, CUHRE_TABLE6, 7>;
}
}  // namespace cuhre
