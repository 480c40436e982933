// Instantiates the rule-evaluation kernels of the 6-D table-driven family.
// The table is read through the cache: staging each block's footprint of it
// in LDS (20 KiB tiles) was measured slower on the 6-D workload (2.46 against
// 2.17 ms of rule evaluation at epsrel 1e-5): the tile cost occupancy and the
// footprint pass barriers, and the reads from L2 were not the limit.
#include "eval_kernels.hpp"
#include "phase2.hpp"

namespace cuhre {
EvalLauncher launcher_table6(int dim, int key)
{
  if (dim != 6) return nullptr;
  return key == 9 ? &launch_eval9<6, CUHRE_TABLE6> : &launch_eval<6, CUHRE_TABLE6>;
}
FusedLauncher fused_table6(int dim) { return dim == 6 ? &launch_fused<6, CUHRE_TABLE6> : nullptr; }
Phase2Launcher p2launcher_table6(int dim, int key)
{
  if (dim != 6) return nullptr;
  return key == 9 ? &launch_phase2<6, CUHRE_TABLE6, 9> : &launch_phase2<6, CUHRE_TABLE6, 7>;
}
}  // namespace cuhre
