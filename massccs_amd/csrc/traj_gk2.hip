// Trajectory-kernel instantiations for the N2 probe.
#include "kernels.hpp"

namespace mccs {
namespace dev {

TrajFn select_traj_gk2(int ft, int L, bool lds, bool count, bool split) {
  if (split) return ft == 6 ? pick_traj_split<2, 6>(L, lds, count) : nullptr;
  switch (ft) {
    case 6: return pick_traj_full<2, 6>(L, lds, count);
    case 1: return pick_traj_basic<2, 1>(L, lds, count);
    case 2: return pick_traj_basic<2, 2>(L, lds, count);
    case 3: return pick_traj_basic<2, 3>(L, lds, count);
    case 4: return pick_traj_basic<2, 4>(L, lds, count);
    case 5: return pick_traj_basic<2, 5>(L, lds, count);
  }
  return nullptr;
}

}  // namespace dev
}  // namespace mccs
