// Trajectory-kernel instantiations for the CO2 probe.
#include "kernels.hpp"

namespace mccs {
namespace dev {

TrajFn select_traj_gk3(int ft, int L, bool lds, bool count) {
  switch (ft) {
    case 6: return pick_traj_full<3, 6>(L, lds, count);
    case 1: return pick_traj_basic<3, 1>(L, lds, count);
    case 2: return pick_traj_basic<3, 2>(L, lds, count);
    case 3: return pick_traj_basic<3, 3>(L, lds, count);
    case 4: return pick_traj_basic<3, 4>(L, lds, count);
    case 5: return pick_traj_basic<3, 5>(L, lds, count);
  }
  return nullptr;
}

}  // namespace dev
}  // namespace mccs
