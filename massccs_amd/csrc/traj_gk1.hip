// Trajectory-kernel instantiations for atomic probes (He, Ar, co2).
#include "kernels.hpp"

namespace mccs {
namespace dev {

TrajFn select_traj_gk1(int ft, int L, bool lds, bool count, bool split) {
  if (split) return ft == 4 ? pick_traj_split<1, 4>(L, lds, count) : nullptr;
  switch (ft) {
    case 2: return pick_traj_full<1, 2>(L, lds, count);
    case 4: return pick_traj_full<1, 4>(L, lds, count);
    case 1: return pick_traj_basic<1, 1>(L, lds, count);
    case 3: return pick_traj_basic<1, 3>(L, lds, count);
  }
  return nullptr;
}

}  // namespace dev
}  // namespace mccs
