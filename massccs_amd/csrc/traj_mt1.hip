// Trajectory-kernel instantiations over several targets, atomic probes.
#include "kernels.hpp"

namespace mccs {
namespace dev {

TrajFn select_traj_multi1(int ft, int L, bool count, bool split) {
  if (split) return ft == 4 ? pick_traj_multi_split<1, 4>(L, count) : nullptr;
  switch (ft) {
    case 2: return pick_traj_multi<1, 2>(L, count);
    case 4: return pick_traj_multi<1, 4>(L, count);
  }
  return nullptr;
}

}  // namespace dev
}  // namespace mccs
