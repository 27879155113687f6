// Trajectory-kernel instantiations over several targets, N2 probe.
#include "kernels.hpp"

namespace mccs {
namespace dev {

TrajFn select_traj_multi2(int ft, int L, bool count, bool split) {
  if (split) return ft == 6 ? pick_traj_multi_split<2, 6>(L, count) : nullptr;
  switch (ft) {
    case 6: return pick_traj_multi<2, 6>(L, count);
  }
  return nullptr;
}

}  // namespace dev
}  // namespace mccs
