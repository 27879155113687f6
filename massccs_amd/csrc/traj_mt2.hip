// Trajectory-kernel instantiations over several targets, N2 probe.
#include "kernels.hpp"

namespace mccs {
namespace dev {

TrajFn select_traj_multi2(int ft, int L, bool count) {
  switch (ft) {
    case 6: return pick_traj_multi<2, 6>(L, count);
  }
  return nullptr;
}

}  // namespace dev
}  // namespace mccs
