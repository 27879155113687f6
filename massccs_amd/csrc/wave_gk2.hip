// Lane-per-trajectory kernel instantiations for the N2 probe.
#include "wave.hpp"

namespace mccs {
namespace dev {

WaveFn select_wave_gk2(int ft, bool lds, bool count) {
  switch (ft) {
    case 6: return pick_wave<2, 6>(lds, count);
    case 1: return lds ? wave_kernel<2, 1, true, true> : wave_kernel<2, 1, false, true>;
    case 2: return lds ? wave_kernel<2, 2, true, true> : wave_kernel<2, 2, false, true>;
    case 3: return lds ? wave_kernel<2, 3, true, true> : wave_kernel<2, 3, false, true>;
    case 4: return lds ? wave_kernel<2, 4, true, true> : wave_kernel<2, 4, false, true>;
    case 5: return lds ? wave_kernel<2, 5, true, true> : wave_kernel<2, 5, false, true>;
  }
  return nullptr;
}

}  // namespace dev
}  // namespace mccs
