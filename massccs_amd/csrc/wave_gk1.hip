// Lane-per-trajectory kernel instantiations for atomic probes (He, Ar, co2).
#include "wave.hpp"

namespace mccs {
namespace dev {

WaveFn select_wave_gk1(int ft, bool lds, bool count) {
  switch (ft) {
    case 2: return pick_wave<1, 2>(lds, count);
    case 4: return pick_wave<1, 4>(lds, count);
    case 1: return lds ? wave_kernel<1, 1, true, true> : wave_kernel<1, 1, false, true>;
    case 3: return lds ? wave_kernel<1, 3, true, true> : wave_kernel<1, 3, false, true>;
  }
  return nullptr;
}

}  // namespace dev
}  // namespace mccs
