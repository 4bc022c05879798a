// fwx_tune.cpp -- (tile, kappa) chosen when a caller passes 0 for either.
#include "fwx_internal.hpp"

namespace fwx {

void tuned_config(int32_t n, int32_t* tile, int32_t* kappa) {
  if (n <= 64) {
    *tile = 32;
    *kappa = 1;
  } else if (n <= 512) {
    *tile = 64;
    *kappa = 2;
  } else {
    *tile = 64;
    *kappa = 4;
  }
}

}  // namespace fwx
