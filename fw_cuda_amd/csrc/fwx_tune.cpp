// fwx_tune.cpp -- (tile, kappa) chosen when a caller passes 0 for either: the
// first size class of the tuned table whose n_max covers n (the table is
// written by fw_cuda_amd/autotune.py from timings on the target GPU).
#include "fwx_internal.hpp"

namespace fwx {

namespace {
struct TuneEntry {
  int32_t n_max, tile, kappa;
};
const TuneEntry kTable[] = {
#include "fwx_tune_table.inc"
};
}  // namespace

void tuned_config(int32_t n, int32_t* tile, int32_t* kappa) {
  for (const TuneEntry& e : kTable) {
    if (n <= e.n_max) {
      *tile = e.tile;
      *kappa = e.kappa;
      return;
    }
  }
  const TuneEntry& last = kTable[sizeof(kTable) / sizeof(kTable[0]) - 1];
  *tile = last.tile;
  *kappa = last.kappa;
}

}  // namespace fwx
