// fw_tiling_gpu.h -- drop-in for include/fw_tiling_gpu.cuh:6.
#pragma once

#include "constants.h"

// Three-phase blocked FW (libfwx fwx_tiling_gpu).  tileSize 32/64/128 is
// honoured; any other value (the reference ignores it) selects the tuned tile.
void fwTilingGPU(WeightType* d_D, int n, int tileSize, int kappa);
