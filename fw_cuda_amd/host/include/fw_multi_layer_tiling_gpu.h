// fw_multi_layer_tiling_gpu.h -- drop-in for include/fw_multi_layer_tiling_gpu.cuh:17.
#pragma once

#include "constants.h"

// Multi-layer blocked FW, kappa k-tiles per stage (libfwx
// fwx_multi_layer_tiling_gpu).  tileSize 32/64/128 and kappa 1..16 are
// honoured; other values select the tuned ones (the reference falls back to
// its defaults the same way, src/fw_multi_layer_tiling_gpu.cu:508-513).
void fwMultiLayerTilingGPU(WeightType* d_D, int n, int tileSize, int kappa);
