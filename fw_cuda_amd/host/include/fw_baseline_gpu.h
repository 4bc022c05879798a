// fw_baseline_gpu.h -- drop-in for include/fw_baseline_gpu.cuh:6.
#pragma once

#include "constants.h"

// One relaxation sweep per k on the GPU (libfwx fwx_baseline_gpu).
void fwBaselineGPU(WeightType* d_D, int n, int tileSize, int kappa);
