// algo_types.h -- drop-in for the reference's include/algo_types.h:6-9.
#pragma once

#include "constants.h"

// CPU algorithm: in place on a host matrix.
using AlgorithmFuncCPU = void (*)(WeightType* D, int n, int tileSize);
// GPU algorithm: in place on a device matrix; returns after the work completed.
using AlgorithmFuncGPU = void (*)(WeightType* d_D, int n, int tileSize, int kappa);
