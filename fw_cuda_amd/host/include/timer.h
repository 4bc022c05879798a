// timer.h -- drop-in for include/timer.cuh:6-30.
#pragma once

#include "algo_types.h"
#include "constants.h"

namespace Timer {
void warmupGPU();
// warmup_runs untimed calls, then one timed call (on the same, already
// processed matrix, as the reference does), seconds.
double measureCPU(AlgorithmFuncCPU func, WeightType* graph, int n, int tileSize, int warmup_runs = 1);
double measureGPU(AlgorithmFuncGPU func, WeightType* d_graph, int n, int tileSize, int kappa,
                  int warmup_runs = 1);
}  // namespace Timer
