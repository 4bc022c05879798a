// algorithm_runner.h -- drop-in for include/algorithm_runner.h:7-19.
#pragma once

#include <string>

#include "algo_types.h"
#include "constants.h"

namespace AlgorithmRunner {
int executeBenchmark(const std::string& algorithmName, int n, int tileSize, int kappa, double density,
                     unsigned int seed);
// Same, with the edge-weight range of the generator ([0, MAX_EDGE_WEIGHT] above).
int executeBenchmark(const std::string& algorithmName, int n, int tileSize, int kappa, double density,
                     unsigned int seed, int wmin, int wmax);
}  // namespace AlgorithmRunner
