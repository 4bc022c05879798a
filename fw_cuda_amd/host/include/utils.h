// utils.h -- drop-in for include/utils.h:9-15.
#pragma once

#include <map>
#include <string>

#include "algo_types.h"
#include "constants.h"

bool hasNegativeCycle(const WeightType* distMatrix, int n);

extern const std::map<std::string, AlgorithmFuncCPU> CPU_ALGORITHMS;

// Throws std::out_of_range for an unknown name.
AlgorithmFuncCPU getCPUAlgorithmFunc(const std::string& algorithmName);
