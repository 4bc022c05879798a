// tests.h -- drop-in for include/tests.h:7-12.  Both return the number of
// failed tests (the reference returns void).
#pragma once

#include <string>

#include "algo_types.h"

int runTestsCPU(AlgorithmFuncCPU algorithm, AlgorithmFuncCPU reference, const std::string& name,
                int tileSize);
int runTestsGPU(AlgorithmFuncGPU algorithm, AlgorithmFuncCPU reference, const std::string& name,
                int tileSize, int kappa);
