// gpu_utils.h -- drop-in for include/cuda_utils.cuh:11-41 (registry and the
// fatal error check), over the C ABI of libfwx.
#pragma once

#include <map>
#include <string>

#include "algo_types.h"

// Prints the libfwx error and exits with EXIT_FAILURE when status != 0
// (the reference's checkCuda behaviour, cuda_utils.cuh:13-21).
void checkFwx(int status, const char* call, const char* file, int line);
#define checkFwxErrors(call) checkFwx((call), #call, __FILE__, __LINE__)

extern const std::map<std::string, AlgorithmFuncGPU> GPU_ALGORITHMS;

// Throws std::out_of_range for an unknown name.
AlgorithmFuncGPU getGPUAlgorithmFunc(const std::string& algorithmName);
bool isGPUAlgorithm(const std::string& algorithmName);
