// gpu_memory.h -- drop-in for include/gpu_memory.cuh:5-20.
#pragma once

#include "constants.h"

namespace GPUMemory {
WeightType* allocate(int n);
void copyToDevice(WeightType* d_graph, const WeightType* h_graph, int n);
void copyToHost(WeightType* h_graph, const WeightType* d_graph, int n);
void free(WeightType* d_graph);
}  // namespace GPUMemory
