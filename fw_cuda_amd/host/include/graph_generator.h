// graph_generator.h -- drop-in for include/graph_generator.h:7-27.
#pragma once

#include <cstdlib>
#include <vector>

#include "constants.h"

#define ALIGNED_MALLOC(size, align) fw_aligned_malloc((size), (align))
#define ALIGNED_FREE(ptr) std::free(ptr)

void* fw_aligned_malloc(size_t size, size_t align);

// Seeded random graph (see fwx_generate_graph in include/fwx.h); weights
// uniform in [0, MAX_EDGE_WEIGHT] as in the reference, or [wmin, wmax].
// Caller frees with ALIGNED_FREE.  nullptr on bad n.
WeightType* generateAlignedGraph1D(int n, double density, unsigned int seed = 0);
WeightType* generateAlignedGraph1D(int n, double density, unsigned int seed, int wmin, int wmax);

std::vector<WeightType> generateGraph1D(int n, double density, unsigned int seed = 0);
std::vector<WeightType> generateGraph1D(int n, double density, unsigned int seed, int wmin, int wmax);
