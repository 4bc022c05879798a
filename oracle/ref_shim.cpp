// ref_shim.cpp -- C-linkage entry points into the reference's own CPU sources
// (src/fw_baseline_cpu.cpp, src/graph_generator.cpp, src/utils.cpp), compiled
// from where they lie by oracle/Makefile into oracle/_ref/libfwref.so.
// TEST INFRASTRUCTURE ONLY: used to pin the oracle and as the CPU baseline.
#include <cstring>
#include <vector>
#include "fw_baseline_cpu.h"
#include "graph_generator.h"
#include "utils.h"

extern "C" {

// fwBaselineCPU(WeightType* D, int n, int tileSize)   include/fw_baseline_cpu.h:6
void ref_fw_baseline_cpu(int* D, int n) { fwBaselineCPU(D, n, 0); }

// generateGraph1D(int n, double density, unsigned seed)   include/graph_generator.h:27
int ref_generate_graph(int* out, int n, double density, unsigned int seed) {
  std::vector<WeightType> v = generateGraph1D(n, density, seed);
  if (v.empty()) return -1;
  std::memcpy(out, v.data(), v.size() * sizeof(WeightType));
  return 0;
}

// hasNegativeCycle(const WeightType*, int)   include/utils.h:9
int ref_has_negative_cycle(const int* D, int n) { return hasNegativeCycle(D, n) ? 1 : 0; }

}
