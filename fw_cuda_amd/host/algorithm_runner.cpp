// algorithm_runner.cpp -- AlgorithmRunner::executeBenchmark with the
// reference's flow and output lines (src/algorithm_runner.cu:12-91), plus one
// machine-readable JSON line after "Execution time".
#include "algorithm_runner.h"

#include <iomanip>
#include <iostream>

#include "gpu_memory.h"
#include "gpu_utils.h"
#include "graph_generator.h"
#include "timer.h"
#include "utils.h"

namespace AlgorithmRunner {

int executeBenchmark(const std::string& algorithmName, int n, int tileSize, int kappa, double density,
                     unsigned int seed, int wmin, int wmax) {
  if (n <= 0 || n > MAX_N) {
    std::cerr << "ERROR: Graph size (n) must be between 1 and " << MAX_N << ". Aborting.\n";
    return 1;
  }
  const bool isGPU = isGPUAlgorithm(algorithmName);

  std::cout << "\n========================================\n";
  std::cout << "BENCHMARK: " << algorithmName << "\n";
  std::cout << "========================================\n";
  std::cout << "Graph size: " << n << " vertices\n";
  std::cout << "Edge density: " << std::fixed << std::setprecision(1) << (density * 100.0) << "%\n";
  std::cout << "Tile size: " << tileSize << "\n";
  std::cout << "Kappa: " << kappa << "\n";
  std::cout << "Algorithm type: " << (isGPU ? "GPU" : "CPU") << "\n";
  if (seed == 0)
    std::cout << "Seed: Random (Non-reproducible)\n";
  else
    std::cout << "Seed: " << seed << " (Reproducible)\n";
  std::cout << "\n";

  std::cout << "Generating graph...\n";
  WeightType* h_graph = generateAlignedGraph1D(n, density, seed, wmin, wmax);
  if (h_graph == nullptr) {
    std::cerr << "ERROR: Failed to generate graph.\n";
    return 1;
  }

  double elapsed = 0.0;
  WeightType* d_graph = nullptr;
  try {
    if (isGPU) {
      AlgorithmFuncGPU gpuFunc = getGPUAlgorithmFunc(algorithmName);
      std::cout << "Warming up GPU...\n";
      Timer::warmupGPU();
      std::cout << "Copying graph to GPU...\n";
      d_graph = GPUMemory::allocate(n);
      GPUMemory::copyToDevice(d_graph, h_graph, n);
      std::cout << "Running " << algorithmName << "...\n";
      elapsed = Timer::measureGPU(gpuFunc, d_graph, n, tileSize, kappa);
      GPUMemory::copyToHost(h_graph, d_graph, n);
      GPUMemory::free(d_graph);
      d_graph = nullptr;
    } else {
      AlgorithmFuncCPU cpuFunc = getCPUAlgorithmFunc(algorithmName);
      std::cout << "Running " << algorithmName << "...\n";
      elapsed = Timer::measureCPU(cpuFunc, h_graph, n, tileSize);
    }
  } catch (const std::exception&) {
    if (d_graph != nullptr) GPUMemory::free(d_graph);
    ALIGNED_FREE(h_graph);
    throw;
  }

  std::cout << "\nResults:\n";
  std::cout << "--------\n";
  std::cout << "Execution time: " << std::fixed << std::setprecision(6) << elapsed << " seconds\n";
  const double n3 = (double)n * (double)n * (double)n;
  std::cout << std::setprecision(6) << std::scientific << "{\"algorithm\": \"" << algorithmName
            << "\", \"n\": " << n << ", \"tile\": " << tileSize << ", \"kappa\": " << kappa
            << ", \"seconds\": " << elapsed << ", \"n3_per_s\": " << (elapsed > 0 ? n3 / elapsed : 0.0)
            << "}\n"
            << std::defaultfloat;

  ALIGNED_FREE(h_graph);
  return 0;
}

int executeBenchmark(const std::string& algorithmName, int n, int tileSize, int kappa, double density,
                     unsigned int seed) {
  return executeBenchmark(algorithmName, n, tileSize, kappa, density, seed, 0, MAX_EDGE_WEIGHT);
}

}  // namespace AlgorithmRunner
