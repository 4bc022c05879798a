// tests.cpp -- the reference's test harness (src/tests.cpp:10-164,
// src/tests_gpu.cu:13-185, src/tests_runner.cu:12-41): six known-answer
// tests plus seeded graphs checked bit-exact against fwBaselineCPU, with the
// same [ PASSED ] / [ FAILED ] / RESULTS lines.  Returns the failure count.
#include "tests.h"

#include <cmath>
#include <iostream>
#include <vector>

#include "fw_baseline_cpu.h"
#include "gpu_memory.h"
#include "gpu_utils.h"
#include "graph_generator.h"
#include "tests_runner.h"
#include "timer.h"
#include "utils.h"

namespace {

struct Tally {
  int passed = 0;
  int total = 0;
};

using Runner = std::vector<WeightType> (*)(const void* algo, const std::vector<WeightType>& in, int n,
                                          int tileSize, int kappa);

void assertEqual(Tally& t, const std::vector<WeightType>& actual, const std::vector<WeightType>& expected,
                 const std::string& testName) {
  t.total++;
  if (actual.size() != expected.size()) {
    std::cerr << "[ FAILED ] " << testName << ": Matrix size mismatch (Expected: " << expected.size()
              << ", Actual: " << actual.size() << ")\n";
    return;
  }
  const int n = (int)std::lround(std::sqrt((double)actual.size()));
  for (size_t i = 0; i < actual.size(); ++i) {
    if (actual[i] != expected[i]) {
      std::cerr << "[ FAILED ] " << testName << ": Mismatch at position [" << i / n << "][" << i % n
                << "]. Expected: " << expected[i] << ", Actual: " << actual[i] << "\n";
      return;
    }
  }
  std::cout << "[ PASSED ] " << testName << "\n";
  t.passed++;
}

std::vector<WeightType> runCPU(const void* algo, const std::vector<WeightType>& in, int n, int tileSize, int) {
  std::vector<WeightType> r = in;
  reinterpret_cast<AlgorithmFuncCPU>(const_cast<void*>(algo))(r.data(), n, tileSize);
  return r;
}

std::vector<WeightType> runGPU(const void* algo, const std::vector<WeightType>& in, int n, int tileSize,
                               int kappa) {
  WeightType* d = GPUMemory::allocate(n);
  GPUMemory::copyToDevice(d, in.data(), n);
  reinterpret_cast<AlgorithmFuncGPU>(const_cast<void*>(algo))(d, n, tileSize, kappa);
  std::vector<WeightType> r(in.size());
  GPUMemory::copyToHost(r.data(), d, n);
  GPUMemory::free(d);
  return r;
}

struct Suite {
  Runner run;
  const void* algo;
  AlgorithmFuncCPU reference;
  int tileSize, kappa;
  Tally t;

  void kat(const std::vector<WeightType>& in, const std::vector<WeightType>& expected, int n,
           const std::string& name) {
    assertEqual(t, run(algo, in, n, tileSize, kappa), expected, name);
  }

  void negativeCycle() {
    t.total++;
    const std::vector<WeightType> in = {0, 2, INF, INF, 0, 1, INF, -3, 0};
    const std::vector<WeightType> r = run(algo, in, 3, tileSize, kappa);
    if (hasNegativeCycle(r.data(), 3)) {
      std::cout << "[ PASSED ] Negative cycle detection (3x3)\n";
      t.passed++;
    } else {
      std::cerr << "[ FAILED ] Negative cycle detection (3x3)\n";
    }
  }

  void againstReference(int n, double density, unsigned seed, const std::string& name) {
    const std::vector<WeightType> in = generateGraph1D(n, density, seed);
    std::vector<WeightType> expected = in;
    reference(expected.data(), n, 0);
    assertEqual(t, run(algo, in, n, tileSize, kappa), expected, name);
  }

  void knownAnswers() {
    kat({0}, {0}, 1, "Single node (1x1)");
    kat({0, INF, INF, 0}, {0, INF, INF, 0}, 2, "No edges (2x2)");
    kat({0, 1, INF, 100, INF, 0, 1, INF, INF, INF, 0, 1, INF, INF, INF, 0},
        {0, 1, 2, 3, INF, 0, 1, 2, INF, INF, 0, 1, INF, INF, INF, 0}, 4, "Simple path optimization (4x4)");
    const std::vector<WeightType> disc = {0, 1, INF, INF, 1, 0, INF, INF, INF, INF, 0, 2, INF, INF, 2, 0};
    kat(disc, disc, 4, "Disconnected components (4x4)");
    kat({0, 5, 10, 0, 0, -1, INF, INF, 0}, {0, 5, 4, 0, 0, -1, INF, INF, 0}, 3, "Negative weights (3x3)");
    negativeCycle();
  }

  int results() {
    std::cout << "========================================\n";
    std::cout << "RESULTS: " << t.passed << "/" << t.total << " passed\n";
    std::cout << "========================================\n\n";
    return t.total - t.passed;
  }
};

}  // namespace

int runTestsCPU(AlgorithmFuncCPU algorithm, AlgorithmFuncCPU reference, const std::string& name, int tileSize) {
  Suite s{runCPU, reinterpret_cast<const void*>(algorithm), reference, tileSize, 0, {}};
  std::cout << "\n========================================\n";
  std::cout << "TESTING: " << name << " (tileSize=" << tileSize << ")\n";
  std::cout << "========================================\n";
  s.knownAnswers();
  const int sizes[4] = {128, 256, 512, 1024};
  const unsigned seeds[4] = {42, 123, 456, 789};
  for (int i = 0; i < 4; ++i)
    s.againstReference(sizes[i], 0.9, seeds[i],
                       "Large graph (N=" + std::to_string(sizes[i]) + ", seed=" + std::to_string(seeds[i]) + ")");
  return s.results();
}

int runTestsGPU(AlgorithmFuncGPU algorithm, AlgorithmFuncCPU reference, const std::string& name, int tileSize,
                int kappa) {
  Suite s{runGPU, reinterpret_cast<const void*>(algorithm), reference, tileSize, kappa, {}};
  std::cout << "\n========================================\n";
  std::cout << "TESTING (GPU): " << name << " (tileSize=" << tileSize << ", kappa=" << kappa << ")\n";
  std::cout << "========================================\n";
  s.knownAnswers();
  const int sizes[8] = {128, 256, 512, 1024, 128, 256, 512, 1024};
  const double dens[8] = {0.9, 0.9, 0.9, 0.9, 0.1, 0.1, 0.1, 0.1};
  const unsigned seeds[8] = {42, 123, 456, 789, 101, 202, 303, 404};
  for (int i = 0; i < 8; ++i)
    s.againstReference(sizes[i], dens[i], seeds[i],
                       "Large graph (N=" + std::to_string(sizes[i]) + ", density=" + std::to_string(dens[i]) +
                           ", seed=" + std::to_string(seeds[i]) + ")");
  return s.results();
}

namespace TestsRunner {

int executeTests(const std::string& algorithmName, int tileSize, int kappa) {
  AlgorithmFuncCPU reference = fwBaselineCPU;
  int failed = 0;
  if (isGPUAlgorithm(algorithmName)) {
    AlgorithmFuncGPU algorithm = nullptr;
    try {
      algorithm = getGPUAlgorithmFunc(algorithmName);
    } catch (const std::out_of_range& e) {
      std::cerr << "ERROR: " << e.what() << "\n";
      return 1;
    }
    std::cout << "Warming up GPU...\n";
    Timer::warmupGPU();
    failed = runTestsGPU(algorithm, reference, algorithmName, tileSize, kappa);
  } else {
    AlgorithmFuncCPU algorithm = nullptr;
    try {
      algorithm = getCPUAlgorithmFunc(algorithmName);
    } catch (const std::out_of_range& e) {
      std::cerr << "ERROR: " << e.what() << "\n";
      return 1;
    }
    failed = runTestsCPU(algorithm, reference, algorithmName, tileSize);
  }
  return failed == 0 ? 0 : 1;
}

}  // namespace TestsRunner
