// bench_runner -- CLI with the reference's argv rules and messages
// (src/bench_runner.cpp:7-103):
//   bench_runner <algorithm_name> <n_size> <density> [seed] [tile_size] [kappa] [wmin wmax]
// The optional trailing weight range [wmin wmax] is an addition (default
// [0, MAX_EDGE_WEIGHT], the reference generator's range).
#include <iostream>
#include <stdexcept>
#include <string>

#include "algorithm_runner.h"
#include "constants.h"

// Usage text in this project's words; the argv rules and the error lines
// below are the reference's contract.
static void printUsage(const char* programName) {
  std::cerr << "Usage: " << programName
            << " <algorithm_name> <n_size> <density> [seed] [tile_size] [kappa] [wmin wmax]\n"
            << "\n"
            << "Times one solve of a generated graph after one warmup call.\n"
            << "  algorithm_name  baseline_cpu, baseline_gpu, tiling_gpu or multi_layer_tiling_gpu\n"
            << "  n_size          vertices, 1 to " << MAX_N << "\n"
            << "  density         probability of an edge i -> j (i != j), 0.0 to 1.0\n"
            << "  seed            generator seed; 0 or omitted draws one at random\n"
            << "  tile_size       32, 64 or 128; 0 or omitted picks the tuned tile for n_size\n"
            << "  kappa           k-tiles per stage (multi_layer_tiling_gpu); 0 picks the tuned value\n"
            << "  wmin wmax       edge weights drawn from [wmin, wmax] (default 0 " << MAX_EDGE_WEIGHT << ")\n"
            << "\n"
            << "e.g. " << programName << " multi_layer_tiling_gpu 8192 0.9 42 0 0\n";
}

int main(int argc, char* argv[]) {
  if (argc < 4 || argc > 9 || argc == 8) {
    printUsage(argv[0]);
    return 1;
  }
  const std::string algorithmName = argv[1];
  int n = 0, tileSize = 0, kappa = 0, wmin = 0, wmax = MAX_EDGE_WEIGHT;
  double density = 0.0;
  unsigned int seed = 0;

  try {
    n = std::stoi(argv[2]);
    if (n <= 0 || n > MAX_N) {
      std::cerr << "ERROR: Invalid graph size (" << n << "). Must be between 1 and " << MAX_N << ".\n";
      return 1;
    }
  } catch (const std::exception&) {
    std::cerr << "ERROR: Invalid N_SIZE provided (" << argv[2] << "). Must be an integer.\n";
    return 1;
  }
  try {
    density = std::stod(argv[3]);
    if (density < 0.0 || density > 1.0) {
      std::cerr << "ERROR: Invalid density (" << density << "). Must be between 0.0 and 1.0.\n";
      return 1;
    }
  } catch (const std::exception&) {
    std::cerr << "ERROR: Invalid DENSITY provided (" << argv[3] << "). Must be a floating point number.\n";
    return 1;
  }
  if (argc >= 5) {
    try {
      seed = std::stoul(argv[4]);
    } catch (const std::exception&) {
      std::cerr << "ERROR: Invalid SEED provided (" << argv[4] << "). Must be an unsigned integer.\n";
      return 1;
    }
  }
  if (argc >= 6) {
    try {
      tileSize = std::stoi(argv[5]);
      if (tileSize < 0 || tileSize > n) {
        std::cerr << "ERROR: Invalid TILE_SIZE (" << tileSize << "). Must be between 0 and " << n << ".\n";
        return 1;
      }
    } catch (const std::exception&) {
      std::cerr << "ERROR: Invalid TILE_SIZE provided (" << argv[5] << "). Must be an integer.\n";
      return 1;
    }
  }
  if (argc >= 7) {
    try {
      kappa = std::stoi(argv[6]);
      if (kappa < 0) {
        std::cerr << "ERROR: Invalid KAPPA (" << kappa << "). Must be non-negative.\n";
        return 1;
      }
    } catch (const std::exception&) {
      std::cerr << "ERROR: Invalid KAPPA provided (" << argv[6] << "). Must be an integer.\n";
      return 1;
    }
  }
  if (argc == 9) {
    try {
      wmin = std::stoi(argv[7]);
      wmax = std::stoi(argv[8]);
      if (wmin > wmax) {
        std::cerr << "ERROR: Invalid weight range (" << wmin << ", " << wmax << ").\n";
        return 1;
      }
    } catch (const std::exception&) {
      std::cerr << "ERROR: Invalid WMIN/WMAX provided. Must be integers.\n";
      return 1;
    }
  }
  try {
    return AlgorithmRunner::executeBenchmark(algorithmName, n, tileSize, kappa, density, seed, wmin, wmax);
  } catch (const std::exception& e) {
    std::cerr << "RUNTIME ERROR: " << e.what() << "\n";
    return 1;
  } catch (...) {
    std::cerr << "RUNTIME ERROR: An unknown error occurred.\n";
    return 1;
  }
}
