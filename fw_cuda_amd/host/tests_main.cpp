// tests_runner -- CLI with the reference's argv rules (src/tests_main.cpp:6-50):
//   tests_runner <algorithm_name> [tile_size] [kappa]
// Exits 1 when any test fails (the reference always exits 0).
#include <iostream>
#include <stdexcept>
#include <string>

#include "tests_runner.h"

// Usage text in this project's words; the argv rules and the error lines
// below are the reference's contract.
static void printUsage(const char* programName) {
  std::cerr << "Usage: " << programName << " <algorithm_name> [tile_size] [kappa]\n"
            << "\n"
            << "Runs the known-answer matrices and seeded graphs against the CPU baseline.\n"
            << "  algorithm_name  baseline_cpu, baseline_gpu, tiling_gpu or multi_layer_tiling_gpu\n"
            << "  tile_size       32, 64 or 128; 0 or omitted picks the tuned tile per graph\n"
            << "  kappa           k-tiles per stage (multi_layer_tiling_gpu); 0 picks the tuned value\n"
            << "\n"
            << "e.g. " << programName << " tiling_gpu 64\n";
}

int main(int argc, char* argv[]) {
  if (argc < 2 || argc > 4) {
    printUsage(argv[0]);
    return 1;
  }
  const std::string algorithmName = argv[1];
  int tileSize = 0;
  if (argc >= 3) {
    try {
      tileSize = std::stoi(argv[2]);
    } catch (const std::exception&) {
      std::cerr << "ERROR: Invalid TILE_SIZE provided (" << argv[2] << "). Must be an integer.\n";
      return 1;
    }
  }
  int kappa = 0;
  if (argc == 4) {
    try {
      kappa = std::stoi(argv[3]);
    } catch (const std::exception&) {
      std::cerr << "ERROR: Invalid KAPPA provided (" << argv[3] << "). Must be an integer.\n";
      return 1;
    }
  }
  return TestsRunner::executeTests(algorithmName, tileSize, kappa);
}
