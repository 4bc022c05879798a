// dropin.cpp -- the reference's C++ API (registries, algorithm functions,
// GPUMemory, Timer, generator, utils) as thin shims over libfwx's C ABI.
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <stdexcept>

#include "../../include/fwx.h"
#include "algo_types.h"
#include "fw_baseline_cpu.h"
#include "fw_baseline_gpu.h"
#include "fw_multi_layer_tiling_gpu.h"
#include "fw_tiling_gpu.h"
#include "gpu_memory.h"
#include "gpu_utils.h"
#include "graph_generator.h"
#include "timer.h"
#include "utils.h"

// ---- errors (cuda_utils.cuh:11-32 behaviour: print, exit) -----------------
void checkFwx(int status, const char* call, const char* file, int line) {
  if (status != FWX_OK) {
    std::cerr << "GPU Error at " << file << ":" << line << "\n";
    std::cerr << "Error code: " << status << "\n";
    std::cerr << "Function: " << call << "\n";
    std::cerr << "Message: " << fwx_last_error() << "\n";
    std::exit(EXIT_FAILURE);
  }
}

// ---- CPU algorithm (fw_baseline_cpu.cpp:3-16 semantics) --------------------
void fwBaselineCPU(WeightType* D, int n, int tileSize) {
  (void)tileSize;
  const size_t N = (size_t)n;
  for (size_t k = 0; k < N; ++k) {
    const WeightType* rowk = D + k * N;
    for (size_t i = 0; i < N; ++i) {
      WeightType* rowi = D + i * N;
      for (size_t j = 0; j < N; ++j) {
        // two's-complement sum, strict compare, store on improvement
        const WeightType np = (WeightType)((uint32_t)rowi[k] + (uint32_t)rowk[j]);
        if (np < rowi[j]) rowi[j] = np;
      }
    }
  }
}

// ---- GPU algorithm shims ------------------------------------------------------
static int32_t sanitize_tile(int tileSize) {
  return (tileSize == 32 || tileSize == 64 || tileSize == 128) ? tileSize : 0;
}
static int32_t sanitize_kappa(int kappa) { return (kappa >= 1 && kappa <= FWX_MAX_KAPPA) ? kappa : 0; }

void fwBaselineGPU(WeightType* d_D, int n, int tileSize, int kappa) {
  (void)tileSize;
  (void)kappa;
  checkFwxErrors(fwx_baseline_gpu(d_D, n, nullptr));
}

void fwTilingGPU(WeightType* d_D, int n, int tileSize, int kappa) {
  (void)kappa;
  checkFwxErrors(fwx_tiling_gpu(d_D, n, sanitize_tile(tileSize), nullptr));
}

void fwMultiLayerTilingGPU(WeightType* d_D, int n, int tileSize, int kappa) {
  checkFwxErrors(fwx_multi_layer_tiling_gpu(d_D, n, sanitize_tile(tileSize), sanitize_kappa(kappa), nullptr));
}

// ---- registries (cuda_utils.cu:7-23, utils.cpp:14-24) ---------------------
const std::map<std::string, AlgorithmFuncGPU> GPU_ALGORITHMS = {
    {"baseline_gpu", fwBaselineGPU},
    {"tiling_gpu", fwTilingGPU},
    {"multi_layer_tiling_gpu", fwMultiLayerTilingGPU}};

AlgorithmFuncGPU getGPUAlgorithmFunc(const std::string& algorithmName) {
  auto it = GPU_ALGORITHMS.find(algorithmName);
  if (it == GPU_ALGORITHMS.end()) throw std::out_of_range("ERROR: Unsupported GPU algorithm: " + algorithmName);
  return it->second;
}

bool isGPUAlgorithm(const std::string& algorithmName) {
  return GPU_ALGORITHMS.find(algorithmName) != GPU_ALGORITHMS.end();
}

const std::map<std::string, AlgorithmFuncCPU> CPU_ALGORITHMS = {{"baseline_cpu", fwBaselineCPU}};

AlgorithmFuncCPU getCPUAlgorithmFunc(const std::string& algorithmName) {
  auto it = CPU_ALGORITHMS.find(algorithmName);
  if (it == CPU_ALGORITHMS.end()) throw std::out_of_range("ERROR: Unsupported CPU algorithm: " + algorithmName);
  return it->second;
}

bool hasNegativeCycle(const WeightType* distMatrix, int n) {
  return fwx_has_negative_cycle(distMatrix, n) != 0;
}

// ---- generator (graph_generator.cpp:7-80 interface) ------------------------
void* fw_aligned_malloc(size_t size, size_t align) {
  void* p = nullptr;
  if (posix_memalign(&p, align, size) != 0) return nullptr;
  return p;
}

WeightType* generateAlignedGraph1D(int n, double density, unsigned int seed, int wmin, int wmax) {
  if (n <= 0 || n > MAX_N) {
    std::cerr << "ERROR: Graph size (n) must be between 1 and " << MAX_N << ". Graph generation failed.\n";
    return nullptr;
  }
  const size_t bytes = (size_t)n * (size_t)n * sizeof(WeightType);
  WeightType* D = static_cast<WeightType*>(ALIGNED_MALLOC(bytes, 32));
  if (D == nullptr) {
    std::cerr << "ERROR: Aligned memory allocation failed.\n";
    return nullptr;
  }
  if (fwx_generate_graph(D, n, density, seed, wmin, wmax) != FWX_OK) {
    std::cerr << "ERROR: " << fwx_last_error() << "\n";
    ALIGNED_FREE(D);
    return nullptr;
  }
  return D;
}

WeightType* generateAlignedGraph1D(int n, double density, unsigned int seed) {
  return generateAlignedGraph1D(n, density, seed, 0, MAX_EDGE_WEIGHT);
}

std::vector<WeightType> generateGraph1D(int n, double density, unsigned int seed, int wmin, int wmax) {
  if (n <= 0 || n > MAX_N) {
    std::cerr << "ERROR: Graph size (n) must be between 1 and " << MAX_N << ". Graph generation failed.\n";
    return {};
  }
  std::vector<WeightType> D((size_t)n * (size_t)n);
  if (fwx_generate_graph(D.data(), n, density, seed, wmin, wmax) != FWX_OK) {
    std::cerr << "ERROR: " << fwx_last_error() << "\n";
    return {};
  }
  return D;
}

std::vector<WeightType> generateGraph1D(int n, double density, unsigned int seed) {
  return generateGraph1D(n, density, seed, 0, MAX_EDGE_WEIGHT);
}

// ---- GPUMemory (gpu_memory.cu:6-27) ----------------------------------------
namespace GPUMemory {
WeightType* allocate(int n) {
  int32_t* d = nullptr;
  checkFwxErrors(fwx_device_alloc(&d, n));
  return d;
}
void copyToDevice(WeightType* d_graph, const WeightType* h_graph, int n) {
  checkFwxErrors(fwx_copy_to_device(d_graph, h_graph, n));
}
void copyToHost(WeightType* h_graph, const WeightType* d_graph, int n) {
  checkFwxErrors(fwx_copy_to_host(h_graph, d_graph, n));
}
void free(WeightType* d_graph) {
  if (d_graph != nullptr) checkFwxErrors(fwx_device_free(d_graph));
}
}  // namespace GPUMemory

// ---- Timer (timer.cu:7-64 protocol) ----------------------------------------
namespace Timer {
void warmupGPU() { checkFwxErrors(fwx_warmup()); }

double measureCPU(AlgorithmFuncCPU func, WeightType* graph, int n, int tileSize, int warmup_runs) {
  for (int i = 0; i < warmup_runs; ++i) func(graph, n, tileSize);
  const auto t0 = std::chrono::steady_clock::now();
  func(graph, n, tileSize);
  const auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

// Events recorded on the default stream before and after the call, as the
// reference does (src/timer.cu:48-63): the GPU's own clock between the two
// markers, seconds.
double measureGPU(AlgorithmFuncGPU func, WeightType* d_graph, int n, int tileSize, int kappa, int warmup_runs) {
  for (int i = 0; i < warmup_runs; ++i) {
    func(d_graph, n, tileSize, kappa);
    checkFwxErrors(fwx_device_synchronize());
  }
  void* start = nullptr;
  void* stop = nullptr;
  checkFwxErrors(fwx_event_create(&start));
  checkFwxErrors(fwx_event_create(&stop));
  checkFwxErrors(fwx_event_record(start, nullptr));
  func(d_graph, n, tileSize, kappa);
  checkFwxErrors(fwx_event_record(stop, nullptr));
  checkFwxErrors(fwx_event_synchronize(stop));
  float ms = 0.0f;
  checkFwxErrors(fwx_event_elapsed_ms(&ms, start, stop));
  checkFwxErrors(fwx_event_destroy(start));
  checkFwxErrors(fwx_event_destroy(stop));
  return ms / 1000.0;
}
}  // namespace Timer
