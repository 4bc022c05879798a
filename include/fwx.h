/*
 * fwx.h -- C ABI of the MI355X (gfx950) Floyd-Warshall engine, libfwx.so.
 *
 * Plain C types only: pointers, sizes, status codes.  `stream` is a
 * hipStream_t passed as void* (NULL = the default stream).  Every entry point
 * returns FWX_OK (0) or a non-zero fwx_status and records a message readable
 * with fwx_last_error() (thread-local).
 *
 * Matrix layout (reference include/constants.h:3-9, src/gpu_memory.cu:6-11):
 * dense row-major n x n int32, D[i*n + j], unpitched, in place; an absent edge
 * is FWX_INF = 1,000,000,000; 1 <= n <= FWX_MAX_N.
 */
#ifndef FWX_H
#define FWX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FWX_INF 1000000000        /* include/constants.h:3 */
#define FWX_MAX_N 32769           /* include/constants.h:5 */
#define FWX_MAX_EDGE_WEIGHT 25000 /* include/constants.h:7 */
#define FWX_MAX_KAPPA 16          /* src/fw_multi_layer_tiling_gpu.cu:5-6 */

typedef enum fwx_status {
  FWX_OK = 0,
  FWX_ERR_INVALID_ARG = 1,  /* bad n, tile, kappa, pointer */
  FWX_ERR_UNKNOWN_ALGO = 2, /* name not in the registry (reference throws std::out_of_range) */
  FWX_ERR_HIP = 3,          /* a HIP runtime call or launch failed */
  FWX_ERR_OOM = 4,          /* a library-owned workspace could not be allocated */
  FWX_ERR_UNSUPPORTED = 5   /* valid request this build does not serve */
} fwx_status;

/* Message of the last failing call on this host thread ("" if none). */
const char* fwx_last_error(void);
/* Library version, major*10000 + minor*100 + patch. */
int fwx_version(void);

/* ---------------- registry (reference src/cuda_utils.cu:7-23) ----------- */

/* Number of GPU algorithms and their names: "baseline_gpu", "tiling_gpu",
 * "multi_layer_tiling_gpu" (reference GPU_ALGORITHMS, src/cuda_utils.cu:7-11). */
int fwx_gpu_algorithm_count(void);
const char* fwx_gpu_algorithm_name(int index);
/* 1 if `name` is a GPU algorithm (reference isGPUAlgorithm, src/cuda_utils.cu:21-23). */
int fwx_is_gpu_algorithm(const char* name);

/* ---------------- plain FW on a device matrix --------------------------- */
/*
 * All four replace a call through AlgorithmFuncGPU
 * `void f(WeightType* d_D, int n, int tileSize, int kappa)`
 * (reference include/algo_types.h:9).  In place on the caller's device
 * buffer; nothing is allocated; the call returns after the work on `stream`
 * has completed (the reference ends with cudaDeviceSynchronize,
 * src/fw_*_gpu.cu).  `tile` = 0 and `kappa` = 0 select the tuned defaults;
 * explicit values must be supported ones (fwx_tiling_gpu: 32, 64, 128;
 * kappa 1..16), others return FWX_ERR_INVALID_ARG.
 */

/* Dispatch by registry name (reference getGPUAlgorithmFunc, src/cuda_utils.cu:13-19). */
int fwx_run(const char* algo, int32_t* d_D, int32_t n, int32_t tile, int32_t kappa, void* stream);

/* Per-k relaxation sweep, one launch per k (reference fwBaselineGPU,
 * include/fw_baseline_gpu.cuh:6, src/fw_baseline_gpu.cu:16-32). */
int fwx_baseline_gpu(int32_t* d_D, int32_t n, void* stream);

/* Three-phase blocked FW (reference fwTilingGPU, include/fw_tiling_gpu.cuh:6,
 * src/fw_tiling_gpu.cu:253-284).  tile 0 = tuned default. */
int fwx_tiling_gpu(int32_t* d_D, int32_t n, int32_t tile, void* stream);

/* Same three-phase algorithm shaped like the reference's kernels
 * (32x32 tiles, 256 threads, 2x2 cells per thread, four launches per
 * k-tile) -- the comparator, src/fw_tiling_gpu.cu:8-284. */
int fwx_tiling_gpu_refshape(int32_t* d_D, int32_t n, void* stream);

/* Multi-layer blocked FW, kappa k-tiles per stage (reference
 * fwMultiLayerTilingGPU, include/fw_multi_layer_tiling_gpu.cuh:17,
 * src/fw_multi_layer_tiling_gpu.cu:507-577).  kappa <= 0 selects the tuned
 * default; kappa > 16 returns FWX_ERR_INVALID_ARG. */
int fwx_multi_layer_tiling_gpu(int32_t* d_D, int32_t n, int32_t tile, int32_t kappa, void* stream);

/* The tuned (tile, kappa) for size n: what tile=0 / kappa=0 resolve to. */
int fwx_tuned_config(int32_t n, int32_t* tile, int32_t* kappa);

/* ---------------- device memory (reference GPUMemory, src/gpu_memory.cu) - */

int fwx_device_alloc(int32_t** d_D, int32_t n);                        /* gpu_memory.cu:6-11 */
int fwx_copy_to_device(int32_t* d_D, const int32_t* h_D, int32_t n);   /* gpu_memory.cu:13-16 */
int fwx_copy_to_host(int32_t* h_D, const int32_t* d_D, int32_t n);     /* gpu_memory.cu:18-21 */
int fwx_device_free(int32_t* d_D);                                     /* gpu_memory.cu:23-27 */
int fwx_device_synchronize(void);
/* Empty launch used to warm the context (reference Timer::warmupGPU, src/timer.cu:7-14). */
int fwx_warmup(void);

/* ---------------- host helpers ----------------------------------------- */

/*
 * Seeded Erdos-Renyi graph, row-major into h_D (n*n int32): diagonal 0; each
 * other cell independently present with probability `density` and then a
 * uniform integer weight in [wmin, wmax], else FWX_INF.  seed 0 = random
 * device.  With (wmin, wmax) = (0, FWX_MAX_EDGE_WEIGHT) the stream of draws
 * and the output are those of the reference generateGraph1D
 * (src/graph_generator.cpp:48-80): std::mt19937, bernoulli_distribution,
 * uniform_int_distribution, in row-major order.
 */
int fwx_generate_graph(int32_t* h_D, int32_t n, double density, uint32_t seed,
                       int32_t wmin, int32_t wmax);
/* 1 if any D[i][i] < 0 (reference hasNegativeCycle, src/utils.cpp:5-12). */
int fwx_has_negative_cycle(const int32_t* h_D, int32_t n);

#ifdef __cplusplus
}
#endif

#endif /* FWX_H */
