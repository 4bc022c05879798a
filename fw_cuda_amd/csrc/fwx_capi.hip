// fwx_capi.hip -- the extern "C" surface of libfwx.so (declared in include/fwx.h).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <random>
#include <string>

#include "fwx_internal.hpp"

namespace fwx {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
  char buf[512];
  std::snprintf(buf, sizeof(buf), "HIP error %d (%s: %s) in %s at %s:%d", (int)e, hipGetErrorName(e),
                hipGetErrorString(e), what, file, line);
  set_error(buf);
  return e == hipErrorOutOfMemory ? FWX_ERR_OOM : FWX_ERR_HIP;
}

__global__ void k_warmup() {}

static const char* const kGpuAlgorithms[] = {"baseline_gpu", "tiling_gpu", "multi_layer_tiling_gpu"};
constexpr int kNumGpuAlgorithms = 3;

static int check_matrix(const void* p, int32_t n) {
  if (p == nullptr) return fail(FWX_ERR_INVALID_ARG, "null matrix pointer");
  if (n <= 0 || n > FWX_MAX_N) return fail(FWX_ERR_INVALID_ARG, "n must be in [1, 32769]");
  return FWX_OK;
}

// tile 0 / kappa <= 0 select the tuned values; kappa 1 passed explicitly
// (tiling_gpu) takes the tile tuned for kappa 1.
static int resolve(int32_t n, int32_t& tile, int32_t& kappa) {
  int32_t t0, k0;
  tuned_config(n, &t0, &k0);
  if (tile == 0) tile = kappa == 1 ? tuned_tiling_tile(n) : t0;
  if (tile != 32 && tile != 64 && tile != 128)
    return fail(FWX_ERR_INVALID_ARG, "tile must be 0 (tuned), 32, 64 or 128");
  if (kappa <= 0) kappa = k0;
  if (kappa > FWX_MAX_KAPPA) return fail(FWX_ERR_INVALID_ARG, "kappa must be <= 16");
  return FWX_OK;
}

static int finish(hipStream_t st) {
  FWX_HIP(hipStreamSynchronize(st));
  return FWX_OK;
}

}  // namespace fwx

using namespace fwx;

extern "C" {

const char* fwx_last_error(void) { return g_last_error.c_str(); }

int fwx_version(void) { return 100; }

int fwx_gpu_algorithm_count(void) { return kNumGpuAlgorithms; }

const char* fwx_gpu_algorithm_name(int index) {
  if (index < 0 || index >= kNumGpuAlgorithms) return nullptr;
  return kGpuAlgorithms[index];
}

int fwx_is_gpu_algorithm(const char* name) {
  if (!name) return 0;
  for (int i = 0; i < kNumGpuAlgorithms; ++i)
    if (std::strcmp(name, kGpuAlgorithms[i]) == 0) return 1;
  return 0;
}

int fwx_tuned_config(int32_t n, int32_t* tile, int32_t* kappa) {
  if (!tile || !kappa) return fail(FWX_ERR_INVALID_ARG, "null output pointer");
  if (n <= 0 || n > FWX_MAX_N) return fail(FWX_ERR_INVALID_ARG, "n must be in [1, 32769]");
  tuned_config(n, tile, kappa);
  return FWX_OK;
}

int fwx_baseline_gpu(int32_t* d_D, int32_t n, void* stream) {
  clear_error();
  if (int rc = check_matrix(d_D, n)) return rc;
  hipStream_t st = as_stream(stream);
  if (int rc = enqueue_baseline(d_D, n, st)) return rc;
  return finish(st);
}

int fwx_tiling_gpu(int32_t* d_D, int32_t n, int32_t tile, void* stream) {
  clear_error();
  if (int rc = check_matrix(d_D, n)) return rc;
  int32_t kappa = 1;
  if (int rc = resolve(n, tile, kappa)) return rc;
  hipStream_t st = as_stream(stream);
  if (int rc = enqueue_blocked(d_D, n, tile, 1, st)) return rc;
  return finish(st);
}

int fwx_tiling_gpu_refshape(int32_t* d_D, int32_t n, void* stream) {
  clear_error();
  if (int rc = check_matrix(d_D, n)) return rc;
  hipStream_t st = as_stream(stream);
  if (int rc = enqueue_refshape(d_D, n, st)) return rc;
  return finish(st);
}

int fwx_multi_layer_tiling_gpu(int32_t* d_D, int32_t n, int32_t tile, int32_t kappa, void* stream) {
  clear_error();
  if (int rc = check_matrix(d_D, n)) return rc;
  if (int rc = resolve(n, tile, kappa)) return rc;
  hipStream_t st = as_stream(stream);
  if (int rc = enqueue_blocked(d_D, n, tile, kappa, st)) return rc;
  return finish(st);
}

int fwx_run(const char* algo, int32_t* d_D, int32_t n, int32_t tile, int32_t kappa, void* stream) {
  clear_error();
  if (!algo) return fail(FWX_ERR_UNKNOWN_ALGO, "null algorithm name");
  if (std::strcmp(algo, "baseline_gpu") == 0) return fwx_baseline_gpu(d_D, n, stream);
  if (std::strcmp(algo, "tiling_gpu") == 0) return fwx_tiling_gpu(d_D, n, tile, stream);
  if (std::strcmp(algo, "multi_layer_tiling_gpu") == 0)
    return fwx_multi_layer_tiling_gpu(d_D, n, tile, kappa, stream);
  return fail(FWX_ERR_UNKNOWN_ALGO, std::string("Unsupported GPU algorithm: ") + algo);
}

int fwx_device_alloc(int32_t** d_D, int32_t n) {
  clear_error();
  if (!d_D) return fail(FWX_ERR_INVALID_ARG, "null output pointer");
  if (n <= 0 || n > FWX_MAX_N) return fail(FWX_ERR_INVALID_ARG, "n must be in [1, 32769]");
  *d_D = nullptr;
  FWX_HIP(hipMalloc(reinterpret_cast<void**>(d_D), (size_t)n * (size_t)n * sizeof(int32_t)));
  return FWX_OK;
}

int fwx_copy_to_device(int32_t* d_D, const int32_t* h_D, int32_t n) {
  clear_error();
  if (int rc = check_matrix(d_D, n)) return rc;
  if (!h_D) return fail(FWX_ERR_INVALID_ARG, "null host pointer");
  FWX_HIP(hipMemcpy(d_D, h_D, (size_t)n * (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice));
  return FWX_OK;
}

int fwx_copy_to_host(int32_t* h_D, const int32_t* d_D, int32_t n) {
  clear_error();
  if (int rc = check_matrix(d_D, n)) return rc;
  if (!h_D) return fail(FWX_ERR_INVALID_ARG, "null host pointer");
  FWX_HIP(hipMemcpy(h_D, d_D, (size_t)n * (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost));
  return FWX_OK;
}

int fwx_device_free(int32_t* d_D) {
  clear_error();
  if (d_D) FWX_HIP(hipFree(d_D));
  return FWX_OK;
}

int fwx_device_synchronize(void) {
  clear_error();
  FWX_HIP(hipDeviceSynchronize());
  return FWX_OK;
}

int fwx_device_memory(uint64_t* free_bytes, uint64_t* total_bytes) {
  clear_error();
  if (!free_bytes || !total_bytes) return fail(FWX_ERR_INVALID_ARG, "null output pointer");
  size_t f = 0, t = 0;
  FWX_HIP(hipMemGetInfo(&f, &t));
  *free_bytes = f;
  *total_bytes = t;
  return FWX_OK;
}

int fwx_warmup(void) {
  clear_error();
  k_warmup<<<1, 1>>>();
  FWX_LAUNCHED("k_warmup");
  FWX_HIP(hipDeviceSynchronize());
  return FWX_OK;
}

int fwx_generate_graph(int32_t* h_D, int32_t n, double density, uint32_t seed, int32_t wmin,
                       int32_t wmax) {
  clear_error();
  if (int rc = check_matrix(h_D, n)) return rc;
  if (!(density >= 0.0 && density <= 1.0)) return fail(FWX_ERR_INVALID_ARG, "density must be in [0, 1]");
  if (wmin > wmax) return fail(FWX_ERR_INVALID_ARG, "wmin must be <= wmax");
  // Same engine, distributions and draw order as src/graph_generator.cpp:25-43,
  // so (wmin, wmax) = (0, 25000) reproduces the reference's matrices.
  std::mt19937 gen(seed != 0 ? seed : std::random_device{}());
  std::uniform_int_distribution<int> weight(wmin, wmax);
  std::bernoulli_distribution present(density);
  const size_t N = (size_t)n;
  for (size_t i = 0; i < N; ++i) {
    int32_t* row = h_D + i * N;
    for (size_t j = 0; j < N; ++j) {
      if (i == j)
        row[j] = 0;
      else if (present(gen))
        row[j] = weight(gen);
      else
        row[j] = FWX_INF;
    }
  }
  return FWX_OK;
}

int fwx_has_negative_cycle(const int32_t* h_D, int32_t n) {
  if (!h_D || n <= 0) return 0;
  for (int32_t i = 0; i < n; ++i)
    if (h_D[(size_t)i * n + i] < 0) return 1;
  return 0;
}

int fwx_relaxation_form(const int32_t* d_D, int32_t n, int32_t* form, void* stream) {
  clear_error();
  if (int rc = check_matrix(d_D, n)) return rc;
  if (!form) return fail(FWX_ERR_INVALID_ARG, "null output pointer");
  bool pk = false;
  if (int rc = values_in_pk64_range(d_D, (int64_t)n * n, as_stream(stream), &pk)) return rc;
  *form = pk ? FWX_FORM_PK64 : FWX_FORM_MIN3;
  return FWX_OK;
}

int fwx_last_solve_form(int32_t* form, int32_t* exact_finish) {
  clear_error();
  if (!form || !exact_finish) return fail(FWX_ERR_INVALID_ARG, "null output pointer");
  last_solve_form(form, exact_finish);
  return FWX_OK;
}

int fwx_last_solve_packed(int32_t* packed) {
  clear_error();
  if (!packed) return fail(FWX_ERR_INVALID_ARG, "null output pointer");
  *packed = last_solve_packed();
  return FWX_OK;
}

// ---- events (the cudaEvent calls of src/timer.cu:48-63) ---------------------
int fwx_event_create(void** ev) {
  clear_error();
  if (!ev) return fail(FWX_ERR_INVALID_ARG, "null output pointer");
  hipEvent_t e = nullptr;
  FWX_HIP(hipEventCreate(&e));
  *ev = e;
  return FWX_OK;
}

int fwx_event_record(void* ev, void* stream) {
  clear_error();
  if (!ev) return fail(FWX_ERR_INVALID_ARG, "null event");
  FWX_HIP(hipEventRecord(static_cast<hipEvent_t>(ev), as_stream(stream)));
  return FWX_OK;
}

int fwx_event_synchronize(void* ev) {
  clear_error();
  if (!ev) return fail(FWX_ERR_INVALID_ARG, "null event");
  FWX_HIP(hipEventSynchronize(static_cast<hipEvent_t>(ev)));
  return FWX_OK;
}

int fwx_event_elapsed_ms(float* ms, void* start, void* stop) {
  clear_error();
  if (!ms || !start || !stop) return fail(FWX_ERR_INVALID_ARG, "null pointer");
  FWX_HIP(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)));
  return FWX_OK;
}

int fwx_event_destroy(void* ev) {
  clear_error();
  if (ev) FWX_HIP(hipEventDestroy(static_cast<hipEvent_t>(ev)));
  return FWX_OK;
}

int fwx_has_negative_cycle_device(const int32_t* d_D, int32_t n, int32_t* result, void* stream) {
  clear_error();
  if (!d_D || !result || n <= 0 || n > FWX_MAX_N)
    return fail(FWX_ERR_INVALID_ARG, "fwx_has_negative_cycle_device: null pointer or n out of range");
  bool neg = false;
  if (int rc = device_negative_cycle(d_D, n, as_stream(stream), &neg)) return rc;
  *result = neg ? 1 : 0;
  return FWX_OK;
}

}  // extern "C"
