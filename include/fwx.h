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
  FWX_ERR_UNSUPPORTED = 5,  /* valid request this build does not serve */
  FWX_ERR_TIMEOUT = 6       /* a device-side wait exceeded its limit / work left unrun */
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

/*
 * The relaxation form a plain solve of d_D would use, from the same one-pass
 * range check every solve starts with: FWX_FORM_PK64 when every cell lies in
 * [0, 2^30) (one VALU instruction per relaxation), else FWX_FORM_MIN3 (1.5).
 * Both are exact; the form bounds the solve's VALU ceiling (CUs x 64 lanes x
 * clock relaxations/s for pk64, two thirds of that for min3).  No reference
 * counterpart.  Synchronises `stream`.
 */
#define FWX_FORM_MIN3 0
#define FWX_FORM_PK64 1
int fwx_relaxation_form(const int32_t* d_D, int32_t n, int32_t* form, void* stream);

/*
 * What the calling thread's last plain blocked solve (tiling_gpu /
 * multi_layer_tiling_gpu) ran in.  *form: FWX_FORM_MIN3, FWX_FORM_PK64, or
 * FWX_FORM_PK16 -- pk64 lead rounds with the 128-wide rest passes in the
 * packed 16-bit form (2 VALU instructions per 4 relaxations on values
 * clamped at FWX_PK16_CAP; exact for every cell below the cap) -- or -1 for
 * the fused small-n rounds (each block picks pk64 or min3 itself).
 * *exact_finish = 1 when cells at or above FWX_PK16_CAP remained after the
 * pk16 run and the solve finished with an exact pk64 run on its result.
 * -1 / 0 before any solve.  No reference counterpart.
 */
#define FWX_FORM_PK16 2
#define FWX_PK16_CAP 15359
int fwx_last_solve_form(int32_t* form, int32_t* exact_finish);
/*
 * *packed = 1 when the calling thread's last plain blocked solve ran its
 * pk16 rest passes on packed stage panels (each stage's panels converted
 * once to the 16-bit operand layout after its lead rounds, then copied into
 * LDS by every rest block): direct launches of a pk16 solve with n % 128 ==
 * 0, 16-deep chunks and a packing workspace (3 x 6 n kappa tile bytes of
 * device memory plus 256 bytes of flags, allocated on first need and
 * kept); FWX_PACK16=0 turns it off.  Such solves on 64-wide tiles with
 * kappa > 1 also run the block lead phase.  0 otherwise.  No reference
 * counterpart.
 */
int fwx_last_solve_packed(int32_t* packed);

/* ---------------- device memory (reference GPUMemory, src/gpu_memory.cu) - */

int fwx_device_alloc(int32_t** d_D, int32_t n);                        /* gpu_memory.cu:6-11 */
int fwx_copy_to_device(int32_t* d_D, const int32_t* h_D, int32_t n);   /* gpu_memory.cu:13-16 */
int fwx_copy_to_host(int32_t* h_D, const int32_t* d_D, int32_t n);     /* gpu_memory.cu:18-21 */
int fwx_device_free(int32_t* d_D);                                     /* gpu_memory.cu:23-27 */
int fwx_device_synchronize(void);
/* Free and total device memory of the current device in bytes (no
 * reference counterpart; the differentiable binding sizes its predecessor
 * storage from it). */
int fwx_device_memory(uint64_t* free_bytes, uint64_t* total_bytes);
/* Empty launch used to warm the context (reference Timer::warmupGPU, src/timer.cu:7-14). */
int fwx_warmup(void);

/* Timing events (opaque hipEvent_t as void*), the cudaEventCreate / Record /
 * Synchronize / ElapsedTime / Destroy calls of Timer::measureGPU
 * (src/timer.cu:48-63).  `stream` NULL = the default stream, as there. */
int fwx_event_create(void** ev);
int fwx_event_record(void* ev, void* stream);
int fwx_event_synchronize(void* ev);
int fwx_event_elapsed_ms(float* ms, void* start, void* stop);
int fwx_event_destroy(void* ev);

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
/* The same check on a device matrix (no copy of the n x n result to the
 * host): *result = 1 if any d_D[i][i] < 0, else 0.  Reads the n diagonal
 * cells on `stream` and synchronises it.  Replaces the host-side
 * hasNegativeCycle call after a GPU solve, which needs the whole matrix
 * copied back first (src/tests_gpu.cu:41-53,139-141). */
int fwx_has_negative_cycle_device(const int32_t* d_D, int32_t n, int32_t* result, void* stream);

/* ---------------- input pipeline (no reference counterpart) ------------- */
/*
 * Graph generated on the GPU into a caller-owned device matrix: Philox4x32-10
 * counter-based draws, one per cell (counter = cell index i*n+j, key = seed),
 * edge with probability `density`, weight uniform in [wmin, wmax], diagonal
 * 0, INF where absent -- the distributions of generateGraph1D
 * (src/graph_generator.cpp:25-43) but NOT its mt19937 stream: a seed gives a
 * different matrix than fwx_generate_graph.  Identical output for any launch
 * shape.  Synchronises `stream`.
 */
int fwx_generate_graph_device(int32_t* d_D, int32_t n, double density, uint64_t seed, int32_t wmin,
                              int32_t wmax, void* stream);
/* Pinned host matrices and asynchronous copies on `stream` (the reference's
 * GPUMemory::copyToDevice/copyToHost are synchronous pageable cudaMemcpy,
 * src/gpu_memory.cu:13-21).  The async copies return at once; pair them with
 * fwx_stream_synchronize.  Pageable host pointers work too, at pageable
 * speed. */
int fwx_host_alloc_pinned(int32_t** h_D, int32_t n);
int fwx_host_free_pinned(int32_t* h_D);
int fwx_copy_to_device_async(int32_t* d_D, const int32_t* h_D, int32_t n, void* stream);
int fwx_copy_to_host_async(int32_t* h_D, const int32_t* d_D, int32_t n, void* stream);
int fwx_stream_synchronize(void* stream);

/* ---------------- predecessors (no reference counterpart) --------------- */
/*
 * Path reconstruction for the int32 engine.  d_W: the input weights (the
 * solve overwrites its buffer, so the caller keeps a copy), d_D: the solved
 * distances, d_P: n x n int32 output.  P[i][j] = the smallest u != j with
 * D[i][u] + W[u][j] == D[i][j] (the vertex before j on a shortest i->j
 * path), P[i][i] = -1, -1 where no u matches.  When every cycle has
 * positive weight, following P from j reaches i along a shortest path.  One
 * n^3 pass on `stream`; synchronises it.
 */
int fwx_predecessors(const int32_t* d_W, const int32_t* d_D, int32_t* d_P, int32_t n, void* stream);

/* ---------------- certificate of solved rows (validation) --------------- */
/*
 * On-device check of claimed shortest-distance rows, without solving
 * anything.  Replaces the host-side checks after a GPU solve: the full
 * comparison with fwBaselineCPU (src/tests.cpp:37-57, out of reach at
 * n=32768) and the host Dijkstra spot check of large solves.  d_W: the input
 * weights (zero diagonal, off-diagonal in [1, FWX_INF]; the INF-completed
 * graph, as src/fw_baseline_cpu.cpp:9 treats it); d_src: nsrc source
 * vertices; d_rows: nsrc x n claimed rows, row q = the claimed distances
 * from d_src[q].  Writes d_ok[q] = 1 iff row q is exactly the shortest
 * distance row (d[s] == 0, every d[v] in [0, INF], and d[v] == min over
 * u != v of d[u] + W[u][v] for v != s), else 0, and *failed = the number of
 * failing rows.  nsrc * n^2 relaxations on `stream`; synchronises it.
 * FWX_ERR_INVALID_ARG if W is outside the domain (then d_ok is undefined).
 */
int fwx_certify_rows(const int32_t* d_W, int32_t n, const int32_t* d_src, int32_t nsrc, const int32_t* d_rows,
                     int32_t* d_ok, int32_t* failed, void* stream);

/* ---------------- batched FW (no reference counterpart) ----------------- */
/*
 * Many independent graphs solved together.  Graph g is an n_g x n_g row-major
 * int32 matrix at d_data + offsets[g] (element offsets; `sizes` and
 * `offsets` are host arrays of `count` entries; d_data 16-byte aligned).
 * The plan builds every graph's multi-layer schedule ((tile, kappa) per
 * graph size class from the batched tuned table), moves each diagonal
 * closure, with the one product item it needs,
 * onto a side branch of the step before it, merges step j of all graphs into
 * one launch per work kind, largest graphs first, and uploads the work lists.
 * Run modes:
 *   FWX_BATCH_DIRECT (0)   the step launches enqueued directly;
 *   FWX_BATCH_GRAPH (1)    the step launches captured into a hipGraph on first
 *                          use and replayed;
 *   FWX_BATCH_PERSISTENT (2) ONE persistent cooperative kernel: a grid of
 *                          resident 512-thread blocks walks the same steps
 *                          (every graph on 128-wide tiles: graphs below
 *                          n = 128 are one closure), each block running the
 *                          tasks the plan assigned it, with a grid-wide
 *                          barrier between steps;
 *   FWX_BATCH_PERSISTENT_GRAPH (3) that one launch captured into a hipGraph
 *                          and replayed.
 * The persistent modes return FWX_ERR_UNSUPPORTED on a device without
 * cooperative launches, or if the runtime cannot capture one into a graph.
 * In the step-launch modes, when every cell of the batch lies in [0, 2^30)
 * the 128-wide products run in the pk16 form (fwx_last_solve_form); after
 * such a run one pass flags every graph that still holds a cell at or above
 * FWX_PK16_CAP, and each flagged graph is finished by an exact pk64 solve of
 * its own matrix.  The persistent modes run the pk64 form.
 * Results equal a plain fwx_multi_layer_tiling_gpu call per graph in every
 * mode.  The plan owns device memory (its work lists), a side stream and two
 * events until fwx_batch_plan_destroy; side branches fork from the plan's
 * stream and join it again within each run, and a run returns after the
 * plan's stream is idle.  One plan runs one batch at a time: concurrent runs
 * of the same plan from several host threads race on its side stream.
 */
typedef struct fwx_batch_plan fwx_batch_plan;
#define FWX_BATCH_DIRECT 0
#define FWX_BATCH_GRAPH 1
#define FWX_BATCH_PERSISTENT 2
#define FWX_BATCH_PERSISTENT_GRAPH 3
int fwx_batch_plan_create(fwx_batch_plan** plan, const int32_t* sizes, const int64_t* offsets,
                          int32_t count, int32_t* d_data, void* stream);
int fwx_batch_plan_run(fwx_batch_plan* plan, int32_t mode);
/* Persistent mode: blocks of the cooperative grid (0 when the device has no
 * cooperative launch), grid-wide steps, and tasks. */
int fwx_batch_plan_persistent_info(const fwx_batch_plan* plan, int32_t* grid, int32_t* steps, int64_t* tasks);
/* Dependent steps, kernel launches per run, and work items of a plan. */
int fwx_batch_plan_stats(const fwx_batch_plan* plan, int32_t* steps, int32_t* launches, int64_t* items);
/* The form of the plan's last run (FWX_FORM_MIN3 / PK64 / PK16; -1 before
 * any run) and how many graphs its pk16 check sent to an exact pk64 finish.
 * No reference counterpart. */
int fwx_batch_plan_last_form(const fwx_batch_plan* plan, int32_t* form, int32_t* exact_graphs);
int fwx_batch_plan_destroy(fwx_batch_plan* plan);
/* The comparator: one plain fwx_multi_layer_tiling_gpu call (with its final
 * stream synchronisation) per graph, in order. */
int fwx_batch_loop(const int32_t* sizes, const int64_t* offsets, int32_t count, int32_t* d_data,
                   void* stream);

/* ---------------- differentiable fp32 FW (no reference counterpart) ----- */
/*
 * Forward, for `batch` contiguous n x n fp32 graphs (1 <= n <= 8192):
 * D = min-plus closure of W with the diagonal taken as 0 and +inf meaning
 * no edge; P[i][j] = vertex before j on the chosen i->j path (i for a direct
 * edge), -1 when there is none.  d_P may be NULL (no predecessors kept).
 * Weights must be positive for the backward's ordering argument.
 */
int fwx_diff_forward(const float* d_W, float* d_D, int32_t* d_P, int32_t n, int32_t batch,
                     void* stream);
/*
 * Backward of L = sum(G * D): dW[u][v] = sum of G[s][t] over every (s,t)
 * whose chosen path uses edge (u,v); d_dW is overwritten.  Per source s each
 * vertex's gradient becomes G[s][t] plus its subtree's: in a tree whose
 * chains all reach s within 32 links each vertex adds G to its ancestors,
 * otherwise the vertices are taken from the deepest level up, each adding
 * its accumulated gradient to its parent.  Each source's final gradient row
 * is then summed into dW by predecessor, per target column.
 * With d_P given, those rows are staged in d_dW itself and nothing is
 * allocated.  d_P = NULL recomputes the predecessors from d_W and d_D, 256
 * sources at a time (a stream-ordered workspace of 256 x n x batch int32
 * plus as many floats, instead of the n x n x batch predecessor matrix).
 * FWX_DIFF_BACK=0 in the environment adds to dW with one global atomic per
 * (source, vertex) instead (the same sums up to fp32 rounding order).
 */
int fwx_diff_backward(const float* d_W, const float* d_D, const int32_t* d_P, const float* d_G,
                      float* d_dW, int32_t n, int32_t batch, void* stream);

/*
 * fwx_diff_backward with recomputed predecessors (d_P = NULL) that allocates
 * nothing.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory -- code under HIP graph capture, tight training
 * loops, harnesses that forbid hipMalloc.  The recomputation workspace is a
 * 16 MiB array in the library's own device memory (one per device, held
 * until the process ends) instead of a stream-ordered allocation per call;
 * sources are processed in chunks of as many rows as fit (at most 256, as
 * fwx_diff_backward), so n x batch may be at most 2,097,152 (1,048,576 with
 * FWX_DIFF_BACK=0); larger sizes return FWX_ERR_INVALID_ARG.  Calls that
 * overlap on one device share that workspace and race: run one at a time
 * per device.  Same results as fwx_diff_backward(d_W, d_D, NULL, ...) for
 * sizes where both take 256-source chunks; synchronises `stream`.
 */
int fwx_diff_backward_noalloc(const float* d_W, const float* d_D, const float* d_G, float* d_dW, int32_t n,
                              int32_t batch, void* stream);

/* ---------------- profiling (HIP events, no reference counterpart) ----- */
/*
 * Between fwx_profile_begin() and fwx_profile_end(), every launch group is
 * bracketed by HIP events on the stream it runs on.  fwx_profile_end fills
 * arrays of FWX_PROF_NCLASS entries (any may be NULL): summed event time in
 * ms, summed algorithmic work (relaxations for REST/LEAD/BAND; bytes for
 * BASELINE: the 4 n^2 bytes every per-k sweep must read -- it writes only
 * the cells that improve), and the number of groups.  The reference
 * times with cudaEvent pairs around a whole call (src/timer.cu:48-63).
 */
#define FWX_PROF_REST 0     /* one k_rest launch per stage */
#define FWX_PROF_LEAD 1     /* the lead rounds of one stage */
#define FWX_PROF_BASELINE 2 /* one per-k sweep */
#define FWX_PROF_BAND 3     /* look-ahead: the rest tiles the next stage's lead reads */
#define FWX_PROF_NCLASS 4
int fwx_profile_begin(void);
int fwx_profile_end(double* ms, double* work, int64_t* groups);

#ifdef __cplusplus
}
#endif

#endif /* FWX_H */
