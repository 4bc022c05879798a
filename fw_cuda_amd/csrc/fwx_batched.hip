// fwx_batched.hip -- many independent graphs of uneven size in one replayable
// launch sequence (no reference counterpart; each graph's result equals the
// plain multi-layer algorithm's, hence the reference's fwBaselineCPU).
//
// Every graph is solved with the same blocked multi-layer schedule as the
// plain path (fwx_blocked.hip): per stage, kappa lead rounds of
// diag -> panel -> cross, then one rest pass.  The schedule of graph g is a
// list of steps; step j of all graphs runs as one launch per work kind
// (closure of a diagonal tile, or a tile product of a given block size),
// whose grid is the union of those graphs' work items.  Items within a step
// are independent (same phase of one graph, or different graphs), so the
// batch needs only max_g(steps) dependent launch groups instead of
// sum_g(steps) launches, and the graphs' work overlaps on the chip.  The
// whole sequence is captured once into a hipGraph and replayed.
//
// Work items live in one device array built at plan creation; each product
// item carries its graph's base offset (64-bit), n, and the origins of its
// C, A and B windows, so one kernel serves panel, cross and rest work.
// Items are ordered largest-graph first within a launch (longest processing
// time first), so the big graphs' tiles start before the small ones' tail.
//
// The persistent mode runs the same steps inside ONE cooperative kernel
// (k_batch_coop below): resident blocks take host-assigned task lists per
// step, and grid-wide barriers stand in for the launch boundaries.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <queue>
#include <vector>

#include "fwx_device.hpp"
#include "fwx_internal.hpp"
#include "fwx_tile.hpp"
#include "fwx_tile16.hpp"

namespace fwx {

namespace cg = cooperative_groups;

struct ProdItem {
  int64_t base;
  int32_t n;
  int32_t cr, cc, ar, ac, br, bc;
  int32_t depth;
  int32_t vec;  // 16-byte global loads are legal for this graph
};

struct DiagItem {
  int64_t base;
  int32_t n;
  int32_t c0;
};

// Work kinds: a diagonal closure or a tile product, by tile size; 128-wide
// products on 512 threads of 8x4 cells (kProd128) or 256 threads of 8x8
// cells (kProd128w), the register micro-tile tuned per size class.
enum Kind {
  kDiag64 = 0,
  kProd64 = 1,
  kProd128 = 2,
  kProd32 = 3,
  kDiag32 = 4,
  kDiag128 = 5,
  kProd128w = 6,
  kKinds = 7
};

template <int T, int NTH>
__global__ __launch_bounds__(NTH) void k_batch_diag(int32_t* __restrict__ data,
                                                    const DiagItem* __restrict__ items) {
  const DiagItem it = items[blockIdx.x];
  close_tile<T, NTH>(make_view(data + it.base, it.n, it.c0, it.c0));
}

// The same closure by pk16 squaring where every cell of the tile is below
// kSq16Small and the tile has no padding (close_tile_sq), else by the T
// steps.
template <int T, int NTH>
__global__ __launch_bounds__(NTH) void k_batch_diag_sq(int32_t* __restrict__ data,
                                                       const DiagItem* __restrict__ items) {
  __shared__ SqSmem16<T, NTH> s;
  const DiagItem it = items[blockIdx.x];
  close_tile_sq<T, NTH>(make_view(data + it.base, it.n, it.c0, it.c0), s);
}

template <int R, bool PK>
__global__ __launch_bounds__(256) void k_batch_prod(int32_t* __restrict__ data,
                                                    const ProdItem* __restrict__ items) {
  __shared__ TileSmem<R, R, 32> s;
  const ProdItem it = items[blockIdx.x];
  int32_t* g = data + it.base;
  const View A = make_view(g, it.n, it.ar, it.ac);
  const View B = make_view(g, it.n, it.br, it.bc);
  const View C = make_view(g, it.n, it.cr, it.cc);
  if (it.vec)
    tile_update<R, R, 32, true, 16, PK>(s, A, B, C, it.depth);
  else
    tile_update<R, R, 32, false, 16, PK>(s, A, B, C, it.depth);
}

// 128 x 128 products on 512 threads (8 x 4 cells each), held to 128 VGPRs
// for two blocks per CU -- the shape of the plain path's rest kernel.
template <bool PK>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_batch_prod512(
    int32_t* __restrict__ data, const ProdItem* __restrict__ items) {
  __shared__ TileSmem<128, 128, 32> s;
  const ProdItem it = items[blockIdx.x];
  int32_t* g = data + it.base;
  const View A = make_view(g, it.n, it.ar, it.ac);
  const View B = make_view(g, it.n, it.br, it.bc);
  const View C = make_view(g, it.n, it.cr, it.cc);
  if (it.vec)
    tile_update<128, 128, 32, true, 32, PK>(s, A, B, C, it.depth);
  else
    tile_update<128, 128, 32, false, 32, PK>(s, A, B, C, it.depth);
}

// R x R products in the pk16 form (fwx_tile16.hpp): 256 threads of 8 x 8
// (R = 128) or 4 x 4 (R = 64) cells, 16-deep chunks; a cell whose result
// reaches the cap keeps its value (the run's check afterwards finds it).
template <int R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_batch_prod16(
    int32_t* __restrict__ data, const ProdItem* __restrict__ items) {
  __shared__ TileSmem16<R, R, 16> s;
  const ProdItem it = items[blockIdx.x];
  int32_t* g = data + it.base;
  const View A = make_view(g, it.n, it.ar, it.ac);
  const View B = make_view(g, it.n, it.br, it.bc);
  const View C = make_view(g, it.n, it.cr, it.cc);
  if (it.vec)
    tile_update16<R, R, 16, true>(s, A, B, C, it.depth);
  else
    tile_update16<R, R, 16, false>(s, A, B, C, it.depth);
}

// After a pk16 run: flags[g] = 1 if graph g holds a cell >= cap.  Block
// (x, g) scans a 1/gridDim.x share of graph g's cells (16-byte loads; every
// graph starts 16-byte aligned).
struct GraphSpan {
  int64_t base;
  int64_t cells;
};

__global__ void k_zero_flags(int32_t* __restrict__ flags, int32_t count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) flags[i] = 0;
}

__global__ __launch_bounds__(256) void k_batch_at_least(const int32_t* __restrict__ data,
                                                        const GraphSpan* __restrict__ spans, int32_t cap,
                                                        int32_t* __restrict__ flags) {
  const GraphSpan sp = spans[blockIdx.y];
  const int32_t* g = data + sp.base;
  const int64_t n4 = sp.cells / 4;
  const int64_t tid = (int64_t)blockIdx.x * 256 + threadIdx.x, stride = (int64_t)gridDim.x * 256;
  int32_t mx = INT32_MIN;
  for (int64_t i = tid; i < n4; i += stride) {
    const int4 v = reinterpret_cast<const int4*>(g)[i];
    mx = max(mx, max(max(v.x, v.y), max(v.z, v.w)));
  }
  for (int64_t i = n4 * 4 + tid; i < sp.cells; i += stride) mx = max(mx, g[i]);
  if (__ballot(mx >= cap) != 0ull && (threadIdx.x & 63) == 0) atomicOr(&flags[blockIdx.y], 1);
}

// ---------------------------------------------------------------------------
// Persistent mode: the whole batch in ONE cooperative launch.
//
// A grid of resident 512-thread blocks (as many as fit on the device, two per
// CU) walks the plan's steps.  In step j the blocks take the step's units
// (one task, or a side branch's product and closure) longest first from a
// shared counter, then every block meets at a grid-wide barrier before step
// j+1.  The steps and their side
// branches are those of the step launches above -- a hoisted closure follows
// its product in one block's list -- so the barriers carry exactly the
// dependencies the launch boundaries carry there, and the barrier's
// device-scope release and acquire make each step's stores visible to every
// XCD.  The cooperative launch guarantees that every block of the grid is
// resident at once (the runtime refuses a grid that does not fit), so every
// barrier is reached by the whole grid and completes.  A launch without the
// cooperative attribute is detected at entry, before any barrier: every block
// sets *status and returns.
struct CoopTask {
  int64_t base;
  int32_t n;
  int16_t kind;  // kDiag128 or kProd128
  int16_t vec;   // 16-byte global loads are legal for this graph
  int32_t cr, cc, ar, ac, br, bc;
  int32_t depth;
};

// The grid barrier (bar[0] = arrivals, bar[64] = generation, zeroed before
// the launch; *gen, in LDS, is the block's own count of barriers passed).  A device-
// scope fence writes back the XCD's whole L2, about 0.2 us per fence and
// serialised per XCD, so only thread 0 of a block fences: every wave first
// waits for its own stores to complete (into its XCD's L2), the block barrier
// orders that before thread 0's release-increment, whose write-back covers
// them; the last arrival resets the count and publishes the next generation
// with a release store; the others poll it with relaxed loads that bypass the
// L2 (an acquire load per poll would invalidate the L2 each time) and then
// acquire once, invalidating the L2 and the CU's L1 -- which every wave of
// the block shares -- before the block barrier lets them load.  Measured on
// MI355X (fw_cuda_amd/microbench/sync_cost.hip), 512 blocks: 12.6 us per
// barrier, against 52 us for cooperative_groups' grid sync, whose every
// thread fences.  Every block passes the same barriers, and the cooperative
// launch keeps them all resident, so each wait ends; a poll that still
// gives up after 2^24 tries sets bit 2 of *status.
__device__ __forceinline__ void coop_barrier(uint32_t* bar, uint32_t blocks, uint32_t* gen, int32_t* status) {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t g = *gen;
    if (__hip_atomic_fetch_add(bar, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == blocks - 1) {
      __hip_atomic_store(bar, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(bar + 64, g + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      uint32_t spins = 0;
      while (__hip_atomic_load(bar + 64, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g) {
        __builtin_amdgcn_s_sleep(1);
        if (++spins > (1u << 24)) {
          atomicOr(status, 2);
          break;
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    *gen = g + 1;
  }
  __syncthreads();
}

// Every persistent task is 128 wide (coop_config below): one product shape --
// the 512-thread, 8 x 4 cells per thread product of the step launches -- and
// one closure, within the same 128 VGPRs.  The products read the thread index
// opaquely (block_tid<true>): with the plain index, their per-thread offsets
// were hoisted out of the task loop, stayed live across it, and the kernel
// spilled ~50 VGPRs inside the products' inner loops.
template <bool PK>
__device__ __forceinline__ void coop_task(TileSmem<128, 128, 32>& s, int32_t* data, const CoopTask& t) {
  int32_t* g = data + t.base;
  if (t.kind == kDiag128) {
    close_tile_in<128, 512>(make_view(g, t.n, t.cr, t.cr), reinterpret_cast<int32_t*>(&s));
    return;
  }
  const View A = make_view(g, t.n, t.ar, t.ac);
  const View B = make_view(g, t.n, t.br, t.bc);
  const View C = make_view(g, t.n, t.cr, t.cc);
  if (t.vec)
    tile_update<128, 128, 32, true, 32, PK, true>(s, A, B, C, t.depth);
  else
    tile_update<128, 128, 32, false, 32, PK, true>(s, A, B, C, t.depth);
}

// Step j's units are [sfirst[j], sfirst[j+1]), longest first; unit u is
// tasks [ufirst[u], ufirst[u+1]) run in order by one block.  A block takes
// the next unit of the step with one device-scope atomic on ctr[j] (zeroed
// before the launch) until the step's units run out, then waits at the grid
// barrier.  (Claiming the next unit while the current one runs, to hide the
// atomic's round trip, measured slower: 19.5 vs 18.9 ms on the 512-graph
// batch.)
template <bool PK>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_batch_coop(
    int32_t* __restrict__ data, const CoopTask* __restrict__ tasks, const int32_t* __restrict__ ufirst,
    const int32_t* __restrict__ sfirst, int32_t steps, int32_t* __restrict__ ctr, uint32_t* __restrict__ bar,
    int32_t* __restrict__ status) {
  __shared__ TileSmem<128, 128, 32> s;
  __shared__ int32_t s_unit;
  __shared__ uint32_t s_gen;
  if (!cg::this_grid().is_valid()) {
    if (threadIdx.x == 0) atomicOr(status, 1);
    return;
  }
  if (threadIdx.x == 0) s_gen = 0;
  for (int32_t j = 0; j < steps; ++j) {
    const int32_t u0 = sfirst[j], u1 = sfirst[j + 1];
    for (;;) {
      // every thread read the previous s_unit before the barrier that ended
      // the previous unit's last task (or the previous step's grid barrier)
      if (threadIdx.x == 0) s_unit = u0 + atomicAdd(&ctr[j], 1);
      __syncthreads();
      const int32_t u = s_unit;
      if (u >= u1) break;
      const int32_t t1 = ufirst[u + 1];
      for (int32_t t = ufirst[u]; t < t1; ++t) {
        coop_task<PK>(s, data, tasks[t]);
        __syncthreads();  // the next task restages the LDS the last one read
      }
    }
    coop_barrier(bar, gridDim.x, &s_gen, status);
  }
}

// Clears the claim counters of every step, the barrier words and the status
// word before a persistent run (a kernel, so that a captured graph holds no
// memset node).
__global__ void k_coop_reset(int32_t* __restrict__ ctr, int32_t steps, uint32_t* __restrict__ bar,
                             int32_t* __restrict__ status) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < steps) ctr[i] = 0;
  if (i < 128) bar[i] = 0;
  if (i == 0) *status = 0;
}

// ---------------------------------------------------------------------------
// plan

namespace {


struct Launch {
  int kind;
  int64_t first;  // index into the kind's item array
  int32_t count;
  bool side;      // on the side stream, beside the step's other launches
};

// (tile, kappa) of one graph inside a batch and the threads of its 128-wide
// products, from the batched table the autotuner writes by timing whole
// batches of the 512-graph mix (fwx_tune.cpp; FWX_BATCH_TABLE overrides it).
// The hand rule it replaced gave graphs with n >= 128 (128, 3) on 512
// threads and smaller ones the plain path's entries.  FWX_BATCH_DEEP="n0,T,
// kappa" (graphs with n >= n0 take (T, kappa), the others the plain table's
// entries) and "off" (the plain table for every graph), both on 512-thread
// 128-wide products, override the table; the tests use them to cover other
// plan shapes.
void batch_config(int32_t n, int32_t* T, int32_t* kappa, int32_t* rest_threads) {
  if (const char* e = std::getenv("FWX_BATCH_DEEP")) {
    tuned_config(n, T, kappa);
    *rest_threads = 512;
    if (std::strcmp(e, "off") == 0) return;
    int a = 0, b = 0, c = 0;
    if (std::sscanf(e, "%d,%d,%d", &a, &b, &c) == 3 && (b == 32 || b == 64 || b == 128) && c >= 1 && c <= 16) {
      if (n >= a) {
        *T = b;
        *kappa = c;
      }
      return;
    }
  }
  batch_tuned_config(n, T, kappa, rest_threads);
}

// FWX_BATCH_DIAG=256 closes 64-wide diagonal tiles with 256 threads
// instead of 512 (read at every run).
bool batch_diag_wide() {
  const char* e = std::getenv("FWX_BATCH_DIAG");
  return !(e && std::strcmp(e, "256") == 0);
}

// FWX_DIAG_SQ=0: every diagonal closure by the T steps (read when a plan
// enqueues or captures its launches).
bool batch_diag_sq() {
  const char* e = std::getenv("FWX_DIAG_SQ");
  return !(e && e[0] == '0');
}

int diag_kind(int T) { return T == 32 ? kDiag32 : T == 64 ? kDiag64 : kDiag128; }
int prod_kind(int R, int rest_threads) {
  return R == 32 ? kProd32 : R == 64 ? kProd64 : rest_threads == 256 ? kProd128w : kProd128;
}

// The schedule of one graph, step by step (mirrors run_blocked).
void schedule(int64_t base, int32_t n, int32_t T, int32_t kappa, std::vector<std::vector<std::pair<int, ProdItem>>>& psteps,
              std::vector<std::vector<std::pair<int, DiagItem>>>& dsteps, int rest_threads = 512, int maxR = 128) {
  const bool vec = (n % 4 == 0) && (base % 4 == 0);
  const int N = (n + T - 1) / T;
  const int span = kappa * T;
  const int R = std::min(maxR, (span % 128 == 0) ? 128 : (span % 64 == 0) ? 64 : 32);
  auto step = [&]() {
    psteps.emplace_back();
    dsteps.emplace_back();
  };
  auto prod = [&](int RR, int cr, int cc, int ar, int ac, int br, int bc, int depth) {
    ProdItem it{base, n, cr, cc, ar, ac, br, bc, depth, vec ? 1 : 0};
    psteps.back().push_back({prod_kind(RR, rest_threads), it});
  };
  for (int kb = 0; kb < N; kb += kappa) {
    const int B = std::min(kappa, N - kb);
    for (int l = 0; l < B; ++l) {
      const int c = kb + l;
      step();
      dsteps.back().push_back({diag_kind(T), DiagItem{base, n, c * T}});
      if (N > 1) {
        step();
        for (int j = 0; j < N; ++j)
          if (j != c) prod(T, c * T, j * T, c * T, c * T, c * T, j * T, T);
        for (int i = 0; i < N; ++i)
          if (i != c) prod(T, i * T, c * T, i * T, c * T, c * T, c * T, T);
      }
      if (B > 1) {
        bool any = false;
        for (int i = 0; i < N; ++i)
          for (int j = 0; j < N; ++j) {
            const bool cross = (i >= kb && i < kb + B) || (j >= kb && j < kb + B);
            if (!cross || i == c || j == c) continue;
            if (!any) step();
            any = true;
            prod(T, i * T, j * T, i * T, c * T, c * T, j * T, T);
          }
      }
    }
    const int s0 = kb * T, S = B * T;
    const int nr = (n + R - 1) / R;
    const int sr0 = s0 / R;
    const int srn = std::min((S + R - 1) / R, nr - sr0);
    if (nr - srn > 0) {
      step();
      for (int I = 0; I < nr; ++I) {
        if (I >= sr0 && I < sr0 + srn) continue;
        for (int J = 0; J < nr; ++J) {
          if (J >= sr0 && J < sr0 + srn) continue;
          prod(R, I * R, J * R, I * R, s0, s0, J * R, S);
        }
      }
    }
  }
}

using ProdSteps = std::vector<std::vector<std::pair<int, ProdItem>>>;
using DiagSteps = std::vector<std::vector<std::pair<int, DiagItem>>>;

int prod_block(int kind) { return kind == kProd32 ? 32 : kind == kProd64 ? 64 : 128; }  // kProd128, kProd128w

// Side branches of one graph's schedule.  A closure step that follows a
// product step needs, from that step, only the item whose C block holds the
// diagonal tile (the next lead round's tile in a cross step, or the next
// stage's first diagonal tile in a rest step).  That item and the closure
// move into a side branch of the product step, and the closure step goes:
// the closure then runs beside the step's other products instead of after
// them.  The side item's C block is outside the row and column panels the
// other items read, and its operands are panels no item of the step writes,
// so the two branches touch disjoint cells; the next step waits for both.
// Writes the side items (sps) and side closures (sds) per step.
void hoist_closures(ProdSteps& ps, DiagSteps& ds, ProdSteps& sps, DiagSteps& sds, int T) {
  sps.assign(ps.size(), {});
  sds.assign(ps.size(), {});
  size_t j = 1;
  while (j < ps.size()) {
    bool moved = false;
    if (ds[j].size() == 1 && ps[j].empty() && sds[j - 1].empty() && !ps[j - 1].empty()) {
      const int c0 = ds[j][0].second.c0;
      auto& prev = ps[j - 1];
      for (size_t i = 0; i < prev.size(); ++i) {
        const ProdItem& it = prev[i].second;
        const int RR = prod_block(prev[i].first);
        if (it.cr <= c0 && c0 + T <= it.cr + RR && it.cc <= c0 && c0 + T <= it.cc + RR) {
          sps[j - 1].push_back(prev[i]);
          prev.erase(prev.begin() + (std::ptrdiff_t)i);
          sds[j - 1].push_back(ds[j][0]);
          ps.erase(ps.begin() + (std::ptrdiff_t)j);
          ds.erase(ds.begin() + (std::ptrdiff_t)j);
          sps.erase(sps.begin() + (std::ptrdiff_t)j);
          sds.erase(sds.begin() + (std::ptrdiff_t)j);
          moved = true;
          break;
        }
      }
    }
    if (!moved) ++j;
  }
}

// FWX_BATCH_HOIST=0 keeps every closure in a step of its own (read at plan
// creation).
bool hoist_enabled() {
  const char* e = std::getenv("FWX_BATCH_HOIST");
  return !(e && std::strcmp(e, "0") == 0);
}

}  // namespace

}  // namespace fwx

struct fwx_batch_plan {
  int32_t* data = nullptr;
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;          // side branches (hoisted closures)
  hipEvent_t fork = nullptr, join = nullptr;
  std::vector<fwx::Launch> launches;  // in order; a step's launches are contiguous, side ones first
  std::vector<int32_t> step_of;       // step index of each launch
  fwx::ProdItem* d_prod[fwx::kKinds] = {};
  fwx::DiagItem* d_diag[fwx::kKinds] = {};
  hipGraph_t graph[3] = {};
  hipGraphExec_t exec[3] = {};  // [0] min3 form, [1] pk64 form, [2] pk64 with pk16 64- and 128-wide products
  // the pk16 check: the graphs with 64- or 128-wide products (the only ones
  // a pk16 run can leave at the cap), their cell spans, a flag each, and
  // their sizes and offsets for the exact finish of the ones it flags
  fwx::GraphSpan* d_spans = nullptr;
  int32_t* d_flags = nullptr;
  std::vector<int32_t> sizes;
  std::vector<int64_t> offsets;
  int32_t last_form = -1;     // form of the last run (FWX_FORM_*)
  int32_t last_finished = 0;  // graphs the last pk16 run finished exactly
  int64_t items = 0;
  int32_t steps = 0;
  int64_t span = 0;  // elements from d_data covering every graph
  // persistent mode (one cooperative launch)
  fwx::CoopTask* d_ctasks = nullptr;
  int32_t* d_ufirst = nullptr; // first task of each unit (units + 1)
  int32_t* d_sfirst = nullptr; // first unit of each step (csteps + 1)
  int32_t* d_ctr = nullptr;    // per-step claim counters (csteps)
  uint32_t* d_bar = nullptr;   // k_batch_coop's grid barrier: arrivals, generation (256 bytes apart)
  int32_t* d_status = nullptr; // bit 1: launched without the cooperative attribute; bit 2: a barrier wait gave up
  int64_t ctasks = 0;
  int32_t csteps = 0;
  int32_t cgrid = 0;           // resident blocks of k_batch_coop (0: no cooperative launch here)
  hipGraph_t pgraph[2] = {};
  hipGraphExec_t pexec[2] = {};
};

namespace fwx {

// Step by step on p->stream; a step's side launches run on p->side, forked
// from p->stream at the step's start and joined back before the next step.
// P16: the 64- and 128-wide products run in the pk16 form.
template <bool PK, bool P16>
static int enqueue_plan(fwx_batch_plan* p) {
  int cur = -1;
  bool open = false;
  for (size_t li = 0; li < p->launches.size(); ++li) {
    const Launch& L = p->launches[li];
    if (p->step_of[li] != cur) {
      if (open) {
        FWX_HIP(hipEventRecord(p->join, p->side));
        FWX_HIP(hipStreamWaitEvent(p->stream, p->join, 0));
        open = false;
      }
      cur = p->step_of[li];
    }
    hipStream_t st = p->stream;
    if (L.side) {
      if (!open) {
        FWX_HIP(hipEventRecord(p->fork, p->stream));
        FWX_HIP(hipStreamWaitEvent(p->side, p->fork, 0));
        open = true;
      }
      st = p->side;
    }
    const bool sq = batch_diag_sq();
    switch (L.kind) {
      case kDiag32:
        if (sq)
          k_batch_diag_sq<32, 256><<<L.count, 256, 0, st>>>(p->data, p->d_diag[kDiag32] + L.first);
        else
          k_batch_diag<32, 256><<<L.count, 256, 0, st>>>(p->data, p->d_diag[kDiag32] + L.first);
        break;
      case kDiag64:
        if (sq)
          k_batch_diag_sq<64, 512><<<L.count, 512, 0, st>>>(p->data, p->d_diag[kDiag64] + L.first);
        else if (batch_diag_wide())
          k_batch_diag<64, 512><<<L.count, 512, 0, st>>>(p->data, p->d_diag[kDiag64] + L.first);
        else
          k_batch_diag<64, 256><<<L.count, 256, 0, st>>>(p->data, p->d_diag[kDiag64] + L.first);
        break;
      case kDiag128:
        if (sq)
          k_batch_diag_sq<128, 512><<<L.count, 512, 0, st>>>(p->data, p->d_diag[kDiag128] + L.first);
        else
          k_batch_diag<128, 512><<<L.count, 512, 0, st>>>(p->data, p->d_diag[kDiag128] + L.first);
        break;
      case kProd32: k_batch_prod<32, PK><<<L.count, 256, 0, st>>>(p->data, p->d_prod[kProd32] + L.first); break;
      case kProd64:
        if (P16)
          k_batch_prod16<64><<<L.count, 256, 0, st>>>(p->data, p->d_prod[kProd64] + L.first);
        else
          k_batch_prod<64, PK><<<L.count, 256, 0, st>>>(p->data, p->d_prod[kProd64] + L.first);
        break;
      case kProd128:
        if (P16)
          k_batch_prod16<128><<<L.count, 256, 0, st>>>(p->data, p->d_prod[kProd128] + L.first);
        else
          k_batch_prod512<PK><<<L.count, 512, 0, st>>>(p->data, p->d_prod[kProd128] + L.first);
        break;
      case kProd128w:
        if (P16)
          k_batch_prod16<128><<<L.count, 256, 0, st>>>(p->data, p->d_prod[kProd128w] + L.first);
        else
          k_batch_prod<128, PK><<<L.count, 256, 0, st>>>(p->data, p->d_prod[kProd128w] + L.first);
        break;
      default: return fail(FWX_ERR_INVALID_ARG, "corrupt batch plan");
    }
    FWX_LAUNCHED("k_batch");
  }
  if (open) {
    FWX_HIP(hipEventRecord(p->join, p->side));
    FWX_HIP(hipStreamWaitEvent(p->stream, p->join, 0));
  }
  return FWX_OK;
}

// form: 0 min3, 1 pk64, 2 pk64 with pk16 64- and 128-wide products
static int enqueue_mode(fwx_batch_plan* p, int form) {
  return form == 2 ? enqueue_plan<true, true>(p) : form ? enqueue_plan<true, false>(p) : enqueue_plan<false, false>(p);
}

// The check after a pk16 run: graphs holding a cell at or above the cap are
// solved once more, each in the exact pk64 form on its own matrix (walk
// lengths of its input graph, each at most its input value, so the closure
// is the input's).  Synchronises the plan's stream.
static int finish_saturated(fwx_batch_plan* p) {
  const int32_t count = (int32_t)p->sizes.size();
  p->last_finished = 0;
  if (count == 0) {
    FWX_HIP(hipStreamSynchronize(p->stream));
    return FWX_OK;
  }
  k_zero_flags<<<(count + 255) / 256, 256, 0, p->stream>>>(p->d_flags, count);
  FWX_LAUNCHED("k_zero_flags");
  k_batch_at_least<<<dim3(16, (unsigned)count), 256, 0, p->stream>>>(p->data, p->d_spans, (int32_t)kP16Cap,
                                                                     p->d_flags);
  FWX_LAUNCHED("k_batch_at_least");
  std::vector<int32_t> flags(count, 0);
  FWX_HIP(hipMemcpyAsync(flags.data(), p->d_flags, count * sizeof(int32_t), hipMemcpyDeviceToHost, p->stream));
  FWX_HIP(hipStreamSynchronize(p->stream));
  p->last_finished = 0;
  for (int32_t g = 0; g < count; ++g) {
    if (!flags[g]) continue;
    ++p->last_finished;
    if (int rc = enqueue_blocked_exact(p->data + p->offsets[g], p->sizes[g], p->stream)) return rc;
  }
  if (p->last_finished) FWX_HIP(hipStreamSynchronize(p->stream));
  return FWX_OK;
}

// pk16 128-wide batch products; FWX_PK16=0 keeps pk64 (read at every run).
static bool batch_pk16_enabled() {
  const char* e = std::getenv("FWX_PK16");
  return !(e && e[0] == '0');
}

template <bool PK>
static int enqueue_persistent(fwx_batch_plan* p) {
  k_coop_reset<<<(std::max(p->csteps, 128) + 255) / 256, 256, 0, p->stream>>>(p->d_ctr, p->csteps, p->d_bar,
                                                                             p->d_status);
  FWX_LAUNCHED("k_coop_reset");
  void* args[] = {&p->data, &p->d_ctasks, &p->d_ufirst, &p->d_sfirst, &p->csteps, &p->d_ctr, &p->d_bar, &p->d_status};
  FWX_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_batch_coop<PK>), dim3(p->cgrid), dim3(512),
                                     args, 0, p->stream));
  return FWX_OK;
}

static int enqueue_persistent_mode(fwx_batch_plan* p, bool pk) {
  return pk ? enqueue_persistent<true>(p) : enqueue_persistent<false>(p);
}

template <class T>
static int upload(T** dst, const T* src, size_t count) {
  FWX_HIP(hipMalloc(reinterpret_cast<void**>(dst), std::max<size_t>(1, count) * sizeof(T)));
  if (count) FWX_HIP(hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice));
  return FWX_OK;
}

// (tile, kappa) of a graph in the persistent plan: 128-wide tiles for every
// graph (the kernel's one product shape, on 512 threads); graphs whose step
// plan entry is 128 wide keep its kappa, the others are one 128-wide closure
// per 128 x 128 diagonal block with kappa 1.
static void coop_config(int32_t n, int32_t* T, int32_t* kappa) {
  int32_t threads = 512;
  batch_config(n, T, kappa, &threads);
  if (*T != 128) {
    *T = 128;
    *kappa = 1;
  }
}

// Estimated microseconds of one task on one of two blocks sharing a CU, to
// order each step's units longest first: 128-wide products at ~6e10
// relaxations/s per block plus a fixed cost, 128-wide closures by their
// measured latency-bound time (128 dependent steps).
static double coop_cost(const CoopTask& t) {
  if (t.kind == kDiag128) return 57.0;
  return 3.0 + 128.0 * 128.0 * t.depth / 6.0e4;
}

static CoopTask coop_prod(int kind, const ProdItem& it) {
  return CoopTask{it.base, it.n, (int16_t)kind, (int16_t)it.vec, it.cr, it.cc, it.ar, it.ac, it.br, it.bc, it.depth};
}

static CoopTask coop_diag(int kind, const DiagItem& d) {
  return CoopTask{d.base, d.n, (int16_t)kind, 0, d.c0, d.c0, d.c0, d.c0, d.c0, d.c0, 0};
}

// The persistent plan: every graph's steps (coop_config, side branches as in
// the step plan), merged by step.  A unit is one task, or a side branch's
// product followed by its closure (kept in order in one block).  Each step's
// units are stored longest first, for the blocks to claim in that order.
// `order` lists the graphs largest first.  cgrid is the number of blocks
// that fit on the device at once (the cooperative grid); it stays 0 when the
// device has no cooperative launch (the persistent modes then return
// FWX_ERR_UNSUPPORTED).
static int build_coop(fwx_batch_plan* p, const int32_t* sizes, const int64_t* offsets,
                      const std::vector<int32_t>& order) {
  int dev = 0, coop = 0, cus = 0;
  FWX_HIP(hipGetDevice(&dev));
  FWX_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
  if (!coop) return FWX_OK;
  FWX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  int o1 = 0, o0 = 0;
  FWX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&o1, reinterpret_cast<const void*>(&k_batch_coop<true>), 512, 0));
  FWX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&o0, reinterpret_cast<const void*>(&k_batch_coop<false>), 512, 0));
  const int G = std::min(o0, o1) * cus;
  if (G <= 0) return FWX_OK;
  const int32_t count = (int32_t)order.size();
  std::vector<ProdSteps> P(count), SP(count);
  std::vector<DiagSteps> Dg(count), SD(count);
  int32_t steps = 0;
  for (int32_t q = 0; q < count; ++q) {
    const int32_t g = order[q];
    int32_t T, kappa;
    coop_config(sizes[g], &T, &kappa);
    schedule(offsets[g], sizes[g], T, kappa, P[q], Dg[q]);
    hoist_closures(P[q], Dg[q], SP[q], SD[q], T);
    steps = std::max<int32_t>(steps, (int32_t)P[q].size());
  }
  std::vector<CoopTask> tasks;
  std::vector<int32_t> ufirst, sfirst;
  for (int32_t j = 0; j < steps; ++j) {
    std::vector<std::vector<CoopTask>> units;
    for (int32_t q = 0; q < count; ++q) {
      if (j >= (int32_t)P[q].size()) continue;
      if (!SP[q][j].empty() || !SD[q][j].empty()) {  // side branch: its products, then its closures
        std::vector<CoopTask> u;
        for (auto& kv : SP[q][j]) u.push_back(coop_prod(kv.first, kv.second));
        for (auto& kv : SD[q][j]) u.push_back(coop_diag(kv.first, kv.second));
        units.push_back(u);
      }
      for (auto& kv : Dg[q][j]) units.push_back({coop_diag(kv.first, kv.second)});
      for (auto& kv : P[q][j]) units.push_back({coop_prod(kv.first, kv.second)});
    }
    std::vector<double> cost(units.size(), 0.0);
    for (size_t u = 0; u < units.size(); ++u)
      for (const CoopTask& t : units[u]) {
        if (t.kind != kDiag128 && t.kind != kProd128)
          return fail(FWX_ERR_INVALID_ARG, "persistent plan: a task that is not 128 wide");
        cost[u] += coop_cost(t);
      }
    std::vector<size_t> idx(units.size());
    for (size_t u = 0; u < idx.size(); ++u) idx[u] = u;
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });
    sfirst.push_back((int32_t)ufirst.size());
    for (size_t u : idx) {
      ufirst.push_back((int32_t)tasks.size());
      tasks.insert(tasks.end(), units[u].begin(), units[u].end());
    }
    if (tasks.size() > (size_t)INT32_MAX / 2) return fail(FWX_ERR_INVALID_ARG, "persistent plan: too many tasks");
  }
  sfirst.push_back((int32_t)ufirst.size());
  ufirst.push_back((int32_t)tasks.size());
  if (int rc = upload(&p->d_ctasks, tasks.data(), tasks.size())) return rc;
  if (int rc = upload(&p->d_ufirst, ufirst.data(), ufirst.size())) return rc;
  if (int rc = upload(&p->d_sfirst, sfirst.data(), sfirst.size())) return rc;
  FWX_HIP(hipMalloc(reinterpret_cast<void**>(&p->d_ctr), (size_t)std::max(1, steps) * sizeof(int32_t)));
  FWX_HIP(hipMalloc(reinterpret_cast<void**>(&p->d_status), sizeof(int32_t)));
  FWX_HIP(hipMalloc(reinterpret_cast<void**>(&p->d_bar), 128 * sizeof(uint32_t)));
  p->ctasks = (int64_t)tasks.size();
  p->csteps = steps;
  p->cgrid = G;
  return FWX_OK;
}

static void destroy_plan(fwx_batch_plan* p) {
  if (!p) return;
  if (p->fork) (void)hipEventDestroy(p->fork);
  if (p->join) (void)hipEventDestroy(p->join);
  if (p->side) (void)hipStreamDestroy(p->side);
  for (int m = 0; m < 3; ++m) {
    if (p->exec[m]) (void)hipGraphExecDestroy(p->exec[m]);
    if (p->graph[m]) (void)hipGraphDestroy(p->graph[m]);
  }
  if (p->d_spans) (void)hipFree(p->d_spans);
  if (p->d_flags) (void)hipFree(p->d_flags);
  for (int k = 0; k < kKinds; ++k) {
    if (p->d_prod[k]) (void)hipFree(p->d_prod[k]);
    if (p->d_diag[k]) (void)hipFree(p->d_diag[k]);
  }
  for (int m = 0; m < 2; ++m) {
    if (p->pexec[m]) (void)hipGraphExecDestroy(p->pexec[m]);
    if (p->pgraph[m]) (void)hipGraphDestroy(p->pgraph[m]);
  }
  if (p->d_ctasks) (void)hipFree(p->d_ctasks);
  if (p->d_ufirst) (void)hipFree(p->d_ufirst);
  if (p->d_sfirst) (void)hipFree(p->d_sfirst);
  if (p->d_ctr) (void)hipFree(p->d_ctr);
  if (p->d_status) (void)hipFree(p->d_status);
  if (p->d_bar) (void)hipFree(p->d_bar);
  delete p;
}

}  // namespace fwx

using namespace fwx;

extern "C" {

int fwx_batch_plan_create(fwx_batch_plan** out, const int32_t* sizes, const int64_t* offsets, int32_t count,
                          int32_t* d_data, void* stream) {
  clear_error();
  if (!out || !sizes || !offsets || !d_data || count <= 0)
    return fail(FWX_ERR_INVALID_ARG, "fwx_batch_plan_create: null argument or empty batch");
  *out = nullptr;
  for (int32_t g = 0; g < count; ++g) {
    if (sizes[g] <= 0 || sizes[g] > FWX_MAX_N) return fail(FWX_ERR_INVALID_ARG, "graph size out of range");
    if (offsets[g] < 0) return fail(FWX_ERR_INVALID_ARG, "negative graph offset");
  }
  if ((reinterpret_cast<uintptr_t>(d_data) & 15u) != 0)
    return fail(FWX_ERR_INVALID_ARG, "batch buffer must be 16-byte aligned");

  // longest processing time first: graphs by n^3, descending
  std::vector<int32_t> order(count);
  for (int32_t g = 0; g < count; ++g) order[g] = g;
  std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return sizes[a] > sizes[b]; });

  std::vector<ProdSteps> P(count), SP(count);
  std::vector<DiagSteps> Dg(count), SD(count);
  const bool hoist = hoist_enabled();
  int32_t steps = 0;
  for (int32_t q = 0; q < count; ++q) {
    const int32_t g = order[q];
    int32_t T, kappa, threads;
    batch_config(sizes[g], &T, &kappa, &threads);
    schedule(offsets[g], sizes[g], T, kappa, P[q], Dg[q], threads);
    if (hoist) {
      hoist_closures(P[q], Dg[q], SP[q], SD[q], T);
    } else {
      SP[q].assign(P[q].size(), {});
      SD[q].assign(P[q].size(), {});
    }
    steps = std::max<int32_t>(steps, (int32_t)P[q].size());
  }

  fwx_batch_plan* p = new fwx_batch_plan();
  p->data = d_data;
  for (int32_t g = 0; g < count; ++g)
    p->span = std::max<int64_t>(p->span, offsets[g] + (int64_t)sizes[g] * sizes[g]);
  p->stream = as_stream(stream);
  p->steps = steps;
  {
    hipError_t e = hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&p->fork, event_flags(false));
    if (e == hipSuccess) e = hipEventCreateWithFlags(&p->join, event_flags(false));
    if (e != hipSuccess) {
      destroy_plan(p);
      return hip_fail(e, "batch plan streams", __FILE__, __LINE__);
    }
  }
  std::vector<ProdItem> prod[kKinds];
  std::vector<DiagItem> diag[kKinds];
  auto add_prod = [&](int k, const std::vector<ProdItem>& v, int32_t j, bool side) {
    if (v.empty()) return;
    p->launches.push_back(Launch{k, (int64_t)prod[k].size(), (int32_t)v.size(), side});
    p->step_of.push_back(j);
    prod[k].insert(prod[k].end(), v.begin(), v.end());
  };
  auto add_diag = [&](int k, const std::vector<DiagItem>& v, int32_t j, bool side) {
    if (v.empty()) return;
    p->launches.push_back(Launch{k, (int64_t)diag[k].size(), (int32_t)v.size(), side});
    p->step_of.push_back(j);
    diag[k].insert(diag[k].end(), v.begin(), v.end());
  };
  for (int32_t j = 0; j < steps; ++j) {
    std::vector<ProdItem> sp[kKinds], ssp[kKinds];
    std::vector<DiagItem> sd[kKinds], ssd[kKinds];
    for (int32_t q = 0; q < count; ++q) {
      if (j >= (int32_t)P[q].size()) continue;
      for (auto& kv : P[q][j]) sp[kv.first].push_back(kv.second);
      for (auto& kv : Dg[q][j]) sd[kv.first].push_back(kv.second);
      for (auto& kv : SP[q][j]) ssp[kv.first].push_back(kv.second);
      for (auto& kv : SD[q][j]) ssd[kv.first].push_back(kv.second);
    }
    // side branch first (its products, then its closures), then the rest
    for (int k = 0; k < kKinds; ++k) add_prod(k, ssp[k], j, true);
    for (int k = 0; k < kKinds; ++k) add_diag(k, ssd[k], j, true);
    for (int k = 0; k < kKinds; ++k) {
      add_diag(k, sd[k], j, false);
      add_prod(k, sp[k], j, false);
    }
  }
  for (int k = 0; k < kKinds; ++k) {
    if (!prod[k].empty()) {
      hipError_t e = hipMalloc(reinterpret_cast<void**>(&p->d_prod[k]), prod[k].size() * sizeof(ProdItem));
      if (e == hipSuccess)
        e = hipMemcpy(p->d_prod[k], prod[k].data(), prod[k].size() * sizeof(ProdItem), hipMemcpyHostToDevice);
      if (e != hipSuccess) {
        destroy_plan(p);
        return hip_fail(e, "batch plan upload", __FILE__, __LINE__);
      }
      p->items += (int64_t)prod[k].size();
    }
    if (!diag[k].empty()) {
      hipError_t e = hipMalloc(reinterpret_cast<void**>(&p->d_diag[k]), diag[k].size() * sizeof(DiagItem));
      if (e == hipSuccess)
        e = hipMemcpy(p->d_diag[k], diag[k].data(), diag[k].size() * sizeof(DiagItem), hipMemcpyHostToDevice);
      if (e != hipSuccess) {
        destroy_plan(p);
        return hip_fail(e, "batch plan upload", __FILE__, __LINE__);
      }
      p->items += (int64_t)diag[k].size();
    }
  }
  if (int rc = build_coop(p, sizes, offsets, order)) {
    destroy_plan(p);
    return rc;
  }
  {
    std::vector<GraphSpan> spans;
    for (int32_t q = 0; q < count; ++q) {
      bool wide = false;
      for (const auto* steps : {&P[q], &SP[q]})
        for (const auto& st : *steps)
          for (const auto& kv : st) wide = wide || kv.first == kProd128 || kv.first == kProd128w || kv.first == kProd64;
      if (!wide) continue;
      const int32_t g = order[q];
      p->sizes.push_back(sizes[g]);
      p->offsets.push_back(offsets[g]);
      spans.push_back(GraphSpan{offsets[g], (int64_t)sizes[g] * sizes[g]});
    }
    int rc = upload(&p->d_spans, spans.data(), spans.size());
    if (!rc) {
      const hipError_t e = hipMalloc(reinterpret_cast<void**>(&p->d_flags), std::max<size_t>(1, spans.size()) * sizeof(int32_t));
      if (e != hipSuccess) rc = hip_fail(e, "batch plan flags", __FILE__, __LINE__);
    }
    if (rc) {
      destroy_plan(p);
      return rc;
    }
  }
  *out = p;
  return FWX_OK;
}

int fwx_batch_plan_run(fwx_batch_plan* p, int32_t mode) {
  clear_error();
  if (!p) return fail(FWX_ERR_INVALID_ARG, "null plan");
  if (mode < FWX_BATCH_DIRECT || mode > FWX_BATCH_PERSISTENT_GRAPH) return fail(FWX_ERR_INVALID_ARG, "unknown batch mode");
  // one range check over the buffer span selects the relaxation form
  bool pk = false;
  if (int rc = values_in_pk64_range(p->data, p->span, p->stream, &pk)) return rc;
  const bool persistent = mode == FWX_BATCH_PERSISTENT || mode == FWX_BATCH_PERSISTENT_GRAPH;
  // the step launches take the pk16 form for their 64- and 128-wide products where
  // pk64 is exact; the persistent kernel keeps pk64
  const int m = !pk ? 0 : (!persistent && batch_pk16_enabled()) ? 2 : 1;
  p->last_form = m;
  p->last_finished = 0;
  const bool use_graph = mode == FWX_BATCH_GRAPH || mode == FWX_BATCH_PERSISTENT_GRAPH;
  if (persistent && p->cgrid <= 0)
    return fail(FWX_ERR_UNSUPPORTED, "persistent batch mode needs a cooperative launch, which this device lacks");
  hipGraph_t* graph = persistent ? &p->pgraph[m] : &p->graph[m];
  hipGraphExec_t* exec = persistent ? &p->pexec[m] : &p->exec[m];
  auto enqueue = [&](fwx_batch_plan* pl) { return persistent ? enqueue_persistent_mode(pl, pk) : enqueue_mode(pl, m); };
  if (use_graph) {
    if (!*exec) {
      // a private capture stream: the caller's stream may be the legacy
      // default stream, which cannot be captured
      hipStream_t cap;
      FWX_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
      hipStream_t saved = p->stream;
      p->stream = cap;
      hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
      int rc = FWX_OK;
      if (e == hipSuccess) rc = enqueue(p);
      hipGraph_t g = nullptr;
      hipError_t e2 = hipStreamEndCapture(cap, &g);
      p->stream = saved;
      (void)hipStreamDestroy(cap);
      if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture", __FILE__, __LINE__);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        if (persistent) {
          (void)hipGetLastError();
          return fail(FWX_ERR_UNSUPPORTED, std::string("the cooperative launch could not be captured into a "
                                                       "hipGraph on this runtime: ") + fwx_last_error());
        }
        return rc;
      }
      if (e2 != hipSuccess) return hip_fail(e2, "hipStreamEndCapture", __FILE__, __LINE__);
      *graph = g;
      FWX_HIP(hipGraphInstantiate(exec, *graph, nullptr, nullptr, 0));
    }
    FWX_HIP(hipGraphLaunch(*exec, p->stream));
  } else {
    if (int rc = enqueue(p)) return rc;
  }
  if (persistent) {
    int32_t status = 0;
    FWX_HIP(hipMemcpyAsync(&status, p->d_status, sizeof(status), hipMemcpyDeviceToHost, p->stream));
    FWX_HIP(hipStreamSynchronize(p->stream));
    if (status & 1)
      return fail(FWX_ERR_UNSUPPORTED, "the persistent kernel ran without a grid barrier (not a cooperative "
                                       "launch); nothing was computed");
    if (status)
      return fail(FWX_ERR_TIMEOUT, "a grid-barrier wait of the persistent kernel gave up; the result is incomplete");
    return FWX_OK;
  }
  if (m == 2) return finish_saturated(p);
  FWX_HIP(hipStreamSynchronize(p->stream));
  return FWX_OK;
}

int fwx_batch_plan_last_form(const fwx_batch_plan* p, int32_t* form, int32_t* exact_graphs) {
  clear_error();
  if (!p || !form || !exact_graphs) return fail(FWX_ERR_INVALID_ARG, "null argument");
  *form = p->last_form;
  *exact_graphs = p->last_finished;
  return FWX_OK;
}

int fwx_batch_plan_persistent_info(const fwx_batch_plan* p, int32_t* grid, int32_t* steps, int64_t* tasks) {
  clear_error();
  if (!p) return fail(FWX_ERR_INVALID_ARG, "null plan");
  if (grid) *grid = p->cgrid;
  if (steps) *steps = p->csteps;
  if (tasks) *tasks = p->ctasks;
  return FWX_OK;
}

int fwx_batch_plan_stats(const fwx_batch_plan* p, int32_t* steps, int32_t* launches, int64_t* items) {
  clear_error();
  if (!p) return fail(FWX_ERR_INVALID_ARG, "null plan");
  if (steps) *steps = p->steps;
  if (launches) *launches = (int32_t)p->launches.size();
  if (items) *items = p->items;
  return FWX_OK;
}

int fwx_batch_plan_destroy(fwx_batch_plan* p) {
  clear_error();
  destroy_plan(p);
  return FWX_OK;
}

int fwx_batch_loop(const int32_t* sizes, const int64_t* offsets, int32_t count, int32_t* d_data,
                   void* stream) {
  clear_error();
  if (!sizes || !offsets || !d_data || count <= 0)
    return fail(FWX_ERR_INVALID_ARG, "fwx_batch_loop: null argument or empty batch");
  for (int32_t g = 0; g < count; ++g) {
    if (int rc = fwx_multi_layer_tiling_gpu(d_data + offsets[g], sizes[g], 0, 0, stream)) return rc;
  }
  return FWX_OK;
}

}  // extern "C"
