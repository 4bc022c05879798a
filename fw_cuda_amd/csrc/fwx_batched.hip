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
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <vector>

#include "fwx_device.hpp"
#include "fwx_internal.hpp"
#include "fwx_tile.hpp"

namespace fwx {

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

// Work kinds: a diagonal closure or a tile product, by tile size.
enum Kind { kDiag64 = 0, kProd64 = 1, kProd128 = 2, kProd32 = 3, kDiag32 = 4, kDiag128 = 5, kKinds = 6 };

template <int T, int NTH>
__global__ __launch_bounds__(NTH) void k_batch_diag(int32_t* __restrict__ data,
                                                    const DiagItem* __restrict__ items) {
  const DiagItem it = items[blockIdx.x];
  close_tile<T, NTH>(make_view(data + it.base, it.n, it.c0, it.c0));
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

// ---------------------------------------------------------------------------
// Persistent mode: the whole batch in ONE kernel launch.
//
// Tasks (a diagonal closure or a tile product of one graph) sit in one list
// per XCD.  Each graph belongs to exactly one XCD's list, and a block takes
// tasks only from the list of the XCD it is running on (read from the
// hardware XCC_ID register), so every dependency is between blocks sharing
// one L2: the producer's stores reach that L2 before it bumps the graph's
// `done` counter (an L2 atomic), and the consumer invalidates its CU's L1
// after seeing the count.  No cross-XCD cache maintenance is needed.
//
// A list is ordered by (step, graph), so a task waits only on tasks earlier
// in its list; those were taken by blocks that are running, and the earliest
// unfinished task always has its dependencies done -- the kernel cannot
// deadlock.  Every wait is bounded in wall-clock time (wait_ticks): a wait
// that exceeds it raises a flag that ends every block's loop, and the host
// reports an error.  A list whose XCD got no block leaves tasks unrun, which
// the host detects from the `finished` count.
struct Task {
  int64_t base;
  int32_t n;
  int16_t kind;  // Kind below
  int16_t vec;   // 16-byte global loads are legal for this graph
  int32_t cr, cc;  // C (product) or diagonal (closure) tile origin
  int32_t k0, depth;
  int32_t graph;  // index into done[]
  int32_t need;   // done[graph] must reach this before the task starts
};

// Every counter lives in a 128-byte line of its own that only one XCD
// touches: the L2s of different XCDs are not coherent with each other during
// a kernel, so two XCDs must never update the same word.
constexpr int kMaxQueues = 16;
constexpr int kLine = 32;  // int32 per 128-byte line
// Spin cap of one dependency wait (each spin: two L2 atomics and an
// s_sleep), a second bound next to the wall-clock limit.
constexpr uint32_t kMaxSpins = 1u << 22;
struct BatchCtl {          // one per queue
  int32_t next;            // next task of the queue
  int32_t error;           // a wait of this queue exceeded its limit
  int32_t finished;        // tasks of this queue completed
  int32_t pad[kLine - 3];
};

// The XCC id each wave of each 256-thread block reports (4 entries per block).
__device__ __forceinline__ uint32_t xcc_id() {
  uint32_t x;
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(x));
  return x & 15u;
}

__global__ __launch_bounds__(256) void k_xcc_probe(int32_t* out) {
  if ((threadIdx.x & 63) == 0) out[blockIdx.x * 4 + (threadIdx.x >> 6)] = (int32_t)xcc_id();
}

// done[] holds one counter per graph, kLine ints apart.
__global__ void k_batch_reset(int32_t* __restrict__ done, int32_t ngraphs, BatchCtl* __restrict__ ctl) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ngraphs) done[i * kLine] = 0;
  if (i < kMaxQueues) {
    ctl[i].next = 0;
    ctl[i].error = 0;
    ctl[i].finished = 0;
  }
}

// L2-coherent read of a counter: a compare-and-swap that never matches is a
// read-modify-write performed in this XCD's L2, never served from the L1.
__device__ __forceinline__ int32_t l2_read(int32_t* p) { return atomicCAS(p, INT32_MIN, INT32_MIN); }

template <int R, bool VEC, bool PK>
__device__ __forceinline__ void task_prod(TileSmem<64, 64, 32>& s, int32_t* g, const Task& t) {
  const View A = make_view(g, t.n, t.cr, t.k0);
  const View B = make_view(g, t.n, t.k0, t.cc);
  const View C = make_view(g, t.n, t.cr, t.cc);
  tile_update<R, R, 32, VEC, 16, PK>(*reinterpret_cast<TileSmem<R, R, 32>*>(&s), A, B, C, t.depth);
}

// 256 threads; tasks of tile width <= 64 (the persistent plan caps the tile
// and the rest block at 64), so register pressure stays that of the 64-wide
// product and four blocks share a CU.
template <bool PK>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) void k_batch_persistent(
    int32_t* __restrict__ data, const Task* __restrict__ tasks, const int32_t* __restrict__ qoff,
    const int32_t* __restrict__ xcc_q, int32_t* __restrict__ done, BatchCtl* __restrict__ ctl,
    int64_t wait_ticks) {
  __shared__ TileSmem<64, 64, 32> s;
  __shared__ int32_t s_idx, s_ok, s_q;
  // one reading of the XCC id for the whole block (thread 0's), so every
  // wave takes the same queue and the same exit
  if (threadIdx.x == 0) s_q = xcc_q[xcc_id()];
  __syncthreads();
  const int q = s_q;
  if (q < 0) return;
  const int begin = qoff[q], end = qoff[q + 1];
  for (;;) {
    if (threadIdx.x == 0) {
      const int i = __hip_atomic_fetch_add(&ctl[q].next, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      int ok = 1;
      if (begin + i < end) {
        const int32_t g = tasks[begin + i].graph, need = tasks[begin + i].need;
        const int64_t t0 = wall_clock64();
        uint32_t spins = 0;
        while (l2_read(&done[g * kLine]) < need) {
          if (l2_read(&ctl[q].error) != 0) {
            ok = 0;
            break;
          }
          // bounded twice: wall-clock time, and a spin count (each spin sleeps)
          if (wall_clock64() - t0 > wait_ticks || ++spins > kMaxSpins) {
            __hip_atomic_fetch_or(&ctl[q].error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            ok = 0;
            break;
          }
          __builtin_amdgcn_s_sleep(2);
        }
      }
      s_idx = i;
      s_ok = ok;
    }
    __syncthreads();
    const int i = s_idx;
    if (begin + i >= end || !s_ok) break;
    // the producers' stores are in this XCD's L2; drop stale L1 lines
    asm volatile("buffer_inv sc0" ::: "memory");
    const Task t = tasks[begin + i];
    int32_t* g = data + t.base;
    int32_t* lines = reinterpret_cast<int32_t*>(&s);
    switch (t.kind) {
      case kDiag32: close_tile_in<32, 256>(make_view(g, t.n, t.cr, t.cr), lines); break;
      case kDiag64: close_tile_in<64, 256>(make_view(g, t.n, t.cr, t.cr), lines); break;
      case kProd32:
        if (t.vec) task_prod<32, true, PK>(s, g, t); else task_prod<32, false, PK>(s, g, t);
        break;
      default:
        if (t.vec) task_prod<64, true, PK>(s, g, t); else task_prod<64, false, PK>(s, g, t);
        break;
    }
    // every store of this block has reached the L2 before the count moves
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    __syncthreads();
    if (threadIdx.x == 0) {
      __hip_atomic_fetch_add(&done[t.graph * kLine], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      __hip_atomic_fetch_add(&ctl[q].finished, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
  }
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

// (tile, kappa) of one graph inside a batch.  Graphs with n >= 128 take
// (128, 3): 128-wide rest items of depth 384 on the 512-thread product and
// 128-wide closures, with far fewer dependent steps than the plain path's
// 64-wide entries for these sizes (tuned for one graph alone, where
// parallelism per launch matters more).  Measured on the 512-graph
// configuration (n log-uniform in [64, 2048]): 18.3 ms per batch on the
// tuned entries, 16.5 with (128, 2); with the closures on side branches
// 15.84 for (128, 2), 15.77 for (128, 3), 15.88 for (128, 4).
// Smaller graphs keep the plain path's tuned entry.  FWX_BATCH_DEEP="n0,T,
// kappa" replaces the rule (graphs with n >= n0 take (T, kappa)); "off" uses
// the tuned entries for every graph.
void batch_config(int32_t n, int32_t* T, int32_t* kappa) {
  tuned_config(n, T, kappa);
  int n0 = 128, t = 128, k = 3;
  if (const char* e = std::getenv("FWX_BATCH_DEEP")) {
    if (std::strcmp(e, "off") == 0) return;
    int a = 0, b = 0, c = 0;
    if (std::sscanf(e, "%d,%d,%d", &a, &b, &c) == 3 && (b == 32 || b == 64 || b == 128) && c >= 1 && c <= 16) {
      n0 = a;
      t = b;
      k = c;
    }
  }
  if (n >= n0) {
    *T = t;
    *kappa = k;
  }
}

// FWX_BATCH_DIAG=256 closes 64-wide diagonal tiles with 256 threads
// instead of 512 (read at every run).
bool batch_diag_wide() {
  const char* e = std::getenv("FWX_BATCH_DIAG");
  return !(e && std::strcmp(e, "256") == 0);
}

int diag_kind(int T) { return T == 32 ? kDiag32 : T == 64 ? kDiag64 : kDiag128; }
int prod_kind(int R) { return R == 32 ? kProd32 : R == 64 ? kProd64 : kProd128; }

// The schedule of one graph, step by step (mirrors run_blocked).
void schedule(int64_t base, int32_t n, int32_t T, int32_t kappa, std::vector<std::vector<std::pair<int, ProdItem>>>& psteps,
              std::vector<std::vector<std::pair<int, DiagItem>>>& dsteps, int maxR = 128) {
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
    psteps.back().push_back({prod_kind(RR), it});
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

int prod_block(int kind) { return kind == kProd32 ? 32 : kind == kProd64 ? 64 : 128; }

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
  hipGraph_t graph[2] = {};
  hipGraphExec_t exec[2] = {};  // [0] min3 form, [1] pk64 form
  int64_t items = 0;
  int32_t steps = 0;
  int64_t span = 0;  // elements from d_data covering every graph
  // persistent mode
  int32_t count = 0;           // graphs (entries of d_done)
  fwx::Task* d_tasks = nullptr;
  int32_t* d_qoff = nullptr;   // per-queue task ranges (nq + 1)
  int32_t* d_xccq = nullptr;   // hardware XCC id -> queue (16 entries, -1 unused)
  int32_t* d_done = nullptr;
  fwx::BatchCtl* d_ctl = nullptr;
  int64_t ntasks = 0;
  int32_t nq = 0;
  int32_t pgrid = 0;           // resident blocks of the persistent kernel
  double wait_s = 10.0;        // limit of one dependency wait
  int64_t wall_khz = 0;        // device wall-clock rate
  hipGraph_t pgraph[2] = {};
  hipGraphExec_t pexec[2] = {};
};

namespace fwx {

// Step by step on p->stream; a step's side launches run on p->side, forked
// from p->stream at the step's start and joined back before the next step.
template <bool PK>
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
    switch (L.kind) {
      case kDiag32: k_batch_diag<32, 256><<<L.count, 256, 0, st>>>(p->data, p->d_diag[kDiag32] + L.first); break;
      case kDiag64:
        if (batch_diag_wide())
          k_batch_diag<64, 512><<<L.count, 512, 0, st>>>(p->data, p->d_diag[kDiag64] + L.first);
        else
          k_batch_diag<64, 256><<<L.count, 256, 0, st>>>(p->data, p->d_diag[kDiag64] + L.first);
        break;
      case kDiag128: k_batch_diag<128, 512><<<L.count, 512, 0, st>>>(p->data, p->d_diag[kDiag128] + L.first); break;
      case kProd32: k_batch_prod<32, PK><<<L.count, 256, 0, st>>>(p->data, p->d_prod[kProd32] + L.first); break;
      case kProd64: k_batch_prod<64, PK><<<L.count, 256, 0, st>>>(p->data, p->d_prod[kProd64] + L.first); break;
      case kProd128: k_batch_prod512<PK><<<L.count, 512, 0, st>>>(p->data, p->d_prod[kProd128] + L.first); break;
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

static int enqueue_mode(fwx_batch_plan* p, bool pk) { return pk ? enqueue_plan<true>(p) : enqueue_plan<false>(p); }

template <bool PK>
static int enqueue_persistent(fwx_batch_plan* p) {
  const int nreset = std::max(p->count, kMaxQueues);
  k_batch_reset<<<(nreset + 255) / 256, 256, 0, p->stream>>>(p->d_done, p->count, p->d_ctl);
  FWX_LAUNCHED("k_batch_reset");
  const int64_t ticks = (int64_t)(p->wait_s * (double)p->wall_khz * 1000.0);
  k_batch_persistent<PK><<<p->pgrid, 256, 0, p->stream>>>(p->data, p->d_tasks, p->d_qoff, p->d_xccq, p->d_done,
                                                           p->d_ctl, ticks);
  FWX_LAUNCHED("k_batch_persistent");
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

// Which hardware XCC ids this device's workgroups run on: one probe launch
// over the persistent grid, every wave of every block reporting.  Returns
// the id -> queue map (16 entries), or nq = 0 when the waves of one block
// disagree (then the persistent mode is unavailable on this device).
static int probe_xcc(int grid, std::vector<int32_t>& xcc_q, int32_t* nq) {
  int32_t* d = nullptr;
  const size_t cnt = (size_t)grid * 4;
  FWX_HIP(hipMalloc(reinterpret_cast<void**>(&d), cnt * sizeof(int32_t)));
  k_xcc_probe<<<grid, 256>>>(d);
  hipError_t e = hipGetLastError();
  std::vector<int32_t> h(cnt, -1);
  if (e == hipSuccess) e = hipMemcpy(h.data(), d, cnt * sizeof(int32_t), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return hip_fail(e, "XCC probe", __FILE__, __LINE__);
  xcc_q.assign(16, -1);
  *nq = 0;
  for (size_t b = 0; b < (size_t)grid; ++b) {
    const int32_t x = h[b * 4];
    for (int w = 1; w < 4; ++w)
      if (h[b * 4 + w] != x) {
        xcc_q.assign(16, -1);
        *nq = 0;
        return FWX_OK;
      }
    if (x >= 0 && x < 16 && xcc_q[x] < 0) xcc_q[x] = (*nq)++;
  }
  return FWX_OK;
}

// Per-queue task lists: graphs (LPT order q) dealt to the least-loaded
// queue by n^3; a queue lists its graphs' tasks by (step, graph), each with
// the number of tasks of its graph in earlier steps.  The schedules are the
// plain ones with the tile capped at 64 (kappa scaled to keep the stage
// width, at most 16) and rest blocks of at most 64.
static int build_queues(fwx_batch_plan* p, const std::vector<int32_t>& nsz, const std::vector<int64_t>& noff) {
  int occ = 0, cus = 0, dev = 0;
  FWX_HIP(hipGetDevice(&dev));
  FWX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)k_batch_persistent<true>, 256, 0));
  FWX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  int khz = 0;
  FWX_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
  p->wall_khz = khz > 0 ? khz : 100000;
  p->pgrid = std::max(1, occ) * std::max(1, cus);
  std::vector<int32_t> xcc_q;
  if (int rc = probe_xcc(p->pgrid, xcc_q, &p->nq)) return rc;
  if (p->nq == 0) return FWX_OK;  // waves of one block reported different XCCs
  const int count = (int)nsz.size();
  std::vector<std::vector<std::vector<std::pair<int, ProdItem>>>> P(count);
  std::vector<std::vector<std::vector<std::pair<int, DiagItem>>>> Dg(count);
  int32_t steps = 0;
  for (int q = 0; q < count; ++q) {
    int32_t T, kappa;
    tuned_config(nsz[q], &T, &kappa);
    if (T > 64) {
      kappa = std::min(16, kappa * (T / 64));
      T = 64;
    }
    schedule(noff[q], nsz[q], T, kappa, P[q], Dg[q], 64);
    steps = std::max<int32_t>(steps, (int32_t)P[q].size());
  }
  std::vector<int> qof(count);
  std::vector<double> load(p->nq, 0.0);
  for (int q = 0; q < count; ++q) {
    int best = 0;
    for (int k = 1; k < p->nq; ++k)
      if (load[k] < load[best]) best = k;
    qof[q] = best;
    load[best] += (double)nsz[q] * nsz[q] * nsz[q];
  }
  std::vector<Task> tasks;
  std::vector<int32_t> qoff(p->nq + 1, 0);
  std::vector<int32_t> prefix(count, 0);
  for (int k = 0; k < p->nq; ++k) {
    qoff[k] = (int32_t)tasks.size();
    for (int32_t j = 0; j < steps; ++j) {
      for (int q = 0; q < count; ++q) {
        if (qof[q] != k || j >= (int32_t)P[q].size()) continue;
        const int32_t need = prefix[q];
        for (const auto& kv : Dg[q][j]) {
          if (kv.first != kDiag32 && kv.first != kDiag64) return fail(FWX_ERR_INVALID_ARG, "persistent plan: tile above 64");
          const DiagItem& d = kv.second;
          tasks.push_back(Task{d.base, d.n, (int16_t)kv.first, 0, d.c0, d.c0, d.c0, 0, q, need});
        }
        for (const auto& kv : P[q][j]) {
          if (kv.first != kProd32 && kv.first != kProd64) return fail(FWX_ERR_INVALID_ARG, "persistent plan: block above 64");
          const ProdItem& it = kv.second;
          tasks.push_back(Task{it.base, it.n, (int16_t)kv.first, (int16_t)it.vec, it.cr, it.cc, it.ac, it.depth, q, need});
        }
        prefix[q] += (int32_t)(Dg[q][j].size() + P[q][j].size());
      }
    }
  }
  qoff[p->nq] = (int32_t)tasks.size();
  p->ntasks = (int64_t)tasks.size();
  p->count = count;
  if (int rc = upload(&p->d_tasks, tasks.data(), tasks.size())) return rc;
  if (int rc = upload(&p->d_qoff, qoff.data(), qoff.size())) return rc;
  if (int rc = upload(&p->d_xccq, xcc_q.data(), xcc_q.size())) return rc;
  FWX_HIP(hipMalloc(reinterpret_cast<void**>(&p->d_done), (size_t)std::max(1, count) * kLine * sizeof(int32_t)));
  FWX_HIP(hipMalloc(reinterpret_cast<void**>(&p->d_ctl), kMaxQueues * sizeof(BatchCtl)));
  return FWX_OK;
}

static void destroy_plan(fwx_batch_plan* p) {
  if (!p) return;
  if (p->fork) (void)hipEventDestroy(p->fork);
  if (p->join) (void)hipEventDestroy(p->join);
  if (p->side) (void)hipStreamDestroy(p->side);
  for (int m = 0; m < 2; ++m) {
    if (p->exec[m]) (void)hipGraphExecDestroy(p->exec[m]);
    if (p->graph[m]) (void)hipGraphDestroy(p->graph[m]);
  }
  for (int k = 0; k < kKinds; ++k) {
    if (p->d_prod[k]) (void)hipFree(p->d_prod[k]);
    if (p->d_diag[k]) (void)hipFree(p->d_diag[k]);
  }
  for (int m = 0; m < 2; ++m) {
    if (p->pexec[m]) (void)hipGraphExecDestroy(p->pexec[m]);
    if (p->pgraph[m]) (void)hipGraphDestroy(p->pgraph[m]);
  }
  if (p->d_tasks) (void)hipFree(p->d_tasks);
  if (p->d_qoff) (void)hipFree(p->d_qoff);
  if (p->d_xccq) (void)hipFree(p->d_xccq);
  if (p->d_done) (void)hipFree(p->d_done);
  if (p->d_ctl) (void)hipFree(p->d_ctl);
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
    int32_t T, kappa;
    batch_config(sizes[g], &T, &kappa);
    schedule(offsets[g], sizes[g], T, kappa, P[q], Dg[q]);
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
  std::vector<int32_t> nsz(count);
  std::vector<int64_t> noff(count);
  for (int32_t q = 0; q < count; ++q) {
    nsz[q] = sizes[order[q]];
    noff[q] = offsets[order[q]];
  }
  // persistent mode switched off (fwx_batch_plan_run): no queues, no probe
  (void)nsz;
  (void)noff;
  (void)build_queues;
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
  const int m = pk ? 1 : 0;
  const bool persistent = mode == FWX_BATCH_PERSISTENT || mode == FWX_BATCH_PERSISTENT_GRAPH;
  const bool use_graph = mode == FWX_BATCH_GRAPH || mode == FWX_BATCH_PERSISTENT_GRAPH;
  // The persistent kernel is switched off: on the GPU its first runs did not
  // complete (a run held the device past the test's time limit although
  // every dependency wait is bounded), and the cause is not established.
  if (persistent)
    return fail(FWX_ERR_UNSUPPORTED, "persistent batch mode is switched off in this build");
  hipGraph_t* graph = persistent ? &p->pgraph[m] : &p->graph[m];
  hipGraphExec_t* exec = persistent ? &p->pexec[m] : &p->exec[m];
  auto enqueue = [&](fwx_batch_plan* pl) { return persistent ? enqueue_persistent_mode(pl, pk) : enqueue_mode(pl, pk); };
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
    BatchCtl h[kMaxQueues];
    FWX_HIP(hipMemcpyAsync(h, p->d_ctl, sizeof(h), hipMemcpyDeviceToHost, p->stream));
    FWX_HIP(hipStreamSynchronize(p->stream));
    int64_t finished = 0;
    int error = 0;
    for (int q = 0; q < p->nq; ++q) {
      finished += h[q].finished;
      error |= h[q].error;
    }
    if (error) return fail(FWX_ERR_TIMEOUT, "batched persistent run: a dependency wait exceeded its limit");
    if (finished != p->ntasks)
      return fail(FWX_ERR_TIMEOUT, "batched persistent run: not every task ran (an XCD queue got no workgroup)");
    return FWX_OK;
  }
  FWX_HIP(hipStreamSynchronize(p->stream));
  return FWX_OK;
}

int fwx_batch_plan_set_wait_limit(fwx_batch_plan* p, double seconds) {
  clear_error();
  if (!p || !(seconds >= 0.0)) return fail(FWX_ERR_INVALID_ARG, "null plan or negative limit");
  if (seconds != p->wait_s) {
    // the limit is a kernel argument: drop captured persistent graphs
    for (int m = 0; m < 2; ++m) {
      if (p->pexec[m]) (void)hipGraphExecDestroy(p->pexec[m]);
      if (p->pgraph[m]) (void)hipGraphDestroy(p->pgraph[m]);
      p->pexec[m] = nullptr;
      p->pgraph[m] = nullptr;
    }
  }
  p->wait_s = seconds;
  return FWX_OK;
}

int fwx_batch_plan_queues(const fwx_batch_plan* p, int32_t* queues, int32_t* grid, int64_t* tasks) {
  clear_error();
  if (!p) return fail(FWX_ERR_INVALID_ARG, "null plan");
  if (queues) *queues = p->nq;
  if (grid) *grid = p->pgrid;
  if (tasks) *tasks = p->ntasks;
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
