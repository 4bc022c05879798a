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

template <int T>
__global__ __launch_bounds__(256) void k_batch_diag(int32_t* __restrict__ data,
                                                    const DiagItem* __restrict__ items) {
  const DiagItem it = items[blockIdx.x];
  close_tile<T>(make_view(data + it.base, it.n, it.c0, it.c0));
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

// ---------------------------------------------------------------------------
// plan

namespace {

enum Kind { kDiag64 = 0, kProd64 = 1, kProd128 = 2, kProd32 = 3, kDiag32 = 4, kDiag128 = 5, kKinds = 6 };

struct Launch {
  int kind;
  int64_t first;  // index into the kind's item array
  int32_t count;
};

int diag_kind(int T) { return T == 32 ? kDiag32 : T == 64 ? kDiag64 : kDiag128; }
int prod_kind(int R) { return R == 32 ? kProd32 : R == 64 ? kProd64 : kProd128; }

// The schedule of one graph, step by step (mirrors run_blocked).
void schedule(int64_t base, int32_t n, int32_t T, int32_t kappa, std::vector<std::vector<std::pair<int, ProdItem>>>& psteps,
              std::vector<std::vector<std::pair<int, DiagItem>>>& dsteps) {
  const bool vec = (n % 4 == 0) && (base % 4 == 0);
  const int N = (n + T - 1) / T;
  const int span = kappa * T;
  const int R = (span % 128 == 0) ? 128 : (span % 64 == 0) ? 64 : 32;
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

}  // namespace

}  // namespace fwx

struct fwx_batch_plan {
  int32_t* data = nullptr;
  hipStream_t stream = nullptr;
  std::vector<fwx::Launch> launches;  // in order; a step's launches are contiguous
  std::vector<int32_t> step_of;       // step index of each launch
  fwx::ProdItem* d_prod[fwx::kKinds] = {};
  fwx::DiagItem* d_diag[fwx::kKinds] = {};
  hipGraph_t graph[2] = {};
  hipGraphExec_t exec[2] = {};  // [0] min3 form, [1] pk64 form
  int64_t items = 0;
  int32_t steps = 0;
  int64_t span = 0;  // elements from d_data covering every graph
};

namespace fwx {

template <bool PK>
static int enqueue_plan(fwx_batch_plan* p) {
  for (const Launch& L : p->launches) {
    switch (L.kind) {
      case kDiag32: k_batch_diag<32><<<L.count, 256, 0, p->stream>>>(p->data, p->d_diag[kDiag32] + L.first); break;
      case kDiag64: k_batch_diag<64><<<L.count, 256, 0, p->stream>>>(p->data, p->d_diag[kDiag64] + L.first); break;
      case kDiag128: k_batch_diag<128><<<L.count, 256, 0, p->stream>>>(p->data, p->d_diag[kDiag128] + L.first); break;
      case kProd32: k_batch_prod<32, PK><<<L.count, 256, 0, p->stream>>>(p->data, p->d_prod[kProd32] + L.first); break;
      case kProd64: k_batch_prod<64, PK><<<L.count, 256, 0, p->stream>>>(p->data, p->d_prod[kProd64] + L.first); break;
      case kProd128: k_batch_prod<128, PK><<<L.count, 256, 0, p->stream>>>(p->data, p->d_prod[kProd128] + L.first); break;
      default: return fail(FWX_ERR_INVALID_ARG, "corrupt batch plan");
    }
    FWX_LAUNCHED("k_batch");
  }
  return FWX_OK;
}

static int enqueue_mode(fwx_batch_plan* p, bool pk) { return pk ? enqueue_plan<true>(p) : enqueue_plan<false>(p); }

static void destroy_plan(fwx_batch_plan* p) {
  if (!p) return;
  for (int m = 0; m < 2; ++m) {
    if (p->exec[m]) (void)hipGraphExecDestroy(p->exec[m]);
    if (p->graph[m]) (void)hipGraphDestroy(p->graph[m]);
  }
  for (int k = 0; k < kKinds; ++k) {
    if (p->d_prod[k]) (void)hipFree(p->d_prod[k]);
    if (p->d_diag[k]) (void)hipFree(p->d_diag[k]);
  }
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

  std::vector<std::vector<std::vector<std::pair<int, ProdItem>>>> P(count);
  std::vector<std::vector<std::vector<std::pair<int, DiagItem>>>> Dg(count);
  int32_t steps = 0;
  for (int32_t q = 0; q < count; ++q) {
    const int32_t g = order[q];
    int32_t T, kappa;
    tuned_config(sizes[g], &T, &kappa);
    schedule(offsets[g], sizes[g], T, kappa, P[q], Dg[q]);
    steps = std::max<int32_t>(steps, (int32_t)P[q].size());
  }

  fwx_batch_plan* p = new fwx_batch_plan();
  p->data = d_data;
  for (int32_t g = 0; g < count; ++g)
    p->span = std::max<int64_t>(p->span, offsets[g] + (int64_t)sizes[g] * sizes[g]);
  p->stream = as_stream(stream);
  p->steps = steps;
  std::vector<ProdItem> prod[kKinds];
  std::vector<DiagItem> diag[kKinds];
  for (int32_t j = 0; j < steps; ++j) {
    std::vector<ProdItem> sp[kKinds];
    std::vector<DiagItem> sd[kKinds];
    for (int32_t q = 0; q < count; ++q) {
      if (j >= (int32_t)P[q].size()) continue;
      for (auto& kv : P[q][j]) sp[kv.first].push_back(kv.second);
      for (auto& kv : Dg[q][j]) sd[kv.first].push_back(kv.second);
    }
    for (int k = 0; k < kKinds; ++k) {
      if (!sd[k].empty()) {
        p->launches.push_back(Launch{k, (int64_t)diag[k].size(), (int32_t)sd[k].size()});
        p->step_of.push_back(j);
        diag[k].insert(diag[k].end(), sd[k].begin(), sd[k].end());
      }
      if (!sp[k].empty()) {
        p->launches.push_back(Launch{k, (int64_t)prod[k].size(), (int32_t)sp[k].size()});
        p->step_of.push_back(j);
        prod[k].insert(prod[k].end(), sp[k].begin(), sp[k].end());
      }
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
  *out = p;
  return FWX_OK;
}

int fwx_batch_plan_run(fwx_batch_plan* p, int32_t use_graph) {
  clear_error();
  if (!p) return fail(FWX_ERR_INVALID_ARG, "null plan");
  // one range check over the buffer span selects the relaxation form
  bool pk = false;
  if (int rc = values_in_pk64_range(p->data, p->span, p->stream, &pk)) return rc;
  const int m = pk ? 1 : 0;
  if (use_graph) {
    if (!p->exec[m]) {
      // a private capture stream: the caller's stream may be the legacy
      // default stream, which cannot be captured
      hipStream_t cap;
      FWX_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
      hipStream_t saved = p->stream;
      p->stream = cap;
      hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
      int rc = FWX_OK;
      if (e == hipSuccess) rc = enqueue_mode(p, pk);
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
      p->graph[m] = g;
      FWX_HIP(hipGraphInstantiate(&p->exec[m], p->graph[m], nullptr, nullptr, 0));
    }
    FWX_HIP(hipGraphLaunch(p->exec[m], p->stream));
  } else {
    if (int rc = enqueue_mode(p, pk)) return rc;
  }
  FWX_HIP(hipStreamSynchronize(p->stream));
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
