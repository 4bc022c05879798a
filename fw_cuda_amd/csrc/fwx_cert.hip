// fwx_cert.hip -- on-device certificate of claimed shortest-distance rows.
//
// The reference checks a GPU result by copying the whole matrix back and
// comparing it with fwBaselineCPU (src/tests.cpp:37-57), which is out of
// reach at n=32768; spot checks of large solves run a host Dijkstra per row.
// This checks any number of rows on the device without solving anything.
// On the INF-completed graph (every off-diagonal cell an edge, as
// src/fw_baseline_cpu.cpp:9 treats it) with a zero diagonal and every
// off-diagonal weight in [1, INF], a row d claimed for source s is exactly
// dist(s, .) iff
//   (a) d[s] == 0 and every d[v] lies in [0, INF], and
//   (b) d[v] == min over u != v of (d[u] + W[u][v])  for every v != s.
// (b) gives d[v] <= d[u] + W[u][v] on every edge, hence d <= dist; and since
// weights are >= 1, the minimising u of (b) has d[u] < d[v], so following
// minimisers from v strictly lowers d and can only stop at s: a real walk of
// weight d[v], hence d >= dist.
//
// (b) is a min-plus product of the claimed rows with W (u == v excluded),
// nsrc * n^2 relaxations: 64 x 64 result tiles through the solver's
// TileProduct in the pk64 form (every operand lies in [0, INF] ⊂ [0, 2^30)
// for rows that pass (a); rows that fail it may produce garbage sums, but
// only in their own results, and they have already failed).  W[v][v] is
// replaced by INF as it is staged, which removes u == v.
#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>

#include "fwx_device.hpp"
#include "fwx_internal.hpp"
#include "fwx_tile.hpp"

namespace fwx {

namespace {

constexpr int kCT = 64;   // result tile: 64 rows x 64 columns
constexpr int kCK = 32;   // depth chunk
constexpr int kCSlots = 64;
__device__ int g_cert_slots[2 * kCSlots];  // per call: [0] domain flag, [1] failing rows

// Any weight outside the domain -> *f = 1.  Blocks stride over rows,
// threads over the row's columns.
__global__ __launch_bounds__(256) void k_cert_domain(const int32_t* __restrict__ W, int32_t n, int* __restrict__ f) {
  bool bad = false;
  for (int64_t u = blockIdx.x; u < n; u += gridDim.x) {
    const int32_t* row = W + u * (int64_t)n;
    for (int32_t v = threadIdx.x; v < n; v += 256) {
      const int32_t w = row[v];
      bad |= (v == u) ? (w != 0) : (w < 1 || w > kInf);
    }
  }
  if (__ballot(bad) != 0ull && (threadIdx.x & 63) == 0) atomicOr(f, 1);
}

// Condition (a), one block per claimed row: ok[q] = 1 or 0.
__global__ __launch_bounds__(256) void k_cert_rows(const int32_t* __restrict__ src, const int32_t* __restrict__ rows,
                                                   int32_t n, int32_t* __restrict__ ok) {
  __shared__ int s_bad;
  const int q = blockIdx.x;
  if (threadIdx.x == 0) s_bad = 0;
  __syncthreads();
  const int32_t s = src[q];
  const int32_t* d = rows + (int64_t)q * n;
  bool bad = (s < 0 || s >= n);
  for (int32_t v = threadIdx.x; v < n; v += 256) {
    const int32_t x = d[v];
    bad |= (x < 0 || x > kInf) || (v == s && x != 0);
  }
  if (__ballot(bad) != 0ull && (threadIdx.x & 63) == 0) atomicOr(&s_bad, 1);
  __syncthreads();
  if (threadIdx.x == 0) ok[q] = s_bad ? 0 : 1;
}

// Condition (b) on one 64 x 64 tile of (claimed rows) x (columns).
template <bool VEC, bool EDGE>
__device__ __forceinline__ void cert_tile(TileSmem<kCT, kCT, kCK>& sm, const int32_t* __restrict__ W,
                                          const int32_t* __restrict__ src, const int32_t* __restrict__ rows,
                                          int32_t n, int32_t nsrc, int32_t q0, int32_t j0, int32_t* __restrict__ ok) {
  using TP = TileProduct<kCT, kCT, kCK, VEC, 16, true, EDGE>;
  // A: the claimed rows q0.. (nsrc x n, row stride n); B: columns j0.. of W
  const View A{const_cast<int32_t*>(rows) + (int64_t)q0 * n, n, nsrc - q0, n};
  const View B{const_cast<int32_t*>(W) + j0, n, n, n - j0};
  int32_t acc[TP::TM][TP::TN];
#pragma unroll
  for (int r = 0; r < TP::TM; ++r)
#pragma unroll
    for (int c = 0; c < TP::TN; ++c) acc[r][c] = 0x7FFFFFFF;
  // B staging unit of this thread (BPASS == 1): k pair row and 4 columns
  constexpr int BQ = kCT / 4;
  static_assert(TP::BPASS == 1 && TP::BU == TP::NT, "one B unit per thread");
  const int brow = 2 * (threadIdx.x / BQ);
  const int bcol = j0 + (threadIdx.x % BQ) * 4;
  TP tp;
  auto exclude = [&](int32_t kc0) {
    const int k0 = kc0 + brow;
    // the staged cells (k0, bcol + e) and (k0 + 1, bcol + e)
    const int e0 = k0 - bcol, e1 = k0 + 1 - bcol;
    int4& r0 = tp.rb0[0];
    int4& r1 = tp.rb1[0];
    r0.x = e0 == 0 ? kInf : r0.x;
    r0.y = e0 == 1 ? kInf : r0.y;
    r0.z = e0 == 2 ? kInf : r0.z;
    r0.w = e0 == 3 ? kInf : r0.w;
    r1.x = e1 == 0 ? kInf : r1.x;
    r1.y = e1 == 1 ? kInf : r1.y;
    r1.z = e1 == 2 ? kInf : r1.z;
    r1.w = e1 == 3 ? kInf : r1.w;
  };
  const int nch = (n + kCK - 1) / kCK;
  tp.fetch(A, B, 0);
  exclude(0);
  tp.stash(sm, 0, A, B);
  __syncthreads();
  for (int ch = 0; ch < nch; ++ch) {
    const int buf = ch & 1;
    if (ch + 1 < nch) tp.fetch(A, B, (ch + 1) * kCK);
    TP::compute(sm, buf, acc);
    if (ch + 1 < nch) {
      exclude((ch + 1) * kCK);
      tp.stash(sm, buf ^ 1, A, B);
    }
    __syncthreads();
  }
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
#pragma unroll
  for (int r = 0; r < TP::TM; ++r) {
    const int q = q0 + ty + 16 * r;
    if (q >= nsrc) continue;
    const int32_t s = src[q];
    const int32_t* d = rows + (int64_t)q * n;
    bool bad = false;
#pragma unroll
    for (int c = 0; c < TP::TN; ++c) {
      const int v = j0 + TP::colmap(tx, c);
      if (v < n && v != s) bad |= acc[r][c] != d[v];
    }
    if (bad) atomicAnd(&ok[q], 0);
  }
}

template <bool VEC>
__global__ __launch_bounds__(256) void k_cert_sweep(const int32_t* __restrict__ W, const int32_t* __restrict__ src,
                                                    const int32_t* __restrict__ rows, int32_t n, int32_t nsrc,
                                                    int32_t* __restrict__ ok) {
  __shared__ TileSmem<kCT, kCT, kCK> sm;
  const int32_t j0 = blockIdx.x * kCT, q0 = blockIdx.y * kCT;
  if (q0 + kCT <= nsrc && j0 + kCT <= n && n % kCK == 0)
    cert_tile<VEC, false>(sm, W, src, rows, n, nsrc, q0, j0, ok);
  else
    cert_tile<VEC, true>(sm, W, src, rows, n, nsrc, q0, j0, ok);
}

__global__ __launch_bounds__(256) void k_cert_count(const int32_t* __restrict__ ok, int32_t nsrc, int* __restrict__ f) {
  int bad = 0;
  for (int32_t q = threadIdx.x; q < nsrc; q += 256) bad += ok[q] == 0;
  atomicAdd(f, bad);
}

__global__ void k_cert_zero(int* f) {
  f[0] = 0;
  f[1] = 0;
}

std::mutex g_mu;
int* g_base[64] = {};
std::atomic<unsigned> g_slot{0};

int cert_slot(int** out) {
  int dev = 0;
  FWX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail(FWX_ERR_UNSUPPORTED, "device index above 63");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_base[dev]) {
    void* p = nullptr;
    FWX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(g_cert_slots)));
    g_base[dev] = static_cast<int*>(p);
  }
  *out = g_base[dev] + 2 * (g_slot.fetch_add(1) % kCSlots);
  return FWX_OK;
}

}  // namespace

}  // namespace fwx

using namespace fwx;

extern "C" {

int fwx_certify_rows(const int32_t* d_W, int32_t n, const int32_t* d_src, int32_t nsrc, const int32_t* d_rows,
                     int32_t* d_ok, int32_t* failed, void* stream) {
  clear_error();
  if (!d_W || !d_src || !d_rows || !d_ok || !failed || n <= 0 || n > FWX_MAX_N || nsrc <= 0)
    return fail(FWX_ERR_INVALID_ARG, "fwx_certify_rows: null pointer, n out of range or nsrc <= 0");
  if ((int64_t)nsrc > 65535LL * kCT) return fail(FWX_ERR_INVALID_ARG, "fwx_certify_rows: nsrc above 4194240");
  hipStream_t st = as_stream(stream);
  int* f = nullptr;
  if (int rc = cert_slot(&f)) return rc;
  k_cert_zero<<<1, 1, 0, st>>>(f);
  FWX_LAUNCHED("k_cert_zero");
  k_cert_domain<<<min(n, 4096), 256, 0, st>>>(d_W, n, f);
  FWX_LAUNCHED("k_cert_domain");
  k_cert_rows<<<nsrc, 256, 0, st>>>(d_src, d_rows, n, d_ok);
  FWX_LAUNCHED("k_cert_rows");
  const dim3 grid((n + kCT - 1) / kCT, (nsrc + kCT - 1) / kCT);
  const bool vec = (n % 4 == 0) && ((reinterpret_cast<uintptr_t>(d_W) & 15u) == 0) &&
                   ((reinterpret_cast<uintptr_t>(d_rows) & 15u) == 0);
  if (vec)
    k_cert_sweep<true><<<grid, 256, 0, st>>>(d_W, d_src, d_rows, n, nsrc, d_ok);
  else
    k_cert_sweep<false><<<grid, 256, 0, st>>>(d_W, d_src, d_rows, n, nsrc, d_ok);
  FWX_LAUNCHED("k_cert_sweep");
  k_cert_count<<<1, 256, 0, st>>>(d_ok, nsrc, f + 1);
  FWX_LAUNCHED("k_cert_count");
  int h[2] = {0, 0};
  FWX_HIP(hipMemcpyAsync(h, f, sizeof(h), hipMemcpyDeviceToHost, st));
  FWX_HIP(hipStreamSynchronize(st));
  if (h[0]) return fail(FWX_ERR_INVALID_ARG, "fwx_certify_rows: W needs a zero diagonal and weights in [1, FWX_INF]");
  *failed = h[1];
  return FWX_OK;
}

}  // extern "C"
