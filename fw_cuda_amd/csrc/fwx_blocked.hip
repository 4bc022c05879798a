// fwx_blocked.hip -- blocked and multi-layer Floyd-Warshall for gfx950.
//
// Reference behaviour replaced: fwTilingGPU (src/fw_tiling_gpu.cu:253-284)
// and fwMultiLayerTilingGPU (src/fw_multi_layer_tiling_gpu.cu:507-577).  Both
// compute the all-pairs shortest paths of the INF-completed graph in place;
// with no negative cycle and no int32 overflow that matrix is unique, so the
// blocked orders below give bit-for-bit the result of the k-i-j loop
// (src/fw_baseline_cpu.cpp:3-16).
//
// The matrix is cut into T x T k-tiles (T = 32, 64, 128; N = ceil(n/T)).
// A stage is kappa consecutive k-tiles K = {kb .. kb+B-1}.  Per stage:
//
//   for each k-tile c in K, in order (the "lead" rounds):
//     diag  : close tile (c,c) with T dependent k steps      1 workgroup
//     panel : row c and column c tiles  C <- min(C, Dcc+ (x) C) or
//             min(C, C (x) Dcc+)  (Dcc+ is closed, so one product is exact)
//     cross : every other tile with a row or column in K
//             C <- min(C, D[i,c] (x) D[c,j])
//   rest  : every tile outside K's rows and columns
//             C <- min(C, D[I,K] (x) D[K,J])    -- depth B*T in one pass
//
// The lead rounds are rounds of classic blocked FW restricted to the cross
// region (rows or columns in K), which is closed under those updates.  The
// rest pass uses the stage's final panels: for i, j outside K every path
// through a vertex of K splits at it into two paths whose lengths the final
// panels bound, so one product of depth B*T gives the same values as B
// rounds.  Each rest tile therefore makes one HBM round trip per kappa
// k-tiles (the point of the reference's multi-layer scheme), with C held in
// VGPRs across the whole depth.  kappa = 1 is the classic three-phase
// algorithm, which is what tiling_gpu runs.
#include <hip/hip_runtime.h>

#include "fwx_device.hpp"
#include "fwx_internal.hpp"
#include "fwx_tile.hpp"

namespace fwx {

constexpr int kChunk = 32;

// ---------------------------------------------------------------------------
// diag: close tile (c, c).  Thread (tx, ty) keeps cells (ty+16r, tx+16q) in
// VGPRs.  Step k: the owners of row k and column k publish them to a
// double-buffered LDS line, one barrier, every thread relaxes its cells.
// Double buffering makes one barrier per step enough: buffer k&1 is next
// rewritten at step k+2, after every thread has passed barrier k+1, hence
// finished reading it.  Steps with k >= n (padding) are skipped, so padded
// vertices are never intermediates.
template <int T>
__global__ __launch_bounds__(256) void k_diag(int32_t* __restrict__ D, int32_t n, int32_t c) {
  constexpr int TM = T / 16;
  __shared__ int32_t rowbuf[2][T];
  __shared__ int32_t colbuf[2][T];
  const int tx = threadIdx.x & 15;
  const int ty = threadIdx.x >> 4;
  const int32_t base = c * T;
  const View V = make_view(D, n, base, base);
  int32_t acc[TM][TM];
#pragma unroll
  for (int r = 0; r < TM; ++r)
#pragma unroll
    for (int q = 0; q < TM; ++q)
      acc[r][q] = V.in(ty + 16 * r, tx + 16 * q) ? *V.at(ty + 16 * r, tx + 16 * q) : kInf;
  const int kmax = min(T, n - base);
#pragma unroll
  for (int kk = 0; kk < TM; ++kk) {
    for (int k16 = 0; k16 < 16; ++k16) {
      const int k = kk * 16 + k16;
      if (k >= kmax) break;
      const int b = k & 1;
      if (ty == k16) {
#pragma unroll
        for (int q = 0; q < TM; ++q) rowbuf[b][tx + 16 * q] = acc[kk][q];
      }
      if (tx == k16) {
#pragma unroll
        for (int r = 0; r < TM; ++r) colbuf[b][ty + 16 * r] = acc[r][kk];
      }
      __syncthreads();
      int32_t rk[TM], ck[TM];
#pragma unroll
      for (int q = 0; q < TM; ++q) rk[q] = rowbuf[b][tx + 16 * q];
#pragma unroll
      for (int r = 0; r < TM; ++r) ck[r] = colbuf[b][ty + 16 * r];
#pragma unroll
      for (int r = 0; r < TM; ++r)
#pragma unroll
        for (int q = 0; q < TM; ++q) acc[r][q] = min(acc[r][q], wadd(ck[r], rk[q]));
    }
  }
#pragma unroll
  for (int r = 0; r < TM; ++r)
#pragma unroll
    for (int q = 0; q < TM; ++q)
      if (V.in(ty + 16 * r, tx + 16 * q)) *V.at(ty + 16 * r, tx + 16 * q) = acc[r][q];
}

// ---------------------------------------------------------------------------
// panel: blocks [0, N-1) take row tile (c, j), blocks [N-1, 2N-2) column tile
// (i, c).  The tile's own pre-update values are the second operand; it is
// staged chunk by chunk before this block stores anything.
template <int T, bool VEC>
__global__ __launch_bounds__(256) void k_panel(int32_t* __restrict__ D, int32_t n, int32_t c,
                                               int32_t N) {
  using P = TileProduct<T, T, kChunk, VEC>;
  __shared__ typename P::Smem s;
  const int b = blockIdx.x;
  const int32_t base = c * T;
  const View diag = make_view(D, n, base, base);
  View A, B, C;
  if (b < N - 1) {
    const int j = b + (b >= c);
    C = make_view(D, n, base, j * T);
    A = diag;
    B = C;
  } else {
    int i = b - (N - 1);
    i += (i >= c);
    C = make_view(D, n, i * T, base);
    A = C;
    B = diag;
  }
  int32_t acc[P::TM][P::TN];
  P::load_c(C, acc);
  P p;
  p.run(s, A, B, T, acc);
  P::store_c(C, acc);
}

// ---------------------------------------------------------------------------
// cross: tiles with a row or a column in the stage, other than row c and
// column c.  Region 1: stage rows (except c) x all columns (except c),
// (B-1)(N-1) tiles.  Region 2: non-stage rows x stage columns (except c),
// (N-B)(B-1) tiles.
template <int T, bool VEC>
__global__ __launch_bounds__(256) void k_cross(int32_t* __restrict__ D, int32_t n, int32_t c,
                                               int32_t kb, int32_t Bst, int32_t N) {
  using P = TileProduct<T, T, kChunk, VEC>;
  __shared__ typename P::Smem s;
  int b = blockIdx.x;
  const int l = c - kb;
  const int r1 = (Bst - 1) * (N - 1);
  int i, j;
  if (b < r1) {
    int m = b / (N - 1);
    const int jj = b % (N - 1);
    m += (m >= l);
    i = kb + m;
    j = jj + (jj >= c);
  } else {
    b -= r1;
    const int ii = b / (Bst - 1);
    int m = b % (Bst - 1);
    m += (m >= l);
    j = kb + m;
    i = ii + (ii >= kb ? Bst : 0);
  }
  const View A = make_view(D, n, i * T, c * T);
  const View B = make_view(D, n, c * T, j * T);
  const View C = make_view(D, n, i * T, j * T);
  int32_t acc[P::TM][P::TN];
  P::load_c(C, acc);
  P p;
  p.run(s, A, B, T, acc);
  P::store_c(C, acc);
}

// ---------------------------------------------------------------------------
// rest: R x R tiles outside the stage rows/columns, depth S = B*T.
// The (m x m) reduced tile grid is split into 8 contiguous ranges, one per
// XCD (workgroups are dealt to the 8 XCDs round-robin), and each range is
// walked in 8-row groups so concurrently resident tiles share A panels
// (same rows) and B panels (nearby columns) in that XCD's L2.
template <int R, bool VEC>
__global__ __launch_bounds__(256) void k_rest(int32_t* __restrict__ D, int32_t n, int32_t s0,
                                              int32_t S, int32_t sr0, int32_t srn, int32_t m,
                                              int32_t per) {
  using P = TileProduct<R, R, kChunk, VEC>;
  __shared__ typename P::Smem s;
  const int b = blockIdx.x;
  const int L = (b & 7) * per + (b >> 3);
  if (L >= m * m) return;
  constexpr int G = 8;
  const int group = L / (G * m);
  const int first = group * G;
  const int rows = min(m - first, G);
  const int within = L - group * G * m;
  const int Ii = first + within % rows;
  const int Jj = within / rows;
  const int I = Ii + (Ii >= sr0 ? srn : 0);
  const int J = Jj + (Jj >= sr0 ? srn : 0);
  const View A = make_view(D, n, I * R, s0);
  const View B = make_view(D, n, s0, J * R);
  const View C = make_view(D, n, I * R, J * R);
  int32_t acc[P::TM][P::TN];
  P::load_c(C, acc);
  P p;
  p.run(s, A, B, S, acc);
  P::store_c(C, acc);
}

// ---------------------------------------------------------------------------
// host driver

template <int T, bool VEC>
static int lead_round(int32_t* d, int32_t n, int32_t c, int32_t kb, int32_t B, int32_t N,
                      hipStream_t st) {
  k_diag<T><<<1, 256, 0, st>>>(d, n, c);
  FWX_LAUNCHED("k_diag");
  if (N > 1) {
    k_panel<T, VEC><<<2 * (N - 1), 256, 0, st>>>(d, n, c, N);
    FWX_LAUNCHED("k_panel");
  }
  if (B > 1) {
    const int cnt = (B - 1) * (2 * N - B - 1);
    if (cnt > 0) {
      k_cross<T, VEC><<<cnt, 256, 0, st>>>(d, n, c, kb, B, N);
      FWX_LAUNCHED("k_cross");
    }
  }
  return FWX_OK;
}

template <int R, bool VEC>
static int rest_pass(int32_t* d, int32_t n, int32_t s0, int32_t S, hipStream_t st) {
  const int nr = (n + R - 1) / R;
  const int sr0 = s0 / R;
  const int srn = min((S + R - 1) / R, nr - sr0);
  const int m = nr - srn;
  if (m <= 0) return FWX_OK;
  const int64_t total = (int64_t)m * m;
  const int per = (int)((total + 7) / 8);
  k_rest<R, VEC><<<8 * per, 256, 0, st>>>(d, n, s0, S, sr0, srn, m, per);
  FWX_LAUNCHED("k_rest");
  return FWX_OK;
}

template <int T, bool VEC>
static int run_blocked(int32_t* d, int32_t n, int32_t kappa, hipStream_t st) {
  const int N = (n + T - 1) / T;
  const int span = kappa * T;
  const int R = (span % 128 == 0) ? 128 : (span % 64 == 0) ? 64 : 32;
  for (int kb = 0; kb < N; kb += kappa) {
    const int B = min(kappa, N - kb);
    const int s0 = kb * T, S = B * T;
    // algorithmic relaxations: every cell x the stage's real k values
    const double L = (double)min(S, n - s0), nn = (double)n;
    hipEvent_t ev;
    if (int rc = prof_open(st, &ev)) return rc;
    for (int l = 0; l < B; ++l) {
      const int rc = lead_round<T, VEC>(d, n, kb + l, kb, B, N, st);
      if (rc) return rc;
    }
    if (int rc = prof_close(st, ev, FWX_PROF_LEAD, (nn * nn - (nn - L) * (nn - L)) * L)) return rc;
    if (int rc = prof_open(st, &ev)) return rc;
    int rc;
    if (R == 128)
      rc = rest_pass<128, VEC>(d, n, s0, S, st);
    else if (R == 64)
      rc = rest_pass<64, VEC>(d, n, s0, S, st);
    else
      rc = rest_pass<32, VEC>(d, n, s0, S, st);
    if (rc) return rc;
    if (int rc2 = prof_close(st, ev, FWX_PROF_REST, (nn - L) * (nn - L) * L)) return rc2;
  }
  return FWX_OK;
}

int enqueue_blocked(int32_t* d, int32_t n, int32_t tile, int32_t kappa, hipStream_t st) {
  const bool vec = (n % 4 == 0) && ((reinterpret_cast<uintptr_t>(d) & 15u) == 0);
  switch (tile) {
    case 32:
      return vec ? run_blocked<32, true>(d, n, kappa, st) : run_blocked<32, false>(d, n, kappa, st);
    case 64:
      return vec ? run_blocked<64, true>(d, n, kappa, st) : run_blocked<64, false>(d, n, kappa, st);
    case 128:
      return vec ? run_blocked<128, true>(d, n, kappa, st) : run_blocked<128, false>(d, n, kappa, st);
    default:
      return fail(FWX_ERR_INVALID_ARG, "tile must be 32, 64 or 128");
  }
}

}  // namespace fwx
