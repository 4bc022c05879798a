/*
 * fw_oracle.c -- CPU restatement of the reference's Floyd-Warshall path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product (libfwx.so and its CLIs) never links it and never falls
 * back to it.
 *
 * Every function cites the reference file:line it restates.  How each one
 * is pinned (golden vectors, closed forms, independent algorithms) is listed
 * next to it; the tests that do the pinning live in tests/test_oracle.py.
 *
 * Integer arithmetic follows the reference's `int` semantics on the
 * relaxation `D[i*n+k] + D[k*n+j]` (src/fw_baseline_cpu.cpp:9): the sum is
 * formed with two's-complement wrap (computed through uint32_t so the C
 * code itself has no signed overflow), then compared signed and strictly.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORACLE_INF 1000000000 /* include/constants.h:3 */

static inline int32_t wrap_add(int32_t a, int32_t b) {
  return (int32_t)((uint32_t)a + (uint32_t)b);
}

/*
 * oracle_fw_serial -- restates fwBaselineCPU (src/fw_baseline_cpu.cpp:3-16):
 * k-i-j order, in place, strict `<`, store only on improvement.
 *
 * The reference re-reads D[i*n+k] on every j because the store to D[i*n+j]
 * may alias it (at j == k).  The split loop below is the same computation:
 * columns j < k see the old D[i][k]; at j == k the cell may change (only when
 * D[k][k] < 0); columns j > k see the value after that step, exactly as the
 * reference's in-order j loop does.  Row i == k aliases the b-row with the
 * output row cell by cell, which an in-order loop handles identically.
 *
 * Pinned by: the reference's KATs (tests/golden/kats.json, taken from
 * src/tests.cpp:59-132), bit-equality with the reference's own
 * fwBaselineCPU built into oracle/_ref, brute-force shortest walks on tiny
 * graphs, and Dijkstra (an independent algorithm) on non-negative graphs.
 */
void oracle_fw_serial(int32_t* D, int32_t n) {
  const size_t N = (size_t)n;
  for (size_t k = 0; k < N; ++k) {
    const int32_t* rowk = D + k * N;
    for (size_t i = 0; i < N; ++i) {
      int32_t* rowi = D + i * N;
      int32_t dik = rowi[k];
      for (size_t j = 0; j < k; ++j) {
        int32_t np = wrap_add(dik, rowk[j]);
        if (np < rowi[j]) rowi[j] = np;
      }
      {
        int32_t np = wrap_add(dik, rowk[k]);
        if (np < rowi[k]) rowi[k] = np;
        dik = rowi[k];
      }
      for (size_t j = k + 1; j < N; ++j) {
        int32_t np = wrap_add(dik, rowk[j]);
        if (np < rowi[j]) rowi[j] = np;
      }
    }
  }
}

/*
 * oracle_fw_rowpar -- the same k-i-j relaxation with the i loop split over
 * threads.  When D[k][k] >= 0, step k leaves row k unchanged, so rows are
 * independent within a step and the result is bit-identical to the serial
 * order.  (With a negative cycle the two may differ; callers use it only for
 * graphs without one, and tests check it equals oracle_fw_serial.)
 */
void oracle_fw_rowpar(int32_t* D, int32_t n, int32_t threads) {
  const long N = (long)n;
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
#else
  (void)threads;
#endif
  for (long k = 0; k < N; ++k) {
    const int32_t* rowk = D + k * N;
#pragma omp parallel for schedule(static) num_threads(threads)
    for (long i = 0; i < N; ++i) {
      int32_t* rowi = D + i * N;
      const int32_t dik = rowi[k];
      if (i == k) {
        for (long j = 0; j < N; ++j) {
          int32_t np = wrap_add(rowi[k], rowk[j]);
          if (np < rowi[j]) rowi[j] = np;
        }
        continue;
      }
      for (long j = 0; j < N; ++j) {
        int32_t np = wrap_add(dik, rowk[j]);
        rowi[j] = np < rowi[j] ? np : rowi[j];
      }
    }
  }
}

/* oracle_has_negative_cycle -- restates hasNegativeCycle (src/utils.cpp:5-12). */
int oracle_has_negative_cycle(const int32_t* D, int32_t n) {
  for (int32_t i = 0; i < n; ++i)
    if (D[(size_t)i * n + i] < 0) return 1;
  return 0;
}

/*
 * oracle_dijkstra_rows -- an algorithm independent of Floyd-Warshall, used to
 * pin the FW oracle and to spot-check rows where an O(n^3) oracle is too slow.
 *
 * It solves single-source shortest paths on the *INF-completed* graph that
 * the reference's FW computes over: every off-diagonal cell, INF included,
 * is an edge with that weight (src/fw_baseline_cpu.cpp:9 never treats INF
 * specially).  Dense O(n^2) Dijkstra; requires every weight >= 0 (returns -1
 * otherwise).  dist[s][s] = min(0, D[s][s]) follows FW's own treatment of a
 * non-negative diagonal (it is never improved).  Sums stay <= 2*INF < 2^31.
 */
int oracle_dijkstra_rows(const int32_t* W, int32_t n, const int32_t* src, int32_t nsrc,
                         int32_t* out, int32_t threads) {
  const size_t N = (size_t)n;
  for (size_t c = 0; c < N * N; ++c)
    if (W[c] < 0) return -1;
  int bad = 0;
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
#else
  (void)threads;
#endif
#pragma omp parallel num_threads(threads)
  {
    int64_t* dist = (int64_t*)malloc(N * sizeof(int64_t));
    unsigned char* done = (unsigned char*)malloc(N);
    if (!dist || !done) {
#pragma omp atomic write
      bad = 1;
    } else {
#pragma omp for schedule(dynamic, 1)
      for (int32_t q = 0; q < nsrc; ++q) {
        const size_t s = (size_t)src[q];
        for (size_t v = 0; v < N; ++v) { dist[v] = INT64_MAX; done[v] = 0; }
        dist[s] = 0;
        for (size_t it = 0; it < N; ++it) {
          size_t u = N;
          int64_t best = INT64_MAX;
          for (size_t v = 0; v < N; ++v)
            if (!done[v] && dist[v] < best) { best = dist[v]; u = v; }
          if (u == N) break;
          done[u] = 1;
          const int32_t* wu = W + u * N;
          for (size_t v = 0; v < N; ++v) {
            if (done[v] || v == u) continue;
            int64_t nd = best + (int64_t)wu[v];
            if (nd < dist[v]) dist[v] = nd;
          }
        }
        int32_t* o = out + (size_t)q * N;
        for (size_t v = 0; v < N; ++v) o[v] = (int32_t)dist[v];
        /* the diagonal keeps its input value when that is <= 0 (it is 0 for
           every generated graph); a positive diagonal is improved to the
           shortest cycle through s, as FW would do. */
        {
          int64_t cyc = (int64_t)W[s * N + s];
          for (size_t v = 0; v < N; ++v) {
            if (v == s) continue;
            int64_t c2 = dist[v] + (int64_t)W[v * N + s];
            if (c2 < cyc) cyc = c2;
          }
          o[s] = (int32_t)cyc;
        }
      }
    }
    free(dist);
    free(done);
  }
  return bad ? -2 : 0;
}

/*
 * oracle_certify_rows -- exact check of claimed distance rows without solving
 * anything: an optimality certificate of single-source shortest paths on the
 * INF-completed graph (every off-diagonal cell an edge, as in
 * src/fw_baseline_cpu.cpp:9).  Requires a zero diagonal and every
 * off-diagonal weight in [1, INF] (returns -1 otherwise).  Row q claims
 * d = dist(s, .) for s = src[q]; it passes iff
 *   (a) d[s] == 0 and every d[v] lies in [0, INF], and
 *   (b) d[v] == min over u != v of (d[u] + W[u][v])  for every v != s.
 * (b) gives d[v] <= d[u] + W[u][v] on every edge, hence d <= dist along any
 * shortest path from d[s] = 0; and since weights are >= 1, the argmin u of
 * (b) has d[u] < d[v], so following argmins from v strictly decreases d and
 * can only stop at s: a real walk of weight d[v], hence d >= dist.  So a row
 * passes exactly when it is the shortest-distance row.  Cost n^2 per row, as
 * Dijkstra, but as blocked min-plus sweeps that vectorise; used for spot
 * checks of many rows at the largest sizes.  Writes ok[q] = 1 / 0; returns
 * the number of failing rows (-1: weights outside the domain, -2: no memory).
 */
#define CERT_SB 32   /* sources per block */
#define CERT_VB 2048 /* columns per block */
__attribute__((target_clones("avx2", "default")))
static void cert_block(const int32_t* W, size_t N, const int32_t* const* d, int nq, size_t v0, size_t v1,
                       int32_t* x /* nq x (v1 - v0) */) {
  const size_t w = v1 - v0;
  for (int q = 0; q < nq; ++q)
    for (size_t v = 0; v < w; ++v) x[(size_t)q * w + v] = INT32_MAX;
  for (size_t u = 0; u < N; ++u) {
    const int32_t* wu = W + u * N + v0;
    for (int q = 0; q < nq; ++q) {
      const int32_t a = d[q][u];
      int32_t* xq = x + (size_t)q * w;
      /* u == v is excluded: split the column range around it */
      const size_t cut = (u >= v0 && u < v1) ? u - v0 : w;
      for (size_t v = 0; v < cut; ++v) {
        const int32_t t = a + wu[v];
        xq[v] = t < xq[v] ? t : xq[v];
      }
      for (size_t v = cut + 1; v < w; ++v) {
        const int32_t t = a + wu[v];
        xq[v] = t < xq[v] ? t : xq[v];
      }
    }
  }
}

int oracle_certify_rows(const int32_t* W, int32_t n, const int32_t* src, int32_t nsrc,
                        const int32_t* rows, int32_t* ok, int32_t threads) {
  const size_t N = (size_t)n;
  for (size_t u = 0; u < N; ++u)
    for (size_t v = 0; v < N; ++v) {
      const int32_t wv = W[u * N + v];
      if (u == v ? wv != 0 : (wv < 1 || wv > ORACLE_INF)) return -1;
    }
  /* (a); rows failing it are excluded from the sweeps (their sums could wrap) */
  for (int32_t q = 0; q < nsrc; ++q) {
    const int32_t* d = rows + (size_t)q * N;
    int good = d[src[q]] == 0;
    for (size_t v = 0; v < N && good; ++v) good = d[v] >= 0 && d[v] <= ORACLE_INF;
    ok[q] = good;
  }
  int nomem = 0;
  const int nsb = (nsrc + CERT_SB - 1) / CERT_SB;
  const int nvb = (int)((N + CERT_VB - 1) / CERT_VB);
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
#else
  (void)threads;
#endif
#pragma omp parallel num_threads(threads)
  {
    int32_t* x = (int32_t*)malloc((size_t)CERT_SB * CERT_VB * sizeof(int32_t));
    const int32_t* d[CERT_SB];
    if (!x) {
#pragma omp atomic write
      nomem = 1;
    }
#pragma omp for schedule(dynamic, 1) collapse(2)
    for (int sb = 0; sb < nsb; ++sb)
      for (int vb = 0; vb < nvb; ++vb) {
        const int q0 = sb * CERT_SB;
        const int nq = (nsrc - q0) < CERT_SB ? (nsrc - q0) : CERT_SB;
        int any = 0;
        for (int q = 0; q < nq; ++q) {
          d[q] = rows + (size_t)(q0 + q) * N;
          int okq;
#pragma omp atomic read
          okq = ok[q0 + q];
          any |= okq;
        }
        if (!any || !x) continue;
        const size_t v0 = (size_t)vb * CERT_VB;
        const size_t v1 = v0 + CERT_VB < N ? v0 + CERT_VB : N;
        cert_block(W, N, d, nq, v0, v1, x);
        for (int q = 0; q < nq; ++q) {
          const size_t s = (size_t)src[q0 + q];
          int good = 1;
          for (size_t v = v0; v < v1 && good; ++v)
            if (v != s && x[(size_t)q * (v1 - v0) + (v - v0)] != d[q][v]) good = 0;
          if (!good) {
#pragma omp atomic write
            ok[q0 + q] = 0;
          }
        }
      }
    free(x);
  }
  if (nomem) return -2;
  int bad = 0;
  for (int32_t q = 0; q < nsrc; ++q) bad += !ok[q];
  return bad;
}

/*
 * oracle_brute_walks -- shortest walks of at most n-1 edges by repeated
 * min-plus squaring of (W with a zero diagonal) -- no FW ordering at all.
 * Used on tiny graphs only (O(n^3 log n)).  Valid when there is no negative
 * cycle; the zero diagonal is taken as the empty walk, which matches FW on
 * graphs with a non-negative diagonal.
 */
void oracle_brute_walks(const int32_t* W, int32_t n, int32_t* out) {
  const size_t N = (size_t)n;
  int64_t* A = (int64_t*)malloc(N * N * sizeof(int64_t));
  int64_t* B = (int64_t*)malloc(N * N * sizeof(int64_t));
  for (size_t c = 0; c < N * N; ++c) A[c] = W[c];
  for (size_t i = 0; i < N; ++i) if (A[i * N + i] > 0) A[i * N + i] = 0;
  for (size_t len = 1; len < N; len *= 2) {
    for (size_t i = 0; i < N; ++i)
      for (size_t j = 0; j < N; ++j) {
        int64_t best = A[i * N + j];
        for (size_t k = 0; k < N; ++k) {
          int64_t v = A[i * N + k] + A[k * N + j];
          if (v < best) best = v;
        }
        B[i * N + j] = best;
      }
    memcpy(A, B, N * N * sizeof(int64_t));
  }
  for (size_t c = 0; c < N * N; ++c) out[c] = (int32_t)A[c];
  free(A);
  free(B);
}

/*
 * oracle_philox4x32_10 / oracle_generate_philox -- restatement of the device
 * generator's definition (no reference counterpart: the reference generates
 * with std::mt19937 on the host).  Philox4x32 with 10 rounds as published by
 * Salmon, Moraes, Dror and Shaw ("Parallel random numbers: as easy as 1, 2,
 * 3", SC'11): each round multiplies counter words 0 and 2 by 0xD2511F53 and
 * 0xCD9E8D57, mixes the high halves with words 1 / 3 and the key, and the
 * key advances by 0x9E3779B9 / 0xBB67AE85 between rounds.  Pinned by the
 * known-answer vectors of the Random123 distribution (tests/test_oracle.py).
 * Cell (i, j): Philox(counter = (cell lo, cell hi, 0, 0), key = (seed lo,
 * seed hi)) = (c0, c1, ..); diagonal 0; edge iff c0 < floor(density 2^32)
 * (always for density 1), weight wmin + floor(c1 (wmax - wmin + 1) / 2^32);
 * else INF.
 */
void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

void oracle_generate_philox(int32_t* out, int32_t n, double density, uint64_t seed, int32_t wmin, int32_t wmax) {
  const uint64_t thresh = density >= 1.0 ? (1ull << 32) : (uint64_t)(density * 4294967296.0);
  const uint64_t range = (uint64_t)((int64_t)wmax - (int64_t)wmin + 1);
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  const size_t N = (size_t)n;
  for (size_t i = 0; i < N; ++i)
    for (size_t j = 0; j < N; ++j) {
      const uint64_t cell = (uint64_t)(i * N + j);
      int32_t v;
      if (i == j) {
        v = 0;
      } else {
        const uint32_t ctr[4] = {(uint32_t)cell, (uint32_t)(cell >> 32), 0u, 0u};
        uint32_t r[4];
        oracle_philox4x32_10(ctr, key, r);
        v = (uint64_t)r[0] < thresh ? wmin + (int32_t)(((uint64_t)r[1] * range) >> 32) : ORACLE_INF;
      }
      out[i * N + j] = v;
    }
}

/*
 * oracle_predecessors -- restatement of the int32 predecessor output (no
 * reference counterpart): P[i][j] = the smallest u != j with
 * D[i][u] + W[u][j] == D[i][j] (sums in int64), P[i][i] = -1, -1 if none.
 * Pinned by a hand-worked graph with ties and by the path invariant (walking
 * P from j reaches i and the weights along the walk sum to D[i][j]) on
 * random graphs with positive cycles (tests/test_oracle.py).
 */
void oracle_predecessors(const int32_t* W, const int32_t* D, int32_t* P, int32_t n) {
  const size_t N = (size_t)n;
  for (size_t i = 0; i < N; ++i)
    for (size_t j = 0; j < N; ++j) {
      int32_t best = -1;
      if (i != j) {
        const int64_t d = D[i * N + j];
        for (size_t u = 0; u < N; ++u) {
          if (u == j) continue;
          if ((int64_t)D[i * N + u] + (int64_t)W[u * N + j] == d) {
            best = (int32_t)u;
            break;
          }
        }
      }
      P[i * N + j] = best;
    }
}

/* ------------------------------------------------------------------------ */
/* Differentiable (fp32) variant.  No reference code exists for it; these   */
/* restate the textbook definitions the product's kernels implement.        */
/* ------------------------------------------------------------------------ */

/*
 * oracle_fw_f32_pred -- fp32 Floyd-Warshall with predecessor tracking, over a
 * weight matrix where an absent edge is +inf and the diagonal is taken as 0.
 *   D[i][j] <- D[i][k] + D[k][j] and P[i][j] <- P[k][j]  iff strictly smaller.
 * P[i][j] is the vertex before j on the chosen i->j path (i for a direct
 * edge), -1 when j is unreachable or j == i.
 * Pinned by: comparison with oracle_fw_serial on integer-valued fp32 weights
 * (exact in fp32), and by path consistency (walking P reproduces D).
 */
void oracle_fw_f32_pred(const float* W, float* D, int32_t* P, int32_t n) {
  const size_t N = (size_t)n;
  for (size_t i = 0; i < N; ++i)
    for (size_t j = 0; j < N; ++j) {
      float w = W[i * N + j];
      if (i == j) { D[i * N + j] = 0.0f; P[i * N + j] = -1; }
      else if (isinf(w)) { D[i * N + j] = INFINITY; P[i * N + j] = -1; }
      else { D[i * N + j] = w; P[i * N + j] = (int32_t)i; }
    }
  for (size_t k = 0; k < N; ++k)
    for (size_t i = 0; i < N; ++i) {
      const float dik = D[i * N + k];
      if (isinf(dik)) continue;
      for (size_t j = 0; j < N; ++j) {
        float np = dik + D[k * N + j];
        if (np < D[i * N + j]) { D[i * N + j] = np; P[i * N + j] = P[k * N + j]; }
      }
    }
}

/*
 * oracle_grad_paths -- path-enumeration gradient of L = sum(G * D):
 * for every reachable (s,t), s != t, walk the predecessor chain t -> s and add
 * G[s][t] to dW of every edge on it.  dW must be zeroed by the caller.
 * This is the definition of dL/dW for a piecewise-linear D (each D[s][t] is
 * the sum of the weights on its path) and does not use distance ordering.
 * Returns -1 if a predecessor chain is longer than n (a corrupt P).
 */
int oracle_grad_paths(const int32_t* P, const float* G, float* dW, int32_t n) {
  const size_t N = (size_t)n;
  for (size_t s = 0; s < N; ++s)
    for (size_t t = 0; t < N; ++t) {
      if (t == s) continue;
      int32_t v = (int32_t)t;
      if (P[s * N + (size_t)v] < 0) continue;
      const float g = G[s * N + t];
      size_t steps = 0;
      while (v != (int32_t)s) {
        int32_t u = P[s * N + (size_t)v];
        if (u < 0 || ++steps > N) return -1;
        dW[(size_t)u * N + (size_t)v] += g;
        v = u;
      }
    }
  return 0;
}

/*
 * oracle_fw_f64 -- float64 min-plus closure used for central finite
 * differences of L = sum(G * D) (absent edges +inf, diagonal 0).  k-i-j order.
 * The weights are float64 so that W +/- eps is formed without fp32 rounding.
 */
void oracle_fw_f64(const double* W, double* D, int32_t n) {
  const size_t N = (size_t)n;
  for (size_t i = 0; i < N; ++i)
    for (size_t j = 0; j < N; ++j)
      D[i * N + j] = (i == j) ? 0.0 : W[i * N + j];
  for (size_t k = 0; k < N; ++k)
    for (size_t i = 0; i < N; ++i) {
      const double dik = D[i * N + k];
      if (isinf(dik)) continue;
      for (size_t j = 0; j < N; ++j) {
        double np = dik + D[k * N + j];
        if (np < D[i * N + j]) D[i * N + j] = np;
      }
    }
}

/*
 * oracle_dijkstra_f64_rows -- float64 single-source shortest path lengths of
 * the fp32 graph the differentiable forward takes (fwx_diff_forward): W[u][v]
 * is an edge weight, +inf means no edge, the diagonal is taken as 0.  Row q
 * holds dist(src[q], .) in float64 (+inf where unreachable).  Dense O(n^2)
 * Dijkstra per source, sources spread over threads; every present weight
 * must be >= 0 (returns -1 otherwise, -2 without memory).  An algorithm
 * independent of Floyd-Warshall, in higher precision than the fp32 kernels:
 * it checks their distances at sizes the O(n^3) oracles cannot reach in a
 * test (n = 4096).
 * Pinned by: equality with oracle_fw_f64 (float64 Floyd-Warshall) on random
 * graphs and on integer-valued weights (where both are exact), and a
 * hand-worked graph in tests/test_oracle.py.
 */
int oracle_dijkstra_f64_rows(const float* W, int32_t n, const int32_t* src, int32_t nsrc, double* out,
                             int32_t threads) {
  const size_t N = (size_t)n;
  for (size_t c = 0; c < N * N; ++c)
    if (W[c] < 0.0f) return -1;
  int bad = 0;
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
#else
  (void)threads;
#endif
#pragma omp parallel num_threads(threads)
  {
    double* dist = (double*)malloc(N * sizeof(double));
    unsigned char* done = (unsigned char*)malloc(N);
    if (!dist || !done) {
#pragma omp atomic write
      bad = 1;
    } else {
#pragma omp for schedule(dynamic, 1)
      for (int32_t q = 0; q < nsrc; ++q) {
        const size_t s = (size_t)src[q];
        for (size_t v = 0; v < N; ++v) { dist[v] = INFINITY; done[v] = 0; }
        dist[s] = 0.0;
        for (size_t it = 0; it < N; ++it) {
          size_t u = N;
          double best = INFINITY;
          for (size_t v = 0; v < N; ++v)
            if (!done[v] && dist[v] < best) { best = dist[v]; u = v; }
          if (u == N) break; /* the rest is unreachable */
          done[u] = 1;
          const float* wu = W + u * N;
          for (size_t v = 0; v < N; ++v) {
            if (done[v] || v == u) continue;
            const double nd = best + (double)wu[v];
            if (nd < dist[v]) dist[v] = nd;
          }
        }
        memcpy(out + (size_t)q * N, dist, N * sizeof(double));
      }
    }
    free(dist);
    free(done);
  }
  return bad ? -2 : 0;
}
