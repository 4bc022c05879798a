// sync_cost.hip -- latency of the building blocks of a small-n solve on the
// GPU (gfx950), each timed with HIP events:
//   * a chain of dependent launches in a hipGraph (1-block and 225-block
//     kernels that do nothing), per launch;
//   * a grid-wide barrier (cooperative launch, cooperative_groups grid sync),
//     per barrier, for 64 .. 512 blocks of 256 threads, with and without each
//     block writing 16 KiB between barriers;
//   * a flag handoff between two workgroups (release store / acquire load at
//     agent scope, the two blocks on different XCDs), per one-way handoff;
//   * the closure of one diagonal tile (close_tile_in of fwx_tile.hpp) and
//     one 64x64x64 tile product (tile_update), looped inside one workgroup, so
//     that the figure holds no launch overhead.
// Every wait is bounded: a poll gives up after 2^22 tries and reports it.
// Usage (repository root, `make -C fw_cuda_amd microbench` first):
//   fw_cuda_amd/bin/sync_cost
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../csrc/fwx_tile.hpp"

namespace cg = cooperative_groups;
using namespace fwx;

#define CK(x)                                                                      \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) {                                                        \
      std::fprintf(stderr, "%s:%d %s: %s\n", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
      std::exit(1);                                                                \
    }                                                                              \
  } while (0)

__global__ void k_empty(int* p) {
  if (p && threadIdx.x == 1023) p[0] = 0;  // never true for the launch shapes used
}

__global__ __launch_bounds__(256) void k_gridsync(int* buf, int iters, int write) {
  cg::grid_group g = cg::this_grid();
  int4* mine = reinterpret_cast<int4*>(buf) + (size_t)blockIdx.x * 1024;  // 16 KiB per block
  for (int it = 0; it < iters; ++it) {
    if (write)
      for (int i = threadIdx.x; i < 1024; i += 256) mine[i] = make_int4(it, i, (int)blockIdx.x, 0);
    g.sync();
  }
}

// A grid barrier of counters in device memory: thread 0 of each block counts
// the block in, the last arrival resets the counter and publishes the next
// generation, the others poll the generation word.  Three variants:
//   V=1  every thread releases (agent scope) before and acquires after; the
//        polls are acquire loads (each one invalidates the XCD's L2);
//   V=2  as 1, but the polls are relaxed loads that bypass the L2, and one
//        acquire follows the poll loop;
//   V=3  as 2, and only thread 0 fences: every wave first waits for its own
//        stores (vmcnt 0), so after the block barrier they are all in the
//        XCD's L2, which thread 0's release writes back; thread 0's acquire
//        invalidates the L2 and the CU's L1, which all waves of the block
//        share, before the block barrier lets them load.
//   V=4  as 3, with two levels of counters: blocks count in per residue of
//        blockIdx mod 8 (the XCD the dispatcher deals a block to), the last
//        of each group counts the group in at the top, and each group polls
//        its own generation word (words 256 bytes apart).
// A poll gives up after 2^22 tries, or as soon as another block gave up (err).
template <int V>
__device__ __forceinline__ void grid_barrier(unsigned* w, unsigned G, unsigned& gen, int* err) {
  if (V >= 3)
    __builtin_amdgcn_s_waitcnt(0);
  else
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned g = gen;
    unsigned* genw = w + 64;
    bool last;
    if (V != 4) {
      last = __hip_atomic_fetch_add(w, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == G - 1;
      if (last) {
        __hip_atomic_store(w, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(genw, g + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      }
    } else {
      const unsigned grp = blockIdx.x & 7, ngrp = G < 8 ? G : 8;
      const unsigned gsize = (G - grp + 7) / 8;
      unsigned* sub = w + 64 * (1 + grp);
      genw = w + 64 * (9 + grp);
      last = false;
      if (__hip_atomic_fetch_add(sub, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == gsize - 1) {
        __hip_atomic_store(sub, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (__hip_atomic_fetch_add(w, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == ngrp - 1) {
          __hip_atomic_store(w, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          for (unsigned k = 0; k < ngrp; ++k)
            __hip_atomic_store(w + 64 * (9 + k), g + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
          last = true;
        }
      }
    }
    if (!last) {
      unsigned spins = 0;
      for (;;) {
        const unsigned cur = V == 1 ? __hip_atomic_load(genw, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT)
                                    : __hip_atomic_load(genw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur != g) break;
        __builtin_amdgcn_s_sleep(1);
        ++spins;
        if (spins > (1u << 22) ||
            ((spins & 255u) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
          atomicOr(err, 1);
          break;
        }
      }
      if (V != 1) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    gen = g + 1;
  }
  __syncthreads();
  if (V < 3) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

template <int V>
__global__ __launch_bounds__(256) void k_mybarrier(int* buf, unsigned* w, int iters, int write, int* err) {
  __shared__ unsigned gen_s;
  unsigned gen = 0;
  int4* mine = reinterpret_cast<int4*>(buf) + (size_t)blockIdx.x * 1024;
  for (int it = 0; it < iters; ++it) {
    if (write)
      for (int i = threadIdx.x; i < 1024; i += 256) mine[i] = make_int4(it, i, (int)blockIdx.x, 0);
    grid_barrier<V>(w, gridDim.x, gen, err);
  }
  if (threadIdx.x == 0) gen_s = gen;
}

// Blocks 0 and 1 hand a counter back and forth: block b acts on values with
// parity b.  Workgroups are dealt to the XCDs round-robin, so the two blocks
// run on different XCDs.
__global__ __launch_bounds__(64) void k_pingpong(int* flag, int rounds, int* err) {
  const int me = blockIdx.x;
  if (threadIdx.x != 0) return;
  for (int r = 0; r < rounds; ++r) {
    const int want = 2 * r + me;
    unsigned spins = 0;
    while (__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != want) {
      if (++spins > (1u << 22)) {
        atomicOr(err, 1);
        return;
      }
    }
    __hip_atomic_store(flag, want + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <int T, int NTH>
__global__ __launch_bounds__(NTH) void k_closures(int32_t* D, int reps) {
  __shared__ int32_t lines[8 * T];
  for (int r = 0; r < reps; ++r) {
    close_tile_in<T, NTH>(make_view(D, T, 0, 0), lines);
    __syncthreads();
  }
}

template <int NTX>
__global__ __launch_bounds__(NTX * 16) void k_products(int32_t* D, int reps) {
  __shared__ TileSmem<64, 64, 32> s;
  // C = the 64x64 tile at (0, 64) of a 64 x 128 matrix, A = the tile at (0, 0), B = C
  for (int r = 0; r < reps; ++r) {
    tile_update<64, 64, 32, true, NTX, true>(s, make_view(D, 128, 0, 0), make_view(D, 128, 0, 64),
                                             make_view(D, 128, 0, 64), 64);
    __syncthreads();
  }
}

static float time_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0;
  CK(hipEventElapsedTime(&ms, a, b));
  return ms;
}

int main() {
  int cus = 0;
  CK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
  std::printf("CUs %d\n", cus);
  hipStream_t st;
  CK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));

  // 1. dependent launches in a graph
  for (int blocks : {1, 225}) {
    constexpr int kLaunches = 48;
    hipGraph_t g;
    hipGraphExec_t x;
    CK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < kLaunches; ++i) k_empty<<<blocks, 256, 0, st>>>(nullptr);
    CK(hipStreamEndCapture(st, &g));
    CK(hipGraphInstantiate(&x, g, nullptr, nullptr, 0));
    for (int w = 0; w < 3; ++w) CK(hipGraphLaunch(x, st));
    CK(hipEventRecord(e0, st));
    for (int rep = 0; rep < 20; ++rep) CK(hipGraphLaunch(x, st));
    CK(hipEventRecord(e1, st));
    CK(hipEventSynchronize(e1));
    std::printf("graph chain of %d empty %d-block launches: %.2f us per launch\n", kLaunches, blocks,
                1e3 * time_ms(e0, e1) / (20.0 * kLaunches));
    CK(hipGraphExecDestroy(x));
    CK(hipGraphDestroy(g));
  }
  // the same on the stream, direct
  {
    for (int i = 0; i < 100; ++i) k_empty<<<1, 256, 0, st>>>(nullptr);
    CK(hipEventRecord(e0, st));
    for (int i = 0; i < 960; ++i) k_empty<<<1, 256, 0, st>>>(nullptr);
    CK(hipEventRecord(e1, st));
    CK(hipEventSynchronize(e1));
    std::printf("direct stream launches of an empty 1-block kernel: %.2f us per launch\n",
                1e3 * time_ms(e0, e1) / 960.0);
  }

  // 2. grid barrier
  int* buf;
  CK(hipMalloc(&buf, (size_t)1024 * 16384));
  for (int blocks : {64, 128, 256, 512}) {
    for (int write : {0, 1}) {
      int iters = 1;
      void* args[] = {&buf, &iters, &write};
      CK(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_gridsync), dim3(blocks), dim3(256), args, 0,
                                    st));
      CK(hipEventRecord(e0, st));
      CK(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_gridsync), dim3(blocks), dim3(256), args, 0,
                                    st));
      CK(hipEventRecord(e1, st));
      CK(hipEventSynchronize(e1));
      const float one = time_ms(e0, e1);
      iters = 1001;
      CK(hipEventRecord(e0, st));
      CK(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_gridsync), dim3(blocks), dim3(256), args, 0,
                                    st));
      CK(hipEventRecord(e1, st));
      CK(hipEventSynchronize(e1));
      const float many = time_ms(e0, e1);
      std::printf("grid barrier, %3d blocks x 256 threads%s: %.2f us per barrier (launch with 1: %.2f us)\n", blocks,
                  write ? ", 16 KiB written per block" : "", 1e3 * (many - one) / 1000.0, 1e3 * one);
    }
  }

  // 2b. the counter barrier above, three variants
  {
    unsigned* w;
    int* err;
    CK(hipMalloc(&w, 64 * 4 * 20));
    CK(hipMalloc(&err, 256));
    for (int levels : {1, 2, 3, 4}) {
      for (int blocks : {64, 256, 512}) {
        for (int write : {0, 1}) {
          CK(hipMemsetAsync(w, 0, 64 * 4 * 20, st));
          CK(hipMemsetAsync(err, 0, 256, st));
          int iters = 1;
          void* args[] = {&buf, &w, &iters, &write, &err};
          const void* f = levels == 1   ? reinterpret_cast<const void*>(k_mybarrier<1>)
                          : levels == 2 ? reinterpret_cast<const void*>(k_mybarrier<2>)
                          : levels == 3 ? reinterpret_cast<const void*>(k_mybarrier<3>)
                                        : reinterpret_cast<const void*>(k_mybarrier<4>);
          CK(hipLaunchCooperativeKernel(f, dim3(blocks), dim3(256), args, 0, st));
          CK(hipMemsetAsync(w, 0, 64 * 4 * 20, st));
          CK(hipEventRecord(e0, st));
          CK(hipLaunchCooperativeKernel(f, dim3(blocks), dim3(256), args, 0, st));
          CK(hipEventRecord(e1, st));
          CK(hipEventSynchronize(e1));
          const float one = time_ms(e0, e1);
          iters = 1001;
          CK(hipMemsetAsync(w, 0, 64 * 4 * 20, st));
          CK(hipEventRecord(e0, st));
          CK(hipLaunchCooperativeKernel(f, dim3(blocks), dim3(256), args, 0, st));
          CK(hipEventRecord(e1, st));
          CK(hipEventSynchronize(e1));
          const float many = time_ms(e0, e1);
          int e = 0;
          CK(hipMemcpy(&e, err, 4, hipMemcpyDeviceToHost));
          std::printf("counter barrier (%s), %3d blocks%s: %.2f us per barrier%s\n",
                      levels == 1   ? "every thread fences, acquire polls"
                      : levels == 2 ? "every thread fences, relaxed polls"
                      : levels == 3 ? "thread 0 fences, relaxed polls"
                                    : "two levels, thread 0 fences, relaxed polls",
                      blocks, write ? ", 16 KiB written per block" : "", 1e3 * (many - one) / 1000.0,
                      e ? " (a poll gave up)" : "");
        }
      }
    }
    CK(hipFree(w));
    CK(hipFree(err));
  }

  // 3. flag handoff
  {
    int *flag, *err;
    CK(hipMalloc(&flag, 256));
    CK(hipMalloc(&err, 256));
    // per handoff from the difference of the two runs
    CK(hipMemsetAsync(flag, 0, 256, st));
    CK(hipMemsetAsync(err, 0, 256, st));
    int rounds = 100;
    void* args[] = {&flag, &rounds, &err};
    CK(hipEventRecord(e0, st));
    CK(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_pingpong), dim3(2), dim3(64), args, 0, st));
    CK(hipEventRecord(e1, st));
    CK(hipEventSynchronize(e1));
    const float a = time_ms(e0, e1);
    CK(hipMemsetAsync(flag, 0, 256, st));
    rounds = 10100;
    CK(hipEventRecord(e0, st));
    CK(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_pingpong), dim3(2), dim3(64), args, 0, st));
    CK(hipEventRecord(e1, st));
    CK(hipEventSynchronize(e1));
    const float b = time_ms(e0, e1);
    int e = 0;
    CK(hipMemcpy(&e, err, 4, hipMemcpyDeviceToHost));
    std::printf("flag handoff across XCDs: %.3f us one way%s\n", 1e3 * (b - a) / (2.0 * 10000), e ? " (gave up)" : "");
    CK(hipFree(flag));
    CK(hipFree(err));
  }

  // 4. closure and product latency inside one workgroup
  {
    std::vector<int32_t> h(128 * 128);
    for (size_t i = 0; i < h.size(); ++i) h[i] = 1 + (int32_t)((i * 2654435761u) % 1000u);
    int32_t* D;
    CK(hipMalloc(&D, h.size() * 4));
    auto run = [&](const char* name, auto launch) {
      CK(hipMemcpy(D, h.data(), h.size() * 4, hipMemcpyHostToDevice));
      CK(hipDeviceSynchronize());
      launch(1);
      CK(hipEventRecord(e0, st));
      launch(1);
      CK(hipEventRecord(e1, st));
      CK(hipEventSynchronize(e1));
      const float one = time_ms(e0, e1);
      CK(hipEventRecord(e0, st));
      launch(201);
      CK(hipEventRecord(e1, st));
      CK(hipEventSynchronize(e1));
      const float many = time_ms(e0, e1);
      std::printf("%s: %.2f us each in a loop (one launch with one: %.2f us)\n", name, 1e3 * (many - one) / 200.0,
                  1e3 * one);
    };
    run("closure 64x64, 256 threads", [&](int r) { k_closures<64, 256><<<1, 256, 0, st>>>(D, r); });
    run("closure 64x64, 512 threads", [&](int r) { k_closures<64, 512><<<1, 512, 0, st>>>(D, r); });
    run("closure 64x64, 1024 threads", [&](int r) { k_closures<64, 1024><<<1, 1024, 0, st>>>(D, r); });
    run("closure 32x32, 256 threads", [&](int r) { k_closures<32, 256><<<1, 256, 0, st>>>(D, r); });
    run("closure 32x32, 512 threads", [&](int r) { k_closures<32, 512><<<1, 512, 0, st>>>(D, r); });
    run("closure 128x128, 512 threads", [&](int r) { k_closures<128, 512><<<1, 512, 0, st>>>(D, r); });
    run("product 64x64x64, 256 threads", [&](int r) { k_products<16><<<1, 256, 0, st>>>(D, r); });
    CK(hipFree(D));
  }
  CK(hipFree(buf));
  CK(hipStreamSynchronize(st));
  std::printf("done\n");
  return 0;
}
