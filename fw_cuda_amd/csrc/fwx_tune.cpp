// fwx_tune.cpp -- configurations chosen per size class from tables that
// fw_cuda_amd/autotune.py writes from timings on the target GPU.
//
// Plain solves: (tile, kappa) when a caller passes 0 for either, and the
// register micro-tile of the 128-wide rest pass (its thread shape) for every
// call of that size class.  Batched plans: (tile, kappa, micro-tile) of each
// graph by its own size, tuned on the mixed 512-graph batch.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "fwx_internal.hpp"

namespace fwx {

namespace {
struct TuneEntry {
  int32_t n_max, tile, kappa, rest, p16kc, tile1, la;
};
// {n_max, tile, kappa, rest, p16kc}: rest = the shape of a pk64 128-wide
// rest pass (2: 512 threads of 8x4 cells; 0: 256 threads of 8x8 cells; 1:
// 256 threads of 8x8 cells on 16-deep chunks); p16kc = the chunk depth of a
// pk16 128-wide rest pass (16 or 32; 0 when the rest pass is narrower);
// tile1 = the tile of tiling_gpu (kappa 1) when the caller passes tile 0;
// la = the two-stream schedule (2 side-band, 3 side-band a stage ahead).
const TuneEntry kTable[] = {
#include "fwx_tune_table.inc"
};
// {n_max, tile, kappa, rest}: rest = threads of the 128-wide products of a
// batched graph (512: 8x4 cells each; 256: 8x8); the other fields unused.
const TuneEntry kBatchTable[] = {
#include "fwx_batch_table.inc"
};

template <size_t K>
const TuneEntry& lookup(const TuneEntry (&t)[K], int32_t n) {
  for (const TuneEntry& e : t)
    if (n <= e.n_max) return e;
  return t[K - 1];
}
}  // namespace

void tuned_config(int32_t n, int32_t* tile, int32_t* kappa) {
  const TuneEntry& e = lookup(kTable, n);
  *tile = e.tile;
  *kappa = e.kappa;
}

int tuned_rest_variant(int32_t n) { return lookup(kTable, n).rest; }

int tuned_p16_chunk(int32_t n) { return lookup(kTable, n).p16kc == 32 ? 32 : 16; }

int tuned_lookahead(int32_t n) {
  const int32_t v = lookup(kTable, n).la;
  return (v == 1 || v == 3) ? v : 2;
}

int tuned_tiling_tile(int32_t n) {
  const int32_t t = lookup(kTable, n).tile1;
  return (t == 32 || t == 64 || t == 128) ? t : lookup(kTable, n).tile;
}

// FWX_BATCH_TABLE="n_max:tile:kappa:threads,..." (ascending n_max) replaces
// the compiled batched table; the autotuner times candidate tables this way.
// Read at every plan creation; a malformed value is ignored.
void batch_tuned_config(int32_t n, int32_t* tile, int32_t* kappa, int32_t* rest_threads) {
  if (const char* env = std::getenv("FWX_BATCH_TABLE")) {
    std::vector<TuneEntry> t;
    const char* p = env;
    bool ok = true;
    while (*p) {
      TuneEntry e{};
      int used = 0;
      if (std::sscanf(p, "%d:%d:%d:%d%n", &e.n_max, &e.tile, &e.kappa, &e.rest, &used) != 4 ||
          (e.tile != 32 && e.tile != 64 && e.tile != 128) || e.kappa < 1 || e.kappa > FWX_MAX_KAPPA ||
          (e.rest != 256 && e.rest != 512)) {
        ok = false;
        break;
      }
      t.push_back(e);
      p += used;
      if (*p == ',') ++p;
    }
    if (ok && !t.empty()) {
      const TuneEntry* pick = &t.back();
      for (const TuneEntry& e : t)
        if (n <= e.n_max) {
          pick = &e;
          break;
        }
      *tile = pick->tile;
      *kappa = pick->kappa;
      *rest_threads = pick->rest;
      return;
    }
  }
  const TuneEntry& e = lookup(kBatchTable, n);
  *tile = e.tile;
  *kappa = e.kappa;
  *rest_threads = e.rest;
}

}  // namespace fwx
