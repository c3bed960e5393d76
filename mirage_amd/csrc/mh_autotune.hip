// Tile autotuner for the length-bucketed persistent form.
//
// For every length bucket of the handle's tile table, each candidate
// (team size, LDS entry capacity) is timed on the caller's reads (the whole
// call, so that work a small tile pushes to the overflow tiers is counted);
// the fastest candidate is installed.  Team sizes 16 and 32 put several
// queries in one wavefront and are tried for tiles of reads up to 1024
// bases.  Results do not depend on the tile: a query whose working set does
// not fit a tile's LDS goes to the overflow tiers, so only speed changes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "mh_internal.hpp"

using namespace mh;

extern "C" mh_status mh_autotune(mh_handle *h, const mh_index *idx, const uint8_t *d_reads,
                                 const uint64_t *d_read_off, uint32_t n_reads, uint32_t iters) {
  if (!h || !idx || !n_reads || !d_reads || !d_read_off) return MH_ERR_ARG;
  if (!iters) iters = 3;
  hipSetDevice(h->device);
  mh_read_result *d_out = nullptr;
  uint32_t *d_top = nullptr;
  const uint64_t top_cap = 16ull * n_reads + 1024;
  MH_HIP(h, hipMalloc(&d_out, n_reads * sizeof(mh_read_result)));
  if (hipMalloc(&d_top, top_cap * 4) != hipSuccess) {
    hipFree(d_out);
    return set_error(h, MH_ERR_NOMEM, "autotune buffers");
  }
  std::vector<mh_tile> tiles;
  {
    std::lock_guard<std::mutex> lk(h->mu);
    tiles = active_tiles(h, idx->view());
  }
  const uint32_t blocks[] = {16, 32, 64, 128, 256, 512};
  const uint32_t ecaps[] = {32, 64, 128, 256, 512, 1024, 2048, 4096};
  mh_status st = MH_OK;
  auto time_bucket = [&](const std::vector<mh_tile> &tt, uint32_t, float &ms) -> mh_status {
    std::vector<float> v;
    for (uint32_t it = 0; it <= iters; it++) {
      std::lock_guard<std::mutex> lk(h->mu);
      h->tiles = tt;
      h->tiles_version++;
      ClassifyArgs a;
      a.ix = idx->view();
      a.luts = h->luts();
      a.rd = Reads{d_reads, d_read_off, n_reads};
      a.flags = MH_CLASSIFY_BATCHED | MH_CLASSIFY_TIMING;
      std::memset(&a.out, 0, sizeof(a.out));
      a.out.top = d_top;
      a.out.top_cap = top_cap;
      a.final_out = d_out;
      mh_status s2 = launch_classify(h, h->ex, a, true);
      if (s2) return s2;
      MH_HIP(h, hipEventSynchronize(h->ex.ev[11]));
      // The whole call: a smaller tile can move work into the overflow tiers.
      float t = 0;
      hipEventElapsedTime(&t, h->ex.ev[0], h->ex.ev[11]);
      if (it) v.push_back(t);  // first call warms up
    }
    std::sort(v.begin(), v.end());
    ms = v[v.size() / 2];
    return MH_OK;
  };
  for (uint32_t b = 0; b < tiles.size() && !st; b++) {
    mh_tile best = tiles[b];
    float best_ms = 0;
    st = time_bucket(tiles, b, best_ms);
    for (uint32_t bl : blocks)
      for (uint32_t ec : ecaps) {
        if (st) break;
        if (bl < 64 && (tiles[b].max_len > 1024 || ec > 128)) continue;
        if (bl >= 64 && ec < 64) continue;
        mh_tile cand{tiles[b].max_len, bl, ec, 0};
        if (cand.block == best.block && cand.entries_cap == best.entries_cap) continue;
        std::vector<mh_tile> tt = tiles;
        tt[b] = cand;
        float ms = 0;
        st = time_bucket(tt, b, ms);
        if (!st && ms < best_ms) {
          best_ms = ms;
          best = cand;
        }
      }
    tiles[b] = best;
  }
  {
    std::lock_guard<std::mutex> lk(h->mu);
    h->tiles = tiles;
    h->tiles_version++;
  }
  hipFree(d_out);
  hipFree(d_top);
  return st;
}
