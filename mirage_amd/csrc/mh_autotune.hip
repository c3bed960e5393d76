// Tile autotuner for the length-bucketed persistent form.
//
// One call with the current tiles counts the strands of every length bucket
// (kernel rows); buckets without strands keep their tile.  For each bucket
// with strands, a local search over (team size, LDS entry capacity): the
// neighbours of the current tile (half and double the team, half and double
// the capacity, and both together) are timed by what the bucket costs: its
// own kernel plus the overflow tiers, which take the strands a smaller tile
// cannot hold (GPU events of one call, best of `iters` calls).  The search
// moves to the best neighbour while one is faster by more than 1 % and
// 0.05 ms (timing noise), at most kMoves times per bucket.  The two LDS
// overflow tiers are searched the same way last, by the tiers' time.  Calls
// classify at most the first kTuneReads reads (a 200 ms call on 10M reads
// becomes a 20 ms one).  Team sizes 16 and 32 put several queries in one
// wavefront and are tried for reads up to 1024 bases and capacities up to
// 128 entries.  Results do not depend on the tiles: a query whose working
// set does not fit a tile's LDS goes to the overflow tiers, so only speed
// changes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "mh_internal.hpp"

using namespace mh;

namespace {

constexpr int kMoves = 4;
constexpr uint32_t kTuneReads = 1u << 20;

bool allowed(const mh_tile &t, bool tier) {
  if (t.block < (tier ? 64u : 16u) || t.block > 1024 || t.entries_cap < 32 ||
      t.entries_cap > 4096)
    return false;
  if (t.block < 64) return t.max_len <= 1024 && t.entries_cap <= 128;
  return t.entries_cap >= 64;
}

std::vector<mh_tile> neighbours(const mh_tile &t, bool tier) {
  std::vector<mh_tile> v;
  const uint32_t bl[3] = {t.block / 2, t.block, t.block * 2};
  const uint32_t ec[3] = {t.entries_cap / 2, t.entries_cap, t.entries_cap * 2};
  for (uint32_t b : bl)
    for (uint32_t e : ec) {
      const mh_tile c{t.max_len, b, e, 0};
      if ((b != t.block || e != t.entries_cap) && allowed(c, tier)) v.push_back(c);
    }
  return v;
}

}  // namespace

extern "C" mh_status mh_autotune(mh_handle *h, const mh_index *idx, const uint8_t *d_reads,
                                 const uint64_t *d_read_off, uint32_t n_reads, uint32_t iters) {
  if (!h || !idx || !n_reads || !d_reads || !d_read_off) return MH_ERR_ARG;
  if (!iters) iters = 2;
  n_reads = std::min(n_reads, kTuneReads);
  hipSetDevice(h->device);
  mh_read_result *d_out = nullptr;
  uint32_t *d_top = nullptr;
  const uint64_t top_cap = 16ull * n_reads + 1024;
  MH_HIP(h, hipMalloc(&d_out, n_reads * sizeof(mh_read_result)));
  if (hipMalloc(&d_top, top_cap * 4) != hipSuccess) {
    hipFree(d_out);
    return set_error(h, MH_ERR_NOMEM, "autotune buffers");
  }
  std::vector<mh_tile> tiles, tiers;
  {
    std::lock_guard<std::mutex> lk(h->mu);
    tiles = active_tiles(h, idx->view());
    tiers = tier_tiles(h, idx->view());
  }
  // One call with tiles `tt` and tiers `tr`; its kernel timings.
  auto run = [&](const std::vector<mh_tile> &tt, const std::vector<mh_tile> &tr,
                 mh_timing &tm) -> mh_status {
    {
      std::lock_guard<std::mutex> lk(h->mu);
      h->tiles = tt;
      h->tier_tiles = tr;
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
      const mh_status s = launch_classify(h, h->ex, a, true);
      if (s) return s;
    }
    return mh_get_timing(h, &tm);
  };
  // Cost of bucket b (b < 0: the overflow tiers alone), best of iters calls.
  auto cost = [&](const std::vector<mh_tile> &tt, const std::vector<mh_tile> &tr, int b,
                  float &ms) -> mh_status {
    ms = 3.4e38f;
    for (uint32_t it = 0; it < iters; it++) {
      mh_timing tm;
      const mh_status s = run(tt, tr, tm);
      if (s) return s;
      ms = std::min(ms, (b >= 0 ? tm.bucket_ms[b] : 0.0f) + tm.big_ms);
    }
    return MH_OK;
  };
  mh_status st = MH_OK;
  // Warm-up call, then which buckets hold strands.
  std::vector<bool> busy(tiles.size(), false);
  {
    mh_timing tm;
    st = run(tiles, tiers, tm);
    mh_counters rows[kKernelRows];
    uint32_t nrows = 0;
    if (!st) st = mh_get_kernel_counters(h, rows, kKernelRows, &nrows);
    for (uint32_t b = 0; !st && b < tiles.size() && b < kTierRow0; b++)
      busy[b] = rows[b].queries + rows[b].overflow > 0;
  }
  for (uint32_t b = 0; b < tiles.size() && !st; b++) {
    if (!busy[b]) continue;
    float best_ms = 0;
    st = cost(tiles, tiers, (int)b, best_ms);
    for (int mv = 0; mv < kMoves && !st; mv++) {
      mh_tile pick = tiles[b];
      float pick_ms = best_ms;
      for (const mh_tile &cand : neighbours(tiles[b], false)) {
        std::vector<mh_tile> tt = tiles;
        tt[b] = cand;
        float ms = 0;
        st = cost(tt, tiers, (int)b, ms);
        if (st) break;
        if (ms < pick_ms) {
          pick_ms = ms;
          pick = cand;
        }
      }
      if (st || !(pick_ms < std::min(0.99f * best_ms, best_ms - 0.05f))) break;
      tiles[b] = pick;
      best_ms = pick_ms;
    }
  }
  for (uint32_t t = 0; t < 2 && !st; t++) {
    float best_ms = 0;
    st = cost(tiles, tiers, -1, best_ms);
    for (int mv = 0; mv < kMoves && !st; mv++) {
      mh_tile pick = tiers[t];
      float pick_ms = best_ms;
      for (const mh_tile &cand : neighbours(tiers[t], true)) {
        std::vector<mh_tile> tr = tiers;
        tr[t] = cand;
        float ms = 0;
        st = cost(tiles, tr, -1, ms);
        if (st) break;
        if (ms < pick_ms) {
          pick_ms = ms;
          pick = cand;
        }
      }
      if (st || !(pick_ms < std::min(0.99f * best_ms, best_ms - 0.05f))) break;
      tiers[t] = pick;
      best_ms = pick_ms;
    }
  }
  {
    std::lock_guard<std::mutex> lk(h->mu);
    h->tiles = tiles;
    h->tier_tiles = tiers;
    h->tiles_version++;
  }
  hipFree(d_out);
  hipFree(d_top);
  return st;
}
