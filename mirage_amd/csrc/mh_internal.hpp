// Internal types shared by the HIP translation units of libmirage_hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "mirage_hip.h"

namespace mh {

constexpr uint32_t kHashBits = 28;                    // MirAge/Hashmap.h:111
constexpr uint32_t kHashMask = (1u << kHashBits) - 1u;
constexpr uint64_t kDirSize = 1ull << kHashBits;      // direct-indexed buckets
constexpr uint32_t kNoHash = 0xFFFFFFFFu;

// Device view of an index.  Buckets are direct-indexed by the 28-bit hash:
// entries of hash h are ent[dir[h], dir[h+1]) as (ref, pos) pairs.
struct IndexView {
  const uint64_t *dir;
  const uint2 *ent;
  const uint8_t *bytes;     // uppercased reference bases
  const uint64_t *ref_off;  // n_refs + 1
  uint32_t n_refs;
  uint32_t empty_hash;      // present bucket with no entries, or kNoHash
  int32_t T, s, k;
  // Copies of the entry (ref 0, pos 0) that bucket 0 holds beyond the one
  // stored in ent[]: the windows dropped after "N"-window cutoffs, which the
  // reference keeps as identical entries of hash 0.  Identical pairs never
  // chain, so each extra copy only adds the one-pair seed (0, j, s, j + s)
  // for every probe j that hit bucket 0; the kernels weight that seed instead
  // of storing the copies.
  uint32_t phantom;
  // Stored entries per non-empty bucket, times 16 (what a probe that finds
  // its bucket gathers on average); sizes the default tiles.
  uint32_t epp_x16;
};

// Tables of glibc logf for the relative score (MatrixResult.cpp:21-29):
// x_log[i] = logf(asfloat(0x3f800000 + i)), smax_log[L] = logf(scoreMax(L)).
struct Luts {
  const float *x_log;
  const float *smax_log;
  uint32_t n_x;
  uint32_t n_smax;
};

struct Reads {
  const uint8_t *bytes;
  const uint64_t *off;
  uint32_t n;
};

// Per read strand (query q = 2 * read + strand).  Up to kTopInline tied
// references are kept inline; larger tie sets go to the tie pool.
constexpr int kTopInline = 4;
struct QueryOut {
  float highest;
  uint32_t n_top;
  uint32_t top_ref[kTopInline];
  uint32_t tie_off;
  uint32_t n_cand;
  uint32_t status;
  uint32_t pad;
};

// Work statistics, accumulated per workgroup and added to one of
// kStatSlots slots (spread over cache lines) when the workgroup ends.
enum Stat { kProbes = 0, kHits, kEntries, kSeeds, kCands, kDirect, kExt, kStrands, kNumStats };
// Per-kernel statistics rows: length tiles 0..7, then the overflow tiers
// (second chance, middle, HBM workspaces, last resort).
constexpr int kKernelRows = 12;
constexpr int kTierRow0 = 8;
constexpr int kStatSlots = 256;
// Slots of each per-kernel row: a launch of one workgroup per query (the
// plain form) flushes once per query, and one address per statistic would
// make those atomics wait in turn.
constexpr int kKernelSlots = 32;

// Device counters, zeroed before each classify call.
struct DevCounters {
  unsigned long long top_used;      // final top-tie pool fill (set by the merge)
  unsigned long long cand_used;
  unsigned long long tie_used;      // scratch tie pool fill
  unsigned long long seed_check;    // first bad pair of mh_score_seeds (pair << 8 | code)
  unsigned long long tier_n[4];     // queries passed on to overflow tier t
  unsigned long long tier_head[4];  // queue heads of the overflow tiers
  unsigned long long queue_head[8]; // queue heads of the length buckets
  unsigned long long stats[kStatSlots][8];
  unsigned long long kstats[kKernelRows][kKernelSlots][8];  // the same, per kernel (kid)
};

struct Outputs {
  QueryOut *q;
  uint32_t *top;
  uint64_t top_cap;
  mh_cand *cands;
  uint64_t cand_cap;
  DevCounters *ctr;
  uint32_t *overflow;      // where queries too large for this tier are appended
  unsigned long long *overflow_cnt;
  uint32_t overflow_cap;
  uint32_t *ties;          // tie pool for tie sets larger than kTopInline
  uint64_t tie_cap;
};

struct Caps {
  uint32_t qcap;   // query bytes
  uint32_t pcap;   // probe hits
  uint32_t ecap;   // bucket entries (power of two)
  uint32_t nslot = 0;  // references scored together (workgroup teams); 0: one at a time
};


// Identity of an index for the lifetime of the process: a freed index's
// address can be reused by the next one, its generation never is.
inline uint64_t next_index_generation() {
  static std::atomic<uint64_t> g{0};
  return ++g;
}
}  // namespace mh

struct mh_index {
  const uint64_t generation = mh::next_index_generation();
  mh_handle *owner = nullptr;
  int device = 0;
  int32_t threshold = 0;
  uint32_t n_refs = 0;
  uint32_t empty_hash = mh::kNoHash;
  uint64_t n_entries = 0;   // entries of the reference's index (stored + phantom)
  uint64_t phantom = 0;     // copies of (0, 0) in bucket 0 not stored in d_ent
  uint64_t n_buckets = 0;
  uint64_t total_bytes = 0;
  uint64_t *d_dir = nullptr;
  uint2 *d_ent = nullptr;
  uint8_t *d_bytes = nullptr;
  uint64_t *d_ref_off = nullptr;
  std::vector<uint64_t> h_ref_off;
  mh::IndexView view() const {
    mh::IndexView v;
    v.dir = d_dir;
    v.ent = d_ent;
    v.bytes = d_bytes;
    v.ref_off = d_ref_off;
    v.n_refs = n_refs;
    v.empty_hash = empty_hash;
    v.T = threshold;
    v.s = threshold / 2;
    v.k = threshold - threshold / 2 + 1;
    v.phantom = (uint32_t)phantom;
    v.epp_x16 = n_buckets ? (uint32_t)std::min<uint64_t>(
                                0xFFFFFFFFull, (16 * (n_entries - phantom) + n_buckets / 2) / n_buckets)
                          : 0;
    return v;
  }
};

namespace mh {
// One execution context: a stream with its own scratch arena and counters.
// A handle has one; the serving loop keeps one per batch in flight.
struct Exec {
  hipStream_t stream = nullptr;
  void *d_scratch = nullptr;
  size_t scratch_bytes = 0;
  DevCounters *d_ctr = nullptr;
  DevCounters *h_ctr = nullptr;  // pinned host copy
  // timing events: [0] start, [1..8] after bucket kernels, [9] before the
  // bucket kernels, [10] after the large kernel, [11] after the merge
  hipEvent_t ev[17] = {};  // 0 start, 1-8 buckets, 9 buckets begin, 10 tiers end,
                           // 11 end, 12 probe pre-pass begin, 13-16 after tiers 0-3
  bool timed_probe = false;
  uint32_t tie_factor = 4;       // tie pool entries per query, grown on overflow
  // captured batched call (MH_CLASSIFY_GRAPH) and the arguments it was made for
  bool capturing = false;
  hipGraphExec_t graph = nullptr;
  uint64_t graph_key[12] = {};
  bool timed = false;
  uint32_t timed_buckets = 0;
  uint32_t bucket_order[8] = {};
};
}  // namespace mh

struct mh_handle {
  int device = 0;
  hipStream_t stream = nullptr;  // == ex.stream
  uint32_t max_read_len = 32768;
  uint32_t cu_count = 256;
  size_t lds_per_block = 65536;
  std::string last_error;
  std::mutex mu;
  // logf tables
  float *d_x_log = nullptr;
  float *d_smax_log = nullptr;
  uint32_t n_x = 0, n_smax = 0;
  mh::Exec ex;
  mh_counters last{};
  bool last_valid = false;
  std::vector<mh_tile> tiles;
  std::vector<mh_tile> tier_tiles;  // the two LDS overflow tiers (empty: defaults)
  uint64_t tiles_version = 0;
  // The default overflow tiers fitted to their kernel's residency, for the
  // first tile capacity and probe stride they were fitted for.
  mutable std::mutex tier_fit_mu;
  mutable std::vector<mh_tile> tier_fit;
  mutable uint32_t tier_fit_cap = 0, tier_fit_s = 0;
  mh::Luts luts() const { return mh::Luts{d_x_log, d_smax_log, n_x, n_smax}; }
};

namespace mh {

// Error helpers (host).
mh_status set_error(mh_handle *h, mh_status st, const std::string &msg);
mh_status hip_error(mh_handle *h, hipError_t e, const char *what);
#define MH_HIP(h, call)                                      \
  do {                                                       \
    hipError_t _e = (call);                                  \
    if (_e != hipSuccess) return ::mh::hip_error((h), _e, #call); \
  } while (0)

// Execution contexts (mh_api.hip).
mh_status exec_init(mh_handle *h, Exec &x);
void exec_free(Exec &x);
// Grows x's scratch arena to at least `bytes` (synchronises x's stream first).
mh_status ensure_scratch(mh_handle *h, Exec &x, size_t bytes);

// Index build (mh_index.hip).
mh_status build_index(mh_handle *h, const uint8_t *h_ref_bytes, const uint64_t *h_ref_off,
                      uint32_t n_refs, int32_t threshold, uint32_t flags, mh_index **out);

// Classification launchers (mh_classify.hip).
struct ClassifyArgs {
  IndexView ix;
  Luts luts;
  Reads rd;
  uint32_t flags;
  Outputs out;
  mh_read_result *final_out;
  // Longest read of the call when the host knows it (the server builds its
  // batches on the host): the batched form then skips the length tiles that
  // no read can reach.  UINT64_MAX: unknown, every tile is launched.
  uint64_t max_len_hint = UINT64_MAX;
};
// A scratch region of fixed size that a classify call uses instead of the
// execution context's arena (nothing is allocated): the HBM tiers run on as
// many workspaces as fit, and the tie pool is as large as fits.
struct FixedScratch {
  unsigned char *base = nullptr;
  size_t bytes = 0;
  uint32_t hbm_blocks = 0;  // workspaces of the first HBM tier
  uint32_t big_blocks = 0;  // workspaces of the last-resort tier
  uint32_t tie_factor = 4;  // tie pool entries per strand
  size_t tie_cap = 0;       // tie pool entries
};
// Fills fs (all but base) for a call over n_reads reads within `bytes`;
// MH_ERR_RANGE when the call cannot fit.
mh_status plan_fixed_scratch(const mh_handle *h, const IndexView &ix, uint32_t n_reads,
                             uint32_t flags, size_t bytes, FixedScratch &fs);
mh_status launch_classify(mh_handle *h, Exec &x, const ClassifyArgs &a, bool batched,
                          const FixedScratch *fixed = nullptr);
// Complement table in constant memory (once per device, before any capture).
hipError_t ensure_comp_table(int device);
// Scratch bytes one classify call needs.
size_t classify_scratch_bytes(const mh_handle *h, const IndexView &ix, uint32_t n_reads,
                              uint32_t tie_factor, uint32_t flags);
// Reads the counters of the last call on x (x's stream must be idle).
void counters_from(const DevCounters &d, uint64_t n_reads, mh_counters &out);
// The per-kernel rows of the last call on x (x's stream must be idle).
void kernel_counters_from(const DevCounters &d, mh_counters *out, uint32_t n);
mh_status launch_score_seeds(mh_handle *h, const IndexView &ix, const Luts &luts,
                             const uint8_t *d_q, const uint64_t *d_qoff, const uint32_t *d_pq,
                             const uint32_t *d_pr, const uint64_t *d_soff, const uint32_t *d_seeds,
                             uint32_t n_pairs, uint32_t flags, float *d_score, float *d_rel,
                             uint64_t *d_sum_r, uint64_t *d_sum_c, uint32_t *d_hits);

// Default tiles (mh_classify.hip).
// Default length tiles for an index at threshold T with epp_x16 (IndexView)
// entries per bucket; the LDS overflow tiers follow the first tile.  A table
// installed with mh_set_tiles / mh_set_overflow_tiles replaces them.
std::vector<mh_tile> default_tiles(uint32_t max_read_len, int T, uint32_t epp_x16);
std::vector<mh_tile> default_tier_tiles(uint32_t max_read_len, uint32_t first_cap);
std::vector<mh_tile> active_tiles(const mh_handle *h, const IndexView &ix);
std::vector<mh_tile> tier_tiles(const mh_handle *h, const IndexView &ix);

}  // namespace mh
