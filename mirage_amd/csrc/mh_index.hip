// Device build of the bucketed k-mer index.
//
// Semantics follow Hashmap::TBuildHashmapForReference
// (MirAge/Hashmap.cpp:2771-2976): stride-1 windows of k = T - T/2 + 1 bases
// at positions [0, n - k) of each reference; a window whose hash equals
// H("N"*k) ends the reference's windows (the hashed pointer stops advancing,
// :2882-2895); buckets group entries by the 28-bit hash
// (ConvertHashesFromThreadToHashmap, :745-822); for T < 20 buckets above
// min(20, sizes[floor(0.3 * #buckets)]) are dropped (SaveHashmapTableV3,
// :1759-1833); the highest-hash bucket can be kept present but empty, as the
// .mhm loader leaves it (:540-554).  Windows dropped after a reference's
// cutoff still count in the size ConvertHashesFromThreadToHashmap converts
// (:1233-1239, :1252); the unfilled tail of its buffer is value-initialised
// (:1026), so bucket 0 holds one (ref 0, pos 0) entry per dropped window.
// Those identical entries are stored once; the count of the others is kept
// beside the index (mh_index::phantom) and weighted by the kernels.
//
// Layout in HBM: dir[2^28 + 1] u64 offsets (direct-indexed by hash) and
// ent[] (u32 ref, u32 pos) pairs, plus the uppercased reference bases.
// Build is a counting sort: hash pass -> per-bucket counts -> exclusive scan
// -> scatter with per-bucket cursors.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <vector>

#include "mh_internal.hpp"

namespace mh {
namespace {

__device__ __forceinline__ uint8_t upper(uint8_t c) {
  return (int8_t)c > 96 ? (uint8_t)(c - 32) : c;  // Sequences.cpp:43-52
}

__device__ __forceinline__ uint32_t hash28(const uint8_t *p, int k) {
  uint32_t h = 2166136261u;  // FNV-1, Hashmap.cpp:3197-3211
  for (int i = 0; i < k; i++) h = (h * 16777619u) ^ p[i];
  return ((h >> kHashBits) ^ h) & kHashMask;  // Hashmap.cpp:3233-3263
}

__device__ __forceinline__ uint32_t ref_of(const uint64_t *ref_off, uint32_t n_refs,
                                           uint64_t g) {
  uint32_t lo = 0, hi = n_refs;  // largest r with ref_off[r] <= g
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (ref_off[mid] <= g) lo = mid;
    else hi = mid;
  }
  return lo;
}

__global__ void k_upper(uint8_t *b, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    b[i] = upper(b[i]);
}

// First window of each reference whose hash equals H("N"*k).
__global__ void k_cutoff(const uint8_t *bytes, const uint64_t *ref_off, uint32_t n_refs,
                         uint64_t total, int k, uint32_t h_nnn, unsigned long long *cut) {
  for (uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < total;
       g += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t r = ref_of(ref_off, n_refs, g);
    uint64_t n = ref_off[r + 1] - ref_off[r];
    uint64_t i = g - ref_off[r];
    if (n <= (uint64_t)k || i >= n - k) continue;
    if (hash28(bytes + g, k) == h_nnn) atomicMin(&cut[r], (unsigned long long)i);
  }
}

__global__ void k_count(const uint8_t *bytes, const uint64_t *ref_off, uint32_t n_refs,
                        uint64_t total, int k, const unsigned long long *cut, uint32_t *cnt) {
  for (uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < total;
       g += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t r = ref_of(ref_off, n_refs, g);
    uint64_t i = g - ref_off[r];
    if (i >= cut[r]) continue;
    atomicAdd(&cnt[hash28(bytes + g, k)], 1u);
  }
}

// Histogram of bucket sizes clamped to 21 (enough for the fragment rule).
__global__ void k_size_hist(const uint32_t *cnt, uint64_t n, unsigned long long *hist) {
  __shared__ unsigned long long loc[22];
  if (threadIdx.x < 22) loc[threadIdx.x] = 0;
  __syncthreads();
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t c = cnt[i];
    if (c) atomicAdd(&loc[c > 21 ? 21 : c], 1ull);
  }
  __syncthreads();
  if (threadIdx.x < 22 && loc[threadIdx.x]) atomicAdd(&hist[threadIdx.x], loc[threadIdx.x]);
}

// Adds the dropped-window entries to bucket 0; flags a count past 32 bits.
__global__ void k_add_phantom(uint32_t *cnt, unsigned long long phantom,
                              unsigned long long *flag) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const unsigned long long v = (unsigned long long)cnt[0] + phantom;
    if (v > 0xFFFFFFFFull) *flag = 1;
    else cnt[0] = (uint32_t)v;
  }
}

__global__ void k_drop_large(uint32_t *cnt, uint64_t n, uint32_t thr) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    if (cnt[i] > thr) cnt[i] = 0;
}

__global__ void k_max_hash(const uint32_t *cnt, uint64_t n, unsigned int *mx,
                           unsigned long long *nonempty) {
  unsigned int best = 0;
  unsigned long long ne = 0;
  bool any = false;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    if (cnt[i]) {
      best = (unsigned int)i;
      any = true;
      ne++;
    }
  if (any) atomicMax(mx, best + 1u);
  if (ne) atomicAdd(nonempty, ne);
}

struct ToU64 {
  __host__ __device__ uint64_t operator()(uint32_t v) const { return v; }
};

__global__ void k_scatter(const uint8_t *bytes, const uint64_t *ref_off, uint32_t n_refs,
                          uint64_t total, int k, const unsigned long long *cut,
                          const uint64_t *dir, uint32_t *cursor, uint2 *ent) {
  for (uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < total;
       g += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t r = ref_of(ref_off, n_refs, g);
    uint64_t i = g - ref_off[r];
    if (i >= cut[r]) continue;
    uint32_t h = hash28(bytes + g, k);
    uint64_t b0 = dir[h], b1 = dir[h + 1];
    if (b1 == b0) continue;  // dropped bucket
    uint32_t slot = atomicAdd(&cursor[h], 1u);
    ent[b0 + slot] = make_uint2(r, (uint32_t)i);
  }
}

uint32_t host_hash28(const uint8_t *p, int k) {
  uint32_t h = 2166136261u;
  for (int i = 0; i < k; i++) h = (h * 16777619u) ^ p[i];
  return ((h >> kHashBits) ^ h) & kHashMask;
}

uint8_t host_upper(uint8_t c) { return (int8_t)c > 96 ? (uint8_t)(c - 32) : c; }

unsigned grid_for(uint64_t n, unsigned block, unsigned cap = 65536) {
  uint64_t g = (n + block - 1) / block;
  return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(g, cap));
}

}  // namespace

mh_status build_index(mh_handle *h, const uint8_t *h_ref_bytes, const uint64_t *h_ref_off,
                      uint32_t n_refs, int32_t threshold, uint32_t flags, mh_index **out) {
  if (!out || !h_ref_off || threshold < 2 || (n_refs && !h_ref_bytes))
    return set_error(h, MH_ERR_ARG, "mh_index_build: bad argument");
  const int s = threshold / 2, k = threshold - s + 1;
  const uint64_t base = h_ref_off[0];
  const uint64_t total = h_ref_off[n_refs] - base;
  hipStream_t st = h->stream;

  mh_index *ix = new mh_index();
  ix->owner = h;
  ix->device = h->device;
  ix->threshold = threshold;
  ix->n_refs = n_refs;
  ix->h_ref_off.resize(n_refs + 1);
  for (uint32_t r = 0; r <= n_refs; r++) ix->h_ref_off[r] = h_ref_off[r] - base;

  auto fail = [&](mh_status stt) {
    mh_index_free(ix);
    return stt;
  };
  auto check = [&](hipError_t e, const char *what) -> mh_status {
    if (e == hipSuccess) return MH_OK;
    return hip_error(h, e, what);
  };
  mh_status stt;
  if ((stt = check(hipMalloc(&ix->d_bytes, std::max<uint64_t>(total, 1) + 64), "hipMalloc bytes")))
    return fail(stt);
  if ((stt = check(hipMalloc(&ix->d_ref_off, (n_refs + 1) * sizeof(uint64_t)), "hipMalloc off")))
    return fail(stt);
  if ((stt = check(hipMalloc(&ix->d_dir, (kDirSize + 1) * sizeof(uint64_t)), "hipMalloc dir")))
    return fail(stt);
  if (total && (stt = check(hipMemcpyAsync(ix->d_bytes, h_ref_bytes + base, total,
                                           hipMemcpyHostToDevice, st), "copy bytes")))
    return fail(stt);
  if ((stt = check(hipMemcpyAsync(ix->d_ref_off, ix->h_ref_off.data(),
                                  (n_refs + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, st),
                   "copy off")))
    return fail(stt);
  if (total) k_upper<<<grid_for(total, 256), 256, 0, st>>>(ix->d_bytes, total);

  // Per-reference window limit: n - k, cut at the first "N"*k hash.
  std::vector<unsigned long long> stop(n_refs);
  for (uint32_t r = 0; r < n_refs; r++) {
    uint64_t n = ix->h_ref_off[r + 1] - ix->h_ref_off[r];
    stop[r] = n > (uint64_t)k ? n - k : 0;
  }
  unsigned long long *d_cut = nullptr;
  uint32_t *d_cnt = nullptr;
  unsigned long long *d_misc = nullptr;  // [0..21] hist, [22] max hash + 1, [23] nonempty,
                                        // [24] bucket-0 count overflow
  auto cleanup_tmp = [&]() {
    if (d_cut) hipFree(d_cut);
    if (d_cnt) hipFree(d_cnt);
    if (d_misc) hipFree(d_misc);
  };
  if ((stt = check(hipMalloc(&d_cut, std::max<uint32_t>(n_refs, 1) * sizeof(uint64_t)), "cut")) ||
      (stt = check(hipMalloc(&d_cnt, (kDirSize + 1) * sizeof(uint32_t)), "cnt")) ||
      (stt = check(hipMalloc(&d_misc, 32 * sizeof(uint64_t)), "misc"))) {
    cleanup_tmp();
    return fail(stt);
  }
  hipMemcpyAsync(d_cut, stop.data(), n_refs * sizeof(uint64_t), hipMemcpyHostToDevice, st);
  hipMemsetAsync(d_cnt, 0, (kDirSize + 1) * sizeof(uint32_t), st);
  hipMemsetAsync(d_misc, 0, 32 * sizeof(uint64_t), st);
  std::vector<uint8_t> nnn(k, 'N');
  const uint32_t h_nnn = host_hash28(nnn.data(), k);
  const unsigned G = grid_for(total, 256);
  uint64_t phantom = 0, cut0 = 0;
  if (total) {
    k_cutoff<<<G, 256, 0, st>>>(ix->d_bytes, ix->d_ref_off, n_refs, total, k, h_nnn, d_cut);
    std::vector<unsigned long long> cut(n_refs);
    hipMemcpyAsync(cut.data(), d_cut, n_refs * sizeof(uint64_t), hipMemcpyDeviceToHost, st);
    if ((stt = check(hipStreamSynchronize(st), "cutoff"))) {
      cleanup_tmp();
      return fail(stt);
    }
    for (uint32_t r = 0; r < n_refs; r++) phantom += stop[r] - cut[r];
    if (n_refs) cut0 = cut[0];
    k_count<<<G, 256, 0, st>>>(ix->d_bytes, ix->d_ref_off, n_refs, total, k, d_cut, d_cnt);
    if (phantom) k_add_phantom<<<1, 64, 0, st>>>(d_cnt, phantom, d_misc + 24);
  }
  if (threshold < 20) {
    k_size_hist<<<grid_for(kDirSize, 256, 4096), 256, 0, st>>>(d_cnt, kDirSize, d_misc);
    unsigned long long hist[22];
    hipMemcpyAsync(hist, d_misc, sizeof(hist), hipMemcpyDeviceToHost, st);
    if ((stt = check(hipStreamSynchronize(st), "fragment hist"))) {
      cleanup_tmp();
      return fail(stt);
    }
    unsigned long long nb = 0;
    for (int i = 1; i < 22; i++) nb += hist[i];
    if (nb) {
      // sizes sorted ascending; value at index floor((nb / 100.0) * 30)
      // (Hashmap.cpp:1801), threshold min(20, value) (:1808).
      size_t frag = (size_t)((nb / 100.0) * 30);
      unsigned long long acc = 0;
      uint32_t val = 21;
      for (int v = 1; v < 22; v++) {
        acc += hist[v];
        if (acc > frag) {
          val = (uint32_t)v;
          break;
        }
      }
      uint32_t thr = std::min<uint32_t>(20, val);
      k_drop_large<<<grid_for(kDirSize, 256, 4096), 256, 0, st>>>(d_cnt, kDirSize, thr);
    }
  }
  k_max_hash<<<grid_for(kDirSize, 256, 4096), 256, 0, st>>>(d_cnt, kDirSize,
                                                          (unsigned int *)(d_misc + 22),
                                                          d_misc + 23);
  unsigned long long mx[3];
  hipMemcpyAsync(mx, d_misc + 22, sizeof(mx), hipMemcpyDeviceToHost, st);
  if ((stt = check(hipStreamSynchronize(st), "max hash"))) {
    cleanup_tmp();
    return fail(stt);
  }
  if (mx[2]) {
    cleanup_tmp();
    return fail(set_error(h, MH_ERR_RANGE, "mh_index_build: bucket 0 would exceed 2^32 entries"));
  }
  const uint32_t mxh = (uint32_t)(mx[0] & 0xFFFFFFFFu);
  ix->n_buckets = mx[1];
  if ((flags & MH_INDEX_LAST_BUCKET_EMPTY) && mxh) {
    ix->empty_hash = mxh - 1;
    hipMemsetAsync(d_cnt + ix->empty_hash, 0, sizeof(uint32_t), st);
  }
  // Bucket 0 keeps one (0, 0) entry: a window of reference 0 at position 0
  // with hash 0, or else one of the dropped-window entries.
  if (phantom) {
    uint32_t c0 = 0;
    hipMemcpyAsync(&c0, d_cnt, sizeof(uint32_t), hipMemcpyDeviceToHost, st);
    if ((stt = check(hipStreamSynchronize(st), "bucket 0"))) {
      cleanup_tmp();
      return fail(stt);
    }
    if (c0 > 0) {
      uint64_t real00 = 0;
      if (n_refs && stop[0] > 0 && cut0 > 0) {
        std::vector<uint8_t> w(k);
        for (int i = 0; i < k; i++) w[i] = host_upper(h_ref_bytes[base + i]);
        real00 = host_hash28(w.data(), k) == 0 ? 1 : 0;
      }
      ix->phantom = phantom + real00 - 1;
      const uint32_t stored = c0 - (uint32_t)ix->phantom;
      hipMemcpyAsync(d_cnt, &stored, sizeof(uint32_t), hipMemcpyHostToDevice, st);
      if ((stt = check(hipStreamSynchronize(st), "bucket 0 count"))) {
        cleanup_tmp();
        return fail(stt);
      }
    }
  }
  // Exclusive scan of counts into u64 offsets.
  {
    hipcub::TransformInputIterator<uint64_t, ToU64, const uint32_t *> in(d_cnt, ToU64());
    size_t tmp_bytes = 0;
    hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, in, ix->d_dir, (int)(kDirSize + 1), st);
    void *d_tmp = nullptr;
    if ((stt = check(hipMalloc(&d_tmp, tmp_bytes), "scan tmp"))) {
      cleanup_tmp();
      return fail(stt);
    }
    hipcub::DeviceScan::ExclusiveSum(d_tmp, tmp_bytes, in, ix->d_dir, (int)(kDirSize + 1), st);
    hipStreamSynchronize(st);
    hipFree(d_tmp);
  }
  uint64_t n_ent = 0, bucket0 = 0;
  hipMemcpyAsync(&n_ent, ix->d_dir + kDirSize, sizeof(uint64_t), hipMemcpyDeviceToHost, st);
  hipMemcpyAsync(&bucket0, ix->d_dir + 1, sizeof(uint64_t), hipMemcpyDeviceToHost, st);
  if ((stt = check(hipStreamSynchronize(st), "scan"))) {
    cleanup_tmp();
    return fail(stt);
  }
  ix->n_entries = n_ent + ix->phantom;
  if ((stt = check(hipMalloc(&ix->d_ent, std::max<uint64_t>(n_ent, 1) * sizeof(uint2)), "ent"))) {
    cleanup_tmp();
    return fail(stt);
  }
  hipMemsetAsync(d_cnt, 0, (kDirSize + 1) * sizeof(uint32_t), st);
  // The scatter fills bucket 0's windows first; the (0, 0) entries after them
  // are zero bytes.
  if (phantom && bucket0) hipMemsetAsync(ix->d_ent, 0, bucket0 * sizeof(uint2), st);
  if (total)
    k_scatter<<<G, 256, 0, st>>>(ix->d_bytes, ix->d_ref_off, n_refs, total, k, d_cut, ix->d_dir,
                                 d_cnt, ix->d_ent);
  if ((stt = check(hipStreamSynchronize(st), "scatter"))) {
    cleanup_tmp();
    return fail(stt);
  }
  cleanup_tmp();
  if (ix->empty_hash != kNoHash) ix->n_buckets -= 1;
  ix->total_bytes = (std::max<uint64_t>(total, 1) + 64) + (n_refs + 1) * 8 +
                    (kDirSize + 1) * 8 + std::max<uint64_t>(n_ent, 1) * 8;
  *out = ix;
  return MH_OK;
}

}  // namespace mh
