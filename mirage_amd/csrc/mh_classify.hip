// Read classification on the device: k-mer probe, successive-occurrence
// filter, diagonal chains, seed scoring with head/tail extension, best hit
// and top-tie set, strand merge.
//
// One workgroup processes one query (one strand of one read) with its whole
// working set in LDS: the query bases, the probe hits, the gathered bucket
// entries and the row-hit bitmap.  Reference semantics, step by step:
//   probe        Analyser.cpp:1453-1520  every s-th window, j < m - T
//   successive   Analyser.cpp:2004-2197  longest run of consecutive hit
//                                        indices per reference; keep refs
//                                        whose run reaches max(1, M - 1)
//   sort/group   Analyser.cpp:2214-2315, :1024-1059  (ref, refPos, qPos)
//   chains       Analyser.cpp:2329-2490  links step s on both axes; only
//                                        refs whose longest chain equals the
//                                        query maximum G >= 2 get seeds
//   score        MatrixCPU.cpp:18-214    tail on the diagonal, head one
//                                        base off it, bitmap popcount
//   best hit     MatrixCPU.cpp:323-370, MatrixResult.cpp:21-29
//   strand merge Analyser.cpp:592-611
// The run and chain steps are done with LDS bitonic sorts and block scans
// instead of the reference's per-reference dense arrays; the head and tail
// comparisons of one seed are spread over the 64 lanes of a wavefront and
// resolved with a ballot.
//
// Queries whose working set does not fit the workgroup's LDS are appended
// to an overflow list and finished by a 1024-thread kernel whose workspace
// lives in HBM.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <vector>

#include "mh_internal.hpp"

namespace mh {
namespace {

__constant__ uint8_t kComp[256];

__device__ __forceinline__ uint8_t upper(uint8_t c) {
  return (int8_t)c > 96 ? (uint8_t)(c - 32) : c;  // Sequences.cpp:43-52
}

__device__ __forceinline__ uint32_t hash28(const uint8_t *p, int k) {
  uint32_t h = 2166136261u;  // Hashmap.cpp:3197-3211
  for (int i = 0; i < k; i++) h = (h * 16777619u) ^ p[i];
  return ((h >> kHashBits) ^ h) & kHashMask;
}

// Cross-lane reads through ds_bpermute with the source address computed once
// for all the values a step moves (each HIP shuffle recomputes it).  A lane
// whose source lane is disabled reads 0, so permutes are issued where every
// lane of the team is active, and results are selected afterwards.
__device__ __forceinline__ uint32_t lane64() { return __lane_id(); }
__device__ __forceinline__ uint32_t src_xor(uint32_t mask) { return (lane64() ^ mask) << 2; }
// source `d` lanes below / above within segments of W lanes (own lane where
// there is none)
template <uint32_t W>
__device__ __forceinline__ uint32_t src_up(uint32_t d) {
  const uint32_t l = lane64();
  return ((l & (W - 1)) >= d ? l - d : l) << 2;
}
template <uint32_t W>
__device__ __forceinline__ uint32_t src_down(uint32_t d) {
  const uint32_t l = lane64();
  return ((l & (W - 1)) + d < W ? l + d : l) << 2;
}
template <uint32_t W>
__device__ __forceinline__ uint32_t src_at(uint32_t u) {  // lane u of the segment
  return ((lane64() & ~(W - 1)) + u) << 2;
}
__device__ __forceinline__ uint32_t bperm(uint32_t src, uint32_t v) {
  return (uint32_t)__builtin_amdgcn_ds_bpermute((int)src, (int)v);
}
__device__ __forceinline__ uint64_t bperm64(uint32_t src, uint64_t v) {
  return ((uint64_t)bperm(src, (uint32_t)(v >> 32)) << 32) | bperm(src, (uint32_t)v);
}

// Segment operations on W consecutive lanes (W = 8, 16, 32 or 64, segments
// aligned to W) through DPP, which moves data between the lanes of a row of
// 16 inside the VALU, and v_readlane across rows: no LDS round trips.  Every
// lane of the segment takes part; lanes that have no source read `old`,
// which is the operation's identity.
// DPP controls: quad_perm 0x00-0xFF, row_shl:n 0x100+n, row_shr:n 0x110+n,
// row_ror:n 0x120+n, row_mirror 0x140, row_half_mirror 0x141, row_bcast:15
// 0x142 (lane 15 of each row to the next row), row_bcast:31 0x143 (lane 31 to
// rows 2 and 3).
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ uint32_t dpp(uint32_t old, uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, CTRL, ROWS, 0xF, false);
}
__device__ __forceinline__ uint32_t readlane(uint32_t v, uint32_t l) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)l);
}

// Inclusive sum over the lanes of the segment up to this one.
template <uint32_t W>
__device__ __forceinline__ uint32_t seg_incl_scan(uint32_t v) {
  static_assert(W == 8 || W == 16 || W == 32 || W == 64, "segment width");
  const uint32_t p = lane64() & (W - 1);
  uint32_t t = dpp<0x111>(0u, v);  // row_shr:1
  v += p >= 1 ? t : 0u;
  t = dpp<0x112>(0u, v);  // row_shr:2
  v += p >= 2 ? t : 0u;
  t = dpp<0x114>(0u, v);  // row_shr:4
  v += p >= 4 ? t : 0u;
  if constexpr (W >= 16) {
    t = dpp<0x118>(0u, v);  // row_shr:8
    v += p >= 8 ? t : 0u;
  }
  if constexpr (W >= 32) v += dpp<0x142, 0xA>(0u, v);  // rows 1, 3 += lane 15 of rows 0, 2
  if constexpr (W == 64) v += dpp<0x143, 0xC>(0u, v);  // rows 2, 3 += lane 31
  return v;
}

// The value of the segment's last lane, on every lane of the segment.
template <uint32_t W>
__device__ __forceinline__ uint32_t seg_last(uint32_t v) {
  if constexpr (W == 64) {
    return readlane(v, 63);
  } else {
    const uint32_t seg = lane64() / W;
    uint32_t r = 0;
#pragma unroll
    for (uint32_t g = 0; g < 64 / W; g++) {
      const uint32_t x = readlane(v, g * W + W - 1);
      r = seg == g ? x : r;
    }
    return r;
  }
}

// Reduction over the segment (op: 0 sum, 1 max, 2 or), on every lane of it.
template <int OP>
__device__ __forceinline__ uint32_t lane_op(uint32_t a, uint32_t b) {
  if constexpr (OP == 0) return a + b;
  else if constexpr (OP == 1) return a > b ? a : b;
  else return a | b;
}
template <uint32_t W, int OP>
__device__ __forceinline__ uint32_t seg_reduce(uint32_t v) {
  static_assert(W == 8 || W == 16 || W == 32 || W == 64, "segment width");
  if constexpr (W >= 16) v = lane_op<OP>(v, dpp<0x128>(0u, v));  // row_ror:8: lane ^ 8
  v = lane_op<OP>(v, dpp<0x141>(0u, v));                         // half-row mirror
  v = lane_op<OP>(v, dpp<0xB1>(0u, v));                          // lane ^ 1
  v = lane_op<OP>(v, dpp<0x4E>(0u, v));                          // lane ^ 2
  if constexpr (W == 32) {
    const uint32_t a = lane_op<OP>(readlane(v, 0), readlane(v, 16));
    const uint32_t b = lane_op<OP>(readlane(v, 32), readlane(v, 48));
    v = lane64() < 32 ? a : b;
  } else if constexpr (W == 64) {
    v = lane_op<OP>(lane_op<OP>(readlane(v, 0), readlane(v, 16)),
                    lane_op<OP>(readlane(v, 32), readlane(v, 48)));
  }
  return v;
}

// The value of lane ^ X (X = 1, 2, 4, 8 inside a row; larger X through
// ds_bpermute).
template <uint32_t X>
__device__ __forceinline__ uint32_t xor_lane(uint32_t v) {
  if constexpr (X == 1) return dpp<0xB1>(0u, v);
  else if constexpr (X == 2) return dpp<0x4E>(0u, v);
  else if constexpr (X == 4) {
    const uint32_t up = dpp<0x114>(0u, v), down = dpp<0x104>(0u, v);  // lane - 4, lane + 4
    return (lane64() & 4) ? up : down;
  } else if constexpr (X == 8) return dpp<0x128>(0u, v);  // row_ror:8
  else return bperm(src_xor(X), v);
}
template <uint32_t X>
__device__ __forceinline__ uint64_t xor_lane64(uint64_t v) {
  return ((uint64_t)xor_lane<X>((uint32_t)(v >> 32)) << 32) | xor_lane<X>((uint32_t)v);
}

// The value of lane - d (lane_up) or lane + d (lane_down) inside segments of
// W lanes, d a power of two below W; lanes with no such lane read garbage.
template <uint32_t W>
__device__ __forceinline__ uint32_t lane_up(uint32_t d, uint32_t v) {
  if constexpr (W <= 16) {
    switch (d) {
      case 1: return dpp<0x111>(0u, v);
      case 2: return dpp<0x112>(0u, v);
      case 4: return dpp<0x114>(0u, v);
      default: return dpp<0x118>(0u, v);
    }
  } else {
    return bperm(src_up<W>(d), v);
  }
}
template <uint32_t W>
__device__ __forceinline__ uint32_t lane_down(uint32_t d, uint32_t v) {
  if constexpr (W <= 16) {
    switch (d) {
      case 1: return dpp<0x101>(0u, v);
      case 2: return dpp<0x102>(0u, v);
      case 4: return dpp<0x104>(0u, v);
      default: return dpp<0x108>(0u, v);
    }
  } else {
    return bperm(src_down<W>(d), v);
  }
}
template <uint32_t W>
__device__ __forceinline__ uint64_t lane_up64(uint32_t d, uint64_t v) {
  return ((uint64_t)lane_up<W>(d, (uint32_t)(v >> 32)) << 32) | lane_up<W>(d, (uint32_t)v);
}
template <uint32_t W>
__device__ __forceinline__ uint64_t lane_down64(uint32_t d, uint64_t v) {
  return ((uint64_t)lane_down<W>(d, (uint32_t)(v >> 32)) << 32) | lane_down<W>(d, (uint32_t)v);
}
__device__ __forceinline__ uint64_t xor_lane64_any(uint32_t x, uint64_t v) {
  switch (x) {
    case 1: return xor_lane64<1>(v);
    case 2: return xor_lane64<2>(v);
    case 4: return xor_lane64<4>(v);
    case 8: return xor_lane64<8>(v);
    default: return bperm64(src_xor(x), v);
  }
}
// The value of lane ^ X with every lane of the wavefront active: DPP inside a
// row for X = 1, 2, 3 (quad_perm [3,2,1,0]), 4, 7 (row_half_mirror), 8 and
// 15 (row_mirror); X = 0 stands for any other mask, through ds_bpermute from
// `src` (the byte address of the source lane).
template <uint32_t X>
__device__ __forceinline__ uint32_t xor_mask(uint32_t src, uint32_t v) {
  if constexpr (X == 3) return dpp<0x1B>(0u, v);
  else if constexpr (X == 7) return dpp<0x141>(0u, v);
  else if constexpr (X == 15) return dpp<0x140>(0u, v);
  else if constexpr (X == 0) return bperm(src, v);
  else return xor_lane<X>(v);
}
template <uint32_t X>
__device__ __forceinline__ uint64_t xor_mask64(uint32_t src, uint64_t v) {
  return ((uint64_t)xor_mask<X>(src, (uint32_t)(v >> 32)) << 32) | xor_mask<X>(src, (uint32_t)v);
}

// The value of lane u of the segment (u the same on every lane of the
// wavefront): v_readlane for whole wavefronts, else ds_bpermute.
template <uint32_t W>
__device__ __forceinline__ uint32_t lane_at(uint32_t u, uint32_t v) {
  if constexpr (W == 64) return readlane(v, u);
  else return bperm(src_at<W>(u), v);
}
template <uint32_t W>
__device__ __forceinline__ uint64_t lane_at64(uint32_t u, uint64_t v) {
  return ((uint64_t)lane_at<W>(u, (uint32_t)(v >> 32)) << 32) | lane_at<W>(u, (uint32_t)v);
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) { return seg_incl_scan<64>(v); }
__device__ __forceinline__ uint32_t wave_max(uint32_t v) { return seg_reduce<64, 1>(v); }
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) { return seg_reduce<64, 0>(v); }

// Exclusive block scan; red holds at least BS/64 words.  All threads get the
// block total.  Contains barriers when BS > 64.
template <int BS>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *red, uint32_t &total) {
  constexpr int NW = BS / 64;
  const uint32_t inc = wave_incl_scan(v);
  if constexpr (NW == 1) {
    total = readlane(inc, 63);
    return inc - v;
  } else {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 63) red[w] = inc;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NW; i++) {
      uint32_t x = red[i];
      off += (i < w) ? x : 0u;
      tot += x;
    }
    total = tot;
    return off + inc - v;
  }
}

template <int BS>
__device__ __forceinline__ uint32_t block_max(uint32_t v, uint32_t *red) {
  constexpr int NW = BS / 64;
  v = wave_max(v);
  if constexpr (NW == 1) {
    return v;
  } else {
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    uint32_t r = 0;
#pragma unroll
    for (int i = 0; i < NW; i++) r = max(r, red[i]);
    return r;
  }
}

template <int BS>
__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t *red) {
  constexpr int NW = BS / 64;
  v = wave_sum(v);
  if constexpr (NW == 1) {
    return v;
  } else {
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    uint32_t r = 0;
#pragma unroll
    for (int i = 0; i < NW; i++) r += red[i];
    return r;
  }
}

// A team processes one query.  BlockTeam<BS>: the whole workgroup.
// LaneTeam<L>: L lanes of a one-wavefront workgroup, 64 / L queries side by
// side.  Its members exchange data through
// cross-lane operations and LDS without workgroup barriers: the LDS
// operations of a wavefront execute in order, and the fences keep the
// compiler from moving memory accesses across a team synchronisation.
template <int BS>
struct BlockTeam {
  static constexpr uint32_t kSize = BS;
  static constexpr uint32_t kWorkerLanes = 64;
  static constexpr uint32_t kPerWave = 1;        // teams per workgroup
  static constexpr uint32_t kSeedLanes = 8;      // lanes that extend one seed
  __device__ static uint32_t rank() { return threadIdx.x; }
  __device__ static uint32_t index() { return 0; }
  __device__ static uint32_t worker() { return threadIdx.x >> 6; }
  __device__ static uint32_t wlane() { return threadIdx.x & 63; }
  __device__ static uint64_t wballot(bool p) { return __ballot(p); }
  __device__ static void sync() { __syncthreads(); }
  __device__ static bool any(int x) { return __syncthreads_or(x) != 0; }
  __device__ static uint32_t excl_scan(uint32_t v, uint32_t *red, uint32_t &tot) {
    return block_excl_scan<BS>(v, red, tot);
  }
  __device__ static uint32_t max(uint32_t v, uint32_t *red) { return block_max<BS>(v, red); }
  __device__ static uint32_t sum(uint32_t v, uint32_t *red) { return block_sum<BS>(v, red); }
};

template <int L>
struct LaneTeam {
  static_assert(L == 8 || L == 16 || L == 32, "lane team size");
  static constexpr uint32_t kSize = L;
  static constexpr uint32_t kWorkerLanes = L;
  static constexpr uint32_t kPerWave = 64 / L;
  static constexpr uint64_t kMask = (1ull << L) - 1;
  static constexpr uint32_t kSeedLanes = L < 16 ? L : 16;
  __device__ static uint32_t rank() { return threadIdx.x & (L - 1); }
  __device__ static uint32_t index() { return threadIdx.x / L; }
  __device__ static uint32_t worker() { return 0; }
  __device__ static uint32_t wlane() { return threadIdx.x & (L - 1); }
  __device__ static uint64_t wballot(bool p) {
    return (__ballot(p) >> (threadIdx.x & (64 - L))) & kMask;
  }
  __device__ static void sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  __device__ static bool any(int x) {
    sync();
    return wballot(x != 0) != 0;
  }
  __device__ static uint32_t excl_scan(uint32_t v, uint32_t *, uint32_t &tot) {
    const uint32_t inc = seg_incl_scan<L>(v);
    tot = seg_last<L>(inc);
    return inc - v;
  }
  __device__ static uint32_t max(uint32_t v, uint32_t *) { return seg_reduce<L, 1>(v); }
  __device__ static uint32_t sum(uint32_t v, uint32_t *) { return seg_reduce<L, 0>(v); }
};

// Seed workers: groups of TM::kSeedLanes lanes of a team, each extending its
// own seeds (four per wavefront), so the memory round trips of several seeds
// overlap.
template <class TM>
__device__ __forceinline__ uint32_t seed_lane() {
  return threadIdx.x & (TM::kSeedLanes - 1);
}
template <class TM>
__device__ __forceinline__ uint64_t seed_ballot(bool p) {
  constexpr uint32_t SL = TM::kSeedLanes;
  return (__ballot(p) >> (threadIdx.x & (64 - SL))) & ((1ull << SL) - 1);
}

__device__ __forceinline__ uint32_t next_pow2(uint32_t x) {
  return x <= 1 ? 1u : 1u << (32 - __clz(x - 1));
}

__device__ __forceinline__ bool pair_less(uint64_t ka, uint32_t va, uint64_t kb, uint32_t vb) {
  return ka < kb || (ka == kb && va < vb);
}

// Elements of the register networks: a u64 key, or a (u64 key, u32 value)
// pair ordered by key, then value.
struct KeyVal {
  uint64_t k;
  uint32_t v;
};
__device__ __forceinline__ bool elem_less(uint64_t a, uint64_t b) { return a < b; }
__device__ __forceinline__ bool elem_less(const KeyVal &a, const KeyVal &b) {
  return pair_less(a.k, a.v, b.k, b.v);
}
template <uint32_t X>
__device__ __forceinline__ uint64_t elem_xor(uint32_t src, uint64_t a) {
  return xor_mask64<X>(src, a);
}
template <uint32_t X>
__device__ __forceinline__ KeyVal elem_xor(uint32_t src, const KeyVal &a) {
  return KeyVal{xor_mask64<X>(src, a.k), xor_mask<X>(src, a.v)};
}
// c ? a : b, field by field (a select of whole structs would go through
// their addresses and keep the network's arrays out of registers)
__device__ __forceinline__ uint64_t elem_sel(bool c, uint64_t a, uint64_t b) { return c ? a : b; }
__device__ __forceinline__ KeyVal elem_sel(bool c, const KeyVal &a, const KeyVal &b) {
  return KeyVal{c ? a.k : b.k, c ? a.v : b.v};
}
template <class El>
__device__ __forceinline__ void cmp_swap(El &a, El &b) {
  const bool sw = elem_less(b, a);
  const El lo = elem_sel(sw, b, a), hi = elem_sel(sw, a, b);
  a = lo;
  b = hi;
}

// The arrays a network loads from and stores to: keys alone, or keys and
// values side by side.
template <class K64>
struct KeyArr {
  using El = uint64_t;
  K64 *key;
  __device__ __forceinline__ El inf() const { return ~0ull; }
  __device__ __forceinline__ El get(uint32_t i) const { return key[i]; }
  __device__ __forceinline__ void put(uint32_t i, const El &x) const { key[i] = x; }
};
template <class K64, class V32>
struct PairArr {
  using El = KeyVal;
  K64 *key;
  V32 *val;
  __device__ __forceinline__ El inf() const { return KeyVal{~0ull, ~0u}; }
  __device__ __forceinline__ El get(uint32_t i) const { return KeyVal{key[i], val[i]}; }
  __device__ __forceinline__ void put(uint32_t i, const El &x) const {
    key[i] = x.k;
    val[i] = x.v;
  }
};

// The network of bitonic_sort over n elements held in registers by a whole
// workgroup, K a thread: thread t holds positions t*K .. t*K+K-1 (n2 = 2^LOGN
// = next_pow2(n) <= workgroup size * K; positions n and beyond hold
// +infinity).  Step (LS, LT) of the merge of size 2^LS pairs position e with
// e ^ M, M = 2^LS - 1 for its first step (LT = LS - 1) and 2^LT for the
// others; the lower position takes the smaller element.  The steps are
// unrolled at compile time, so each one's partner rule and lane masks are
// constants: partners inside a thread swap in registers; partners inside a
// wavefront come through DPP or ds_bpermute, every element's partner fetched
// before any is compared; partners in another wavefront (M of 64*K and more:
// 3 of the 55 steps for 1,024 keys on 256 threads) through the arrays in LDS
// between two barriers.
template <int K, int LS, int LT, class A>
__device__ __forceinline__ void net_step(typename A::El (&x)[K], const A &arr, uint32_t e0,
                                         uint32_t lane, uint32_t n) {
  using El = typename A::El;
  constexpr uint32_t LK = K == 1 ? 0 : (K == 2 ? 1 : 2);
  constexpr uint32_t WK = 64 * K;  // positions per wavefront
  constexpr bool mirror = LT == LS - 1;
  constexpr uint32_t stride = 1u << LT;
  constexpr uint32_t m = mirror ? (2u << LT) - 1 : stride;
  if constexpr (m < (uint32_t)K) {
    if constexpr (m == 1) {
      cmp_swap(x[0], x[1]);
      if constexpr (K == 4) cmp_swap(x[2], x[3]);
    } else if constexpr (m == 2) {
      cmp_swap(x[0], x[2]);
      cmp_swap(x[1], x[3]);
    } else {
      cmp_swap(x[0], x[3]);
      cmp_swap(x[1], x[2]);
    }
  } else if constexpr (m < WK) {
    // partner lane ^ (m >> LK); its register j ^ (K - 1) in a mirror step
    // (the low bits of m are all set), else register j
    constexpr uint32_t lm = m >> LK;
    constexpr uint32_t X =
        (lm == 1 || lm == 2 || lm == 3 || lm == 4 || lm == 7 || lm == 8 || lm == 15) ? lm : 0;
    const uint32_t src = X == 0 ? src_xor(lm) : 0u;
    const bool lower = (lane & (stride >> LK)) == 0;
    El y[K];
#pragma unroll
    for (int j = 0; j < K; j++) y[j] = elem_xor<X>(src, x[mirror ? K - 1 - j : j]);
#pragma unroll
    for (int j = 0; j < K; j++) {
      const bool lt = elem_less(y[j], x[j]);
      x[j] = elem_sel(lower == lt, y[j], x[j]);
    }
  } else {
    // Positions n and beyond keep +infinity: they are neither stored nor
    // read, and a position below n whose partner is one of them keeps its
    // element (the partner is the higher position).
#pragma unroll
    for (int j = 0; j < K; j++)
      if (e0 + j < n) arr.put(e0 + j, x[j]);
    __syncthreads();
    const bool lower = (e0 & stride) == 0;
    El y[K];
#pragma unroll
    for (int j = 0; j < K; j++) {
      const uint32_t p = (e0 + j) ^ m;
      if (p < n) y[j] = arr.get(p);
      else y[j] = arr.inf();
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < K; j++) {
      const bool lt = elem_less(y[j], x[j]);
      x[j] = elem_sel(lower == lt, y[j], x[j]);
    }
  }
}

// Steps LT .. 0 of the merge of size 2^LS, then the merges up to 2^LOGN.
template <int K, int LS, int LT, class A>
__device__ __forceinline__ void net_merge(typename A::El (&x)[K], const A &arr, uint32_t e0,
                                          uint32_t lane, uint32_t n) {
  net_step<K, LS, LT>(x, arr, e0, lane, n);
  if constexpr (LT > 0) net_merge<K, LS, LT - 1>(x, arr, e0, lane, n);
}
template <int K, int LS, int LOGN, class A>
__device__ __forceinline__ void net_sizes(typename A::El (&x)[K], const A &arr, uint32_t e0,
                                          uint32_t lane, uint32_t n) {
  net_merge<K, LS, LS - 1>(x, arr, e0, lane, n);
  if constexpr (LS < LOGN) net_sizes<K, LS + 1, LOGN>(x, arr, e0, lane, n);
}

template <int K, int LOGN, class A>
__device__ __forceinline__ void reg_sort(const A &arr, uint32_t n) {
  static_assert(K == 1 || K == 2 || K == 4, "elements per thread");
  const uint32_t e0 = threadIdx.x * K;
  const uint32_t lane = lane64();
  typename A::El x[K];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < K; j++) {
    if (e0 + j < n) x[j] = arr.get(e0 + j);
    else x[j] = arr.inf();
  }
  net_sizes<K, 1, LOGN>(x, arr, e0, lane, n);
#pragma unroll
  for (int j = 0; j < K; j++)
    if (e0 + j < n) arr.put(e0 + j, x[j]);
  __syncthreads();
}

// reg_sort for n2 = 2^lg, lg from LOGN up to LB + 2 (LB = log2 of the
// workgroup size): K = 1 up to one element a thread, then 2, then 4.  False
// when lg is beyond.
template <int LB, int LOGN, class A>
__device__ __forceinline__ bool reg_sort_pow2(const A &arr, uint32_t n, uint32_t lg) {
  if constexpr (LOGN > LB + 2) {
    return false;
  } else {
    if (lg == (uint32_t)LOGN) {
      reg_sort<(LOGN <= LB ? 1 : (LOGN == LB + 1 ? 2 : 4)), LOGN>(arr, n);
      return true;
    }
    return reg_sort_pow2<LB, LOGN + 1>(arr, n, lg);
  }
}

// log2 of a whole-workgroup team's size
template <class TM>
constexpr int team_log2() {
  return TM::kSize == 64 ? 6 : TM::kSize == 128 ? 7 : TM::kSize == 256 ? 8 : TM::kSize == 512 ? 9 : 10;
}

// Ascending sort of n (key, val) pairs, any n.  Whole-workgroup teams sort in
// registers (reg_sort above).  Otherwise up to one wavefront's worth
// of elements (64, or the team's lanes) are ranked through shuffles: each
// lane counts the smaller elements and writes its element to its rank.
// Larger arrays use a bitonic network in LDS whose comparators all put the
// smaller element first: each merge of size k compares i with its mirror
// i ^ (k - 1), then with i ^ d for d = k/4 .. 1.  Positions n and beyond act
// as +infinity and never move, so the comparisons that would touch them are
// skipped and nothing is padded.
// Synchronisation before a network step.  Thread t of a team handles pairs
// p = t, t + kSize, ...; a wavefront's pairs P..P+63 (P a multiple of 64)
// touch elements 2P..2P+127 in a step of stride up to 64, and in a mirror
// step of size up to 128, so between two such steps each wavefront only
// waits for itself (its LDS accesses execute in order); any other step, and
// the step after it, needs the whole team.
__device__ __forceinline__ bool wave_local_step(uint32_t size, uint32_t stride) {
  return stride == (size >> 1) ? size <= 128 : stride <= 64;
}
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
template <class TM>
__device__ __forceinline__ void step_sync(bool local, bool &prev_local) {
  if (local && prev_local) wave_sync();
  else TM::sync();
  prev_local = local;
}

template <class TM, class K64, class V32>
__device__ __noinline__ void bitonic_sort_in(K64 *key, V32 *val, uint32_t n) {
  if constexpr (TM::kPerWave == 1) {
    // whole-workgroup teams: the register network from two pairs up (one
    // wavefront: from 65, as for keys)
    static_assert((1 << team_log2<TM>()) == (int)TM::kSize, "workgroup size");
    constexpr int LO = TM::kSize == 64 ? 7 : 1;
    if (n >= 2 && reg_sort_pow2<team_log2<TM>(), LO>(PairArr<K64, V32>{key, val}, n,
                                                     31 - __clz(next_pow2(n))))
      return;
  }
  constexpr uint32_t W = TM::kWorkerLanes;
  if (n <= W) {
    TM::sync();
    if (TM::worker() == 0) {
      const uint32_t lane = TM::wlane();
      uint64_t k = lane < n ? key[lane] : ~0ull;
      uint32_t v = lane < n ? val[lane] : ~0u;
      uint32_t rank = 0;
      for (uint32_t u = 0; u < n; u++) {
        const uint64_t ku = lane_at64<W>(u, k);
        const uint32_t vu = lane_at<W>(u, v);
        rank += (pair_less(ku, vu, k, v) || (ku == k && vu == v && u < lane)) ? 1u : 0u;
      }
      __builtin_amdgcn_wave_barrier();
      if (lane < n) {
        key[rank] = k;
        val[rank] = v;
      }
    }
    TM::sync();
    return;
  }
  const uint32_t n2 = next_pow2(n);
  bool prev_local = false;
  for (uint32_t size = 2; size <= n2; size <<= 1) {
    for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      step_sync<TM>(wave_local_step(size, stride), prev_local);
      const uint32_t lg = 31 - __clz(stride);
      const bool mirror = stride == (size >> 1);
      for (uint32_t t = TM::rank(); t < (n2 >> 1); t += TM::kSize) {
        const uint32_t i = ((t >> lg) << (lg + 1)) + (t & (stride - 1));
        const uint32_t j = mirror ? (i ^ (size - 1)) : (i + stride);
        if (j >= n) continue;  // j > i: +infinity at j, nothing moves
        const uint64_t ka = key[i], kb = key[j];
        const uint32_t va = val[i], vb = val[j];
        if (pair_less(kb, vb, ka, va)) {
          key[i] = kb;
          key[j] = ka;
          val[i] = vb;
          val[j] = va;
        }
      }
    }
  }
  TM::sync();
}


// Sorts n distinct u64 keys in place: whole-workgroup teams in registers
// (reg_sort, up to 4 keys a thread; a one-wavefront team from 65 keys);
// otherwise the rank of each key among the first n by shuffles for one
// worker's worth, and the network of bitonic_sort without a value array
// beyond that.
template <class TM, class K64>
__device__ __noinline__ void bitonic_sort_keys_in(K64 *key, uint32_t n) {
  if constexpr (TM::kPerWave == 1) {
    // Whole-workgroup teams: the register network from two keys up.  A
    // one-wavefront team ranks up to 64 keys by shuffles (measured faster:
    // it has no other wavefront to leave idle).
    static_assert((1 << team_log2<TM>()) == (int)TM::kSize, "workgroup size");
    constexpr int LO = TM::kSize == 64 ? 7 : 1;
    if (n >= 2 && reg_sort_pow2<team_log2<TM>(), LO>(KeyArr<K64>{key}, n, 31 - __clz(next_pow2(n))))
      return;
  }
  constexpr uint32_t W = TM::kWorkerLanes;
  if (n <= W) {
    TM::sync();
    if (TM::worker() == 0) {
      const uint32_t lane = TM::wlane();
      const uint64_t k = lane < n ? key[lane] : ~0ull;
      uint32_t rank = 0;
      for (uint32_t u = 0; u < n; u++) rank += lane_at64<W>(u, k) < k ? 1u : 0u;
      __builtin_amdgcn_wave_barrier();
      if (lane < n) key[rank] = k;
    }
    TM::sync();
    return;
  }
  const uint32_t n2 = next_pow2(n);
  bool prev_local = false;
  for (uint32_t size = 2; size <= n2; size <<= 1) {
    for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      step_sync<TM>(wave_local_step(size, stride), prev_local);
      const uint32_t lg = 31 - __clz(stride);
      const bool mirror = stride == (size >> 1);
      for (uint32_t t = TM::rank(); t < (n2 >> 1); t += TM::kSize) {
        const uint32_t i = ((t >> lg) << (lg + 1)) + (t & (stride - 1));
        const uint32_t j = mirror ? (i ^ (size - 1)) : (i + stride);
        if (j >= n) continue;  // j > i: +infinity at j, nothing moves
        const uint64_t ka = key[i], kb = key[j];
        if (kb < ka) {
          key[i] = kb;
          key[j] = ka;
        }
      }
    }
  }
  TM::sync();
}

// The sorts are called, not inlined (inlined, their registers spill in every
// tile kernel), so a pointer reaches them without its address space and
// would compile to flat accesses.  Whole-workgroup teams on LDS workspaces
// get an instantiation on LDS pointers (ds_read / ds_write); HBM workspaces
// and lane teams the generic one (lane teams sort a few dozen keys, and
// their tile kernel measured faster that way: 2.44 against 2.52 ms on 2M
// strands of 150 bp).
using lds_u64 = __attribute__((address_space(3))) uint64_t;
using lds_u32 = __attribute__((address_space(3))) uint32_t;

__device__ __forceinline__ bool in_lds(const void *p) {
  return __builtin_amdgcn_is_shared((const __attribute__((address_space(0))) void *)p);
}

template <class TM>
__device__ __forceinline__ void bitonic_sort_keys(uint64_t *key, uint32_t n) {
  if constexpr (TM::kPerWave == 1) {
    if (in_lds(key)) {
      bitonic_sort_keys_in<TM>((lds_u64 *)key, n);
      return;
    }
  }
  bitonic_sort_keys_in<TM>(key, n);
}

template <class TM>
__device__ __forceinline__ void bitonic_sort(uint64_t *key, uint32_t *val, uint32_t n) {
  if constexpr (TM::kPerWave == 1) {
    if (in_lds(key)) {
      bitonic_sort_in<TM>((lds_u64 *)key, (lds_u32 *)val, n);
      return;
    }
  }
  bitonic_sort_in<TM>(key, val, n);
}

// Ascending sort of n u64 keys in global memory through an LDS buffer of
// lcap keys (a power of two), for workspaces that live in HBM.  Same network
// as bitonic_sort; every run of its steps whose stride is below the chunk
// size m = min(next_pow2(n), lcap) is done one m-key chunk at a time in LDS
// (a chunk's tail past n padded with ~0), the larger strides in global
// memory, where positions n and beyond act as +infinity.  For n2 = 2^17 and
// lcap = 2^12 that is 15 global steps and 6 LDS passes in place of 153
// global steps.  Keys need not be distinct.
template <class TM>
__device__ void staged_sort_keys(uint64_t *key, uint32_t n, uint64_t *stage, uint32_t lcap) {
  lds_u64 *lds = (lds_u64 *)stage;  // the buffer is in LDS: ds_read / ds_write
  if (n <= 1) {
    TM::sync();
    return;
  }
  const uint32_t n2 = next_pow2(n);
  const uint32_t m = min(n2, lcap);
  auto step = [&](uint64_t *a, uint32_t len, uint32_t size, uint32_t stride, bool mirror) {
    const uint32_t lg = 31 - __clz(stride);
    for (uint32_t t = TM::rank(); t < (len >> 1); t += TM::kSize) {
      const uint32_t i = ((t >> lg) << (lg + 1)) + (t & (stride - 1));
      const uint32_t j = mirror ? (i ^ (size - 1)) : (i + stride);
      if (j >= n) continue;
      const uint64_t ka = a[i], kb = a[j];
      if (kb < ka) {
        a[i] = kb;
        a[j] = ka;
      }
    }
  };
  // full: sizes 2..m; else the strides m/2..1 of a merge already past them
  auto chunks = [&](bool full) {
    for (uint32_t c0 = 0; c0 < n; c0 += m) {
      TM::sync();
      for (uint32_t i = TM::rank(); i < m; i += TM::kSize)
        lds[i] = c0 + i < n ? key[c0 + i] : ~0ull;
      for (uint32_t size = full ? 2 : m; size <= m; size <<= 1)
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
          TM::sync();
          // inside the chunk every position is loaded (pads are ~0)
          const uint32_t lg = 31 - __clz(stride);
          const bool mirror = full && stride == (size >> 1);
          for (uint32_t t = TM::rank(); t < (m >> 1); t += TM::kSize) {
            const uint32_t i = ((t >> lg) << (lg + 1)) + (t & (stride - 1));
            const uint32_t j = mirror ? (i ^ (size - 1)) : (i + stride);
            const uint64_t ka = lds[i], kb = lds[j];
            if (kb < ka) {
              lds[i] = kb;
              lds[j] = ka;
            }
          }
        }
      TM::sync();
      for (uint32_t i = TM::rank(); i < m && c0 + i < n; i += TM::kSize) key[c0 + i] = lds[i];
    }
    TM::sync();
  };
  chunks(true);
  for (uint32_t size = m << 1; size <= n2; size <<= 1) {
    for (uint32_t stride = size >> 1; stride >= m; stride >>= 1) {
      step(key, n2, size, stride, stride == (size >> 1));
      TM::sync();
    }
    chunks(false);
  }
}

// Workspace carved from LDS (or HBM for the large kernel).
struct Ws {
  uint8_t *q;          // qcap query bases
  uint32_t *hit_j;     // pcap
  uint64_t *hit_lo;    // pcap
  uint32_t *hit_base;  // pcap + 1
  uint64_t *k1;        // ecap
  uint32_t *v1;        // ecap
  uint64_t *k2;        // ecap
  uint32_t *v2;        // ecap
  uint32_t *t1, *t2, *t3, *rref;  // ecap each, one contiguous region (4 * ecap)
  uint32_t *bitmap;    // qcap / 32 + 1
  uint32_t *red;       // kRedWords words: block-scan partials, then broadcasts
  unsigned long long *acc;  // 2
  unsigned long long *stats;  // kNumStats, per workgroup
  struct RefSlot *slots;  // nslot references scored together
  uint32_t *slot_bm;      // nslot row bitmaps of qcap / 32 + 1 words
  uint64_t *stage;        // LDS sort buffer of stage_cap keys, or null
  uint32_t stage_cap;     // a power of two
};

// One reference of a multi-reference scoring batch.
struct RefSlot {
  unsigned long long acc[2];  // sumR, sumC
  uint64_t ro, len;           // the reference's byte offset and length
  uint32_t r, R;              // rank, reference id
  uint32_t c0, nchain;        // its first chain and chain count
  uint32_t nseed, base;       // seeds (chains + unstored copies of (0, 0)); first batch item
  uint32_t pad[2];
};

__host__ __device__ inline size_t al8(size_t x) { return (x + 7) & ~size_t(7); }

// red[0..15]: per-wavefront partials of block scans (up to 1024 threads);
// red[kBcast..]: values one lane hands to the team.
constexpr uint32_t kRedWords = 24;
constexpr uint32_t kBcast = 16;

__host__ __device__ inline size_t ws_bytes(const Caps &c) {
  size_t b = 0;
  b += al8(c.qcap);
  b += al8(c.pcap * 4ull) + al8(c.pcap * 8ull) + al8((c.pcap + 1) * 4ull);
  b += 2 * (al8(c.ecap * 8ull) + al8(c.ecap * 4ull));
  b += al8(c.ecap * 16ull);
  b += al8((c.qcap / 32 + 1) * 4ull);
  b += kRedWords * 4 + 16 + 64;
  b += c.nslot * (sizeof(RefSlot) + (c.qcap / 32 + 1) * 4ull);
  return b;
}

// Bytes between the workspaces of the teams of one workgroup.
__host__ __device__ inline size_t ws_stride_lds(const Caps &c) {
  return (ws_bytes(c) + 15) & ~size_t(15);
}

__device__ inline Ws ws_carve(unsigned char *base, const Caps &c) {
  Ws w;
  size_t o = 0;
  auto take = [&](size_t n) {
    unsigned char *p = base + o;
    o += al8(n);
    return p;
  };
  w.q = take(c.qcap);
  w.hit_j = (uint32_t *)take(c.pcap * 4ull);
  w.hit_lo = (uint64_t *)take(c.pcap * 8ull);
  w.hit_base = (uint32_t *)take((c.pcap + 1) * 4ull);
  w.k1 = (uint64_t *)take(c.ecap * 8ull);
  w.v1 = (uint32_t *)take(c.ecap * 4ull);
  w.k2 = (uint64_t *)take(c.ecap * 8ull);
  w.v2 = (uint32_t *)take(c.ecap * 4ull);
  w.t1 = (uint32_t *)take(c.ecap * 16ull);
  w.t2 = w.t1 + c.ecap;
  w.t3 = w.t1 + 2ull * c.ecap;
  w.rref = w.t1 + 3ull * c.ecap;
  w.bitmap = (uint32_t *)take((c.qcap / 32 + 1) * 4ull);
  w.red = (uint32_t *)take(kRedWords * 4);
  w.acc = (unsigned long long *)take(16);
  w.stats = (unsigned long long *)take(64);
  w.slots = (RefSlot *)take(c.nslot * sizeof(RefSlot));
  w.slot_bm = (uint32_t *)take(c.nslot * (c.qcap / 32 + 1) * 4ull);
  w.stage = nullptr;
  w.stage_cap = 0;
  return w;
}

// A workspace in HBM comes with an LDS sort buffer (ws.stage).  Only the
// HBM-workspace kernels' teams (256 and 1,024 threads) can have one, so the
// other teams compile without the staged paths (every test of the buffer
// goes through here, so a team always takes one form throughout).
template <class TM>
__device__ __forceinline__ bool has_stage(const Ws &ws) {
  if constexpr (TM::kPerWave == 1 && (TM::kSize == 256 || TM::kSize == 1024))
    return ws.stage != nullptr;
  else
    return false;
}

struct Ctx {
  IndexView ix;
  Luts lut;
  Reads rd;
  uint32_t flags;
  uint32_t max_len;  // longest read the handle accepts
  // Probe records written by k_probe for this launch's strands of length
  // up to rec_maxlen, or null (the kernel stages, checks, hashes and looks
  // up itself): rec[q * rec_stride] is kRecBad for a strand with a base
  // outside the alphabet, else 0; rec[q * rec_stride + 1 + p] is probe p.
  const uint64_t *rec;
  uint32_t rec_stride;
  uint32_t rec_maxlen;
  // Entries beyond which a strand skips the LDS overflow tiers (0: none).
  // A strand with more is passed on flagged (kBigFlag), and the LDS tiers
  // pass it on unread to the HBM-workspace tier.
  uint32_t lds_ecap_max;
  uint32_t kid;  // statistics row of this launch (kKernelRows)
  Outputs out;
};

// Probe record: lo << 24 | cnt of the probe's bucket; kRecAbsent when the
// hash has no bucket; cnt kRecCntMax when the count does not fit 24 bits.
constexpr uint64_t kRecAbsent = ~0ull;
constexpr uint64_t kRecBad = ~0ull - 1;
constexpr uint32_t kRecCntMax = 0xFFFFFFu;

enum : uint32_t { kDone = 0, kOverflow = 1, kOverflowBig = 2 };
// Overflow-list flag: the strand needs more entries than any LDS tier holds.
constexpr uint32_t kBigFlag = 1u << 31;

__device__ __forceinline__ float f32_from_bits(uint32_t u) { return __uint_as_float(u); }

// Relative score (MatrixResult::GetRelativeScore, MatrixResult.cpp:21-29)
// from the logf tables; returns false when the value is outside the tables.
__device__ __forceinline__ bool rel_score(const Luts &lut, uint32_t m, uint64_t ref_len,
                                          float score, float &rel) {
  const float x = score + 1.0f;
  const uint32_t xi = __float_as_uint(x) - 0x3f800000u;
  const uint64_t L = (uint64_t)m < ref_len ? (uint64_t)m : ref_len;
  if (xi >= lut.n_x || L >= lut.n_smax) return false;
  rel = lut.x_log[xi] / lut.smax_log[L];
  return true;
}

// The same with L = min(m, reference length) given.
__device__ __forceinline__ bool rel_score_l(const Luts &lut, uint32_t L, float score, float &rel) {
  const uint32_t xi = __float_as_uint(score + 1.0f) - 0x3f800000u;
  if (xi >= lut.n_x || L >= lut.n_smax) return false;
  rel = lut.x_log[xi] / lut.smax_log[L];
  return true;
}

// A scored reference whose relative score is still to be looked up is
// marked kRelPending + L in valid_of.
constexpr uint32_t kRelPending = 2;

// Score = (sumR / sumC) / 100 * percentageMatch (MatrixCPU.cpp:195-213).
__device__ __forceinline__ float matrix_score(uint64_t sum_r, uint64_t sum_c, uint32_t hits,
                                              uint32_t m) {
  if (sum_c == 0) return -2.0f;
  const float sf = (float)sum_r / (float)sum_c;
  const float pct = (float)((100.0 / (double)(float)m) * (double)(float)hits);
  return (float)(((double)sf / 100.0) * (double)pct);
}

// BL (4 or 8) consecutive bytes from p (any alignment) as BL / 4
// little-endian words, read as the aligned dwords that hold them (one more
// than BL / 4); dwords below lo read as 0.
template <uint32_t BL>
__device__ __forceinline__ void load_bytes(const uint8_t *p, uintptr_t lo, uint32_t (&w)[BL / 4]) {
  constexpr uint32_t ND = BL / 4 + 1;
  const uintptr_t a = (uintptr_t)p & ~(uintptr_t)3;
  const uint32_t sh = (uint32_t)((uintptr_t)p & 3);
  uint32_t d[ND];
#pragma unroll
  for (uint32_t i = 0; i < ND; i++) d[i] = a + 4 * i >= lo ? *(const uint32_t *)(a + 4 * i) : 0u;
#pragma unroll
  for (uint32_t i = 0; i < BL / 4; i++) w[i] = __builtin_amdgcn_alignbyte(d[i + 1], d[i], sh);
}

// Bit 8j + 7 set where byte j of the BL bytes in a and b differs.
template <uint32_t BL>
__device__ __forceinline__ uint64_t byte_diff(const uint32_t (&a)[BL / 4], const uint32_t (&b)[BL / 4]) {
  uint64_t nz = 0;
#pragma unroll
  for (uint32_t i = 0; i < BL / 4; i++) {
    const uint32_t x = a[i] ^ b[i];
    nz |= (uint64_t)((((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x) & 0x80808080u) << (32 * i);
  }
  return nz;
}

// Head/tail extension of one seed by one seed worker (MatrixCPU.cpp:93-172).
// Returns the matched length tot and the head length h; all lanes agree.
// Tail and head are independent, so each round loads a chunk of both before
// the ballots find the first mismatch: every lane takes 64 / kSeedLanes
// consecutive bases in each direction, so one memory round trip covers 64
// bases (the whole extension at T <= 64).  A lane reads its reference and
// query bases as words (the aligned dwords that hold them; a reference's
// bytes are followed by 64 bytes of padding, and nothing below the bytes
// buffer `lo` is read) and finds its first differing base with byte
// arithmetic (query words below the query's first byte, which is 4-byte
// aligned and followed by more of its workspace, read as 0).  The head's
// bases l = l0 .. l0 + BL - 1 lie at descending
// positions, so its word block holds base l at byte BL - 1 - (l - l0).  The
// first mismatch is the lowest position of the first lane whose bases hold
// one; that lane's position is read with a cross-lane permute.
template <class TM>
__device__ __forceinline__ void extend_seed(const uint8_t *q, uint32_t m, const uint8_t *rb,
                                            uint64_t ref_len, uint32_t sref, uint32_t rs,
                                            uint32_t lref, uint32_t re, int T, bool ultrafast,
                                            const uint8_t *lo, uint32_t &tot, uint32_t &head,
                                            uint32_t &compares) {
  constexpr uint32_t W = TM::kSeedLanes;
  const uint32_t lane = seed_lane<TM>();
  const uint32_t mid = re - rs;
  uint32_t t = 0, h = 0;
  compares = 0;
  if (!ultrafast) {
    // Tail: at most min(T, ref_len - lref) bases, inside the query.
    uint64_t tl64 = (uint64_t)lref < ref_len ? ref_len - lref : 0;
    tl64 = tl64 < (uint64_t)T ? tl64 : (uint64_t)T;
    const uint64_t qroom = re < m ? (uint64_t)(m - re) : 0;
    const uint32_t tl = (uint32_t)(tl64 < qroom ? tl64 : qroom);
    // Head: query row rs-1-l against reference byte sref-l for l <= lb; the
    // base that meets a boundary is not counted (h stays lb without a
    // mismatch).
    uint32_t hl = 0, lb = 0;
    if (rs > 0) {
      const uint32_t hs = sref < (uint32_t)T ? sref : (uint32_t)T;
      if (hs > 0) {
        lb = (rs - 1) < (hs - 1) ? (rs - 1) : (hs - 1);
        hl = lb + 1;
      }
    }
    t = tl;
    h = lb;
    bool tdone = tl == 0, hdone = hl == 0;
    constexpr uint32_t BL = 64 / W;  // bases per lane, direction and round
    static_assert(BL == 4 || BL == 8, "4 or 8 bases per lane");
    constexpr uint32_t CH = BL * W;
    constexpr uint64_t kAll = BL == 8 ? ~0ull : 0xFFFFFFFFull;
    for (uint32_t b = 0; !(tdone && hdone); b += CH) {
      const uint32_t l0 = b + BL * lane;
      uint64_t tm = 0, hm = 0;  // bit 8j + 7: base j of the lane's block differs
      if (!tdone && l0 < tl) {
        uint32_t rw[BL / 4], qw[BL / 4];
        load_bytes<BL>(rb + (uint64_t)lref + l0, (uintptr_t)lo, rw);
        load_bytes<BL>(q + re + l0, (uintptr_t)q, qw);
        const uint32_t nv = tl - l0 < BL ? tl - l0 : BL;
        tm = byte_diff<BL>(rw, qw) & (nv >= BL ? kAll : ((1ull << (8 * nv)) - 1));
      }
      if (!hdone && l0 < hl) {
        // reference bytes sref-l0-BL+1 .. sref-l0, query rs-l0-BL .. rs-1-l0
        uint32_t rw[BL / 4], qw[BL / 4];
        load_bytes<BL>(rb + (int64_t)sref - (int64_t)l0 - (int64_t)(BL - 1), (uintptr_t)lo, rw);
        load_bytes<BL>(q + (int64_t)rs - (int64_t)l0 - (int64_t)BL, (uintptr_t)q, qw);
        const uint32_t nv = hl - l0 < BL ? hl - l0 : BL;  // bytes BL-nv .. BL-1
        hm = byte_diff<BL>(rw, qw) & (kAll & ~((1ull << (8 * (BL - nv))) - 1));
      }
      const unsigned long long bt = seed_ballot<TM>(tm != 0);
      const unsigned long long bh = seed_ballot<TM>(hm != 0);
      // each lane's first differing base (read from the first lane that has one)
      const uint32_t pt = l0 + (tm ? (uint32_t)(__ffsll((long long)tm) - 1) >> 3 : 0u);
      const uint32_t ph = l0 + (hm ? BL - 1 - ((63u - (uint32_t)__clzll((long long)hm)) >> 3) : 0u);
      if (!tdone) {
        if (bt) {
          t = bperm(src_at<W>((uint32_t)__ffsll((long long)bt) - 1), pt);
          tdone = true;
        } else if (b + CH >= tl) {
          tdone = true;
        }
      }
      if (!hdone) {
        if (bh) {
          h = bperm(src_at<W>((uint32_t)__ffsll((long long)bh) - 1), ph);
          hdone = true;
        } else if (b + CH >= hl) {
          hdone = true;
        }
      }
    }
    // the reference compares until the first mismatch or the limit
    compares = t + (t < tl ? 1u : 0u) + (hl ? h + 1 : 0u);
  }
  tot = t + h + mid;
  head = h;
}

template <class TM>
__device__ __forceinline__ void set_bits(uint32_t *bm, uint32_t a, uint32_t len) {
  const uint32_t lane = seed_lane<TM>();
  const uint32_t b = a + len;  // [a, b)
  const uint32_t w0 = a >> 5, w1 = (b - 1) >> 5;
  for (uint32_t w = w0 + lane; w <= w1; w += TM::kSeedLanes) {
    uint32_t mask = 0xFFFFFFFFu;
    if (w == w0) mask &= 0xFFFFFFFFu << (a & 31);
    if (w == w1) mask &= 0xFFFFFFFFu >> (31 - ((b - 1) & 31));
    atomicOr(&bm[w], mask);
  }
}

// Scores one reference from its seeds: every wave takes seeds in turn.
// get_seed(i, sref, sq, lref, lq, wt) supplies seed i and its number of
// identical copies wt, whose sums count wt times (the row bitmap once).
// Returns on all threads.
template <class TM, class GetSeed>
__device__ void score_ref(const uint8_t *q, uint32_t m, const uint8_t *rb, uint64_t ref_len,
                          const uint8_t *lo, uint32_t n_seeds, int T, bool ultrafast, const Ws &ws,
                          GetSeed get_seed, uint64_t &sum_r, uint64_t &sum_c, uint32_t &hits) {
  constexpr uint32_t NW = TM::kSize / TM::kSeedLanes, BS = TM::kSize;
  const uint32_t wid = TM::rank() / TM::kSeedLanes, lane = seed_lane<TM>();
  const uint32_t nw = (m + 31) >> 5;
  TM::sync();
  for (uint32_t w = TM::rank(); w < nw; w += BS) ws.bitmap[w] = 0;
  if (TM::rank() == 0) {
    ws.acc[0] = 0;
    ws.acc[1] = 0;
  }
  TM::sync();
  unsigned long long sr = 0, sc = 0, ncmp = 0;
  for (uint32_t i = wid; i < n_seeds; i += NW) {
    uint32_t sref, sq, lref, lq, wt;  // wt: copies of this seed
    get_seed(i, sref, sq, lref, lq, wt);
    const uint32_t rs = sq, re = lq;
    if (rs > re) continue;  // MatrixCPU.cpp:57-58
    uint32_t tot, h, cmp;
    extend_seed<TM>(q, m, rb, ref_len, sref, rs, lref, re, T, ultrafast, lo, tot, h, cmp);
    ncmp += (unsigned long long)wt * cmp;
    if (tot >= (uint32_t)T) {  // MatrixCPU.cpp:178-184
      sc += (unsigned long long)wt * tot;
      sr += (unsigned long long)wt * ((unsigned long long)tot * (tot + 1) / 2);
      set_bits<TM>(ws.bitmap, rs - h, tot);
    }
  }
  if (lane == 0 && sc) {
    atomicAdd(&ws.acc[0], sr);
    atomicAdd(&ws.acc[1], sc);
  }
  if (lane == 0 && ncmp && ws.stats) atomicAdd(&ws.stats[kExt], ncmp);
  TM::sync();
  uint32_t pc = 0;
  for (uint32_t w = TM::rank(); w < nw; w += BS) pc += __popc(ws.bitmap[w]);
  hits = TM::sum(pc, ws.red);
  TM::sync();
  sum_r = ws.acc[0];
  sum_c = ws.acc[1];
}

// Scores every rank whose longest chain equals G, for teams of whole
// workgroups: up to nslot references at a time get a slot (sums, row bitmap),
// and the seeds of all of them are spread over the seed workers, so the
// memory round trips of different references overlap.  Per reference the
// sums, bitmap and score are those of score_ref (integer sums and a bitwise
// OR do not depend on the order).  Sets rel_of / valid_of per rank (0 for
// ranks not scored); seeds_total (rank 0's share; the other threads add
// theirs to the statistics), seeded and lut_bad are meaningful on rank 0.
template <class TM>
__device__ __forceinline__ void score_multi(const Ctx &c, const Ws &ws, const Caps &cap,
                                            uint32_t m, uint32_t nranks, uint32_t nchains,
                                            uint32_t G, uint32_t nhits, uint32_t nph,
                                            bool ultrafast, uint32_t read, uint32_t strand,
                                            unsigned long long &seeds_total, uint32_t &seeded,
                                            uint32_t &lut_bad) {
  constexpr uint32_t BS = TM::kSize, NW = BS / TM::kSeedLanes;
  static_assert(BS >= 64 && BS % 64 == 0, "whole-workgroup teams");
  const uint32_t tid = TM::rank(), wid = tid / TM::kSeedLanes, lane = seed_lane<TM>();
  const IndexView &ix = c.ix;
  const int T = ix.T, s = ix.s;
  const uint32_t *cfl = (const uint32_t *)ws.k1;
  float *rel_of = (float *)ws.t2;
  uint32_t *valid_of = ws.t1;
  const uint32_t nw = (m + 31) >> 5, bw = nw + 1;  // row bitmap words for this read
  // The slot region (cap.nslot slots whose bitmaps fit reads of cap.qcap
  // bases) holds more slots for a shorter read: up to one per thread, so that
  // fewer rounds score a strand's references (a 150 bp read gets 69 in the
  // 1 kb second-chance tier instead of 31).
  const uint32_t region = cap.nslot * (uint32_t)(sizeof(RefSlot) + (cap.qcap / 32 + 1) * 4);
  const uint32_t nslot = min(BS, region / (uint32_t)(sizeof(RefSlot) + bw * 4));
  uint32_t *const slot_bm = (uint32_t *)(ws.slots + nslot);
  seeds_total = 0;
  seeded = 0;
  unsigned long long seeds_part = 0;  // this thread's slots' seeds
  for (uint32_t r = tid; r < nranks; r += BS) valid_of[r] = 0;
  if (tid == 0) ws.red[kBcast + 3] = 0;  // a relative score outside the tables
  for (uint32_t base = 0; base < nranks; base += BS) {
    const uint32_t r = base + tid;
    const bool f = r < nranks && ws.t3[r] == G;
    // This thread's reference, where its bytes lie (loaded once, before the
    // rounds, so that the loads of every round's references are in flight
    // together), and its seeds: its chains, and for reference 0 the unstored
    // copies of (0, 0) as one-pair seeds, one per probe that hit bucket 0
    // (Analyser.cpp:2421-2481).
    uint32_t R = 0, c0 = 0, nchain = 0, nseed = 0;
    uint64_t ro = 0, rlen = 0;
    if (f) {
      R = ws.rref[r];
      ro = ix.ref_off[R];
      rlen = ix.ref_off[R + 1] - ro;
      c0 = ws.v1[r];
      nchain = ((r + 1 < nranks) ? ws.v1[r + 1] : nchains) - c0;
      nseed = nchain + (R == 0 ? nph : 0u);
    }
    // One scan numbers the candidates (bits 21 and up) and the seeds before
    // each (bits 0-20: a team's seeds stay below 2^21, its candidates below
    // 2^11).  When they all fit one round, each slot's first item is its
    // seed offset from this scan.
    uint32_t sums;
    const uint32_t ex = TM::excl_scan(f ? (1u << 21) | nseed : 0u, ws.red, sums);
    const uint32_t idx = ex >> 21, tot = sums >> 21;
    const bool one_round = tot <= nslot;
    for (uint32_t g = 0; g < tot; g += nslot) {
      const uint32_t nb = min(nslot, tot - g);
      TM::sync();
      if (f && idx >= g && idx < g + nb) {
        RefSlot &sl = ws.slots[idx - g];
        sl.acc[0] = 0;
        sl.acc[1] = 0;
        sl.ro = ro;
        sl.len = rlen;
        sl.r = r;
        sl.R = R;
        sl.c0 = c0;
        sl.nchain = nchain;
        sl.nseed = nseed;
        if (one_round) {
          sl.base = ex & 0x1FFFFFu;
          seeds_part += nchain + (unsigned long long)(nseed - nchain) * ix.phantom;
        }
      }
      for (uint32_t w = tid; w < nb * bw; w += BS) slot_bm[w] = 0;
      TM::sync();
      uint32_t nitems;
      if (one_round) {
        nitems = sums & 0x1FFFFFu;
      } else {
        // each slot's first item: a scan of the round's seed counts (nb <= BS)
        const uint32_t ns = tid < nb ? ws.slots[tid].nseed : 0u;
        const uint32_t base_t = TM::excl_scan(ns, ws.red, nitems);
        if (tid < nb) {
          ws.slots[tid].base = base_t;
          seeds_part += ws.slots[tid].nchain +
                        (unsigned long long)(ws.slots[tid].nseed - ws.slots[tid].nchain) *
                            ix.phantom;
        }
        TM::sync();
      }
      seeded += nb;
      unsigned long long ncmp = 0;
      for (uint32_t i = wid; i < nitems; i += NW) {
        uint32_t b = 0, hi = nb;  // the slot of item i: the last with base <= i
        while (hi - b > 1) {
          const uint32_t mid = (b + hi) >> 1;
          if (ws.slots[mid].base <= i) b = mid;
          else hi = mid;
        }
        const RefSlot &sl = ws.slots[b];
        const uint32_t jj = i - sl.base;
        uint32_t sref, sq, lref, lq, wt;
        if (jj < sl.nchain) {
          const uint32_t cid = sl.c0 + jj;
          const uint32_t f0 = cfl[2 * cid], l0 = cfl[2 * cid + 1];
          sref = (uint32_t)ws.k2[f0];
          sq = ws.v2[f0];
          lref = (uint32_t)ws.k2[l0] + (uint32_t)s;  // Analyser.cpp:2465-2467
          lq = ws.v2[l0] + (uint32_t)s;
          wt = 1;
        } else {
          uint32_t v = jj - sl.nchain, j = 0;
          for (uint32_t t = 0; t < nhits; t++)
            if (ws.hit_lo[t] == 0 && ws.hit_base[t + 1] > ws.hit_base[t]) {
              if (v == 0) {
                j = ws.hit_j[t];
                break;
              }
              v--;
            }
          sref = 0;
          sq = j;
          lref = (uint32_t)s;
          lq = j + (uint32_t)s;
          wt = ix.phantom;
        }
        const uint32_t rs = sq, re = lq;
        if (rs > re) continue;  // MatrixCPU.cpp:57-58
        uint32_t tl, h, cmp;
        extend_seed<TM>(ws.q, m, ix.bytes + sl.ro, sl.len, sref, rs, lref, re, T, ultrafast,
                        ix.bytes, tl, h, cmp);
        ncmp += (unsigned long long)wt * cmp;
        if (tl >= (uint32_t)T) {  // MatrixCPU.cpp:178-184
          if (lane == 0) {
            atomicAdd(&ws.slots[b].acc[0],
                      (unsigned long long)wt * ((unsigned long long)tl * (tl + 1) / 2));
            atomicAdd(&ws.slots[b].acc[1], (unsigned long long)wt * tl);
          }
          set_bits<TM>(slot_bm + b * bw, rs - h, tl);
        }
      }
      if (lane == 0 && ncmp && ws.stats) atomicAdd(&ws.stats[kExt], ncmp);
      TM::sync();
      // one thread per slot: row hits, Score; without the candidate output
      // the relative score waits for the pass after the last round (the table
      // lookups of every scored reference in flight together)
      for (uint32_t b = tid; b < nb; b += BS) {
        const uint32_t *bm = slot_bm + b * bw;
        uint32_t hits = 0;
        for (uint32_t w = 0; w < nw; w++) hits += __popc(bm[w]);
        if (!c.out.cands) {
          const RefSlot &sl = ws.slots[b];
          const float score = matrix_score(sl.acc[0], sl.acc[1], hits, m);
          rel_of[sl.r] = score;
          // MatrixCPU.cpp:356; pending with L = min(m, reference length)
          valid_of[sl.r] = score > -1.0f ? kRelPending + (uint32_t)min((uint64_t)m, sl.len) : 0u;
        } else {
          const RefSlot &sl = ws.slots[b];
          const float score = matrix_score(sl.acc[0], sl.acc[1], hits, m);
          float rel = -1.0f;
          uint32_t valid = 0;
          if (score > -1.0f) {  // MatrixCPU.cpp:356
            if (rel_score(c.lut, m, sl.len, score, rel)) valid = 1;
            else ws.red[kBcast + 3] = 1;
          }
          rel_of[sl.r] = rel;
          valid_of[sl.r] = valid;
          const unsigned long long slot = atomicAdd(&c.out.ctr->cand_used, 1ull);
          if (slot < c.out.cand_cap) {
            mh_cand cd;
            cd.read = read;
            cd.strand = strand;
            cd.ref = sl.R;
            cd.hits = hits;
            cd.sum_r = sl.acc[0];
            cd.sum_c = sl.acc[1];
            cd.score = score;
            cd.rel = rel;
            cd.n_seeds = (uint32_t)(sl.nchain +
                                    (unsigned long long)(sl.nseed - sl.nchain) * ix.phantom);
            cd.pad = 0;
            c.out.cands[slot] = cd;
          }
        }
      }
    }
  }
  if (!c.out.cands) {
    // relative scores of the pending references (MatrixResult.cpp:21-29)
    TM::sync();
    for (uint32_t r = tid; r < nranks; r += BS) {
      const uint32_t v = valid_of[r];
      if (v >= kRelPending) {
        float rel = -1.0f;
        if (rel_score_l(c.lut, v - kRelPending, rel_of[r], rel)) {
          valid_of[r] = 1;
        } else {
          valid_of[r] = 0;
          ws.red[kBcast + 3] = 1;
        }
        rel_of[r] = rel;
      }
    }
  }
  // rank 0 adds the count of seeds to the statistics with the others'
  if (seeds_part && ws.stats && tid != 0) atomicAdd(&ws.stats[kSeeds], seeds_part);
  if (tid == 0) seeds_total = seeds_part;
  TM::sync();
  lut_bad = ws.red[kBcast + 3];
}

template <class TM>
__device__ __forceinline__ void write_query(const Ctx &c, uint32_t q, const QueryOut &r) {
  if (TM::rank() == 0) c.out.q[q] = r;
}

constexpr uint32_t kEmpty = 0xFFFFFFFFu;

__device__ __forceinline__ uint32_t hash_slot(uint32_t key, uint32_t mask) {
  return (key * 0x9E3779B1u) & mask;
}

// Inserts key into an open-addressing set; returns true when it was new.
__device__ __forceinline__ bool set_insert(uint32_t *set, uint32_t mask, uint32_t key) {
  for (uint32_t s = hash_slot(key, mask);; s = (s + 1) & mask) {
    const uint32_t old = atomicCAS(&set[s], kEmpty, key);
    if (old == kEmpty) return true;
    if (old == key) return false;
  }
}

// Slot of key in an open-addressing map, inserting it when absent.
__device__ __forceinline__ uint32_t map_slot(uint32_t *keys, uint32_t mask, uint32_t key) {
  for (uint32_t s = hash_slot(key, mask);; s = (s + 1) & mask) {
    const uint32_t old = atomicCAS(&keys[s], kEmpty, key);
    if (old == kEmpty || old == key) return s;
  }
}

// Runs of consecutive hit indices over sorted pair keys: p = key >> KS
// holds the hit index in its low RS bits and the reference above them.
// Returns, per element, the length of its run in `ids` (distinct hits; a
// repeat of one hit in a bucket does not lengthen a run, Analyser.cpp:2057-
// 2058), using `len` as scratch.  All threads get the maximum.
template <int KS, int RS>
__device__ __forceinline__ uint32_t run_flags(const uint64_t *key, uint32_t t) {
  if (t == 0) return 3u;
  const uint64_t p = key[t] >> KS, pp = key[t - 1] >> KS;
  const bool same = (p >> RS) == (pp >> RS);
  return ((!same || (p != pp && p != pp + 1)) ? 1u : 0u) | ((!same || p != pp) ? 2u : 0u);
}

template <class TM, int KS, int RS>
__device__ uint32_t run_lengths(const Ws &ws, const uint64_t *key, uint32_t n, uint32_t *ids,
                                uint32_t *len) {
  constexpr uint32_t BS = TM::kSize;
  const uint32_t tid = TM::rank();
  uint32_t nruns = 0;
  for (uint32_t base = 0; base < n; base += BS) {
    const uint32_t t = base + tid;
    const uint32_t st = t < n ? (run_flags<KS, RS>(key, t) & 1u) : 0u;
    uint32_t tot;
    const uint32_t ex = TM::excl_scan(st, ws.red, tot);
    if (t < n) ids[t] = nruns + ex + st - 1;
    nruns += tot;
  }
  TM::sync();
  for (uint32_t r = tid; r < nruns; r += BS) len[r] = 0;
  TM::sync();
  for (uint32_t t = tid; t < n; t += BS)
    if (run_flags<KS, RS>(key, t) & 2u) atomicAdd(&len[ids[t]], 1u);
  TM::sync();
  uint32_t mx = 0;
  for (uint32_t t = tid; t < n; t += BS) {
    const uint32_t L = len[ids[t]];
    mx = max(mx, L);
    ids[t] = L;  // per element from here on
  }
  const uint32_t m = TM::max(mx, ws.red);
  TM::sync();
  return m;
}

// Successive-occurrence filter with hashing.  A reference with a single
// bucket entry has a run of 1; only references with two or more entries
// (the read's own reference and repeats, few) are sorted by (ref, hit) to
// measure their runs.  The reference map has its values in t1..t2 and its
// keys in t3..rref (2E slots each); the keys are free again after the
// inserts and hold the run scratch.
template <class TM>
__device__ void filter_hashed(const Ws &ws, uint32_t E, uint32_t ecap, uint32_t &nP) {
  constexpr uint32_t BS = TM::kSize;
  const uint32_t tid = TM::rank();
  // 2E slots, at most the largest power of two in the 2 * ecap words (more
  // than E slots either way)
  const uint32_t room = 1u << (31 - __clz(2 * ecap));
  const uint32_t cap = min(next_pow2(2 * E), room), mask = cap - 1;
  uint32_t *rval = ws.t1, *rkeys = ws.t3;
  for (uint32_t i = tid; i < cap; i += BS) {
    rkeys[i] = kEmpty;
    rval[i] = 0;
  }
  TM::sync();
  for (uint32_t e = tid; e < E; e += BS) {
    const uint32_t r = (uint32_t)(ws.k1[e] >> 32);
    const uint32_t slot = map_slot(rkeys, mask, r);
    atomicAdd(&rval[slot], 1u);
    ws.v2[e] = slot;  // map slot of each pair
  }
  TM::sync();
  // Pairs of references with two or more entries -> k2 as
  // ((ref << 12 | hit) << 32 | slot).
  uint32_t nm = 0;
  for (uint32_t base = 0; base < E; base += BS) {
    const uint32_t e = base + tid;
    uint32_t multi = 0, slot = 0;
    uint64_t key = 0;
    if (e < E) {
      slot = ws.v2[e];
      key = ws.k1[e];
      multi = rval[slot] >= 2 ? 1u : 0u;
    }
    uint32_t tot;
    const uint32_t ex = TM::excl_scan(multi, ws.red, tot);
    if (multi)
      ws.k2[nm + ex] =
          ((uint64_t)(((uint32_t)(key >> 32) << 12) | (uint32_t)key) << 32) | slot;
    nm += tot;
  }
  TM::sync();
  // Every reference starts with a run of 1; the sorted pass raises it.
  for (uint32_t i = tid; i < cap; i += BS) rval[i] = 1;
  TM::sync();
  uint32_t maxrun = 0;
  if (nm) {
    if (has_stage<TM>(ws))
      staged_sort_keys<TM>(ws.k2, nm, ws.stage, ws.stage_cap);
    else
      bitonic_sort<TM>(ws.k2, ws.t3, nm);  // t3: value scratch, unused
    maxrun = run_lengths<TM, 32, 12>(ws, ws.k2, nm, ws.t3, ws.rref);
    for (uint32_t t = tid; t < nm; t += BS) atomicMax(&rval[(uint32_t)ws.k2[t]], ws.t3[t]);
    TM::sync();
  }
  const uint32_t cut = maxrun >= 2 ? maxrun - 1 : 1;
  // Kept pairs, compacted in place into k2/v2 (position <= source index;
  // each chunk is read before it is written).
  uint32_t nkeep = 0;
  for (uint32_t base = 0; base < E; base += BS) {
    const uint32_t e = base + tid;
    uint32_t keep = 0, pos = 0, j = 0;
    uint64_t key = 0;
    if (e < E) {
      key = ws.k1[e];
      pos = ws.v1[e];
      if (!has_stage<TM>(ws)) j = ws.hit_j[(uint32_t)key];
      keep = rval[ws.v2[e]] >= cut ? 1u : 0u;
    }
    uint32_t tot;
    const uint32_t ex = TM::excl_scan(keep, ws.red, tot);
    TM::sync();
    if (keep) {
      if (has_stage<TM>(ws)) {
        // packed for a key-only sort: ref << 44 | refPos << 12 | hit
        ws.k2[nkeep + ex] = ((key >> 32) << 44) | ((uint64_t)pos << 12) | (uint32_t)key;
      } else {
        ws.k2[nkeep + ex] = (key & 0xFFFFFFFF00000000ull) | pos;
        ws.v2[nkeep + ex] = j;
      }
    }
    nkeep += tot;
  }
  nP = nkeep;
  TM::sync();
}

// Successive-occurrence filter for at most 64 hits: each reference's slot in
// an LDS hash map collects the mask of the hits it occurs in (a repeat of one
// hit sets one bit, Analyser.cpp:2057-2058); its longest run of set bits is
// its longest run of consecutive hit indices.  Same decisions as
// filter_hashed without its sort.  Kept pairs go to k2 packed as
// ref << 38 | refPos << 6 | hit, unsorted.  The map has `cap` slots, a power
// of two above E (every reference then finds one), in the t1..rref region:
// masks (MT: 32 bits for up to 32 hits, else 64), then the keys; runs are
// recomputed from the masks.  References < 2^26.
__device__ __forceinline__ uint32_t longest_run(uint64_t x) {
  uint32_t r = 0;
  while (x) {
    x &= x >> 1;
    r++;
  }
  return r;
}

// Map slots filter_masked can use: twice the entries when the region has
// room, else the largest power of two it holds; 0 when that is not above E.
__device__ __forceinline__ uint32_t masked_slots(uint32_t E, uint32_t nhits, uint32_t ecap) {
  const uint32_t words = 4 * ecap;                       // t1..rref
  const uint32_t per_slot = nhits <= 32 ? 2u : 3u;       // mask words + key
  const uint32_t room = 1u << (31 - __clz(words / per_slot));
  const uint32_t cap = min(next_pow2(2 * E), room);
  return cap > E ? cap : 0u;
}

template <class TM, class MT>
__device__ uint32_t filter_masked(const Ws &ws, uint32_t E, uint32_t cap, uint32_t &nP) {
  constexpr uint32_t BS = TM::kSize;
  const uint32_t tid = TM::rank();
  const uint32_t mask = cap - 1;
  MT *hm = reinterpret_cast<MT *>(ws.t1);
  uint32_t *rkeys = ws.t1 + cap * (uint32_t)(sizeof(MT) / 4);
  for (uint32_t i = tid; i < cap; i += BS) {
    rkeys[i] = kEmpty;
    hm[i] = 0;
  }
  TM::sync();
  for (uint32_t e = tid; e < E; e += BS) {
    const uint64_t k = ws.k1[e];
    const uint32_t slot = map_slot(rkeys, mask, (uint32_t)(k >> 32));
    atomicOr(&hm[slot], (MT)1 << ((uint32_t)k & 63));
    ws.v2[e] = slot;
  }
  TM::sync();
  // each entry's run, from its reference's mask (v2 takes the run)
  uint32_t mr = 0;
  for (uint32_t e = tid; e < E; e += BS) {
    const uint32_t r = longest_run(hm[ws.v2[e]]);
    ws.v2[e] = r;
    mr = max(mr, r);
  }
  const uint32_t maxrun = TM::max(mr, ws.red);
  nP = 0;
  if (maxrun < 2) return maxrun;  // cut 1 keeps every pair, and none links
  const uint32_t cut = maxrun - 1;
  auto pack = [&](uint32_t e) {
    const uint64_t k = ws.k1[e];
    return ((k >> 32) << 38) | ((uint64_t)ws.v1[e] << 6) | ((uint32_t)k & 63);
  };
  if constexpr (TM::kPerWave == 1 && BS > 64) {
    // contiguous runs of entries, one team scan (two barriers) for all of them
    const uint32_t per = (E + BS - 1) / BS;
    const uint32_t lo = min(E, tid * per), hi = min(E, lo + per);
    uint32_t cnt = 0;
    for (uint32_t e = lo; e < hi; e++) cnt += ws.v2[e] >= cut ? 1u : 0u;
    uint32_t tot;
    uint32_t at = TM::excl_scan(cnt, ws.red, tot);
    for (uint32_t e = lo; e < hi; e++)
      if (ws.v2[e] >= cut) ws.k2[at++] = pack(e);
    nP = tot;
  } else {
    uint32_t nkeep = 0;
    for (uint32_t base = 0; base < E; base += BS) {
      const uint32_t e = base + tid;
      const uint32_t keep = (e < E && ws.v2[e] >= cut) ? 1u : 0u;
      uint32_t tot;
      const uint32_t ex = TM::excl_scan(keep, ws.red, tot);
      if (keep) ws.k2[nkeep + ex] = pack(e);
      nkeep += tot;
    }
    nP = nkeep;
  }
  TM::sync();
  return maxrun;
}

// Successive-occurrence filter by sorting every pair by (ref, hit); used
// when (ref, hit) does not pack into 32 bits.
template <class TM>
__device__ void filter_sorted(const Ws &ws, uint32_t E, uint32_t &nP) {
  constexpr uint32_t BS = TM::kSize;
  const uint32_t tid = TM::rank();
  bitonic_sort<TM>(ws.k1, ws.v1, E);
  const uint32_t maxrun = run_lengths<TM, 0, 32>(ws, ws.k1, E, ws.t2, ws.t3);
  const uint32_t cut = maxrun >= 2 ? maxrun - 1 : 1;
  // longest run per reference segment
  for (uint32_t t = tid; t < E; t += BS)
    ws.t1[t] = (t == 0 || (ws.k1[t - 1] >> 32) != (ws.k1[t] >> 32)) ? 1u : 0u;
  TM::sync();
  uint32_t nseg = 0;
  for (uint32_t base = 0; base < E; base += BS) {
    const uint32_t t = base + tid;
    const uint32_t st = t < E ? ws.t1[t] : 0u;
    uint32_t tot;
    const uint32_t ex = TM::excl_scan(st, ws.red, tot);
    if (t < E) ws.t3[t] = nseg + ex + st - 1;
    nseg += tot;
  }
  TM::sync();
  for (uint32_t r = tid; r < nseg; r += BS) ws.rref[r] = 0;
  TM::sync();
  for (uint32_t t = tid; t < E; t += BS) atomicMax(&ws.rref[ws.t3[t]], ws.t2[t]);
  TM::sync();
  uint32_t nkeep = 0;
  for (uint32_t base = 0; base < E; base += BS) {
    const uint32_t t = base + tid;
    const uint32_t keep = (t < E && ws.rref[ws.t3[t]] >= cut) ? 1u : 0u;
    uint32_t tot;
    const uint32_t ex = TM::excl_scan(keep, ws.red, tot);
    if (keep) {
      const uint64_t key = ws.k1[t];
      ws.k2[nkeep + ex] = (key & 0xFFFFFFFF00000000ull) | ws.v1[t];
      ws.v2[nkeep + ex] = ws.hit_j[(uint32_t)key];
    }
    nkeep += tot;
  }
  nP = nkeep;
  TM::sync();
}

enum : uint32_t { kSmallNo = 0, kSmallDirect, kSmallSorted };

template <class TM>
__device__ __forceinline__ uint64_t team_or64(uint64_t x) {
  constexpr uint32_t W = TM::kWorkerLanes;
  return ((uint64_t)seg_reduce<W, 2>((uint32_t)(x >> 32)) << 32) |
         seg_reduce<W, 2>((uint32_t)x);
}

// Bitonic network of N2 elements held two to a lane (element i: slot i & 1
// of lane i >> 1), fully unrolled so that every partner distance is a
// constant (cross-lane moves through DPP up to 8 lanes apart).
template <uint32_t N2>
__device__ __forceinline__ void lane_bitonic(uint64_t (&P)[2], uint32_t lane) {
#pragma unroll
  for (uint32_t k = 2; k <= N2; k <<= 1) {
    const bool asc = ((2 * lane) & k) == 0;
#pragma unroll
    for (uint32_t d = k >> 1; d > 1; d >>= 1) {
      const uint32_t lx = d >> 1;  // partner lane = lane ^ lx
      // the lower lane of a pair keeps the minimum in an ascending run
      const bool want_min = ((lane & lx) == 0) == asc;
#pragma unroll
      for (int j = 0; j < 2; j++) {
        const uint64_t pk = xor_lane64_any(lx, P[j]);
        P[j] = ((pk < P[j]) == want_min) ? pk : P[j];
      }
    }
    // d == 1: the two slots of a lane
    const uint64_t lo = P[0] < P[1] ? P[0] : P[1], hi = P[0] < P[1] ? P[1] : P[0];
    P[0] = asc ? lo : hi;
    P[1] = asc ? hi : lo;
  }
}

// Small working sets (at most 64 entries and 64 hits, fewer than 2^26
// references, a team inside one wavefront) in registers: element i of the
// team's array is slot i & 1 of lane i >> 1.  Gives the decisions of the LDS
// path below.
//  * Direct hit (whole-wavefront teams): tested on the gathered entries with
//    ballots (r* from the first bucket: as many entries as hits, all on its
//    diagonal; no other reference of the second bucket present in every
//    bucket 1..N-2).
//  * Otherwise the entries are sorted by (ref, refPos, qPos) with a bitonic
//    network across the lanes; a segmented OR-scan over each reference's
//    run of elements gives its hit mask (bit i: an entry at hit index i),
//    whose longest run of set bits is the reference's longest run of
//    consecutive hit indices (repeats of one hit set one bit).  Kept pairs
//    (run >= cut) are compacted in sorted order into k2/v2.
// Results: kSmallDirect (P, ranks and chains set up as the direct path
// does), kSmallSorted (kept pairs sorted in k2/v2, nP set).
template <class TM>
__device__ uint32_t small_path(const Ctx &c, const Ws &ws, uint32_t E, uint32_t nhits,
                               uint32_t nprobe, uint32_t &nP) {
  constexpr uint32_t W = TM::kWorkerLanes;
  static_assert(TM::kSize == W, "one-wavefront teams");
  const uint32_t lane = TM::rank();
  uint32_t ref[2], pos[2], qp[2], hit[2];
  bool ok[2];
#pragma unroll
  for (int j = 0; j < 2; j++) {
    const uint32_t e = 2 * lane + j;
    ok[j] = e < E;
    ref[j] = pos[j] = qp[j] = 0xFFFFFFFFu;
    hit[j] = 63;
    if (ok[j]) {
      const uint64_t k = ws.k1[e];
      ref[j] = (uint32_t)(k >> 32);
      hit[j] = (uint32_t)k;
      pos[j] = ws.v1[e];
      qp[j] = ws.hit_j[hit[j]];
    }
  }
  // (Lane teams skip the direct test: with their few entries the register
  // sort below costs less than the test; 2.49 against 2.55 ms for 1M 150 bp
  // reads on a 5000-reference index.)
  if (W == 64 && !(c.flags & MH_CLASSIFY_NO_DIRECT) && nhits == nprobe && nhits >= 3) {
    const uint32_t N = nhits;
    const uint32_t c0 = ws.hit_base[1];  // entries of the first bucket
    uint32_t nstar = 0, rstar = 0;
    for (uint32_t b = 0; b < c0; b++) {
      const uint32_t r = (uint32_t)(ws.k1[b] >> 32), p0 = ws.v1[b];  // hit 0: qPos 0
      uint32_t cm = 0, con = 0;
#pragma unroll
      for (int j = 0; j < 2; j++) {
        const bool m = ok[j] && ref[j] == r;
        cm += __popcll(TM::wballot(m));
        con += __popcll(TM::wballot(m && pos[j] - qp[j] == p0));
      }
      if (cm == N && con == N) {  // one entry per bucket, all on r's diagonal
        nstar++;
        rstar = r;
      }
    }
    bool direct = nstar == 1;
    if (direct) {
      const uint64_t full = N == 64 ? ~0ull : ((1ull << N) - 1);
      const uint64_t mid = full & ~1ull & ~(1ull << (N - 1));
      for (uint32_t b = ws.hit_base[1]; b < ws.hit_base[2] && direct; b++) {
        const uint32_t r = (uint32_t)(ws.k1[b] >> 32);
        if (r == rstar) continue;
        uint64_t bits = 0;
#pragma unroll
        for (int j = 0; j < 2; j++) bits |= (ok[j] && ref[j] == r) ? (1ull << hit[j]) : 0ull;
        bits = team_or64<TM>(bits);
        if ((bits & mid) == mid) direct = false;
      }
    }
    if (direct) {
      TM::sync();
#pragma unroll
      for (int j = 0; j < 2; j++)
        if (ok[j] && ref[j] == rstar) {  // r*'s pairs in hit order, rank 0
          ws.k2[hit[j]] = (uint64_t)pos[j];
          ws.v2[hit[j]] = qp[j];
        }
      if (lane == 0) {
        ws.rref[0] = rstar;
        ws.t3[0] = N;                // chain maximum of rank 0
        ws.v1[0] = 0;                // chains of rank 0 start at chain 0
        ((uint32_t *)ws.k1)[0] = 0;  // chain 0 spans pairs [0, N-1]
        ((uint32_t *)ws.k1)[1] = N - 1;
        ws.stats[kDirect] += 1;
      }
      TM::sync();
      nP = N;
      return kSmallDirect;
    }
  }
  // Bitonic sort of next_pow2(E) elements by (ref, refPos, qPos), packed
  // into one u64 as ref << 38 | refPos << 6 | hit (hit order is qPos order;
  // the caller guarantees ref < 2^26 and hit < 64, and no two entries are
  // equal).  Padding is ~0, after every element.
  uint64_t P[2];
#pragma unroll
  for (int j = 0; j < 2; j++)
    P[j] = ok[j] ? ((uint64_t)ref[j] << 38) | ((uint64_t)pos[j] << 6) | hit[j] : ~0ull;
  switch (next_pow2(E < 2 ? 2 : E)) {
    case 2: lane_bitonic<2>(P, lane); break;
    case 4: lane_bitonic<4>(P, lane); break;
    case 8: lane_bitonic<8>(P, lane); break;
    case 16: lane_bitonic<16>(P, lane); break;
    case 32: lane_bitonic<32>(P, lane); break;
    default: lane_bitonic<64>(P, lane); break;
  }
  uint64_t key[2];  // ref << 32 | refPos
#pragma unroll
  for (int j = 0; j < 2; j++)
    key[j] = P[j] == ~0ull ? ~0ull : ((P[j] >> 38) << 32) | ((P[j] >> 6) & 0xFFFFFFFFull);
  // Segment heads (a new reference), then each reference's hit mask: a
  // segmented OR over its elements, forward and backward.
  const uint32_t pk_hi = lane_up<W>(1, (uint32_t)(key[1] >> 32));
  bool head[2], valid[2];
  uint64_t bit[2];
#pragma unroll
  for (int j = 0; j < 2; j++) {
    valid[j] = key[j] != ~0ull;
    bit[j] = valid[j] ? (1ull << (P[j] & 63)) : 0ull;
  }
  head[0] = lane == 0 || pk_hi != (uint32_t)(key[0] >> 32);
  head[1] = (key[0] >> 32) != (key[1] >> 32);
  // forward inclusive
  uint64_t f1 = head[1] ? bit[1] : (bit[0] | bit[1]);
  bool ff = head[0] || head[1];
  uint64_t fv = f1;
#pragma unroll
  for (uint32_t d = 1; d < W; d <<= 1) {
    const uint64_t pv = lane_up64<W>(d, fv);
    const uint32_t pf = lane_up<W>(d, ff ? 1u : 0u);
    const bool in = lane >= d;
    fv = (in && !ff) ? (fv | pv) : fv;
    ff = ff || (in && pf);
  }
  // every lane takes part in a permute (a disabled source lane reads as 0)
  const uint64_t fv_up = lane_up64<W>(1, fv);
  const uint64_t before = lane == 0 ? 0ull : fv_up;
  const uint64_t inc0 = head[0] ? bit[0] : (before | bit[0]);
  const uint64_t inc1 = head[1] ? bit[1] : (inc0 | bit[1]);
  // backward inclusive (a segment ends where the next element is a head)
  const uint32_t nh0 = lane_down<W>(1, head[0] ? 1u : 0u);
  const bool tail1 = lane == W - 1 || nh0;
  const bool tail0 = head[1];
  uint64_t bv = tail0 ? bit[0] : (bit[0] | bit[1]);
  bool bf = tail0 || tail1;
#pragma unroll
  for (uint32_t d = 1; d < W; d <<= 1) {
    const uint64_t pv = lane_down64<W>(d, bv);
    const uint32_t pf = lane_down<W>(d, bf ? 1u : 0u);
    const bool in = lane + d < W;
    bv = (in && !bf) ? (bv | pv) : bv;
    bf = bf || (in && pf);
  }
  const uint64_t bv_down = lane_down64<W>(1, bv);
  const uint64_t after = lane == W - 1 ? 0ull : bv_down;
  const uint64_t suf1 = tail1 ? bit[1] : (after | bit[1]);
  const uint64_t suf0 = tail0 ? bit[0] : (suf1 | bit[0]);
  const uint64_t mask[2] = {inc0 | suf0, inc1 | suf1};
  // Filter: runs, M, cut, keep.
  uint32_t run[2], mr = 0;
#pragma unroll
  for (int j = 0; j < 2; j++) {
    uint64_t x = mask[j];
    uint32_t r = 0;
    while (x) {
      x &= x >> 1;
      r++;
    }
    run[j] = r;
    mr = max(mr, r);
  }
  const uint32_t maxrun = TM::max(mr, ws.red);
  const uint32_t cut = maxrun >= 2 ? maxrun - 1 : 1;
  const bool keep0 = valid[0] && run[0] >= cut, keep1 = valid[1] && run[1] >= cut;
  uint32_t tot;
  const uint32_t at = TM::excl_scan((keep0 ? 1u : 0u) + (keep1 ? 1u : 0u), ws.red, tot);
  TM::sync();
  if (keep0) {
    ws.k2[at] = key[0];
    ws.v2[at] = ws.hit_j[P[0] & 63];  // qPos of the hit
  }
  if (keep1) {
    ws.k2[at + (keep0 ? 1 : 0)] = key[1];
    ws.v2[at + (keep0 ? 1 : 0)] = ws.hit_j[P[1] & 63];
  }
  nP = tot;
  TM::sync();
  return kSmallSorted;
}

// Where a query's read lies in the read buffer, when the caller has loaded
// it already (known), else loaded by process_query.
struct QueryLoc {
  uint64_t r0 = 0;
  uint64_t len = 0;
  bool known = false;
};

// Processes query q.  Returns kOverflow (uniformly) when the working set
// exceeds the workspace capacities; nothing is written in that case.
template <class TM>
__device__ __forceinline__ uint32_t process_query(const Ctx &c, const Ws &ws, const Caps &cap,
                                                  uint32_t q, bool last_resort,
                                                  const uint8_t *comp,
                                                  const QueryLoc &loc = QueryLoc()) {
  constexpr uint32_t BS = TM::kSize;
  const uint32_t tid = TM::rank();
  const uint32_t read = q >> 1, strand = q & 1;
  const IndexView &ix = c.ix;
  const int T = ix.T, s = ix.s, k = ix.k;
  const uint64_t r0 = loc.known ? loc.r0 : c.rd.off[read];
  const uint64_t m64 = loc.known ? loc.len : c.rd.off[read + 1] - r0;
  QueryOut res;
  res.highest = -1.0f;
  res.n_top = 0;
  for (int i = 0; i < kTopInline; i++) res.top_ref[i] = 0;
  res.tie_off = 0;
  res.n_cand = 0;
  res.status = 0;
  res.pad = 0;
  if (m64 > c.max_len) {
    res.status = MH_READ_TOO_LONG;
    write_query<TM>(c, q, res);
    return kDone;
  }
  if (m64 > cap.qcap) {
    if (!last_resort) return kOverflow;
    res.status = MH_READ_TOO_LONG;
    write_query<TM>(c, q, res);
    return kDone;
  }
  const uint32_t m = (uint32_t)m64;

  // Stage the strand: uppercase, reverse complement for strand 1
  // (Sequences.cpp:43-52, :795-841).  With probe records the base check is
  // done and staging waits until a seed needs the bases.
  // The read is loaded as the 4-byte aligned dwords that hold its bytes,
  // kStageWords per thread issued together, so that a read of up to
  // 16 * BS bytes takes one memory round trip.  (Each dword holds a byte of
  // the read, so it lies in the same memory page as that byte: the bytes
  // beside the read that it also holds are never used.)  comp is the
  // complement table in LDS, zero outside the nucleotide alphabet.
  constexpr uint32_t kStageWords = 4;
  auto stage = [&]() {
    const uint8_t *first = c.rd.bytes + r0;
    const uintptr_t a0 = (uintptr_t)first & ~(uintptr_t)3;
    const uint32_t off = (uint32_t)((uintptr_t)first - a0);
    const uint32_t nw = (off + m + 3) >> 2;
    int b = 0;
    for (uint32_t wb = 0; wb < nw; wb += kStageWords * BS) {  // team-uniform
      const uint32_t w0 = wb + tid;
      uint32_t word[kStageWords];
#pragma unroll
      for (uint32_t u = 0; u < kStageWords; u++) {
        if (wb + u * BS >= nw) break;
        const uint32_t w = w0 + u * BS;
        word[u] = w < nw ? *(const uint32_t *)(a0 + 4u * w) : 0u;
      }
#pragma unroll
      for (uint32_t u = 0; u < kStageWords; u++) {
        if (wb + u * BS >= nw) break;
        const uint32_t w = w0 + u * BS;
#pragma unroll
        for (uint32_t t = 0; t < 4; t++) {
          const uint32_t x = 4 * w + t - off;  // position in the read (wraps below 0)
          if (w < nw && x < m) {
            const uint8_t ch = upper((uint8_t)(word[u] >> (8 * t)));
            const uint8_t cc = comp[ch];
            b |= (cc == 0);
            if (strand) ws.q[m - 1 - x] = cc;
            else ws.q[x] = ch;
          }
        }
      }
    }
    return TM::any(b);
  };
  const bool use_rec = c.rec != nullptr && m <= c.rec_maxlen;
  const uint64_t *rec = use_rec ? c.rec + (uint64_t)q * c.rec_stride : nullptr;
  bool staged = false, bad;
  // With records: this thread's probe record of the first pass, loaded with
  // the status word (one memory round trip); a record row holds every probe
  // of a read of up to rec_maxlen bases.
  uint64_t rec_first = kRecAbsent;
  if (use_rec) {
    const uint32_t np0 = m > (uint32_t)T ? (m - T + s - 1) / s : 0u;
    if (tid < np0 && tid + 1 < c.rec_stride) rec_first = rec[1 + tid];
    bad = rec[0] == kRecBad;
  } else {
    bad = stage();
    staged = true;
  }
  if (bad) {
    res.status = MH_READ_BADBASE;
    write_query<TM>(c, q, res);
    return kDone;
  }
  if (m <= (uint32_t)T) {  // no well-defined probe (the reference reads past the end)
    write_query<TM>(c, q, res);
    return kDone;
  }
  const uint32_t nprobe = (m - T + s - 1) / s;
  if (nprobe > cap.pcap) {
    if (!last_resort) return kOverflow;
    res.status = MH_READ_CAPACITY;
    write_query<TM>(c, q, res);
    return kDone;
  }

  // Probe every s-th window; compact present hashes into the hit list.
  // With probe records, the lookups were made beforehand by k_probe.
  // elim: the entry count past which the strand leaves this workspace
  // (the LDS tiers' largest counts too, which route an overflowing strand).
  const uint32_t elim = max(cap.ecap, c.lds_ecap_max);
  uint32_t nhits = 0, big = 0, ph = 0;
  for (uint32_t base = 0; base < nprobe; base += BS) {
    const uint32_t p = base + tid;
    uint32_t present = 0, cnt = 0;
    uint64_t lo = 0;
    if (p < nprobe) {
      if (use_rec) {
        const uint64_t r = base == 0 ? rec_first : rec[1 + p];
        present = r != kRecAbsent;
        if (present) {
          lo = r >> 24;
          cnt = (uint32_t)(r & kRecCntMax);
          big |= cnt == kRecCntMax ? 1u : 0u;
        }
      } else {
        const uint32_t h = hash28(ws.q + p * s, k);
        if (h == ix.empty_hash) {
          present = 1;
        } else {
          lo = ix.dir[h];
          const uint64_t hi = ix.dir[h + 1];
          present = hi > lo;
          // A bucket larger than every LDS workspace overflows on its own:
          // clamp, so that the 32-bit entry sums below cannot wrap.
          cnt = (uint32_t)min(hi - lo, (uint64_t)elim + 1);
        }
      }
    }
    uint32_t tot;
    const uint32_t idx = TM::excl_scan(present, ws.red, tot);
    if (present) {
      ws.hit_j[nhits + idx] = p * s;
      ws.hit_lo[nhits + idx] = lo;
      ws.hit_base[nhits + idx] = cnt;
    }
    nhits += tot;
    ph += (lo == 0 && cnt > 0) ? 1u : 0u;  // bucket 0, the only one at offset 0
  }
  // Probes that hit bucket 0 while it holds unstored copies of (0, 0).
  const uint32_t nph = ix.phantom ? TM::sum(ph, ws.red) : 0u;
  TM::sync();
  // Entry total E and the hit_base prefix sums.  Every count is at most
  // elim + 1 <= 2^19 + 1 (probe records saturate at kRecCntMax), so one round
  // adds less than 2^30 to an E that is at most elim: once E exceeds elim
  // the strand overflows (E saturates) and the sums stop before they wrap.
  uint32_t E = 0;
  for (uint32_t base = 0; base < nhits; base += BS) {
    const uint32_t i = base + tid;
    const uint32_t cnt = i < nhits ? min(ws.hit_base[i], elim + 1) : 0;
    uint32_t tot;
    const uint32_t ex = TM::excl_scan(cnt, ws.red, tot);
    if (i < nhits) ws.hit_base[i] = E + ex;
    E += tot;
    if (E > elim) {
      E = 0xFFFFFFFFu;
      break;
    }
  }
  if (tid == 0 && E <= cap.ecap) ws.hit_base[nhits] = E;
  if (use_rec && TM::any(big)) E = 0xFFFFFFFFu;  // a saturated count: overflow
  TM::sync();
  if (E > cap.ecap) {
    if (!last_resort)
      return (c.lds_ecap_max && E > c.lds_ecap_max && q < kBigFlag) ? kOverflowBig : kOverflow;
    res.status = MH_READ_CAPACITY;
    write_query<TM>(c, q, res);
    return kDone;
  }
  if (tid == 0) {
    ws.stats[kProbes] += nprobe;
    ws.stats[kHits] += nhits;
    ws.stats[kEntries] += E + (unsigned long long)nph * ix.phantom;
  }
  if (nhits == 0 || E == 0) {  // Analyser.cpp:1989-1990, :2012-2013
    write_query<TM>(c, q, res);
    return kDone;
  }

  // Gather bucket entries: k1 = ref << 32 | hit index, v1 = ref position.
  // Four entries per thread at a time, their loads issued together: one
  // memory round trip for up to 4 * BS entries.
  constexpr uint32_t kGather = 4;
  for (uint32_t e0 = tid; e0 < E; e0 += kGather * BS) {
    uint2 en[kGather];
    uint32_t hit[kGather];
#pragma unroll
    for (uint32_t u = 0; u < kGather; u++) {
      const uint32_t e = e0 + u * BS;
      hit[u] = 0;
      en[u] = make_uint2(0, 0);
      if (e < E) {
        uint32_t lo = 0, hi = nhits;  // largest i with hit_base[i] <= e
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) >> 1;
          if (ws.hit_base[mid] <= e) lo = mid;
          else hi = mid;
        }
        hit[u] = lo;
        en[u] = ix.ent[ws.hit_lo[lo] + (e - ws.hit_base[lo])];
      }
    }
#pragma unroll
    for (uint32_t u = 0; u < kGather; u++) {
      const uint32_t e = e0 + u * BS;
      if (e < E) {
        ws.k1[e] = ((uint64_t)en[u].x << 32) | hit[u];
        ws.v1[e] = en[u].y;
      }
    }
  }
  TM::sync();

  uint32_t nP = 0, nranks = 0, nchains = 0, G = 0;
  bool direct = false, sorted = false, small = false;
  if constexpr (BS <= 64) {
    if (E <= 2 * BS && E <= 64 && nhits <= 64 && nph == 0 && ix.n_refs <= (1u << 26)) {
      small = true;
      const uint32_t r = small_path<TM>(c, ws, E, nhits, nprobe, nP);
      if (r == kSmallDirect) {
        direct = true;
        nranks = 1;
        nchains = 1;
        G = nP;
      }
      sorted = r == kSmallSorted;
    }
  }
  // (A strand that hit the copies of (0, 0) takes the general path, where
  // their seeds are weighted.)
  // (Tried only while buckets average at most 4 entries: the test scans
  // every bucket per entry of the first, and with large buckets one
  // reference alone in every bucket is rare; the full path below gives the
  // same result either way.)
  if (!small && !(c.flags & MH_CLASSIFY_NO_DIRECT) && nhits == nprobe && nhits >= 3 &&
      nph == 0 && E <= 4 * nhits) {
    // Direct hit.  Every probe found a bucket, and one reference r* has
    // exactly one entry in each bucket, all on one diagonal.  Then r* has the
    // longest run N = nhits, the cut is N - 1, and if no other reference
    // occurs in every bucket 1..N-2 (needed for a run of N - 1), r* is the
    // only kept reference: its N pairs form one chain, G = N, one seed.  The
    // sorts and scans are skipped; the result is the same as the full path.
    const uint32_t N = nhits;
    const uint32_t b0_lo = ws.hit_base[0], b0_hi = ws.hit_base[1];
    uint32_t found = 0, rstar = 0;
    for (uint32_t e = b0_lo + tid; e < b0_hi; e += BS) {
      const uint32_t r = (uint32_t)(ws.k1[e] >> 32);
      const uint32_t diag = ws.v1[e];  // hit 0 has query position 0
      bool ok = true;
      for (uint32_t i = 0; i < N && ok; i++) {
        uint32_t cnt = 0;
        bool on = false;
        for (uint32_t f = ws.hit_base[i]; f < ws.hit_base[i + 1]; f++)
          if ((uint32_t)(ws.k1[f] >> 32) == r) {
            cnt++;
            on = ws.v1[f] - ws.hit_j[i] == diag;
          }
        ok = cnt == 1 && on;
      }
      if (ok) {
        found++;
        rstar = r;
      }
    }
    const uint32_t nfound = TM::sum(found, ws.red);
    if (nfound == 1) {
      const uint32_t rs_ = TM::max(found ? rstar : 0u, ws.red);
      // Any other reference with a run of N - 1 occurs in buckets 1..N-2.
      int other = 0;
      for (uint32_t e = ws.hit_base[1] + tid; e < ws.hit_base[2]; e += BS) {
        const uint32_t r = (uint32_t)(ws.k1[e] >> 32);
        if (r == rs_) continue;
        bool all = true;
        for (uint32_t i = 2; i + 1 < N && all; i++) {
          bool in = false;
          for (uint32_t f = ws.hit_base[i]; f < ws.hit_base[i + 1] && !in; f++)
            in = (uint32_t)(ws.k1[f] >> 32) == r;
          all = in;
        }
        other |= all ? 1 : 0;
      }
      direct = TM::any(other) == 0;
      if (direct) {
        // P = r*'s entries in hit order (already sorted by refPos, qPos).
        if (tid == 0) {
          uint32_t t = 0;
          for (uint32_t i = 0; i < N; i++)
            for (uint32_t f = ws.hit_base[i]; f < ws.hit_base[i + 1]; f++)
              if ((uint32_t)(ws.k1[f] >> 32) == rs_) {
                ws.k2[t] = (uint64_t)ws.v1[f];  // rank 0
                ws.v2[t] = ws.hit_j[i];
                t++;
              }
          ws.rref[0] = rs_;
          ws.t3[0] = N;                      // chain maximum of rank 0
          ws.v1[0] = 0;                      // chains of rank 0 start at chain 0
          ((uint32_t *)ws.k1)[0] = 0;        // chain 0 spans pairs [0, N-1]
          ((uint32_t *)ws.k1)[1] = N - 1;
          ws.stats[kDirect] += 1;
        }
        nP = N;
        nranks = 1;
        nchains = 1;
        G = N;
      }
      TM::sync();
    }
  }

  // The hit-mask filter (below) needs at most 64 hits and a map of more
  // slots than entries; it also tells when no reference occurs in two
  // consecutive hits, which the link pre-check would find at a higher cost.
  const uint32_t mslots = (!direct && !sorted && nhits <= 64 && ix.n_refs <= (1u << 26))
                              ? masked_slots(E, nhits, cap.ecap)
                              : 0u;
  if (!direct && !sorted && !mslots) {
    // A seed needs two pairs of one reference one stride apart on both axes
    // (a chain link).  Links only join pairs of consecutive probes, so when
    // no entry of a hit has its successor in the next probe's bucket, every
    // chain has length 1 whatever the filters keep, and the query gets no
    // seeds (Analyser.cpp:2398-2403): stop before the sorts.
    int link = 0;
    for (uint32_t e = tid; e < E; e += BS) {
      const uint32_t i = (uint32_t)ws.k1[e];
      if (i + 1 < nhits && ws.hit_j[i + 1] == ws.hit_j[i] + (uint32_t)s) {
        const uint32_t r = (uint32_t)(ws.k1[e] >> 32), p = ws.v1[e] + (uint32_t)s;
        for (uint32_t f = ws.hit_base[i + 1]; f < ws.hit_base[i + 2]; f++)
          link |= ((uint32_t)(ws.k1[f] >> 32) == r && ws.v1[f] == p) ? 1 : 0;
      }
    }
    if (!TM::any(link)) {
      write_query<TM>(c, q, res);
      return kDone;
    }
  }

  if (!direct) {
    // Successive-occurrence filter (Analyser.cpp:2004-2197): the longest run
    // of consecutive hit indices of each reference, M the longest of all
    // (counted only from 2, :2067-2072), cut = max(1, M - 1) (:2099); every
    // pair of a reference whose run reaches the cut is kept.  Result: kept
    // pairs in k2/v2 as (ref << 32 | refPos, qPos), unsorted.
    // (Done in registers by small_path for small working sets.)
    if (!sorted) {
      if (mslots) {
        const uint32_t maxrun = nhits <= 32 ? filter_masked<TM, uint32_t>(ws, E, mslots, nP)
                                            : filter_masked<TM, unsigned long long>(ws, E, mslots, nP);
        if (maxrun < 2) {  // no reference in two consecutive hits: no link, no seeds
          write_query<TM>(c, q, res);
          return kDone;
        }
        // Sort by (reference, refPos, qPos) (Analyser.cpp:1024-1059): the
        // packed keys order so (hit order is qPos order); then unpack.
        if (has_stage<TM>(ws))
          staged_sort_keys<TM>(ws.k2, nP, ws.stage, ws.stage_cap);
        else
          bitonic_sort_keys<TM>(ws.k2, nP);
        for (uint32_t t = tid; t < nP; t += BS) {
          const uint64_t pk = ws.k2[t];
          ws.k2[t] = ((pk >> 38) << 32) | ((pk >> 6) & 0xFFFFFFFFull);
          ws.v2[t] = ws.hit_j[pk & 63];
        }
        TM::sync();
      } else {
        const bool hashed = nhits < (1u << 12) && ix.n_refs < (1u << 20);
        if (hashed)
          filter_hashed<TM>(ws, E, cap.ecap, nP);
        else
          filter_sorted<TM>(ws, E, nP);
        // Sort by (reference, refPos, qPos) (Analyser.cpp:1024-1059).
        TM::sync();
        if (hashed && has_stage<TM>(ws)) {
          // filter_hashed packed the pairs as ref << 44 | refPos << 12 | hit
          staged_sort_keys<TM>(ws.k2, nP, ws.stage, ws.stage_cap);
          for (uint32_t t = tid; t < nP; t += BS) {
            const uint64_t pk = ws.k2[t];
            ws.k2[t] = ((pk >> 44) << 32) | ((pk >> 12) & 0xFFFFFFFFull);
            ws.v2[t] = ws.hit_j[pk & 0xFFF];
          }
          TM::sync();
        } else {
          bitonic_sort<TM>(ws.k2, ws.v2, nP);
        }
      }
    }
    // Candidate references in ascending order: rank = index among the
    // distinct references; k2 becomes rank << 32 | refPos.  Chains:
    // neighbours link when both coordinates step by s (Analyser.cpp:2351-2353,
    // u32 arithmetic); cfl[2c] and cfl[2c + 1] are the first and last pair of
    // chain c, v1[rank] the rank's first chain, t3[rank] zeroed.
    uint32_t *cfl = (uint32_t *)ws.k1;  // chain first/last element
    if constexpr (TM::kPerWave == 1) {
      // Whole-workgroup teams: each thread takes a contiguous run of `per`
      // pairs, so that one team scan numbers them all; the scan's barriers
      // order every read of a neighbour's pair before the pairs are
      // rewritten (a one-wavefront team syncs instead).
      const uint32_t per = (nP + BS - 1) / BS;
      const uint32_t lo = min(nP, tid * per), hi = min(nP, lo + per);
      {
        uint32_t cnt = 0;
        uint32_t prev = lo > 0 ? (uint32_t)(ws.k2[lo - 1] >> 32) : 0u;
        for (uint32_t t = lo; t < hi; t++) {
          const uint32_t r = (uint32_t)(ws.k2[t] >> 32);
          const uint32_t st = (t == 0 || prev != r) ? 1u : 0u;
          ws.t1[t] = st;
          cnt += st;
          prev = r;
        }
        uint32_t tot;
        uint32_t rank = TM::excl_scan(cnt, ws.red, tot) - 1;  // wraps to 0 at a run's start
        nranks = tot;
        if constexpr (BS <= 64) TM::sync();
        for (uint32_t t = lo; t < hi; t++) {
          const uint64_t key = ws.k2[t];
          if (ws.t1[t]) {
            rank++;
            ws.rref[rank] = (uint32_t)(key >> 32);
            ws.t3[rank] = 0;
          }
          ws.k2[t] = ((uint64_t)rank << 32) | (uint32_t)key;
        }
      }
      TM::sync();
      {  // chains
        uint32_t cnt = 0;
        uint64_t pa = lo > 0 ? ws.k2[lo - 1] : 0ull;
        uint32_t pq = lo > 0 ? ws.v2[lo - 1] : 0u;
        for (uint32_t t = lo; t < hi; t++) {
          const uint64_t b = ws.k2[t];
          const uint32_t bq = ws.v2[t];
          const bool link = t > 0 && (pa >> 32) == (b >> 32) &&
                            (uint32_t)pa == (uint32_t)b - (uint32_t)s && pq == bq - (uint32_t)s;
          const uint32_t st = link ? 0u : 1u;
          ws.t1[t] = st;
          cnt += st;
          pa = b;
          pq = bq;
        }
        uint32_t tot;
        uint32_t cid = TM::excl_scan(cnt, ws.red, tot) - 1;
        nchains = tot;
        if constexpr (BS <= 64) TM::sync();  // the next run's first flag, written above
        for (uint32_t t = lo; t < hi; t++) {
          const uint32_t st = ws.t1[t];
          cid += st;
          if (st) cfl[2 * cid] = t;
          if (t + 1 == nP || ws.t1[t + 1]) cfl[2 * cid + 1] = t;
          const uint32_t rank = (uint32_t)(ws.k2[t] >> 32);
          if (t == 0 || (uint32_t)(ws.k2[t - 1] >> 32) != rank) ws.v1[rank] = cid;
        }
      }
      TM::sync();
    } else {
      // Lane teams: pair-strided passes, one team scan per kSize pairs
      // (fewer registers live than the contiguous runs above).
      for (uint32_t t = tid; t < nP; t += BS)
        ws.t1[t] = (t == 0 || (ws.k2[t - 1] >> 32) != (ws.k2[t] >> 32)) ? 1u : 0u;
      TM::sync();
      uint32_t nrank = 0;
      for (uint32_t base = 0; base < nP; base += BS) {
        const uint32_t t = base + tid;
        const uint32_t st = t < nP ? ws.t1[t] : 0u;
        uint32_t tot;
        const uint32_t ex = TM::excl_scan(st, ws.red, tot);
        if (t < nP) {
          const uint64_t key = ws.k2[t];
          const uint32_t rank = nrank + ex + st - 1;
          if (st) ws.rref[rank] = (uint32_t)(key >> 32);
          ws.k2[t] = ((uint64_t)rank << 32) | (uint32_t)key;
        }
        nrank += tot;
      }
      nranks = nrank;
      TM::sync();
      for (uint32_t t = tid; t < nP; t += BS) {  // chains
        uint32_t st = 1;
        if (t > 0) {
          const uint64_t a = ws.k2[t - 1], b = ws.k2[t];
          st = ((a >> 32) == (b >> 32) && (uint32_t)a == (uint32_t)b - (uint32_t)s &&
                ws.v2[t - 1] == ws.v2[t] - (uint32_t)s)
                   ? 0u
                   : 1u;
        }
        ws.t1[t] = st;
      }
      TM::sync();
      uint32_t nch = 0;
      for (uint32_t base = 0; base < nP; base += BS) {
        const uint32_t t = base + tid;
        const uint32_t st = t < nP ? ws.t1[t] : 0u;
        uint32_t tot;
        const uint32_t ex = TM::excl_scan(st, ws.red, tot);
        if (t < nP) ws.t2[t] = nch + ex + st - 1;  // chain id
        nch += tot;
      }
      nchains = nch;
      TM::sync();
      for (uint32_t t = tid; t < nP; t += BS) {
        const uint32_t cid = ws.t2[t];
        if (ws.t1[t]) cfl[2 * cid] = t;
        if (t + 1 == nP || ws.t1[t + 1]) cfl[2 * cid + 1] = t;
        const uint32_t rank = (uint32_t)(ws.k2[t] >> 32);
        if (t == 0 || (uint32_t)(ws.k2[t - 1] >> 32) != rank) ws.v1[rank] = cid;
      }
      for (uint32_t r = tid; r < nranks; r += BS) ws.t3[r] = 0;
      TM::sync();
    }
    for (uint32_t cid = tid; cid < nchains; cid += BS) {
      const uint32_t f = cfl[2 * cid], l = cfl[2 * cid + 1];
      atomicMax(&ws.t3[(uint32_t)(ws.k2[f] >> 32)], l - f + 1);
    }
    TM::sync();
    uint32_t gm = 0;
    for (uint32_t r = tid; r < nranks; r += BS) gm = max(gm, ws.t3[r]);
    G = TM::max(gm, ws.red);
    TM::sync();
  }

  if (G < 2) {  // Analyser.cpp:2398-2403: no seeds, every candidate scores -2
    write_query<TM>(c, q, res);
    return kDone;
  }

  if (!staged) stage();
  // Score every reference whose longest chain equals G (Analyser.cpp:2429-2436).
  const bool ultrafast = (c.flags & MH_CLASSIFY_ULTRAFAST) != 0;
  const uint32_t *cfl = (const uint32_t *)ws.k1;
  float *rel_of = (float *)ws.t2;
  uint32_t *valid_of = ws.t1;
  unsigned long long seeds_total = 0;
  uint32_t seeded = 0;
  bool multi = false;
  if constexpr (TM::kPerWave == 1 && BS >= 128) {  // slots need 128 threads or more
    if (cap.nslot != 0 && nranks >= 2) {
      multi = true;
      uint32_t lut_bad = 0;
      score_multi<TM>(c, ws, cap, m, nranks, nchains, G, nhits, nph, ultrafast, read, strand,
                      seeds_total, seeded, lut_bad);
      if (lut_bad) res.status |= MH_READ_CAPACITY;
    }
  }
  for (uint32_t r = 0; r < nranks && !multi; r++) {
    if (ws.t3[r] != G) {
      if (tid == 0) valid_of[r] = 0;
      continue;
    }
    const uint32_t c0 = ws.v1[r];
    const uint32_t c1 = (r + 1 < nranks) ? ws.v1[r + 1] : nchains;
    const uint32_t R = ws.rref[r];
    const uint64_t ro = ix.ref_off[R];
    const uint64_t ref_len = ix.ref_off[R + 1] - ro;
    const uint8_t *rb = ix.bytes + ro;
    // Reference 0 also gets, per probe j that hit bucket 0, ix.phantom more
    // one-pair seeds (0, j, s, j + s): the unstored copies of (0, 0, j), each
    // a chain of its own (Analyser.cpp:2421-2481).
    const uint32_t nvirt = R == 0 ? nph : 0u;
    auto seed = [&](uint32_t i, uint32_t &sref, uint32_t &sq, uint32_t &lref, uint32_t &lq,
                    uint32_t &wt) {
      if (i < c1 - c0) {
        const uint32_t cid = c0 + i;
        const uint32_t f = cfl[2 * cid], l = cfl[2 * cid + 1];
        sref = (uint32_t)ws.k2[f];
        sq = ws.v2[f];
        lref = (uint32_t)ws.k2[l] + (uint32_t)s;  // Analyser.cpp:2465-2467
        lq = ws.v2[l] + (uint32_t)s;
        wt = 1;
        return;
      }
      uint32_t v = i - (c1 - c0), j = 0;
      for (uint32_t t = 0; t < nhits; t++)
        if (ws.hit_lo[t] == 0 && ws.hit_base[t + 1] > ws.hit_base[t]) {
          if (v == 0) {
            j = ws.hit_j[t];
            break;
          }
          v--;
        }
      sref = 0;
      sq = j;
      lref = (uint32_t)s;
      lq = j + (uint32_t)s;
      wt = ix.phantom;
    };
    uint64_t sum_r, sum_c;
    uint32_t hits;
    score_ref<TM>(ws.q, m, rb, ref_len, ix.bytes, c1 - c0 + nvirt, T, ultrafast, ws, seed, sum_r,
                  sum_c, hits);
    const float score = matrix_score(sum_r, sum_c, hits, m);
    float rel = -1.0f;
    uint32_t valid = 0, lut_bad = 0;
    if (score > -1.0f) {  // MatrixCPU.cpp:356
      if (rel_score(c.lut, m, ref_len, score, rel)) valid = 1;
      else lut_bad = 1;
    }
    seeds_total += (c1 - c0) + (unsigned long long)nvirt * ix.phantom;
    seeded++;
    if (tid == 0) {
      rel_of[r] = rel;
      valid_of[r] = valid;
      if (lut_bad) res.status |= MH_READ_CAPACITY;
      if (c.out.cands) {
        const unsigned long long slot = atomicAdd(&c.out.ctr->cand_used, 1ull);
        if (slot < c.out.cand_cap) {
          mh_cand cd;
          cd.read = read;
          cd.strand = strand;
          cd.ref = R;
          cd.hits = hits;
          cd.sum_r = sum_r;
          cd.sum_c = sum_c;
          cd.score = score;
          cd.rel = rel;
          cd.n_seeds = (uint32_t)((c1 - c0) + (unsigned long long)nvirt * ix.phantom);
          cd.pad = 0;
          c.out.cands[slot] = cd;
        }
      }
    }
  }
  TM::sync();
  if (tid == 0) {
    ws.stats[kSeeds] += seeds_total;
    ws.stats[kCands] += seeded;
  }

  // Best hit: first strict maximum of the relative scores (all >= 0), and
  // the references tied with it, in ascending reference order.
  uint32_t nv = 0, hb = 0;
  for (uint32_t r = tid; r < nranks; r += BS)
    if (valid_of[r]) {
      nv++;
      hb = max(hb, __float_as_uint(rel_of[r]));
    }
  const uint32_t n_cand = TM::sum(nv, ws.red);
  const uint32_t best = TM::max(hb, ws.red);
  res.n_cand = n_cand;
  if (n_cand > 0) {
    res.highest = __uint_as_float(best);
    uint32_t ntop = 0;
    for (uint32_t base = 0; base < nranks; base += BS) {
      const uint32_t r = base + tid;
      const uint32_t tie = (r < nranks && valid_of[r] && __float_as_uint(rel_of[r]) == best);
      uint32_t tot;
      const uint32_t ex = TM::excl_scan(tie, ws.red, tot);
      if (tie) ws.t3[ntop + ex] = ws.rref[r];  // t3 is free once seeding is done
      ntop += tot;
    }
    TM::sync();
    res.n_top = ntop;
    if (ntop <= (uint32_t)kTopInline) {
      // constant indices keep res in registers
#pragma unroll
      for (int i = 0; i < kTopInline; i++) res.top_ref[i] = (uint32_t)i < ntop ? ws.t3[i] : 0u;
    } else {
      if (tid == 0) {
        const unsigned long long off = atomicAdd(&c.out.ctr->tie_used, (unsigned long long)ntop);
        ws.red[kBcast] = (uint32_t)off;
        ws.red[kBcast + 1] = (off + ntop <= c.out.tie_cap) ? 1u : 0u;
      }
      TM::sync();
      res.tie_off = ws.red[kBcast];
      if (ws.red[kBcast + 1])
        for (uint32_t i = tid; i < ntop; i += BS) c.out.ties[res.tie_off + i] = ws.t3[i];
    }
  }
  write_query<TM>(c, q, res);
  TM::sync();
  return kDone;
}

template <class TM>
__device__ __forceinline__ void stats_begin(const Ws &ws) {
  if (TM::rank() < kNumStats) ws.stats[TM::rank()] = 0;
  TM::sync();
}

// Adds the team's statistics to slot `slot` of the device counters and to
// the launch's row.
template <class TM>
__device__ __forceinline__ void stats_flush(const Ws &ws, DevCounters *ctr, uint32_t slot,
                                            uint32_t kid) {
  TM::sync();
  if (TM::rank() < kNumStats && ws.stats[TM::rank()]) {
    atomicAdd(&ctr->stats[slot & (kStatSlots - 1)][TM::rank()], ws.stats[TM::rank()]);
    atomicAdd(&ctr->kstats[kid < kKernelRows ? kid : 0][slot & (kKernelSlots - 1)][TM::rank()],
              ws.stats[TM::rank()]);
  }
}

template <class TM>
__device__ __forceinline__ void count_done(const Ws &ws, uint32_t r) {
  if (r == kDone && TM::rank() == 0) ws.stats[kStrands] += 1;
}

template <class TM>
__device__ __forceinline__ void pass_on(const Ctx &c, uint32_t q) {
  if (TM::rank() == 0) {
    const unsigned long long slot = atomicAdd(c.out.overflow_cnt, 1ull);
    if (slot < c.out.overflow_cap) c.out.overflow[slot] = q;
  }
}

// Handles process_query's result for q: overflowing strands go to the next
// tier, flagged when no LDS tier can hold their entries.
template <class TM>
__device__ __forceinline__ void after_query(const Ctx &c, uint32_t q, uint32_t r) {
  if (r == kOverflow) pass_on<TM>(c, q);
  else if (r == kOverflowBig) pass_on<TM>(c, q | kBigFlag);
}

// Most queries a team takes from its queue at a time.
constexpr uint32_t kQueueChunk = 16;
// No query to pass on (query ids are below 2^31 - 2: fewer than 2^30 reads).
constexpr uint32_t kNoPass = 0xFFFFFFFFu;

// Length-tile kernel.  Each team has its own workspace in LDS.  Plain form:
// one query per team; batched form: a persistent loop over a queue of
// queries, taken a chunk per workgroup at a time: up to kQueueChunk per team
// while the queue holds at least four chunks per team, one per team for
// short queues (the overflow tiers).  Queries that do not fit are passed on
// to the next tier.
// Every tile kernel is held to 128 VGPRs (4 wavefronts per SIMD): their
// throughput follows the number of strands in flight (the 64-thread teams
// of 150 bp reads on an 8 Gbp index: 81 ms at 128 VGPRs against 91 ms at
// 136, which leaves 3 wavefronts per SIMD).
template <class TM>
__global__ __launch_bounds__(TM::kSize * TM::kPerWave)
__attribute__((amdgpu_waves_per_eu(4, 8))) void k_classify(
    Ctx c, Caps cap, const uint32_t *qlist, const unsigned long long *qstart, uint32_t nq,
    const unsigned long long *nq_dev, unsigned long long *head) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr uint32_t G = TM::kPerWave;
  const uint32_t team = TM::index();
  const Ws ws = ws_carve(smem + team * ws_stride_lds(cap), cap);
  __shared__ uint32_t s_base;
  __shared__ uint8_t s_comp[256];
  // lane teams: a chunk's queries, read offsets and lengths
  constexpr uint32_t kLoc = G > 1 ? kQueueChunk * G : 1;
  __shared__ uint32_t s_q[kLoc], s_len[kLoc];
  __shared__ uint64_t s_r0[kLoc];
  for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) s_comp[i] = kComp[i];
  __syncthreads();
  if (qlist && qstart) qlist += *qstart;
  stats_begin<TM>(ws);
  const uint32_t slot = blockIdx.x * G + team;
  if (head == nullptr) {
    if (slot < nq) {
      const uint32_t q = qlist ? qlist[slot] : slot;
      const uint32_t r = process_query<TM>(c, ws, cap, q, false, s_comp);
      count_done<TM>(ws, r);
      after_query<TM>(c, q, r);
    }
    stats_flush<TM>(ws, c.out.ctr, slot, c.kid);
    return;
  }
  const unsigned long long n = nq_dev ? *nq_dev : nq;
  const unsigned long long per_team =
      n / ((unsigned long long)gridDim.x * G * 4ull * kQueueChunk) ? kQueueChunk
      : max(1ull, n / ((unsigned long long)gridDim.x * G * 4ull));
  const uint32_t chunk = (uint32_t)per_team * G;
  for (;;) {
    if (n == 0) break;  // an empty bucket or tier
    __syncthreads();
    if (threadIdx.x == 0) {
      const unsigned long long i = atomicAdd(head, (unsigned long long)chunk);
      s_base = i < n ? (uint32_t)i : 0xFFFFFFFFu;
    }
    __syncthreads();
    const uint32_t base = s_base;
    if (base == 0xFFFFFFFFu) break;
    const uint32_t end = (uint32_t)min((unsigned long long)base + chunk, n);
    // The chunk's queries and where their reads lie are loaded at once, so
    // that a query does not wait for these two dependent loads of its own:
    // whole-workgroup teams hold them in registers, lane k of every
    // wavefront the chunk's k-th query (chunk <= kQueueChunk <= 64); lane
    // teams in LDS.
    uint32_t c_q = kNoPass, c_len = 0;
    uint64_t c_r0 = 0;
    if constexpr (G == 1) {
      const uint32_t k = threadIdx.x & 63;
      if (base + k < end) {
        c_q = qlist ? qlist[base + k] : base + k;
        if (!(c_q & kBigFlag)) {
          c_r0 = c.rd.off[c_q >> 1];
          c_len = (uint32_t)min(c.rd.off[(c_q >> 1) + 1] - c_r0, (uint64_t)0xFFFFFFFFu);
        }
      }
    } else {
      for (uint32_t k = threadIdx.x; k < chunk; k += blockDim.x) {
        const uint32_t i = base + k;
        uint32_t q = 0;
        uint64_t r0 = 0, len = 0;
        if (i < end) {
          q = qlist ? qlist[i] : i;
          if (!(q & kBigFlag)) {
            r0 = c.rd.off[q >> 1];
            len = c.rd.off[(q >> 1) + 1] - r0;
          }
        }
        s_q[k] = q;
        s_r0[k] = r0;
        s_len[k] = (uint32_t)min(len, (uint64_t)0xFFFFFFFFu);
      }
      __syncthreads();
    }
    for (uint32_t i = base + team; i < end; i += G) {
      QueryLoc loc;
      uint32_t q;
      if constexpr (G == 1) {
        const uint32_t k = i - base;  // uniform
        q = readlane(c_q, k);
        loc.r0 = ((uint64_t)readlane((uint32_t)(c_r0 >> 32), k) << 32) | readlane((uint32_t)c_r0, k);
        loc.len = readlane(c_len, k);
      } else {
        q = s_q[i - base];
        loc.r0 = s_r0[i - base];
        loc.len = s_len[i - base];
      }
      loc.known = true;
      if (q & kBigFlag) {  // more entries than an LDS tier holds: straight on
        if constexpr (G > 1) pass_on<TM>(c, q);  // workgroup teams: lane k keeps it
        continue;
      }
      const uint32_t r = process_query<TM>(c, ws, cap, q, false, s_comp, loc);
      count_done<TM>(ws, r);
      if constexpr (G == 1) {
        // lane k of every wavefront now holds what query k passes on
        if ((threadIdx.x & 63) == i - base)
          c_q = r == kOverflow ? q : r == kOverflowBig ? q | kBigFlag : kNoPass;
      } else {
        after_query<TM>(c, q, r);
      }
    }
    if constexpr (G == 1) {
      // The chunk's queries for the next tier, appended with one atomic per
      // chunk: one counter takes every overflow of the launch, and a
      // returning atomic per query there made the queries wait in turn.
      if (threadIdx.x < 64) {
        const bool pass = c_q != kNoPass;
        const uint64_t bal = __ballot(pass);
        if (bal) {
          unsigned long long at = 0;
          if (threadIdx.x == 0)
            at = atomicAdd(c.out.overflow_cnt, (unsigned long long)__popcll(bal));
          at = ((unsigned long long)readlane((uint32_t)(at >> 32), 0) << 32) |
               readlane((uint32_t)at, 0);
          const uint32_t before =
              __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                        __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
          if (pass && at + before < c.out.overflow_cap) c.out.overflow[at + before] = c_q;
        }
      }
    }
  }
  stats_flush<TM>(ws, c.out.ctr, slot, c.kid);
}

// Queries that overflowed the LDS tiers, with the workspace in HBM: one
// workspace per workgroup, queries taken one at a time from `list`.  The
// last resort flags what it cannot hold; the tier before it passes such
// queries on.
constexpr uint32_t kStageKeys = 4096;

template <int BS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(4))) void k_classify_hbm(Ctx c, Caps cap, unsigned char *scratch,
                                                     size_t ws_stride, const uint32_t *list,
                                                     const unsigned long long *count,
                                                     unsigned long long *head, int last_resort) {
  using TM = BlockTeam<BS>;
  // 32 KiB of LDS per workgroup for the sorts: four resident 256-thread
  // workgroups per CU fit in its 160 KiB
  __shared__ uint64_t s_stage[kStageKeys];
  Ws ws = ws_carve(scratch + blockIdx.x * ws_stride, cap);
  ws.stage = s_stage;
  ws.stage_cap = kStageKeys;
  __shared__ uint32_t slot;
  __shared__ uint8_t s_comp[256];
  for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) s_comp[i] = kComp[i];
  __syncthreads();
  stats_begin<TM>(ws);
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) {
      const unsigned long long n = *count;
      const unsigned long long i = n ? atomicAdd(head, 1ull) : 0ull;
      slot = (i < n && i < c.out.overflow_cap) ? (uint32_t)i : 0xFFFFFFFFu;
    }
    __syncthreads();
    if (slot == 0xFFFFFFFFu) break;
    const uint32_t q = list[slot] & ~kBigFlag;
    const uint32_t r = process_query<TM>(c, ws, cap, q, last_resort != 0, s_comp);
    count_done<TM>(ws, r);
    if (r != kDone) pass_on<TM>(c, q);
  }
  stats_flush<TM>(ws, c.out.ctr, blockIdx.x, c.kid);
}

// Probe records for the strands of the shortest length bucket.  An 8-lane
// team per strand, 32 strands per 256-thread workgroup: the team copies its
// read into an LDS row with dword loads, checks the bases against the
// alphabet and writes the strand (uppercase, reverse complement for strand
// 1) into a second row; then each lane hashes one probe window and looks up
// its bucket.  All lookups of a workgroup are in flight together, so the
// pass runs at the rate HBM serves random 16-byte reads.  Strands longer
// than `maxlen` are left to the tile kernel.
constexpr uint32_t kProbeLanes = 8;
constexpr uint32_t kProbeTeams = 32;
__global__ __launch_bounds__(kProbeLanes * kProbeTeams) void k_probe(
    IndexView ix, Reads rd, const uint32_t *binned, const unsigned long long *start,
    const unsigned long long *count, uint32_t stride, uint32_t row_words, uint32_t maxlen,
    uint64_t *rec) {
  extern __shared__ __align__(16) unsigned char smem[];
  uint8_t *comp = smem;  // 256-byte complement table
  const uint32_t team = threadIdx.x / kProbeLanes, lane = threadIdx.x % kProbeLanes;
  uint32_t *raw = (uint32_t *)(smem + 256) + team * 2 * row_words;  // the read's dwords
  uint8_t *str = (uint8_t *)(raw + row_words);                       // the strand's bases
  for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) comp[i] = kComp[i];
  const uint32_t *list = binned + *start;
  const uint64_t n = *count;
  const uint8_t *buf_end = rd.bytes + rd.off[rd.n];
  const int T = ix.T, s = ix.s, k = ix.k;
  __syncthreads();
  for (uint64_t base = (uint64_t)blockIdx.x * kProbeTeams; base < n;
       base += (uint64_t)gridDim.x * kProbeTeams) {
    const uint64_t i = base + team;
    if (i >= n) continue;  // team-uniform; no workgroup barrier below
    const uint32_t q = list[i];
    const uint64_t r0 = rd.off[q >> 1];
    const uint64_t m = rd.off[(q >> 1) + 1] - r0;
    if (m > maxlen) continue;
    const uint32_t mm = (uint32_t)m, strand = q & 1;
    // dwords at 4-byte aligned addresses covering the read
    const uint8_t *first = rd.bytes + r0;
    const uintptr_t a0 = (uintptr_t)first & ~(uintptr_t)3;
    const uint32_t off = (uint32_t)((uintptr_t)first - a0);
    const uint32_t nw = (off + mm + 3) >> 2;
    for (uint32_t w = lane; w < nw; w += kProbeLanes) {
      const uintptr_t at = a0 + 4u * w;
      uint32_t word = 0;
      if (at >= (uintptr_t)first && at + 4 <= (uintptr_t)buf_end) {
        word = *(const uint32_t *)at;
      } else {  // the first or last dword: only bytes of the buffer
        for (uint32_t t = 0; t < 4; t++)
          if (at + t >= (uintptr_t)rd.bytes && at + t < (uintptr_t)buf_end)
            word |= (uint32_t)(*(const uint8_t *)(at + t)) << (8 * t);
      }
      raw[w] = word;
    }
    LaneTeam<kProbeLanes>::sync();
    const uint8_t *b = (const uint8_t *)raw + off;
    int bad = 0;
    for (uint32_t x = lane; x < mm; x += kProbeLanes) {
      const uint8_t ch = upper(b[strand ? mm - 1 - x : x]);
      const uint8_t cc = comp[ch];  // zero outside the nucleotide alphabet
      bad |= cc == 0;
      str[x] = strand ? cc : ch;
    }
    uint64_t *out = rec + (uint64_t)q * stride;
    bad = LaneTeam<kProbeLanes>::any(bad);
    if (lane == 0) out[0] = bad ? kRecBad : 0ull;
    if (bad || mm <= (uint32_t)T) continue;
    const uint32_t nprobe = (mm - T + s - 1) / s;
    for (uint32_t p = lane; p < nprobe; p += kProbeLanes) {
      const uint32_t h = hash28(str + p * s, k);
      uint64_t r = 0;  // the empty bucket: present, no entries
      if (h != ix.empty_hash) {
        const uint64_t lo = ix.dir[h], hi = ix.dir[h + 1];
        r = hi > lo ? (lo << 24) | (uint64_t)min(hi - lo, (uint64_t)kRecCntMax) : kRecAbsent;
      }
      out[1 + p] = r;
    }
    LaneTeam<kProbeLanes>::sync();  // rows are reused by the team's next strand
  }
}

// Strand merge (Analyser.cpp:592-611): the reverse strand wins only when its
// HighestScore is strictly higher; a read is reported only with candidates
// and a non-zero HighestScore (Analyser.cpp:388).  Pass 1 writes the result
// and the number of reported ids; an exclusive scan gives each read its
// place in the top-tie pool; pass 2 copies the ids.
__global__ void k_merge_count(const QueryOut *qo, uint32_t n_reads, mh_read_result *out,
                              uint32_t *cnt) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n_reads;
       r += gridDim.x * blockDim.x) {
    const QueryOut &f = qo[2 * r], &b = qo[2 * r + 1];
    const bool rc = f.highest < b.highest;
    const QueryOut &w = rc ? b : f;
    mh_read_result o;
    o.highest = w.highest;
    o.strand = rc ? 1u : 0u;
    o.n_cand = w.n_cand;
    const bool report = w.n_cand > 0 && w.highest != 0.0f;
    o.n_top = report ? w.n_top : 0u;
    o.top_off = 0;
    o.status = f.status | b.status;
    out[r] = o;
    cnt[r] = o.n_top;
  }
}

// One read per lane; a tie set of up to kTopInline references is copied by
// its lane, a longer one (from the tie pool) by the whole wavefront in turn,
// in consecutive words.  The grid is a whole number of wavefronts.
__global__ void k_merge_write(const QueryOut *qo, uint32_t n_reads, const uint32_t *off,
                              const uint32_t *ties, uint64_t tie_cap, mh_read_result *out,
                              uint32_t *top, uint64_t top_cap, DevCounters *ctr) {
  const uint32_t lane = threadIdx.x & 63;
  for (uint32_t base = (blockIdx.x * blockDim.x + threadIdx.x) & ~63u; base < n_reads;
       base += gridDim.x * blockDim.x) {
    const uint32_t r = base + lane;
    uint32_t n = 0, o = 0, src = 0;  // a pooled set: ties[src, src + n)
    if (r < n_reads) {
      n = out[r].n_top;
      o = off[r];
      out[r].top_off = o;
      if (n && (uint64_t)o + n <= top_cap) {
        const QueryOut &w = qo[2 * r + out[r].strand];
        if (n <= (uint32_t)kTopInline) {
          for (uint32_t i = 0; i < n; i++) top[o + i] = w.top_ref[i];
          n = 0;
        } else if ((uint64_t)w.tie_off + n <= tie_cap) {
          src = w.tie_off;
        } else {
          n = 0;
        }
      } else {
        n = 0;
      }
    }
    for (uint64_t pend = __ballot(n != 0); pend; pend &= pend - 1) {
      const uint32_t j = (uint32_t)__ffsll((long long)pend) - 1;
      const uint32_t nj = readlane(n, j), oj = readlane(o, j), sj = readlane(src, j);
      for (uint32_t i = lane; i < nj; i += 64) top[oj + i] = ties[sj + i];
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) ctr->top_used = off[n_reads];
}

__global__ void k_zero_u32(uint32_t *p, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    p[i] = 0;
}

__global__ void k_zero(unsigned long long *p, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    p[i] = 0;
}

// Length buckets for the batched form: bucket of each read by its length.
// One read per thread; counts are gathered in LDS and added once per block.
struct Lens {
  uint32_t v[8];
};

__device__ __forceinline__ uint32_t bucket_of(const Reads &rd, uint32_t r, const Lens &max_len,
                                              uint32_t nb) {
  const uint64_t m = rd.off[r + 1] - rd.off[r];
  uint32_t b = 0;
  while (b + 1 < nb && m > max_len.v[b]) b++;
  return b;
}

// Binning: each workgroup takes a contiguous chunk of reads; a wavefront
// counts its reads per bucket with ballots, one LDS atomic per bucket and
// wavefront, and the workgroup makes one global atomic per bucket.
constexpr uint32_t kBinBlocks = 512;

__device__ __forceinline__ void bin_chunk(const Reads &rd, uint32_t &lo, uint32_t &hi) {
  const uint32_t per = (rd.n + gridDim.x - 1) / gridDim.x;
  lo = min(rd.n, blockIdx.x * per);
  hi = min(rd.n, lo + per);
}

__global__ void k_bin_count(Reads rd, Lens max_len, uint32_t nb, unsigned long long *cnt) {
  __shared__ unsigned int loc[8];
  if (threadIdx.x < 8) loc[threadIdx.x] = 0;
  __syncthreads();
  uint32_t lo, hi;
  bin_chunk(rd, lo, hi);
  const uint32_t lane = threadIdx.x & 63;
  for (uint32_t r0 = lo; r0 < hi; r0 += blockDim.x) {
    const uint32_t r = r0 + threadIdx.x;
    const uint32_t b = r < hi ? bucket_of(rd, r, max_len, nb) : 0xFFu;
    for (uint32_t k = 0; k < nb; k++) {
      const unsigned long long m = __ballot(b == k);
      if (m && lane == (uint32_t)__ffsll((long long)m) - 1) atomicAdd(&loc[k], 2u * __popcll(m));
    }
  }
  __syncthreads();
  if (threadIdx.x < nb && loc[threadIdx.x])
    atomicAdd(&cnt[threadIdx.x], (unsigned long long)loc[threadIdx.x]);
}

__global__ void k_bin_scatter(Reads rd, Lens max_len, uint32_t nb,
                              const unsigned long long *start, unsigned long long *cursor,
                              uint32_t *list) {
  __shared__ unsigned int loc[8];
  __shared__ unsigned long long base[8];
  if (threadIdx.x < 8) loc[threadIdx.x] = 0;
  __syncthreads();
  uint32_t lo, hi;
  bin_chunk(rd, lo, hi);
  const uint32_t lane = threadIdx.x & 63;
  const unsigned long long below = (1ull << lane) - 1;
  // pass 1: the chunk's count per bucket
  for (uint32_t r0 = lo; r0 < hi; r0 += blockDim.x) {
    const uint32_t r = r0 + threadIdx.x;
    const uint32_t b = r < hi ? bucket_of(rd, r, max_len, nb) : 0xFFu;
    for (uint32_t k = 0; k < nb; k++) {
      const unsigned long long m = __ballot(b == k);
      if (m && lane == (uint32_t)__ffsll((long long)m) - 1) atomicAdd(&loc[k], 2u * __popcll(m));
    }
  }
  __syncthreads();
  if (threadIdx.x < nb)
    base[threadIdx.x] = loc[threadIdx.x] ? start[threadIdx.x] + atomicAdd(&cursor[threadIdx.x],
                                                                         (unsigned long long)loc[threadIdx.x])
                                         : 0ull;
  __syncthreads();
  if (threadIdx.x < 8) loc[threadIdx.x] = 0;
  __syncthreads();
  // pass 2: places within the chunk's ranges
  for (uint32_t r0 = lo; r0 < hi; r0 += blockDim.x) {
    const uint32_t r = r0 + threadIdx.x;
    const uint32_t b = r < hi ? bucket_of(rd, r, max_len, nb) : 0xFFu;
    uint64_t at = 0;
    for (uint32_t k = 0; k < nb; k++) {
      const unsigned long long m = __ballot(b == k);
      uint32_t first = 0;
      if (m && lane == (uint32_t)__ffsll((long long)m) - 1) first = atomicAdd(&loc[k], 2u * __popcll(m));
      first = __shfl(first, __ffsll((long long)m) - 1 < 0 ? 0 : __ffsll((long long)m) - 1, 64);
      if (b == k) at = base[k] + first + 2u * __popcll(m & below);
    }
    if (r < hi) {
      list[at] = 2 * r;
      list[at + 1] = 2 * r + 1;
    }
  }
}

// Per-bucket queue start offsets into the binned list.
__global__ void k_bin_offsets(const unsigned long long *cnt, uint32_t nb, unsigned long long *start) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    unsigned long long a = 0;
    for (uint32_t i = 0; i < nb; i++) {
      start[i] = a;
      a += cnt[i];
    }
  }
}

// Seed scoring only (drop-in for the scoring role of CUDAGetHighestScore).
__global__ __launch_bounds__(64) void k_score_seeds(IndexView ix, Luts lut, const uint8_t *qb,
                                                   const uint64_t *qoff, const uint32_t *pq,
                                                   const uint32_t *pr, const uint64_t *soff,
                                                   const uint32_t *seeds, uint32_t n_pairs,
                                                   uint32_t qcap, uint32_t flags, float *o_score,
                                                   float *o_rel, uint64_t *o_sr, uint64_t *o_sc,
                                                   uint32_t *o_hits) {
  extern __shared__ __align__(16) unsigned char smem[];
  // Every piece starts on an 8-byte boundary: the 64-bit LDS atomics and
  // paired 64-bit accesses on acc need natural alignment.
  Ws ws;
  size_t o = 0;
  ws.q = smem;
  o += al8(qcap);
  ws.bitmap = (uint32_t *)(smem + o);
  o += al8((qcap / 32 + 1) * 4ull);
  ws.red = (uint32_t *)(smem + o);
  o += 64 * 4;
  ws.acc = (unsigned long long *)(smem + o);
  ws.stats = nullptr;
  for (uint32_t p = blockIdx.x; p < n_pairs; p += gridDim.x) {
    const uint32_t qi = pq[p], R = pr[p];
    const uint64_t q0 = qoff[qi];
    const uint64_t m64 = qoff[qi + 1] - q0;
    float score = -2.0f, rel = -1.0f;
    uint64_t sum_r = 0, sum_c = 0;
    uint32_t hits = 0;
    if (m64 <= qcap && R < ix.n_refs) {
      const uint32_t m = (uint32_t)m64;
      __syncthreads();
      for (uint32_t x = threadIdx.x; x < m; x += 64) ws.q[x] = qb[q0 + x];
      const uint64_t s0 = soff[p], s1 = soff[p + 1];
      const uint64_t ro = ix.ref_off[R];
      const uint64_t ref_len = ix.ref_off[R + 1] - ro;
      auto seed = [&](uint32_t i, uint32_t &sref, uint32_t &sq, uint32_t &lref, uint32_t &lq,
                      uint32_t &wt) {
        const uint32_t *sd = seeds + 4 * (s0 + i);
        sref = sd[0];
        sq = sd[1];
        lref = sd[2];
        lq = sd[3];
        wt = 1;
      };
      if (s1 > s0) {
        score_ref<BlockTeam<64>>(ws.q, m, ix.bytes + ro, ref_len, ix.bytes, (uint32_t)(s1 - s0),
                                 ix.T, (flags & MH_CLASSIFY_ULTRAFAST) != 0, ws, seed, sum_r,
                                 sum_c, hits);
        score = matrix_score(sum_r, sum_c, hits, m);
        if (score > -1.0f && !rel_score(lut, m, ref_len, score, rel)) rel = -1.0f;
      }
    }
    if (threadIdx.x == 0) {
      o_score[p] = score;
      if (o_rel) o_rel[p] = rel;
      if (o_sr) o_sr[p] = sum_r;
      if (o_sc) o_sc[p] = sum_c;
      if (o_hits) o_hits[p] = hits;
    }
  }
}

// Seed check of mh_score_seeds, before any scoring: the scoring kernel
// trusts every seed it extends.  A seed (start, last) with start.query <=
// last.query is scored (MatrixCPU.cpp:57-58 skips the others), so it must
// lie inside its query (last.query <= m: the row bitmap holds m bits) and
// its reference (start.ref < ref_len: the head extension reads the
// reference byte at start.ref; last.ref <= ref_len).  The first offending
// pair is kept as pair << 8 | code (atomicMin: the lowest pair wins).
enum SeedCheck : uint32_t {
  kSeedOk = 0, kSeedRefId = 1, kSeedQueryLen = 2, kSeedOffsets = 3, kSeedNull = 4,
  kSeedQuery = 5, kSeedRef = 6
};
__global__ __launch_bounds__(256) void k_check_seeds(uint32_t n_refs, const uint64_t *ref_off,
                                                     const uint64_t *qoff, const uint32_t *pq,
                                                     const uint32_t *pr, const uint64_t *soff,
                                                     const uint32_t *seeds, uint32_t n_pairs,
                                                     uint32_t max_len,
                                                     unsigned long long *first_bad) {
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n_pairs;
       p += gridDim.x * blockDim.x) {
    uint32_t code = kSeedOk;
    const uint32_t qi = pq[p], R = pr[p];
    const uint64_t m = qoff[qi + 1] - qoff[qi];
    const uint64_t s0 = soff[p], s1 = soff[p + 1];
    if (R >= n_refs) code = kSeedRefId;
    else if (m > max_len) code = kSeedQueryLen;
    else if (s1 < s0) code = kSeedOffsets;
    else if (s1 > s0 && seeds == nullptr) code = kSeedNull;
    else {
      const uint64_t ref_len = ref_off[R + 1] - ref_off[R];
      for (uint64_t i = s0; i < s1 && code == kSeedOk; i++) {
        const uint32_t *sd = seeds + 4 * i;  // start.ref, start.query, last.ref, last.query
        const uint32_t sref = sd[0], sq = sd[1], lref = sd[2], lq = sd[3];
        if (sq > lq) continue;
        if (lq > m) code = kSeedQuery;
        else if (sref >= ref_len || lref > ref_len) code = kSeedRef;
      }
    }
    if (code != kSeedOk) atomicMin(first_bad, ((unsigned long long)p << 8) | code);
  }
}

__global__ void k_set_u64(unsigned long long *p, unsigned long long v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

bool g_comp_ready[64] = {};

}  // namespace

hipError_t ensure_comp_table(int device) {
  if (device >= 0 && device < 64 && g_comp_ready[device]) return hipSuccess;
  uint8_t t[256] = {};
  const char *from = "ATUGCYRSWKMBDHVN-~";
  const char *to = "TAACGRYSWMKVHDBN-~";
  for (int i = 0; from[i]; i++) t[(uint8_t)from[i]] = (uint8_t)to[i];
  hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(kComp), t, sizeof(t));
  if (e == hipSuccess && device >= 0 && device < 64) g_comp_ready[device] = true;
  return e;
}

namespace {

// Teams per workgroup for a tile's block field: 8, 16 and 32 are lane teams
// of a one-wavefront workgroup; 64..1024 are whole workgroups.
uint32_t teams_of(uint32_t block) { return block < 64 ? 64 / block : 1; }
uint32_t threads_of(uint32_t block) { return block < 64 ? 64 : block; }

template <class TM>
hipError_t launch_team(const Ctx &c, const Caps &cap, const uint32_t *qlist,
                       const unsigned long long *qstart, uint32_t nq,
                       const unsigned long long *nq_dev, unsigned long long *head, uint32_t grid,
                       hipStream_t st) {
  const size_t lds = TM::kPerWave * ws_stride_lds(cap);
  hipFuncSetAttribute((const void *)k_classify<TM>, hipFuncAttributeMaxDynamicSharedMemorySize,
                      (int)lds);
  if (!head) grid = (nq + TM::kPerWave - 1) / TM::kPerWave;  // plain: a query per team
  hipLaunchKernelGGL((k_classify<TM>), dim3(grid), dim3(TM::kSize * TM::kPerWave), lds, st, c,
                     cap, qlist, qstart, nq, nq_dev, head);
  return hipGetLastError();
}

hipError_t launch_tile(uint32_t block, const Ctx &c, const Caps &cap, const uint32_t *qlist,
                       const unsigned long long *qstart, uint32_t nq,
                       const unsigned long long *nq_dev, unsigned long long *head, uint32_t grid,
                       hipStream_t st) {
  switch (block) {
    case 8: return launch_team<LaneTeam<8>>(c, cap, qlist, qstart, nq, nq_dev, head, grid, st);
    case 16: return launch_team<LaneTeam<16>>(c, cap, qlist, qstart, nq, nq_dev, head, grid, st);
    case 32: return launch_team<LaneTeam<32>>(c, cap, qlist, qstart, nq, nq_dev, head, grid, st);
    case 64: return launch_team<BlockTeam<64>>(c, cap, qlist, qstart, nq, nq_dev, head, grid, st);
    case 128: return launch_team<BlockTeam<128>>(c, cap, qlist, qstart, nq, nq_dev, head, grid, st);
    case 256: return launch_team<BlockTeam<256>>(c, cap, qlist, qstart, nq, nq_dev, head, grid, st);
    case 512: return launch_team<BlockTeam<512>>(c, cap, qlist, qstart, nq, nq_dev, head, grid, st);
    default:
      return launch_team<BlockTeam<1024>>(c, cap, qlist, qstart, nq, nq_dev, head, grid, st);
  }
}

template <class TM>
uint32_t resident(size_t lds) {
  int n = 0;
  const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &n, k_classify<TM>, TM::kSize * TM::kPerWave, lds);
  return (e == hipSuccess && n > 0) ? (uint32_t)n : 1u;
}

// Workgroups of one tile kernel resident per CU with this LDS per workgroup.
uint32_t resident_blocks(uint32_t block, size_t lds) {
  switch (block) {
    case 8: return resident<LaneTeam<8>>(lds);
    case 16: return resident<LaneTeam<16>>(lds);
    case 32: return resident<LaneTeam<32>>(lds);
    case 64: return resident<BlockTeam<64>>(lds);
    case 128: return resident<BlockTeam<128>>(lds);
    case 256: return resident<BlockTeam<256>>(lds);
    case 512: return resident<BlockTeam<512>>(lds);
    default: return resident<BlockTeam<1024>>(lds);
  }
}

__global__ void k_max_len(Reads rd, unsigned long long *out) {
  uint32_t mx = 0;
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < rd.n; r += gridDim.x * blockDim.x)
    mx = max(mx, (uint32_t)min(rd.off[r + 1] - rd.off[r], (uint64_t)0xFFFFFFFFu));
  mx = wave_max(mx);
  if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)mx);
}

// LDS capacities of a tile, shrunk until the workspace fits `limit` bytes:
// first the entry capacity (to what fits, multiples of 16, down to 256),
// then the query length.  Queries beyond the capacities go to the overflow
// tiers.
// References scored together by teams of 128 or more threads and by the HBM
// tiers, for strands with two or more candidate references: each takes a
// 64-byte slot and a row bitmap of the query capacity.  In LDS only
// workspaces of reads up to 2 kb get them (16 slots cost at most 5.2 KB
// there; longer reads score one reference at a time); HBM workspaces get
// them at any length (about 66 KB each for 32 kb reads).
constexpr uint32_t kRefSlots = 16;
constexpr uint32_t kRefSlotMaxLen = 2048;
// LDS tiles give the slots about 4 KB (more would cost the second-chance
// tier entries at four workgroups per CU): 20 slots sized for 1 kb reads,
// which multi-reference scoring turns into 44 for a 150 bp read.
constexpr size_t kSlotBytes = 4096;
constexpr uint32_t kMaxRefSlots = 64;

uint32_t lds_slots(uint32_t qcap) {
  const size_t per = sizeof(RefSlot) + (qcap / 32 + 1) * 4ull;
  return (uint32_t)std::min<size_t>(kMaxRefSlots, std::max<size_t>(kRefSlots, kSlotBytes / per));
}

Caps caps_for(const mh_tile &t, int s, size_t limit) {
  Caps c;
  c.nslot = (t.block >= 128 && t.max_len <= kRefSlotMaxLen)
                ? lds_slots((t.max_len + 63) & ~63u)
                : 0;
  const uint32_t sd = (uint32_t)(s > 0 ? s : 1);
  c.qcap = (t.max_len + 63) & ~63u;
  c.pcap = c.qcap / sd + 2;
  c.ecap = t.entries_cap;
  if (ws_bytes(c) > limit) {
    Caps z = c;
    z.ecap = 0;
    const size_t fixed = ws_bytes(z);
    const size_t fit = fixed < limit ? (limit - fixed) / 40 : 0;  // 40 bytes per entry
    c.ecap = (uint32_t)std::max<size_t>(256, std::min<size_t>(c.ecap, fit) & ~size_t(15));
    while (ws_bytes(c) > limit && c.ecap > 256) c.ecap -= 16;
  }
  while (ws_bytes(c) > limit && c.qcap > 64) {
    c.qcap = ((c.qcap / 2) + 63) & ~63u;
    c.pcap = c.qcap / sd + 2;
  }
  return c;
}

// HBM tiers: a first one with moderate workspaces over many workgroups,
// then the last resort with the largest entry capacity over a few.  They run
// one after the other and share one region.  The first tier's strands wait on
// global memory, so it runs as many 256-thread workgroups as are resident at
// once (four per CU at its register count): 1.3 GiB of workspaces for 32 kb
// reads.
constexpr uint32_t kHbmBlocks = 1024;
constexpr uint32_t kHbmEcap = 1u << 15;
constexpr uint32_t kBigBlocks = 32;
constexpr uint32_t kBigEcap = 1u << 19;
// In a call of at most this many reads the first HBM tier holds about one
// strand per workgroup, and its duration is one strand's latency: such calls
// run it with 1024-thread teams, which finish a strand sooner (served stream
// of workload 5 at 0.25 load: service p50 0.87 -> 0.79 ms).
constexpr uint32_t kHbmWideCall = 1u << 17;

// Capacities of a tile's team workspace: the workgroup's LDS is shared by
// its teams.
Caps tile_caps(const mh_tile &t, int s, size_t lds_per_block) {
  // 1 KB stays free for the kernel's static LDS and the 16-byte rounding of
  // each team's workspace, and lane teams' query chunk (16 bytes a query)
  const uint32_t teams = teams_of(t.block);
  const size_t reserve = 1024 + (teams > 1 ? 16ull * kQueueChunk * teams : 0);
  const size_t usable = lds_per_block > 2 * reserve ? lds_per_block - reserve : lds_per_block / 2;
  return caps_for(t, s, usable / teams);
}

// LDS of one workgroup of a tile.
size_t tile_lds(const mh_tile &t, const Caps &c) { return teams_of(t.block) * ws_stride_lds(c); }

Caps hbm_caps(uint32_t max_read_len, int s, uint32_t ecap) {
  Caps c;
  c.nslot = kRefSlots;
  c.qcap = (max_read_len + 63) & ~63u;
  c.pcap = c.qcap / (uint32_t)(s > 0 ? s : 1) + 2;
  c.ecap = ecap;
  return c;
}

size_t ws_stride(const Caps &c) { return (ws_bytes(c) + 255) & ~size_t(255); }

}  // namespace

std::vector<mh_tile> default_tiles(uint32_t max_read_len, int T, uint32_t epp_x16) {
  // Entry capacities follow the expected bucket entries of each length: the
  // base table fits about two entries per probe (a 5000-reference index at
  // k = 16), and short reads run 2 or 4 to a wavefront.  An index with more
  // entries per bucket widens a tile whose expected entries (the probes of
  // a read of 3/4 its longest length times the index's entries per bucket,
  // plus 8) exceed the base capacity: capacity the next power of two, at
  // least 64 threads per 256 entries.  (On 8e9 entries, 29.8 per bucket, the
  // 150 bp reads at T = 38 take 64-thread teams of 256 entries, the fastest
  // measured: 542 ms for 10M reads against 1,012 with 16-lane teams of 32.)
  const mh_tile base[7] = {{200, 16, 32, 0},    {400, 32, 64, 0},    {1024, 128, 512, 0},
                           {2048, 128, 512, 0},  {4096, 256, 1024, 0}, {8192, 512, 1024, 0},
                           {0xFFFFFFFFu, 1024, 2048, 0}};
  const uint32_t s = (uint32_t)std::max(1, T / 2);
  std::vector<mh_tile> t;
  for (mh_tile b : base) {
    const bool last = b.max_len >= max_read_len;
    if (last) b.max_len = max_read_len;
    if (T > 0 && epp_x16) {
      const uint32_t len = b.max_len / 4 * 3;
      const uint64_t probes = len > (uint32_t)T ? (len - (uint32_t)T) / s + 1 : 1;
      const uint64_t want = (probes * epp_x16 + 15) / 16 + 8;
      if (want > b.entries_cap) {
        uint32_t cap = b.entries_cap;
        while (cap < want && cap < 4096) cap *= 2;
        b.entries_cap = cap;
        b.block = std::max<uint32_t>(b.block, std::min<uint32_t>(1024, cap / 4));
      }
    }
    t.push_back(b);
    if (last) break;
  }
  return t;
}

std::vector<mh_tile> default_tier_tiles(uint32_t max_read_len, uint32_t first_cap) {
  // The second-chance tier holds four times the first tile's entries (at
  // least 512), a 64-thread wavefront per 256 entries: on 8e9 entries the
  // reads of shared segments (about 900 entries per strand at 150 bp) then
  // stay in LDS (1,024 entries, 256 threads: 385 against 543 ms for 10M
  // reads with the 512-entry tier, whose overflow went to HBM workspaces).
  const uint32_t cap0 = std::max<uint32_t>(512, std::min<uint32_t>(4096, 4 * first_cap));
  return {mh_tile{std::min<uint32_t>(1024u, max_read_len), std::max<uint32_t>(128, cap0 / 4),
                  cap0, 0},
          mh_tile{max_read_len, 256, 4096, 0}};
}

uint32_t entries_per_bucket_x16(const IndexView &ix) { return ix.epp_x16; }

std::vector<mh_tile> active_tiles(const mh_handle *h, const IndexView &ix) {
  return h->tiles.empty() ? default_tiles(h->max_read_len, ix.T, ix.epp_x16) : h->tiles;
}

// The second-chance tier's entry capacity, lowered to the largest multiple of
// 16 at which one more of its workgroups is resident per CU, when that costs
// at most a quarter of it.  On workload 5 (150 bp reads, 35 entries per
// probe) 1,024 -> 848 entries takes the tier from 3 to 4 workgroups per CU:
// 33.4 -> 25.6 ms per 4M reads, with 1.1 % of its strands passed on to the
// HBM-workspace tier (1.7 ms); 864 entries leaves 3 (59.6 ms per call).
void fit_tier_residency(const mh_handle *h, int s, mh_tile &t) {
  auto resident_at = [&](uint32_t ecap) {
    mh_tile u = t;
    u.entries_cap = ecap;
    return resident_blocks(u.block, tile_lds(u, tile_caps(u, s, h->lds_per_block)));
  };
  const uint32_t n0 = resident_at(t.entries_cap);
  uint32_t lo = ((t.entries_cap * 3 / 4) + 15) & ~15u, hi = t.entries_cap - 16;
  if (lo > hi || resident_at(lo) <= n0) return;
  while (lo < hi) {  // resident_at(lo) > n0; residency falls as the capacity grows
    const uint32_t mid = ((lo + hi + 16) / 2) & ~15u;
    if (resident_at(mid) > n0) lo = mid;
    else hi = mid - 16;
  }
  t.entries_cap = lo;
}

std::vector<mh_tile> tier_tiles(const mh_handle *h, const IndexView &ix) {
  if (h->tier_tiles.size() == 2) return h->tier_tiles;
  const uint32_t cap = active_tiles(h, ix)[0].entries_cap;
  std::lock_guard<std::mutex> lk(h->tier_fit_mu);
  if (h->tier_fit.size() != 2 || h->tier_fit_cap != cap || h->tier_fit_s != (uint32_t)ix.s) {
    std::vector<mh_tile> t = default_tier_tiles(h->max_read_len, cap);
    fit_tier_residency(h, ix.s, t[0]);
    h->tier_fit = t;
    h->tier_fit_cap = cap;
    h->tier_fit_s = (uint32_t)ix.s;
  }
  return h->tier_fit;
}

// Probe records per strand of the first length tile (0: that tile stages,
// hashes and looks up itself): with MH_CLASSIFY_PREPROBE, a status word and
// one record per probe the tile can hold, for tiles of reads up to 1 kb.
uint32_t probe_stride(const std::vector<mh_tile> &tiles, int s, size_t lds_limit,
                      uint32_t flags) {
  if (!(flags & MH_CLASSIFY_PREPROBE)) return 0;
  const Caps c0 = tile_caps(tiles[0], s, lds_limit);
  return c0.qcap <= 1024 ? c0.pcap + 1 : 0u;
}

// Scratch layout for one classify call (bytes from the handle's arena):
//   QueryOut[2n] | overflow lists u32[2n] x 2 | binned list u32[2n] |
//   merge counts u32[n+1] | merge offsets u32[n+1] | tie pool u32[f*2n] |
//   bucket words u64[64] | scan temp | probe records u64[2n * stride] |
//   HBM-tier workspaces
struct Layout {
  size_t qo, overflow, overflow2, binned, cnt, off, ties, bw, scan, rec, big, total;
  size_t scan_bytes, tie_cap;
};

Layout layout(uint32_t n_reads, uint32_t max_read_len, int s, uint32_t tie_factor,
              uint32_t rec_stride, uint32_t hbm_blocks = kHbmBlocks,
              uint32_t big_blocks = kBigBlocks) {
  const size_t nq = 2ull * n_reads;
  Layout L;
  size_t o = 0;
  auto take = [&](size_t b) {
    size_t at = o;
    o = (o + b + 255) & ~size_t(255);
    return at;
  };
  L.qo = take(nq * sizeof(QueryOut));
  L.overflow = take(nq * 4);
  L.overflow2 = take(nq * 4);
  L.binned = take(nq * 4);
  L.cnt = take((n_reads + 1ull) * 4);
  L.off = take((n_reads + 1ull) * 4);
  L.tie_cap = (size_t)tie_factor * nq + 256;
  L.ties = take(L.tie_cap * 4);
  L.bw = take(64 * 8);
  L.scan_bytes = 0;
  hipcub::DeviceScan::ExclusiveSum(nullptr, L.scan_bytes, (const uint32_t *)nullptr,
                                   (uint32_t *)nullptr, (int)(n_reads + 1));
  L.scan = take(L.scan_bytes);
  L.rec = take(nq * rec_stride * 8ull);
  L.big = take(std::max(hbm_blocks * ws_stride(hbm_caps(max_read_len, s, kHbmEcap)),
                        big_blocks * ws_stride(hbm_caps(max_read_len, s, kBigEcap))));
  L.total = o;
  return L;
}

size_t classify_scratch_bytes(const mh_handle *h, const IndexView &ix, uint32_t n_reads,
                              uint32_t tie_factor, uint32_t flags) {
  const uint32_t stride = probe_stride(active_tiles(h, ix), ix.s, h->lds_per_block, flags);
  return layout(n_reads, h->max_read_len, ix.s, tie_factor, stride).total;
}

mh_status plan_fixed_scratch(const mh_handle *h, const IndexView &ix, uint32_t n_reads,
                             uint32_t flags, size_t bytes, FixedScratch &fs) {
  const int s = ix.s;
  const uint32_t stride = probe_stride(active_tiles(h, ix), s, h->lds_per_block, flags);
  // Per-read parts with the smallest tie pool, and one workspace of each
  // HBM tier: the least a call needs.
  const Layout least = layout(n_reads, h->max_read_len, s, 4, stride, 1, 1);
  if (least.total > bytes) return MH_ERR_RANGE;
  const size_t hs = ws_stride(hbm_caps(h->max_read_len, s, kHbmEcap));
  const size_t bs = ws_stride(hbm_caps(h->max_read_len, s, kBigEcap));
  const size_t region = bytes - least.big;  // least.big: bytes before the HBM-tier region
  // HBM-tier workspaces take what the per-read parts leave (they run one
  // tier after the other in one region), up to the allocating path's counts.
  fs.hbm_blocks = (uint32_t)std::min<size_t>(kHbmBlocks, region / hs);
  fs.big_blocks = (uint32_t)std::min<size_t>(kBigBlocks, region / bs);
  if (!fs.hbm_blocks || !fs.big_blocks) return MH_ERR_RANGE;
  // Then the largest tie pool that fits, at most the allocating path's
  // largest default.
  fs.tie_factor = 4;
  while (fs.tie_factor < 64 &&
         layout(n_reads, h->max_read_len, s, fs.tie_factor * 2, stride, fs.hbm_blocks,
                fs.big_blocks).total <= bytes)
    fs.tie_factor *= 2;
  fs.bytes = bytes;
  fs.tie_cap = (size_t)fs.tie_factor * 2ull * n_reads + 256;
  return MH_OK;
}

mh_status launch_classify(mh_handle *h, Exec &x, const ClassifyArgs &a, bool batched,
                          const FixedScratch *fixed) {
  hipError_t e = hipSuccess;
  const uint32_t n = a.rd.n;
  const uint64_t nq = 2ull * n;
  const std::vector<mh_tile> tiles = active_tiles(h, a.ix);
  const uint32_t rec_stride =
      batched ? probe_stride(tiles, a.ix.s, h->lds_per_block, a.flags) : 0u;
  const uint32_t hbm_blocks = fixed ? fixed->hbm_blocks : kHbmBlocks;
  const uint32_t big_blocks = fixed ? fixed->big_blocks : kBigBlocks;
  const Layout L = layout(n, h->max_read_len, a.ix.s, fixed ? fixed->tie_factor : x.tie_factor,
                          rec_stride, hbm_blocks, big_blocks);
  unsigned char *base;
  if (fixed) {
    // A caller-provided region of fixed size: nothing is allocated.
    if (L.total > fixed->bytes)
      return set_error(h, MH_ERR_RANGE, "fixed scratch smaller than the call's layout");
    base = fixed->base;
  } else {
    mh_status st = ensure_scratch(h, x, L.total);
    if (st) return st;
    base = (unsigned char *)x.d_scratch;
  }
  QueryOut *qo = (QueryOut *)(base + L.qo);
  uint32_t *overflow = (uint32_t *)(base + L.overflow);
  uint32_t *overflow2 = (uint32_t *)(base + L.overflow2);
  uint32_t *binned = (uint32_t *)(base + L.binned);
  uint32_t *mcnt = (uint32_t *)(base + L.cnt);
  uint32_t *moff = (uint32_t *)(base + L.off);
  // bucket words: [0..7] count, [8..15] cursor, [16..23] start, [24] longest
  // read, [32..35] bucket length limits (u32)
  unsigned long long *bw = (unsigned long long *)(base + L.bw);
  unsigned char *big_ws = base + L.big;

  Ctx c;
  c.ix = a.ix;
  c.lut = a.luts;
  c.rd = a.rd;
  c.flags = a.flags;
  c.max_len = h->max_read_len;
  c.rec = nullptr;
  c.rec_stride = 0;
  c.rec_maxlen = 0;
  c.lds_ecap_max = 0;
  c.kid = 0;
  c.out = a.out;
  c.out.q = qo;
  c.out.ctr = x.d_ctr;
  c.out.overflow = overflow;
  c.out.overflow_cnt = &x.d_ctr->tier_n[0];
  c.out.overflow_cap = (uint32_t)nq;
  c.out.ties = (uint32_t *)(base + L.ties);
  c.out.tie_cap = L.tie_cap;
  hipStream_t s = x.stream;
  const uint32_t zn = sizeof(DevCounters) / 8;
  const bool timed = (a.flags & MH_CLASSIFY_TIMING) != 0 && !x.capturing;
  x.timed = false;
  x.timed_buckets = 0;
  x.timed_probe = false;
  if (timed) hipEventRecord(x.ev[0], s);
  hipLaunchKernelGGL(k_zero, dim3(8), dim3(256), 0, s, (unsigned long long *)x.d_ctr, zn);
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, s, bw, 64u);
  if (nq) {
    const unsigned gr = (unsigned)((n + 255) / 256);
    // Probe records of the first length tile's strands, also read by the LDS
    // overflow tiers for the strands that tile passes on (none without
    // MH_CLASSIFY_PREPROBE).
    // A strand needing more entries than the second-chance tier holds goes
    // flagged to the HBM-workspace tier.  Its 1,024 workgroups (four per
    // CU) finish such strands sooner than the middle LDS tier, whose 32 kb
    // query and ~3,200-entry workspace leave one workgroup per CU (2M
    // mixed-length reads: tiers 10.5 -> 9.2 ms; workload 5 in the plain form,
    // 2M reads: 42.9 -> 39 ms).  The middle tier keeps strands whose entries
    // fit the second-chance tier but whose read is longer than its 1,024
    // bases.
    c.lds_ecap_max = tile_caps(tier_tiles(h, a.ix)[0], a.ix.s, h->lds_per_block).ecap;
    Ctx crec = c;
    // The persistent kernels take strands from queues, so a grid of more
    // workgroups than the call has strands would only dispatch idle ones.
    const uint32_t max_grid = (uint32_t)std::min<uint64_t>(nq, 1u << 30);
    const uint32_t hbm_grid = std::min(hbm_blocks, max_grid);
    if (!batched) {
      // Plain form: one launch per batch, one workgroup per query, with the
      // single tile that covers the longest read of the batch.
      hipLaunchKernelGGL(k_max_len, dim3(std::min<uint32_t>(gr, 1024)), dim3(256), 0, s, a.rd,
                         bw + 24);
      unsigned long long mlen = 0;
      e = hipMemcpyAsync(&mlen, bw + 24, 8, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      if (e != hipSuccess) return hip_error(h, e, "max read length");
      size_t ti = 0;
      while (ti + 1 < tiles.size() && mlen > tiles[ti].max_len) ti++;
      // whole-workgroup teams here: the plain form is the simple one
      mh_tile t0 = tiles[ti];
      t0.block = std::max<uint32_t>(t0.block, 64);
      const Caps cap = tile_caps(t0, a.ix.s, h->lds_per_block);
      if (timed) hipEventRecord(x.ev[9], s);
      e = launch_tile(t0.block, c, cap, nullptr, nullptr, (uint32_t)nq, nullptr, nullptr, 0, s);
      if (e != hipSuccess) return hip_error(h, e, "k_classify");
      if (timed) {
        hipEventRecord(x.ev[1], s);
        x.bucket_order[0] = 0;
        x.timed_buckets = 1;
      }
    } else {
      const uint32_t nb = (uint32_t)std::min<size_t>(tiles.size(), 8);
      Lens d_maxlen{};
      for (uint32_t i = 0; i < nb; i++) d_maxlen.v[i] = tiles[i].max_len;
      const Reads rd = a.rd;
      const uint32_t bin_grid = std::min<uint32_t>(gr, kBinBlocks);
      hipLaunchKernelGGL(k_bin_count, dim3(bin_grid), dim3(256), 0, s, rd, d_maxlen, nb, bw);
      hipLaunchKernelGGL(k_bin_offsets, dim3(1), dim3(64), 0, s, bw, nb, bw + 16);
      hipLaunchKernelGGL(k_bin_scatter, dim3(bin_grid), dim3(256), 0, s, rd, d_maxlen, nb, bw + 16,
                         bw + 8, binned);
      // Probe records for the first tile's strands.
      Ctx c0 = c;
      if (rec_stride) {
        if (timed) {
          hipEventRecord(x.ev[12], s);
          x.timed_probe = true;
        }
        uint64_t *rec = (uint64_t *)(base + L.rec);
        const Caps c0cap = tile_caps(tiles[0], a.ix.s, h->lds_per_block);
        const uint32_t maxlen = c0cap.qcap;
        const uint32_t row_words = (maxlen + 3 + 3) / 4;  // the read, any alignment
        const size_t lds = 256 + (size_t)kProbeTeams * 2 * row_words * 4;
        const uint32_t blocks =
            (uint32_t)std::min<uint64_t>((nq + kProbeTeams - 1) / kProbeTeams,
                                         (uint64_t)h->cu_count * 64);
        hipFuncSetAttribute((const void *)k_probe, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds);
        hipLaunchKernelGGL(k_probe, dim3(blocks), dim3(kProbeLanes * kProbeTeams), lds, s, a.ix,
                           rd, (const uint32_t *)binned, (const unsigned long long *)(bw + 16),
                           (const unsigned long long *)bw, rec_stride, row_words, maxlen, rec);
        c0.rec = rec;
        c0.rec_stride = rec_stride;
        c0.rec_maxlen = maxlen;
        // a tier's strand has a record when it was in the first tile's bin
        crec.rec = rec;
        crec.rec_stride = rec_stride;
        crec.rec_maxlen = std::min<uint32_t>(maxlen, tiles[0].max_len);
      }
      // Long buckets first so that their tail overlaps with the short ones.
      if (timed) hipEventRecord(x.ev[9], s);
      for (int b = (int)nb - 1; b >= 0; b--) {
        // A read longer than tiles[b - 1].max_len is the only kind that bins
        // into tile b or later.
        if (b > 0 && a.max_len_hint <= tiles[b - 1].max_len) continue;
        const mh_tile &t = tiles[b];
        const Caps cap = tile_caps(t, a.ix.s, h->lds_per_block);
        uint32_t grid = t.grid;
        if (!grid) grid = h->cu_count * resident_blocks(t.block, tile_lds(t, cap));
        grid = std::min(grid, max_grid);
        // queue b: binned[start[b], start[b] + cnt[b]), head queue_head[b]
        Ctx cb = b == 0 ? c0 : c;
        cb.kid = (uint32_t)b;
        e = launch_tile(t.block, cb, cap, binned, bw + 16 + b, 0, bw + b,
                        &x.d_ctr->queue_head[b], grid, s);
        if (e != hipSuccess) return hip_error(h, e, "k_classify batched");
        if (timed) {
          hipEventRecord(x.ev[1 + x.timed_buckets], s);
          x.bucket_order[x.timed_buckets++] = (uint32_t)b;
        }
      }
    }
    // Overflow tiers, each a persistent kernel over the list the one before
    // it filled (two lists in turn):
    //   0  second chance (LDS, default 128-thread teams, 1024 bases, 512 entries)
    //   1  middle (LDS, default 256-thread teams with the largest workspace)
    //   2  HBM workspaces of 2^15 entries over kHbmBlocks workgroups
    //   3  last resort: HBM workspaces of 2^19 entries, flags what is left
    uint32_t *list[2] = {overflow, overflow2};
    DevCounters *d = x.d_ctr;
    const std::vector<mh_tile> lds_tiers = tier_tiles(h, a.ix);  // lds_ecap_max: tier 0's cap
    for (int t = 0; t < 2; t++) {
      const Caps tc = tile_caps(lds_tiers[t], a.ix.s, h->lds_per_block);
      Ctx ct = crec;
      ct.kid = kTierRow0 + (uint32_t)t;
      ct.out.overflow = list[(t + 1) & 1];
      ct.out.overflow_cnt = &d->tier_n[t + 1];
      const uint32_t grid = std::min(
          max_grid, h->cu_count * resident_blocks(lds_tiers[t].block, tile_lds(lds_tiers[t], tc)));
      e = launch_tile(lds_tiers[t].block, ct, tc, list[t & 1], nullptr, 0, &d->tier_n[t],
                      &d->tier_head[t], grid, s);
      if (e != hipSuccess) return hip_error(h, e, "k_classify overflow tier");
      if (timed) hipEventRecord(x.ev[13 + t], s);
    }
    {
      const Caps ca = hbm_caps(h->max_read_len, a.ix.s, kHbmEcap);
      Ctx ch = c;
      ch.kid = kTierRow0 + 2;
      ch.out.overflow = list[1];
      ch.out.overflow_cnt = &d->tier_n[3];
      if (n <= kHbmWideCall)
        hipLaunchKernelGGL(k_classify_hbm<1024>, dim3(hbm_grid), dim3(1024), 0, s, ch, ca,
                           big_ws, ws_stride(ca), (const uint32_t *)list[0],
                           (const unsigned long long *)&d->tier_n[2], &d->tier_head[2], 0);
      else
        hipLaunchKernelGGL(k_classify_hbm<256>, dim3(hbm_grid), dim3(256), 0, s, ch, ca,
                           big_ws, ws_stride(ca), (const uint32_t *)list[0],
                           (const unsigned long long *)&d->tier_n[2], &d->tier_head[2], 0);
      if (timed) hipEventRecord(x.ev[15], s);
      const Caps cb = hbm_caps(h->max_read_len, a.ix.s, kBigEcap);
      Ctx cl = c;
      cl.kid = kTierRow0 + 3;
      hipLaunchKernelGGL(k_classify_hbm<1024>, dim3(std::min(big_blocks, max_grid)), dim3(1024), 0,
                         s, cl, cb, big_ws,
                         ws_stride(cb), (const uint32_t *)list[1],
                         (const unsigned long long *)&d->tier_n[3], &d->tier_head[3], 1);
    }
    if (timed) hipEventRecord(x.ev[10], s);
    hipLaunchKernelGGL(k_merge_count, dim3(std::min<uint32_t>(gr, 4096)), dim3(256), 0, s,
                       (const QueryOut *)qo, n, a.final_out, mcnt);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, mcnt + n, 1u);
    size_t sb = L.scan_bytes;
    e = hipcub::DeviceScan::ExclusiveSum((void *)(base + L.scan), sb, mcnt, moff, (int)(n + 1),
                                         s);
    if (e != hipSuccess) return hip_error(h, e, "merge scan");
    hipLaunchKernelGGL(k_merge_write, dim3(std::min<uint32_t>(gr, 4096)), dim3(256), 0, s,
                       (const QueryOut *)qo, n, (const uint32_t *)moff,
                       (const uint32_t *)c.out.ties, (uint64_t)c.out.tie_cap, a.final_out,
                       a.out.top, a.out.top_cap, x.d_ctr);
    if (timed) {
      hipEventRecord(x.ev[11], s);
      x.timed = true;
    }
    e = hipGetLastError();
    if (e != hipSuccess) return hip_error(h, e, "classify launch");
  }
  return MH_OK;
}

void kernel_counters_from(const DevCounters &d, mh_counters *o, uint32_t n) {
  for (uint32_t k = 0; k < n && k < (uint32_t)kKernelRows; k++) {
    unsigned long long t[kNumStats] = {};
    for (int sl = 0; sl < kKernelSlots; sl++)
      for (int i = 0; i < kNumStats; i++) t[i] += d.kstats[k][sl][i];
    o[k].queries = t[kStrands];
    o[k].probes = t[kProbes];
    o[k].hits = t[kHits];
    o[k].entries = t[kEntries];
    o[k].seeds = t[kSeeds];
    o[k].cands = t[kCands];
    o[k].direct = t[kDirect];
    o[k].overflow = k >= (uint32_t)kTierRow0 ? d.tier_n[k - kTierRow0] : 0;
    o[k].ext_compares = t[kExt];
  }
}

void counters_from(const DevCounters &d, uint64_t n_reads, mh_counters &o) {
  unsigned long long t[kNumStats] = {};
  for (int sl = 0; sl < kStatSlots; sl++)
    for (int i = 0; i < kNumStats; i++) t[i] += d.stats[sl][i];
  o.queries = 2 * n_reads;
  o.probes = t[kProbes];
  o.hits = t[kHits];
  o.entries = t[kEntries];
  o.seeds = t[kSeeds];
  o.cands = t[kCands];
  o.direct = t[kDirect];
  o.overflow = d.tier_n[0] + d.tier_n[1] + d.tier_n[2] + d.tier_n[3];
  o.ext_compares = t[kExt];
}

mh_status launch_score_seeds(mh_handle *h, const IndexView &ix, const Luts &luts,
                             const uint8_t *d_q, const uint64_t *d_qoff, const uint32_t *d_pq,
                             const uint32_t *d_pr, const uint64_t *d_soff, const uint32_t *d_seeds,
                             uint32_t n_pairs, uint32_t flags, float *d_score, float *d_rel,
                             uint64_t *d_sum_r, uint64_t *d_sum_c, uint32_t *d_hits) {
  if (!n_pairs) return MH_OK;
  {
    // Every pair and seed is checked before anything is scored.
    Exec &x = h->ex;
    unsigned long long *bad = &x.d_ctr->seed_check;
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, x.stream, bad, ~0ull);
    const uint32_t cgrid = std::min<uint32_t>((n_pairs + 255) / 256, 4096);
    hipLaunchKernelGGL(k_check_seeds, dim3(cgrid), dim3(256), 0, x.stream, ix.n_refs,
                       ix.ref_off, d_qoff, d_pq, d_pr, d_soff, d_seeds, n_pairs,
                       h->max_read_len, bad);
    MH_HIP(h, hipGetLastError());
    MH_HIP(h, hipMemcpyAsync(&x.h_ctr->seed_check, bad, 8, hipMemcpyDeviceToHost, x.stream));
    MH_HIP(h, hipStreamSynchronize(x.stream));
    const unsigned long long v = x.h_ctr->seed_check;
    if (v != ~0ull) {
      static const char *what[] = {
          "", "reference id not below the index's reference count",
          "query longer than the handle's maximum read length",
          "seed offsets decrease", "seeds are NULL but the pair has seeds",
          "a seed ends past its query (last.query > query length)",
          "a seed lies outside its reference (start.ref >= length or last.ref > length)"};
      const uint32_t code = (uint32_t)(v & 0xFF);
      const uint64_t pair = v >> 8;
      return set_error(h, code == kSeedQueryLen ? MH_ERR_RANGE : MH_ERR_ARG,
                       "mh_score_seeds: pair " + std::to_string(pair) + ": " +
                           (code < 7 ? what[code] : "bad pair"));
    }
  }
  const uint32_t qcap = (h->max_read_len + 63) & ~63u;
  const size_t lds = al8(qcap) + al8((qcap / 32 + 1) * 4ull) + 64 * 4 + 16;
  const uint32_t grid = std::min<uint32_t>(n_pairs, 65536);
  hipLaunchKernelGGL(k_score_seeds, dim3(grid), dim3(64), lds, h->stream, ix, luts, d_q, d_qoff,
                     d_pq, d_pr, d_soff, d_seeds, n_pairs, qcap, flags, d_score, d_rel, d_sum_r,
                     d_sum_c, d_hits);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_error(h, e, "k_score_seeds");
  return MH_OK;
}

}  // namespace mh
