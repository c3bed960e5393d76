// Seeded synthetic references and reads for the classification workloads.
//
// Every draw comes from xoshiro256** seeded through SplitMix64 with
// (seed, stream id), one stream per reference or read, so the output does
// not depend on the number of threads.  Host-only helper of the benchmark and
// tests; the classification path does not use it.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

uint64_t splitmix64(uint64_t &x) {
  uint64_t z = (x += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

struct Rng {
  uint64_t s[4];
  Rng(uint64_t seed, uint64_t stream) {
    uint64_t x = seed ^ (stream * 0xd1b54a32d192ed03ull);
    for (auto &v : s) v = splitmix64(x);
  }
  static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
  uint64_t next() {  // xoshiro256**
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (next() >> 11) * 0x1.0p-53; }
  uint64_t below(uint64_t n) { return n ? next() % n : 0; }
  double normal() {
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

const char kBases[4] = {'A', 'C', 'G', 'T'};

template <class F>
void par(uint64_t n, F f) {
  unsigned th = std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
  std::atomic<uint64_t> next(0);
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < th; t++)
    pool.emplace_back([&] {
      for (;;) {
        uint64_t i = next.fetch_add(64);
        if (i >= n) break;
        for (uint64_t j = i; j < std::min(n, i + 64); j++) f(j);
      }
    });
  for (auto &p : pool) p.join();
}

inline uint8_t base_gc(Rng &r, double gc) {
  double u = r.uniform();
  if (u < gc * 0.5) return 'G';
  if (u < gc) return 'C';
  if (u < gc + (1 - gc) * 0.5) return 'A';
  return 'T';
}

}  // namespace

extern "C" {

// Length draws.  mode 0: fixed a; 1: round(LogNormal(ln a, b)) capped at cap,
// at least lo; 2: uniform integer in [a, b]; 3: mixture, with probability
// `mix` uniform [a, b], else round(LogNormal(ln c, d)) capped at cap.
int mhg_lengths(uint64_t n, int mode, double a, double b, double c, double d, double mix,
                uint64_t lo, uint64_t cap, uint64_t seed, uint64_t *out) {
  par(n, [&](uint64_t i) {
    Rng r(seed, 0x100000000ull + i);
    double v;
    switch (mode) {
      case 0: v = a; break;
      case 1: v = std::round(std::exp(std::log(a) + b * r.normal())); break;
      case 2: v = (double)((uint64_t)a + r.below((uint64_t)(b - a + 1))); break;
      default:
        if (r.uniform() < mix) v = (double)((uint64_t)a + r.below((uint64_t)(b - a + 1)));
        else v = std::round(std::exp(std::log(c) + d * r.normal()));
    }
    if (v > (double)cap) v = (double)cap;
    if (v < (double)lo) v = (double)lo;
    out[i] = (uint64_t)v;
  });
  return 0;
}

// Reference bases.  Per-reference GC fraction uniform in [gc_lo, gc_hi].
// When seg_pool > 0, a pool of seg_pool segments with lengths in
// [seg_lo, seg_hi] is drawn, and a fraction seg_frac of the references each
// receive 1-3 of them written over random positions.
int mhg_refs(uint64_t n_refs, const uint64_t *off, double gc_lo, double gc_hi, uint64_t seg_pool,
             uint64_t seg_lo, uint64_t seg_hi, double seg_frac, uint64_t seed, uint8_t *bytes) {
  std::vector<std::vector<uint8_t>> pool(seg_pool);
  par(seg_pool, [&](uint64_t p) {
    Rng r(seed, 0x300000000ull + p);
    uint64_t len = seg_lo + r.below(seg_hi - seg_lo + 1);
    pool[p].resize(len);
    for (auto &x : pool[p]) x = kBases[r.next() >> 62];
  });
  par(n_refs, [&](uint64_t i) {
    Rng r(seed, 0x200000000ull + i);
    double gc = gc_lo + (gc_hi - gc_lo) * r.uniform();
    uint8_t *p = bytes + off[i];
    const uint64_t n = off[i + 1] - off[i];
    for (uint64_t j = 0; j < n; j++) p[j] = base_gc(r, gc);
    if (seg_pool && r.uniform() < seg_frac) {
      int ins = 1 + (int)r.below(3);
      for (int t = 0; t < ins; t++) {
        const std::vector<uint8_t> &sg = pool[r.below(seg_pool)];
        if (sg.size() >= n) continue;
        uint64_t at = r.below(n - sg.size() + 1);
        std::memcpy(p + at, sg.data(), sg.size());
      }
    }
  });
  return 0;
}

// Reads.  Each read draws a kind: exact substring (p_exact), uniform random
// (p_random), or sampled with errors.  Sampled reads walk a uniformly chosen
// reference from a uniform start, substituting each copied base with
// probability sub_rate and starting a 1-3 base insertion or deletion with
// probability indel_rate per base.  With probability rc_frac the read is
// reverse-complemented.  truth receives (ref, start, kind | strand << 8) per
// read, kind 0 sampled, 1 exact, 2 random; ref = UINT32_MAX for random reads.
int mhg_reads(const uint8_t *ref_bytes, const uint64_t *ref_off, uint64_t n_refs, uint64_t n_reads,
              const uint64_t *read_off, double p_exact, double p_random, double sub_rate,
              double indel_rate, double rc_frac, uint64_t seed, uint8_t *out, uint32_t *truth) {
  par(n_reads, [&](uint64_t i) {
    Rng r(seed, 0x400000000ull + i);
    uint8_t *q = out + read_off[i];
    const uint64_t m = read_off[i + 1] - read_off[i];
    const double u = r.uniform();
    uint32_t kind = u < p_exact ? 1 : (u < p_exact + p_random ? 2 : 0);
    uint32_t ref = UINT32_MAX, start = 0;
    if (kind == 2 || n_refs == 0) {
      kind = 2;
      for (uint64_t j = 0; j < m; j++) q[j] = kBases[r.next() >> 62];
    } else {
      uint64_t rid = r.below(n_refs), len = ref_off[rid + 1] - ref_off[rid];
      for (int t = 0; t < 64 && len < m + 8; t++) {
        rid = r.below(n_refs);
        len = ref_off[rid + 1] - ref_off[rid];
      }
      const uint8_t *src = ref_bytes + ref_off[rid];
      uint64_t st = len > m ? r.below(len - m + 1) : 0;
      ref = (uint32_t)rid;
      start = (uint32_t)st;
      uint64_t j = 0, p = st;
      if (kind == 1) {
        for (; j < m && p < len; j++, p++) q[j] = src[p];
      } else {
        while (j < m && p < len) {
          if (indel_rate > 0 && r.uniform() < indel_rate) {
            uint64_t el = 1 + r.below(3);
            if (r.uniform() < 0.5) {
              for (uint64_t t = 0; t < el && j < m; t++) q[j++] = kBases[r.next() >> 62];
            } else {
              p += el;
            }
            continue;
          }
          uint8_t b = src[p++];
          if (sub_rate > 0 && r.uniform() < sub_rate) {
            uint8_t nb;
            do nb = kBases[r.next() >> 62]; while (nb == b);
            b = nb;
          }
          q[j++] = b;
        }
      }
      for (; j < m; j++) q[j] = kBases[r.next() >> 62];  // reference ran out
    }
    uint32_t strand = 0;
    if (rc_frac > 0 && r.uniform() < rc_frac) {
      strand = 1;
      std::reverse(q, q + m);
      for (uint64_t j = 0; j < m; j++) {
        uint8_t c = q[j];
        q[j] = c == 'A' ? 'T' : c == 'T' ? 'A' : c == 'C' ? 'G' : c == 'G' ? 'C' : c;
      }
    }
    if (truth) {
      truth[3 * i] = ref;
      truth[3 * i + 1] = start;
      truth[3 * i + 2] = kind | (strand << 8);
    }
  });
  return 0;
}

}  // extern "C"
