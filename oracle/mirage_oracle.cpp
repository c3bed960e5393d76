/*
 * CPU oracle for MirAge's read-classification path (mirdesign/mirage).
 *
 * TEST INFRASTRUCTURE ONLY: a plain restatement of the reference's CPU
 * algorithm, used by tests/ as the checker, by __graft_entry__.smoke() and by
 * bench.py as the timed CPU baseline.  The product never links or calls it.
 *
 * Parity pins (each is a test under tests/):
 *   hash (a1)        FNV-1 32-bit known-answer vectors and the 28-bit fold;
 *                    H("N"*12) = 0x0995d49e; the reference's own CreateHash24
 *                    through oracle/ref_harness when oracle/_ref is built.
 *   index (a2,a3)    brute-force enumeration on tiny references; the
 *                    "N"-window cutoff; the fragment rule at T < 20;
 *                    entry-count invariant; bucket-for-bucket equality with
 *                    .mhm files the reference wrote (tests/golden).
 *   .mhm (f1)        oracle/mhm.py decodes those files and writes them back
 *                    byte for byte.
 *   candidates       brute-force transcription on tiny inputs plus
 *   (a5-a9)          end-to-end golden outputs written by the reference's own
 *                    binary (JSON fixtures and their generator in tests/golden).
 *   score (a10)      closed forms (one exact seed: R = L(L+1)/2, C = H = L),
 *                    random seeds against the reference's CalculateScore
 *                    through oracle/ref_harness.
 *   relative (a12)   the reference's GetRelativeScore through ref_harness, and
 *                    glibc logf on the same host.
 *   strand merge,    constructed ties and the reference binary's goldens.
 *   tie set (a13-14)
 *
 * Defined behaviour where the reference is undefined:
 *   - a read of length m <= T hashes past its end (Analyser.cpp:1459-1463 with
 *     the probe loop at :1474-1476); such reads get no candidates here.
 *   - a reference of length n <= k hashes past its end
 *     (Hashmap.cpp:2839-2841); it contributes no index entries here.
 *
 * The index equals, bucket for bucket, what the reference writes to its .mhm
 * file and loads back (checked against files the reference wrote).
 */
#include "mirage_oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <unordered_map>
#include <vector>

#include <immintrin.h>

namespace {

constexpr uint32_t kHashBits = 28;                  // Hashmap.h:111
constexpr uint32_t kHashMask = (1u << kHashBits) - 1u;

/* Sequences::SequenceToUpperCase (Sequences.cpp:43-52): char is signed, so
 * only bytes 97..127 are shifted. */
inline uint8_t to_upper(uint8_t c) {
  return (int8_t)c > 96 ? (uint8_t)(c - 32) : c;
}

/* Complement table of Sequences::reverseComplementDNA (Sequences.cpp:800-822);
 * 0 marks a byte the reference rejects with invalid_argument. */
inline uint8_t complement(uint8_t c) {
  switch (c) {
    case 'A': return 'T';
    case 'T': return 'A';
    case 'U': return 'A';
    case 'G': return 'C';
    case 'C': return 'G';
    case 'Y': return 'R';
    case 'R': return 'Y';
    case 'S': return 'S';
    case 'W': return 'W';
    case 'K': return 'M';
    case 'M': return 'K';
    case 'B': return 'V';
    case 'D': return 'H';
    case 'H': return 'D';
    case 'V': return 'B';
    case 'N': return 'N';
    case '-': return '-';
    case '~': return '~';
    default: return 0;
  }
}

struct Entry {
  uint32_t hash, ref, pos;
};

}  // namespace

/* Hashmap::CreateHash32 (Hashmap.cpp:3197-3211): FNV-1, multiply then xor. */
extern "C" uint32_t mo_hash32(const uint8_t *p, int32_t len) {
  uint32_t h = 2166136261u;
  for (int32_t i = 0; i < len; i++) h = (h * 16777619u) ^ p[i];
  return h;
}

/* Hashmap::CreateHash24 (Hashmap.cpp:3233-3263) with HashBits = 28. */
extern "C" uint32_t mo_hash28(const uint8_t *p, int32_t len) {
  uint32_t h = mo_hash32(p, len);
  return ((h >> kHashBits) ^ h) & kHashMask;
}

struct mo_index {
  int32_t threshold = 0, stride = 0, k = 0;
  std::vector<uint8_t> bytes;        // uppercased reference bytes
  std::vector<uint64_t> ref_off;     // n_refs + 1
  std::vector<uint64_t> cutoff;      // windows [0, cutoff) indexed per ref
  std::vector<Entry> entries;        // sorted by (hash, ref, pos)
  std::vector<uint32_t> keys;        // distinct present hashes, ascending
  std::vector<uint64_t> key_begin;   // keys.size()+1 offsets into entries
  std::vector<uint32_t> dir;         // optional direct table: hash -> key idx+1
};

namespace {

int hw_threads(int32_t n) {
  if (n > 0) return n;
  unsigned h = std::thread::hardware_concurrency();
  return h ? (int)h : 1;
}

template <class F>
void parallel_for(uint64_t n, int threads, F f) {
  if (threads <= 1 || n < 2) {
    for (uint64_t i = 0; i < n; i++) f(i, 0);
    return;
  }
  std::atomic<uint64_t> next(0);
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; t++)
    pool.emplace_back([&, t] {
      for (;;) {
        uint64_t i = next.fetch_add(1);
        if (i >= n) break;
        f(i, t);
      }
    });
  for (auto &th : pool) th.join();
}

/* Lookup: returns -1 when absent, else the key index. */
int64_t find_key(const mo_index &ix, uint32_t h) {
  if (!ix.dir.empty()) return (int64_t)ix.dir[h] - 1;
  auto it = std::lower_bound(ix.keys.begin(), ix.keys.end(), h);
  if (it == ix.keys.end() || *it != h) return -1;
  return it - ix.keys.begin();
}

}  // namespace

/*
 * Index build.  Hashmap::TBuildHashmapForReference (Hashmap.cpp:2771-2976):
 * stride-1 windows of length k = T - T/2 + 1 over each reference, window
 * positions i in [0, n-k) when k < n.  A window whose hash equals H("N"*k) is
 * dropped without advancing the hashed pointer (:2882-2895), so every later
 * window of that reference hashes the same bytes and is dropped as well: the
 * reference is indexed only up to its first such window.  The dropped
 * windows still count in the size passed to ConvertHashesFromThreadToHashmap
 * (:1233-1239, :1252), so bucket 0 gains one (ref 0, pos 0) entry per
 * dropped window (the value-initialised tail of the build buffer, :1026).
 * ConvertHashesFromThreadToHashmap (:745-822) groups by hash; the fragment
 * rule of SaveHashmapTableV3 (:1759-1809, :1831-1833) drops buckets larger
 * than min(20, sizes[floor(0.30 * #buckets)]) when T < 20.  With
 * MO_FLAG_LAST_BUCKET_EMPTY the bucket with the highest hash keeps its key but
 * no entries, as LoadHashMapFromHashmapDB (:540-554) leaves its BufferSize at
 * 0 and the decode loop (Analyser.cpp:1613-1614) then skips it.
 */
extern "C" mo_index *mo_index_build(const uint8_t *ref_bytes,
                                    const uint64_t *ref_off, uint32_t n_refs,
                                    int32_t threshold, uint32_t flags) {
  if (threshold < 2 || (!ref_bytes && n_refs) || !ref_off) return nullptr;
  mo_index *ix = new mo_index();
  ix->threshold = threshold;
  ix->stride = threshold / 2;                     // Hashmap.cpp:2789
  ix->k = threshold - ix->stride + 1;             // Hashmap.cpp:2790
  const int k = ix->k;
  const uint64_t total = ref_off[n_refs] - ref_off[0];
  ix->bytes.resize(total);
  for (uint64_t i = 0; i < total; i++)
    ix->bytes[i] = to_upper(ref_bytes[ref_off[0] + i]);
  ix->ref_off.resize(n_refs + 1);
  for (uint32_t r = 0; r <= n_refs; r++) ix->ref_off[r] = ref_off[r] - ref_off[0];

  std::vector<uint8_t> nnn(k, 'N');
  const uint32_t h_nnn = mo_hash28(nnn.data(), k);  // Hashmap.cpp:2834

  const int threads = hw_threads(0);
  ix->cutoff.assign(n_refs, 0);
  std::vector<uint64_t> nwin(n_refs + 1, 0);
  parallel_for(n_refs, threads, [&](uint64_t r, int) {
    const uint64_t n = ix->ref_off[r + 1] - ix->ref_off[r];
    if (n <= (uint64_t)k) return;                 // undefined in the reference
    const uint64_t stop = n - k;                  // Hashmap.cpp:2839-2841
    const uint8_t *p = ix->bytes.data() + ix->ref_off[r];
    uint64_t i = 0;
    for (; i < stop; i++)
      if (mo_hash28(p + i, k) == h_nnn) break;
    ix->cutoff[r] = i;
  });
  for (uint32_t r = 0; r < n_refs; r++) nwin[r + 1] = nwin[r] + ix->cutoff[r];
  // Windows dropped after a reference's cutoff still count towards the size
  // handed to ConvertHashesFromThreadToHashmap (Hashmap.cpp:1233-1239,
  // :1252), whose tail past the filled entries holds value-initialised
  // records (resize at :1026): one (ref 0, pos 0) entry of hash 0 each.
  uint64_t phantom = 0;
  for (uint32_t r = 0; r < n_refs; r++) {
    const uint64_t n = ix->ref_off[r + 1] - ix->ref_off[r];
    if (n > (uint64_t)k) phantom += (n - k) - ix->cutoff[r];
  }
  const uint64_t n_entries = nwin[n_refs] + phantom;

  // Partition by the top 8 hash bits, then sort each partition.
  constexpr int kParts = 256;
  std::vector<std::vector<uint64_t>> cnt(threads, std::vector<uint64_t>(kParts, 0));
  std::vector<Entry> tmp(n_entries);
  parallel_for(n_refs, threads, [&](uint64_t r, int) {
    const uint8_t *p = ix->bytes.data() + ix->ref_off[r];
    Entry *out = tmp.data() + nwin[r];
    for (uint64_t i = 0; i < ix->cutoff[r]; i++)
      out[i] = Entry{mo_hash28(p + i, k), (uint32_t)r, (uint32_t)i};
  });
  for (uint64_t i = nwin[n_refs]; i < n_entries; i++) tmp[i] = Entry{0u, 0u, 0u};
  std::vector<uint64_t> part_cnt(kParts + 1, 0);
  for (uint64_t i = 0; i < n_entries; i++) part_cnt[(tmp[i].hash >> 20) + 1]++;
  for (int b = 0; b < kParts; b++) part_cnt[b + 1] += part_cnt[b];
  ix->entries.resize(n_entries);
  {
    std::vector<uint64_t> cur(part_cnt.begin(), part_cnt.end() - 1);
    for (uint64_t i = 0; i < n_entries; i++) ix->entries[cur[tmp[i].hash >> 20]++] = tmp[i];
  }
  std::vector<Entry>().swap(tmp);
  parallel_for(kParts, threads, [&](uint64_t b, int) {
    std::sort(ix->entries.begin() + part_cnt[b], ix->entries.begin() + part_cnt[b + 1],
              [](const Entry &a, const Entry &c) {
                if (a.hash != c.hash) return a.hash < c.hash;
                if (a.ref != c.ref) return a.ref < c.ref;
                return a.pos < c.pos;
              });
  });

  // Group into buckets.
  std::vector<uint32_t> keys;
  std::vector<uint64_t> begin;
  for (uint64_t i = 0; i < n_entries; i++)
    if (i == 0 || ix->entries[i].hash != ix->entries[i - 1].hash) {
      keys.push_back(ix->entries[i].hash);
      begin.push_back(i);
    }
  begin.push_back(n_entries);

  // Fragment rule (T < 20).
  std::vector<char> keep(keys.size(), 1);
  if (threshold < 20 && !keys.empty()) {
    std::vector<uint64_t> sizes(keys.size());
    for (size_t b = 0; b < keys.size(); b++) sizes[b] = begin[b + 1] - begin[b];
    std::sort(sizes.begin(), sizes.end());
    size_t frag = (size_t)((sizes.size() / 100.0) * 30);  // Hashmap.cpp:1801
    uint64_t thr = std::min<uint64_t>(20, sizes[frag]);    // :1802,1808
    for (size_t b = 0; b < keys.size(); b++)
      if (begin[b + 1] - begin[b] > thr) keep[b] = 0;      // :1831-1833
  }
  if (threshold < 20) {
    std::vector<Entry> kept;
    std::vector<uint32_t> k2;
    std::vector<uint64_t> b2;
    for (size_t b = 0; b < keys.size(); b++) {
      if (!keep[b]) continue;
      k2.push_back(keys[b]);
      b2.push_back(kept.size());
      kept.insert(kept.end(), ix->entries.begin() + begin[b], ix->entries.begin() + begin[b + 1]);
    }
    b2.push_back(kept.size());
    ix->entries.swap(kept);
    keys.swap(k2);
    begin.swap(b2);
  }
  ix->keys.swap(keys);
  ix->key_begin.swap(begin);
  if ((flags & MO_FLAG_LAST_BUCKET_EMPTY) && !ix->keys.empty()) {
    // The last key keeps its place; its entries are dropped from view.
    size_t b = ix->keys.size() - 1;
    ix->entries.resize(ix->key_begin[b]);
    ix->key_begin[b + 1] = ix->key_begin[b];
  }
  if (ix->entries.size() > (1u << 22)) {
    ix->dir.assign((size_t)1 << kHashBits, 0);
    for (size_t b = 0; b < ix->keys.size(); b++) ix->dir[ix->keys[b]] = (uint32_t)(b + 1);
  }
  return ix;
}

namespace {

/* FNV-1 28-bit hashes of the 8 windows p[0..k), p[1..k+1), ..., p[7..k+7):
 * the same arithmetic as mo_hash28, eight independent chains at once.  With
 * AVX2 the chains are the lanes of one register (the 8 bytes p[j..j+7] are
 * zero-extended into them); otherwise plain loops.  Reads p[0..k+7). */
__attribute__((target("avx2"))) void hash28_x8_avx2(const uint8_t *p, int k, uint32_t *out) {
  __m256i h = _mm256_set1_epi32((int)2166136261u);
  const __m256i prime = _mm256_set1_epi32(16777619);
  for (int j = 0; j < k; j++) {
    uint64_t w;
    std::memcpy(&w, p + j, 8);
    const __m256i b = _mm256_cvtepu8_epi32(_mm_cvtsi64_si128((long long)w));
    h = _mm256_xor_si256(_mm256_mullo_epi32(h, prime), b);
  }
  const __m256i f = _mm256_and_si256(_mm256_xor_si256(_mm256_srli_epi32(h, kHashBits), h),
                                     _mm256_set1_epi32((int)kHashMask));
  _mm256_storeu_si256((__m256i *)out, f);
}

void hash28_x8_plain(const uint8_t *p, int k, uint32_t *out) {
  uint32_t h[8];
  for (int t = 0; t < 8; t++) h[t] = 2166136261u;
  for (int j = 0; j < k; j++)
    for (int t = 0; t < 8; t++) h[t] = (h[t] * 16777619u) ^ p[j + t];
  for (int t = 0; t < 8; t++) out[t] = ((h[t] >> kHashBits) ^ h[t]) & kHashMask;
}

}  // namespace

/*
 * The index restricted to the buckets a read set probes: for every hash that
 * some probe window of the reads (both strands, windows j = 0, s, 2s, ... <
 * m - T, Analyser.cpp:1459-1476) produces, the bucket holds exactly what
 * mo_index_build gives it; every other hash is absent.  Classifying those
 * reads (and only those) then gives the same results as the full index,
 * without holding the entries of buckets no probe reads.  The reference
 * windows are scanned as in mo_index_build, including the "N"-window cutoff,
 * bucket 0's (ref 0, pos 0) entries for the windows dropped after a cutoff,
 * and, with MO_FLAG_LAST_BUCKET_EMPTY, the emptied bucket of the highest hash
 * of the whole index (kept only if a probe reads it).  The fragment rule
 * needs every bucket's size, so T < 20 is refused (NULL).
 */
extern "C" mo_index *mo_index_build_for_reads(const uint8_t *ref_bytes, const uint64_t *ref_off,
                                              uint32_t n_refs, int32_t threshold, uint32_t flags,
                                              const uint8_t *reads, const uint64_t *read_off,
                                              uint32_t n_reads) {
  if (threshold < 20 || (!ref_bytes && n_refs) || !ref_off || (!reads && n_reads) ||
      (!read_off && n_reads))
    return nullptr;
  mo_index *ix = new mo_index();
  ix->threshold = threshold;
  ix->stride = threshold / 2;                     // Hashmap.cpp:2789
  ix->k = threshold - ix->stride + 1;             // Hashmap.cpp:2790
  const int k = ix->k;
  const uint32_t T = (uint32_t)threshold, s = (uint32_t)ix->stride;
  const int threads = hw_threads(0);

  // Hashes the reads probe, as a 2^28-bit set.
  std::vector<std::atomic<uint64_t>> probed((size_t)1 << (kHashBits - 6));
  for (auto &w : probed) w.store(0, std::memory_order_relaxed);
  std::atomic<int> bad(0);
  parallel_for(n_reads, threads, [&](uint64_t r, int) {
    const uint64_t m = read_off[r + 1] - read_off[r];
    if (m <= T) return;
    std::vector<uint8_t> q(2 * m);
    const uint8_t *rd = reads + read_off[r];
    for (uint64_t i = 0; i < m; i++) q[i] = to_upper(rd[i]);
    for (uint64_t i = 0; i < m; i++) {
      const uint8_t c = complement(q[m - 1 - i]);
      if (!c) {
        bad.store(1);
        return;
      }
      q[m + i] = c;
    }
    for (int strand = 0; strand < 2; strand++)
      for (uint64_t j = 0; j < m - T; j += s) {
        const uint32_t h = mo_hash28(q.data() + strand * m + j, k);
        probed[h >> 6].fetch_or(1ull << (h & 63), std::memory_order_relaxed);
      }
  });
  auto is_probed = [&](uint32_t h) {
    return (probed[h >> 6].load(std::memory_order_relaxed) >> (h & 63)) & 1;
  };

  const uint64_t total = ref_off[n_refs] - ref_off[0];
  ix->bytes.resize(total);
  parallel_for((total + (1u << 24) - 1) >> 24, threads, [&](uint64_t c, int) {
    const uint64_t a = c << 24, b = std::min<uint64_t>(total, a + (1u << 24));
    for (uint64_t i = a; i < b; i++) ix->bytes[i] = to_upper(ref_bytes[ref_off[0] + i]);
  });
  ix->ref_off.resize(n_refs + 1);
  for (uint32_t r = 0; r <= n_refs; r++) ix->ref_off[r] = ref_off[r] - ref_off[0];

  std::vector<uint8_t> nnn(k, 'N');
  const uint32_t h_nnn = mo_hash28(nnn.data(), k);  // Hashmap.cpp:2834
  ix->cutoff.assign(n_refs, 0);
  const bool avx2 = __builtin_cpu_supports("avx2");
  // Per thread: kept entries and the largest hash of any indexed window.
  std::vector<std::vector<Entry>> kept(threads);
  std::vector<uint32_t> top(threads, 0);
  std::vector<char> any(threads, 0);
  parallel_for(n_refs, threads, [&](uint64_t r, int t) {
    const uint64_t n = ix->ref_off[r + 1] - ix->ref_off[r];
    if (n <= (uint64_t)k) return;                 // undefined in the reference
    const uint64_t stop = n - k;                  // Hashmap.cpp:2839-2841
    const uint8_t *p = ix->bytes.data() + ix->ref_off[r];
    uint32_t hv[8];
    uint64_t i = 0;
    bool cut = false;
    uint32_t top_r = 0;  // kept in registers: the per-thread slots share cache lines
    std::vector<Entry> &out = kept[t];
    while (i < stop && !cut) {
      const uint64_t nb = std::min<uint64_t>(8, stop - i);
      if (nb == 8) {  // windows i..i+7 end at i + 7 + k <= stop - 1 + k < n
        if (avx2) hash28_x8_avx2(p + i, k, hv);
        else hash28_x8_plain(p + i, k, hv);
      } else {
        for (uint64_t t2 = 0; t2 < nb; t2++) hv[t2] = mo_hash28(p + i + t2, k);
      }
      for (uint64_t t2 = 0; t2 < nb; t2++) {
        const uint32_t h = hv[t2];
        if (h == h_nnn) {  // the reference stops indexing this one here
          cut = true;
          break;
        }
        top_r = std::max(top_r, h);
        if (is_probed(h)) out.push_back(Entry{h, (uint32_t)r, (uint32_t)i});
        i++;  // window i + 1 is hv[t2 + 1]
      }
    }
    ix->cutoff[r] = i;
    if (i > 0) {
      any[t] = 1;
      top[t] = std::max(top[t], top_r);
    }
  });
  uint64_t phantom = 0;
  for (uint32_t r = 0; r < n_refs; r++) {
    const uint64_t n = ix->ref_off[r + 1] - ix->ref_off[r];
    if (n > (uint64_t)k) phantom += (n - k) - ix->cutoff[r];
  }
  bool have = phantom > 0;
  uint32_t highest = 0;  // the highest hash of the full index
  for (int t = 0; t < threads; t++)
    if (any[t]) {
      highest = have ? std::max(highest, top[t]) : top[t];
      have = true;
    }
  size_t n_kept = 0;
  for (auto &v : kept) n_kept += v.size();
  const bool zero = is_probed(0) && phantom > 0;
  ix->entries.reserve(n_kept + (zero ? phantom : 0));
  for (auto &v : kept) {
    ix->entries.insert(ix->entries.end(), v.begin(), v.end());
    std::vector<Entry>().swap(v);
  }
  if (zero)
    for (uint64_t i = 0; i < phantom; i++) ix->entries.push_back(Entry{0u, 0u, 0u});
  std::sort(ix->entries.begin(), ix->entries.end(), [](const Entry &a, const Entry &c) {
    if (a.hash != c.hash) return a.hash < c.hash;
    if (a.ref != c.ref) return a.ref < c.ref;
    return a.pos < c.pos;
  });
  for (uint64_t i = 0; i < ix->entries.size(); i++)
    if (i == 0 || ix->entries[i].hash != ix->entries[i - 1].hash) {
      ix->keys.push_back(ix->entries[i].hash);
      ix->key_begin.push_back(i);
    }
  ix->key_begin.push_back(ix->entries.size());
  if ((flags & MO_FLAG_LAST_BUCKET_EMPTY) && have && !ix->keys.empty() &&
      ix->keys.back() == highest) {
    // The highest hash of the whole index keeps its key, not its entries.
    const size_t b = ix->keys.size() - 1;
    ix->entries.resize(ix->key_begin[b]);
    ix->key_begin[b + 1] = ix->key_begin[b];
  }
  if (ix->entries.size() > (1u << 22)) {
    ix->dir.assign((size_t)1 << kHashBits, 0);
    for (size_t b = 0; b < ix->keys.size(); b++) ix->dir[ix->keys[b]] = (uint32_t)(b + 1);
  }
  (void)bad;
  return ix;
}

extern "C" void mo_index_free(mo_index *ix) { delete ix; }
extern "C" uint64_t mo_index_num_entries(const mo_index *ix) { return ix->entries.size(); }
extern "C" uint64_t mo_index_num_buckets(const mo_index *ix) { return ix->keys.size(); }
extern "C" uint64_t mo_index_ref_cutoff(const mo_index *ix, uint32_t r) {
  return r < ix->cutoff.size() ? ix->cutoff[r] : 0;
}

extern "C" int64_t mo_index_bucket(const mo_index *ix, uint32_t hash,
                                   uint32_t *refs, uint32_t *pos, uint64_t cap) {
  int64_t b = find_key(*ix, hash);
  if (b < 0) return -1;
  uint64_t lo = ix->key_begin[b], hi = ix->key_begin[b + 1];
  for (uint64_t i = lo; i < hi && i - lo < cap; i++) {
    if (refs) refs[i - lo] = ix->entries[i].ref;
    if (pos) pos[i - lo] = ix->entries[i].pos;
  }
  return (int64_t)(hi - lo);
}

namespace {

struct Seed {
  uint32_t sref, sq, lref, lq;  // start (Reference, Query), last (Reference, Query)
};

struct Pair {
  uint32_t ref, rpos, qpos;     // FoundRefID_Pos_Pairs element (Analyser.cpp:2187)
};

struct Work {
  std::vector<int> occ, occ_max, last;  // InitialiseOccureBuffer (Analyser.cpp:2619-2644)
  std::vector<uint32_t> hit_j;
  std::vector<int64_t> hit_key;
  std::vector<Pair> pairs;
  std::vector<uint8_t> row_hit;
  std::vector<uint8_t> query;
  std::vector<uint32_t> cand_refs;
  std::vector<std::vector<Seed>> cand_seeds;
  mo_work counters{};
};

/*
 * Candidate generation for one query (one strand of one read):
 * Analyser::TPrepareGPUAnalysis, Analyser.cpp:1317-2587, in Extreme memory
 * mode with a single index file.
 *   probe            :1453-1520  hash every s-th window, j < iStop
 *   successive       :2004-2197  longest run of consecutive hit indices per ref
 *   sort and group   :2214-2315  (ref, refPos, qPos) order, distinct refs
 *   chains           :2329-2490  diagonal links, max chain G, seed emit
 * cand_refs receives every distinct reference of the kept pairs; cand_seeds
 * the seeds of each (empty for references whose chain maximum is below G).
 */
void gen_candidates(const mo_index &ix, const uint8_t *q, uint64_t m, Work &w) {
  w.cand_refs.clear();
  w.cand_seeds.clear();
  const uint32_t T = (uint32_t)ix.threshold;
  const uint32_t s = (uint32_t)ix.stride;
  const int k = ix.k;
  if (m <= T) return;  // defined here: the reference hashes past the read end
  const uint64_t i_stop = m - T;                       // :1459-1463

  // Probe.
  w.hit_j.clear();
  w.hit_key.clear();
  for (uint64_t j = 0; j < i_stop; j += s) {           // :1474-1476, :1498-1517
    w.counters.probes++;
    int64_t b = find_key(ix, mo_hash28(q + j, k));
    if (b < 0) continue;
    w.counters.hits++;
    w.hit_j.push_back((uint32_t)j);
    w.hit_key.push_back(b);
  }
  const size_t n_hits = w.hit_j.size();
  if (n_hits == 0) return;                             // :1989-1990

  // Reference-id range over the hits (:1994-2009).
  uint32_t lo = UINT32_MAX, hi = 0;
  bool any = false;
  for (size_t i = 0; i < n_hits; i++) {
    uint64_t b0 = ix.key_begin[w.hit_key[i]], b1 = ix.key_begin[w.hit_key[i] + 1];
    w.counters.entries += b1 - b0;
    for (uint64_t e = b0; e < b1; e++) {
      lo = std::min(lo, ix.entries[e].ref);
      hi = std::max(hi, ix.entries[e].ref);
      any = true;
    }
  }
  if (!any) return;                                    // :2012-2013
  const size_t range = (size_t)(hi - lo) + 1;
  if (w.occ.size() < range) {
    w.occ.resize(range);
    w.occ_max.resize(range);
    w.last.resize(range);
  }
  std::fill(w.occ.begin(), w.occ.begin() + range, 0);
  std::fill(w.occ_max.begin(), w.occ_max.begin() + range, 0);
  std::fill(w.last.begin(), w.last.begin() + range, -1);

  // Successive occurrence (:2034-2095).
  uint64_t highest = 0;
  for (size_t i = 0; i < n_hits; i++) {
    uint64_t b0 = ix.key_begin[w.hit_key[i]], b1 = ix.key_begin[w.hit_key[i] + 1];
    for (uint64_t e = b0; e < b1; e++) {
      const uint32_t r = ix.entries[e].ref - lo;
      if (i == 0) {
        w.occ[r] = 1;
        w.occ_max[r] = 1;
        w.last[r] = 0;
        continue;
      }
      if (w.last[r] == (int)i) continue;
      if (w.last[r] == (int)i - 1) {
        w.occ[r]++;
        w.occ_max[r] = std::max(w.occ_max[r], w.occ[r]);
        highest = std::max<uint64_t>(highest, (uint64_t)w.occ_max[r]);
      } else {
        w.occ[r] = 1;
        w.occ_max[r] = std::max(w.occ_max[r], 1);
      }
      w.last[r] = (int)i;
    }
  }
  const uint64_t cut = std::max<uint64_t>(1, highest != 0 ? highest - 1 : 0);  // :2099

  // Keep pairs of references whose longest run reaches the cut (:2115-2191).
  w.pairs.clear();
  for (size_t i = 0; i < n_hits; i++) {
    uint64_t b0 = ix.key_begin[w.hit_key[i]], b1 = ix.key_begin[w.hit_key[i] + 1];
    for (uint64_t e = b0; e < b1; e++) {
      const Entry &en = ix.entries[e];
      if ((uint64_t)w.occ_max[en.ref - lo] >= cut)
        w.pairs.push_back(Pair{en.ref, en.pos, w.hit_j[i]});
    }
  }
  if (w.pairs.empty()) return;                         // :2206-2207

  // Sort by (ref, refPos, qPos) (:1024-1059).
  std::sort(w.pairs.begin(), w.pairs.end(), [](const Pair &a, const Pair &b) {
    if (a.ref != b.ref) return a.ref < b.ref;
    if (a.rpos != b.rpos) return a.rpos < b.rpos;
    return a.qpos < b.qpos;
  });

  // Distinct references, in order (:2290-2308).
  std::vector<size_t> seg;  // segment starts
  for (size_t t = 0; t < w.pairs.size(); t++)
    if (t == 0 || w.pairs[t].ref != w.pairs[t - 1].ref) {
      seg.push_back(t);
      w.cand_refs.push_back(w.pairs[t].ref);
    }
  seg.push_back(w.pairs.size());
  const size_t n_cand = w.cand_refs.size();
  w.cand_seeds.assign(n_cand, {});

  // Longest diagonal chain per reference (:2335-2385).  Neighbours link when
  // both coordinates step by exactly s (u32 arithmetic, :2351-2353).
  auto links = [&](const Pair &a, const Pair &b) {
    return a.rpos == (uint32_t)(b.rpos - s) && a.qpos == (uint32_t)(b.qpos - s);
  };
  std::vector<uint64_t> chain_max(n_cand, 0);
  uint64_t G = 0;
  for (size_t c = 0; c < n_cand; c++) {
    uint64_t best = 0, cur = 1;
    for (size_t t = seg[c] + 1; t < seg[c + 1]; t++) {
      if (!links(w.pairs[t - 1], w.pairs[t])) {
        best = std::max(best, cur);
        cur = 0;
      }
      cur++;
    }
    best = std::max(best, cur);
    chain_max[c] = best;
    G = std::max(G, best);
  }
  if (G < 2) return;                                   // :2398-2403

  // Seeds for references whose chain maximum equals G (:2421-2481).
  for (size_t c = 0; c < n_cand; c++) {
    if (chain_max[c] < G || chain_max[c] < 2) continue;
    std::vector<Seed> &sv = w.cand_seeds[c];
    size_t first = seg[c];
    for (size_t t = seg[c] + 1; t <= seg[c + 1]; t++) {
      if (t == seg[c + 1] || !links(w.pairs[t - 1], w.pairs[t])) {
        const Pair &a = w.pairs[first], &z = w.pairs[t - 1];
        sv.push_back(Seed{a.rpos, a.qpos, z.rpos + s, z.qpos + s});
        first = t;
      }
    }
  }
}

struct ScoreOut {
  float score;
  uint64_t sum_r, sum_c;
  uint32_t hits;
};

/*
 * MatrixCPU::CalculateScore (MatrixCPU.cpp:18-214), Extreme memory mode.
 *   tail   :93-112   along the seed diagonal from last, at most T bases
 *   head   :116-172  from query row start-1 against reference byte
 *                    start.Reference (one base off the diagonal), at most
 *                    min(T, start.Reference) bases; the base that meets a
 *                    boundary is not counted
 *   accumulate :175-184, score :195-213.
 */
ScoreOut score_seeds(const uint8_t *q, uint64_t m, const uint8_t *ref,
                     uint64_t ref_len, const Seed *seeds, size_t n_seeds,
                     uint32_t T, bool ultrafast, std::vector<uint8_t> &row_hit,
                     mo_work *ctr) {
  ScoreOut o{-2.0f, 0, 0, 0};
  if (n_seeds == 0) return o;                          // :24-26
  row_hit.assign(m, 0);
  uint64_t sum_r = 0, sum_c = 0;
  for (size_t c = 0; c < n_seeds; c++) {
    const Seed &sd = seeds[c];
    const uint64_t row_start = sd.sq, row_end = sd.lq;
    if (row_start > row_end) continue;                 // :57-58
    const uint64_t n_mid = row_end - row_start;
    uint64_t n_head = 0, n_tail = 0;
    if (!ultrafast) {
      // Tail.
      const uint64_t tail_size = std::min<uint64_t>(T, ref_len - (uint64_t)sd.lref);
      uint64_t row = row_end, rp = sd.lref, idx = 0;
      while (true) {
        if (row >= m || rp >= ref_len || idx >= tail_size) break;
        if (ctr) ctr->ext_compares++;
        if (q[row] != ref[rp]) break;
        row++, rp++, idx++;
      }
      n_tail = idx;
      // Head.
      if (row_start > 0) {
        const uint64_t head_size = std::min<uint64_t>(T, (uint64_t)sd.sref);
        if (head_size > 0) {
          uint64_t qrow = row_start - 1, rp2 = sd.sref, hidx = head_size - 1, rev = 1;
          while (true) {
            if (ctr) ctr->ext_compares++;
            if (q[qrow] != ref[rp2]) {
              n_head += rev - 1;
              break;
            }
            if (qrow == 0 || rp2 == 0 || hidx == 0) {
              n_head += rev - 1;
              break;
            }
            qrow--, rp2--, hidx--, rev++;
          }
        }
      }
    }
    const uint64_t total = n_tail + n_head + n_mid;
    if (total >= T) {
      sum_c += total;
      std::memset(row_hit.data() + (row_start - n_head), 1, total);
      sum_r += (total * (total + 1)) / 2;
    }
  }
  o.sum_r = sum_r;
  o.sum_c = sum_c;
  if (sum_c > 0) {
    uint32_t h = 0;
    for (uint64_t i = 0; i < m; i++) h += row_hit[i] ? 1 : 0;
    o.hits = h;
    const float sf = (float)sum_r / (float)sum_c;                          // :209
    const float pct = (float)((100.0 / (double)(float)m) * (double)(float)h);  // :211
    o.score = (float)(((double)sf / 100.0) * (double)pct);                  // :212
  }
  return o;
}

float relative_score(float query_size, float ref_size, float score) {
  // MatrixResult::GetRelativeScore (MatrixResult.cpp:21-29); log(float)
  // resolves to the float overload, i.e. logf.
  const float len_max = (float)(query_size < ref_size ? query_size : ref_size);
  const float score_max =
      (float)((((double)len_max * ((double)len_max + 1.0)) / 2.0) / (double)len_max + 1);
  score += 1;
  return logf(score) / logf(score_max);
}

struct StrandResult {
  float highest;                 // CreateResultsScores (MatrixCPU.cpp:333-369)
  std::vector<uint32_t> refs;    // candidates with Score > -1, in order
  std::vector<float> rels;
};

/* One strand: candidates, scores (MatrixCPU.cpp:248-318), relative scores
 * and the first strict maximum (MatrixCPU.cpp:323-370). */
void classify_strand(const mo_index &ix, const uint8_t *q, uint64_t m, bool ultrafast,
                     uint32_t read, uint32_t strand, Work &w, StrandResult &res,
                     std::vector<mo_cand> *cands) {
  gen_candidates(ix, q, m, w);
  res.highest = -1;
  res.refs.clear();
  res.rels.clear();
  for (size_t c = 0; c < w.cand_refs.size(); c++) {
    const std::vector<Seed> &sv = w.cand_seeds[c];
    if (sv.empty()) continue;  // Score stays -2 (MatrixCPU.cpp:21-26)
    const uint32_t r = w.cand_refs[c];
    const uint64_t rl = ix.ref_off[r + 1] - ix.ref_off[r];
    ScoreOut so = score_seeds(q, m, ix.bytes.data() + ix.ref_off[r], rl, sv.data(), sv.size(),
                              (uint32_t)ix.threshold, ultrafast, w.row_hit, &w.counters);
    w.counters.seeds += sv.size();
    w.counters.cands++;
    float rel = -1;
    if (so.score > -1) {
      rel = relative_score((float)m, (float)rl, so.score);
      if (rel > res.highest) res.highest = rel;
      res.refs.push_back(r);
      res.rels.push_back(rel);
    }
    if (cands)
      cands->push_back(mo_cand{read, strand, r, so.hits, so.sum_r, so.sum_c, so.score, rel,
                               (uint32_t)sv.size(), 0});
  }
}

int classify_read(const mo_index &ix, const uint8_t *read, uint64_t m, uint32_t rid,
                  uint32_t flags, Work &w, mo_read_result &out, std::vector<uint32_t> &top,
                  std::vector<mo_cand> *cands) {
  const bool ultrafast = (flags & MO_FLAG_ULTRAFAST) != 0;
  // Query bytes are uppercased on load (Sequences.cpp:127,175) and the
  // reverse complement is appended as a second query (Sequences.cpp:843-885).
  w.query.resize(2 * m);
  for (uint64_t i = 0; i < m; i++) w.query[i] = to_upper(read[i]);
  for (uint64_t i = 0; i < m; i++) {
    uint8_t c = complement(w.query[m - 1 - i]);
    if (!c) return -1;
    w.query[m + i] = c;
  }
  StrandResult fwd, rc;
  classify_strand(ix, w.query.data(), m, ultrafast, rid, 0, w, fwd, cands);
  classify_strand(ix, w.query.data() + m, m, ultrafast, rid, 1, w, rc, cands);
  // Strand merge (Analyser.cpp:592-611): reverse wins only when strictly higher.
  const StrandResult &win = (fwd.highest < rc.highest) ? rc : fwd;
  out.highest = win.highest;
  out.strand = (&win == &rc) ? 1u : 0u;
  out.n_cand = (uint32_t)win.refs.size();
  top.clear();
  // Reported only when Scores is non-empty and HighestScore != 0
  // (Analyser.cpp:388); the group is every reference whose relative score
  // equals the top one (Analyser.cpp:401-424 with output_firstresults_only).
  if (!win.refs.empty() && win.highest != 0) {
    for (size_t i = 0; i < win.refs.size(); i++)
      if (win.rels[i] == win.highest) top.push_back(win.refs[i]);
    std::sort(top.begin(), top.end());
  }
  out.n_top = (uint32_t)top.size();
  return 0;
}

}  // namespace

extern "C" float mo_relative_score(float query_size, float ref_size, float score) {
  return relative_score(query_size, ref_size, score);
}

extern "C" float mo_score_seeds(const uint8_t *query, uint64_t qlen, const uint8_t *ref,
                                uint64_t ref_len, const uint32_t *seeds, uint64_t n_seeds,
                                int32_t threshold, uint32_t flags, uint64_t *sum_r,
                                uint64_t *sum_c, uint32_t *hits) {
  std::vector<Seed> sv(n_seeds);
  for (uint64_t i = 0; i < n_seeds; i++)
    sv[i] = Seed{seeds[4 * i], seeds[4 * i + 1], seeds[4 * i + 2], seeds[4 * i + 3]};
  std::vector<uint8_t> rh;
  ScoreOut o = score_seeds(query, qlen, ref, ref_len, sv.data(), sv.size(), (uint32_t)threshold,
                           (flags & MO_FLAG_ULTRAFAST) != 0, rh, nullptr);
  if (sum_r) *sum_r = o.sum_r;
  if (sum_c) *sum_c = o.sum_c;
  if (hits) *hits = o.hits;
  return o.score;
}

extern "C" int32_t mo_classify(const mo_index *ix, const uint8_t *reads, const uint64_t *read_off,
                               uint32_t n_reads, int32_t threshold, uint32_t flags,
                               int32_t n_threads, mo_read_result *out, uint32_t *top_refs,
                               uint32_t top_cap, mo_cand *cands, uint64_t cand_cap,
                               uint64_t *n_cands) {
  if (!ix || !out || threshold != ix->threshold) return -2;
  const int threads = hw_threads(n_threads);
  std::vector<Work> work(threads);
  std::vector<std::vector<mo_cand>> per_read;
  if (cands || n_cands) per_read.resize(n_reads);
  std::atomic<int> err(0);
  std::vector<std::vector<uint32_t>> tops(threads);
  parallel_for(n_reads, threads, [&](uint64_t r, int t) {
    const uint64_t m = read_off[r + 1] - read_off[r];
    std::vector<mo_cand> *cv = per_read.empty() ? nullptr : &per_read[r];
    if (classify_read(*ix, reads + read_off[r], m, (uint32_t)r, flags, work[t], out[r], tops[t],
                      cv) != 0) {
      err.store(-1);
      out[r] = mo_read_result{-1, 0, 0, 0};
      return;
    }
    if (top_refs)
      for (uint32_t i = 0; i < out[r].n_top && i < top_cap; i++)
        top_refs[(uint64_t)r * top_cap + i] = tops[t][i];
  });
  if (!per_read.empty()) {
    uint64_t n = 0;
    for (auto &v : per_read)
      for (auto &c : v) {
        if (cands && n < cand_cap) cands[n] = c;
        n++;
      }
    if (n_cands) *n_cands = n;
  }
  return err.load();
}

extern "C" int32_t mo_count_work(const mo_index *ix, const uint8_t *reads,
                                 const uint64_t *read_off, uint32_t n_reads, int32_t threshold,
                                 uint32_t flags, int32_t n_threads, mo_work *work_out) {
  if (!ix || !work_out || threshold != ix->threshold) return -2;
  const int threads = hw_threads(n_threads);
  std::vector<Work> work(threads);
  std::vector<std::vector<uint32_t>> tops(threads);
  parallel_for(n_reads, threads, [&](uint64_t r, int t) {
    mo_read_result rr;
    const uint64_t m = read_off[r + 1] - read_off[r];
    classify_read(*ix, reads + read_off[r], m, (uint32_t)r, flags, work[t], rr, tops[t], nullptr);
  });
  mo_work sum{};
  for (auto &w : work) {
    sum.probes += w.counters.probes;
    sum.hits += w.counters.hits;
    sum.entries += w.counters.entries;
    sum.ext_compares += w.counters.ext_compares;
    sum.seeds += w.counters.seeds;
    sum.cands += w.counters.cands;
  }
  *work_out = sum;
  return 0;
}

/*
 * The two library sorts the reference's result logs depend on, as GNU
 * libstdc++ (the image's GCC 11) performs them, with the reference's
 * comparators: candidates by relative score, descending
 * (MirAge/Analyser.cpp:404), and a group's references by taxonomy
 * occurrence (:252-293, not a strict weak ordering: id -1 goes first
 * against anything).  They pin the Python restatement in oracle/stdsort.py.
 */
extern "C" void mo_std_sort_desc(float *score, uint32_t *id, uint64_t n) {
  struct RS {  // MatrixResult::ResultsScores (MatrixResult.h:30-38)
    float score;
    uint64_t referenceID;
  };
  std::vector<RS> v(n);
  for (uint64_t i = 0; i < n; i++) v[i] = RS{score[i], id[i]};
  std::sort(v.begin(), v.end(), [](const RS &l, const RS &r) { return l.score > r.score; });
  for (uint64_t i = 0; i < n; i++) {
    score[i] = v[i].score;
    id[i] = (uint32_t)v[i].referenceID;
  }
}

extern "C" void mo_std_stable_sort_tax(uint64_t *ids, uint64_t n, const int32_t *taxids) {
  std::vector<uint64_t> v(ids, ids + n);
  std::unordered_map<uint64_t, size_t> occ;
  for (uint64_t r : v) occ[(uint64_t)(int64_t)taxids[r]]++;
  std::stable_sort(v.begin(), v.end(), [&](uint64_t a, uint64_t b) {
    const uint64_t ta = (uint64_t)(int64_t)taxids[a], tb = (uint64_t)(int64_t)taxids[b];
    if (ta == (uint64_t)-1) return true;
    if (tb == (uint64_t)-1) return false;
    if (ta == tb) return false;
    if (occ[ta] == occ[tb]) return false;
    if (occ[ta] > occ[tb]) return true;
    if (occ[ta] < occ[tb]) return false;
    return ta < tb;
  });
  std::copy(v.begin(), v.end(), ids);
}
