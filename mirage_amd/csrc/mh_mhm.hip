// MirAge .mhm hashmap files, version 4: host parsing of the header and the
// tables of contents, device decode of the bucket records into the
// direct-indexed index, and the writer.
//
// Layout (little-endian; HB = hash width 4 or 8, SB = stream-position
// width), as SaveHashmapTableV3 writes it (MirAge/Hashmap.cpp:1835-2540) and
// LoadHashMapFromHashmapDB / ReadHashmapFile_Future read it (:174-580):
//   73-byte indicator, u64 TOC entries, i32 version 4, i32 HB, i32 T, then
//   one byte each: SB, width of the largest bucket size, width of the longest
//   sequence, width of the longest header;
//   TOC: per bucket in ascending hash order, hash [HB] and the file position
//   of its record [SB];
//   taxonomy ids: count [HB], count x i32;
//   sequence TOC: position after it [SB], count [HB], sizes [HB each],
//   positions [SB each]; the header TOC likewise; sequence bytes; header bytes;
//   records (decoded as Analyser.cpp:1614-1720 does): widths bPO bID bPOS
//   bNS (a byte each), refIDMin [bPO], refPOSMin [bPO], n [bNS], n ids minus
//   refIDMin [bID each], bRep [1], number of position runs [bPOS], per run
//   count [bRep] and position minus refPOSMin [bPOS].  Ids pair index-wise
//   with the run-expanded positions.
//
// Device decode: one pass reads each record's entry count into the bucket
// counts (their scan is the directory); a second pass decodes a record of at
// most kSmallRecord entries per thread, and a third takes the larger records
// one workgroup each (ids in parallel; position runs 256 at a time, their
// counts scanned in LDS and the chunk's outputs spread over the workgroup by
// binary search, so one run of any length is shared by all 256 threads).
#include <fcntl.h>
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "mh_internal.hpp"

namespace mh {
namespace {

constexpr char kIndicator[] =
    "MirAge - Maarten Pater - www.mirdesign.nl - mirage@mirdesign.nl - Hashmap";
constexpr uint64_t kIndicatorLen = sizeof(kIndicator) - 1;  // 73
constexpr int32_t kVersion = 4;                              // Hashmap.cpp:433-442
constexpr uint64_t kSmallRecord = 64;  // entries decoded by one thread
constexpr unsigned kBigThreads = 256;

__host__ __device__ inline uint64_t get_le(const uint8_t *p, uint32_t w) {
  uint64_t v = 0;
  for (uint32_t i = 0; i < w; i++) v |= (uint64_t)p[i] << (8 * i);
  return v;
}

// Parsed fixed part of one record; ok is false when a width is outside
// 1..8 or a field would run past the record's extent.
struct RecHead {
  uint32_t bpo, bid, bpos, bns, brep;
  uint64_t id_min, pos_min, n, nrun;
  uint64_t ids_at, runs_at;
  bool ok;
};

__host__ __device__ inline bool width_ok(uint32_t w) { return w >= 1 && w <= 8; }

__host__ __device__ inline RecHead rec_head(const uint8_t *r, uint64_t len) {
  RecHead h;
  h.ok = false;
  h.bpo = h.bid = h.bpos = h.bns = h.brep = 0;
  h.id_min = h.pos_min = h.n = h.nrun = h.ids_at = h.runs_at = 0;
  if (len < 4) return h;
  h.bpo = r[0];
  h.bid = r[1];
  h.bpos = r[2];
  h.bns = r[3];
  if (!width_ok(h.bpo) || !width_ok(h.bid) || !width_ok(h.bpos) || !width_ok(h.bns)) return h;
  uint64_t o = 4;
  if (2ull * h.bpo + h.bns > len - o) return h;
  h.id_min = get_le(r + o, h.bpo);
  o += h.bpo;
  h.pos_min = get_le(r + o, h.bpo);
  o += h.bpo;
  h.n = get_le(r + o, h.bns);
  o += h.bns;
  h.ids_at = o;
  if (h.pos_min > 0xFFFFFFFFull) return h;
  if (h.n == 0 || h.n > (len - o) / h.bid) return h;
  o += h.n * h.bid;
  if (o >= len) return h;
  h.brep = r[o];
  o += 1;
  if (!width_ok(h.brep) || h.bpos > len - o) return h;
  h.nrun = get_le(r + o, h.bpos);
  o += h.bpos;
  h.runs_at = o;
  if (h.nrun == 0 || h.nrun > (len - o) / (h.brep + h.bpos)) return h;
  h.ok = true;
  return h;
}

// Reference id v (stored minus id_min) of a record, or ~0 when it is not
// below n_refs.
__host__ __device__ inline uint64_t ref_id(uint64_t v, uint64_t id_min, uint32_t n_refs) {
  if (id_min >= n_refs || v >= n_refs - id_min) return ~0ull;
  return v + id_min;
}

// A position is usable when it lies inside its reference (0 for an empty
// one): the extension reads at most the bytes [1, pos] before it and stops
// at the reference's end after it.
__host__ __device__ inline bool pos_ok(uint64_t pos, uint64_t ref_len) {
  return pos < (ref_len > 0 ? ref_len : 1);
}

__device__ inline void flag_bad(unsigned long long *err, uint64_t i) {
  atomicMin(err, (unsigned long long)i);
}

// Pass 1: entry counts of the records into the hash-indexed counts; records
// with more than kSmallRecord entries are listed for the workgroup pass.
__global__ void k_mhm_count(const uint8_t *rec, const uint64_t *sp, const uint32_t *hash,
                            uint64_t first, uint64_t n, uint64_t skip_last, uint32_t *cnt,
                            unsigned long long *err, uint32_t *big, unsigned int *n_big) {
  for (uint64_t i = first + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    if (skip_last && i == n - 1) continue;
    const RecHead hd = rec_head(rec + sp[i], sp[i + 1] - sp[i]);
    if (!hd.ok || hd.n > 0xFFFFFFFFull) {
      flag_bad(err, i);
      continue;
    }
    cnt[hash[i]] = (uint32_t)hd.n;
    if (hd.n > kSmallRecord) big[atomicAdd(n_big, 1u)] = (uint32_t)i;
  }
}

// Pass 2: a thread decodes a record of at most kSmallRecord entries.
__global__ void k_mhm_decode_small(const uint8_t *rec, const uint64_t *sp, const uint32_t *hash,
                                   uint64_t first, uint64_t n, uint64_t skip_last,
                                   const uint64_t *dir, const uint64_t *ref_off, uint32_t n_refs,
                                   uint2 *ent, unsigned long long *err) {
  for (uint64_t i = first + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    if (skip_last && i == n - 1) continue;
    const uint8_t *r = rec + sp[i];
    const RecHead hd = rec_head(r, sp[i + 1] - sp[i]);
    if (!hd.ok || hd.n > kSmallRecord) continue;  // flagged in pass 1, or pass 3's
    uint2 *e = ent + dir[hash[i]];
    bool bad = false;
    for (uint64_t j = 0; j < hd.n; j++) {
      const uint64_t id = ref_id(get_le(r + hd.ids_at + j * hd.bid, hd.bid), hd.id_min, n_refs);
      if (id == ~0ull) bad = true;
      e[j].x = (uint32_t)(bad ? 0 : id);
    }
    uint64_t o = 0;
    const uint32_t step = hd.brep + hd.bpos;
    for (uint64_t t = 0; t < hd.nrun && !bad; t++) {
      const uint8_t *q = r + hd.runs_at + t * step;
      const uint64_t c = get_le(q, hd.brep);
      const uint64_t pv = get_le(q + hd.brep, hd.bpos);
      if (c > hd.n - o || pv > 0xFFFFFFFFull - hd.pos_min) {
        bad = true;
        break;
      }
      for (uint64_t u = 0; u < c; u++) e[o + u].y = (uint32_t)(pv + hd.pos_min);
      o += c;
    }
    if (o != hd.n) bad = true;
    for (uint64_t j = 0; j < hd.n && !bad; j++) {
      const uint32_t id = e[j].x;
      if (!pos_ok(e[j].y, ref_off[id + 1] - ref_off[id])) bad = true;
    }
    if (bad) flag_bad(err, i);
  }
}

// Pass 3: a workgroup per record with more than kSmallRecord entries.
__global__ __launch_bounds__(kBigThreads) void k_mhm_decode_big(
    const uint8_t *rec, const uint64_t *sp, const uint32_t *hash, const uint32_t *big,
    const unsigned int *n_big, const uint64_t *dir, const uint64_t *ref_off, uint32_t n_refs,
    uint2 *ent, unsigned long long *err) {
  __shared__ uint64_t pre[kBigThreads];
  __shared__ uint64_t pval[kBigThreads];
  __shared__ int s_bad;
  const unsigned tid = threadIdx.x;
  const unsigned nb = *n_big;
  for (unsigned b = blockIdx.x; b < nb; b += gridDim.x) {
    const uint64_t i = big[b];
    const uint8_t *r = rec + sp[i];
    const RecHead hd = rec_head(r, sp[i + 1] - sp[i]);  // ok: checked in pass 1
    uint2 *e = ent + dir[hash[i]];
    if (tid == 0) s_bad = 0;
    __syncthreads();
    for (uint64_t j = tid; j < hd.n; j += kBigThreads) {
      const uint64_t id = ref_id(get_le(r + hd.ids_at + j * hd.bid, hd.bid), hd.id_min, n_refs);
      if (id == ~0ull) s_bad = 1;
      e[j].x = (uint32_t)(id == ~0ull ? 0 : id);
    }
    const uint32_t step = hd.brep + hd.bpos;
    uint64_t base = 0;
    for (uint64_t c0 = 0; c0 < hd.nrun; c0 += kBigThreads) {
      const uint64_t t = c0 + tid;
      uint64_t c = 0, pv = 0;
      if (t < hd.nrun) {
        const uint8_t *q = r + hd.runs_at + t * step;
        c = get_le(q, hd.brep);
        pv = get_le(q + hd.brep, hd.bpos);
        if (c > hd.n || pv > 0xFFFFFFFFull - hd.pos_min) {
          s_bad = 1;
          c = 0;
        }
      }
      pre[tid] = c;
      pval[tid] = pv + hd.pos_min;
      __syncthreads();
      // inclusive scan of the counts (each is at most n < 2^32, the sum of
      // 256 of them fits easily)
      for (unsigned d = 1; d < kBigThreads; d <<= 1) {
        const uint64_t add = tid >= d ? pre[tid - d] : 0;
        __syncthreads();
        pre[tid] += add;
        __syncthreads();
      }
      const uint64_t total = pre[kBigThreads - 1];
      if (total > hd.n - base) {
        if (tid == 0) s_bad = 1;
      } else {
        for (uint64_t o = tid; o < total; o += kBigThreads) {
          unsigned lo = 0, hi = kBigThreads - 1;  // first run whose inclusive sum exceeds o
          while (lo < hi) {
            const unsigned mid = (lo + hi) >> 1;
            if (pre[mid] > o) hi = mid;
            else lo = mid + 1;
          }
          e[base + o].y = (uint32_t)pval[lo];
        }
        base += total;
      }
      // every thread reads the flag between two barriers, before any thread
      // can set it again in the next chunk
      __syncthreads();
      const int stop = s_bad;
      __syncthreads();
      if (stop) break;
    }
    if (tid == 0 && base != hd.n) s_bad = 1;
    __syncthreads();
    if (!s_bad) {
      for (uint64_t j = tid; j < hd.n; j += kBigThreads) {
        const uint32_t id = e[j].x;
        if (!pos_ok(e[j].y, ref_off[id + 1] - ref_off[id])) s_bad = 1;
      }
    }
    __syncthreads();
    if (tid == 0 && s_bad) flag_bad(err, i);
    __syncthreads();
  }
}

// TOC entries (hash [hb], record position [sb]) into aligned columns;
// positions become offsets into the record bytes, sp[n] their total.
__global__ void k_mhm_toc(const uint8_t *toc, uint64_t n, uint32_t hb, uint32_t sb,
                          uint64_t rec0, uint64_t rec_bytes, uint32_t *hash, uint64_t *sp) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i <= n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    if (i == n) {
      sp[n] = rec_bytes;
      continue;
    }
    const uint8_t *e = toc + i * (hb + sb);
    hash[i] = (uint32_t)get_le(e, hb);
    sp[i] = get_le(e + hb, sb) - rec0;
  }
}

__global__ void k_upper_bytes(uint8_t *b, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint8_t c = b[i];
    b[i] = (int8_t)c > 96 ? (uint8_t)(c - 32) : c;  // Sequences.cpp:43-52
  }
}

// Non-empty buckets of an index, ascending: flag per hash for the selection.
struct NonEmpty {
  const uint64_t *dir;
  __host__ __device__ bool operator()(uint32_t h) const { return dir[h + 1] > dir[h]; }
};

__global__ void k_gather_dir(const uint64_t *dir, const uint32_t *hashes, uint64_t n,
                             uint64_t *begin) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    begin[i] = dir[hashes[i]];
}

struct ToU64 {
  __host__ __device__ uint64_t operator()(uint32_t v) const { return v; }
};

unsigned grid_of(uint64_t n, unsigned block, unsigned cap = 65536) {
  const uint64_t g = (n + block - 1) / block;
  return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(g, cap));
}

// Bytes the writer spends on a value: ((countBits(v) - 1) / 8) + 1 with
// countBits(0) = 0 (Hashmap.cpp:1534-1549, :1638-1651).
uint32_t nbytes(uint64_t v) {
  uint32_t bits = 0;
  while (v) {
    bits++;
    v >>= 1;
  }
  return bits == 0 ? 1 : (bits - 1) / 8 + 1;
}

uint8_t *put_le(uint8_t *p, uint64_t v, uint32_t w) {
  for (uint32_t i = 0; i < w; i++) *p++ = (uint8_t)(v >> (8 * i));
  return p;
}

// f(i) for i < n on up to 16 host threads, which claim `grain` items at a time.
template <class F>
void parallel_for(uint64_t n, F &&f, uint64_t grain = 1024) {
  const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n <= grain || nt == 1) {
    for (uint64_t i = 0; i < n; i++) f(i);
    return;
  }
  std::atomic<uint64_t> next{0};
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; t++)
    th.emplace_back([&]() {
      for (;;) {
        const uint64_t b = next.fetch_add(grain);
        if (b >= n) return;
        const uint64_t e = std::min<uint64_t>(n, b + grain);
        for (uint64_t i = b; i < e; i++) f(i);
      }
    });
  for (auto &t : th) t.join();
}

// Host array left uninitialised (every byte is written before it is read):
// the writer's buffers run to gigabytes, and zeroing them first costs seconds.
template <class T>
struct HostBuf {
  std::unique_ptr<T[]> p;
  uint64_t n = 0;
  explicit HostBuf(uint64_t count) : p(new T[count ? count : 1]), n(count) {}
  T *data() { return p.get(); }
  const T *data() const { return p.get(); }
  uint64_t size() const { return n; }
  T *begin() { return p.get(); }
};

// Size of one bucket's record for entries sorted by position
// (GetSaveHashmapPreCalculates, Hashmap.cpp:1620-1697, written at :2435-2516).
struct RecPlan {
  uint32_t bpo, bid, bpos, bns, brep;
  uint64_t id_min, pos_min, nrun, bytes;
};

RecPlan plan_record(const uint2 *e, uint64_t n) {
  RecPlan p;
  uint64_t id_max = e[0].x;
  p.id_min = e[0].x;
  for (uint64_t j = 1; j < n; j++) {
    p.id_min = std::min<uint64_t>(p.id_min, e[j].x);
    id_max = std::max<uint64_t>(id_max, e[j].x);
  }
  p.pos_min = e[0].y;
  const uint64_t pos_max = e[n - 1].y;
  p.nrun = 1;
  for (uint64_t j = 1; j < n; j++) p.nrun += e[j].y != e[j - 1].y;
  p.bid = nbytes(id_max - p.id_min);
  p.bpos = nbytes(pos_max - p.pos_min);
  p.bpo = nbytes(std::max(p.id_min, p.pos_min));
  p.bns = nbytes(n);
  p.brep = nbytes(n);
  p.bytes = 4 + 2ull * p.bpo + p.bns + n * p.bid + 1 + p.bpos + p.nrun * (p.brep + p.bpos);
  return p;
}

void write_record(uint8_t *o, const uint2 *e, uint64_t n, const RecPlan &p) {
  *o++ = (uint8_t)p.bpo;
  *o++ = (uint8_t)p.bid;
  *o++ = (uint8_t)p.bpos;
  *o++ = (uint8_t)p.bns;
  o = put_le(o, p.id_min, p.bpo);
  o = put_le(o, p.pos_min, p.bpo);
  o = put_le(o, n, p.bns);
  for (uint64_t j = 0; j < n; j++) o = put_le(o, e[j].x - p.id_min, p.bid);
  *o++ = (uint8_t)p.brep;
  o = put_le(o, p.nrun, p.bpos);  // truncated to bPOS bytes, as the writer does
  uint64_t j = 0;
  while (j < n) {
    uint64_t c = 1;
    while (j + c < n && e[j + c].y == e[j].y) c++;
    o = put_le(o, c, p.brep);
    o = put_le(o, e[j].y - p.pos_min, p.bpos);
    j += c;
  }
}

}  // namespace
}  // namespace mh

using namespace mh;

struct mh_mhm {
  int fd = -1;
  const uint8_t *p = nullptr;
  uint64_t size = 0;
  int32_t threshold = 0;
  uint32_t hb = 4, sb = 4;
  uint64_t n = 0;        // TOC entries
  uint64_t toc_off = 0;  // file position of the first TOC entry
  // TOC entry i, read from the mapped file (validated at open: ascending
  // hashes below 2^28, ascending positions inside the file).
  uint64_t hash(uint64_t i) const { return get_le(p + toc_off + i * (hb + sb), hb); }
  uint64_t sp(uint64_t i) const {  // sp(n) is the file size
    return i == n ? size : get_le(p + toc_off + i * (hb + sb) + hb, sb);
  }
  std::vector<int32_t> taxids;
  std::vector<uint64_t> seq_size, seq_sp, hdr_size, hdr_sp;
  uint64_t ref_bytes = 0, hdr_bytes = 0;
  ~mh_mhm() {
    if (p) munmap((void *)p, size);
    if (fd >= 0) close(fd);
  }
};

namespace {

struct Cursor {
  const uint8_t *p;
  uint64_t size, o;
  bool ok;
  uint64_t get(uint32_t w) {
    if (!ok || w > 8 || w > size - o) {
      ok = false;
      return 0;
    }
    const uint64_t v = get_le(p + o, w);
    o += w;
    return v;
  }
};

mh_status bad_file(const std::string &msg) {
  return set_error(nullptr, MH_ERR_ARG, "mh_mhm_open: " + msg);
}

mh_status parse(mh_mhm *m) {
  Cursor c{m->p, m->size, 0, true};
  if (m->size < kIndicatorLen + 24 || std::memcmp(m->p, kIndicator, kIndicatorLen) != 0)
    return bad_file("not a MirAge hashmap file (indicator string)");
  c.o = kIndicatorLen;
  m->n = c.get(8);
  const int32_t version = (int32_t)c.get(4);
  const int32_t hb = (int32_t)c.get(4);
  m->threshold = (int32_t)c.get(4);
  m->sb = (uint32_t)c.get(1);
  c.get(3);  // widths of the largest bucket, sequence and header: not needed to read
  if (version != kVersion) return bad_file("unsupported hashmap version " + std::to_string(version));
  if (hb != 4 && hb != 8) return bad_file("unsupported hash width " + std::to_string(hb));
  if (m->sb < 1 || m->sb > 8) return bad_file("stream position width " + std::to_string(m->sb));
  if (m->threshold < 2) return bad_file("threshold " + std::to_string(m->threshold));
  m->hb = (uint32_t)hb;
  if (m->n > (m->size - c.o) / (m->hb + m->sb)) return bad_file("table of contents past the end");
  m->toc_off = c.o;
  c.o += m->n * (m->hb + m->sb);
  const uint64_t ntax = c.get(m->hb);
  if (!c.ok || ntax > (m->size - c.o) / 4) return bad_file("taxonomy ids past the end");
  m->taxids.resize(ntax);
  for (uint64_t i = 0; i < ntax; i++) m->taxids[i] = (int32_t)c.get(4);
  for (int part = 0; part < 2; part++) {
    std::vector<uint64_t> &sz = part ? m->hdr_size : m->seq_size;
    std::vector<uint64_t> &at = part ? m->hdr_sp : m->seq_sp;
    const uint64_t after = c.get(m->sb);
    const uint64_t cnt = c.get(m->hb);
    if (!c.ok || cnt != ntax || cnt >= 0xFFFFFFFFull)
      return bad_file("reference count differs between the tables");
    if (cnt > (m->size - c.o) / (m->hb + m->sb)) return bad_file("reference table past the end");
    if (after != c.o + cnt * (m->hb + m->sb)) return bad_file("reference table position");
    sz.resize(cnt);
    at.resize(cnt);
    for (uint64_t i = 0; i < cnt; i++) sz[i] = c.get(m->hb);
    for (uint64_t i = 0; i < cnt; i++) at[i] = c.get(m->sb);
    uint64_t total = 0;
    for (uint64_t i = 0; i < cnt; i++) {
      if (at[i] > m->size || sz[i] > m->size - at[i]) return bad_file("reference data past the end");
      if (!part && sz[i] > 0xFFFFFFFFull) return bad_file("reference longer than 2^32 bases");
      total += sz[i];
    }
    (part ? m->hdr_bytes : m->ref_bytes) = total;
  }
  if (!c.ok) return bad_file("truncated file");
  if (m->n && m->sp(0) < c.o) return bad_file("record positions inside the tables");
  if (m->n && m->sp(m->n - 1) >= m->size) return bad_file("record positions past the end");
  // The TOC, in parallel slices: hashes below 2^28 and both columns ascending.
  std::atomic<int> bad{0};  // 1 hash range, 2 hash order, 3 position order
  parallel_for((m->n + 65535) / 65536, [&](uint64_t b) {
    const uint64_t i0 = b * 65536, i1 = std::min<uint64_t>(m->n, i0 + 65536);
    uint64_t ph = i0 ? m->hash(i0 - 1) : 0, ps = i0 ? m->sp(i0 - 1) : 0;
    for (uint64_t i = i0; i < i1; i++) {
      const uint64_t h = m->hash(i), sp = m->sp(i);
      int e = 0;
      if (h > kHashMask) e = 1;
      else if (i && h <= ph) e = 2;
      else if (i && sp <= ps) e = 3;
      if (e) {
        int z = 0;
        bad.compare_exchange_strong(z, e);
        return;
      }
      ph = h;
      ps = sp;
    }
  }, 1);
  if (bad == 1) return bad_file("hash above 28 bits in the table of contents");
  if (bad == 2) return bad_file("table of contents not in ascending hash order");
  if (bad == 3) return bad_file("record positions not ascending");
  return MH_OK;
}

// Host decode of TOC entry i into (id, pos) pairs; false on a bad record.
bool decode_host(const mh_mhm *m, uint64_t i, std::vector<uint2> &out) {
  const uint8_t *r = m->p + m->sp(i);
  const RecHead hd = rec_head(r, m->sp(i + 1) - m->sp(i));
  if (!hd.ok || hd.n > 0xFFFFFFFFull) return false;
  const uint32_t n_refs = (uint32_t)m->seq_size.size();
  out.assign(hd.n, make_uint2(0, 0));
  for (uint64_t j = 0; j < hd.n; j++) {
    const uint64_t id = ref_id(get_le(r + hd.ids_at + j * hd.bid, hd.bid), hd.id_min, n_refs);
    if (id == ~0ull) return false;
    out[j].x = (uint32_t)id;
  }
  uint64_t o = 0;
  for (uint64_t t = 0; t < hd.nrun; t++) {
    const uint8_t *q = r + hd.runs_at + t * (hd.brep + hd.bpos);
    const uint64_t c = get_le(q, hd.brep);
    const uint64_t pv = get_le(q + hd.brep, hd.bpos);
    if (c > hd.n - o || pv > 0xFFFFFFFFull - hd.pos_min) return false;
    for (uint64_t u = 0; u < c; u++) out[o + u].y = (uint32_t)(pv + hd.pos_min);
    o += c;
  }
  if (o != hd.n) return false;
  for (const uint2 &e : out)
    if (!pos_ok(e.y, m->seq_size[e.x])) return false;
  return true;
}

// Host decode of the record of hash 0 (TOC entry 0) with its identical
// (0, 0) entries counted instead of stored: out receives the other entries
// and one (0, 0) when there is any; *copies the number of (0, 0) entries.
bool reduce_bucket0(const mh_mhm *m, std::vector<uint2> &out, uint64_t *copies) {
  const uint8_t *r = m->p + m->sp(0);
  const RecHead hd = rec_head(r, m->sp(1) - m->sp(0));
  if (!hd.ok || hd.n > 0xFFFFFFFFull) return false;
  const uint32_t n_refs = (uint32_t)m->seq_size.size();
  out.clear();
  *copies = 0;
  uint64_t j = 0;
  for (uint64_t t = 0; t < hd.nrun; t++) {
    const uint8_t *q = r + hd.runs_at + t * (hd.brep + hd.bpos);
    const uint64_t c = get_le(q, hd.brep);
    const uint64_t pv = get_le(q + hd.brep, hd.bpos);
    if (c > hd.n - j || pv > 0xFFFFFFFFull - hd.pos_min) return false;
    const uint32_t pos = (uint32_t)(pv + hd.pos_min);
    for (uint64_t u = 0; u < c; u++, j++) {
      const uint64_t id = ref_id(get_le(r + hd.ids_at + j * hd.bid, hd.bid), hd.id_min, n_refs);
      if (id == ~0ull || !pos_ok(pos, m->seq_size[id])) return false;
      if (id == 0 && pos == 0) {
        if ((*copies)++ == 0) out.push_back(make_uint2(0, 0));
      } else {
        out.push_back(make_uint2((uint32_t)id, pos));
      }
    }
  }
  return j == hd.n;
}

// Copies file bytes [off, off + len) to device memory through pinned
// buffers: host threads read 16 MiB chunks with pread, each issuing its
// copies on a stream of its own, with two buffers per thread so that reading
// the next chunk overlaps the transfer of the last.  Returns 0, -1 when the
// file could not be read, or the first HIP error.
int upload_span(const mh_mhm *m, int device, uint64_t off, uint64_t len, uint8_t *d_dst) {
  if (len == 0) return 0;
  constexpr uint64_t kChunk = 16ull << 20;
  const uint64_t nchunk = (len + kChunk - 1) / kChunk;
  const unsigned hw = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
  const unsigned nt = (unsigned)std::min<uint64_t>(nchunk, hw);
  std::atomic<int> err{0};
  auto note = [&](int e) {
    int z = 0;
    err.compare_exchange_strong(z, e);
  };
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; t++)
    th.emplace_back([&, t]() {
      hipSetDevice(device);
      hipStream_t s = nullptr;
      uint8_t *buf[2] = {nullptr, nullptr};
      hipEvent_t ev[2] = {nullptr, nullptr};
      bool used[2] = {false, false};
      hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
      for (int b = 0; b < 2 && e == hipSuccess; b++) {
        e = hipHostMalloc((void **)&buf[b], kChunk, hipHostMallocDefault);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ev[b], hipEventDisableTiming);
      }
      int par = 0;
      for (uint64_t c = t; c < nchunk && e == hipSuccess && err.load() == 0; c += nt) {
        if (used[par] && (e = hipEventSynchronize(ev[par])) != hipSuccess) break;
        const uint64_t o = c * kChunk, want = std::min(kChunk, len - o);
        uint64_t got = 0;
        while (got < want) {
          const ssize_t r = pread(m->fd, buf[par] + got, want - got, (off_t)(off + o + got));
          if (r <= 0) break;
          got += (uint64_t)r;
        }
        if (got < want) {
          note(-1);
          break;
        }
        e = hipMemcpyAsync(d_dst + o, buf[par], want, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipEventRecord(ev[par], s);
        used[par] = true;
        par ^= 1;
      }
      if (s) {
        const hipError_t e2 = hipStreamSynchronize(s);
        if (e == hipSuccess) e = e2;
      }
      for (int b = 0; b < 2; b++) {
        if (ev[b]) hipEventDestroy(ev[b]);
        if (buf[b]) hipHostFree(buf[b]);
      }
      if (s) hipStreamDestroy(s);
      if (e != hipSuccess) note((int)e);
    });
  for (auto &x : th) x.join();
  return err.load();
}

}  // namespace

extern "C" {

mh_status mh_mhm_open(const char *path, mh_mhm **out) {
  if (!path || !out) return set_error(nullptr, MH_ERR_ARG, "mh_mhm_open: bad argument");
  *out = nullptr;
  mh_mhm *m = new mh_mhm();
  m->fd = open(path, O_RDONLY | O_CLOEXEC);
  struct stat stt;
  if (m->fd < 0 || fstat(m->fd, &stt) != 0) {
    delete m;
    return set_error(nullptr, MH_ERR_ARG, std::string("mh_mhm_open: cannot open ") + path);
  }
  m->size = (uint64_t)stt.st_size;
  if (m->size > 0) {
    void *p = mmap(nullptr, m->size, PROT_READ, MAP_PRIVATE, m->fd, 0);
    if (p == MAP_FAILED) {
      delete m;
      return set_error(nullptr, MH_ERR_ARG, std::string("mh_mhm_open: cannot map ") + path);
    }
    m->p = (const uint8_t *)p;
  }
  const mh_status st = parse(m);
  if (st != MH_OK) {
    delete m;
    return st;
  }
  *out = m;
  return MH_OK;
}

mh_status mh_mhm_close(mh_mhm *m) {
  delete m;
  return MH_OK;
}

mh_status mh_mhm_get_info(const mh_mhm *m, mh_mhm_info *out) {
  if (!m || !out) return set_error(nullptr, MH_ERR_ARG, "mh_mhm_get_info: bad argument");
  out->n_buckets = m->n;
  out->file_bytes = m->size;
  out->record_bytes = m->n ? m->size - m->sp(0) : 0;
  out->ref_bytes = m->ref_bytes;
  out->header_bytes = m->hdr_bytes;
  out->n_refs = (uint32_t)m->seq_size.size();
  out->threshold = m->threshold;
  out->hash_bytes = m->hb;
  out->streampos_bytes = m->sb;
  return MH_OK;
}

mh_status mh_mhm_references(const mh_mhm *m, uint8_t *h_bytes, uint64_t *h_off,
                            uint8_t *h_headers, uint64_t *h_header_off, int32_t *h_taxids) {
  if (!m) return set_error(nullptr, MH_ERR_ARG, "mh_mhm_references: bad argument");
  const uint64_t nr = m->seq_size.size();
  for (int part = 0; part < 2; part++) {
    uint8_t *dst = part ? h_headers : h_bytes;
    uint64_t *off = part ? h_header_off : h_off;
    const std::vector<uint64_t> &sz = part ? m->hdr_size : m->seq_size;
    const std::vector<uint64_t> &at = part ? m->hdr_sp : m->seq_sp;
    uint64_t o = 0;
    for (uint64_t i = 0; i < nr; i++) {
      if (off) off[i] = o;
      if (dst) std::memcpy(dst + o, m->p + at[i], sz[i]);
      o += sz[i];
    }
    if (off) off[nr] = o;
  }
  if (h_taxids) std::copy(m->taxids.begin(), m->taxids.end(), h_taxids);
  return MH_OK;
}

mh_status mh_mhm_bucket(const mh_mhm *m, uint32_t hash, uint32_t flags, uint32_t *h_refs,
                        uint32_t *h_pos, uint64_t cap, int64_t *count) {
  if (!m || !count) return set_error(nullptr, MH_ERR_ARG, "mh_mhm_bucket: bad argument");
  uint64_t lo = 0, hi = m->n;  // first entry with hash >= the key
  while (lo < hi) {
    const uint64_t mid = (lo + hi) / 2;
    if (m->hash(mid) < hash) lo = mid + 1;
    else hi = mid;
  }
  if (lo == m->n || m->hash(lo) != hash) {
    *count = -1;
    return MH_OK;
  }
  const uint64_t i = lo;
  if ((flags & MH_INDEX_LAST_BUCKET_EMPTY) && i == m->n - 1) {
    *count = 0;
    return MH_OK;
  }
  std::vector<uint2> e;
  if (!decode_host(m, i, e))
    return set_error(nullptr, MH_ERR_ARG,
                     "mh_mhm_bucket: corrupt record of hash " + std::to_string(hash));
  for (uint64_t j = 0; j < e.size() && j < cap; j++) {
    if (h_refs) h_refs[j] = e[j].x;
    if (h_pos) h_pos[j] = e[j].y;
  }
  *count = (int64_t)e.size();
  return MH_OK;
}

mh_status mh_index_from_mhm(mh_handle *h, const mh_mhm *m, int32_t threshold, uint32_t flags,
                            mh_index **out) {
  if (!h || !m || !out) return set_error(h, MH_ERR_ARG, "mh_index_from_mhm: bad argument");
  if (threshold == 0) threshold = m->threshold;
  if (threshold < m->threshold)
    return set_error(h, MH_ERR_ARG,
                     "mh_index_from_mhm: threshold " + std::to_string(threshold) +
                         " is below the file's " + std::to_string(m->threshold));
  std::lock_guard<std::mutex> lk(h->mu);
  hipSetDevice(h->device);
  hipStream_t st = h->stream;
  const uint32_t n_refs = (uint32_t)m->seq_size.size();
  const uint64_t total = m->ref_bytes;
  const uint64_t n = m->n;
  const bool skip_last = (flags & MH_INDEX_LAST_BUCKET_EMPTY) && n > 0;

  mh_index *ix = new mh_index();
  ix->owner = h;
  ix->device = h->device;
  ix->threshold = threshold;
  ix->n_refs = n_refs;
  ix->h_ref_off.resize(n_refs + 1);
  ix->h_ref_off[0] = 0;
  bool contiguous = true;
  for (uint32_t r = 0; r < n_refs; r++) {
    ix->h_ref_off[r + 1] = ix->h_ref_off[r] + m->seq_size[r];
    if (m->seq_sp[r] != m->seq_sp[0] + ix->h_ref_off[r]) contiguous = false;
  }

  uint8_t *d_rec = nullptr, *d_toc = nullptr;
  uint64_t *d_sp = nullptr;
  uint32_t *d_hash = nullptr, *d_cnt = nullptr, *d_big = nullptr;
  unsigned long long *d_misc = nullptr;  // [0] first bad record, [1] big-record count
  void *d_tmp = nullptr;
  auto cleanup = [&]() {
    if (d_rec) hipFree(d_rec);
    if (d_toc) hipFree(d_toc);
    if (d_sp) hipFree(d_sp);
    if (d_hash) hipFree(d_hash);
    if (d_cnt) hipFree(d_cnt);
    if (d_big) hipFree(d_big);
    if (d_misc) hipFree(d_misc);
    if (d_tmp) hipFree(d_tmp);
  };
  auto fail = [&](mh_status s) {
    cleanup();
    hipStreamSynchronize(st);
    mh_index_free(ix);
    return s;
  };
#define MHM_HIP(call)                                          \
  do {                                                         \
    hipError_t e_ = (call);                                    \
    if (e_ != hipSuccess) return fail(hip_error(h, e_, #call)); \
  } while (0)

  const uint64_t rec0 = n ? m->sp(0) : m->size;
  const uint64_t rec_bytes = m->size - rec0;
  const uint64_t toc_bytes = n * (m->hb + m->sb);
  MHM_HIP(hipMalloc(&ix->d_bytes, std::max<uint64_t>(total, 1) + 64));
  MHM_HIP(hipMalloc(&ix->d_ref_off, (n_refs + 1) * sizeof(uint64_t)));
  MHM_HIP(hipMalloc(&ix->d_dir, (kDirSize + 1) * sizeof(uint64_t)));
  MHM_HIP(hipMalloc(&d_rec, std::max<uint64_t>(rec_bytes, 1)));
  MHM_HIP(hipMalloc(&d_toc, std::max<uint64_t>(toc_bytes, 1)));
  MHM_HIP(hipMalloc(&d_sp, (n + 1) * sizeof(uint64_t)));
  MHM_HIP(hipMalloc(&d_hash, std::max<uint64_t>(n, 1) * sizeof(uint32_t)));
  MHM_HIP(hipMalloc(&d_big, std::max<uint64_t>(n, 1) * sizeof(uint32_t)));
  MHM_HIP(hipMalloc(&d_cnt, (kDirSize + 1) * sizeof(uint32_t)));
  MHM_HIP(hipMalloc(&d_misc, 4 * sizeof(unsigned long long)));
  MHM_HIP(hipMemcpyAsync(ix->d_ref_off, ix->h_ref_off.data(), (n_refs + 1) * sizeof(uint64_t),
                         hipMemcpyHostToDevice, st));
  // The bulk of the file goes up as stored: sequences, TOC, records.
  {
    int e = 0;
    if (total && contiguous) {
      e = upload_span(m, h->device, m->seq_sp[0], total, ix->d_bytes);
    } else {
      for (uint32_t r = 0; r < n_refs && e == 0; r++)
        if (m->seq_size[r])
          e = (int)hipMemcpy(ix->d_bytes + ix->h_ref_off[r], m->p + m->seq_sp[r],
                             m->seq_size[r], hipMemcpyHostToDevice);
    }
    if (e == 0) e = upload_span(m, h->device, m->toc_off, toc_bytes, d_toc);
    if (e == 0) e = upload_span(m, h->device, rec0, rec_bytes, d_rec);
    if (e == -1) return fail(set_error(h, MH_ERR_ARG, "mh_index_from_mhm: file read failed"));
    if (e) return fail(hip_error(h, (hipError_t)e, "mh_index_from_mhm: upload"));
  }
  if (total) k_upper_bytes<<<grid_of(total, 256), 256, 0, st>>>(ix->d_bytes, total);
  k_mhm_toc<<<grid_of(n + 1, 256), 256, 0, st>>>(d_toc, n, m->hb, m->sb, rec0, rec_bytes, d_hash,
                                                  d_sp);
  MHM_HIP(hipMemsetAsync(d_cnt, 0, (kDirSize + 1) * sizeof(uint32_t), st));
  const unsigned long long init[4] = {~0ull, 0, 0, 0};
  MHM_HIP(hipMemcpyAsync(d_misc, init, sizeof(init), hipMemcpyHostToDevice, st));
  // Bucket 0 is decoded here: its identical (0, 0) entries, one per window
  // the reference dropped after an "N"-window cutoff, are stored once and
  // counted in ix->phantom.
  const uint64_t first = (n && m->hash(0) == 0 && !(skip_last && n == 1)) ? 1 : 0;
  std::vector<uint2> b0;
  if (first) {
    uint64_t copies = 0;
    if (!reduce_bucket0(m, b0, &copies))
      return fail(set_error(h, MH_ERR_ARG, "mh_index_from_mhm: corrupt record 0 (hash 0)"));
    ix->phantom = copies > 1 ? copies - 1 : 0;
    const uint32_t c0 = (uint32_t)b0.size();
    MHM_HIP(hipMemcpyAsync(d_cnt, &c0, sizeof(uint32_t), hipMemcpyHostToDevice, st));
  }
  if (n > first)
    k_mhm_count<<<grid_of(n - first, 256), 256, 0, st>>>(d_rec, d_sp, d_hash, first, n,
                                                         skip_last ? 1 : 0, d_cnt, d_misc, d_big,
                                                         (unsigned int *)(d_misc + 1));
  {
    hipcub::TransformInputIterator<uint64_t, ToU64, const uint32_t *> in(d_cnt, ToU64());
    size_t tmp_bytes = 0;
    hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, in, ix->d_dir, (int)(kDirSize + 1), st);
    MHM_HIP(hipMalloc(&d_tmp, tmp_bytes));
    hipcub::DeviceScan::ExclusiveSum(d_tmp, tmp_bytes, in, ix->d_dir, (int)(kDirSize + 1), st);
  }
  unsigned long long misc[4];
  uint64_t n_ent = 0;
  MHM_HIP(hipMemcpyAsync(misc, d_misc, sizeof(misc), hipMemcpyDeviceToHost, st));
  MHM_HIP(hipMemcpyAsync(&n_ent, ix->d_dir + kDirSize, sizeof(uint64_t), hipMemcpyDeviceToHost,
                         st));
  MHM_HIP(hipStreamSynchronize(st));
  auto corrupt = [&](unsigned long long i) {
    return fail(set_error(h, MH_ERR_ARG,
                          "mh_index_from_mhm: corrupt record " + std::to_string(i) +
                              " (hash " + std::to_string(m->hash(i)) + ")"));
  };
  if (misc[0] != ~0ull) return corrupt(misc[0]);
  ix->n_entries = n_ent + ix->phantom;
  MHM_HIP(hipMalloc(&ix->d_ent, std::max<uint64_t>(n_ent, 1) * sizeof(uint2)));
  if (!b0.empty())  // bucket 0 starts at entry 0
    MHM_HIP(hipMemcpyAsync(ix->d_ent, b0.data(), b0.size() * sizeof(uint2),
                           hipMemcpyHostToDevice, st));
  if (n > first) {
    k_mhm_decode_small<<<grid_of(n - first, 256), 256, 0, st>>>(
        d_rec, d_sp, d_hash, first, n, skip_last ? 1 : 0, ix->d_dir, ix->d_ref_off, n_refs,
        ix->d_ent, d_misc);
    const unsigned n_big = (unsigned)misc[1];
    if (n_big)
      k_mhm_decode_big<<<std::min(n_big, 4096u), kBigThreads, 0, st>>>(
          d_rec, d_sp, d_hash, d_big, (const unsigned int *)(d_misc + 1), ix->d_dir,
          ix->d_ref_off, n_refs, ix->d_ent, d_misc);
  }
  MHM_HIP(hipMemcpyAsync(misc, d_misc, sizeof(misc), hipMemcpyDeviceToHost, st));
  MHM_HIP(hipStreamSynchronize(st));
  if (misc[0] != ~0ull) return corrupt(misc[0]);
#undef MHM_HIP
  cleanup();
  ix->n_buckets = skip_last ? n - 1 : n;
  if (skip_last) ix->empty_hash = (uint32_t)m->hash(n - 1);
  ix->total_bytes = (std::max<uint64_t>(total, 1) + 64) + (n_refs + 1) * 8 +
                    (kDirSize + 1) * 8 + std::max<uint64_t>(n_ent, 1) * 8;
  *out = ix;
  return MH_OK;
}

mh_status mh_index_save_mhm(mh_handle *h, const mh_index *idx, const char *path,
                            const uint8_t *h_headers, const uint64_t *h_header_off,
                            const int32_t *h_taxids, uint32_t hash_bytes,
                            uint32_t streampos_bytes) {
  if (!h || !idx || !path || ((h_headers == nullptr) != (h_header_off == nullptr)))
    return set_error(h, MH_ERR_ARG, "mh_index_save_mhm: bad argument");
  const uint32_t hb = hash_bytes ? hash_bytes : 4;
  const uint32_t sb = streampos_bytes ? streampos_bytes : 4;
  if ((hb != 4 && hb != 8) || (sb != 4 && sb != 8))
    return set_error(h, MH_ERR_ARG, "mh_index_save_mhm: widths must be 4 or 8");
  if (idx->empty_hash != kNoHash)
    return set_error(h, MH_ERR_STATE,
                     "mh_index_save_mhm: the index dropped its last bucket "
                     "(built with MH_INDEX_LAST_BUCKET_EMPTY)");
  std::lock_guard<std::mutex> lk(h->mu);
  hipSetDevice(h->device);
  hipStream_t st = h->stream;
  const uint64_t nb = idx->n_buckets;
  const uint32_t n_refs = idx->n_refs;
  const uint64_t total = idx->h_ref_off[n_refs];

  // Non-empty buckets in ascending hash order with their entry offsets.
  const uint64_t n_stored = idx->n_entries - idx->phantom;
  std::vector<uint32_t> hashes(nb);
  std::vector<uint64_t> begin(nb + 1);
  HostBuf<uint2> ent(n_stored);
  HostBuf<uint8_t> bytes(total);
  {
    uint32_t *d_sel = nullptr;
    uint64_t *d_begin = nullptr;
    unsigned long long *d_num = nullptr;
    void *d_tmp = nullptr;
    auto release = [&]() {
      if (d_sel) hipFree(d_sel);
      if (d_begin) hipFree(d_begin);
      if (d_num) hipFree(d_num);
      if (d_tmp) hipFree(d_tmp);
    };
    hipError_t e = hipMalloc(&d_sel, std::max<uint64_t>(nb, 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&d_begin, std::max<uint64_t>(nb, 1) * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc(&d_num, sizeof(unsigned long long));
    hipcub::CountingInputIterator<uint32_t> hs(0);
    size_t tmp_bytes = 0;
    if (e == hipSuccess)
      e = hipcub::DeviceSelect::If(nullptr, tmp_bytes, hs, d_sel, d_num, (int)kDirSize,
                                   NonEmpty{idx->d_dir}, st);
    if (e == hipSuccess) e = hipMalloc(&d_tmp, tmp_bytes);
    if (e == hipSuccess)
      e = hipcub::DeviceSelect::If(d_tmp, tmp_bytes, hs, d_sel, d_num, (int)kDirSize,
                                   NonEmpty{idx->d_dir}, st);
    unsigned long long got = 0;
    if (e == hipSuccess)
      e = hipMemcpyAsync(&got, d_num, sizeof(got), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess && got != nb) {
      release();
      return set_error(h, MH_ERR_STATE, "mh_index_save_mhm: bucket count mismatch");
    }
    if (e == hipSuccess && nb) {
      k_gather_dir<<<grid_of(nb, 256), 256, 0, st>>>(idx->d_dir, d_sel, nb, d_begin);
      e = hipMemcpyAsync(hashes.data(), d_sel, nb * sizeof(uint32_t), hipMemcpyDeviceToHost, st);
      if (e == hipSuccess)
        e = hipMemcpyAsync(begin.data(), d_begin, nb * sizeof(uint64_t), hipMemcpyDeviceToHost,
                           st);
    }
    if (e == hipSuccess && n_stored)
      e = hipMemcpyAsync(ent.data(), idx->d_ent, n_stored * sizeof(uint2),
                         hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && total)
      e = hipMemcpyAsync(bytes.data(), idx->d_bytes, total, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    release();
    if (e != hipSuccess) return hip_error(h, e, "mh_index_save_mhm: copy from device");
  }
  begin[nb] = n_stored;
  // Bucket 0 with its identical (0, 0) copies written out, as the reference
  // holds them.
  std::vector<uint2> b0;
  const bool expand0 = nb && hashes[0] == 0 && idx->phantom;
  if (expand0) {
    b0.assign(ent.begin() + begin[0], ent.begin() + begin[1]);
    b0.resize(b0.size() + idx->phantom, make_uint2(0, 0));
  }
  auto bucket = [&](uint64_t b) -> uint2 * {
    return (b == 0 && expand0) ? b0.data() : ent.data() + begin[b];
  };
  auto bucket_n = [&](uint64_t b) -> uint64_t {
    return (b == 0 && expand0) ? b0.size() : begin[b + 1] - begin[b];
  };

  // Records: each bucket sorted by position (SortAndCorrectPairVector,
  // Hashmap.cpp:1561-1573; ties by reference id), sized, then written.
  std::vector<RecPlan> plan(nb);
  uint64_t max_bucket = 0;
  for (uint64_t b = 0; b < nb; b++) max_bucket = std::max(max_bucket, bucket_n(b));
  parallel_for(nb, [&](uint64_t b) {
    uint2 *e = bucket(b);
    const uint64_t n = bucket_n(b);
    if (n > 1)
      std::sort(e, e + n, [](const uint2 &a, const uint2 &c) {
        return a.y < c.y || (a.y == c.y && a.x < c.x);
      });
    plan[b] = plan_record(e, n);
  });

  // Headers as MirAge stores them, and their sizes.
  std::vector<uint8_t> hdr;
  std::vector<uint64_t> hdr_off(n_refs + 1, 0);
  if (h_headers) {
    for (uint32_t r = 0; r <= n_refs; r++) hdr_off[r] = h_header_off[r] - h_header_off[0];
    hdr.assign(h_headers + h_header_off[0], h_headers + h_header_off[n_refs]);
  } else {
    for (uint32_t r = 0; r < n_refs; r++) {
      const std::string s = ">" + std::to_string(r);
      hdr.insert(hdr.end(), s.begin(), s.end());
      hdr.push_back(0);
      hdr_off[r + 1] = hdr.size();
    }
  }
  uint64_t max_seq = 0, max_hdr = 0;
  for (uint32_t r = 0; r < n_refs; r++) {
    max_seq = std::max(max_seq, idx->h_ref_off[r + 1] - idx->h_ref_off[r]);
    max_hdr = std::max(max_hdr, hdr_off[r + 1] - hdr_off[r]);
  }

  // Record positions (Hashmap.cpp:1886-1920, :2064-2091).
  std::vector<uint64_t> rec_at(nb + 1, 0);
  for (uint64_t b = 0; b < nb; b++) rec_at[b + 1] = rec_at[b] + plan[b].bytes;

  // Everything before the records (Hashmap.cpp:1870-2240): the header, the
  // TOC (filled in parallel at fixed offsets), the taxonomy ids and the two
  // reference tables.
  const uint64_t head0 = kIndicatorLen + 8 + 12 + 4;
  const uint64_t meta = hb + 4ull * n_refs + 2 * (sb + hb + (uint64_t)n_refs * (hb + sb));
  const uint64_t data0 = head0 + nb * (hb + sb) + meta;
  const uint64_t rec0 = data0 + total + hdr.size();
  if (sb == 4 && nb && rec0 + rec_at[nb - 1] > 0xFFFFFFFFull)
    return set_error(h, MH_ERR_RANGE,
                     "mh_index_save_mhm: file too large for 4-byte stream positions");
  HostBuf<uint8_t> head(data0);
  uint8_t *o = head.data();
  std::memcpy(o, kIndicator, kIndicatorLen);
  o += kIndicatorLen;
  o = put_le(o, nb, 8);
  o = put_le(o, (uint32_t)kVersion, 4);
  o = put_le(o, hb, 4);
  o = put_le(o, (uint32_t)idx->threshold, 4);
  *o++ = (uint8_t)sb;
  *o++ = (uint8_t)nbytes(max_bucket);
  *o++ = (uint8_t)nbytes(max_seq);
  *o++ = (uint8_t)nbytes(max_hdr);
  parallel_for(nb, [&](uint64_t b) {
    uint8_t *e = head.data() + head0 + b * (hb + sb);
    e = put_le(e, hashes[b], hb);
    put_le(e, rec0 + rec_at[b], sb);
  }, 65536);
  o = head.data() + head0 + nb * (hb + sb);
  o = put_le(o, n_refs, hb);
  for (uint32_t r = 0; r < n_refs; r++) o = put_le(o, (uint32_t)(h_taxids ? h_taxids[r] : -1), 4);
  uint64_t dpos = data0;
  for (int part = 0; part < 2; part++) {
    const uint64_t *off = part ? hdr_off.data() : idx->h_ref_off.data();
    o = put_le(o, (uint64_t)(o - head.data()) + sb + hb + (uint64_t)n_refs * (hb + sb), sb);
    o = put_le(o, n_refs, hb);
    for (uint32_t r = 0; r < n_refs; r++) o = put_le(o, off[r + 1] - off[r], hb);
    for (uint32_t r = 0; r < n_refs; r++) {
      o = put_le(o, dpos, sb);
      dpos += off[r + 1] - off[r];
    }
  }
  if ((uint64_t)(o - head.data()) != data0 || dpos != rec0)
    return set_error(h, MH_ERR_STATE, "mh_index_save_mhm: layout mismatch");

  HostBuf<uint8_t> recs(rec_at[nb]);
  parallel_for(nb, [&](uint64_t b) {
    write_record(recs.data() + rec_at[b], bucket(b), bucket_n(b), plan[b]);
  });

  // Four segments written at their offsets by parallel pwrite calls.
  const int fd = open(path, O_WRONLY | O_CREAT | O_TRUNC | O_CLOEXEC, 0644);
  if (fd < 0) return set_error(h, MH_ERR_ARG, std::string("mh_index_save_mhm: cannot create ") + path);
  struct Seg {
    const uint8_t *p;
    uint64_t n, at;
  };
  const Seg segs[4] = {{head.data(), head.size(), 0},
                       {bytes.data(), bytes.size(), data0},
                       {hdr.data(), hdr.size(), data0 + total},
                       {recs.data(), recs.size(), rec0}};
  constexpr uint64_t kPiece = 64ull << 20;
  std::vector<Seg> pieces;
  for (const Seg &g : segs)
    for (uint64_t x = 0; x < g.n; x += kPiece)
      pieces.push_back({g.p + x, std::min(kPiece, g.n - x), g.at + x});
  std::atomic<int> werr{0};
  parallel_for(pieces.size(), [&](uint64_t i) {
    const Seg &g = pieces[i];
    uint64_t done = 0;
    while (done < g.n) {
      const ssize_t r = pwrite(fd, g.p + done, g.n - done, (off_t)(g.at + done));
      if (r <= 0) {
        werr = 1;
        return;
      }
      done += (uint64_t)r;
    }
  }, 1);
  const bool ok = (close(fd) == 0) && werr == 0;
  if (!ok) return set_error(h, MH_ERR_ARG, std::string("mh_index_save_mhm: write failed: ") + path);
  return MH_OK;
}

}  // extern "C"
