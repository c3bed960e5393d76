// FASTA text to sequences and headers on the device, as MirAge reads it
// (Sequences::ReadFastaSequencesFromMem / TReadFastaSequencesFromMem,
// MirAge/Sequences.cpp:57-211, :283-562): the text must start with '>'
// (otherwise the reference finds no header and stops, :286-298); a header
// line is kept with its '>' and a NUL in place of a final '\r', or appended
// (:117-127); every other line is sequence, appended without a final '\r'
// (:143-151) and uppercased (SequenceToUpperCase, :40-52); a last line
// without a newline is read too.
// Accepted texts: each header followed by a non-empty sequence line, no
// empty line, no line starting with '@', no NUL byte.  The reference reads
// past its buffers or pairs headers and sequences by thread layout outside
// that set (:102-118, :356-380); here such a text is an error naming the
// line.
//
// Passes: newline counts per 16 KiB wave tile -> scan -> newline positions
// in order; per line its kind and output sizes, scanned per block of 1024
// lines, then the blocks -> the copy.  Every pass is wavefront-local (no
// workgroup barrier): a wavefront steps through its tile 1 KiB at a time,
// each lane on 16 consecutive bytes read in one load, and a wave scan of the
// newline counts gives every lane the line its bytes start in.  Scratch comes
// from the handle's arena.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <string>
#include <vector>

#include "mh_internal.hpp"

namespace mh {
namespace {

constexpr unsigned kFaThreads = 256;
constexpr unsigned kFaWaves = kFaThreads / 64;
constexpr unsigned kFaVec = 16;                      // bytes per lane and step
constexpr unsigned kFaSteps = 16;
constexpr uint64_t kFaStep = 64 * kFaVec;            // 1 KiB per wavefront step
constexpr uint64_t kFaTile = kFaStep * kFaSteps;     // 16 KiB per wavefront
constexpr unsigned kLineShift = 10;                  // lines per line block: 1024
constexpr uint64_t kLineBlock = 1ull << kLineShift;  // (64 lanes x 16 steps)

enum : uint32_t { kLineSkip = 0, kLineHeader = 1, kLineSeq = 2 };
enum : uint32_t { kErrNul = 1, kErrAt = 2, kErrEmpty = 3, kErrHeaderAlone = 4, kErrNoHeader = 5 };

// Per line: bytes of sequence, bytes of header (with its NUL), 1 for a header
// line.
struct Len3 {
  unsigned long long s, h, r;
};
__device__ __forceinline__ Len3 operator+(const Len3 &a, const Len3 &b) {
  return Len3{a.s + b.s, a.h + b.h, a.r + b.r};
}

__device__ inline void fa_error(unsigned long long *err, uint64_t line, uint32_t code) {
  atomicMin(err, (unsigned long long)(line << 3 | code));
}

// 0x80 in every byte of x that is zero, 0 elsewhere (exact: no borrow
// crosses a byte).
__device__ __forceinline__ uint32_t zero_bytes(uint32_t x) {
  return ~(((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x | 0x7F7F7F7Fu);
}

// Bit i set where byte i of the 16 equals the byte in `pattern` (repeated 4x).
__device__ __forceinline__ uint32_t match16(const uint32_t w[4], uint32_t pattern) {
  uint32_t m = 0;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const uint32_t z = zero_bytes(w[i] ^ pattern);
    m |= (((z >> 7) & 1u) | ((z >> 14) & 2u) | ((z >> 21) & 4u) | ((z >> 28) & 8u)) << (4 * i);
  }
  return m;
}

// The 16 bytes at t + p; bytes at n and beyond read as 'A' (no newline, no
// NUL).  `aligned`: t is 16-byte aligned, so a whole chunk is one load.
__device__ __forceinline__ void load16(const uint8_t *t, uint64_t n, uint64_t p, bool aligned,
                                       uint32_t w[4]) {
  if (aligned && p + kFaVec <= n) {
    const uint4 v = *reinterpret_cast<const uint4 *>(t + p);
    w[0] = v.x;
    w[1] = v.y;
    w[2] = v.z;
    w[3] = v.w;
    return;
  }
#pragma unroll
  for (int i = 0; i < 4; i++) {
    uint32_t x = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const uint64_t q = p + 4 * i + j;
      x |= (uint32_t)(q < n ? t[q] : (uint8_t)'A') << (8 * j);
    }
    w[i] = x;
  }
}

__device__ __forceinline__ uint8_t byte_of(const uint32_t w[4], unsigned i) {
  return (uint8_t)(w[i >> 2] >> (8 * (i & 3)));
}

// Exclusive scans over the wavefront; `total` gets the sum.
template <class T>
__device__ __forceinline__ T wave_scan(T c, T &total) {
  const unsigned lane = threadIdx.x & 63;
  T x = c;
#pragma unroll
  for (unsigned d = 1; d < 64; d <<= 1) {
    const T y = __shfl_up(x, d);
    if (lane >= d) x += y;
  }
  total = __shfl(x, 63);
  return x - c;
}
__device__ __forceinline__ Len3 wave_scan3(const Len3 &v, Len3 &total) {
  Len3 e;
  e.s = wave_scan(v.s, total.s);
  e.h = wave_scan(v.h, total.h);
  e.r = wave_scan(v.r, total.r);
  return e;
}

// Newlines per wave tile; the first NUL byte's offset.
__global__ __launch_bounds__(kFaThreads) void k_fa_count(const uint8_t *t, uint64_t n,
                                                         uint64_t n_tiles, int aligned,
                                                         unsigned long long *tile_nl,
                                                         unsigned long long *nul_at) {
  const uint64_t wt = (uint64_t)blockIdx.x * kFaWaves + (threadIdx.x >> 6);
  if (wt >= n_tiles) return;
  const uint64_t p0 = wt * kFaTile + (threadIdx.x & 63) * kFaVec;
  uint32_t c = 0;
  for (unsigned s = 0; s < kFaSteps; s++) {
    const uint64_t p = p0 + s * kFaStep;
    if (p >= n) break;
    uint32_t w[4];
    load16(t, n, p, aligned, w);
    c += __popc(match16(w, 0x0A0A0A0Au));
    const uint32_t z = match16(w, 0);
    if (z) atomicMin(nul_at, (unsigned long long)(p + __builtin_ctz(z)));
  }
  for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o);
  if ((threadIdx.x & 63) == 0) tile_nl[wt] = c;
}

// Newline positions in text order, from the tile's base (the scan of the
// tile counts) and a wave scan per step.
__global__ __launch_bounds__(kFaThreads) void k_fa_newlines(const uint8_t *t, uint64_t n,
                                                            uint64_t n_tiles, int aligned,
                                                            const unsigned long long *tile_base,
                                                            uint64_t *nl) {
  const uint64_t wt = (uint64_t)blockIdx.x * kFaWaves + (threadIdx.x >> 6);
  if (wt >= n_tiles) return;
  const uint64_t p0 = wt * kFaTile + (threadIdx.x & 63) * kFaVec;
  uint64_t base = tile_base[wt];
  for (unsigned s = 0; s < kFaSteps; s++) {
    const uint64_t p = p0 + s * kFaStep;
    if (wt * kFaTile + s * kFaStep >= n) break;  // uniform over the wavefront
    uint32_t w[4];
    load16(t, n, p, aligned, w);
    uint32_t m = p < n ? match16(w, 0x0A0A0A0Au) : 0;
    uint32_t total;
    uint64_t o = base + wave_scan<uint32_t>(__popc(m), total);
    while (m) {
      nl[o++] = p + __builtin_ctz(m);
      m &= m - 1;
    }
    base += total;
  }
}

__device__ inline void fa_line(const uint64_t *nl, uint64_t n_nl, uint64_t n, uint64_t l,
                               uint64_t &a, uint64_t &e) {
  a = l == 0 ? 0 : nl[l - 1] + 1;
  e = l < n_nl ? nl[l] : n;
}

// Per line: info = body length << 2 | kind, errors by line, and the line's
// Len3 scanned within its block of 1024 lines (a wavefront per block); the
// block sums go to blk.
__global__ __launch_bounds__(kFaThreads) void k_fa_lines(const uint8_t *t, const uint64_t *nl,
                                                         uint64_t n_nl, uint64_t n,
                                                         uint64_t n_lines, uint64_t *info,
                                                         Len3 *pos, Len3 *blk,
                                                         unsigned long long *err) {
  const uint64_t b = (uint64_t)blockIdx.x * kFaWaves + (threadIdx.x >> 6);
  const uint64_t l0 = b << kLineShift;
  if (l0 >= n_lines) return;
  Len3 carry{0, 0, 0};
  for (unsigned s = 0; s < kFaSteps; s++) {
    if (l0 + s * 64 >= n_lines) break;  // uniform over the wavefront
    const uint64_t l = l0 + s * 64 + (threadIdx.x & 63);
    Len3 v{0, 0, 0};
    if (l < n_lines) {
      uint64_t a, e;
      fa_line(nl, n_nl, n, l, a, e);
      uint64_t body = e - a;
      if (body && t[e - 1] == '\r') body--;
      uint32_t k = kLineSkip;
      if (body == 0) {
        fa_error(err, l, kErrEmpty);
      } else if (t[a] == '@') {
        fa_error(err, l, kErrAt);
      } else if (t[a] == '>') {
        k = kLineHeader;
        v.h = body + 1;
        v.r = 1;
        // the next line must be sequence
        if (l + 1 >= n_lines) {
          fa_error(err, l, kErrHeaderAlone);
        } else {
          uint64_t a2, e2;
          fa_line(nl, n_nl, n, l + 1, a2, e2);
          if (e2 > a2 && t[a2] == '>') fa_error(err, l, kErrHeaderAlone);
        }
      } else {
        k = kLineSeq;
        v.s = body;
      }
      info[l] = body << 2 | k;
    }
    Len3 tot;
    const Len3 ex = wave_scan3(v, tot);
    if (l < n_lines) pos[l] = carry + ex;
    carry = carry + tot;
  }
  if ((threadIdx.x & 63) == 0) blk[b] = carry;
}

// Exclusive scan of the line blocks' sums in place (one workgroup); the
// total goes to *sum.
__global__ __launch_bounds__(kFaThreads) void k_fa_scan_blocks(Len3 *blk, uint64_t nb,
                                                               Len3 *sum) {
  __shared__ Len3 part[kFaWaves];
  Len3 carry{0, 0, 0};
  for (uint64_t b0 = 0; b0 < nb; b0 += kFaThreads) {
    const uint64_t b = b0 + threadIdx.x;
    const Len3 v = b < nb ? blk[b] : Len3{0, 0, 0};
    Len3 wt;
    const Len3 ex = wave_scan3(v, wt);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = wt;
    __syncthreads();
    Len3 before{0, 0, 0}, tot{0, 0, 0};
    for (unsigned i = 0; i < kFaWaves; i++) {
      if (i < (threadIdx.x >> 6)) before = before + part[i];
      tot = tot + part[i];
    }
    __syncthreads();  // part is written again by the next chunk
    if (b < nb) blk[b] = carry + before + ex;
    carry = carry + tot;
  }
  if (threadIdx.x == 0) *sum = carry;
}

__device__ __forceinline__ uint8_t fa_upper(uint8_t c) {
  return (int8_t)c > 96 ? (uint8_t)(c - 32) : c;
}

// Orders one wavefront's LDS writes before its later LDS reads (the
// wavefront runs in lockstep; this keeps the compiler from reordering).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// What the copy needs of line l, relative to the window starting at w0:
// `off` such that the byte at window offset j goes to offset off + j of the
// sequence or header output, and packed
// 12-bit fields: the window offsets of the body end and of the line start
// (each clamped to [-1, 2047] and stored + 1), and the kind.  `rec` holds a
// header's record number and offsets.
struct LineSlot {
  uint64_t off;
  uint32_t packed;
  uint32_t pad;
};
__device__ __forceinline__ uint32_t clamp12(int64_t v) {
  return (uint32_t)(v < -1 ? 0 : v > 2047 ? 2048 : v + 1);
}
__device__ __forceinline__ void line_slot(const uint64_t *nl, const uint64_t *info,
                                          const Len3 *pos, const Len3 *blk, uint64_t l,
                                          uint64_t w0, LineSlot &x, Len3 &rec) {
  const uint64_t a = l == 0 ? 0 : nl[l - 1] + 1;
  const uint64_t f = info[l];
  rec = pos[l] + blk[l >> kLineShift];
  const uint32_t kind = (uint32_t)(f & 3);
  const uint64_t body = f >> 2;
  x.off = (kind == kLineSeq ? rec.s : rec.h) + (w0 - a);
  x.packed = clamp12((int64_t)(a + body - w0)) | clamp12((int64_t)(a - w0)) << 12 | kind << 24;
  x.pad = 0;
}

// The copy, a wavefront per 16 KiB tile and 1 KiB window per step: the
// window's bytes go to LDS, and the lines it touches (up to 64; more are read
// from memory) to an LDS table, one lane per line.  Then 16 rounds of 64
// consecutive bytes: a ballot of the newlines gives each lane its line
// (mbcnt), so consecutive lanes store to consecutive output bytes.
// Sequence bytes are uppercased; a header gets a NUL after its body, and
// its record's offsets are written at its first byte.
__global__ __launch_bounds__(kFaThreads) void k_fa_copy(
    const uint8_t *t, uint64_t n, uint64_t n_tiles, int aligned,
    const unsigned long long *tile_base, const uint64_t *nl, uint64_t n_lines,
    const uint64_t *info, const Len3 *pos, const Len3 *blk, uint8_t *seq, uint8_t *hdr,
    uint64_t *seq_off, uint64_t *hdr_off) {
  __shared__ uint4 win_all[kFaWaves][64];
  __shared__ LineSlot slot_all[kFaWaves][64];
  __shared__ Len3 rec_all[kFaWaves][64];
  const unsigned lane = threadIdx.x & 63;
  const uint64_t wt = (uint64_t)blockIdx.x * kFaWaves + (threadIdx.x >> 6);
  if (wt >= n_tiles) return;
  uint4 *win = win_all[threadIdx.x >> 6];
  LineSlot *slots = slot_all[threadIdx.x >> 6];
  Len3 *recs = rec_all[threadIdx.x >> 6];
  const uint8_t *win_bytes = reinterpret_cast<const uint8_t *>(win);
  uint64_t base = tile_base[wt];  // lines before the window
  uint32_t w[4], wn[4];
  load16(t, n, wt * kFaTile + lane * kFaVec, aligned, w);
  for (unsigned s = 0; s < kFaSteps; s++) {
    const uint64_t w0 = wt * kFaTile + s * kFaStep;  // window start
    if (w0 >= n) break;                               // uniform over the wavefront
    if (s + 1 < kFaSteps) load16(t, n, w0 + kFaStep + lane * kFaVec, aligned, wn);
    const uint32_t lim = n - w0 < kFaStep ? (uint32_t)(n - w0) : (uint32_t)kFaStep;
    uint32_t total;
    wave_scan<uint32_t>(__popc(match16(w, 0x0A0A0A0Au)), total);  // newlines in the window
    if (lane <= total && base + lane < n_lines)
      line_slot(nl, info, pos, blk, base + lane, w0, slots[lane], recs[lane]);
    win[lane] = make_uint4(w[0], w[1], w[2], w[3]);
    wave_sync();
    uint32_t run = 0;  // newlines in the rounds before
    for (unsigned r = 0; r < kFaVec; r++) {
      const uint32_t j = r * 64 + lane;  // window offset
      const uint8_t x = win_bytes[j];
      const bool valid = j < lim;
      const uint64_t nlm = __ballot(valid && x == '\n');
      const uint32_t before =
          run + __builtin_amdgcn_mbcnt_hi((uint32_t)(nlm >> 32),
                                          __builtin_amdgcn_mbcnt_lo((uint32_t)nlm, 0));
      run += (uint32_t)__popcll(nlm);
      if (!valid) continue;
      LineSlot ls;
      Len3 rc;
      if (before < 64) {
        ls = slots[before];
      } else {
        line_slot(nl, info, pos, blk, base + before, w0, ls, rc);
      }
      const uint32_t kind = ls.packed >> 24;
      const int32_t end = (int32_t)(ls.packed & 0xFFF) - 1;
      const int32_t jj = (int32_t)j;
      const bool is_seq = kind == kLineSeq, is_hdr = kind == kLineHeader;
      // one store per byte at most: a body byte, or a header's NUL
      if ((is_seq || is_hdr) && (jj < end || (is_hdr && jj == end)))
        (is_seq ? seq : hdr)[ls.off + j] = is_seq ? fa_upper(x) : jj < end ? x : (uint8_t)0;
      if (is_hdr && jj == (int32_t)((ls.packed >> 12) & 0xFFF) - 1) {
        if (before < 64) rc = recs[before];
        seq_off[rc.r] = rc.s;
        hdr_off[rc.r] = rc.h;
      }
    }
    wave_sync();  // the window and tables are written again next step
    base += total;
#pragma unroll
    for (int k = 0; k < 4; k++) w[k] = wn[k];
  }
}

__global__ void k_fa_ends(uint64_t *seq_off, uint64_t *hdr_off, uint64_t records,
                          uint64_t seq_bytes, uint64_t hdr_bytes) {
  seq_off[records] = seq_bytes;
  hdr_off[records] = hdr_bytes;
}

unsigned fa_blocks(uint64_t waves) { return (unsigned)((waves + kFaWaves - 1) / kFaWaves); }

size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace
}  // namespace mh

using namespace mh;

extern "C" mh_status mh_fasta_parse(mh_handle *h, const uint8_t *d_text, uint64_t n_bytes,
                                    uint8_t *d_seq, uint64_t seq_cap, uint64_t *d_seq_off,
                                    uint8_t *d_hdr, uint64_t hdr_cap, uint64_t *d_hdr_off,
                                    uint64_t rec_cap, uint64_t *n_records, uint64_t *seq_bytes,
                                    uint64_t *hdr_bytes) {
  if (!h || (n_bytes && !d_text) || !n_records || !seq_bytes || !hdr_bytes)
    return set_error(h, MH_ERR_ARG, "mh_fasta_parse: bad argument");
  std::lock_guard<std::mutex> lk(h->mu);
  MH_HIP(h, hipSetDevice(h->device));
  Exec &x = h->ex;
  hipStream_t st = x.stream;
  const uint64_t n = n_bytes;
  if (n == 0) return set_error(h, MH_ERR_ARG, "mh_fasta_parse: the text does not start with '>'");
  const uint64_t tiles = (n + kFaTile - 1) / kFaTile;
  if (tiles >= (1ull << 31)) return set_error(h, MH_ERR_ARG, "mh_fasta_parse: text too large");
  const int aligned = ((uintptr_t)d_text & 15) == 0;

  // Scratch, from the handle's arena: stage 1 (tile counts and bases, error
  // words and the line total, scan storage), then stage 2 (newlines, line
  // info, line offsets, line block offsets) once the line count is known.
  size_t tmp1 = 0;
  MH_HIP(h, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp1, (unsigned long long *)nullptr,
                                             (unsigned long long *)nullptr, (int)tiles, st));
  const size_t s1 = 2 * up256(tiles * 8) + 256 + up256(tmp1);
  mh_status rs = ensure_scratch(h, x, s1);
  if (rs) return rs;
  unsigned long long *tile_nl = nullptr, *tile_base = nullptr, *misc = nullptr;
  Len3 *sum = nullptr;
  void *d_tmp1 = nullptr;
  auto carve1 = [&]() {
    char *b = (char *)x.d_scratch;
    tile_nl = (unsigned long long *)b;
    tile_base = (unsigned long long *)(b + up256(tiles * 8));
    misc = (unsigned long long *)(b + 2 * up256(tiles * 8));  // [0] NUL offset, [1] error
    sum = (Len3 *)(misc + 2);
    d_tmp1 = b + 2 * up256(tiles * 8) + 256;
  };
  auto stage1 = [&]() -> hipError_t {
    hipError_t e = hipMemsetAsync(misc, 0xFF, 2 * sizeof(unsigned long long), st);
    if (e != hipSuccess) return e;
    k_fa_count<<<fa_blocks(tiles), kFaThreads, 0, st>>>(d_text, n, tiles, aligned, tile_nl, misc);
    size_t tb = tmp1;
    return hipcub::DeviceScan::ExclusiveSum(d_tmp1, tb, tile_nl, tile_base, (int)tiles, st);
  };
  carve1();
  MH_HIP(h, stage1());
  unsigned long long last_count = 0, last_base = 0, nul_at = 0;
  uint8_t first_byte = 0, last_byte = 0;
  MH_HIP(h, hipMemcpyAsync(&last_count, tile_nl + tiles - 1, 8, hipMemcpyDeviceToHost, st));
  MH_HIP(h, hipMemcpyAsync(&last_base, tile_base + tiles - 1, 8, hipMemcpyDeviceToHost, st));
  MH_HIP(h, hipMemcpyAsync(&nul_at, misc, 8, hipMemcpyDeviceToHost, st));
  MH_HIP(h, hipMemcpyAsync(&first_byte, d_text, 1, hipMemcpyDeviceToHost, st));
  MH_HIP(h, hipMemcpyAsync(&last_byte, d_text + n - 1, 1, hipMemcpyDeviceToHost, st));
  MH_HIP(h, hipStreamSynchronize(st));
  MH_HIP(h, hipGetLastError());
  if (nul_at != ~0ull)
    return set_error(h, MH_ERR_ARG, "mh_fasta_parse: NUL byte at offset " + std::to_string(nul_at));
  if (first_byte != '>')  // the reference finds 0 headers and stops
    return set_error(h, MH_ERR_ARG, "mh_fasta_parse: the text does not start with '>'");
  const uint64_t n_nl = last_count + last_base;
  // Lines: one per newline, plus the text after the last newline.
  const uint64_t n_lines = n_nl + (last_byte != '\n' ? 1 : 0);
  const uint64_t n_lblocks = (n_lines + kLineBlock - 1) >> kLineShift;

  const size_t s2 = up256(std::max<uint64_t>(n_nl, 1) * 8) + up256(n_lines * 8) +
                    up256(n_lines * sizeof(Len3)) + up256(n_lblocks * sizeof(Len3));
  if (x.scratch_bytes < s1 + s2) {  // a new arena: stage 1 again into it
    rs = ensure_scratch(h, x, s1 + s2);
    if (rs) return rs;
    carve1();
    MH_HIP(h, stage1());
  }
  char *b2 = (char *)x.d_scratch + s1;
  uint64_t *nl = (uint64_t *)b2;
  b2 += up256(std::max<uint64_t>(n_nl, 1) * 8);
  uint64_t *info = (uint64_t *)b2;
  b2 += up256(n_lines * 8);
  Len3 *pos = (Len3 *)b2;
  b2 += up256(n_lines * sizeof(Len3));
  Len3 *blk = (Len3 *)b2;

  k_fa_newlines<<<fa_blocks(tiles), kFaThreads, 0, st>>>(d_text, n, tiles, aligned, tile_base,
                                                         nl);
  k_fa_lines<<<fa_blocks(n_lblocks), kFaThreads, 0, st>>>(d_text, nl, n_nl, n, n_lines, info, pos,
                                                          blk, misc + 1);
  k_fa_scan_blocks<<<1, kFaThreads, 0, st>>>(blk, n_lblocks, sum);
  unsigned long long err = 0;
  Len3 total{};
  MH_HIP(h, hipMemcpyAsync(&err, misc + 1, 8, hipMemcpyDeviceToHost, st));
  MH_HIP(h, hipMemcpyAsync(&total, sum, sizeof(Len3), hipMemcpyDeviceToHost, st));
  MH_HIP(h, hipStreamSynchronize(st));
  MH_HIP(h, hipGetLastError());
  if (err != ~0ull) {
    static const char *what[] = {"", "NUL byte", "line starting with '@'", "empty line",
                                 "header without a sequence line", "no header"};
    return set_error(h, MH_ERR_ARG, "mh_fasta_parse: line " + std::to_string((err >> 3) + 1) +
                                        ": " + what[err & 7]);
  }
  *n_records = total.r;
  *seq_bytes = total.s;
  *hdr_bytes = total.h;
  if (!d_seq || !d_seq_off || !d_hdr || !d_hdr_off) return MH_OK;  // sizes only
  if (seq_cap < total.s || hdr_cap < total.h || rec_cap < total.r)
    return set_error(h, MH_ERR_OVERFLOW, "mh_fasta_parse: output too small");
  k_fa_copy<<<fa_blocks(tiles), kFaThreads, 0, st>>>(d_text, n, tiles, aligned, tile_base, nl,
                                                     n_lines, info, pos, blk, d_seq, d_hdr,
                                                     d_seq_off, d_hdr_off);
  k_fa_ends<<<1, 1, 0, st>>>(d_seq_off, d_hdr_off, total.r, total.s, total.h);
  MH_HIP(h, hipStreamSynchronize(st));
  MH_HIP(h, hipGetLastError());
  return MH_OK;
}
