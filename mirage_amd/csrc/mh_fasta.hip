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
// Passes: newline counts per 64 KiB tile -> scan -> newline positions in
// order (each thread a contiguous 256-byte slice); per line its kind and
// output sizes; scans over lines give each line's output offsets and record
// index; a wavefront per line copies it (lanes over the bytes, so a long
// unwrapped sequence line is shared by 64 lanes).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <string>
#include <vector>

#include "mh_internal.hpp"

namespace mh {
namespace {

constexpr unsigned kFaThreads = 256;
constexpr uint64_t kFaSlice = 256;                        // bytes per thread
constexpr uint64_t kFaTile = kFaThreads * kFaSlice;       // bytes per workgroup

enum : uint32_t { kLineSkip = 0, kLineHeader = 1, kLineSeq = 2 };
enum : uint32_t { kErrNul = 1, kErrAt = 2, kErrEmpty = 3, kErrHeaderAlone = 4, kErrNoHeader = 5 };

__device__ inline void fa_error(unsigned long long *err, uint64_t line, uint32_t code) {
  atomicMin(err, (unsigned long long)(line << 3 | code));
}

// Newlines per tile; a NUL byte is an error (its line is found later).
__global__ __launch_bounds__(kFaThreads) void k_fa_count(const uint8_t *t, uint64_t n,
                                                         unsigned long long *tile_nl,
                                                         unsigned long long *nul_at) {
  __shared__ unsigned long long part[kFaThreads / 64];
  const uint64_t b0 = blockIdx.x * kFaTile + threadIdx.x * kFaSlice;
  unsigned long long c = 0;
  for (uint64_t i = b0; i < b0 + kFaSlice && i < n; i++) {
    const uint8_t x = t[i];
    c += x == '\n';
    if (x == 0) atomicMin(nul_at, (unsigned long long)i);
  }
  for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long s = 0;
    for (unsigned w = 0; w < kFaThreads / 64; w++) s += part[w];
    tile_nl[blockIdx.x] = s;
  }
}

// Newline positions in file order: tile base from the scan, then an
// exclusive scan of the threads' slice counts.
__global__ __launch_bounds__(kFaThreads) void k_fa_newlines(const uint8_t *t, uint64_t n,
                                                            const unsigned long long *tile_base,
                                                            uint64_t *nl) {
  __shared__ unsigned long long sc[kFaThreads];
  const uint64_t b0 = blockIdx.x * kFaTile + threadIdx.x * kFaSlice;
  unsigned long long c = 0;
  for (uint64_t i = b0; i < b0 + kFaSlice && i < n; i++) c += t[i] == '\n';
  sc[threadIdx.x] = c;
  __syncthreads();
  for (unsigned d = 1; d < kFaThreads; d <<= 1) {
    const unsigned long long a = threadIdx.x >= d ? sc[threadIdx.x - d] : 0;
    __syncthreads();
    sc[threadIdx.x] += a;
    __syncthreads();
  }
  unsigned long long o = tile_base[blockIdx.x] + sc[threadIdx.x] - c;
  for (uint64_t i = b0; i < b0 + kFaSlice && i < n; i++)
    if (t[i] == '\n') nl[o++] = i;
}

__device__ inline void fa_line(const uint64_t *nl, uint64_t n_nl, uint64_t n, uint64_t l,
                               uint64_t &a, uint64_t &e) {
  a = l == 0 ? 0 : nl[l - 1] + 1;
  e = l < n_nl ? nl[l] : n;
}

// First line that starts with '>' (reading starts there).
__global__ void k_fa_first(const uint8_t *t, const uint64_t *nl, uint64_t n_nl, uint64_t n,
                           uint64_t n_lines, unsigned long long *first) {
  for (uint64_t l = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; l < n_lines;
       l += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t a, e;
    fa_line(nl, n_nl, n, l, a, e);
    if (e > a && t[a] == '>') atomicMin(first, (unsigned long long)l);
  }
}

// Per line: kind, bytes of sequence, bytes of header (with its NUL), and
// 1 for a header (the record count).
__global__ void k_fa_lines(const uint8_t *t, const uint64_t *nl, uint64_t n_nl, uint64_t n,
                           uint64_t n_lines, uint64_t first, uint32_t *kind,
                           unsigned long long *seq_len, unsigned long long *hdr_len,
                           unsigned long long *is_hdr, unsigned long long *err) {
  for (uint64_t l = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; l < n_lines;
       l += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t a, e;
    fa_line(nl, n_nl, n, l, a, e);
    uint64_t body = e - a;
    if (body && t[e - 1] == '\r') body--;
    uint32_t k = kLineSkip;
    unsigned long long sl = 0, hl = 0, ih = 0;
    if (l >= first) {
      if (body == 0) {
        fa_error(err, l, kErrEmpty);
      } else if (t[a] == '@') {
        fa_error(err, l, kErrAt);
      } else if (t[a] == '>') {
        k = kLineHeader;
        hl = body + 1;
        ih = 1;
        // the next line must be sequence
        uint64_t a2, e2;
        if (l + 1 >= n_lines) {
          fa_error(err, l, kErrHeaderAlone);
        } else {
          fa_line(nl, n_nl, n, l + 1, a2, e2);
          if (e2 > a2 && t[a2] == '>') fa_error(err, l, kErrHeaderAlone);
        }
      } else {
        k = kLineSeq;
        sl = body;
      }
    }
    kind[l] = k;
    seq_len[l] = sl;
    hdr_len[l] = hl;
    is_hdr[l] = ih;
  }
}

__device__ __forceinline__ uint8_t fa_upper(uint8_t c) {
  return (int8_t)c > 96 ? (uint8_t)(c - 32) : c;
}

// A wavefront per line: sequence bytes uppercased, header bytes with a NUL.
__global__ __launch_bounds__(256) void k_fa_copy(
    const uint8_t *t, const uint64_t *nl, uint64_t n_nl, uint64_t n, uint64_t n_lines,
    const uint32_t *kind, const unsigned long long *seq_pos, const unsigned long long *hdr_pos,
    const unsigned long long *rec, uint8_t *seq, uint8_t *hdr, uint64_t *seq_off,
    uint64_t *hdr_off) {
  const unsigned lane = threadIdx.x & 63;
  const uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
  const uint64_t waves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t l = wave; l < n_lines; l += waves) {
    const uint32_t k = kind[l];
    if (k == kLineSkip) continue;
    uint64_t a, e;
    fa_line(nl, n_nl, n, l, a, e);
    uint64_t body = e - a;
    if (body && t[e - 1] == '\r') body--;
    if (k == kLineSeq) {
      uint8_t *o = seq + seq_pos[l];
      for (uint64_t i = lane; i < body; i += 64) o[i] = fa_upper(t[a + i]);
    } else {
      uint8_t *o = hdr + hdr_pos[l];
      for (uint64_t i = lane; i < body; i += 64) o[i] = t[a + i];
      if (lane == 0) {
        o[body] = 0;
        const uint64_t r = rec[l];
        seq_off[r] = seq_pos[l];
        hdr_off[r] = hdr_pos[l];
      }
    }
  }
}

unsigned fa_grid(uint64_t n, unsigned block, unsigned cap = 65536) {
  const uint64_t g = (n + block - 1) / block;
  return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(g, cap));
}

struct ToU64 {
  __host__ __device__ unsigned long long operator()(unsigned long long v) const { return v; }
};

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
  hipSetDevice(h->device);
  hipStream_t st = h->stream;
  const uint64_t n = n_bytes;
  const uint64_t tiles = std::max<uint64_t>(1, (n + kFaTile - 1) / kFaTile);

  // scratch: tile counts and bases, misc words, then the line arrays
  unsigned long long *d_tiles = nullptr, *d_misc = nullptr;
  uint64_t *d_nl = nullptr;
  uint32_t *d_kind = nullptr;
  unsigned long long *d_len = nullptr;  // 3 x n_lines: seq, hdr, record; then 3 x positions
  void *d_tmp = nullptr;
  auto cleanup = [&]() {
    hipStreamSynchronize(st);
    for (void *p : {(void *)d_tiles, (void *)d_misc, (void *)d_nl, (void *)d_kind,
                    (void *)d_len, d_tmp})
      if (p) hipFree(p);
  };
  auto fail = [&](mh_status s) {
    cleanup();
    return s;
  };
#define FA_HIP(call)                                           \
  do {                                                         \
    hipError_t e_ = (call);                                    \
    if (e_ != hipSuccess) return fail(hip_error(h, e_, #call)); \
  } while (0)

  FA_HIP(hipMalloc(&d_tiles, 2 * tiles * sizeof(unsigned long long)));
  FA_HIP(hipMalloc(&d_misc, 4 * sizeof(unsigned long long)));
  const unsigned long long init[4] = {~0ull, ~0ull, ~0ull, 0};  // nul, first header, error
  FA_HIP(hipMemcpyAsync(d_misc, init, sizeof(init), hipMemcpyHostToDevice, st));
  if (n)
    k_fa_count<<<(unsigned)tiles, kFaThreads, 0, st>>>(d_text, n, d_tiles, d_misc);
  else
    FA_HIP(hipMemsetAsync(d_tiles, 0, tiles * sizeof(unsigned long long), st));
  {
    size_t tb = 0;
    hipcub::DeviceScan::ExclusiveSum(nullptr, tb, d_tiles, d_tiles + tiles, (int)tiles, st);
    FA_HIP(hipMalloc(&d_tmp, std::max<size_t>(tb, 16)));
    hipcub::DeviceScan::ExclusiveSum(d_tmp, tb, d_tiles, d_tiles + tiles, (int)tiles, st);
  }
  unsigned long long last_count = 0, last_base = 0, nul_at = 0;
  FA_HIP(hipMemcpyAsync(&last_count, d_tiles + tiles - 1, 8, hipMemcpyDeviceToHost, st));
  FA_HIP(hipMemcpyAsync(&last_base, d_tiles + 2 * tiles - 1, 8, hipMemcpyDeviceToHost, st));
  FA_HIP(hipMemcpyAsync(&nul_at, d_misc, 8, hipMemcpyDeviceToHost, st));
  FA_HIP(hipStreamSynchronize(st));
  const uint64_t n_nl = last_count + last_base;
  if (nul_at != ~0ull)
    return fail(set_error(h, MH_ERR_ARG,
                          "mh_fasta_parse: NUL byte at offset " + std::to_string(nul_at)));
  // Lines: one per newline, plus the text after the last newline.
  uint8_t tail = 0;
  if (n) FA_HIP(hipMemcpy(&tail, d_text + n - 1, 1, hipMemcpyDeviceToHost));
  const uint64_t n_lines = n_nl + ((n && tail != '\n') ? 1 : 0);
  if (n_lines == 0) return fail(set_error(h, MH_ERR_ARG, "mh_fasta_parse: no header"));
  FA_HIP(hipMalloc(&d_nl, std::max<uint64_t>(n_nl, 1) * sizeof(uint64_t)));
  FA_HIP(hipMalloc(&d_kind, n_lines * sizeof(uint32_t)));
  FA_HIP(hipMalloc(&d_len, 6 * n_lines * sizeof(unsigned long long)));
  if (n) k_fa_newlines<<<(unsigned)tiles, kFaThreads, 0, st>>>(d_text, n, d_tiles + tiles, d_nl);
  k_fa_first<<<fa_grid(n_lines, 256), 256, 0, st>>>(d_text, d_nl, n_nl, n, n_lines, d_misc + 1);
  unsigned long long first = 0;
  FA_HIP(hipMemcpyAsync(&first, d_misc + 1, 8, hipMemcpyDeviceToHost, st));
  FA_HIP(hipStreamSynchronize(st));
  if (first != 0)
    return fail(set_error(h, MH_ERR_ARG, "mh_fasta_parse: the text does not start with '>'"));
  unsigned long long *seq_len = d_len, *hdr_len = d_len + n_lines, *is_hdr = d_len + 2 * n_lines;
  unsigned long long *seq_pos = d_len + 3 * n_lines, *hdr_pos = d_len + 4 * n_lines,
                     *rec = d_len + 5 * n_lines;
  k_fa_lines<<<fa_grid(n_lines, 256), 256, 0, st>>>(d_text, d_nl, n_nl, n, n_lines, first, d_kind,
                                                    seq_len, hdr_len, is_hdr, d_misc + 2);
  {
    hipFree(d_tmp);
    d_tmp = nullptr;
    size_t tb = 0;
    hipcub::DeviceScan::ExclusiveSum(nullptr, tb, seq_len, seq_pos, (int)n_lines, st);
    FA_HIP(hipMalloc(&d_tmp, std::max<size_t>(tb, 16)));
    for (int k = 0; k < 3; k++)
      hipcub::DeviceScan::ExclusiveSum(d_tmp, tb, d_len + k * n_lines, d_len + (3 + k) * n_lines,
                                       (int)n_lines, st);
  }
  unsigned long long misc[4], tot[6];
  FA_HIP(hipMemcpyAsync(misc, d_misc, sizeof(misc), hipMemcpyDeviceToHost, st));
  for (int k = 0; k < 3; k++) {
    FA_HIP(hipMemcpyAsync(&tot[k], d_len + k * n_lines + n_lines - 1, 8, hipMemcpyDeviceToHost,
                          st));
    FA_HIP(hipMemcpyAsync(&tot[3 + k], d_len + (3 + k) * n_lines + n_lines - 1, 8,
                          hipMemcpyDeviceToHost, st));
  }
  FA_HIP(hipStreamSynchronize(st));
  if (misc[2] != ~0ull) {
    static const char *what[] = {"", "NUL byte", "line starting with '@'", "empty line",
                                 "header without a sequence line", "no header"};
    const uint64_t line = misc[2] >> 3;
    return fail(set_error(h, MH_ERR_ARG, "mh_fasta_parse: line " + std::to_string(line + 1) +
                                             ": " + what[misc[2] & 7]));
  }
  const uint64_t total_seq = tot[0] + tot[3], total_hdr = tot[1] + tot[4],
                 records = tot[2] + tot[5];
  *n_records = records;
  *seq_bytes = total_seq;
  *hdr_bytes = total_hdr;
  if (!d_seq || !d_seq_off || !d_hdr || !d_hdr_off) return fail(MH_OK);  // sizes only
  if (seq_cap < total_seq || hdr_cap < total_hdr || rec_cap < records)
    return fail(set_error(h, MH_ERR_OVERFLOW, "mh_fasta_parse: output too small"));
  k_fa_copy<<<fa_grid(n_lines * 64, 256, 1 << 16), 256, 0, st>>>(
      d_text, d_nl, n_nl, n, n_lines, d_kind, seq_pos, hdr_pos, rec, d_seq, d_hdr, d_seq_off,
      d_hdr_off);
  const uint64_t ends[2] = {total_seq, total_hdr};
  FA_HIP(hipMemcpyAsync(d_seq_off + records, &ends[0], 8, hipMemcpyHostToDevice, st));
  FA_HIP(hipMemcpyAsync(d_hdr_off + records, &ends[1], 8, hipMemcpyHostToDevice, st));
#undef FA_HIP
  cleanup();
  return MH_OK;
}
