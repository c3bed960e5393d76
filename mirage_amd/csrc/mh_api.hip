// C ABI of libmirage_hip: handles, index lifetime, classification entry
// points.  No exceptions cross this boundary; every call returns mh_status.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "mh_internal.hpp"

namespace mh {

// Error text of calls made without a handle, per calling thread.
static thread_local std::string t_last_error;

mh_status set_error(mh_handle *h, mh_status st, const std::string &msg) {
  if (h) h->last_error = msg;
  else t_last_error = msg;
  return st;
}

mh_status hip_error(mh_handle *h, hipError_t e, const char *what) {
  const mh_status st = (e == hipErrorOutOfMemory) ? MH_ERR_NOMEM : MH_ERR_HIP;
  return set_error(h, st, std::string(what) + ": " + hipGetErrorString(e));
}

mh_status exec_init(mh_handle *h, Exec &x) {
  MH_HIP(h, hipStreamCreateWithFlags(&x.stream, hipStreamNonBlocking));
  MH_HIP(h, hipMalloc(&x.d_ctr, sizeof(DevCounters)));
  MH_HIP(h, hipHostMalloc(&x.h_ctr, sizeof(DevCounters), hipHostMallocDefault));
  MH_HIP(h, hipMemsetAsync(x.d_ctr, 0, sizeof(DevCounters), x.stream));
  std::memset(x.h_ctr, 0, sizeof(DevCounters));
  for (auto &e : x.ev) MH_HIP(h, hipEventCreate(&e));
  return MH_OK;
}

void exec_free(Exec &x) {
  if (x.stream) hipStreamSynchronize(x.stream);
  if (x.graph) hipGraphExecDestroy(x.graph);
  if (x.d_scratch) hipFree(x.d_scratch);
  if (x.d_ctr) hipFree(x.d_ctr);
  if (x.h_ctr) hipHostFree(x.h_ctr);
  for (auto &e : x.ev)
    if (e) hipEventDestroy(e);
  if (x.stream) hipStreamDestroy(x.stream);
  x = Exec();
}

mh_status ensure_scratch(mh_handle *h, Exec &x, size_t bytes) {
  if (x.scratch_bytes >= bytes) return MH_OK;
  MH_HIP(h, hipStreamSynchronize(x.stream));
  // A captured call holds addresses inside the old arena: capture again.
  if (x.graph) {
    hipGraphExecDestroy(x.graph);
    x.graph = nullptr;
  }
  const size_t want = std::max(bytes, x.scratch_bytes * 5 / 4);
  if (x.d_scratch) hipFree(x.d_scratch);
  x.d_scratch = nullptr;
  x.scratch_bytes = 0;
  MH_HIP(h, hipMalloc(&x.d_scratch, want));
  x.scratch_bytes = want;
  return MH_OK;
}

namespace {

uint64_t tiles_version(const mh_handle *h) { return h->tiles_version; }

// scoreMax of MatrixResult::GetRelativeScore (MatrixResult.cpp:23-24).
float score_max(uint32_t L) {
  const float len_max = (float)L;
  return (float)((((double)len_max * ((double)len_max + 1.0)) / 2.0) / (double)len_max + 1);
}

// The relative score divides logf(Score + 1) by logf(scoreMax).  Both logf
// values come from tables filled here with the host's libm, so the device
// result is the same float the reference computes.  Score <= (m + 1) / 2 for
// a read of length m, so Score + 1 stays below the next power of two of
// m / 2 + 8.
mh_status build_luts(mh_handle *h) {
  const uint32_t L = h->max_read_len;
  const double xmax = (double)L / 2.0 + 8.0;
  int e2 = 1;
  while (std::ldexp(1.0, e2) < xmax) e2++;
  const float top = std::ldexp(1.0f, e2);
  uint32_t top_bits;
  std::memcpy(&top_bits, &top, 4);
  const uint32_t n_x = top_bits - 0x3f800000u;
  std::vector<float> xs(n_x), sm(L + 1, 0.0f);
  const unsigned th = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < th; t++)
    pool.emplace_back([&, t] {
      const uint64_t lo = (uint64_t)n_x * t / th, hi = (uint64_t)n_x * (t + 1) / th;
      for (uint64_t i = lo; i < hi; i++) {
        const uint32_t b = 0x3f800000u + (uint32_t)i;
        float x;
        std::memcpy(&x, &b, 4);
        xs[i] = logf(x);
      }
    });
  for (auto &p : pool) p.join();
  for (uint32_t l = 1; l <= L; l++) sm[l] = logf(score_max(l));
  MH_HIP(h, hipMalloc(&h->d_x_log, (size_t)n_x * 4));
  MH_HIP(h, hipMalloc(&h->d_smax_log, (size_t)(L + 1) * 4));
  MH_HIP(h, hipMemcpy(h->d_x_log, xs.data(), (size_t)n_x * 4, hipMemcpyHostToDevice));
  MH_HIP(h, hipMemcpy(h->d_smax_log, sm.data(), (size_t)(L + 1) * 4, hipMemcpyHostToDevice));
  h->n_x = n_x;
  h->n_smax = L + 1;
  return MH_OK;
}

bool valid_tiles(const mh_tile *t, uint32_t n, uint32_t max_len) {
  if (n == 0 || n > 8) return false;
  for (uint32_t i = 0; i < n; i++) {
    const uint32_t b = t[i].block;
    if (b != 8 && b != 16 && b != 32 && b != 64 && b != 128 && b != 256 && b != 512 &&
        b != 1024)
      return false;
    if (t[i].entries_cap < 32 || (t[i].entries_cap & (t[i].entries_cap - 1))) return false;
    if (i && t[i].max_len <= t[i - 1].max_len) return false;
  }
  return t[n - 1].max_len >= max_len;
}

// Scratch of mh_classify_noalloc: a fixed device array instead of the
// handle's arena, for callers that cannot allocate.  Each device the library
// runs on holds its own copy until the process ends.  In host code the
// array's name is not a device address: kernels and copies use the address
// hipGetSymbolAddress gives for the current device (g_noalloc_scratch).
static __device__ __align__(256) unsigned char scratch[MH_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
unsigned char *g_noalloc_scratch[kMaxDevices] = {};

// Looks up the device address of `scratch` on h's device (the current one),
// once per device.
mh_status noalloc_init(mh_handle *h) {
  if (h->device < 0 || h->device >= kMaxDevices)
    return set_error(h, MH_ERR_ARG, "device number above the no-allocation table");
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (g_noalloc_scratch[h->device]) return MH_OK;
  void *ptr = nullptr;
  MH_HIP(h, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr || ((uintptr_t)ptr & 255))
    return set_error(h, MH_ERR_HIP, "no-allocation scratch: bad symbol address");
  g_noalloc_scratch[h->device] = (unsigned char *)ptr;
  return MH_OK;
}

}  // namespace
}  // namespace mh

using namespace mh;

extern "C" {

mh_status mh_open(int32_t device, uint32_t max_read_len, mh_handle **out) {
  if (!out) return MH_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return MH_ERR_ARG;
  mh_handle *h = new mh_handle();
  h->device = device;
  h->max_read_len = max_read_len ? max_read_len : 32768;
  if (h->max_read_len > (1u << 24)) {
    delete h;
    return MH_ERR_ARG;
  }
  mh_status st = MH_OK;
  if (hipSetDevice(device) != hipSuccess) st = MH_ERR_HIP;
  if (!st) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device) == hipSuccess) {
      h->cu_count = (uint32_t)p.multiProcessorCount;
      h->lds_per_block = p.sharedMemPerBlock;
    }
    st = exec_init(h, h->ex);
    h->stream = h->ex.stream;
  }
  if (!st && ensure_comp_table(device) != hipSuccess) st = MH_ERR_HIP;
  if (!st) st = build_luts(h);
  if (!st) st = noalloc_init(h);
  if (st) {
    mh_close(h);
    return st;
  }
  *out = h;
  return MH_OK;
}

mh_status mh_close(mh_handle *h) {
  if (!h) return MH_ERR_ARG;
  hipSetDevice(h->device);
  exec_free(h->ex);
  if (h->d_x_log) hipFree(h->d_x_log);
  if (h->d_smax_log) hipFree(h->d_smax_log);
  delete h;
  return MH_OK;
}

const char *mh_last_error(const mh_handle *h) {
  return h ? h->last_error.c_str() : t_last_error.c_str();
}

void *mh_stream(mh_handle *h) { return h ? (void *)h->stream : nullptr; }

uint32_t mh_max_read_len(const mh_handle *h) { return h ? h->max_read_len : 0; }

uint64_t mh_max_batch_reads(const mh_handle *h) {
  // A batch of this size keeps every CU busy for several waves of queries
  // while its scratch stays well under 1 GiB.
  return h ? (uint64_t)h->cu_count * 4096 : 0;
}

mh_status mh_index_build(mh_handle *h, const uint8_t *h_ref_bytes, const uint64_t *h_ref_off,
                         uint32_t n_refs, int32_t threshold, uint32_t flags, mh_index **out) {
  if (!h) return MH_ERR_ARG;
  std::lock_guard<std::mutex> lk(h->mu);
  hipSetDevice(h->device);
  return build_index(h, h_ref_bytes, h_ref_off, n_refs, threshold, flags, out);
}

mh_status mh_index_free(mh_index *idx) {
  if (!idx) return MH_ERR_ARG;
  hipSetDevice(idx->device);
  if (idx->d_dir) hipFree(idx->d_dir);
  if (idx->d_ent) hipFree(idx->d_ent);
  if (idx->d_bytes) hipFree(idx->d_bytes);
  if (idx->d_ref_off) hipFree(idx->d_ref_off);
  delete idx;
  return MH_OK;
}

mh_status mh_index_get_info(const mh_index *idx, mh_index_info *out) {
  if (!idx || !out) return MH_ERR_ARG;
  out->n_entries = idx->n_entries;
  out->n_buckets = idx->n_buckets;
  out->device_bytes = idx->total_bytes;
  out->n_refs = idx->n_refs;
  out->threshold = idx->threshold;
  out->empty_hash = idx->empty_hash;
  out->pad = 0;
  return MH_OK;
}

mh_status mh_index_bucket(mh_handle *h, const mh_index *idx, uint32_t hash, uint32_t *h_refs,
                          uint32_t *h_pos, uint64_t cap, int64_t *count) {
  if (!h || !idx || !count || hash > kHashMask) return MH_ERR_ARG;
  std::lock_guard<std::mutex> lk(h->mu);
  hipSetDevice(h->device);
  uint64_t b[2];
  MH_HIP(h, hipMemcpy(b, idx->d_dir + hash, sizeof(b), hipMemcpyDeviceToHost));
  if (b[1] == b[0]) {
    *count = (hash == idx->empty_hash) ? 0 : -1;
    return MH_OK;
  }
  const uint64_t stored = b[1] - b[0];
  const uint64_t n = stored + (hash == 0 ? idx->phantom : 0);  // identical (0, 0) copies
  std::vector<uint2> e(n, make_uint2(0, 0));
  MH_HIP(h, hipMemcpy(e.data(), idx->d_ent + b[0], stored * sizeof(uint2),
                      hipMemcpyDeviceToHost));
  std::sort(e.begin(), e.end(), [](const uint2 &a, const uint2 &c) {
    return a.x < c.x || (a.x == c.x && a.y < c.y);
  });
  for (uint64_t i = 0; i < n && i < cap; i++) {
    if (h_refs) h_refs[i] = e[i].x;
    if (h_pos) h_pos[i] = e[i].y;
  }
  *count = (int64_t)n;
  return MH_OK;
}

static mh_status classify_on(mh_handle *h, Exec &x, const mh_index *idx, const uint8_t *d_reads,
                             const uint64_t *d_read_off, uint32_t n_reads, uint32_t flags,
                             mh_read_result *d_out, uint32_t *d_top_pool, uint64_t top_cap,
                             mh_cand *d_cands, uint64_t cand_cap, bool wait,
                             uint64_t *top_needed, uint64_t *cand_needed,
                             const FixedScratch *fixed = nullptr) {
  // A strand's id (2 * read + strand) and its overflow flag share 32 bits.
  if (n_reads >= MH_MAX_CALL_READS)
    return set_error(h, MH_ERR_ARG,
                     "classify: " + std::to_string(n_reads) + " reads, at most " +
                         std::to_string(MH_MAX_CALL_READS - 1) + " per call");
  ClassifyArgs a;
  a.ix = idx->view();
  a.luts = h->luts();
  a.rd = Reads{d_reads, d_read_off, n_reads};
  a.flags = flags;
  std::memset(&a.out, 0, sizeof(a.out));
  a.out.top = d_top_pool;
  a.out.top_cap = d_top_pool ? top_cap : 0;
  a.out.cands = d_cands;
  a.out.cand_cap = d_cands ? cand_cap : 0;
  a.final_out = d_out;
  mh_status st;
  if ((flags & MH_CLASSIFY_GRAPH) && !fixed) {
    // Replay the captured batched call when every argument is the same;
    // otherwise capture it again.  The scratch arena is sized before the
    // capture (no allocation or synchronisation may happen inside it).
    a.flags = (flags | MH_CLASSIFY_BATCHED) & ~(MH_CLASSIFY_TIMING | MH_CLASSIFY_GRAPH);
    // The index enters by its generation: a new index may reuse a freed
    // one's address, and the captured kernels hold the old one's buffers.
    const uint64_t key[12] = {idx->generation, (uint64_t)d_reads, (uint64_t)d_read_off, n_reads,
                              a.flags, (uint64_t)d_out, (uint64_t)d_top_pool, top_cap,
                              (uint64_t)d_cands, cand_cap, tiles_version(h), x.tie_factor};
    if (!x.graph || std::memcmp(key, x.graph_key, sizeof(key)) != 0) {
      if (x.graph) {
        hipGraphExecDestroy(x.graph);
        x.graph = nullptr;
      }
      st = ensure_scratch(h, x, classify_scratch_bytes(h, a.ix, n_reads, x.tie_factor, a.flags));
      if (st) return st;
      hipGraph_t g = nullptr;
      MH_HIP(h, hipStreamBeginCapture(x.stream, hipStreamCaptureModeThreadLocal));
      x.capturing = true;
      st = launch_classify(h, x, a, true);
      x.capturing = false;
      hipError_t e = hipStreamEndCapture(x.stream, &g);
      if (st) {
        if (g) hipGraphDestroy(g);
        return st;
      }
      if (e != hipSuccess) return hip_error(h, e, "graph capture");
      e = hipGraphInstantiate(&x.graph, g, nullptr, nullptr, 0);
      hipGraphDestroy(g);
      if (e != hipSuccess) {
        x.graph = nullptr;
        return hip_error(h, e, "graph instantiate");
      }
      std::memcpy(x.graph_key, key, sizeof(key));
    }
    MH_HIP(h, hipGraphLaunch(x.graph, x.stream));
    st = MH_OK;
  } else {
    st = launch_classify(h, x, a, (flags & MH_CLASSIFY_BATCHED) != 0, fixed);
  }
  if (st) return st;
  if (!wait) return MH_OK;
  MH_HIP(h, hipMemcpyAsync(x.h_ctr, x.d_ctr, sizeof(DevCounters), hipMemcpyDeviceToHost,
                           x.stream));
  MH_HIP(h, hipStreamSynchronize(x.stream));
  if (top_needed) *top_needed = x.h_ctr->top_used;
  if (cand_needed) *cand_needed = x.h_ctr->cand_used;
  if (fixed && x.h_ctr->tie_used > fixed->tie_cap)
    return set_error(h, MH_ERR_OVERFLOW,
                     "tie sets larger than the fixed scratch holds; classify fewer reads");
  if (!fixed && x.h_ctr->tie_used > (unsigned long long)x.tie_factor * 2ull * n_reads + 256) {
    // Tie sets larger than the scratch pool: grow it for the next call.
    while ((unsigned long long)x.tie_factor * 2ull * n_reads + 256 < x.h_ctr->tie_used)
      x.tie_factor *= 2;
    return set_error(h, MH_ERR_OVERFLOW, "tie pool grown; call again");
  }
  if (x.h_ctr->top_used > a.out.top_cap || (d_cands && x.h_ctr->cand_used > cand_cap))
    return set_error(h, MH_ERR_OVERFLOW, "output pool too small");
  return MH_OK;
}

mh_status mh_classify(mh_handle *h, const mh_index *idx, const uint8_t *d_reads,
                      const uint64_t *d_read_off, uint32_t n_reads, uint32_t flags,
                      mh_read_result *d_out, uint32_t *d_top_pool, uint64_t top_cap,
                      mh_cand *d_cands, uint64_t cand_cap, uint64_t *top_needed,
                      uint64_t *cand_needed) {
  if (!h || !idx || (n_reads && (!d_reads || !d_read_off || !d_out))) return MH_ERR_ARG;
  if (idx->owner != h && idx->device != h->device) return set_error(h, MH_ERR_ARG, "index device");
  std::lock_guard<std::mutex> lk(h->mu);
  hipSetDevice(h->device);
  h->last_valid = false;
  mh_status st = classify_on(h, h->ex, idx, d_reads, d_read_off, n_reads, flags, d_out,
                             d_top_pool, top_cap, d_cands, cand_cap,
                             top_needed != nullptr || cand_needed != nullptr, top_needed,
                             cand_needed);
  h->last.queries = 2ull * n_reads;
  return st;
}

mh_status mh_classify_noalloc(mh_handle *h, const mh_index *idx, const uint8_t *d_reads,
                              const uint64_t *d_read_off, uint32_t n_reads, uint32_t flags,
                              mh_read_result *d_out, uint32_t *d_top_pool, uint64_t top_cap,
                              mh_cand *d_cands, uint64_t cand_cap, uint64_t *top_needed,
                              uint64_t *cand_needed) {
  if (!h || !idx || (n_reads && (!d_reads || !d_read_off || !d_out))) return MH_ERR_ARG;
  if (idx->owner != h && idx->device != h->device) return set_error(h, MH_ERR_ARG, "index device");
  if (flags & MH_CLASSIFY_GRAPH)
    return set_error(h, MH_ERR_ARG, "mh_classify_noalloc: MH_CLASSIFY_GRAPH is not supported");
  if (n_reads > MH_NOALLOC_MAX_READS)
    return set_error(h, MH_ERR_RANGE,
                     "mh_classify_noalloc: " + std::to_string(n_reads) + " reads, at most " +
                         std::to_string(MH_NOALLOC_MAX_READS) + " per call");
  std::lock_guard<std::mutex> lk(h->mu);
  hipSetDevice(h->device);
  h->last_valid = false;
  unsigned char *base = g_noalloc_scratch[h->device];
  if (!base) return set_error(h, MH_ERR_STATE, "mh_classify_noalloc: scratch not set up");
  FixedScratch fs;
  if (plan_fixed_scratch(h, idx->view(), n_reads, flags, MH_NOALLOC_SCRATCH_BYTES, fs))
    return set_error(h, MH_ERR_RANGE,
                     "mh_classify_noalloc: " + std::to_string(n_reads) +
                         " reads do not fit the fixed scratch of " +
                         std::to_string(MH_NOALLOC_SCRATCH_BYTES >> 20) +
                         " MiB at this maximum read length");
  fs.base = base;
  mh_status st = classify_on(h, h->ex, idx, d_reads, d_read_off, n_reads, flags, d_out,
                             d_top_pool, top_cap, d_cands, cand_cap,
                             top_needed != nullptr || cand_needed != nullptr, top_needed,
                             cand_needed, &fs);
  h->last.queries = 2ull * n_reads;
  return st;
}

mh_status mh_classify_host(mh_handle *h, const mh_index *idx, const uint8_t *h_reads,
                           const uint64_t *h_read_off, uint32_t n_reads, uint32_t flags,
                           mh_read_result *h_out, uint32_t *h_top_pool, uint64_t top_cap,
                           mh_cand *h_cands, uint64_t cand_cap, uint64_t *top_needed,
                           uint64_t *cand_needed) {
  if (!h || !idx || (n_reads && (!h_reads || !h_read_off || !h_out))) return MH_ERR_ARG;
  std::lock_guard<std::mutex> lk(h->mu);
  hipSetDevice(h->device);
  h->last_valid = false;
  const uint64_t b0 = n_reads ? h_read_off[0] : 0;
  const uint64_t nbytes = n_reads ? h_read_off[n_reads] - b0 : 0;
  std::vector<uint64_t> off(n_reads + 1);
  for (uint32_t i = 0; i <= n_reads; i++) off[i] = n_reads ? h_read_off[i] - b0 : 0;
  for (uint32_t i = 0; i < n_reads; i++)
    if (off[i + 1] - off[i] > h->max_read_len) {
      // Longer reads cannot be scored (the reference CUDA form caps them too,
      // Sequences.cpp:183-194); flagged per read by the kernels.
      break;
    }
  uint8_t *d_reads = nullptr;
  uint64_t *d_off = nullptr;
  mh_read_result *d_out = nullptr;
  uint32_t *d_top = nullptr;
  mh_cand *d_cands = nullptr;
  auto cleanup = [&]() {
    if (d_reads) hipFree(d_reads);
    if (d_off) hipFree(d_off);
    if (d_out) hipFree(d_out);
    if (d_top) hipFree(d_top);
    if (d_cands) hipFree(d_cands);
  };
  hipError_t e = hipMalloc(&d_reads, std::max<uint64_t>(nbytes, 1));
  if (e == hipSuccess) e = hipMalloc(&d_off, (n_reads + 1) * 8ull);
  if (e == hipSuccess) e = hipMalloc(&d_out, std::max<uint64_t>(n_reads, 1) * sizeof(mh_read_result));
  if (e == hipSuccess) e = hipMalloc(&d_top, std::max<uint64_t>(top_cap, 1) * 4);
  if (e == hipSuccess && h_cands) e = hipMalloc(&d_cands, std::max<uint64_t>(cand_cap, 1) * sizeof(mh_cand));
  if (e == hipSuccess && nbytes)
    e = hipMemcpyAsync(d_reads, h_reads + b0, nbytes, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d_off, off.data(), (n_reads + 1) * 8ull, hipMemcpyHostToDevice, h->stream);
  if (e != hipSuccess) {
    cleanup();
    return hip_error(h, e, "mh_classify_host staging");
  }
  uint64_t tn = 0, cn = 0;
  mh_status st = classify_on(h, h->ex, idx, d_reads, d_off, n_reads, flags, d_out, d_top, top_cap,
                             d_cands, cand_cap, true, &tn, &cn);
  if (top_needed) *top_needed = tn;
  if (cand_needed) *cand_needed = cn;
  if (st == MH_OK || st == MH_ERR_OVERFLOW) {
    hipError_t e2 = hipMemcpy(h_out, d_out, n_reads * sizeof(mh_read_result), hipMemcpyDeviceToHost);
    if (e2 == hipSuccess && h_top_pool && top_cap)
      e2 = hipMemcpy(h_top_pool, d_top, std::min(tn, top_cap) * 4, hipMemcpyDeviceToHost);
    if (e2 == hipSuccess && h_cands && cand_cap)
      e2 = hipMemcpy(h_cands, d_cands, std::min(cn, cand_cap) * sizeof(mh_cand),
                     hipMemcpyDeviceToHost);
    if (e2 != hipSuccess) st = hip_error(h, e2, "mh_classify_host readback");
  }
  cleanup();
  h->last.queries = 2ull * n_reads;
  return st;
}

mh_status mh_get_counters(mh_handle *h, mh_counters *out) {
  if (!h || !out) return MH_ERR_ARG;
  std::lock_guard<std::mutex> lk(h->mu);
  hipSetDevice(h->device);
  MH_HIP(h, hipMemcpyAsync(h->ex.h_ctr, h->ex.d_ctr, sizeof(DevCounters), hipMemcpyDeviceToHost,
                           h->ex.stream));
  MH_HIP(h, hipStreamSynchronize(h->ex.stream));
  counters_from(*h->ex.h_ctr, h->last.queries / 2, h->last);
  *out = h->last;
  return MH_OK;
}

mh_status mh_get_kernel_counters(mh_handle *h, mh_counters *out, uint32_t cap, uint32_t *n) {
  if (!h || !n || (cap && !out)) return MH_ERR_ARG;
  std::lock_guard<std::mutex> lk(h->mu);
  hipSetDevice(h->device);
  MH_HIP(h, hipMemcpyAsync(h->ex.h_ctr, h->ex.d_ctr, sizeof(DevCounters), hipMemcpyDeviceToHost,
                           h->ex.stream));
  MH_HIP(h, hipStreamSynchronize(h->ex.stream));
  *n = (uint32_t)kKernelRows;
  kernel_counters_from(*h->ex.h_ctr, out, std::min<uint32_t>(cap, (uint32_t)kKernelRows));
  return MH_OK;
}

mh_status mh_get_timing(mh_handle *h, mh_timing *out) {
  if (!h || !out) return MH_ERR_ARG;
  std::lock_guard<std::mutex> lk(h->mu);
  hipSetDevice(h->device);
  std::memset(out, 0, sizeof(*out));
  Exec &x = h->ex;
  if (!x.timed) return set_error(h, MH_ERR_STATE, "last call was not timed");
  MH_HIP(h, hipEventSynchronize(x.ev[11]));
  auto ms = [&](int a, int b) {
    float v = 0;
    hipEventElapsedTime(&v, x.ev[a], x.ev[b]);
    return v;
  };
  out->total_ms = ms(0, 11);
  out->n_buckets = x.timed_buckets;
  int prev = 9;
  for (uint32_t i = 0; i < x.timed_buckets; i++) {
    out->bucket_ms[x.bucket_order[i]] = ms(prev, 1 + (int)i);
    prev = 1 + (int)i;
  }
  out->big_ms = ms(prev, 10);
  // each tier from the end of the one before (the last length tile first)
  int at = prev;
  for (int t = 0; t < 4; t++) {
    const int ev = t < 3 ? 13 + t : 10;
    out->tier_ms[t] = ms(at, ev);
    at = ev;
  }
  if (x.timed_probe) out->probe_ms = ms(12, 9);
  out->merge_ms = ms(10, 11);
  return MH_OK;
}

mh_status mh_score_seeds(mh_handle *h, const mh_index *idx, const uint8_t *d_queries,
                         const uint64_t *d_query_off, const uint32_t *d_pair_query,
                         const uint32_t *d_pair_ref, const uint64_t *d_seed_off,
                         const uint32_t *d_seeds, uint32_t n_pairs, uint32_t flags,
                         float *d_score, float *d_rel, uint64_t *d_sum_r, uint64_t *d_sum_c,
                         uint32_t *d_hits) {
  if (!h || !idx || (n_pairs && (!d_queries || !d_query_off || !d_pair_query || !d_pair_ref ||
                                 !d_seed_off || !d_score)))
    return MH_ERR_ARG;
  std::lock_guard<std::mutex> lk(h->mu);
  hipSetDevice(h->device);
  return launch_score_seeds(h, idx->view(), h->luts(), d_queries, d_query_off, d_pair_query,
                            d_pair_ref, d_seed_off, d_seeds, n_pairs, flags, d_score, d_rel,
                            d_sum_r, d_sum_c, d_hits);
}

mh_status mh_set_tiles(mh_handle *h, const mh_tile *tiles, uint32_t n_tiles) {
  if (!h) return MH_ERR_ARG;
  if (!tiles && !n_tiles) {
    std::lock_guard<std::mutex> lk(h->mu);
    h->tiles.clear();
    h->tiles_version++;
    return MH_OK;
  }
  if (!tiles || !valid_tiles(tiles, n_tiles, h->max_read_len))
    return set_error(h, MH_ERR_ARG,
                     "mh_set_tiles: 1-8 tiles of ascending max_len, the last at least "
                     "max_read_len; block 8..1024 (a power of two); entries_cap a power of "
                     "two >= 32");
  std::lock_guard<std::mutex> lk(h->mu);
  h->tiles.assign(tiles, tiles + n_tiles);
  h->tiles_version++;
  return MH_OK;
}

mh_status mh_set_overflow_tiles(mh_handle *h, const mh_tile *tiers) {
  if (!h) return MH_ERR_ARG;
  if (!tiers) {
    std::lock_guard<std::mutex> lk(h->mu);
    h->tier_tiles.clear();
    h->tiles_version++;
    return MH_OK;
  }
  for (int i = 0; i < 2; i++) {
    const mh_tile &t = tiers[i];
    const uint32_t b = t.block;
    if ((b != 64 && b != 128 && b != 256 && b != 512 && b != 1024) || t.entries_cap < 32 ||
        (t.entries_cap & 15) || t.max_len < 64 || t.max_len > h->max_read_len)
      return set_error(h, MH_ERR_ARG,
                       "mh_set_overflow_tiles: two tiles; block 64..1024 (a power of two), "
                       "entries_cap a multiple of 16 >= 32, max_len 64..max_read_len");
  }
  std::lock_guard<std::mutex> lk(h->mu);
  h->tier_tiles.assign(tiers, tiers + 2);
  h->tiles_version++;
  return MH_OK;
}

mh_status mh_get_tiles(const mh_handle *h, const mh_index *idx, mh_tile *tiles, uint32_t cap,
                       uint32_t *n_tiles, mh_tile *tiers) {
  if (!h || !idx || !n_tiles) return MH_ERR_ARG;
  const IndexView v = idx->view();
  const std::vector<mh_tile> t = active_tiles(h, v);
  *n_tiles = (uint32_t)t.size();
  for (uint32_t i = 0; i < cap && i < t.size(); i++) tiles[i] = t[i];
  if (tiers) {
    const std::vector<mh_tile> tt = tier_tiles(h, v);
    tiers[0] = tt[0];
    tiers[1] = tt[1];
  }
  return MH_OK;
}

}  // extern "C"
