// Serving loop: a resident index, requests admitted by HBM budget, and
// continuous batching over several execution contexts at once.
//
// A scheduler thread owns max_inflight slots.  Each slot is an execution
// context (its own stream, scratch arena and counters) with device and
// pinned host buffers for one coalesced batch.  Whenever a slot is free, the
// scheduler packs the oldest pending requests - as many as fit the slot's
// read and byte capacity and the remaining HBM budget - into one batch and
// hands it to a thread of the slot, which stages it into pinned memory,
// copies it up, launches the batched classification and copies the results
// back on the slot's stream, then waits for the stream, splits the results
// per request and wakes the waiters.  Requests that arrive while batches are running are admitted as
// soon as a slot and budget free up, so batches overlap on the device.
// Per-read results do not depend on batch composition: every query is
// processed independently and the tie pool is compacted per batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <vector>

#include "mh_internal.hpp"

using namespace mh;

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

struct Request {
  uint64_t id = 0;
  std::vector<uint8_t> reads;
  std::vector<uint64_t> off;  // n + 1, from 0
  uint32_t n = 0;
  double submit_s = 0, admit_s = 0, done_s = 0;
  int32_t status = MH_OK;
  bool done = false;
  std::vector<mh_read_result> out;
  std::vector<uint32_t> top;
  uint64_t bytes() const { return reads.size() + off.size() * 8 + (uint64_t)n * 64; }
};

struct Slot {
  Exec ex;
  uint8_t *d_reads = nullptr;
  uint64_t *d_off = nullptr;
  mh_read_result *d_out = nullptr;
  uint32_t *d_top = nullptr;
  uint8_t *h_reads = nullptr;  // pinned staging
  uint64_t *h_off = nullptr;
  mh_read_result *h_out = nullptr;
  uint32_t *h_top = nullptr;
  uint64_t cap_bytes = 0, cap_reads = 0, top_cap = 0;
  std::vector<Request *> batch;
  uint64_t batch_bytes = 0;
  uint32_t batch_reads = 0;
  bool busy = false;
  bool launched_ok = false;
  std::thread waiter;
};

}  // namespace

struct mh_server {
  mh_handle *h = nullptr;
  const mh_index *idx = nullptr;
  mh_server_config cfg{};
  std::mutex mu;
  std::condition_variable cv_sched, cv_done;
  std::deque<Request *> pending;
  std::map<uint64_t, std::unique_ptr<Request>> reqs;
  std::vector<std::unique_ptr<Slot>> slots;
  uint64_t next_id = 1;
  uint64_t inflight_bytes = 0;
  bool stop = false;
  std::thread sched;
  std::string error;
};

namespace {

void free_slot(Slot &s) {
  exec_free(s.ex);
  if (s.d_reads) hipFree(s.d_reads);
  if (s.d_off) hipFree(s.d_off);
  if (s.d_out) hipFree(s.d_out);
  if (s.d_top) hipFree(s.d_top);
  if (s.h_reads) hipHostFree(s.h_reads);
  if (s.h_off) hipHostFree(s.h_off);
  if (s.h_out) hipHostFree(s.h_out);
  if (s.h_top) hipHostFree(s.h_top);
}

mh_status init_slot(mh_handle *h, Slot &s, uint64_t cap_bytes, uint64_t cap_reads,
                    const IndexView &ix, uint32_t flags) {
  mh_status st = exec_init(h, s.ex);
  if (st) return st;
  s.cap_bytes = cap_bytes;
  s.cap_reads = cap_reads;
  s.top_cap = 8 * cap_reads + 1024;
  MH_HIP(h, hipMalloc(&s.d_reads, cap_bytes));
  MH_HIP(h, hipMalloc(&s.d_off, (cap_reads + 1) * 8));
  MH_HIP(h, hipMalloc(&s.d_out, cap_reads * sizeof(mh_read_result)));
  MH_HIP(h, hipMalloc(&s.d_top, s.top_cap * 4));
  MH_HIP(h, hipHostMalloc(&s.h_reads, cap_bytes, hipHostMallocDefault));
  MH_HIP(h, hipHostMalloc(&s.h_off, (cap_reads + 1) * 8, hipHostMallocDefault));
  MH_HIP(h, hipHostMalloc(&s.h_out, cap_reads * sizeof(mh_read_result), hipHostMallocDefault));
  MH_HIP(h, hipHostMalloc(&s.h_top, s.top_cap * 4, hipHostMallocDefault));
  return ensure_scratch(h, s.ex, classify_scratch_bytes(h, ix, (uint32_t)cap_reads, 4,
                                                        flags));
}

// Launch the slot's batch, from the slot's thread without the server lock:
// the slot is busy and its batch's requests are no longer pending.
mh_status launch_slot(mh_server *sv, Slot &s) {
  mh_handle *h = sv->h;
  uint64_t pos = 0, longest = 0;
  uint32_t r = 0;
  s.h_off[0] = 0;
  for (Request *q : s.batch) {
    std::memcpy(s.h_reads + pos, q->reads.data(), q->reads.size());
    for (uint32_t i = 0; i < q->n; i++) {
      s.h_off[r + i + 1] = pos + q->off[i + 1];
      longest = std::max<uint64_t>(longest, q->off[i + 1] - q->off[i]);
    }
    pos += q->reads.size();
    r += q->n;
  }
  hipStream_t st = s.ex.stream;
  MH_HIP(h, hipMemcpyAsync(s.d_reads, s.h_reads, std::max<uint64_t>(pos, 1),
                           hipMemcpyHostToDevice, st));
  MH_HIP(h, hipMemcpyAsync(s.d_off, s.h_off, (r + 1) * 8ull, hipMemcpyHostToDevice, st));
  ClassifyArgs a;
  a.ix = sv->idx->view();
  a.luts = h->luts();
  a.rd = Reads{s.d_reads, s.d_off, r};
  a.flags = sv->cfg.flags | MH_CLASSIFY_BATCHED;
  std::memset(&a.out, 0, sizeof(a.out));
  a.out.top = s.d_top;
  a.out.top_cap = s.top_cap;
  a.final_out = s.d_out;
  a.max_len_hint = longest;
  mh_status stt;
  {
    std::lock_guard<std::mutex> lk(h->mu);
    hipSetDevice(h->device);
    stt = launch_classify(h, s.ex, a, true);
  }
  if (stt) return stt;
  MH_HIP(h, hipMemcpyAsync(s.h_out, s.d_out, r * sizeof(mh_read_result), hipMemcpyDeviceToHost,
                           st));
  MH_HIP(h, hipMemcpyAsync(s.ex.h_ctr, s.ex.d_ctr, sizeof(DevCounters), hipMemcpyDeviceToHost,
                           st));
  s.batch_reads = r;
  return MH_OK;
}

// Waits for the slot's stream, hands results to the requests, frees the slot.
void complete_slot(mh_server *sv, Slot *s) {
  hipSetDevice(sv->h->device);
  hipError_t e = hipStreamSynchronize(s->ex.stream);
  // The top-tie ids: only as many as the batch produced.
  if (e == hipSuccess && s->launched_ok && s->ex.h_ctr->top_used <= s->top_cap &&
      s->ex.h_ctr->top_used)
    e = hipMemcpy(s->h_top, s->d_top, s->ex.h_ctr->top_used * 4, hipMemcpyDeviceToHost);
  const double t = now_s();
  // The batch's requests belong to this slot until they are marked done
  // under the lock, so the slot-private checks, the pool growth and the
  // results split below run without it: other slots keep completing and the
  // scheduler keeps admitting meanwhile.
  int32_t st = (e == hipSuccess && s->launched_ok) ? MH_OK : MH_ERR_HIP;
  bool overflow = false;
  if (st == MH_OK) {
    overflow = s->ex.h_ctr->top_used > s->top_cap ||
               s->ex.h_ctr->tie_used > (unsigned long long)s->ex.tie_factor * 2ull *
                                               s->batch_reads + 256;
    if (s->ex.h_ctr->tie_used > (unsigned long long)s->ex.tie_factor * 2ull * s->batch_reads + 256)
      while ((unsigned long long)s->ex.tie_factor * 2ull * s->batch_reads + 256 <
             s->ex.h_ctr->tie_used)
        s->ex.tie_factor *= 2;
  }
  if (overflow && s->ex.h_ctr->top_used > s->top_cap) {
    // Grow the top-tie pool to what the batch needed (the stream is idle).
    const uint64_t want = s->ex.h_ctr->top_used + s->ex.h_ctr->top_used / 4 + 1024;
    uint32_t *nd = nullptr, *nh = nullptr;
    if (hipMalloc(&nd, want * 4) == hipSuccess &&
        hipHostMalloc(&nh, want * 4, hipHostMallocDefault) == hipSuccess) {
      hipFree(s->d_top);
      hipHostFree(s->h_top);
      s->d_top = nd;
      s->h_top = nh;
      s->top_cap = want;
    } else {
      if (nd) hipFree(nd);
      st = MH_ERR_NOMEM;
      overflow = false;
    }
  }
  if (!overflow) {
    uint32_t r = 0;
    for (Request *q : s->batch) {
      q->status = st;
      q->out.assign(s->h_out + r, s->h_out + r + q->n);
      if (st == MH_OK) {
        uint64_t ntop = 0;
        for (auto &o : q->out) ntop += o.n_top;
        q->top.resize(ntop);
        uint64_t w = 0;
        for (auto &o : q->out) {
          std::memcpy(q->top.data() + w, s->h_top + o.top_off, o.n_top * 4ull);
          o.top_off = (uint32_t)w;
          w += o.n_top;
        }
      }
      q->done_s = t;
      r += q->n;
    }
  }
  std::unique_lock<std::mutex> lk(sv->mu);
  if (overflow) {
    // Rare: more tied references than the pools held.  Requeue the batch at
    // the front; the pools have grown to what it needs.
    for (auto it = s->batch.rbegin(); it != s->batch.rend(); ++it) sv->pending.push_front(*it);
  } else {
    for (Request *q : s->batch) q->done = true;
  }
  sv->inflight_bytes -= s->batch_bytes;
  s->batch.clear();
  s->batch_bytes = 0;
  s->busy = false;
  lk.unlock();
  sv->cv_done.notify_all();
  sv->cv_sched.notify_all();
}

void scheduler(mh_server *sv) {
  hipSetDevice(sv->h->device);
  std::unique_lock<std::mutex> lk(sv->mu);
  while (!sv->stop) {
    bool launched = false;
    for (auto &sp : sv->slots) {
      Slot &s = *sp;
      if (s.busy || sv->pending.empty()) continue;
      if (s.waiter.joinable()) s.waiter.join();
      // Admit the oldest requests that fit this slot and the HBM budget.
      uint64_t bytes = 0, reads = 0;
      while (!sv->pending.empty()) {
        Request *q = sv->pending.front();
        const uint64_t qb = q->bytes();
        const bool first = s.batch.empty();
        if (!first && (reads + q->n > s.cap_reads || bytes + q->reads.size() > s.cap_bytes))
          break;
        if (!first && sv->inflight_bytes + qb > sv->cfg.hbm_budget) break;
        if (first && sv->inflight_bytes > 0 && sv->inflight_bytes + qb > sv->cfg.hbm_budget)
          break;
        if (q->n > s.cap_reads || q->reads.size() > s.cap_bytes) {
          // Larger than any batch: fail it rather than block the queue.
          sv->pending.pop_front();
          q->status = MH_ERR_RANGE;
          q->done = true;
          q->done_s = now_s();
          sv->cv_done.notify_all();
          continue;
        }
        sv->pending.pop_front();
        q->admit_s = now_s();
        s.batch.push_back(q);
        bytes += q->reads.size();
        reads += q->n;
        sv->inflight_bytes += qb;
        s.batch_bytes += qb;
      }
      if (s.batch.empty()) continue;
      s.busy = true;
      // The batch belongs to the slot now: the slot's own thread stages it
      // (host copies into pinned memory), launches it and completes it, so
      // the batches of several slots are staged in parallel and the
      // scheduler goes on admitting meanwhile.
      Slot *sp_ = &s;
      s.waiter = std::thread([sv, sp_] {
        hipSetDevice(sv->h->device);
        sp_->launched_ok = launch_slot(sv, *sp_) == MH_OK;
        complete_slot(sv, sp_);
      });
      launched = true;
    }
    if (!launched) sv->cv_sched.wait_for(lk, std::chrono::milliseconds(50));
  }
}

}  // namespace

extern "C" {

mh_status mh_server_start(mh_handle *h, const mh_index *idx, const mh_server_config *cfg,
                          mh_server **out) {
  if (!h || !idx || !out) return MH_ERR_ARG;
  hipSetDevice(h->device);
  auto *sv = new mh_server();
  sv->h = h;
  sv->idx = idx;
  if (cfg) sv->cfg = *cfg;
  if (!sv->cfg.hbm_budget) sv->cfg.hbm_budget = 8ull << 30;
  if (!sv->cfg.max_inflight) sv->cfg.max_inflight = 3;
  if (!sv->cfg.max_batch_reads) sv->cfg.max_batch_reads = 262144;
  const uint64_t cap_reads = sv->cfg.max_batch_reads;
  const uint64_t cap_bytes =
      std::min<uint64_t>(sv->cfg.hbm_budget / sv->cfg.max_inflight, 512ull << 20);
  for (uint32_t i = 0; i < sv->cfg.max_inflight; i++) {
    sv->slots.emplace_back(new Slot());
    mh_status st = init_slot(h, *sv->slots.back(), cap_bytes, cap_reads, idx->view(),
                             sv->cfg.flags | MH_CLASSIFY_BATCHED);
    if (st) {
      for (auto &s : sv->slots) free_slot(*s);
      delete sv;
      return st;
    }
  }
  sv->sched = std::thread(scheduler, sv);
  *out = sv;
  return MH_OK;
}

mh_status mh_server_submit(mh_server *sv, const uint8_t *h_reads, const uint64_t *h_read_off,
                           uint32_t n_reads, uint64_t *request_id) {
  if (!sv || !request_id || (n_reads && (!h_reads || !h_read_off))) return MH_ERR_ARG;
  auto q = std::make_unique<Request>();
  q->n = n_reads;
  q->submit_s = now_s();
  const uint64_t b0 = n_reads ? h_read_off[0] : 0;
  const uint64_t nb = n_reads ? h_read_off[n_reads] - b0 : 0;
  q->reads.assign(h_reads + b0, h_reads + b0 + nb);
  q->off.resize(n_reads + 1);
  for (uint32_t i = 0; i <= n_reads; i++) q->off[i] = n_reads ? h_read_off[i] - b0 : 0;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->stop) return MH_ERR_STATE;
    q->id = sv->next_id++;
    *request_id = q->id;
    sv->pending.push_back(q.get());
    sv->reqs[q->id] = std::move(q);
  }
  sv->cv_sched.notify_all();
  return MH_OK;
}

mh_status mh_server_wait(mh_server *sv, uint64_t request_id, double timeout_s,
                         mh_read_result *h_out, uint32_t *h_top, uint64_t top_cap,
                         uint64_t *top_needed, mh_request_info *info) {
  if (!sv) return MH_ERR_ARG;
  std::unique_lock<std::mutex> lk(sv->mu);
  auto it = sv->reqs.find(request_id);
  if (it == sv->reqs.end()) return MH_ERR_ARG;
  Request *q = it->second.get();
  auto ready = [&] { return q->done; };
  if (timeout_s < 0) sv->cv_done.wait(lk, ready);
  else if (!sv->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready))
    return MH_ERR_STATE;
  if (top_needed) *top_needed = q->top.size();
  if (info) {
    info->request_id = q->id;
    info->n_reads = q->n;
    info->status = q->status;
    info->submit_s = q->submit_s;
    info->admit_s = q->admit_s;
    info->done_s = q->done_s;
  }
  if (q->top.size() > top_cap && h_top) return MH_ERR_OVERFLOW;
  const mh_status st = q->status;
  // The request leaves the table here; its results are copied out after the
  // lock is released.
  std::unique_ptr<Request> own = std::move(it->second);
  sv->reqs.erase(it);
  lk.unlock();
  if (h_out && !own->out.empty())
    std::memcpy(h_out, own->out.data(), own->out.size() * sizeof(mh_read_result));
  if (h_top && !own->top.empty()) std::memcpy(h_top, own->top.data(), own->top.size() * 4);
  return st;
}

mh_status mh_server_stop(mh_server *sv) {
  if (!sv) return MH_ERR_ARG;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stop = true;
  }
  sv->cv_sched.notify_all();
  if (sv->sched.joinable()) sv->sched.join();
  for (auto &s : sv->slots)
    if (s->waiter.joinable()) s->waiter.join();
  {
    // Requests never admitted finish with MH_ERR_STATE.
    std::lock_guard<std::mutex> lk(sv->mu);
    for (Request *q : sv->pending) {
      q->status = MH_ERR_STATE;
      q->done = true;
    }
    sv->pending.clear();
  }
  sv->cv_done.notify_all();
  hipSetDevice(sv->h->device);
  for (auto &s : sv->slots) free_slot(*s);
  delete sv;
  return MH_OK;
}

}  // extern "C"
