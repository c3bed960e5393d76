// C++ call surface over the C ABI (see matrix_hip.hpp).
#include "matrix_hip.hpp"

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstring>

namespace mirage_hip {

namespace {

void check(mh_handle *h, mh_status st, const char *what) {
  if (st != MH_OK)
    throw Error(st, std::string(what) + ": " + (h ? mh_last_error(h) : "no handle"));
}

template <class T>
struct DevBuf {
  T *p = nullptr;
  explicit DevBuf(size_t n) {
    if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess)
      throw Error(MH_ERR_NOMEM, "device allocation");
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  void up(const T *src, size_t n) {
    if (n && hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
      throw Error(MH_ERR_HIP, "upload");
  }
  void down(T *dst, size_t n) const {
    if (n && hipMemcpy(dst, p, n * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess)
      throw Error(MH_ERR_HIP, "download");
  }
};

}  // namespace

float MatrixResult::GetRelativeScore(float querySize, float referenceSize, float score) {
  const float lenMax = querySize < referenceSize ? querySize : referenceSize;
  const float scoreMax =
      (float)((((double)lenMax * ((double)lenMax + 1.0)) / 2.0) / (double)lenMax + 1);
  score += 1;
  return logf(score) / logf(scoreMax);
}

Device::Device(int device, uint32_t max_read_len) {
  check(nullptr, mh_open(device, max_read_len, &h_), "mh_open");
}

Device::~Device() {
  if (ix_) mh_index_free(ix_);
  if (h_) mh_close(h_);
}

void Device::LoadReferences(const std::vector<std::vector<char>> &refs, int32_t T) {
  std::vector<uint64_t> off(refs.size() + 1, 0);
  for (size_t i = 0; i < refs.size(); i++) off[i + 1] = off[i] + refs[i].size();
  std::vector<uint8_t> bytes(off.back());
  for (size_t i = 0; i < refs.size(); i++)
    if (!refs[i].empty()) std::memcpy(bytes.data() + off[i], refs[i].data(), refs[i].size());
  if (ix_) mh_index_free(ix_);
  ix_ = nullptr;
  refLens_.clear();
  check(h_, mh_index_build(h_, bytes.data(), off.data(), (uint32_t)refs.size(), T,
                           MH_INDEX_LAST_BUCKET_EMPTY, &ix_),
        "mh_index_build");
  for (const auto &r : refs) refLens_.push_back(r.size());
}

size_t Device::MaxReadLength() const { return mh_max_read_len(h_); }

size_t Device::MaxBlockSize() const { return mh_max_batch_reads(h_); }

std::vector<MatrixResult> Device::GetHighestScore(
    const std::vector<std::vector<char> *> &seqQuery,
    const std::vector<std::vector<SEQ_ID_t>> &qrc,
    const std::vector<std::vector<std::vector<Position>>> &cp,
    const std::vector<std::vector<std::vector<Position>>> &cpl, bool ultraFast) {
  if (!ix_) throw Error(MH_ERR_STATE, "no reference index loaded");
  const size_t nq = seqQuery.size();
  {
    std::vector<size_t> qlen(nq);
    for (size_t i = 0; i < nq; i++) {
      if (!seqQuery[i])
        throw Error(MH_ERR_ARG, "GetHighestScore: query " + std::to_string(i) + " is NULL");
      qlen[i] = seqQuery[i]->size();
    }
    const std::string bad = CheckScoringInputs(qlen, refLens_, MaxReadLength(), qrc, cp, cpl);
    if (!bad.empty()) throw Error(MH_ERR_ARG, "GetHighestScore: " + bad);
  }
  std::vector<uint64_t> qoff(nq + 1, 0);
  for (size_t i = 0; i < nq; i++) qoff[i + 1] = qoff[i] + seqQuery[i]->size();
  std::vector<uint8_t> qbytes(qoff.back());
  for (size_t i = 0; i < nq; i++)
    if (!seqQuery[i]->empty())
      std::memcpy(qbytes.data() + qoff[i], seqQuery[i]->data(), seqQuery[i]->size());
  // Flatten (query, candidate) pairs and their seeds.
  std::vector<uint32_t> pq, pr, seeds;
  std::vector<uint64_t> soff(1, 0);
  for (size_t q = 0; q < nq; q++)
    for (size_t c = 0; c < qrc[q].size(); c++) {
      pq.push_back((uint32_t)q);
      pr.push_back(qrc[q][c]);
      const auto &st = cp[q][c];
      const auto &la = cpl[q][c];
      for (size_t i = 0; i < st.size(); i++) {
        seeds.push_back(st[i].Reference);
        seeds.push_back(st[i].Query);
        seeds.push_back(la[i].Reference);
        seeds.push_back(la[i].Query);
      }
      soff.push_back(seeds.size() / 4);
    }
  const size_t np = pq.size();
  DevBuf<uint8_t> d_q(qbytes.size());
  DevBuf<uint64_t> d_qoff(nq + 1), d_soff(np + 1);
  DevBuf<uint32_t> d_pq(np), d_pr(np), d_seeds(seeds.size());
  DevBuf<float> d_score(np), d_rel(np);
  d_q.up(qbytes.data(), qbytes.size());
  d_qoff.up(qoff.data(), nq + 1);
  d_soff.up(soff.data(), np + 1);
  d_pq.up(pq.data(), np);
  d_pr.up(pr.data(), np);
  d_seeds.up(seeds.data(), seeds.size());
  check(h_, mh_score_seeds(h_, ix_, d_q.p, d_qoff.p, d_pq.p, d_pr.p, d_soff.p, d_seeds.p,
                           (uint32_t)np, ultraFast ? MH_CLASSIFY_ULTRAFAST : 0u, d_score.p,
                           d_rel.p, nullptr, nullptr, nullptr),
        "mh_score_seeds");
  if (hipStreamSynchronize((hipStream_t)mh_stream(h_)) != hipSuccess)
    throw Error(MH_ERR_HIP, "mh_score_seeds");
  std::vector<float> score(np), rel(np);
  d_score.down(score.data(), np);
  d_rel.down(rel.data(), np);
  std::vector<MatrixResult> out(nq);
  size_t p = 0;
  for (size_t q = 0; q < nq; q++) {
    out[q].QueryID = q;
    out[q].HighestScore = -1;
    for (size_t c = 0; c < qrc[q].size(); c++, p++) {
      if (!(score[p] > -1)) continue;
      if (rel[p] > out[q].HighestScore) out[q].HighestScore = rel[p];
      out[q].Scores.push_back({rel[p], pr[p]});
    }
  }
  return out;
}

std::vector<MatrixResult> Device::Classify(const std::vector<std::vector<char>> &reads,
                                           bool ultraFast) {
  if (!ix_) throw Error(MH_ERR_STATE, "no reference index loaded");
  const size_t n = reads.size();
  std::vector<uint64_t> off(n + 1, 0);
  for (size_t i = 0; i < n; i++) off[i + 1] = off[i] + reads[i].size();
  std::vector<uint8_t> bytes(off.back());
  for (size_t i = 0; i < n; i++)
    if (!reads[i].empty()) std::memcpy(bytes.data() + off[i], reads[i].data(), reads[i].size());
  std::vector<mh_read_result> res(n);
  uint64_t cap = 8 * n + 64, need = 0, cneed = 0;
  std::vector<uint32_t> top;
  for (int attempt = 0; attempt < 4; attempt++) {
    top.assign(cap, 0);
    const mh_status st = mh_classify_host(
        h_, ix_, bytes.data(), off.data(), (uint32_t)n,
        MH_CLASSIFY_BATCHED | (ultraFast ? MH_CLASSIFY_ULTRAFAST : 0u), res.data(), top.data(),
        cap, nullptr, 0, &need, &cneed);
    if (st == MH_ERR_OVERFLOW) {
      cap = std::max<uint64_t>(cap, need + 64);
      continue;
    }
    check(h_, st, "mh_classify_host");
    break;
  }
  std::vector<MatrixResult> out(n);
  for (size_t i = 0; i < n; i++) {
    out[i].QueryID = i;
    out[i].HighestScore = res[i].highest;
    for (uint32_t t = 0; t < res[i].n_top; t++)
      out[i].Scores.push_back({res[i].highest, top[res[i].top_off + t]});
  }
  return out;
}

}  // namespace mirage_hip
