// MatrixGPUHIP.cpp - the reference's GPU engine entry points on the MI355X
// path.  A MirAge build compiles this file in place of MatrixGPUCU.cu and
// links libmirage_hip.so; Analyser calls it exactly where it calls the CUDA
// engine (Analyser::TGetHighestScoreUsingGPU, MirAge/Analyser.cpp:785).
//
// Entry points (declared by the reference):
//   CUDAGetHighestScore        MirAge/Analyser.h:228 (MatrixGPUCU.cu:571)
//   InitialiseCuda             MirAge/Analyser.h:231 (MatrixGPUCU.cu:481)
//   Get_LOCAL_DATA_rowhit_SIZE MirAge/Analyser.h:232 (MatrixGPUCU.cu:39)
//   CUDAGetMaxBlockSize        MirAge/Hardware.h:16  (MatrixGPUCU.cu:549)
// and MirageHipLoadReferences, which the host calls once after reading the
// reference FASTA (Sequences::referenceSequences) to build the device index.
//
// Errors arrive as std::runtime_error (mirage_hip::Error), which
// TGetHighestScoreUsingGPU catches to shrink the batch and retry
// (Analyser.cpp:841-939).
#include <memory>
#include <stdexcept>
#include <vector>

#include "Analyser.h"
#include "MatrixGPUHIP.h"
#include "matrix_hip.hpp"

namespace {
std::vector<std::unique_ptr<mirage_hip::Device>> g_dev;

mirage_hip::Device &device(int id) {
  if (id < 0 || (size_t)id >= g_dev.size())
    throw std::runtime_error("MatrixGPUHIP: device not initialised");
  return *g_dev[(size_t)id];
}
}  // namespace

void InitialiseCuda(int numDevices) {
  g_dev.clear();
  for (int i = 0; i < numDevices; i++) g_dev.emplace_back(new mirage_hip::Device(i));
}

size_t Get_LOCAL_DATA_rowhit_SIZE() { return device(0).MaxReadLength(); }

size_t CUDAGetMaxBlockSize(int cudaDeviceID) { return device(cudaDeviceID).MaxBlockSize(); }

void MirageHipLoadReferences(const vector<vector<char>> &refs, int32_t considerThreshold) {
  for (auto &d : g_dev) d->LoadReferences(refs, considerThreshold);
}

// Scores the slice [queriesToAnalyse_Offset, +queriesToAnalyse_SizePerAnalysis)
// of seqQuery.  queryReferenceCombinations, ...CombiPositions and
// ..._LastPositionInRange point at the slice's first query and are indexed
// per query (MatrixGPUCU.cu:659-681); candidates outside the reference
// window [referencesToAnalyse_Offset, Offset + SizePerAnalysis] are left out
// as the CUDA engine leaves them out (MatrixGPUCU.cu:703-704).  The result
// holds one MatrixResult per query with at least one scored candidate
// (MatrixGPUCU.cu:1712, 1817), in ascending query order: QueryID is the
// query's index within the slice, as the CUDA engine reports it (the block
// index, MatrixGPUCU.cu:453); Scores is the (relative score, reference) of
// every candidate whose score is above -1, in candidate order, and
// HighestScore their maximum.  The CUDA engine lists a query's scores in
// the order its threads finished; Analyser sorts the results by QueryID
// (Analyser.cpp:595) and the log picks by score, so only the order inside
// Scores differs.  The position lists and head/tail sequences that the
// reference's older kernels read are not used.
vector<MatrixResult> CUDAGetHighestScore(
    vector<vector<char> *> seqQuery, Sequences *sequences, bool debug, int cudaDeviceID,
    size_t ConsiderThreshold, vector<SEQ_ID_t> *queryReferenceCombinations,
    vector<vector<SEQ_POS_t>> *queryReferenceCombinationsRefPositions,
    vector<vector<SEQ_POS_t>> *queryReferenceCombinationsQueryPositions,
    vector<vector<Enum::Position>> *queryReferenceCombinationsCombiPositions,
    vector<vector<Enum::Position>> *queryReferenceCombinationsCombiPositions_LastPositionInRange,
    int threadSize, Display *display, MultiThreading *multiThreading,
    SEQ_AMOUNT_t referencesToAnalyse_Offset, SEQ_AMOUNT_t referencesToAnalyse_SizePerAnalysis,
    SEQ_AMOUNT_t queriesToAnalyse_Offset, SEQ_AMOUNT_t queriesToAnalyse_SizePerAnalysis,
    vector<vector<Enum::HeadTailSequence>> *matchedReferencesPerQuery_HeadTailSequences) {
  (void)sequences;
  (void)debug;
  (void)ConsiderThreshold;
  (void)queryReferenceCombinationsRefPositions;
  (void)queryReferenceCombinationsQueryPositions;
  (void)threadSize;
  (void)display;
  (void)multiThreading;
  (void)matchedReferencesPerQuery_HeadTailSequences;
  mirage_hip::Device &dev = device(cudaDeviceID);
  const size_t nq = queriesToAnalyse_SizePerAnalysis;
  const uint64_t ref_lo = referencesToAnalyse_Offset;
  const uint64_t ref_hi = ref_lo + referencesToAnalyse_SizePerAnalysis;  // inclusive, as :703
  std::vector<std::vector<char> *> queries(nq);
  std::vector<std::vector<mirage_hip::SEQ_ID_t>> cand(nq);
  std::vector<std::vector<std::vector<mirage_hip::Position>>> start(nq), last(nq);
  for (size_t i = 0; i < nq; i++) {
    queries[i] = seqQuery[queriesToAnalyse_Offset + i];
    const vector<SEQ_ID_t> &qrc = queryReferenceCombinations[i];
    const vector<vector<Enum::Position>> &cp = queryReferenceCombinationsCombiPositions[i];
    const vector<vector<Enum::Position>> &cpl =
        queryReferenceCombinationsCombiPositions_LastPositionInRange[i];
    for (size_t c = 0; c < qrc.size(); c++) {
      if (!((uint64_t)qrc[c] >= ref_lo && (uint64_t)qrc[c] <= ref_hi)) continue;
      cand[i].push_back((mirage_hip::SEQ_ID_t)qrc[c]);
      start[i].emplace_back();
      last[i].emplace_back();
      for (size_t s = 0; s < cp[c].size(); s++) {
        start[i].back().push_back(
            {(mirage_hip::SEQ_POS_t)cp[c][s].Reference, (mirage_hip::SEQ_POS_t)cp[c][s].Query});
        last[i].back().push_back({(mirage_hip::SEQ_POS_t)cpl[c][s].Reference,
                                  (mirage_hip::SEQ_POS_t)cpl[c][s].Query});
      }
    }
  }
  const std::vector<mirage_hip::MatrixResult> res = dev.GetHighestScore(queries, cand, start, last);
  vector<MatrixResult> out;
  out.reserve(nq);
  for (size_t i = 0; i < nq; i++) {
    if (res[i].Scores.empty()) continue;
    vector<MatrixResult::ResultsScores> scores;
    scores.reserve(res[i].Scores.size());
    for (const auto &x : res[i].Scores) scores.emplace_back(x.score, (SEQ_ID_t)x.referenceID);
    out.emplace_back((SEQ_ID_t)i, res[i].HighestScore, scores);
  }
  return out;
}
