// C++ call surface of the HIP matrix path, shaped like the reference's
// engines so that Analyser can call it where it calls
//   MatrixCPU::CPUGetHighestScore   (MirAge/MatrixCPU.h:41)
//   CUDAGetHighestScore             (MirAge/Analyser.h:228, MatrixGPUCU.cu:571)
// and receive vector<MatrixResult> (MirAge/MatrixResult.h:18-62).  Errors are
// thrown as std::runtime_error, which Analyser::TGetHighestScoreUsingGPU
// catches to shrink the batch and retry (MirAge/Analyser.cpp:841-939).
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "mirage_hip.h"

namespace mirage_hip {

using SEQ_ID_t = uint32_t;   // MirAge/Enum.h:10-22 (32-bit build)
using SEQ_POS_t = uint32_t;

// Enum::Position (MirAge/Enum.h:64-76).
struct Position {
  SEQ_POS_t Reference;
  SEQ_POS_t Query;
};

// MatrixResult (MirAge/MatrixResult.h:18-62).
struct MatrixResult {
  struct ResultsScores {
    float score;
    SEQ_ID_t referenceID;
  };
  size_t QueryID = 0;
  float HighestScore = -1;
  std::vector<ResultsScores> Scores;
  // MatrixResult::GetRelativeScore (MatrixResult.cpp:21-29).
  static float GetRelativeScore(float querySize, float referenceSize, float score);
};

// Checks the scoring inputs of Device::GetHighestScore on the host before
// anything is uploaded: one candidate list and one seed list per query, as
// many last positions as start positions, every candidate a reference id
// below refLens.size(), and every seed that is scored (start.Query <=
// last.Query; MatrixCPU.cpp:57-58 skips the others) inside its query
// (last.Query <= query length) and its reference (start.Reference < length,
// last.Reference <= length).  Queries longer than maxQueryLen are refused
// too.  Returns "" when the inputs are well formed, else a message naming the
// query, candidate and seed.  Host only; it touches no device.
std::string CheckScoringInputs(const std::vector<size_t> &queryLens,
                               const std::vector<uint64_t> &refLens, size_t maxQueryLen,
                               const std::vector<std::vector<SEQ_ID_t>> &queryReferenceCombinations,
                               const std::vector<std::vector<std::vector<Position>>> &combiPositions,
                               const std::vector<std::vector<std::vector<Position>>> &combiPositionsLast);

class Error : public std::runtime_error {
 public:
  Error(mh_status st, const std::string &what) : std::runtime_error(what), status(st) {}
  mh_status status;
};

// One device with a resident reference index.
class Device {
 public:
  // InitialiseCuda (MatrixGPUCU.cu:481-488) for one device.
  Device(int device, uint32_t max_read_len = 0);
  ~Device();
  Device(const Device &) = delete;
  Device &operator=(const Device &) = delete;

  // Builds the index (Hashmap::BuildHashmapTable semantics) on the device.
  void LoadReferences(const std::vector<std::vector<char>> &refs, int32_t considerThreshold);

  // Get_LOCAL_DATA_rowhit_SIZE (MatrixGPUCU.cu:39-42): longest read accepted.
  size_t MaxReadLength() const;
  // CUDAGetMaxBlockSize (MatrixGPUCU.cu:549-556): reads per batch.
  size_t MaxBlockSize() const;

  // Scoring role of CUDAGetHighestScore / CPUGetHighestScore: queries (the
  // forward reads and their reverse complements, as the caller built them)
  // with their candidate references and seeds
  // (matchedReferencesPerQuery, _CombiPositions, _LastPositionInRange,
  // Analyser.h:94-98).  Returns one MatrixResult per query whose Scores hold
  // (relative score, reference) for every candidate with Score > -1, in
  // candidate order, and HighestScore the first strict maximum
  // (MatrixCPU::CreateResultsScores, MatrixCPU.cpp:323-370).
  std::vector<MatrixResult> GetHighestScore(
      const std::vector<std::vector<char> *> &seqQuery,
      const std::vector<std::vector<SEQ_ID_t>> &queryReferenceCombinations,
      const std::vector<std::vector<std::vector<Position>>> &combiPositions,
      const std::vector<std::vector<std::vector<Position>>> &combiPositionsLast,
      bool ultraFast = false);

  // The whole path on the device for forward reads: candidate generation,
  // scoring, best hit and strand merge (Analyser.cpp:592-611).  One
  // MatrixResult per read (QueryID = read index); Scores holds the top-tie
  // group, each with the HighestScore, ascending by reference.
  std::vector<MatrixResult> Classify(const std::vector<std::vector<char>> &reads,
                                     bool ultraFast = false);

  mh_handle *handle() { return h_; }
  mh_index *index() { return ix_; }

 private:
  mh_handle *h_ = nullptr;
  mh_index *ix_ = nullptr;
  std::vector<uint64_t> refLens_;  // lengths of the indexed references
};

}  // namespace mirage_hip
