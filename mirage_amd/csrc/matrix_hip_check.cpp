// Host-side check of the scoring inputs of the C++ call surface (see
// matrix_hip.hpp).  No HIP here: tests compile this file on machines
// without a GPU.
#include <string>

#include "matrix_hip.hpp"

namespace mirage_hip {

std::string CheckScoringInputs(const std::vector<size_t> &queryLens,
                               const std::vector<uint64_t> &refLens, size_t maxQueryLen,
                               const std::vector<std::vector<SEQ_ID_t>> &qrc,
                               const std::vector<std::vector<std::vector<Position>>> &cp,
                               const std::vector<std::vector<std::vector<Position>>> &cpl) {
  const size_t nq = queryLens.size();
  if (qrc.size() != nq || cp.size() != nq || cpl.size() != nq)
    return "one candidate list and one seed list per query are needed (" + std::to_string(nq) +
           " queries, " + std::to_string(qrc.size()) + " candidate lists, " +
           std::to_string(cp.size()) + " start lists, " + std::to_string(cpl.size()) +
           " last lists)";
  for (size_t q = 0; q < nq; q++) {
    const std::string at = "query " + std::to_string(q);
    const uint64_t m = queryLens[q];
    if (m > maxQueryLen)
      return at + ": length " + std::to_string(m) + " above the maximum " +
             std::to_string(maxQueryLen);
    if (cp[q].size() != qrc[q].size() || cpl[q].size() != qrc[q].size())
      return at + ": " + std::to_string(qrc[q].size()) + " candidates but " +
             std::to_string(cp[q].size()) + " start and " + std::to_string(cpl[q].size()) +
             " last seed lists";
    for (size_t c = 0; c < qrc[q].size(); c++) {
      const std::string atc = at + ", candidate " + std::to_string(c);
      const SEQ_ID_t r = qrc[q][c];
      if (r >= refLens.size())
        return atc + ": reference id " + std::to_string(r) + " not below " +
               std::to_string(refLens.size());
      if (cpl[q][c].size() != cp[q][c].size())
        return atc + ": " + std::to_string(cp[q][c].size()) + " start but " +
               std::to_string(cpl[q][c].size()) + " last positions";
      const uint64_t rl = refLens[r];
      for (size_t i = 0; i < cp[q][c].size(); i++) {
        const Position &s = cp[q][c][i], &l = cpl[q][c][i];
        if (s.Query > l.Query) continue;  // skipped, as MatrixCPU.cpp:57-58 does
        const std::string ats = atc + ", seed " + std::to_string(i);
        if (l.Query > m)
          return ats + ": last.Query " + std::to_string(l.Query) + " past the query's " +
                 std::to_string(m) + " bases";
        if (s.Reference >= rl || l.Reference > rl)
          return ats + ": reference positions " + std::to_string(s.Reference) + ".." +
                 std::to_string(l.Reference) + " outside the reference's " + std::to_string(rl) +
                 " bases";
      }
    }
  }
  return "";
}

}  // namespace mirage_hip
