// CPU test driver of mirage_hip::CheckScoringInputs (matrix_hip_check.cpp):
// builds well-formed and malformed scoring inputs and prints one line per
// case: the case name and the message ("ok" for an empty one).  No HIP.
#include <cstdio>
#include <string>
#include <vector>

#include "matrix_hip.hpp"

using namespace mirage_hip;

int main() {
  const std::vector<uint64_t> refLens = {100, 50};
  struct Case {
    const char *name;
    std::vector<size_t> qlen;
    std::vector<std::vector<SEQ_ID_t>> qrc;
    std::vector<std::vector<std::vector<Position>>> cp, cpl;
  };
  std::vector<Case> cases;
  // (start.Reference, start.Query) / (last.Reference, last.Query)
  cases.push_back({"good", {40}, {{0, 1}}, {{{{10, 0}}, {{5, 2}}}}, {{{{30, 20}}, {{25, 22}}}}});
  cases.push_back({"skipped_seed", {40}, {{0}}, {{{{500, 30}}}}, {{{{600, 10}}}}});
  cases.push_back({"edge_fits", {40}, {{1}}, {{{{49, 0}}}}, {{{{50, 40}}}}});
  cases.push_back({"bad_ref_id", {40}, {{2}}, {{{}}}, {{{}}}});
  cases.push_back({"last_query_past_read", {40}, {{0}}, {{{{10, 0}}}}, {{{{51, 41}}}}});
  cases.push_back({"start_ref_past_ref", {40}, {{1}}, {{{{50, 0}}}}, {{{{60, 10}}}}});
  cases.push_back({"last_ref_past_ref", {40}, {{1}}, {{{{40, 0}}}}, {{{{51, 10}}}}});
  cases.push_back({"uneven_seed_lists", {40}, {{0}}, {{{{10, 0}, {20, 10}}}}, {{{{30, 20}}}}});
  cases.push_back({"uneven_cand_lists", {40}, {{0, 1}}, {{{}}}, {{{}, {}}}});
  cases.push_back({"missing_query_lists", {40, 30}, {{0}}, {{{}}}, {{{}}}});
  cases.push_back({"query_too_long", {5000}, {{0}}, {{{}}}, {{{}}}});
  for (const Case &c : cases) {
    const std::string m = CheckScoringInputs(c.qlen, refLens, 4096, c.qrc, c.cp, c.cpl);
    std::printf("%s\t%s\n", c.name, m.empty() ? "ok" : m.c_str());
  }
  return 0;
}
