// The reference's result logs, written from the classification results:
// one line per read (per group of equal scores with MH_RESULTS_ALL_GROUPS)
// in results_Raw.log, and in results_HighSensitivity.log /
// results_HighPrecision.log / results_Balanced.log when the group's score
// reaches that file's threshold (MirAge/Analyser.cpp:378-491; line
// :226-376; files LogManager.cpp:294-404, header line :390-404).
//
// Line: InputID;read header;score;taxonomy ids;occurrences;reference
// headers.  Headers are passed as MirAge keeps them and printed without
// their last byte (the NUL, Analyser.cpp:240, :369).  The references of a
// group are ordered exactly as the reference's own calls leave them:
// std::sort of the read's scored candidates by score, descending (:404),
// then std::stable_sort by taxonomy occurrence with the comparator of
// :266-293 as written; both are the same libstdc++ algorithms here, so the
// same inputs give the same permutation.  The scored candidates are those
// of the surviving strand with Score > -1, in ascending reference order
// (MatrixCPU.cpp:323-370); they come from the classification's candidate
// output.  Without it the top-tie set stands in for them, which gives the
// same order whenever a read's scored candidates are its top ties or number
// at most 16 (std::sort is then an insertion sort).
#include <stdint.h>

#include <algorithm>
#include <cstdio>
#include <string>
#include <unordered_map>
#include <vector>

#include "mh_internal.hpp"

namespace {

struct ResultsScores {  // MatrixResult::ResultsScores (MatrixResult.h:30-38)
  float score;
  uint64_t referenceID;
};

const char *const kKinds[4] = {"Raw", "HighSensitivity", "HighPrecision", "Balanced"};
const char kHeaderLine[] = "InputID;InputSequence;Confident;TaxIDs;Occurence_Per_TaxID;Headers\n";

void append_header(std::string &s, const uint8_t *h, const uint64_t *off, uint64_t i) {
  const uint64_t a = off[i], b = off[i + 1];
  if (b > a) s.append((const char *)h + a, (size_t)(b - a - 1));
}

// LogResultsToFileManager (Analyser.cpp:226-376) for one group.
void group_line(std::string &s, uint64_t input_id, const uint8_t *read_headers,
                const uint64_t *read_header_off, uint64_t read, float rel,
                std::vector<uint64_t> &refs, const uint8_t *ref_headers,
                const uint64_t *ref_header_off, const int32_t *taxids) {
  auto tax = [&](uint64_t r) -> int32_t { return taxids ? taxids[r] : -1; };
  s += std::to_string(input_id);
  s += ';';
  append_header(s, read_headers, read_header_off, read);
  s += ';';
  s += std::to_string(rel);
  s += ';';
  if (refs.empty()) {
    s += "?;";
    return;
  }
  std::unordered_map<uint64_t, size_t> occ;  // keyed as SEQ_ID_t
  for (uint64_t r : refs) occ[(uint64_t)(int64_t)tax(r)]++;
  std::stable_sort(refs.begin(), refs.end(), [&](uint64_t a, uint64_t b) {
    const uint64_t ta = (uint64_t)(int64_t)tax(a), tb = (uint64_t)(int64_t)tax(b);
    if (ta == (uint64_t)-1) return true;
    if (tb == (uint64_t)-1) return false;
    if (ta == tb) return false;
    if (occ[ta] == occ[tb]) return false;
    if (occ[ta] > occ[tb]) return true;
    if (occ[ta] < occ[tb]) return false;
    return ta < tb;
  });
  int32_t cur = tax(refs[0]);
  const size_t max_occ = occ[(uint64_t)(int64_t)cur];
  for (int column = 0; column < 2; column++) {  // ids, then occurrences; cur carries over
    for (size_t i = 0; i < refs.size(); i++) {
      const int32_t t = tax(refs[i]);
      if (i > 0 && t == cur) continue;
      const size_t n = occ[(uint64_t)(int64_t)t];
      if (n != max_occ) break;
      cur = t;
      if (i != 0) s += ',';
      s += column == 0 ? std::to_string(cur) : std::to_string(n);
    }
    s += ';';
  }
  for (size_t i = 0; i < refs.size(); i++) {
    if (occ[(uint64_t)(int64_t)tax(refs[i])] != max_occ) break;
    if (i != 0) s += ',';
    append_header(s, ref_headers, ref_header_off, refs[i]);
  }
}

}  // namespace

extern "C" mh_status mh_write_results(const char *dir, uint64_t first_read_id, uint32_t n_reads,
                                      const uint8_t *h_read_headers,
                                      const uint64_t *h_read_header_off,
                                      const mh_read_result *h_results,
                                      const uint32_t *h_top_pool, const mh_cand *h_cands,
                                      uint64_t n_cands, uint32_t n_refs,
                                      const uint8_t *h_ref_headers,
                                      const uint64_t *h_ref_header_off, const int32_t *h_taxids,
                                      const mh_log_thresholds *thresholds, uint32_t flags) {
  using mh::set_error;
  if (!dir || (n_reads && (!h_read_headers || !h_read_header_off || !h_results || !h_top_pool)) ||
      !h_ref_headers || !h_ref_header_off || (n_cands && !h_cands))
    return set_error(nullptr, MH_ERR_ARG, "mh_write_results: bad argument");
  const mh_log_thresholds thr = thresholds ? *thresholds : mh_log_thresholds{0.4f, 1.0f, 0.7f};
  const float limit[4] = {0.0f, thr.high_sensitivity, thr.high_precision, thr.balanced};

  // Scored candidates of each read's surviving strand, by read then reference.
  std::vector<uint64_t> first(n_reads + 1, 0);
  std::vector<ResultsScores> scored;
  if (n_cands) {
    auto keep = [&](const mh_cand &c) {
      return c.read < n_reads && c.strand == h_results[c.read].strand && c.score > -1.0f;
    };
    for (uint64_t i = 0; i < n_cands; i++)
      if (keep(h_cands[i])) first[h_cands[i].read + 1]++;
    for (uint32_t r = 0; r < n_reads; r++) first[r + 1] += first[r];
    scored.resize(first[n_reads]);
    std::vector<uint64_t> fill(first.begin(), first.end() - 1);
    for (uint64_t i = 0; i < n_cands; i++) {
      const mh_cand &c = h_cands[i];
      if (!keep(c)) continue;
      if (c.ref >= n_refs) return set_error(nullptr, MH_ERR_ARG, "mh_write_results: reference id");
      scored[fill[c.read]++] = ResultsScores{c.rel, c.ref};
    }
    for (uint32_t r = 0; r < n_reads; r++)
      std::sort(scored.begin() + first[r], scored.begin() + first[r + 1],
                [](const ResultsScores &a, const ResultsScores &b) {
                  return a.referenceID < b.referenceID;
                });
  }

  std::string out[4];
  std::vector<ResultsScores> sc;
  std::vector<uint64_t> refs;
  std::string line;
  for (uint32_t r = 0; r < n_reads; r++) {
    const mh_read_result &res = h_results[r];
    if (res.n_top == 0 || res.highest == 0.0f) continue;  // Analyser.cpp:388
    sc.clear();
    if (first[r + 1] > first[r]) {
      sc.assign(scored.begin() + first[r], scored.begin() + first[r + 1]);
    } else {
      for (uint32_t t = 0; t < res.n_top; t++) {
        const uint32_t id = h_top_pool[res.top_off + t];
        if (id >= n_refs) return set_error(nullptr, MH_ERR_ARG, "mh_write_results: reference id");
        sc.push_back(ResultsScores{res.highest, id});
      }
    }
    std::sort(sc.begin(), sc.end(), [](const ResultsScores &a, const ResultsScores &b) {
      return a.score > b.score;  // Analyser.cpp:404
    });
    size_t g = 0;
    while (g < sc.size()) {
      const float rel = sc[g].score;
      refs.clear();
      size_t e = g;
      for (; e < sc.size() && sc[e].score == rel; e++) refs.push_back(sc[e].referenceID);
      line.clear();
      group_line(line, first_read_id + r, h_read_headers, h_read_header_off, r, rel, refs,
                 h_ref_headers, h_ref_header_off, h_taxids);
      line += '\n';
      for (int k = 0; k < 4; k++)
        if (k == 0 || rel >= limit[k]) out[k] += line;
      if (!(flags & MH_RESULTS_ALL_GROUPS)) break;  // output_firstresults_only
      g = e;
    }
  }

  for (int k = 0; k < 4; k++) {
    const std::string path = std::string(dir) + "/results_" + kKinds[k] + ".log";
    FILE *f = std::fopen(path.c_str(), (flags & MH_RESULTS_APPEND) ? "ab" : "wb");
    if (!f) return set_error(nullptr, MH_ERR_ARG, "mh_write_results: cannot open " + path);
    bool ok = true;
    if (!(flags & MH_RESULTS_APPEND))
      ok = std::fwrite(kHeaderLine, 1, sizeof(kHeaderLine) - 1, f) == sizeof(kHeaderLine) - 1;
    ok = ok && std::fwrite(out[k].data(), 1, out[k].size(), f) == out[k].size();
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return set_error(nullptr, MH_ERR_ARG, "mh_write_results: write failed: " + path);
  }
  return MH_OK;
}
