// Drives the C++ call surface (mirage_amd/csrc/matrix_hip.hpp) for the GPU
// tests.  Input (text, stdin):
//   T n_refs n_reads n_queries
//   n_refs lines of reference bases, n_reads lines of read bases
//   per query: bases, then "n_cand", then per candidate "ref n_seeds" and
//   4*n_seeds integers (start.ref start.query last.ref last.query)
// Output: per read "C highest_bits n_top ref..." (Device::Classify), then per
// query "S highest_bits n_scores rel_bits ref ..." (Device::GetHighestScore).
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "matrix_hip.hpp"

using namespace mirage_hip;

static uint32_t bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}

int main() {
  int T;
  size_t nr, nq, ns;
  std::cin >> T >> nr >> nq >> ns;
  std::vector<std::vector<char>> refs(nr), reads(nq);
  std::string s;
  for (auto &r : refs) {
    std::cin >> s;
    r.assign(s.begin(), s.end());
  }
  for (auto &r : reads) {
    std::cin >> s;
    r.assign(s.begin(), s.end());
  }
  std::vector<std::vector<char>> qs(ns);
  std::vector<std::vector<SEQ_ID_t>> qrc(ns);
  std::vector<std::vector<std::vector<Position>>> cp(ns), cpl(ns);
  for (size_t q = 0; q < ns; q++) {
    std::cin >> s;
    qs[q].assign(s.begin(), s.end());
    size_t nc;
    std::cin >> nc;
    qrc[q].resize(nc);
    cp[q].resize(nc);
    cpl[q].resize(nc);
    for (size_t c = 0; c < nc; c++) {
      size_t nsd;
      std::cin >> qrc[q][c] >> nsd;
      for (size_t i = 0; i < nsd; i++) {
        Position a, b;
        std::cin >> a.Reference >> a.Query >> b.Reference >> b.Query;
        cp[q][c].push_back(a);
        cpl[q][c].push_back(b);
      }
    }
  }
  try {
    Device dev(0, 4096);
    dev.LoadReferences(refs, T);
    for (const auto &m : dev.Classify(reads)) {
      std::printf("C %08x %zu", bits(m.HighestScore), m.Scores.size());
      for (const auto &x : m.Scores) std::printf(" %u", x.referenceID);
      std::printf("\n");
    }
    std::vector<std::vector<char> *> qp;
    for (auto &q : qs) qp.push_back(&q);
    for (const auto &m : dev.GetHighestScore(qp, qrc, cp, cpl)) {
      std::printf("S %08x %zu", bits(m.HighestScore), m.Scores.size());
      for (const auto &x : m.Scores) std::printf(" %08x %u", bits(x.score), x.referenceID);
      std::printf("\n");
    }
  } catch (const Error &e) {
    std::fprintf(stderr, "error %d: %s\n", e.status, e.what());
    return 1;
  }
  return 0;
}
