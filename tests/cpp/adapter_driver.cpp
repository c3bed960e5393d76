// CPU driver for the reference-side adapter (mirage_amd/csrc/adapter/
// MatrixGPUHIP.cpp).  The adapter is linked against a recording stand-in of
// mirage_hip::Device defined here, so its type conversion, query slicing,
// reference window and result assembly run without a GPU.  Prints one
// "key<TAB>value" line per check.
#include <cstdio>
#include <sstream>
#include <string>

#include "Analyser.h"
#include "MatrixGPUHIP.h"
#include "matrix_hip.hpp"

namespace {
int g_devices = 0;
size_t g_loaded_refs = 0, g_loaded_bases = 0;
int32_t g_loaded_T = 0;
bool g_throw = false;
std::string g_seen;  // what GetHighestScore received
}  // namespace

namespace mirage_hip {
Device::Device(int device, uint32_t) { g_devices = device + 1; }
Device::~Device() {}
void Device::LoadReferences(const std::vector<std::vector<char>> &refs, int32_t T) {
  g_loaded_refs = refs.size();
  g_loaded_bases = 0;
  for (const auto &r : refs) g_loaded_bases += r.size();
  g_loaded_T = T;
}
size_t Device::MaxReadLength() const { return 4096; }
size_t Device::MaxBlockSize() const { return 123456; }
// Scores candidate c of query q as 10*q + c, except reference 7, which gets
// no score; records the queries, candidates and seeds it was given.
std::vector<MatrixResult> Device::GetHighestScore(
    const std::vector<std::vector<char> *> &q, const std::vector<std::vector<SEQ_ID_t>> &qrc,
    const std::vector<std::vector<std::vector<Position>>> &cp,
    const std::vector<std::vector<std::vector<Position>>> &cpl, bool) {
  if (g_throw) throw Error(MH_ERR_NOMEM, "device allocation");
  std::ostringstream os;
  std::vector<MatrixResult> out(q.size());
  for (size_t i = 0; i < q.size(); i++) {
    os << std::string(q[i]->begin(), q[i]->end()) << ":";
    out[i].QueryID = i;
    for (size_t c = 0; c < qrc[i].size(); c++) {
      os << qrc[i][c] << "(";
      for (size_t s = 0; s < cp[i][c].size(); s++)
        os << cp[i][c][s].Reference << "," << cp[i][c][s].Query << "-" << cpl[i][c][s].Reference
           << "," << cpl[i][c][s].Query << ";";
      os << ")";
      if (qrc[i][c] == 7) continue;
      const float sc = (float)(10 * i + c);
      if (sc > out[i].HighestScore) out[i].HighestScore = sc;
      out[i].Scores.push_back({sc, qrc[i][c]});
    }
    os << "|";
  }
  g_seen = os.str();
  return out;
}
}  // namespace mirage_hip

static std::string show(const vector<MatrixResult> &r) {
  std::ostringstream os;
  for (const auto &m : r) {
    os << m.QueryID << "=" << m.HighestScore << "[";
    for (const auto &s : m.Scores) os << s.referenceID << ":" << s.score << ",";
    os << "]";
  }
  return os.str();
}

int main() {
  InitialiseCuda(2);
  printf("devices\t%d\n", g_devices);
  printf("rowhit\t%zu\n", Get_LOCAL_DATA_rowhit_SIZE());
  printf("blocksize\t%zu\n", CUDAGetMaxBlockSize(1));

  vector<vector<char>> refs{{'A', 'C', 'G'}, {'T', 'T'}};
  MirageHipLoadReferences(refs, 31);
  printf("loaded\t%zu %zu %d\n", g_loaded_refs, g_loaded_bases, g_loaded_T);

  // Five queries; the call scores the slice of queries 1..3 and the
  // reference window [2, 2 + 5] = references 2..7.
  vector<vector<char>> qs{{'a'}, {'b'}, {'c'}, {'d'}, {'e'}};
  vector<vector<char> *> seqQuery;
  for (auto &x : qs) seqQuery.push_back(&x);
  // Per-query lists as Analyser holds them, starting at query 0.
  vector<vector<SEQ_ID_t>> qrc{{0}, {1, 2, 7, 9}, {8}, {3, 4}, {5}};
  vector<vector<vector<Enum::Position>>> cp(5), cpl(5);
  for (size_t q = 0; q < 5; q++)
    for (size_t c = 0; c < qrc[q].size(); c++) {
      cp[q].push_back({Enum::Position(100 * q + c, q), Enum::Position(100 * q + c + 1, q + 1)});
      cpl[q].push_back(
          {Enum::Position(100 * q + c + 30, q + 30), Enum::Position(100 * q + c + 31, q + 31)});
    }
  vector<vector<vector<SEQ_POS_t>>> unused(5);
  vector<vector<vector<Enum::HeadTailSequence>>> ht(5);
  const SEQ_AMOUNT_t qoff = 1, qn = 3, roff = 2, rn = 5;
  vector<MatrixResult> res = CUDAGetHighestScore(
      seqQuery, nullptr, false, 1, 31, &qrc[qoff], &unused[qoff], &unused[qoff], &cp[qoff],
      &cpl[qoff], 512, nullptr, nullptr, roff, rn, qoff, qn, &ht[qoff]);
  printf("seen\t%s\n", g_seen.c_str());
  printf("results\t%s\n", show(res).c_str());

  g_throw = true;
  try {
    CUDAGetHighestScore(seqQuery, nullptr, false, 0, 31, &qrc[qoff], &unused[qoff],
                        &unused[qoff], &cp[qoff], &cpl[qoff], 512, nullptr, nullptr, roff, rn,
                        qoff, qn, &ht[qoff]);
    printf("error\tnone\n");
  } catch (const std::runtime_error &e) {
    printf("error\t%s\n", e.what());
  }
  try {
    CUDAGetMaxBlockSize(2);
    printf("bad_device\tnone\n");
  } catch (const std::runtime_error &e) {
    printf("bad_device\t%s\n", e.what());
  }
  return 0;
}
