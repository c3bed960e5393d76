// Driver that calls the reference's own scoring, relative-score and hash
// functions (built from /root/reference by Makefile.ref) so that tests can
// check the oracle's restatement of them value by value.  Test tool only.
//
// Text protocol on stdin, one request per line group, one answer per line:
//   S T ultrafast qlen reflen nseeds      then lines: query, reference, and
//                                         4*nseeds integers (start.ref,
//                                         start.query, last.ref, last.query)
//     -> "<score bits hex> <percentageMatch bits hex>"
//        from CalculateScore (MatrixCPU.cpp:18-214)
//   R qsize rsize scorebits               -> relative score bits
//        from MatrixResult::GetRelativeScore (MatrixResult.cpp:21-29)
//   H len sequence                        -> CreateHash24 (Hashmap.cpp:3233)
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "Analyser.h"
#include "Hashmap.h"
#include "MatrixCPU.h"
#include "MatrixResult.h"
#include "MultiThreading.h"
#include "Sequences.h"

void CalculateScore(vector<char> *query, vector<Enum::HeadTailSequence> *HeadTailSequence,
                    vector<Enum::Position> *combiStartPositions,
                    vector<Enum::Position> *combiStartPositions_LastPositionInRange,
                    int32_t threshold, MultiThreading *multiThreading,
                    MatrixResult::Result *retVal, bool *HitBuffer, Sequences *sequences,
                    SEQ_ID_t referenceID, size_t HashmapID);

static uint32_t bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}

int main() {
  MultiThreading mt;
  mt.SetGPUActive_NumberOfGPUs(0);
  Hashmap::Hashmap_Load_Reference_Memory_UsageMode = Enum::Hashmap_MemoryUsageMode::Extreme;
  Hashmap::referenceSequences_First_HashStreamPos.assign(1, 0);
  Sequences seqs(nullptr, nullptr, &mt, nullptr, nullptr);
  std::string cmd;
  while (std::cin >> cmd) {
    if (cmd == "S") {
      int T, uf;
      size_t qlen, rlen, ns;
      std::cin >> T >> uf >> qlen >> rlen >> ns;
      std::string q, r;
      std::cin >> q >> r;
      std::vector<Enum::Position> st(ns), la(ns);
      for (size_t i = 0; i < ns; i++) {
        uint32_t a, b, c, d;
        std::cin >> a >> b >> c >> d;
        st[i] = Enum::Position(a, b);
        la[i] = Enum::Position(c, d);
      }
      Analyser::SpeedMode = uf ? Enum::Analysis_Modes::UltraFast : Enum::Analysis_Modes::Sensitive;
      std::vector<char> query(q.begin(), q.end());
      std::vector<char> refbuf(r.begin(), r.end());
      seqs.referenceHashStreamPos.assign(1, 0);
      seqs.referenceSequencesSizes.assign(1, (SEQ_ID_t)rlen);
      seqs.referenceSequences_Buffer.assign(1, refbuf.data());
      std::vector<Enum::HeadTailSequence> ht(ns + 1);
      std::vector<char> hit(qlen + 1, 0);
      MatrixResult::Result res;
      res.percentageMatch = 0;
      CalculateScore(&query, &ht, &st, &la, T, &mt, &res, (bool *)hit.data(), &seqs, 0, 0);
      std::printf("%08x %08x\n", bits(res.Score), bits(res.percentageMatch));
    } else if (cmd == "R") {
      float qs, rs;
      uint32_t sb;
      std::cin >> qs >> rs >> std::hex >> sb >> std::dec;
      float sc;
      std::memcpy(&sc, &sb, 4);
      std::printf("%08x\n", bits(MatrixResult::GetRelativeScore(qs, rs, sc)));
    } else if (cmd == "H") {
      int len;
      std::string s;
      std::cin >> len >> s;
      std::printf("%u\n", Hashmap::CreateHash24((unsigned char *)s.data(), len));
    }
    std::fflush(stdout);
  }
  return 0;
}
