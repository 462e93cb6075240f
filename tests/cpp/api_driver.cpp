// Test driver of the reference-compatible C++ API (include/pbig.h): reads
// boxes from a file, calls pbig_init / pbig_check / pbig_check_stream /
// pbig_decode_stream / pbig_finalize and prints the pair counts.
//   api_driver <file> <n> <range> <cs>
// The file holds n records of 2*DIMS coordinates (float, or unsigned with
// PBIG_COMPAT_U32).
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "pbig.h"
#include "rect.h"
#include "report.h"

int main(int argc, char** argv) {
  if (argc != 5) return 2;
  const unsigned n = (unsigned)atol(argv[2]), range = (unsigned)atol(argv[3]), cs = (unsigned)atol(argv[4]);
  const int per = (int)(sizeof(CoordInfo) / sizeof(CType));
  std::vector<CType> coords((size_t)n * per);
  FILE* f = fopen(argv[1], "rb");
  if (!f || fread(coords.data(), sizeof(CType), coords.size(), f) != coords.size()) return 3;
  fclose(f);
  std::vector<RectInfo> rects(n);
  for (unsigned i = 0; i < n; ++i) {
    CType* c = (CType*)&rects[i].coordinate;
    for (int d = 0; d < per; ++d) c[d] = coords[(size_t)i * per + d];
    rects[i].rid = i;
  }
  pbig_init(n, range, cs, 0, 0);
  report r1;
  pbig_check(rects.data(), &r1);
  std::vector<unsigned long long> words;
  std::vector<unsigned int> start, recs;
  pbig_check_stream(rects.data(), &words, &start, &recs);
  report r2;
  pbig_decode_stream(words, start, recs, &r2);
  pbig_finalize();
  printf("pairs %llu stream_pairs %llu words %zu\n", r1.count(), r2.count(), words.size());
  return 0;
}
