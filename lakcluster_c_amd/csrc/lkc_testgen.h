// lkc_testgen.h -- the draws of the reference's self-test generators (bcc_init, dcc_init):
// glibc rand() from the state a program that never seeds it starts in, rand01, and the
// cropped draw of the reference's maximum / minimum macros.
#pragma once
#include <stdint.h>

namespace lkc {

// glibc's rand(): the additive generator of random() seeded as srand(1).
// r_i = 16807 r_{i-1} mod (2^31 - 1) for i < 31 (r_0 = 1), r_i = r_{i-31} for i = 31..33,
// then r_i = r_{i-31} + r_{i-3} mod 2^32; the k-th output is r_{k+344} >> 1.
class GlibcRand {
  uint32_t r[34]; unsigned long long pos = 34;
  uint32_t step() { uint32_t v = r[(pos - 31) % 34] + r[(pos - 3) % 34]; r[pos % 34] = v; pos++; return v; }
 public:
  GlibcRand()
  {
    r[0] = 1;
    for (int k = 1; k < 31; k++) { long long v = (16807LL * (int32_t)r[k - 1]) % 2147483647LL; if (v < 0) v += 2147483647LL; r[k] = (uint32_t)v; }
    for (int k = 31; k < 34; k++) r[k] = r[k - 31];
    for (int k = 34; k < 344; k++) step();
  }
  int next() { return (int)(step() >> 1); }
  double r01() { return (double)next() / (double)2147483647; }   // rand01 (lakcluster_header.h:261)
};

// maximum(0.1, minimum(0.9, rand01)) with the reference's macros (lakcluster_header.h:257-258),
// which evaluate the chosen argument again: 2 to 4 draws
inline double crop_draw(GlibcRand& g)
{
  double m1 = 0.9 < g.r01() ? 0.9 : g.r01();
  if (0.1 > m1) return 0.1;
  return 0.9 < g.r01() ? 0.9 : g.r01();
}

}  // namespace lkc
