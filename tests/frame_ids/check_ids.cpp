// Prints c3d::frame_ids_fit for the triangle and line-fragment counts given
// as argument pairs ("ntri nline ..."), one 0/1 per pair.
#include <cstdio>
#include <cstdlib>

#include "c3d_geom.h"

int main(int argc, char** argv) {
  for (int a = 1; a + 1 < argc; a += 2) {
    const unsigned long long t = std::strtoull(argv[a], nullptr, 0), l = std::strtoull(argv[a + 1], nullptr, 0);
    std::printf("%d\n", c3d::frame_ids_fit(t, l) ? 1 : 0);
  }
  return 0;
}
