// The 1D index builder (evp1d_index_count / evp1d_index_build) on 400 random
// masks: random sizes (1..40 x 1..30 cells inside 1- or 2-cell rims), ocean
// densities from 0 to 100 %, and in a third of them a wet north rim.  Built
// host-only with AddressSanitizer and UBSan by tests/test_tiles.py, so any
// out-of-bounds access or undefined behaviour in the builder fails the test.
// Prints "fails N"; a mask the count accepts must also build.
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "evp1d.h"

extern "C" int evp1d__set_error(int c, const char *) { return c; }

int main() {
  std::mt19937 rng(7);
  int fails = 0, built = 0;
  for (int it = 0; it < 400; ++it) {
    const int nx = 1 + (int)(rng() % 40), ny = 1 + (int)(rng() % 30), ng = 1 + (int)(rng() % 2);
    const int NX = nx + 2 * ng, NY = ny + 2 * ng;
    std::vector<int32_t> m((size_t)NX * NY, 0);
    const int dens = (int)(rng() % 101);
    for (int j = ng; j < NY - ng; ++j)
      for (int i = ng; i < NX - ng; ++i) m[(size_t)j * NX + i] = (int)(rng() % 100) < dens;
    if (it % 3 == 0)
      for (int i = ng; i < NX; ++i) m[(size_t)(NY - 1) * NX + i] = (int)(rng() % 2);
    int32_t na = 0, navel = 0;
    if (evp1d_index_count(NX, NY, ng, m.data(), &na, &navel)) continue;
    std::vector<int32_t> ti(na), tj(na), tij(navel), ee(na), ne(na), se(na), nw(na), sw(na), sse(na);
    const int rc = evp1d_index_build(NX, NY, ng, m.data(), na, navel, ti.data(), tj.data(), tij.data(), ee.data(),
                                     ne.data(), se.data(), nw.data(), sw.data(), sse.data());
    if (rc) {
      ++fails;
      printf("build failed on %d x %d (rim %d): %d\n", NX, NY, ng, rc);
    } else {
      ++built;
    }
  }
  printf("built %d, fails %d\n", built, fails);
  return fails != 0;
}
