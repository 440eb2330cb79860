// Each corruption of a valid fused-subcycle tile plan must be reported by
// evp1d::check_plan, the host check every plan passes before the fused kernel
// trusts it (tests/test_tiles.py builds and runs this on a synthetic grid).
// Input file: int32 na, navel; six int32[na] neighbour lists (0-based
// caller positions, ee ne se nw sw sse); uint8[na] skip flags.
#include "evp1d_tiles.hpp"
#include <cstdio>
#include <vector>
#include <functional>
int main(int argc, char **argv) {
  if (argc < 2) return 2;
  FILE *f = fopen(argv[1], "rb");
  if (!f) return 2;
  int hdr[2]; if (fread(hdr, 4, 2, f) != 2) return 1;
  int na = hdr[0], navel = hdr[1];
  std::vector<int32_t> idx[6]; const int32_t *ip[6];
  for (int k = 0; k < 6; ++k) { idx[k].resize(na); if (fread(idx[k].data(), 4, na, f) != (size_t)na) return 1; ip[k] = idx[k].data(); }
  std::vector<uint8_t> skip(na); if (fread(skip.data(), 1, na, f) != (size_t)na) return 1; fclose(f);
  const auto good = evp1d::plan_tiles(na, navel, ip, skip.data(), 32, 8, 2560);
  std::vector<std::pair<const char *, std::function<void(evp1d::TilePlan &)>>> cases = {
    {"swap two uloc entries", [](evp1d::TilePlan &p) { std::swap(p.uloc[400], p.uloc[401]); }},
    {"uloc entry off by one", [](evp1d::TilePlan &p) { p.uloc[1000] += 1; }},
    {"tile b range shortened", [](evp1d::TilePlan &p) { p.tiles[7].b1 -= 1; }},
    {"tile a range overlaps next", [](evp1d::TilePlan &p) { for (auto &t : p.tiles) if (t.a1 > t.a0) { t.a1 += 1; break; } }},
    {"c range dropped", [](evp1d::TilePlan &p) { for (auto &t : p.tiles) if (t.c1 > t.c0) { t.c0 = t.c1; break; } }},
    {"halo entry replaced", [](evp1d::TilePlan &p) { p.list[10] = p.list[11]; }},
    {"hid wrong", [](evp1d::TilePlan &p) { p.hid[5] += 1; }},
    {"shared prefix too long", [](evp1d::TilePlan &p) { for (auto &t : p.tiles) if (t.sb > 0 && t.sb < t.b1 - t.b0) { t.sb += 1; break; } }},
    {"perm swap across classes", [](evp1d::TilePlan &p) { std::swap(p.perm[0], p.perm[p.perm.size() - 1]); }},
    {"perm duplicate", [](evp1d::TilePlan &p) { p.perm[3] = p.perm[4]; }},
    {"konst entry in T range", [](evp1d::TilePlan &p) { if (!p.klist.empty()) p.klist[0] = 0; }},
  };
  int fails = 0;
  printf("good: '%s'\n", evp1d::check_plan(good, na, navel, ip, skip.data()).c_str());
  for (auto &c : cases) {
    auto p = good;
    c.second(p);
    auto r = evp1d::check_plan(p, na, navel, ip, skip.data());
    printf("%-28s -> %s\n", c.first, r.empty() ? "NOT CAUGHT" : r.c_str());
    fails += r.empty();
  }
  return fails;
}
extern "C" int evp1d__set_error(int c, const char *) { return c; }
