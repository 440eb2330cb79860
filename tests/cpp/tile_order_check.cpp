// Tile visiting order of evp1d_tiles.cpp (tile_ids, order 1 = Hilbert curve):
//  1. on full n x n grids of single-cell tiles (n a power of two) the ids
//     are a permutation of 0..n*n-1 and consecutive ids are 4-neighbours,
//     the defining property of the Hilbert curve (a wrong turn breaks it),
//     and the curve starts at (0, 0) and ends at (n - 1, 0);
//  2. on a non-square grid, and on a sparse frame of two far-apart
//     components (the sorted-key path), the ids are dense and every tile's
//     ids are those of the row-major order permuted (same tile membership);
//  3. on a real grid (input file as for pair_plan_check.cpp) the Hilbert
//     ordered plan and its pair tables pass check_plan / check_pairs.
#include "evp1d_tiles.hpp"
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

static evp1d::GridCoords full_grid(int w, int h) {
  evp1d::GridCoords c;
  c.ok = true;
  c.width = w;
  c.height = h;
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      c.gx.push_back(x);
      c.gy.push_back(y);
    }
  return c;
}

static int fails = 0;
#define CHECK(cond, ...)                  \
  do {                                    \
    if (!(cond)) {                        \
      std::printf("FAIL: " __VA_ARGS__);  \
      std::printf("\n");                  \
      ++fails;                            \
    }                                     \
  } while (0)

int main(int argc, char **argv) {
  for (int n : {2, 4, 8, 16, 64}) {
    const evp1d::GridCoords c = full_grid(n, n);
    const evp1d::TileIds t = evp1d::tile_ids(c, 1, 1, 1);
    CHECK(t.ntiles == n * n, "n=%d ntiles %d", n, t.ntiles);
    std::vector<int> at((size_t)n * n, -1);   // id -> cell
    for (int q = 0; q < n * n; ++q) {
      const int id = t.tid[(size_t)q];
      CHECK(id >= 0 && id < n * n && at[(size_t)id] < 0, "n=%d id %d repeated or out of range", n, id);
      if (id >= 0 && id < n * n) at[(size_t)id] = q;
    }
    for (int k = 1; k < n * n; ++k) {
      const int a = at[(size_t)k - 1], b = at[(size_t)k];
      const int d = std::abs(c.gx[(size_t)a] - c.gx[(size_t)b]) + std::abs(c.gy[(size_t)a] - c.gy[(size_t)b]);
      CHECK(d == 1, "n=%d ids %d and %d are not neighbours", n, k - 1, k);
    }
    CHECK(at[0] == 0, "n=%d curve does not start at (0,0)", n);
    CHECK(at[(size_t)n * n - 1] == n - 1, "n=%d curve does not end at (n-1,0)", n);
  }
  {
    const int w = 93, h = 205, tw = 31, th = 15;   // 3 x 14 tiles, ragged bottom row
    const evp1d::GridCoords c = full_grid(w, h);
    const evp1d::TileIds r = evp1d::tile_ids(c, tw, th, 0), t = evp1d::tile_ids(c, tw, th, 1);
    CHECK(r.ntiles == t.ntiles && t.ntiles == 3 * 14, "ntiles %d %d", r.ntiles, t.ntiles);
    std::vector<int> map((size_t)t.ntiles, -1);   // row-major id -> Hilbert id
    for (size_t q = 0; q < c.gx.size(); ++q) {
      int &m = map[(size_t)r.tid[q]];
      if (m < 0) m = t.tid[q];
      CHECK(m == t.tid[q], "cell %zu: tile membership differs between orders", q);
    }
    std::vector<int> seen((size_t)t.ntiles, 0);
    for (int m : map) if (m >= 0 && m < t.ntiles) seen[(size_t)m]++;
    for (int k = 0; k < t.ntiles; ++k) CHECK(seen[(size_t)k] == 1, "Hilbert id %d not used once", k);
    CHECK(evp1d::tile_ids(c, tw, th, 2).ntiles == 0, "order 2 accepted");
  }
  {
    // a sparse frame (two small components far apart, as a batch lays out
    // its members): the sorted-key path, same tile membership in both orders
    evp1d::GridCoords c;
    c.ok = true;
    c.width = 200000;
    c.height = 40;
    for (int y = 0; y < 40; ++y)
      for (int x = 0; x < 30; ++x) {
        c.gx.push_back(x);
        c.gy.push_back(y);
        c.gx.push_back(199970 + x);
        c.gy.push_back(y);
      }
    const evp1d::TileIds r = evp1d::tile_ids(c, 8, 8, 0), t = evp1d::tile_ids(c, 8, 8, 1);
    CHECK(r.ntiles == t.ntiles && t.ntiles == 2 * 4 * 5, "sparse ntiles %d %d", r.ntiles, t.ntiles);
    std::vector<int> map((size_t)std::max(r.ntiles, 1), -1), seen((size_t)std::max(t.ntiles, 1), 0);
    for (size_t q = 0; q < c.gx.size(); ++q) {
      int &m = map[(size_t)r.tid[q]];
      if (m < 0) {
        m = t.tid[q];
        if (m >= 0 && m < t.ntiles) seen[(size_t)m]++;
      }
      CHECK(m == t.tid[q], "sparse cell %zu: tile membership differs between orders", q);
    }
    for (int k = 0; k < t.ntiles; ++k) CHECK(seen[(size_t)k] == 1, "sparse Hilbert id %d not used once", k);
  }
  if (argc >= 2) {
    FILE *f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    int hdr[2];
    if (std::fread(hdr, 4, 2, f) != 2) return 1;
    const int na = hdr[0], navel = hdr[1];
    std::vector<int32_t> idx[6];
    const int32_t *ip[6];
    for (int k = 0; k < 6; ++k) {
      idx[k].resize((size_t)na);
      if (std::fread(idx[k].data(), 4, (size_t)na, f) != (size_t)na) return 1;
      ip[k] = idx[k].data();
    }
    std::vector<uint8_t> skip((size_t)na);
    if (std::fread(skip.data(), 1, (size_t)na, f) != (size_t)na) return 1;
    std::fclose(f);
    const evp1d::GridCoords c = evp1d::grid_coords(na, navel, ip);
    CHECK(c.ok, "grid_coords: %s", c.why.c_str());
    for (int o : {0, 1}) {
      const evp1d::TileIds t = evp1d::tile_ids(c, 31, 15, o);
      const evp1d::TilePlan p = evp1d::plan_tiles(na, navel, ip, skip.data(), t, 2560);
      CHECK(p.ok, "order %d plan: %s", o, p.why.c_str());
      const std::string bad = evp1d::check_plan(p, na, navel, ip, skip.data());
      CHECK(bad.empty(), "order %d check_plan: %s", o, bad.c_str());
      const evp1d::PairPlan q = evp1d::plan_pairs(p, na, navel, ip, skip.data(), 2560, 10240, 1024);
      CHECK(q.ok, "order %d pairs: %s", o, q.why.c_str());
      const std::string bad2 = evp1d::check_pairs(q, p, na, navel, ip, skip.data());
      CHECK(bad2.empty(), "order %d check_pairs: %s", o, bad2.c_str());
      std::printf("order %d: %zu tiles, ok\n", o, p.tiles.size());
    }
  }
  std::printf(fails ? "FAILED %d\n" : "all good\n", fails);
  return fails ? 1 : 0;
}

// the library's error hook (evp1d_device.hip), unused by these checks
extern "C" int evp1d__set_error(int c, const char *) { return c; }
