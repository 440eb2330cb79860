// Tiling of the 1D point list for the fused subcycle kernel (host only).
// The layout and the roles of halo and constant entries are described in
// evp1d_tiles.hpp.
#include "evp1d_tiles.hpp"

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <string>
#include <thread>

namespace evp1d {

namespace {

// stress only | stress and stepu | stepu only | neither
inline int point_class(uint8_t b) {
  const bool t = !(b & 1), u = !(b & 2);
  return t && !u ? 0 : (t && u ? 1 : (u ? 2 : 3));
}

// grid offset of each neighbour list: ee ne se nw sw sse
constexpr int kDx[6] = {-1, -1, 0, 1, 1, 0};
constexpr int kDy[6] = {0, -1, -1, 0, 1, 1};

// class order with caller order inside each class (the untiled layout)
void class_order(int na, const uint8_t *skip, TilePlan &p) {
  std::fill(p.cnt, p.cnt + 4, 0);
  for (int q = 0; q < na; ++q) ++p.cnt[point_class(skip[q])];
  int start[4] = {0, p.cnt[0], p.cnt[0] + p.cnt[1], p.cnt[0] + p.cnt[1] + p.cnt[2]};
  p.perm.assign((size_t)na, 0);
  for (int q = 0; q < na; ++q) p.perm[(size_t)start[point_class(skip[q])]++] = q;
}

// Grid coordinates of every point from the neighbour lists, one frame per
// connected component, components placed side by side.  Returns false when
// the lists are not those of a grid (a point reached at two positions).
bool grid_coordinates(int na, const int32_t *const idx[6], std::vector<int64_t> &gx, std::vector<int64_t> &gy,
                      std::string &why) {
  std::vector<int64_t> x((size_t)na), y((size_t)na);
  std::vector<int32_t> comp((size_t)na, -1), queue;
  queue.reserve((size_t)na);
  gx.assign((size_t)na, 0);
  gy.assign((size_t)na, 0);
  int64_t frame_x = 0;
  int ncomp = 0;
  for (int r = 0; r < na; ++r) {
    if (comp[(size_t)r] >= 0) continue;
    queue.clear();
    queue.push_back(r);
    comp[(size_t)r] = ncomp;
    x[(size_t)r] = 0;
    y[(size_t)r] = 0;
    int64_t xmin = 0, xmax = 0, ymin = 0;
    for (size_t h = 0; h < queue.size(); ++h) {
      const int p = queue[h];
      for (int k = 0; k < 6; ++k) {
        const int32_t q = idx[k][p];
        if (q < 0 || q >= na) continue;   // ghost slot: no coordinates needed
        const int64_t qx = x[(size_t)p] + kDx[k], qy = y[(size_t)p] + kDy[k];
        if (comp[(size_t)q] < 0) {
          comp[(size_t)q] = ncomp;
          x[(size_t)q] = qx;
          y[(size_t)q] = qy;
          xmin = std::min(xmin, qx);
          xmax = std::max(xmax, qx);
          ymin = std::min(ymin, qy);
          queue.push_back(q);
        } else if (x[(size_t)q] != qx || y[(size_t)q] != qy) {
          char buf[160];
          snprintf(buf, sizeof buf, "point %d is reached from point %d at two grid positions", q + 1, p + 1);
          why = buf;
          return false;
        }
      }
    }
    for (int p : queue) {
      gx[(size_t)p] = x[(size_t)p] - xmin + frame_x;
      gy[(size_t)p] = y[(size_t)p] - ymin;
    }
    frame_x += xmax - xmin + 2;
    ++ncomp;
  }
  return true;
}

}  // namespace

GridCoords grid_coords(int na, int navel, const int32_t *const idx[6]) {
  GridCoords c;
  if (na == 0) {
    c.why = "empty list";
    return c;
  }
  if (!idx) {
    c.why = "neighbour lists not set";
    return c;
  }
  for (int k = 0; k < 6; ++k)
    if (!idx[k]) {
      c.why = "neighbour lists not set";
      return c;
    }
  for (int k = 0; k < 6; ++k)
    for (int q = 0; q < na; ++q)
      if (idx[k][q] < 0 || idx[k][q] >= navel) {
        c.why = "neighbour index out of range";
        return c;
      }
  std::vector<int64_t> gx, gy;
  if (!grid_coordinates(na, idx, gx, gy, c.why)) return c;
  for (int q = 0; q < na; ++q) {
    c.width = std::max(c.width, gx[(size_t)q] + 1);
    c.height = std::max(c.height, gy[(size_t)q] + 1);
  }
  if (c.width > INT32_MAX || c.height > INT32_MAX) {
    c.why = "grid frame too large";
    return c;
  }
  c.gx.assign(gx.begin(), gx.end());
  c.gy.assign(gy.begin(), gy.end());
  c.ok = true;
  return c;
}

// Position of (x, y) along the Hilbert curve that fills the n x n square (n a
// power of two): the classic bit-pair recursion (Hilbert 1891; the iterative
// form of Butz 1971), which visits each quadrant in the order the curve enters
// it and turns the coordinates so the sub-curve starts where the last ended.
static int64_t hilbert_index(int64_t n, int64_t x, int64_t y) {
  int64_t d = 0;
  for (int64_t s = n / 2; s > 0; s /= 2) {
    const int64_t rx = (x & s) > 0 ? 1 : 0, ry = (y & s) > 0 ? 1 : 0;
    d += s * s * ((3 * rx) ^ ry);
    if (ry == 0) {
      if (rx == 1) {
        x = n - 1 - x;
        y = n - 1 - y;
      }
      std::swap(x, y);
    }
  }
  return d;
}

TileIds tile_ids(const GridCoords &c, int tile_w, int tile_h, int order) {
  TileIds t;
  t.tile_w = tile_w;
  t.tile_h = tile_h;
  t.order = order;
  if (!c.ok) {
    t.why = c.why;
    return t;
  }
  if (tile_w < 1 || tile_h < 1) {
    t.why = "tile size < 1";
    return t;
  }
  if (order != 0 && order != 1) {
    t.why = "tile order must be 0 or 1";
    return t;
  }
  const int na = (int)c.gx.size();
  // dense tile ids in (row of tiles, column of tiles) order, or along a
  // Hilbert curve over the square of side hn that holds the tile grid
  const int64_t ntx = (c.width + tile_w - 1) / tile_w, nty = (c.height + tile_h - 1) / tile_h;
  int64_t hn = 1;
  while (hn < ntx || hn < nty) hn *= 2;
  const int64_t nkey = order == 1 ? hn * hn : ntx * nty;
  auto key_of = [&](int q) {
    const int64_t tx = c.gx[(size_t)q] / tile_w, ty = c.gy[(size_t)q] / tile_h;
    return order == 1 ? hilbert_index(hn, tx, ty) : ty * ntx + tx;
  };
  t.tid.resize((size_t)na);
  if (nkey <= 4 * (int64_t)na + 1024 && nkey <= 4 * ntx * nty + 1024) {   // a table over all tile positions
    std::vector<int32_t> dense((size_t)nkey, -1);
    for (int q = 0; q < na; ++q) dense[(size_t)key_of(q)] = 0;
    int32_t n = 0;
    for (int32_t &d : dense)
      if (d == 0) d = n++;
    for (int q = 0; q < na; ++q) t.tid[(size_t)q] = dense[(size_t)key_of(q)];
    t.ntiles = n;
  } else {   // sparse frames (many small components): sort the keys
    std::vector<int64_t> keys((size_t)na);
    for (int q = 0; q < na; ++q) keys[(size_t)q] = key_of(q);
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    for (int q = 0; q < na; ++q)
      t.tid[(size_t)q] = (int32_t)(std::lower_bound(keys.begin(), keys.end(), key_of(q)) - keys.begin());
    t.ntiles = (int)keys.size();
  }
  return t;
}

TileIds tile_ids(int na, int navel, const int32_t *const idx[6], int tile_w, int tile_h) {
  return tile_ids(grid_coords(na, navel, idx), tile_w, tile_h);
}

TilePlan plan_tiles(int na, int navel, const int32_t *const idx[6], const uint8_t *skip, int tile_w, int tile_h,
                    int max_entries) {
  return plan_tiles(na, navel, idx, skip, tile_ids(na, navel, idx, tile_w, tile_h), max_entries);
}

namespace {

// Host threads for the planner and its check: the machine's, at most 16.
int plan_threads() {
  const unsigned hw = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(hw ? hw : 1u, 16u));
}

// Run body(w) for w = 0..T-1, on T-1 new threads and the caller; when a
// thread cannot be started (a process thread limit) its share runs on the
// caller, so the work is always done.  An exception thrown by a share (a
// host allocation failure) is caught on its thread and the first one is
// rethrown on the caller after every thread has joined; the C entry points
// turn it into a status code (no_throw).
template <class B>
void run_workers(int T, B body) {
  std::vector<std::exception_ptr> err((size_t)T);
  auto guarded = [&body, &err](int w) {
    try {
      body(w);
    } catch (...) {
      err[(size_t)w] = std::current_exception();
    }
  };
  std::vector<std::thread> th;
  std::vector<int> left;
  try {
    th.reserve((size_t)T);
    left.reserve((size_t)T);
  } catch (...) {
    for (int w = 0; w < T; ++w) guarded(w);   // no room for the bookkeeping: all on the caller
    for (const auto &e : err)
      if (e) std::rethrow_exception(e);
    return;
  }
  for (int w = 1; w < T; ++w) {
    try {
      th.emplace_back([&guarded, w] { guarded(w); });
    } catch (...) {
      left.push_back(w);
    }
  }
  guarded(0);
  for (int w : left) guarded(w);
  for (auto &t : th) t.join();
  for (const auto &e : err)
    if (e) std::rethrow_exception(e);
}

// f(lo, hi, w) over [0, n) in contiguous chunks, one per worker w, in order.
template <class F>
void par_chunks(int64_t n, int64_t grain, F f) {
  const int T = (int)std::max<int64_t>(1, std::min<int64_t>(plan_threads(), n / std::max<int64_t>(grain, 1)));
  if (T <= 1) {
    f((int64_t)0, n, 0);
    return;
  }
  run_workers(T, [&](int w) { f(n * w / T, n * (w + 1) / T, w); });
}

// f(i, w) for every i in [0, n), items handed out dynamically (uneven work)
template <class F>
void par_items(int n, F f) {
  const int T = std::max(1, std::min(plan_threads(), n / 4));
  if (T <= 1) {
    for (int i = 0; i < n; ++i) f(i, 0);
    return;
  }
  std::atomic<int> next{0};
  run_workers(T, [&](int w) {
    for (int i = next.fetch_add(8); i < n; i = next.fetch_add(8))
      for (int j = i; j < std::min(n, i + 8); ++j) f(j, w);
  });
}

// Small open-addressing map slot -> local entry number, cleared per tile;
// it doubles when half full, so any tile size works.
struct SlotMap {
  std::vector<int32_t> key, val, used;
  explicit SlotMap(int cap_pow2 = 1 << 13) : key((size_t)cap_pow2, -1), val((size_t)cap_pow2, 0) {}
  void grow() {
    std::vector<int32_t> k2, v2;
    for (int32_t h : used) { k2.push_back(key[(size_t)h]); v2.push_back(val[(size_t)h]); }
    key.assign(key.size() * 2, -1);
    val.assign(val.size(), 0);
    used.clear();
    bool a;
    for (size_t i = 0; i < k2.size(); ++i) find_or_add(k2[i], v2[i], a);
  }
  int find_or_add(int32_t k, int32_t v_new, bool &added) {
    if (2 * (used.size() + 1) > key.size()) grow();
    const size_t m = key.size() - 1;
    size_t h = ((uint32_t)k * 2654435761u) & m;
    while (key[h] != -1 && key[h] != k) h = (h + 1) & m;
    added = key[h] == -1;
    if (added) {
      key[h] = k;
      val[h] = v_new;
      used.push_back((int32_t)h);
    }
    return val[h];
  }
  int get(int32_t k) const {
    const size_t m = key.size() - 1;
    size_t h = ((uint32_t)k * 2654435761u) & m;
    while (key[h] != k) h = (h + 1) & m;
    return val[h];
  }
  void clear() {
    for (int32_t h : used) key[(size_t)h] = -1;
    used.clear();
  }
};

struct TileWork {
  std::vector<int32_t> halo, konst;
  int entries = 0;
};

}  // namespace

// The plan is a pure function of the lists, the flags and the tile ids; the
// work is split over host threads, and every part is computed in an order
// that does not depend on the number of threads, so the plan does not either.
TilePlan plan_tiles(int na, int navel, const int32_t *const idx[6], const uint8_t *skip, const TileIds &ids,
                    int max_entries) {
  TilePlan p;
  p.tile_w = ids.tile_w;
  p.tile_h = ids.tile_h;
  if (ids.ntiles == 0 || (int)ids.tid.size() != na || !idx) {
    class_order(na, skip, p);
    p.why = ids.why.empty() ? "no tiles" : ids.why;
    return p;
  }
  const int nt = ids.ntiles;
  const std::vector<int32_t> &tid = ids.tid;

  // Shared points: active T points that a stepu point of another tile reads
  // (nw, sse, sw), i.e. that appear in some tile's halo.  Every writer stores 1.
  std::vector<uint8_t> shf((size_t)na, 0);
  par_chunks(na, 1 << 15, [&](int64_t lo, int64_t hi, int) {
    for (int64_t q = lo; q < hi; ++q) {
      const int c = point_class(skip[q]);
      if (c != 1 && c != 2) continue;
      for (int n : {3, 5, 4}) {
        const int32_t r = idx[n][q];
        if (r >= 0 && r < na && point_class(skip[r]) <= 1 && tid[(size_t)r] != tid[(size_t)q])
          __atomic_store_n(&shf[(size_t)r], (uint8_t)1, __ATOMIC_RELAXED);
      }
    }
  });

  // counting sort by (class, tile, shared first), caller order inside:
  // per-chunk histograms, then each chunk scatters its points in order
  const size_t nkeys = (size_t)8 * nt;
  auto key = [&](int64_t q) {
    return ((size_t)point_class(skip[q]) * nt + tid[(size_t)q]) * 2 + (shf[(size_t)q] ? 0 : 1);
  };
  const int T = std::max(1, std::min(plan_threads(), na / (1 << 15)));
  std::vector<std::vector<int32_t>> hist((size_t)T, std::vector<int32_t>(nkeys, 0));
  auto chunk_lo = [&](int w) { return (int64_t)na * w / T; };
  run_workers(T, [&](int w) {
    for (int64_t q = chunk_lo(w); q < chunk_lo(w + 1); ++q) ++hist[(size_t)w][key(q)];
  });
  std::vector<int32_t> start(nkeys + 1, 0);
  {
    int64_t run = 0;
    for (size_t k = 0; k < nkeys; ++k) {
      start[k] = (int32_t)run;
      for (int w = 0; w < T; ++w) {
        const int32_t c = hist[(size_t)w][k];
        hist[(size_t)w][k] = (int32_t)run;   // this chunk's first position for key k
        run += c;
      }
    }
    start[nkeys] = (int32_t)run;
  }
  p.perm.assign((size_t)na, 0);
  run_workers(T, [&](int w) {
    std::vector<int32_t> &pos = hist[(size_t)w];
    for (int64_t q = chunk_lo(w); q < chunk_lo(w + 1); ++q) p.perm[(size_t)pos[key(q)]++] = (int32_t)q;
  });
  for (int c = 0; c < 4; ++c) p.cnt[c] = start[(size_t)(c + 1) * nt * 2] - start[(size_t)c * nt * 2];
  std::vector<int32_t> inv((size_t)na);
  par_chunks(na, 1 << 16, [&](int64_t lo, int64_t hi, int) {
    for (int64_t s = lo; s < hi; ++s) inv[(size_t)p.perm[(size_t)s]] = (int32_t)s;
  });
  auto slot = [&](int32_t q) { return q < na ? inv[(size_t)q] : q; };

  const int t_end = p.cnt[0] + p.cnt[1], u_begin = p.cnt[0];
  const int nu = p.cnt[1] + p.cnt[2];
  p.uloc.resize((size_t)4 * nu);   // every entry is written below when the plan succeeds
  std::vector<TileDesc> desc((size_t)nt);
  std::vector<TileWork> work((size_t)nt);
  Table<int32_t> lidx((size_t)t_end);   // owned slot -> entry in its tile's table (written before read)
  std::vector<SlotMap> maps;
  maps.reserve((size_t)plan_threads());
  for (int w = 0; w < plan_threads(); ++w) maps.emplace_back();
  par_items(nt, [&](int t, int w) {
    auto at = [&](int c, int half) { return start[((size_t)c * nt + t) * 2 + half]; };
    TileDesc &d = desc[(size_t)t];
    d = TileDesc{};
    d.a0 = at(0, 0);
    d.a1 = at(0, 2);
    d.sa = at(0, 1) - d.a0;
    d.b0 = at(1, 0);
    d.b1 = at(1, 2);
    d.sb = at(1, 1) - d.b0;
    d.c0 = at(2, 0);
    d.c1 = at(2, 2);
    const int nA = d.a1 - d.a0, nB = d.b1 - d.b0, nO = nA + nB, nUt = nB + (d.c1 - d.c0);
    if (nO + (d.c1 - d.c0) == 0) return;
    for (int s = d.a0; s < d.a1; ++s) lidx[(size_t)s] = s - d.a0;
    for (int s = d.b0; s < d.b1; ++s) lidx[(size_t)s] = nA + (s - d.b0);
    auto owned = [&](int32_t r) { return (r >= d.a0 && r < d.a1) || (r >= d.b0 && r < d.b1); };
    // the four slots each stepu point reads (own, nw: str2/str7, sse: str3/str6,
    // sw: str4/str8), B then C points; non-owned ones get halo (active T) or
    // constant entries in first-visit order
    SlotMap &m = maps[(size_t)w];
    TileWork &tw = work[(size_t)t];
    auto refs = [&](int k, int32_t out[4]) {
      const int s = k < nB ? d.b0 + k : d.c0 + (k - nB);
      const int32_t q = p.perm[(size_t)s];
      out[0] = s;
      out[1] = slot(idx[3][q]);
      out[2] = slot(idx[5][q]);
      out[3] = slot(idx[4][q]);
      return s;
    };
    for (int k = 0; k < nUt; ++k) {
      int32_t r[4];
      refs(k, r);
      for (int c = 0; c < 4; ++c) {
        if (owned(r[c])) continue;
        bool added = false;
        std::vector<int32_t> &lst = r[c] < t_end ? tw.halo : tw.konst;
        m.find_or_add(r[c], (int32_t)lst.size() | (r[c] < t_end ? 0 : (int32_t)0x40000000), added);
        if (added) lst.push_back(r[c]);
      }
    }
    const int nH = (int)tw.halo.size(), nK = (int)tw.konst.size();
    tw.entries = nO + nH + nK;
    if (tw.entries <= max_entries && tw.entries <= 65535) {
      for (int k = 0; k < nUt; ++k) {
        int32_t r[4];
        const int s = refs(k, r);
        for (int c = 0; c < 4; ++c) {
          int e;
          if (owned(r[c])) {
            e = lidx[(size_t)r[c]];
          } else {
            const int v = m.get(r[c]);
            e = (v & 0x40000000) ? nO + nH + (v & 0x3fffffff) : nO + v;
          }
          p.uloc[(size_t)4 * (s - u_begin) + c] = (uint16_t)e;
        }
      }
    }
    m.clear();
  });
  // tiles with work, in tile order; the first that overflows LDS fails the plan
  for (int t = 0; t < nt; ++t) {
    TileDesc d = desc[(size_t)t];
    const TileWork &tw = work[(size_t)t];
    if ((d.a1 - d.a0) + (d.b1 - d.b0) + (d.c1 - d.c0) == 0) continue;
    if (tw.entries > max_entries || tw.entries > 65535) {
      char buf[160];
      snprintf(buf, sizeof buf, "tile %d needs %d LDS entries (limit %d)", t, tw.entries, max_entries);
      p.why = buf;
      p.tiles.clear();
      p.list.clear();
      p.klist.clear();
      return p;
    }
    p.max_entries = std::max(p.max_entries, tw.entries);
    d.h0 = (int32_t)p.list.size();
    d.nh = (int32_t)tw.halo.size();
    d.k0 = (int32_t)p.klist.size();
    d.nk = (int32_t)tw.konst.size();
    p.list.insert(p.list.end(), tw.halo.begin(), tw.halo.end());
    p.klist.insert(p.klist.end(), tw.konst.begin(), tw.konst.end());
    p.tiles.push_back(d);
  }
  // shared ids: the flagged stress slots in slot order (exactly the slots of
  // the halo lists; the check re-derives that), per-entry and per-slot maps
  // (two passes over chunks of slots: count the flagged ones, then number
  // them in slot order)
  {
    const int Ts = std::max(1, std::min(plan_threads(), t_end / (1 << 15)));
    auto lo_of = [&](int w) { return (int64_t)t_end * w / Ts; };
    std::vector<int64_t> first((size_t)Ts + 1, 0);
    run_workers(Ts, [&](int w) {
      int64_t c = 0;
      for (int64_t s = lo_of(w); s < lo_of(w + 1); ++s) c += shf[(size_t)p.perm[(size_t)s]] ? 1 : 0;
      first[(size_t)w + 1] = c;
    });
    for (int w = 0; w < Ts; ++w) first[(size_t)w + 1] += first[(size_t)w];
    p.sid.resize((size_t)t_end);
    p.shared.resize((size_t)first[(size_t)Ts]);
    run_workers(Ts, [&](int w) {
      int64_t id = first[(size_t)w];
      for (int64_t s = lo_of(w); s < lo_of(w + 1); ++s) {
        if (shf[(size_t)p.perm[(size_t)s]]) {
          p.sid[(size_t)s] = (int32_t)id;
          p.shared[(size_t)id++] = (int32_t)s;
        } else {
          p.sid[(size_t)s] = -1;
        }
      }
    });
  }
  p.hid.resize(p.list.size());
  par_chunks((int64_t)p.list.size(), 1 << 15, [&](int64_t lo, int64_t hi, int) {
    for (int64_t k = lo; k < hi; ++k) p.hid[(size_t)k] = p.sid[(size_t)p.list[(size_t)k]];
  });
  for (TileDesc &d : p.tiles) {
    d.ia = d.sa > 0 ? p.sid[(size_t)d.a0] : 0;
    d.ib = d.sb > 0 ? p.sid[(size_t)d.b0] : 0;
  }
  p.ok = true;
  return p;
}

namespace {
// every non-empty range [lo, hi) of `r`, sorted, must tile [from, to) exactly
bool ranges_partition(std::vector<std::pair<int32_t, int32_t>> r, int32_t from, int32_t to) {
  std::sort(r.begin(), r.end());
  int32_t at = from;
  for (const auto &x : r) {
    if (x.first == x.second) continue;
    if (x.first != at) return false;
    at = x.second;
  }
  return at == to;
}
}  // namespace

std::string check_plan(const TilePlan &p, int na, int navel, const int32_t *const idx[6], const uint8_t *skip) {
  if ((int)p.perm.size() != na) return "perm has the wrong length";
  std::vector<int32_t> inv((size_t)na, -1);
  std::atomic<bool> bad_perm{false};
  par_chunks(na, 1 << 16, [&](int64_t lo, int64_t hi, int) {
    for (int64_t s = lo; s < hi; ++s) {
      const int32_t q = p.perm[(size_t)s];
      int32_t expect = -1;
      if (q < 0 || q >= na ||
          !__atomic_compare_exchange_n(&inv[(size_t)q], &expect, (int32_t)s, false, __ATOMIC_RELAXED,
                                       __ATOMIC_RELAXED)) {
        bad_perm = true;
        return;
      }
    }
  });
  if (bad_perm) return "perm is not a permutation";
  {
    int lo = 0;
    for (int c = 0; c < 4; ++c) {
      if (p.cnt[c] < 0 || lo + p.cnt[c] > na) return "class sizes do not add up to na";
      lo += p.cnt[c];
    }
    if (lo != na) return "class sizes do not add up to na";
  }
  std::atomic<bool> bad_class{false};
  par_chunks(na, 1 << 16, [&](int64_t lo, int64_t hi, int) {
    int c = 0, edge = p.cnt[0];
    for (int64_t s = lo; s < hi; ++s) {
      while (s >= edge && c < 3) edge += p.cnt[++c];
      if (point_class(skip[p.perm[(size_t)s]]) != c) {
        bad_class = true;
        return;
      }
    }
  });
  if (bad_class) return "slot outside its class range";
  if (!p.ok) return "";
  const int t_end = p.cnt[0] + p.cnt[1], u_begin = p.cnt[0], u_end = p.cnt[0] + p.cnt[1] + p.cnt[2];
  if ((int)p.uloc.size() != 4 * (u_end - u_begin)) return "uloc has the wrong length";
  if ((int)p.sid.size() != t_end || p.hid.size() != p.list.size()) return "shared-id tables have the wrong size";
  auto slot = [&](int32_t q) { return q < na ? inv[(size_t)q] : q; };
  // ownership: the owned stress ranges of all tiles partition the stress
  // slots, and the stepu ranges partition the stepu slots
  std::vector<std::pair<int32_t, int32_t>> ra, rb, rc;
  for (const TileDesc &d : p.tiles) {
    if (!(0 <= d.a0 && d.a0 <= d.a1 && d.a1 <= p.cnt[0])) return "stress-only range outside its class";
    if (!(p.cnt[0] <= d.b0 && d.b0 <= d.b1 && d.b1 <= t_end)) return "stress-and-stepu range outside its class";
    if (!(t_end <= d.c0 && d.c0 <= d.c1 && d.c1 <= u_end)) return "stepu-only range outside its class";
    ra.emplace_back(d.a0, d.a1);
    rb.emplace_back(d.b0, d.b1);
    rc.emplace_back(d.c0, d.c1);
  }
  if (!ranges_partition(ra, 0, p.cnt[0]) || !ranges_partition(rb, p.cnt[0], t_end))
    return "active T slots not owned exactly once";
  if (!ranges_partition(rc, t_end, u_end)) return "active U slots not stepped exactly once";
  // every tile's table: bounds, entries and the LDS entry of every stepu read
  std::vector<std::string> err((size_t)p.tiles.size());
  std::vector<std::vector<int32_t>> emaps((size_t)plan_threads());
  par_items((int)p.tiles.size(), [&](int t, int w) {
    const TileDesc &d = p.tiles[(size_t)t];
    std::string &e = err[(size_t)t];
    if (d.h0 < 0 || d.nh < 0 || (size_t)d.h0 + d.nh > p.list.size()) { e = "halo list range"; return; }
    if (d.k0 < 0 || d.nk < 0 || (size_t)d.k0 + d.nk > p.klist.size()) { e = "constant list range"; return; }
    const int nA = d.a1 - d.a0, nO = nA + (d.b1 - d.b0), ne = nO + d.nh + d.nk;
    if (ne > p.max_entries) { e = "tile exceeds max_entries"; return; }
    // owned shared points first, with consecutive shared ids
    if (d.sa < 0 || d.sa > nA || d.sb < 0 || d.sb > d.b1 - d.b0) { e = "shared prefix longer than its range"; return; }
    for (int s = d.a0; s < d.a1; ++s)
      if (p.sid[(size_t)s] != (s - d.a0 < d.sa ? d.ia + (s - d.a0) : -1)) { e = "owned slot with a wrong shared id"; return; }
    for (int s = d.b0; s < d.b1; ++s)
      if (p.sid[(size_t)s] != (s - d.b0 < d.sb ? d.ib + (s - d.b0) : -1)) { e = "owned slot with a wrong shared id"; return; }
    std::vector<int32_t> &emap = emaps[(size_t)w];
    emap.assign((size_t)ne, -1);
    for (int s = d.a0; s < d.a1; ++s) emap[(size_t)(s - d.a0)] = s;
    for (int s = d.b0; s < d.b1; ++s) emap[(size_t)(nA + s - d.b0)] = s;
    for (int k = 0; k < d.nh; ++k) {
      const int32_t g = p.list[(size_t)d.h0 + k];
      if (g < 0 || g >= t_end) { e = "halo entry is not an active T slot"; return; }
      if ((g >= d.a0 && g < d.a1) || (g >= d.b0 && g < d.b1)) { e = "halo entry owned by its own tile"; return; }
      emap[(size_t)(nO + k)] = g;
    }
    for (int k = 0; k < d.nk; ++k) {
      const int32_t g = p.klist[(size_t)d.k0 + k];
      if (g < t_end || g >= navel) { e = "constant entry is an active T slot or out of range"; return; }
      emap[(size_t)(nO + d.nh + k)] = g;
    }
    auto check_u = [&](int s) -> bool {
      const int32_t q = p.perm[(size_t)s];
      const int32_t want[4] = {s, slot(idx[3][q]), slot(idx[5][q]), slot(idx[4][q])};
      for (int k = 0; k < 4; ++k) {
        const int en = p.uloc[(size_t)4 * (s - u_begin) + k];
        if (en >= ne || emap[(size_t)en] != want[k]) return false;
      }
      return true;
    };
    for (int s = d.b0; s < d.b1; ++s)
      if (!check_u(s)) { e = "stepu slot " + std::to_string(s) + ": bad LDS entry"; return; }
    for (int s = d.c0; s < d.c1; ++s)
      if (!check_u(s)) { e = "stepu slot " + std::to_string(s) + ": bad LDS entry"; return; }
  });
  for (const std::string &e : err)
    if (!e.empty()) return e;
  // shared ids: every halo entry's id names its slot, ids are distinct
  for (size_t k = 0; k < p.shared.size(); ++k) {
    const int32_t g = p.shared[k];
    if (g < 0 || g >= t_end || p.sid[(size_t)g] != (int32_t)k) return "shared id table inconsistent";
  }
  std::atomic<bool> bad_hid{false};
  par_chunks((int64_t)p.list.size(), 1 << 15, [&](int64_t lo, int64_t hi, int) {
    for (int64_t k = lo; k < hi; ++k) {
      const int32_t id = p.hid[(size_t)k];
      if (id < 0 || id >= (int32_t)p.shared.size() || p.shared[(size_t)id] != p.list[(size_t)k]) {
        bad_hid = true;
        return;
      }
    }
  });
  if (bad_hid) return "halo entry with a wrong shared id";
  for (int s = 0; s < t_end; ++s) {
    const int32_t id = p.sid[(size_t)s];
    if (id >= (int32_t)p.shared.size() || (id >= 0 && p.shared[(size_t)id] != s)) return "stress slot with a wrong shared id";
  }
  return "";
}

// ---------------------------------------------------------------------------
// two subcycles per workgroup
// ---------------------------------------------------------------------------
namespace {
constexpr int32_t kConstTag = 0x40000000;

std::vector<int32_t> inverse_perm(const TilePlan &p, int na) {
  std::vector<int32_t> inv((size_t)na);
  par_chunks(na, 1 << 16, [&](int64_t lo, int64_t hi, int) {
    for (int64_t s = lo; s < hi; ++s) inv[(size_t)p.perm[(size_t)s]] = (int32_t)s;
  });
  return inv;
}

struct PairWork {
  std::vector<int32_t> x, k, e, tn, fx;
  std::vector<uint16_t> uloc, vloc;
  int ns = 0, ne_s = 0, nu = 0;
  std::string err;
};
}  // namespace

// ghost slot -> copy-list source slot (-1: zero), or -2 when not listed
std::vector<int32_t> halo_map(const HaloList &halo, int na, int navel) {
  std::vector<int32_t> m((size_t)std::max(navel - na, 0), -2);
  for (int k = 0; k < halo.n; ++k) {
    const int32_t d = halo.dst[k];
    if (d >= na && d < navel) m[(size_t)(d - na)] = halo.src[k] < 0 ? -1 : halo.src[k];
  }
  return m;
}

PairPlan plan_pairs(const TilePlan &p, int na, int navel, const int32_t *const idx[6], const uint8_t *skip,
                    int cap_s, int cap_v, int max_t2, const HaloList &halo) {
  (void)skip;
  PairPlan r;
  if (!p.ok || !idx) {
    r.why = p.ok ? "neighbour lists not set" : "no tiling (" + p.why + ")";
    return r;
  }
  const int t_end = p.cnt[0] + p.cnt[1], u_begin = p.cnt[0], u_end = t_end + p.cnt[2];
  const std::vector<int32_t> inv = inverse_perm(p, na);
  const std::vector<int32_t> hmap = halo_map(halo, na, navel);
  auto slot = [&](int32_t q) { return q < na ? inv[(size_t)q] : q; };
  const int nt = (int)p.tiles.size();
  std::vector<PairWork> work((size_t)nt);
  std::vector<SlotMap> smaps, vmaps;
  for (int w = 0; w < plan_threads(); ++w) {
    smaps.emplace_back();
    vmaps.emplace_back();
  }
  par_items(nt, [&](int t, int w) {
    const TileDesc &d = p.tiles[(size_t)t];
    PairWork &wk = work[(size_t)t];
    SlotMap &sm = smaps[(size_t)w], &vm = vmaps[(size_t)w];
    const int nA = d.a1 - d.a0, nB = d.b1 - d.b0, nO = nA + nB, nS = nO + d.nh, nUo = nB + (d.c1 - d.c0);
    auto tslot = [&](int k) { return k < nA ? d.a0 + k : k < nO ? d.b0 + (k - nA) : p.list[(size_t)d.h0 + (k - nO)]; };
    bool added;
    for (int k = 0; k < nS; ++k) sm.find_or_add(tslot(k), k, added);
    for (int k = 0; k < nUo; ++k) vm.find_or_add(k < nB ? d.b0 + k : d.c0 + (k - nB), k, added);
    // level-2 stresses gather cc, ee, ne, se: active U points come from the
    // velocity table (U1-only entries in first-visit order), others from
    // memory; a ghost on the copy list reads its source as the copy made
    // after level 1 leaves it (the source's level-1 velocity, its unchanging
    // velocity, or zero)
    wk.vloc.resize((size_t)4 * nS);
    for (int k = 0; k < nS; ++k) {
      const int32_t g = tslot(k), q = p.perm[(size_t)g];
      const int32_t src[4] = {g, slot(idx[0][q]), slot(idx[1][q]), slot(idx[2][q])};
      for (int c = 0; c < 4; ++c) {
        uint16_t en = PairPlan::kNoEntry;
        int32_t s = src[c];
        const bool ghost = s >= na && hmap[(size_t)(s - na)] != -2;
        if (ghost) s = hmap[(size_t)(s - na)];
        if (s >= u_begin && s < u_end) {
          const int v = vm.find_or_add(s, nUo + (int)wk.e.size(), added);
          if (added) wk.e.push_back(s);
          en = (uint16_t)std::min(v, 0xfffe);
        } else if (ghost) {
          int j = 0;
          while (j < (int)wk.fx.size() && wk.fx[(size_t)j] != s) ++j;
          if (j == (int)wk.fx.size()) wk.fx.push_back(s);
          if (j >= 0xffff - PairPlan::kFixBase && wk.err.empty())
            wk.err = "tile " + std::to_string(t) + " has too many copy-list sources";
          en = (uint16_t)(PairPlan::kFixBase + std::min(j, 0xffff - PairPlan::kFixBase - 1));
        }
        wk.vloc[(size_t)4 * k + c] = en;
      }
    }
    // level-1 stepu points gather own, nw, sse, sw: active T points outside
    // T2 become T1-only entries, the rest constant entries
    const int nU = nUo + (int)wk.e.size();
    std::vector<int32_t> tag((size_t)4 * nU);
    for (int k = 0; k < nU; ++k) {
      const int32_t s = k < nB ? d.b0 + k : k < nUo ? d.c0 + (k - nB) : wk.e[(size_t)(k - nUo)];
      const int32_t q = p.perm[(size_t)s];
      const int32_t src[4] = {s, slot(idx[3][q]), slot(idx[5][q]), slot(idx[4][q])};
      for (int c = 0; c < 4; ++c) {
        int v;
        if (src[c] < t_end) {
          v = sm.find_or_add(src[c], nS + (int)wk.x.size(), added);
          if (added) wk.x.push_back(src[c]);
          if (k < nUo && v >= nS && wk.err.empty()) wk.err = "owned stepu point reads a T point outside its halo";
        } else {
          v = sm.find_or_add(src[c], kConstTag | (int)wk.k.size(), added);
          if (added) wk.k.push_back(src[c]);
        }
        tag[(size_t)4 * k + c] = v;
      }
    }
    const int nT = nS + (int)wk.x.size(), nE = nT + (int)wk.k.size();
    wk.tn.resize((size_t)4 * nT);
    for (int k = 0; k < nT; ++k) {
      const int32_t g = k < nS ? tslot(k) : wk.x[(size_t)(k - nS)], q = p.perm[(size_t)g];
      wk.tn[(size_t)4 * k] = g;
      for (int c = 0; c < 3; ++c) wk.tn[(size_t)4 * k + 1 + c] = slot(idx[c][q]);
    }
    wk.ns = nS;
    wk.ne_s = nE;
    wk.nu = nU;
    if (wk.err.empty() && nS > max_t2)
      wk.err = "tile " + std::to_string(t) + " has " + std::to_string(nS) + " level-2 stress points (limit " +
               std::to_string(max_t2) + ")";
    if (wk.err.empty() && (nE > cap_s || nE > 0xfffe))
      wk.err = "tile " + std::to_string(t) + " needs " + std::to_string(nE) + " str entries (limit " +
               std::to_string(cap_s) + ")";
    if (wk.err.empty() && (nU > cap_v || nU >= PairPlan::kFixBase))
      wk.err = "tile " + std::to_string(t) + " needs " + std::to_string(nU) + " velocity entries (limit " +
               std::to_string(cap_v) + ")";
    if (wk.err.empty()) {
      wk.uloc.resize(tag.size());
      for (size_t j = 0; j < tag.size(); ++j)
        wk.uloc[j] = (uint16_t)((tag[j] & kConstTag) ? nT + (tag[j] & (kConstTag - 1)) : tag[j]);
    }
    sm.clear();
    vm.clear();
  });
  for (int t = 0; t < nt; ++t) {
    const PairWork &wk = work[(size_t)t];
    if (!wk.err.empty()) {
      r.why = wk.err;
      return r;
    }
  }
  r.tiles.resize((size_t)nt);
  size_t nx = 0, nk = 0, ne = 0, nu = 0, nv = 0, nn = 0, nf = 0;
  for (int t = 0; t < nt; ++t) {
    const PairWork &wk = work[(size_t)t];
    PairTile &pt = r.tiles[(size_t)t];
    pt.t = p.tiles[(size_t)t];
    pt.x0 = (int32_t)nx;
    pt.nx = (int32_t)wk.x.size();
    pt.q0 = (int32_t)nk;
    pt.nq = (int32_t)wk.k.size();
    pt.e0 = (int32_t)ne;
    pt.ne = (int32_t)wk.e.size();
    pt.u0 = (int32_t)nu;
    pt.v0 = (int32_t)nv;
    pt.n0 = (int32_t)nn;
    nn += wk.tn.size() / 4;
    pt.f0 = (int32_t)nf;
    pt.nf = (int32_t)wk.fx.size();
    nf += wk.fx.size();
    nx += wk.x.size();
    nk += wk.k.size();
    ne += wk.e.size();
    nu += wk.uloc.size() / 4;
    nv += wk.vloc.size() / 4;
    r.cap_s = std::max(r.cap_s, wk.ne_s);
    r.cap_v = std::max(r.cap_v, wk.nu);
    r.max_t2 = std::max(r.max_t2, wk.ns);
  }
  if (nu > 0x7fffffff || nv > 0x7fffffff || nn > 0x7fffffff) {
    r.why = "tables too large";
    r.tiles.clear();
    return r;
  }
  r.xlist.resize(nx);
  r.klist.resize(nk);
  r.elist.resize(ne);
  r.uloc.resize(4 * nu);
  r.vloc.resize(4 * nv);
  r.tnbr.resize(4 * nn);
  r.fix.resize(nf);
  par_items(nt, [&](int t, int) {
    const PairWork &wk = work[(size_t)t];
    const PairTile &pt = r.tiles[(size_t)t];
    std::copy(wk.x.begin(), wk.x.end(), r.xlist.begin() + pt.x0);
    std::copy(wk.k.begin(), wk.k.end(), r.klist.begin() + pt.q0);
    std::copy(wk.e.begin(), wk.e.end(), r.elist.begin() + pt.e0);
    std::copy(wk.uloc.begin(), wk.uloc.end(), r.uloc.begin() + (size_t)4 * pt.u0);
    std::copy(wk.vloc.begin(), wk.vloc.end(), r.vloc.begin() + (size_t)4 * pt.v0);
    std::copy(wk.tn.begin(), wk.tn.end(), r.tnbr.begin() + (size_t)4 * pt.n0);
    std::copy(wk.fx.begin(), wk.fx.end(), r.fix.begin() + pt.f0);
  });
  r.ok = true;
  return r;
}

std::string check_pairs(const PairPlan &q, const TilePlan &p, int na, int navel, const int32_t *const idx[6],
                        const uint8_t *skip, const HaloList &halo) {
  (void)skip;
  if (!q.ok) return "";
  if (!p.ok || !idx) return "pair plan without a tiling";
  if (q.tiles.size() != p.tiles.size()) return "pair plan and tiling differ in tiles";
  const int t_end = p.cnt[0] + p.cnt[1], u_begin = p.cnt[0], u_end = t_end + p.cnt[2];
  const std::vector<int32_t> inv = inverse_perm(p, na);
  const std::vector<int32_t> hmap = halo_map(halo, na, navel);
  auto slot = [&](int32_t r) { return r < na ? inv[(size_t)r] : r; };
  std::vector<std::string> err(q.tiles.size());
  std::vector<std::vector<int32_t>> smaps((size_t)plan_threads()), vmaps((size_t)plan_threads());
  par_items((int)q.tiles.size(), [&](int t, int w) {
    const PairTile &pt = q.tiles[(size_t)t];
    const TileDesc &d = pt.t, &o = p.tiles[(size_t)t];
    std::string &e = err[(size_t)t];
    if (d.a0 != o.a0 || d.a1 != o.a1 || d.b0 != o.b0 || d.b1 != o.b1 || d.c0 != o.c0 || d.c1 != o.c1 ||
        d.h0 != o.h0 || d.nh != o.nh) {
      e = "pair tile differs from its tile";
      return;
    }
    const int nA = d.a1 - d.a0, nB = d.b1 - d.b0, nO = nA + nB, nS = nO + d.nh, nUo = nB + (d.c1 - d.c0);
    const int nT = nS + pt.nx, nE = nT + pt.nq, nU = nUo + pt.ne;
    if (pt.nx < 0 || pt.x0 < 0 || (size_t)pt.x0 + pt.nx > q.xlist.size()) { e = "T1-only list range"; return; }
    if (pt.nq < 0 || pt.q0 < 0 || (size_t)pt.q0 + pt.nq > q.klist.size()) { e = "constant list range"; return; }
    if (pt.ne < 0 || pt.e0 < 0 || (size_t)pt.e0 + pt.ne > q.elist.size()) { e = "U1-only list range"; return; }
    if (pt.u0 < 0 || (size_t)4 * ((size_t)pt.u0 + nU) > q.uloc.size()) { e = "uloc range"; return; }
    if (pt.v0 < 0 || (size_t)4 * ((size_t)pt.v0 + nS) > q.vloc.size()) { e = "vloc range"; return; }
    if (pt.n0 < 0 || (size_t)4 * ((size_t)pt.n0 + nT) > q.tnbr.size()) { e = "tnbr range"; return; }
    if (nS > q.max_t2 || nE > q.cap_s || nU > q.cap_v) { e = "tile exceeds the plan's limits"; return; }
    std::vector<int32_t> &sm = smaps[(size_t)w], &vm = vmaps[(size_t)w];
    sm.assign((size_t)nE, -1);
    vm.assign((size_t)nU, -1);
    for (int k = 0; k < nA; ++k) sm[(size_t)k] = d.a0 + k;
    for (int k = 0; k < nB; ++k) sm[(size_t)(nA + k)] = d.b0 + k;
    for (int k = 0; k < d.nh; ++k) sm[(size_t)(nO + k)] = p.list[(size_t)d.h0 + k];
    for (int k = 0; k < pt.nx; ++k) {
      const int32_t g = q.xlist[(size_t)pt.x0 + k];
      if (g < 0 || g >= t_end) { e = "T1-only entry is not an active T slot"; return; }
      sm[(size_t)(nS + k)] = g;
    }
    for (int k = 0; k < pt.nq; ++k) {
      const int32_t g = q.klist[(size_t)pt.q0 + k];
      if (g < t_end || g >= navel) { e = "constant entry is an active T slot or out of range"; return; }
      sm[(size_t)(nT + k)] = g;
    }
    for (int k = 0; k < nB; ++k) vm[(size_t)k] = d.b0 + k;
    for (int k = 0; k < d.c1 - d.c0; ++k) vm[(size_t)(nB + k)] = d.c0 + k;
    for (int k = 0; k < pt.ne; ++k) {
      const int32_t g = q.elist[(size_t)pt.e0 + k];
      if (g < u_begin || g >= u_end) { e = "U1-only entry is not an active U slot"; return; }
      vm[(size_t)(nUo + k)] = g;
    }
    // every level-1 stepu read names its source's entry; the owned points'
    // (read again at level 2) never name a T1-only entry
    for (int k = 0; k < nU; ++k) {
      const int32_t s = vm[(size_t)k], r0 = p.perm[(size_t)s];
      const int32_t want[4] = {s, slot(idx[3][r0]), slot(idx[5][r0]), slot(idx[4][r0])};
      for (int c = 0; c < 4; ++c) {
        const int en = q.uloc[(size_t)4 * (pt.u0 + k) + c];
        if (en >= nE || sm[(size_t)en] != want[c] || (k < nUo && en >= nS && en < nT)) {
          e = "stepu entry " + std::to_string(k) + " of tile " + std::to_string(t) + ": bad str entry";
          return;
        }
      }
    }
    // every level-1 stress entry's slot and ee, ne, se neighbour slots
    for (int k = 0; k < nT; ++k) {
      const int32_t g = sm[(size_t)k], r0 = p.perm[(size_t)g];
      const int32_t *tn = &q.tnbr[(size_t)4 * (pt.n0 + k)];
      if (tn[0] != g || tn[1] != slot(idx[0][r0]) || tn[2] != slot(idx[1][r0]) || tn[3] != slot(idx[2][r0])) {
        e = "stress entry " + std::to_string(k) + " of tile " + std::to_string(t) + ": bad neighbour slots";
        return;
      }
    }
    // every level-2 stress read names the level-1 velocity of an active U
    // point, and only those; a copy-list ghost is read as its source
    if (pt.nf < 0 || pt.f0 < 0 || (size_t)pt.f0 + pt.nf > q.fix.size()) { e = "fixed-source range"; return; }
    for (int k = 0; k < nS; ++k) {
      const int32_t g = sm[(size_t)k], r0 = p.perm[(size_t)g];
      const int32_t want[4] = {g, slot(idx[0][r0]), slot(idx[1][r0]), slot(idx[2][r0])};
      for (int c = 0; c < 4; ++c) {
        const int en = q.vloc[(size_t)4 * (pt.v0 + k) + c];
        const bool ghost = want[c] >= na && hmap[(size_t)(want[c] - na)] != -2;
        const int32_t s = ghost ? hmap[(size_t)(want[c] - na)] : want[c];
        const bool active = s >= u_begin && s < u_end;
        bool ok;
        if (en == PairPlan::kNoEntry)
          ok = !active && !ghost;
        else if (en >= PairPlan::kFixBase)
          ok = ghost && !active && en - PairPlan::kFixBase < pt.nf && q.fix[(size_t)pt.f0 + (en - PairPlan::kFixBase)] == s;
        else
          ok = active && en < nU && vm[(size_t)en] == s;
        if (!ok) {
          e = "stress entry " + std::to_string(k) + " of tile " + std::to_string(t) + ": bad velocity entry";
          return;
        }
      }
    }
  });
  for (const std::string &e : err)
    if (!e.empty()) return e;
  return "";
}

}  // namespace evp1d

#include "../../include/evp1d.h"

extern "C" int evp1d_tile_plan(int32_t na, int32_t navel, const int32_t *ee, const int32_t *ne, const int32_t *se,
                               const int32_t *nw, const int32_t *sw, const int32_t *sse, const int32_t *skipT,
                               const int32_t *skipU, int32_t tile_w, int32_t tile_h, int32_t max_entries,
                               int32_t *perm, int32_t *counts, int32_t *ntiles, int32_t *max_used,
                               int32_t *tiles, int32_t tiles_cap, int32_t *list, int32_t list_cap, int32_t *nlist,
                               int32_t *klist, int32_t klist_cap, int32_t *nklist, int32_t *uloc) {
  return evp1d::no_throw("evp1d_tile_plan", [&]() -> int {
  if (na < 0 || navel < na) return evp1d__set_error(EVP1D_EINVAL, "evp1d_tile_plan: need 0 <= na <= navel");
  const int32_t *in[6] = {ee, ne, se, nw, sw, sse};
  for (int k = 0; k < 6; ++k)
    if (na > 0 && !in[k]) return evp1d__set_error(EVP1D_EINVAL, "evp1d_tile_plan: NULL neighbour list");
  if (na > 0 && (!skipT || !skipU)) return evp1d__set_error(EVP1D_EINVAL, "evp1d_tile_plan: NULL flags");
  std::vector<int32_t> z[6];
  const int32_t *zp[6];
  for (int k = 0; k < 6; ++k) {
    z[k].resize((size_t)na);
    for (int q = 0; q < na; ++q) {
      const int32_t v = in[k][q];
      if (v < 1 || v > navel) return evp1d__set_error(EVP1D_EINDEX, "evp1d_tile_plan: neighbour index out of range");
      z[k][(size_t)q] = v - 1;
    }
    zp[k] = z[k].data();
  }
  std::vector<uint8_t> skip((size_t)na);
  for (int q = 0; q < na; ++q) skip[(size_t)q] = (uint8_t)((skipT[q] ? 1 : 0) | (skipU[q] ? 2 : 0));
  const evp1d::TilePlan p = evp1d::plan_tiles(na, navel, zp, skip.data(), tile_w, tile_h, max_entries);
  const std::string bad = evp1d::check_plan(p, na, navel, zp, skip.data());
  if (!bad.empty()) return evp1d__set_error(EVP1D_EINDEX, ("evp1d_tile_plan: " + bad).c_str());
  if (perm)
    for (int s = 0; s < na; ++s) perm[s] = p.perm[(size_t)s] + 1;
  if (counts)
    for (int c = 0; c < 4; ++c) counts[c] = p.cnt[c];
  const int nt = p.ok ? (int)p.tiles.size() : 0;
  if (ntiles) *ntiles = nt;
  if (max_used) *max_used = p.max_entries;
  if (nlist) *nlist = (int32_t)p.list.size();
  if (nklist) *nklist = (int32_t)p.klist.size();
  if (tiles && tiles_cap >= nt)
    for (int t = 0; t < nt; ++t) {
      const evp1d::TileDesc &d = p.tiles[(size_t)t];
      const int32_t v[14] = {d.a0, d.a1, d.b0, d.b1, d.c0, d.c1, d.h0, d.nh, d.k0, d.nk, d.sa, d.ia, d.sb, d.ib};
      for (int k = 0; k < 14; ++k) tiles[14 * t + k] = v[k];
    }
  if (list && list_cap >= (int32_t)p.list.size())
    for (size_t k = 0; k < p.list.size(); ++k) list[k] = p.list[k];
  if (klist && klist_cap >= (int32_t)p.klist.size())
    for (size_t k = 0; k < p.klist.size(); ++k) klist[k] = p.klist[k];
  if (uloc && p.ok)
    for (size_t k = 0; k < p.uloc.size(); ++k) uloc[k] = p.uloc[k];
  return EVP1D_OK;
  });
}
