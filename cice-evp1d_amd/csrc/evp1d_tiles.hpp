// Tiling of the 1D point list for the fused subcycle kernel (host only).
//
// One subcycle is stress on every active T point, then stepu on every active
// U point, where stepu at q reads str1..8 at q and at its neighbours nw(q),
// sse(q), sw(q) (bench_v2b.F90:526-531).  The fused kernel runs both halves
// for one tile in one workgroup: it evaluates stress for the tile's own T
// points plus the "halo" T points its U points reference, keeps str1..8 in
// LDS, synchronises the workgroup, then runs stepu for the tile's U points.
// Halo stresses are evaluated twice (by their owner and by the neighbour
// tile) but written only by their owner, so every output is produced once.
//
// Slot order.  Points are ordered by class (stress only | stress and stepu |
// stepu only | neither), then by tile, then shared points (those some other
// tile evaluates as halo) before the rest, then by caller position, so a
// tile's owned T points and its U points are each two contiguous slot ranges
// and the class ranges that the unfused kernels iterate stay contiguous.  Tiles
// are W x H boxes of grid coordinates recovered from the neighbour lists:
// ee=(i-1,j) ne=(i-1,j-1) se=(i,j-1) nw=(i+1,j) sw=(i+1,j+1) sse=(i,j+1)
// (ice_dyn_evp1d.F90:766-772).  Connected components get disjoint frames.
#pragma once
#include <cstdint>
#include <cstdio>
#include <exception>
#include <memory>
#include <new>
#include <string>
#include <utility>
#include <vector>

extern "C" int evp1d__set_error(int code, const char *msg);

namespace evp1d {

// Runs f() (which returns a status code) so that no C++ exception crosses the
// C ABI: a host allocation failure becomes EVP1D_ENOMEM (4), any other
// exception EVP1D_EINVAL (1), with the message kept for evp1d_last_error.
// The message is formatted into a fixed buffer (nothing is allocated).
template <class F>
int no_throw(const char *who, F &&f) {
  char buf[256];
  try {
    return f();
  } catch (const std::bad_alloc &) {
    std::snprintf(buf, sizeof buf, "%s: host memory allocation failed", who);
    return evp1d__set_error(4, buf);
  } catch (const std::exception &e) {
    std::snprintf(buf, sizeof buf, "%s: %s", who, e.what());
    return evp1d__set_error(1, buf);
  } catch (...) {
    std::snprintf(buf, sizeof buf, "%s: unexpected exception", who);
    return evp1d__set_error(1, buf);
  }
}

// std::vector whose resize leaves new elements uninitialised, for the large
// planner tables that are written in full (in parallel) before any read: the
// value-initialisation of std::vector would be a serial pass over them.
template <class T>
struct UninitAlloc : std::allocator<T> {
  template <class U>
  struct rebind {
    using other = UninitAlloc<U>;
  };
  UninitAlloc() = default;
  template <class U>
  UninitAlloc(const UninitAlloc<U> &) {}
  template <class U>
  void construct(U *p) noexcept {
    ::new (static_cast<void *>(p)) U;
  }
  template <class U, class... A>
  void construct(U *p, A &&...a) {
    ::new (static_cast<void *>(p)) U(std::forward<A>(a)...);
  }
};
template <class T>
using Table = std::vector<T, UninitAlloc<T>>;


// Slot ranges of one tile; entries of its LDS table are numbered
// [owned stress-only | owned stress-and-stepu | halo stress | constants].
struct TileDesc {
  int32_t a0, a1;   // owned stress-only slots
  int32_t b0, b1;   // owned stress-and-stepu slots
  int32_t c0, c1;   // stepu-only slots
  int32_t h0, nh;   // halo stress slots: TilePlan::list[h0, h0 + nh)
  int32_t k0, nk;   // constant slots (str1..8 never written): TilePlan::klist[k0, k0 + nk)
  int32_t sa, ia;   // [a0, a0 + sa) are shared, with shared ids ia, ia + 1, ...
  int32_t sb, ib;   // [b0, b0 + sb) are shared, with shared ids ib, ib + 1, ...
};

struct TilePlan {
  bool ok = false;          // false: no tiling (perm is still the class order)
  std::string why;          // reason when !ok
  int tile_w = 0, tile_h = 0;
  int cnt[4] = {0, 0, 0, 0};          // class sizes
  std::vector<int32_t> perm;          // slot -> caller position, slots [0, na)
  std::vector<TileDesc> tiles;        // tiles with work, in launch order
  std::vector<int32_t> list;          // halo stress slots, per tile
  std::vector<int32_t> klist;         // constant slots, per tile
  Table<uint16_t> uloc;               // 4 per stepu slot (slot - cnt[0]): own, nw, sse, sw entries
  // Stresses of shared points (halo of some tile) live, during a fused run,
  // in a double-buffered copy indexed by a shared id: every reader sees the
  // previous subcycle's values while their owner writes the next ones.
  Table<int32_t> shared;              // shared id -> slot (sorted slots of all halo entries)
  Table<int32_t> sid;                 // stress slot -> shared id, or -1
  std::vector<int32_t> hid;           // per halo entry (parallel to list): shared id
  int max_entries = 0;
};

// Grid coordinates of every point, recovered from the neighbour lists (one
// frame per connected component, components side by side).  ok = false (and
// `why` set) when the lists are not those of a grid.
struct GridCoords {
  bool ok = false;
  std::string why;
  std::vector<int32_t> gx, gy;
  int64_t width = 0, height = 0;
};
GridCoords grid_coords(int na, int navel, const int32_t *const idx[6]);

// Tile of every point: depends on the neighbour lists and the tile size only,
// so a caller can keep it while skip flags change.  ntiles = 0 (and `why`
// set) when the lists are not those of a grid.
struct TileIds {
  int tile_w = 0, tile_h = 0;
  int order = 0;              // tile launch order: 0 rows of tiles, 1 Hilbert curve over the tiles
  int ntiles = 0;
  std::vector<int32_t> tid;   // caller position -> tile, in launch order
  std::string why;
};
TileIds tile_ids(const GridCoords &c, int tile_w, int tile_h, int order = 0);
TileIds tile_ids(int na, int navel, const int32_t *const idx[6], int tile_w, int tile_h);

// idx: ee, ne, se, nw, sw, sse as 0-based caller positions (values in
// [0, navel); positions >= na are ghost slots).  skip: bit0 skipT, bit1 skipU.
// idx may be null (no tiling, class order only).
TilePlan plan_tiles(int na, int navel, const int32_t *const idx[6], const uint8_t *skip, int tile_w, int tile_h,
                    int max_entries);
// The same from tile ids made by tile_ids for these lists.
TilePlan plan_tiles(int na, int navel, const int32_t *const idx[6], const uint8_t *skip, const TileIds &ids,
                    int max_entries);

// Re-check every bound the fused kernel relies on; returns "" when the plan
// is consistent, else the first violation.
std::string check_plan(const TilePlan &p, int na, int navel, const int32_t *const idx[6], const uint8_t *skip);

// Two subcycles per workgroup (the "pair" form), on the slot order and
// tiles of a TilePlan.  Level 2 is the one-subcycle tile: stress on the
// owned and halo T points ("T2"), stepu on the owned U points.  Level 1
// produces what level 2 reads: stepu on every active U point whose velocity
// a T2 stress gathers (cc, ee, ne, se) plus the owned U points ("U1"), and
// stress on every active T point whose str1..8 a U1 stepu gathers ("T1",
// T2 first).  Level 1 reads the input state only, level 2 reads level 1's
// results from LDS, so the tile writes the state two subcycles on.
//
// Entries of a tile's str table: [owned T | halo T | T1-only | constants];
// of its velocity table: [owned U (B then C) | U1-only].
struct PairTile {
  TileDesc t;               // owned ranges and halo list (h0, nh) of the TilePlan tile
  int32_t x0, nx;           // T1-only stress slots: PairPlan::xlist[x0, x0 + nx)
  int32_t q0, nq;           // constant slots (str1..8 never written): PairPlan::klist[q0, q0 + nq)
  int32_t e0, ne;           // U1-only stepu slots: PairPlan::elist[e0, e0 + ne)
  int32_t u0, v0;           // first uloc entry (per U1 entry), first vloc entry (per T2 entry)
  int32_t n0;               // first tnbr entry (per stress entry of level 1, T2 entries first)
  int32_t f0, nf;           // fixed-source entries: PairPlan::fix[f0, f0 + nf)
};

struct PairPlan {
  bool ok = false;
  std::string why;
  std::vector<PairTile> tiles;        // parallel to TilePlan::tiles
  Table<int32_t> xlist, klist, elist;
  // 4 per U1 entry: str table entries of own, nw, sse, sw
  Table<uint16_t> uloc;
  // 4 per T2 entry: velocity table entries of cc, ee, ne, se; kNoEntry when
  // that velocity never changes (not an active U point: read from memory)
  Table<uint16_t> vloc;
  // 4 per level-1 stress entry (T2 then T1-only): its slot and the slots of
  // its ee, ne, se neighbours, so one load gives every address the entry's
  // stress gathers
  Table<int32_t> tnbr;
  // Ghost copy list (evp1d_set_halo): between the two subcycles of a launch
  // a listed ghost slot holds the copy of its source.  A level-2 read of such
  // a ghost names the source's level-1 velocity entry when the source is an
  // active U point; otherwise vloc holds kFixBase + j and fix[f0 + j] is the
  // source slot (a velocity no subcycle changes) or -1 (the copy is zero).
  Table<int32_t> fix;
  int cap_s = 0, cap_v = 0;           // largest str / velocity table
  int max_t2 = 0;                     // largest T2 (level-2 stresses, one per thread)
  static constexpr uint16_t kNoEntry = 0xffff;
  static constexpr uint16_t kFixBase = 0xc000;   // velocity table entries stay below it
};

// Ghost copy list in slot space: ghost slot dst[k] (>= na) takes the value
// of slot src[k] (< na) or zero (src[k] < 0) after every subcycle.
struct HaloList {
  int n = 0;
  const int32_t *dst = nullptr, *src = nullptr;
};

// cap_s, cap_v, max_t2: limits on the tables and on T2 (plan fails beyond).
PairPlan plan_pairs(const TilePlan &p, int na, int navel, const int32_t *const idx[6], const uint8_t *skip,
                    int cap_s, int cap_v, int max_t2, const HaloList &halo = HaloList());
// Re-derives every table of a pair plan from the lists; "" when consistent.
std::string check_pairs(const PairPlan &q, const TilePlan &p, int na, int navel, const int32_t *const idx[6],
                        const uint8_t *skip, const HaloList &halo = HaloList());

}  // namespace evp1d
