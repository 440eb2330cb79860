// evp1d_index.cpp - 1D point list and neighbour positions from a 2D T mask.
//
// Produces exactly what the reference's init path produces
// (integration/ice_dyn_evp1d.F90):
//   na      = count of wet T cells over the whole block       (calc_nActiveTU 494-519)
//   1D order: j outer, i inner, from 1+nghost                 (calc_2d_indices_init 559-589)
//   linear index of (i,j) = i + (j-1)*nx, and of the six neighbours
//     ee=(i-1,j) ne=(i-1,j-1) se=(i,j-1) nw=(i+1,j) sw=(i+1,j+1) sse=(i,j+1)   (766-772)
//   navel   = |union of the seven linear-index lists|          (calc_navel 1047-1087)
//   ghost slots na+1..navel = sorted (union \ target points)   (setdiff 786-790)
//   ee..sse = 1-based slot of each neighbour's linear index    (findXinY 792-797)
// The linear-index arithmetic is kept as the reference has it, including the
// wrap of i+1 at the east rim into the next row: every 1D slot corresponds to
// the same linear index as in the reference.
//
// Build decision: the seven lists are merged with one sort of their
// concatenation and the positions found by a single forward walk per list
// (each list is sorted because the target list is), instead of six pairwise
// unions.  The result is the same set in the same order.
#include <algorithm>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/evp1d.h"
#include "evp1d_tiles.hpp"

// error channel shared with the device translation unit
extern "C" int evp1d__set_error(int code, const char *msg);

namespace {

int ifail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return evp1d__set_error(code, buf);
}

struct Lists {
  std::vector<int64_t> in, nb[6];   // target and neighbour linear indices
  std::vector<int32_t> ti, tj;
};

int collect(int32_t nx, int32_t ny, int32_t nghost, const int32_t *tmask, Lists &L, int64_t &na_all) {
  if (nx < 1 || ny < 1 || nghost < 0) return ifail(EVP1D_EINVAL, "evp1d_index: bad sizes nx=%d ny=%d nghost=%d", nx, ny, nghost);
  if (!tmask) return ifail(EVP1D_EINVAL, "evp1d_index: tmask is NULL");
  na_all = 0;
  for (int64_t k = 0; k < (int64_t)nx * ny; ++k) na_all += tmask[k] != 0;
  for (int32_t j = 1 + nghost; j <= ny; ++j)
    for (int32_t i = 1 + nghost; i <= nx; ++i)
      if (tmask[(int64_t)(i - 1) + (int64_t)(j - 1) * nx]) {
        L.ti.push_back(i);
        L.tj.push_back(j);
      }
  if ((int64_t)L.ti.size() != na_all)
    return ifail(EVP1D_EINVAL,
                 "evp1d_index: %lld wet T cells lie in the west/south rim (i or j <= nghost=%d); "
                 "the 1D list only covers i,j > nghost",
                 (long long)(na_all - (int64_t)L.ti.size()), nghost);
  const size_t na = L.ti.size();
  L.in.resize(na);
  for (int k = 0; k < 6; ++k) L.nb[k].resize(na);
  for (size_t w = 0; w < na; ++w) {
    const int64_t i = L.ti[w], j = L.tj[w];
    L.in[w] = i + (j - 1) * nx;
    L.nb[0][w] = i - 1 + (j - 1) * nx;   // ee
    L.nb[1][w] = i - 1 + (j - 2) * nx;   // ne
    L.nb[2][w] = i + (j - 2) * nx;       // se
    L.nb[3][w] = i + 1 + (j - 1) * nx;   // nw
    L.nb[4][w] = i + 1 + (j - 0) * nx;   // sw
    L.nb[5][w] = i + (j - 0) * nx;       // sse
  }
  return EVP1D_OK;
}

std::vector<int64_t> ghosts_of(const Lists &L) {
  std::vector<int64_t> all;
  all.reserve(L.in.size() * 6);
  for (int k = 0; k < 6; ++k) all.insert(all.end(), L.nb[k].begin(), L.nb[k].end());
  std::sort(all.begin(), all.end());
  all.erase(std::unique(all.begin(), all.end()), all.end());
  std::vector<int64_t> g;
  g.reserve(all.size());
  std::set_difference(all.begin(), all.end(), L.in.begin(), L.in.end(), std::back_inserter(g));
  return g;
}

}  // namespace

extern "C" {

int evp1d_index_count(int32_t nx, int32_t ny, int32_t nghost, const int32_t *tmask, int32_t *na,
                      int32_t *navel) {
  return evp1d::no_throw("evp1d_index_count", [&]() -> int {
  Lists L;
  int64_t na_all = 0;
  int rc = collect(nx, ny, nghost, tmask, L, na_all);
  if (rc) return rc;
  const std::vector<int64_t> g = ghosts_of(L);
  if (na) *na = (int32_t)L.in.size();
  if (navel) *navel = (int32_t)(L.in.size() + g.size());
  return EVP1D_OK;
  });
}

int evp1d_index_build(int32_t nx, int32_t ny, int32_t nghost, const int32_t *tmask, int32_t na,
                      int32_t navel, int32_t *indxti, int32_t *indxtj, int32_t *indxtij, int32_t *ee,
                      int32_t *ne, int32_t *se, int32_t *nw, int32_t *sw, int32_t *sse) {
  return evp1d::no_throw("evp1d_index_build", [&]() -> int {
  Lists L;
  int64_t na_all = 0;
  int rc = collect(nx, ny, nghost, tmask, L, na_all);
  if (rc) return rc;
  const std::vector<int64_t> g = ghosts_of(L);
  if ((int64_t)na != (int64_t)L.in.size() || (int64_t)navel != (int64_t)(L.in.size() + g.size()))
    return ifail(EVP1D_EINVAL, "evp1d_index_build: sizes (%d, %d) differ from evp1d_index_count (%lld, %lld)", na,
                 navel, (long long)L.in.size(), (long long)(L.in.size() + g.size()));
  int32_t *out[6] = {ee, ne, se, nw, sw, sse};
  for (int k = 0; k < 6; ++k)
    if (na > 0 && !out[k]) return ifail(EVP1D_EINVAL, "evp1d_index_build: neighbour output %d is NULL", k);
  if (na > 0 && (!indxti || !indxtj)) return ifail(EVP1D_EINVAL, "evp1d_index_build: indxti/indxtj NULL");
  for (int32_t w = 0; w < na; ++w) {
    indxti[w] = L.ti[(size_t)w];
    indxtj[w] = L.tj[(size_t)w];
  }
  if (indxtij) {
    for (int32_t w = 0; w < na; ++w) indxtij[w] = (int32_t)L.in[(size_t)w];
    for (size_t q = 0; q < g.size(); ++q) indxtij[(size_t)na + q] = (int32_t)g[q];
  }
  // forward walk: each neighbour list is sorted, so both cursors only advance
  for (int k = 0; k < 6; ++k) {
    size_t p1 = 0, p2 = 0;
    for (int32_t w = 0; w < na; ++w) {
      const int64_t x = L.nb[k][(size_t)w];
      while (p1 < L.in.size() && L.in[p1] < x) ++p1;
      while (p2 < g.size() && g[p2] < x) ++p2;
      if (p1 < L.in.size() && L.in[p1] == x)
        out[k][w] = (int32_t)(p1 + 1);
      else if (p2 < g.size() && g[p2] == x)
        out[k][w] = (int32_t)(na + p2 + 1);
      else
        return ifail(EVP1D_EINDEX, "evp1d_index_build: neighbour %d of point %d not found", k, w + 1);
    }
  }
  return EVP1D_OK;
  });
}

}  // extern "C"
