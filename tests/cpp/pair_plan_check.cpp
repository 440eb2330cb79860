// The two-subcycle ("pair") tables of evp1d_tiles.cpp on a real grid:
//  1. for several tile shapes the plan exists, passes check_pairs and stays
//     within the ring bounds (T2 <= (w+1)(h+1), str table <= (w+3)(h+3),
//     velocity table <= (w+2)(h+2));
//  2. a CPU model of the pair kernel's schedule, run through the tables with
//     integer hash "kernels" in place of stress and stepu, reproduces two
//     plain subcycles of the same hashes bit for bit (every owned output);
//  3. each corruption of a valid pair plan is reported by check_pairs.
// Steps 1 and 2 run with and without a ghost copy list (evp1d_set_halo):
// the reference then copies after each plain subcycle, the model after the
// launch, and level-2 reads of listed ghosts follow the plan's vloc.
// Input file: int32 na, navel; six int32[na] neighbour lists (0-based caller
// positions, ee ne se nw sw sse); uint8[na] skip flags (bit0 skipT, bit1 skipU).
#include "evp1d_tiles.hpp"
#include <cstdio>
#include <functional>
#include <vector>

using u64 = uint64_t;
static u64 mix(u64 a, u64 b) {
  u64 x = a * 0x9e3779b97f4a7c15ull ^ (b + 0x632be59bd9b4e019ull + (a << 6) + (a >> 2));
  x ^= x >> 31;
  return x * 0xbf58476d1ce4e5b9ull;
}
// toy stress at a T point: new state and its str from state and 4 velocities
static void tstress(u64 s, u64 cc, u64 ee, u64 ne, u64 se, u64 &s_out, u64 &str) {
  s_out = mix(mix(mix(mix(s, cc), ee), ne), se);
  str = mix(s_out, 17);
}
// toy stepu at a U point: new velocity from own velocity and 4 str
static u64 tstepu(u64 v, u64 a, u64 b, u64 c, u64 e) { return mix(mix(mix(mix(v, a), b), c), e); }

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  FILE *f = fopen(argv[1], "rb");
  if (!f) return 2;
  int hdr[2];
  if (fread(hdr, 4, 2, f) != 2) return 1;
  const int na = hdr[0], navel = hdr[1];
  std::vector<int32_t> idx[6];
  const int32_t *ip[6];
  for (int k = 0; k < 6; ++k) {
    idx[k].resize(na);
    if (fread(idx[k].data(), 4, na, f) != (size_t)na) return 1;
    ip[k] = idx[k].data();
  }
  std::vector<uint8_t> skip(na);
  if (fread(skip.data(), 1, na, f) != (size_t)na) return 1;
  fclose(f);
  int fails = 0;
  const int shapes[5][2] = {{32, 16}, {16, 16}, {64, 8}, {5, 7}, {1, 1}};
  for (const auto &sh : shapes) {
    const int W = sh[0], H = sh[1];
    const evp1d::TilePlan p = evp1d::plan_tiles(na, navel, ip, skip.data(), W, H, 2560);
    // slot space: classes, inverse permutation and slot-space lists
    const int t_end = p.cnt[0] + p.cnt[1], u_begin = p.cnt[0], u_end = t_end + p.cnt[2];
    std::vector<int32_t> inv(na);
    for (int s = 0; s < na; ++s) inv[p.perm[s]] = s;
    auto slot = [&](int32_t r) { return r < na ? inv[r] : r; };
    std::vector<int32_t> nb[6];
    for (int k = 0; k < 6; ++k) {
      nb[k].resize(na);
      for (int s = 0; s < na; ++s) nb[k][s] = slot(idx[k][p.perm[s]]);
    }
    // a ghost copy list over every third ghost slot: sources an active U
    // point, a point whose velocity never changes, or zero
    std::vector<int32_t> hdst, hsrc;
    for (int g = na, k = 0; g < navel; g += 3, ++k) {
      hdst.push_back(g);
      hsrc.push_back(k % 3 == 0 ? u_begin + (k * 7919) % std::max(u_end - u_begin, 1)
                                : k % 3 == 1 ? (k * 104729) % std::max(u_begin, 1) : -1);
    }
    for (int with_halo = 0; with_halo < 2; ++with_halo) {
      evp1d::HaloList hl;
      if (with_halo) {
        hl.n = (int)hdst.size();
        hl.dst = hdst.data();
        hl.src = hsrc.data();
      }
      const evp1d::PairPlan q = evp1d::plan_pairs(p, na, navel, ip, skip.data(), 4096, 4096, 1024, hl);
      const std::string chk = evp1d::check_pairs(q, p, na, navel, ip, skip.data(), hl);
      printf("tile %dx%d%s: ok=%d why='%s' check='%s' max_t2=%d cap_s=%d cap_v=%d fix=%zu\n", W, H,
             with_halo ? " with copy list" : "", (int)q.ok, q.why.c_str(), chk.c_str(), q.max_t2, q.cap_s, q.cap_v,
             q.fix.size());
      if (!q.ok || !chk.empty() ||
          (!with_halo && (q.max_t2 > (W + 1) * (H + 1) || q.cap_s > (W + 3) * (H + 3) || q.cap_v > (W + 2) * (H + 2)))) {
        ++fails;
        continue;
      }
      std::vector<u64> sig(navel), vel(navel), str(navel);
      for (int s = 0; s < navel; ++s) {
        sig[s] = mix(s, 1);
        vel[s] = mix(s, 2);
        str[s] = mix(s, 3);   // points stress never writes keep theirs
      }
      auto copy_list = [&](std::vector<u64> &v) {
        for (int k = 0; k < hl.n; ++k) v[hl.dst[k]] = hl.src[k] < 0 ? 0 : v[hl.src[k]];
      };
      // reference: two plain subcycles, each followed by the copy list
      std::vector<u64> rs = sig, rv = vel, rstr = str;
      for (int sub = 0; sub < 2; ++sub) {
        std::vector<u64> ns = rs, nv = rv;
        for (int s = 0; s < t_end; ++s) tstress(rs[s], rv[s], rv[nb[0][s]], rv[nb[1][s]], rv[nb[2][s]], ns[s], rstr[s]);
        for (int s = u_begin; s < u_end; ++s)
          nv[s] = tstepu(rv[s], rstr[s], rstr[nb[3][s]], rstr[nb[5][s]], rstr[nb[4][s]]);
        rs.swap(ns);
        rv.swap(nv);
        copy_list(rv);
      }
      // the pair kernel's schedule, tile by tile, from the input state only,
      // then the copy list after the launch
      std::vector<u64> os = sig, ov = vel;
      for (const evp1d::PairTile &pt : q.tiles) {
        const evp1d::TileDesc &t = pt.t;
        const int nA = t.a1 - t.a0, nB = t.b1 - t.b0, nO = nA + nB, nS = nO + t.nh, nT = nS + pt.nx, nE = nT + pt.nq;
        const int nUo = nB + (t.c1 - t.c0), nU = nUo + pt.ne;
        auto tslot = [&](int k) {
          return k < nA ? t.a0 + k : k < nO ? t.b0 + (k - nA) : k < nS ? p.list[t.h0 + (k - nO)] : q.xlist[pt.x0 + (k - nS)];
        };
        auto uslot = [&](int k) { return k < nB ? t.b0 + k : k < nUo ? t.c0 + (k - nB) : q.elist[pt.e0 + (k - nUo)]; };
        std::vector<u64> S(nE), sg(nS), V(nU);
        for (int k = 0; k < nE; ++k) {
          if (k < nT) {
            // as the kernel: the slot of owned entries by position, of the
            // others and every entry's ee, ne, se from the tnbr table
            const int32_t *tn = &q.tnbr[4 * ((size_t)pt.n0 + k)];
            const int g = k < nO ? tslot(k) : tn[0];
            u64 s1;
            tstress(sig[g], vel[g], vel[tn[1]], vel[tn[2]], vel[tn[3]], s1, S[k]);
            if (k < nS) sg[k] = s1;
          } else {
            S[k] = str[q.klist[pt.q0 + (k - nT)]];
          }
        }
        auto ustep = [&](int k, u64 v) {
          const uint16_t *w = &q.uloc[4 * ((size_t)pt.u0 + k)];
          return tstepu(v, S[w[0]], S[w[1]], S[w[2]], S[w[3]]);
        };
        for (int k = 0; k < nU; ++k) V[k] = ustep(k, vel[uslot(k)]);
        for (int k = 0; k < nS; ++k) {
          const int g = tslot(k);
          const uint16_t *w = &q.vloc[4 * ((size_t)pt.v0 + k)];
          const int32_t src[4] = {g, nb[0][g], nb[1][g], nb[2][g]};
          u64 v[4];
          for (int c = 0; c < 4; ++c) {
            if (w[c] == evp1d::PairPlan::kNoEntry) {
              v[c] = vel[src[c]];
            } else if (w[c] >= evp1d::PairPlan::kFixBase) {
              const int32_t s = q.fix[(size_t)pt.f0 + (w[c] - evp1d::PairPlan::kFixBase)];
              v[c] = s < 0 ? 0 : vel[s];
            } else {
              v[c] = V[w[c]];
            }
          }
          u64 s2;
          tstress(sg[k], v[0], v[1], v[2], v[3], s2, S[k]);
          if (k < nO) os[g] = s2;
        }
        for (int k = 0; k < nUo; ++k) ov[uslot(k)] = ustep(k, V[k]);
      }
      copy_list(ov);
      int bad = 0;
      for (int s = 0; s < t_end; ++s) bad += os[s] != rs[s];
      for (int s = u_begin; s < u_end; ++s) bad += ov[s] != rv[s];
      for (int s = na; s < navel; ++s) bad += ov[s] != rv[s];
      printf("  schedule vs two plain subcycles: %d differing outputs\n", bad);
      fails += bad != 0;
    }
  }
  // corruptions of a valid plan
  const evp1d::TilePlan p = evp1d::plan_tiles(na, navel, ip, skip.data(), 32, 16, 2560);
  const evp1d::PairPlan good = evp1d::plan_pairs(p, na, navel, ip, skip.data(), 4096, 4096, 1024);
  auto first_tile = [&](const evp1d::PairPlan &q, std::function<bool(const evp1d::PairTile &)> pred) {
    for (size_t t = 0; t < q.tiles.size(); ++t)
      if (pred(q.tiles[t])) return (int)t;
    return -1;
  };
  std::vector<std::pair<const char *, std::function<void(evp1d::PairPlan &)>>> cases = {
      {"swap two uloc entries", [](evp1d::PairPlan &q) { std::swap(q.uloc[400], q.uloc[401]); }},
      {"uloc entry off by one", [](evp1d::PairPlan &q) { q.uloc[1002] += 1; }},
      {"vloc entry off by one", [](evp1d::PairPlan &q) {
         for (auto &v : q.vloc) if (v != evp1d::PairPlan::kNoEntry && v > 0) { v -= 1; break; } }},
      {"vloc active source dropped", [](evp1d::PairPlan &q) {
         for (auto &v : q.vloc) if (v != evp1d::PairPlan::kNoEntry) { v = evp1d::PairPlan::kNoEntry; break; } }},
      {"vloc constant source named", [](evp1d::PairPlan &q) {
         for (auto &v : q.vloc) if (v == evp1d::PairPlan::kNoEntry) { v = 0; break; } }},
      {"T1-only entry replaced", [](evp1d::PairPlan &q) { q.xlist[7] = q.xlist[8]; }},
      {"U1-only entry replaced", [](evp1d::PairPlan &q) { q.elist[7] = q.elist[8]; }},
      {"constant entry is a T slot", [](evp1d::PairPlan &q) { q.klist[0] = 0; }},
      {"owned stepu reads T1-only", [&](evp1d::PairPlan &q) {
         const int t = first_tile(q, [](const evp1d::PairTile &x) { return x.nx > 0 && x.t.b1 > x.t.b0; });
         const auto &d = q.tiles[t].t;
         q.uloc[4 * (size_t)q.tiles[t].u0 + 1] = (uint16_t)(d.a1 - d.a0 + d.b1 - d.b0 + d.nh);
       }},
      {"tnbr neighbour slot changed", [](evp1d::PairPlan &q) { q.tnbr[4 * 300 + 2] += 1; }},
      {"tnbr entry slot changed", [&](evp1d::PairPlan &q) {
         const int t = first_tile(q, [](const evp1d::PairTile &x) { return x.nx > 0; });
         const auto &d = q.tiles[t].t;
         const int k = d.a1 - d.a0 + d.b1 - d.b0 + d.nh;   // the first T1-only entry
         q.tnbr[4 * ((size_t)q.tiles[t].n0 + k)] = q.tnbr[4 * ((size_t)q.tiles[t].n0 + k) + 1];
       }},
      {"T2 above the thread limit", [](evp1d::PairPlan &q) { q.max_t2 -= 1; }},
      {"T1-only range past the list", [](evp1d::PairPlan &q) { q.tiles.back().nx += 1000000; }},
      {"tile owned range changed", [](evp1d::PairPlan &q) { q.tiles[3].t.b1 -= 1; }},
  };
  {
    // with a ghost copy list: a changed fixed source must be reported
    std::vector<int32_t> hd, hs;
    for (int g = na, k = 0; g < navel; g += 2, ++k) {
      hd.push_back(g);
      hs.push_back(k % 2 ? -1 : (k * 31) % std::max(p.cnt[0], 1));
    }
    evp1d::HaloList hl;
    hl.n = (int)hd.size();
    hl.dst = hd.data();
    hl.src = hs.data();
    evp1d::PairPlan qh = evp1d::plan_pairs(p, na, navel, ip, skip.data(), 4096, 4096, 1024, hl);
    const std::string r0 = evp1d::check_pairs(qh, p, na, navel, ip, skip.data(), hl);
    size_t j = 0;
    while (j < qh.fix.size() && qh.fix[j] < 0) ++j;
    std::string r1 = "no fixed source";
    if (j < qh.fix.size()) {
      qh.fix[j] = -1;
      r1 = evp1d::check_pairs(qh, p, na, navel, ip, skip.data(), hl);
    }
    printf("copy list: good '%s'; %-22s -> %s\n", r0.c_str(), "fixed source changed", r1.empty() ? "NOT CAUGHT" : r1.c_str());
    fails += !r0.empty() || r1.empty();
  }
  printf("good: '%s'\n", evp1d::check_pairs(good, p, na, navel, ip, skip.data()).c_str());
  for (auto &c : cases) {
    auto q = good;
    c.second(q);
    const std::string r = evp1d::check_pairs(q, p, na, navel, ip, skip.data());
    printf("%-30s -> %s\n", c.first, r.empty() ? "NOT CAUGHT" : r.c_str());
    fails += r.empty();
  }
  return fails;
}
extern "C" int evp1d__set_error(int c, const char *) { return c; }
