// The division-free second term of visc_replpress (capping_term with
// capping_fast) against the reference expression
// tmpcalc = capping * (strength / max(Delta, DminArea))
//         + (1 - capping) * (strength / (Delta + DminArea))
// (src/bench_v2b.F90:431-432): every output of visc_replpress must have the
// same bits on edge values (signed zeros, subnormals, huge, infinities,
// NaN) and on random values, for capping = 1 (where the fast path applies)
// and for other cappings (where it must not).  Prints "ok N" or the first
// mismatch.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "evp1d_point.hpp"

static uint64_t bits(double x) {
  uint64_t u;
  std::memcpy(&u, &x, 8);
  return u;
}

int main() {
  const double inf = INFINITY, nan = NAN, dmax = 1.7976931348623157e308, den = 4.9e-324;
  const std::vector<double> S = {0.0, -0.0, den, -den, 1e-300, 1.0, 3.5e4, 1e300, dmax, -1.0, -dmax, inf, -inf, nan};
  const std::vector<double> D = {1.0, 1e-8, 1e-300, den, 0.0, -0.0, -1.0, 0x1p30, 0x1p200, dmax, inf, nan};
  const std::vector<double> L = {0.0, -0.0, den, 1e-20, 0.5, 1.0, 1e10, 1e300, dmax, inf, nan};
  std::vector<double> caps = {1.0, 0.5, 0.0, 2.0};
  std::mt19937_64 rng(12345);
  std::uniform_real_distribution<double> u(-8.0, 8.0);
  long n = 0;
  auto check = [&](const evp1d::Scalars &s, double st, double dm, double dl) -> bool {
    double z0, e0, r0, z1, e1, r1;
    evp1d::visc_replpress(s, st, dm, dl, z0, e0, r0, false);                          // the division
    evp1d::visc_replpress(s, st, dm, dl, z1, e1, r1, evp1d::capping_fast(s, st, dm));   // as the kernels call it
    ++n;
    if (bits(z0) != bits(z1) || bits(e0) != bits(e1) || bits(r0) != bits(r1)) {
      std::printf("mismatch capping=%g strength=%a DminArea=%a Delta=%a: zeta %a/%a eta %a/%a rep %a/%a\n",
                  s.capping, st, dm, dl, z0, z1, e0, e1, r0, r1);
      return false;
    }
    return true;
  };
  for (double c : caps) {
    evp1d::Scalars s{};
    s.capping = c;
    s.Ktens = 0.0;
    s.epp2i = 0.25;
    for (double st : S)
      for (double dm : D)
        for (double dl : L)
          if (!check(s, st, dm, dl)) return 1;
    for (int k = 0; k < 200000; ++k) {
      const double st = std::ldexp(std::fabs(u(rng)), (int)(u(rng) * 40)) * (k % 97 == 0 ? -1.0 : 1.0);
      const double dm = std::ldexp(std::fabs(u(rng)), (int)(u(rng) * 40) - 20);
      const double dl = std::sqrt(std::ldexp(std::fabs(u(rng)), (int)(u(rng) * 60)));
      if (!check(s, st, dm, dl)) return 1;
    }
  }
  std::printf("ok %ld\n", n);
  return 0;
}
