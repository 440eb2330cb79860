// evp1d_point.hpp - per-point EVP arithmetic shared by every kernel form.
//
// Each function evaluates exactly the expression tree of the reference's
// serial Fortran (left-to-right association, parentheses as written, no
// contraction in the exact build) so that the GPU result is bit-identical to
// the reference f0 build.  Citations are to src/bench_v2b.F90.
//
// Unary minus: Fortran parses -a*b as -(a*b); C parses it as (-a)*b.  IEEE
// multiplication is sign-symmetric, so both round to the same value.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace evp1d {

// ice_constants (src/ice.F90:25-68): the p-constants are c1/c9 etc., i.e.
// the correctly rounded quotients, which is what these literal divisions give.
constexpr double c1 = 1.0;
constexpr double c1p5 = 1.5;
constexpr double p5 = 0.5;
constexpr double p25 = 0.25;
constexpr double p111 = 1.0 / 9.0;
constexpr double p166 = 1.0 / 6.0;
constexpr double p222 = 2.0 / 9.0;
constexpr double p333 = 1.0 / 3.0;
constexpr double p055 = p111 * p5;
constexpr double p027 = p055 * p5;

struct Scalars {
  double capping, e_factor, epp2i, deltaminEVP;
  double arlx1i, denom1, brlx, revp, Ktens;
  double u0, cosw, sinw, rhow;
};

// Forward-mode (tangent-linear) number: value and directional derivative.
// Each operation computes its value part with the same single IEEE operation
// as on double, so a Dual evaluation reproduces the double results bit for
// bit while propagating derivatives.  sqrt has derivative 0 at 0; comparisons
// look at values only.
struct Dual {
  double v, d;
};
using ::copysign;
using ::sqrt;
__host__ __device__ inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__host__ __device__ inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__host__ __device__ inline Dual operator-(Dual a) { return {-a.v, -a.d}; }
__host__ __device__ inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__host__ __device__ inline Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
__host__ __device__ inline Dual operator+(double a, Dual b) { return {a + b.v, b.d}; }
__host__ __device__ inline Dual operator+(Dual a, double b) { return {a.v + b, a.d}; }
__host__ __device__ inline Dual operator-(double a, Dual b) { return {a - b.v, -b.d}; }
__host__ __device__ inline Dual operator-(Dual a, double b) { return {a.v - b, a.d}; }
__host__ __device__ inline Dual operator*(double a, Dual b) { return {a * b.v, a * b.d}; }
__host__ __device__ inline Dual operator*(Dual a, double b) { return {a.v * b, a.d * b}; }
__host__ __device__ inline Dual operator/(Dual a, double b) { return {a.v / b, a.d / b}; }
__host__ __device__ inline Dual operator/(double a, Dual b) {
  const double q = a / b.v;
  return {q, -q * b.d / b.v};
}
__host__ __device__ inline bool operator>(Dual a, Dual b) { return a.v > b.v; }
__host__ __device__ inline Dual sqrt(Dual a) {
  const double r = ::sqrt(a.v);
  return {r, r > 0.0 ? a.d / (2.0 * r) : 0.0};
}
__host__ __device__ inline Dual copysign(double mag, Dual sgn) { return {::copysign(mag, sgn.v), 0.0}; }
__host__ __device__ inline double value_of(double x) { return x; }
__host__ __device__ inline double value_of(Dual x) { return x.v; }

// Outputs of the stress phase at one T point: the 12 updated stresses and
// str1..8 (bench_v2b.F90:187-310).
template <class R>
struct StressOutT {
  R sp[4], sm[4], s12[4];
  R str[8];
};
using StressOut = StressOutT<double>;

// The second term of tmpcalc in visc_replpress,
// (c1 - capping) * (strength / (Delta + DminArea)), without its division
// where that gives the same IEEE value.  `fast` (capping_fast, once per
// point) holds when capping = 1, DminArea > 0 and strength is finite with
// |strength| <= DminArea * 2^1000.  The term is then +0 * q.  Delta is a
// square root, so unless it is a NaN it is +-0, positive or +inf; the
// divisor is then at least DminArea > 0 and q is finite (|q| <= 2^1000) with
// strength's sign, so +0 * q is the zero with strength's sign.  A NaN Delta
// (or any other case) takes the division.  The tangent-linear type always
// takes the division.
template <class R>
__host__ __device__ inline bool capping_fast(const Scalars &, R, R) {
  return false;
}
__host__ __device__ inline bool capping_fast(const Scalars &s, double strength, double DminArea) {
  return s.capping == c1 && DminArea > 0.0 && fabs(strength) <= fmin(__DBL_MAX__, DminArea * 0x1p1000);
}
template <class R>
__host__ __device__ inline R capping_term(const Scalars &s, bool, R strength, R DminArea, R Delta) {
  return (c1 - s.capping) * (strength / (Delta + DminArea));
}
__host__ __device__ inline double capping_term(const Scalars &s, bool fast, double strength, double DminArea,
                                               double Delta) {
  if (fast && Delta == Delta) return copysign(0.0, strength);
  return (c1 - s.capping) * (strength / (Delta + DminArea));
}

// visc_replpress (bench_v2b.F90:415-437); fast: capping_fast(s, strength, DminArea)
template <class R>
__host__ __device__ inline void visc_replpress(const Scalars &s, R strength, R DminArea,
                                               R Delta, R &zetax2, R &etax2,
                                               R &rep_prs, bool fast = false) {
  const R dmax = (Delta > DminArea) ? Delta : DminArea;   // max(Delta, DminArea)
  const R tmpcalc = s.capping * (strength / dmax) + capping_term(s, fast, strength, DminArea, Delta);
  zetax2 = (c1 + s.Ktens) * tmpcalc;
  rep_prs = (c1 - s.Ktens) * tmpcalc * Delta;
  etax2 = s.epp2i * zetax2;
}

// stress for one T point (bench_v2b.F90:129-310, strain_rates 321-397).
// Stress at one T point from its metric terms (the 2D v0 kernel's inputs
// cxp, cyp, cxm, cym, DminTarea, dxhy, dyhx, bench_v0.F90:62-331; the 1D
// kernel computes them first, stress_point below).  sig_in: the 12 stresses
// at the start of the subcycle in the order stressp_1..4, stressm_1..4,
// stress12_1..4.  Velocities: cc = (i,j), ee = (i-1,j), se = (i,j-1),
// ne = (i-1,j-1).
template <class R>
__host__ __device__ inline void stress_core(const Scalars &s,
    R uvel_cc, R vvel_cc, R uvel_ee, R vvel_ee,
    R uvel_se, R vvel_se, R uvel_ne, R vvel_ne,
    R dxT, R dyT, R strength,
    R cxp, R cyp, R cxm, R cym, R DminTarea, R dxhy, R dyhx,
    const R *sig_in, StressOutT<R> &o) {

  // strain rates (bench_v2b.F90:363-390)
  const R divune = cyp * uvel_cc - dyT * uvel_ee + cxp * vvel_cc - dxT * vvel_se;
  const R divunw = cym * uvel_ee + dyT * uvel_cc + cxp * vvel_ee - dxT * vvel_ne;
  const R divusw = cym * uvel_ne + dyT * uvel_se + cxm * vvel_ne + dxT * vvel_ee;
  const R divuse = cyp * uvel_se - dyT * uvel_ne + cxm * vvel_se + dxT * vvel_cc;

  const R tensionne = -cym * uvel_cc - dyT * uvel_ee + cxm * vvel_cc + dxT * vvel_se;
  const R tensionnw = -cyp * uvel_ee + dyT * uvel_cc + cxm * vvel_ee + dxT * vvel_ne;
  const R tensionsw = -cyp * uvel_ne + dyT * uvel_se + cxp * vvel_ne - dxT * vvel_ee;
  const R tensionse = -cym * uvel_se - dyT * uvel_ne + cxp * vvel_se - dxT * vvel_cc;

  const R shearne = -cym * vvel_cc - dyT * vvel_ee - cxm * uvel_cc - dxT * uvel_se;
  const R shearnw = -cyp * vvel_ee + dyT * vvel_cc - cxm * uvel_ee - dxT * uvel_ne;
  const R shearsw = -cyp * vvel_ne + dyT * vvel_se - cxp * uvel_ne + dxT * uvel_ee;
  const R shearse = -cym * vvel_se - dyT * vvel_ne - cxp * uvel_se + dxT * uvel_cc;

  // Delta (bench_v2b.F90:392-395)
  const R Deltane = sqrt(divune * divune + s.e_factor * (tensionne * tensionne + shearne * shearne));
  const R Deltanw = sqrt(divunw * divunw + s.e_factor * (tensionnw * tensionnw + shearnw * shearnw));
  const R Deltasw = sqrt(divusw * divusw + s.e_factor * (tensionsw * tensionsw + shearsw * shearsw));
  const R Deltase = sqrt(divuse * divuse + s.e_factor * (tensionse * tensionse + shearse * shearse));

  R zetax2ne, etax2ne, rep_prsne, zetax2nw, etax2nw, rep_prsnw;
  R zetax2sw, etax2sw, rep_prssw, zetax2se, etax2se, rep_prsse;
  const bool fast = capping_fast(s, strength, DminTarea);
  visc_replpress(s, strength, DminTarea, Deltane, zetax2ne, etax2ne, rep_prsne, fast);
  visc_replpress(s, strength, DminTarea, Deltanw, zetax2nw, etax2nw, rep_prsnw, fast);
  visc_replpress(s, strength, DminTarea, Deltasw, zetax2sw, etax2sw, rep_prssw, fast);
  visc_replpress(s, strength, DminTarea, Deltase, zetax2se, etax2se, rep_prsse, fast);

  // stress relaxation (bench_v2b.F90:187-212)
  const double keep = c1 - s.arlx1i * s.revp;
  const R sp1 = (sig_in[0] * keep + s.arlx1i * (zetax2ne * divune - rep_prsne)) * s.denom1;
  const R sp2 = (sig_in[1] * keep + s.arlx1i * (zetax2nw * divunw - rep_prsnw)) * s.denom1;
  const R sp3 = (sig_in[2] * keep + s.arlx1i * (zetax2sw * divusw - rep_prssw)) * s.denom1;
  const R sp4 = (sig_in[3] * keep + s.arlx1i * (zetax2se * divuse - rep_prsse)) * s.denom1;

  const R sm1 = (sig_in[4] * keep + s.arlx1i * etax2ne * tensionne) * s.denom1;
  const R sm2 = (sig_in[5] * keep + s.arlx1i * etax2nw * tensionnw) * s.denom1;
  const R sm3 = (sig_in[6] * keep + s.arlx1i * etax2sw * tensionsw) * s.denom1;
  const R sm4 = (sig_in[7] * keep + s.arlx1i * etax2se * tensionse) * s.denom1;

  const R s121 = (sig_in[8] * keep + s.arlx1i * p5 * etax2ne * shearne) * s.denom1;
  const R s122 = (sig_in[9] * keep + s.arlx1i * p5 * etax2nw * shearnw) * s.denom1;
  const R s123 = (sig_in[10] * keep + s.arlx1i * p5 * etax2sw * shearsw) * s.denom1;
  const R s124 = (sig_in[11] * keep + s.arlx1i * p5 * etax2se * shearse) * s.denom1;

  o.sp[0] = sp1; o.sp[1] = sp2; o.sp[2] = sp3; o.sp[3] = sp4;
  o.sm[0] = sm1; o.sm[1] = sm2; o.sm[2] = sm3; o.sm[3] = sm4;
  o.s12[0] = s121; o.s12[1] = s122; o.s12[2] = s123; o.s12[3] = s124;

  // combinations of the stresses (bench_v2b.F90:217-260)
  const R ssigpn = sp1 + sp2;
  const R ssigps = sp3 + sp4;
  const R ssigpe = sp1 + sp4;
  const R ssigpw = sp2 + sp3;
  const R ssigp1 = (sp1 + sp3) * p055;
  const R ssigp2 = (sp2 + sp4) * p055;

  const R ssigmn = sm1 + sm2;
  const R ssigms = sm3 + sm4;
  const R ssigme = sm1 + sm4;
  const R ssigmw = sm2 + sm3;
  const R ssigm1 = (sm1 + sm3) * p055;
  const R ssigm2 = (sm2 + sm4) * p055;

  const R ssig12n = s121 + s122;
  const R ssig12s = s123 + s124;
  const R ssig12e = s121 + s124;
  const R ssig12w = s122 + s123;
  const R ssig121 = (s121 + s123) * p111;
  const R ssig122 = (s122 + s124) * p111;

  const R csigpne = p111 * sp1 + ssigp2 + p027 * sp3;
  const R csigpnw = p111 * sp2 + ssigp1 + p027 * sp4;
  const R csigpsw = p111 * sp3 + ssigp2 + p027 * sp1;
  const R csigpse = p111 * sp4 + ssigp1 + p027 * sp2;

  const R csigmne = p111 * sm1 + ssigm2 + p027 * sm3;
  const R csigmnw = p111 * sm2 + ssigm1 + p027 * sm4;
  const R csigmsw = p111 * sm3 + ssigm2 + p027 * sm1;
  const R csigmse = p111 * sm4 + ssigm1 + p027 * sm2;

  const R csig12ne = p222 * s121 + ssig122 + p055 * s123;
  const R csig12nw = p222 * s122 + ssig121 + p055 * s124;
  const R csig12sw = p222 * s123 + ssig122 + p055 * s121;
  const R csig12se = p222 * s124 + ssig121 + p055 * s122;

  const R str12ew = p5 * dxT * (p333 * ssig12e + p166 * ssig12w);
  const R str12we = p5 * dxT * (p333 * ssig12w + p166 * ssig12e);
  const R str12ns = p5 * dyT * (p333 * ssig12n + p166 * ssig12s);
  const R str12sn = p5 * dyT * (p333 * ssig12s + p166 * ssig12n);

  // dF/dx (bench_v2b.F90:265-285)
  R strp_tmp = p25 * dyT * (p333 * ssigpn + p166 * ssigps);
  R strm_tmp = p25 * dyT * (p333 * ssigmn + p166 * ssigms);
  o.str[0] = -strp_tmp - strm_tmp - str12ew + dxhy * (-csigpne + csigmne) + dyhx * csig12ne;
  o.str[1] = strp_tmp + strm_tmp - str12we + dxhy * (-csigpnw + csigmnw) + dyhx * csig12nw;
  strp_tmp = p25 * dyT * (p333 * ssigps + p166 * ssigpn);
  strm_tmp = p25 * dyT * (p333 * ssigms + p166 * ssigmn);
  o.str[2] = -strp_tmp - strm_tmp + str12ew + dxhy * (-csigpse + csigmse) + dyhx * csig12se;
  o.str[3] = strp_tmp + strm_tmp + str12we + dxhy * (-csigpsw + csigmsw) + dyhx * csig12sw;

  // dF/dy (bench_v2b.F90:290-310)
  strp_tmp = p25 * dxT * (p333 * ssigpe + p166 * ssigpw);
  strm_tmp = p25 * dxT * (p333 * ssigme + p166 * ssigmw);
  o.str[4] = -strp_tmp + strm_tmp - str12ns - dyhx * (csigpne + csigmne) + dxhy * csig12ne;
  o.str[5] = strp_tmp - strm_tmp - str12sn - dyhx * (csigpse + csigmse) + dxhy * csig12se;
  strp_tmp = p25 * dxT * (p333 * ssigpw + p166 * ssigpe);
  strm_tmp = p25 * dxT * (p333 * ssigmw + p166 * ssigme);
  o.str[6] = -strp_tmp + strm_tmp + str12ns - dyhx * (csigpnw + csigmnw) + dxhy * csig12nw;
  o.str[7] = strp_tmp - strm_tmp + str12sn - dyhx * (csigpsw + csigmsw) + dxhy * csig12sw;
}

// Stress at one point of the 1D list (bench_v2b.F90:125-311): the metric
// terms from HTE, HTN and their i-1 / j-1 neighbours (139-147), then the core.
template <class R>
__host__ __device__ inline void stress_point(const Scalars &s,
    R uvel_cc, R vvel_cc, R uvel_ee, R vvel_ee,
    R uvel_se, R vvel_se, R uvel_ne, R vvel_ne,
    R dxT, R dyT, R strength,
    R hte, R htn, R htem1, R htnm1,
    const R *sig_in, StressOutT<R> &o) {
  const R cxp = c1p5 * htn - p5 * htnm1;
  const R cyp = c1p5 * hte - p5 * htem1;
  const R cxm = -(c1p5 * htnm1 - p5 * htn);
  const R cym = -(c1p5 * htem1 - p5 * hte);
  const R tmparea = dxT * dyT;
  const R DminTarea = s.deltaminEVP * tmparea;
  const R dxhy = p5 * (hte - htem1);
  const R dyhx = p5 * (htn - htnm1);
  stress_core(s, uvel_cc, vvel_cc, uvel_ee, vvel_ee, uvel_se, vvel_se, uvel_ne, vvel_ne, dxT, dyT, strength,
              cxp, cyp, cxm, cym, DminTarea, dxhy, dyhx, sig_in, o);
}

template <class R>
struct StepuOutT {
  R uvel, vvel, strintx, strinty, Cb;
};
using StepuOut = StepuOutT<double>;

// stepu for one U point (bench_v2b.F90:509-544).  str1/str5 are the point's
// own values; str2/str7 come from nw, str3/str6 from sse, str4/str8 from sw.
template <class R>
__host__ __device__ inline void stepu_point(const Scalars &s,
    R uold, R vold, R Cw, R aiX, R uocn, R vocn,
    R waterx, R watery, R forcex, R forcey, R umassdti,
    R fm, R uarear, R uvel_init, R vvel_init, R Tbu,
    R str1, R str2_nw, R str3_sse, R str4_sw,
    R str5, R str6_sse, R str7_nw, R str8_sw, StepuOutT<R> &o) {
  const R du = uocn - uold;
  const R dv = vocn - vold;
  const R vrel = aiX * s.rhow * Cw * sqrt(du * du + dv * dv);
  const R taux = vrel * waterx;
  const R tauy = vrel * watery;
  const R Cb = Tbu / (sqrt(uold * uold + vold * vold) + s.u0);
  const R cca = (s.brlx + s.revp) * umassdti + vrel * s.cosw + Cb;
  const R ccb = fm + copysign(c1, fm) * vrel * s.sinw;   // sign(c1, fm)
  const R ab2 = cca * cca + ccb * ccb;
  const R strintx = uarear * (str1 + str2_nw + str3_sse + str4_sw);
  const R strinty = uarear * (str5 + str6_sse + str7_nw + str8_sw);
  const R cc1 = strintx + forcex + taux + umassdti * (s.brlx * uold + s.revp * uvel_init);
  const R cc2 = strinty + forcey + tauy + umassdti * (s.brlx * vold + s.revp * vvel_init);
  o.uvel = (cca * cc1 + ccb * cc2) / ab2;
  o.vvel = (cca * cc2 - ccb * cc1) / ab2;
  o.strintx = strintx;
  o.strinty = strinty;
  o.Cb = Cb;
}

}  // namespace evp1d
