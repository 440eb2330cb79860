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

// Outputs of the stress phase at one T point: the 12 updated stresses and
// str1..8 (bench_v2b.F90:187-310).
struct StressOut {
  double sp[4], sm[4], s12[4];
  double str[8];
};

// visc_replpress (bench_v2b.F90:415-437)
__host__ __device__ inline void visc_replpress(const Scalars &s, double strength, double DminArea,
                                               double Delta, double &zetax2, double &etax2,
                                               double &rep_prs) {
  const double dmax = (Delta > DminArea) ? Delta : DminArea;   // max(Delta, DminArea)
  const double tmpcalc = s.capping * (strength / dmax) + (c1 - s.capping) * (strength / (Delta + DminArea));
  zetax2 = (c1 + s.Ktens) * tmpcalc;
  rep_prs = (c1 - s.Ktens) * tmpcalc * Delta;
  etax2 = s.epp2i * zetax2;
}

// stress for one T point (bench_v2b.F90:129-310, strain_rates 321-397).
// sig_in: the 12 stresses at the start of the subcycle in the order
// stressp_1..4, stressm_1..4, stress12_1..4.
__host__ __device__ inline void stress_point(const Scalars &s,
    double uvel_cc, double vvel_cc, double uvel_ee, double vvel_ee,
    double uvel_se, double vvel_se, double uvel_ne, double vvel_ne,
    double dxT, double dyT, double strength,
    double hte, double htn, double htem1, double htnm1,
    const double *sig_in, StressOut &o) {
  // metric terms (bench_v2b.F90:139-147)
  const double cxp = c1p5 * htn - p5 * htnm1;
  const double cyp = c1p5 * hte - p5 * htem1;
  const double cxm = -(c1p5 * htnm1 - p5 * htn);
  const double cym = -(c1p5 * htem1 - p5 * hte);
  const double tmparea = dxT * dyT;
  const double DminTarea = s.deltaminEVP * tmparea;
  const double dxhy = p5 * (hte - htem1);
  const double dyhx = p5 * (htn - htnm1);

  // strain rates (bench_v2b.F90:363-390)
  const double divune = cyp * uvel_cc - dyT * uvel_ee + cxp * vvel_cc - dxT * vvel_se;
  const double divunw = cym * uvel_ee + dyT * uvel_cc + cxp * vvel_ee - dxT * vvel_ne;
  const double divusw = cym * uvel_ne + dyT * uvel_se + cxm * vvel_ne + dxT * vvel_ee;
  const double divuse = cyp * uvel_se - dyT * uvel_ne + cxm * vvel_se + dxT * vvel_cc;

  const double tensionne = -cym * uvel_cc - dyT * uvel_ee + cxm * vvel_cc + dxT * vvel_se;
  const double tensionnw = -cyp * uvel_ee + dyT * uvel_cc + cxm * vvel_ee + dxT * vvel_ne;
  const double tensionsw = -cyp * uvel_ne + dyT * uvel_se + cxp * vvel_ne - dxT * vvel_ee;
  const double tensionse = -cym * uvel_se - dyT * uvel_ne + cxp * vvel_se - dxT * vvel_cc;

  const double shearne = -cym * vvel_cc - dyT * vvel_ee - cxm * uvel_cc - dxT * uvel_se;
  const double shearnw = -cyp * vvel_ee + dyT * vvel_cc - cxm * uvel_ee - dxT * uvel_ne;
  const double shearsw = -cyp * vvel_ne + dyT * vvel_se - cxp * uvel_ne + dxT * uvel_ee;
  const double shearse = -cym * vvel_se - dyT * vvel_ne - cxp * uvel_se + dxT * uvel_cc;

  // Delta (bench_v2b.F90:392-395)
  const double Deltane = sqrt(divune * divune + s.e_factor * (tensionne * tensionne + shearne * shearne));
  const double Deltanw = sqrt(divunw * divunw + s.e_factor * (tensionnw * tensionnw + shearnw * shearnw));
  const double Deltasw = sqrt(divusw * divusw + s.e_factor * (tensionsw * tensionsw + shearsw * shearsw));
  const double Deltase = sqrt(divuse * divuse + s.e_factor * (tensionse * tensionse + shearse * shearse));

  double zetax2ne, etax2ne, rep_prsne, zetax2nw, etax2nw, rep_prsnw;
  double zetax2sw, etax2sw, rep_prssw, zetax2se, etax2se, rep_prsse;
  visc_replpress(s, strength, DminTarea, Deltane, zetax2ne, etax2ne, rep_prsne);
  visc_replpress(s, strength, DminTarea, Deltanw, zetax2nw, etax2nw, rep_prsnw);
  visc_replpress(s, strength, DminTarea, Deltasw, zetax2sw, etax2sw, rep_prssw);
  visc_replpress(s, strength, DminTarea, Deltase, zetax2se, etax2se, rep_prsse);

  // stress relaxation (bench_v2b.F90:187-212)
  const double keep = c1 - s.arlx1i * s.revp;
  const double sp1 = (sig_in[0] * keep + s.arlx1i * (zetax2ne * divune - rep_prsne)) * s.denom1;
  const double sp2 = (sig_in[1] * keep + s.arlx1i * (zetax2nw * divunw - rep_prsnw)) * s.denom1;
  const double sp3 = (sig_in[2] * keep + s.arlx1i * (zetax2sw * divusw - rep_prssw)) * s.denom1;
  const double sp4 = (sig_in[3] * keep + s.arlx1i * (zetax2se * divuse - rep_prsse)) * s.denom1;

  const double sm1 = (sig_in[4] * keep + s.arlx1i * etax2ne * tensionne) * s.denom1;
  const double sm2 = (sig_in[5] * keep + s.arlx1i * etax2nw * tensionnw) * s.denom1;
  const double sm3 = (sig_in[6] * keep + s.arlx1i * etax2sw * tensionsw) * s.denom1;
  const double sm4 = (sig_in[7] * keep + s.arlx1i * etax2se * tensionse) * s.denom1;

  const double s121 = (sig_in[8] * keep + s.arlx1i * p5 * etax2ne * shearne) * s.denom1;
  const double s122 = (sig_in[9] * keep + s.arlx1i * p5 * etax2nw * shearnw) * s.denom1;
  const double s123 = (sig_in[10] * keep + s.arlx1i * p5 * etax2sw * shearsw) * s.denom1;
  const double s124 = (sig_in[11] * keep + s.arlx1i * p5 * etax2se * shearse) * s.denom1;

  o.sp[0] = sp1; o.sp[1] = sp2; o.sp[2] = sp3; o.sp[3] = sp4;
  o.sm[0] = sm1; o.sm[1] = sm2; o.sm[2] = sm3; o.sm[3] = sm4;
  o.s12[0] = s121; o.s12[1] = s122; o.s12[2] = s123; o.s12[3] = s124;

  // combinations of the stresses (bench_v2b.F90:217-260)
  const double ssigpn = sp1 + sp2;
  const double ssigps = sp3 + sp4;
  const double ssigpe = sp1 + sp4;
  const double ssigpw = sp2 + sp3;
  const double ssigp1 = (sp1 + sp3) * p055;
  const double ssigp2 = (sp2 + sp4) * p055;

  const double ssigmn = sm1 + sm2;
  const double ssigms = sm3 + sm4;
  const double ssigme = sm1 + sm4;
  const double ssigmw = sm2 + sm3;
  const double ssigm1 = (sm1 + sm3) * p055;
  const double ssigm2 = (sm2 + sm4) * p055;

  const double ssig12n = s121 + s122;
  const double ssig12s = s123 + s124;
  const double ssig12e = s121 + s124;
  const double ssig12w = s122 + s123;
  const double ssig121 = (s121 + s123) * p111;
  const double ssig122 = (s122 + s124) * p111;

  const double csigpne = p111 * sp1 + ssigp2 + p027 * sp3;
  const double csigpnw = p111 * sp2 + ssigp1 + p027 * sp4;
  const double csigpsw = p111 * sp3 + ssigp2 + p027 * sp1;
  const double csigpse = p111 * sp4 + ssigp1 + p027 * sp2;

  const double csigmne = p111 * sm1 + ssigm2 + p027 * sm3;
  const double csigmnw = p111 * sm2 + ssigm1 + p027 * sm4;
  const double csigmsw = p111 * sm3 + ssigm2 + p027 * sm1;
  const double csigmse = p111 * sm4 + ssigm1 + p027 * sm2;

  const double csig12ne = p222 * s121 + ssig122 + p055 * s123;
  const double csig12nw = p222 * s122 + ssig121 + p055 * s124;
  const double csig12sw = p222 * s123 + ssig122 + p055 * s121;
  const double csig12se = p222 * s124 + ssig121 + p055 * s122;

  const double str12ew = p5 * dxT * (p333 * ssig12e + p166 * ssig12w);
  const double str12we = p5 * dxT * (p333 * ssig12w + p166 * ssig12e);
  const double str12ns = p5 * dyT * (p333 * ssig12n + p166 * ssig12s);
  const double str12sn = p5 * dyT * (p333 * ssig12s + p166 * ssig12n);

  // dF/dx (bench_v2b.F90:265-285)
  double strp_tmp = p25 * dyT * (p333 * ssigpn + p166 * ssigps);
  double strm_tmp = p25 * dyT * (p333 * ssigmn + p166 * ssigms);
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

struct StepuOut {
  double uvel, vvel, strintx, strinty, Cb;
};

// stepu for one U point (bench_v2b.F90:509-544).  str1/str5 are the point's
// own values; str2/str7 come from nw, str3/str6 from sse, str4/str8 from sw.
__host__ __device__ inline void stepu_point(const Scalars &s,
    double uold, double vold, double Cw, double aiX, double uocn, double vocn,
    double waterx, double watery, double forcex, double forcey, double umassdti,
    double fm, double uarear, double uvel_init, double vvel_init, double Tbu,
    double str1, double str2_nw, double str3_sse, double str4_sw,
    double str5, double str6_sse, double str7_nw, double str8_sw, StepuOut &o) {
  const double du = uocn - uold;
  const double dv = vocn - vold;
  const double vrel = aiX * s.rhow * Cw * sqrt(du * du + dv * dv);
  const double taux = vrel * waterx;
  const double tauy = vrel * watery;
  const double Cb = Tbu / (sqrt(uold * uold + vold * vold) + s.u0);
  const double cca = (s.brlx + s.revp) * umassdti + vrel * s.cosw + Cb;
  const double ccb = fm + copysign(c1, fm) * vrel * s.sinw;   // sign(c1, fm)
  const double ab2 = cca * cca + ccb * ccb;
  const double strintx = uarear * (str1 + str2_nw + str3_sse + str4_sw);
  const double strinty = uarear * (str5 + str6_sse + str7_nw + str8_sw);
  const double cc1 = strintx + forcex + taux + umassdti * (s.brlx * uold + s.revp * uvel_init);
  const double cc2 = strinty + forcey + tauy + umassdti * (s.brlx * vold + s.revp * vvel_init);
  o.uvel = (cca * cc1 + ccb * cc2) / ab2;
  o.vvel = (cca * cc2 - ccb * cc1) / ab2;
  o.strintx = strintx;
  o.strinty = strinty;
  o.Cb = Cb;
}

}  // namespace evp1d
