// evp1d_adjoint.hpp - hand-written reverse mode of the per-point EVP kernels.
//
// Differentiates the loss-carrying state (the 12 stresses, uvel, vvel) of one
// subcycle with respect to its inputs, point by point, for the gradient of a
// final-velocity misfit with respect to strength, the ocean drag coefficient
// Cw (cdn_ocn) and the forcing forcex/forcey.  The forward statements being
// reversed are those of evp1d_point.hpp (src/bench_v2b.F90:129-310, 509-544).
// Branches follow the forward values: max(Delta, DminArea) differentiates the
// branch taken, sqrt at 0 and |.| at 0 take derivative 0, sign(1, fm) is
// constant.  Derivatives are exact products of the partials (no finite
// differences); rounding differs from the forward, so no bit-exactness claim.
#pragma once

#include "evp1d_point.hpp"

namespace evp1d {

// ---------------------------------------------------------------------------
// stepu (bench_v2b.F90:509-544)
// ---------------------------------------------------------------------------
struct StepuAdj {
  double uold_b, vold_b;   // adjoint of the old velocity (stepu part)
  double gx, gy;           // uarear * adjoint of cc1 / cc2: the adjoint of every str term
  double fx_b, fy_b;       // d/d forcex, d/d forcey
  double cw_b;             // d/d Cw
};

__host__ __device__ inline void stepu_adjoint(const Scalars &s,
    double uold, double vold, double Cw, double aiX, double uocn, double vocn,
    double waterx, double watery, double forcex, double forcey, double umassdti,
    double fm, double uarear, double uvel_init, double vvel_init, double Tbu,
    double str1, double str2_nw, double str3_sse, double str4_sw,
    double str5, double str6_sse, double str7_nw, double str8_sw,
    double u_b, double v_b, StepuAdj &o) {
  // forward recompute
  const double du = uocn - uold, dv = vocn - vold;
  const double R = sqrt(du * du + dv * dv);
  const double A = aiX * s.rhow * Cw;
  const double vrel = A * R;
  const double S = sqrt(uold * uold + vold * vold);
  const double den = S + s.u0;
  const double Cb = Tbu / den;
  const double sg = copysign(c1, fm);
  const double cca = (s.brlx + s.revp) * umassdti + vrel * s.cosw + Cb;
  const double ccb = fm + sg * vrel * s.sinw;
  const double ab2 = cca * cca + ccb * ccb;
  const double strintx = uarear * (str1 + str2_nw + str3_sse + str4_sw);
  const double strinty = uarear * (str5 + str6_sse + str7_nw + str8_sw);
  const double cc1 = strintx + forcex + vrel * waterx + umassdti * (s.brlx * uold + s.revp * uvel_init);
  const double cc2 = strinty + forcey + vrel * watery + umassdti * (s.brlx * vold + s.revp * vvel_init);
  const double u = (cca * cc1 + ccb * cc2) / ab2;
  const double v = (cca * cc2 - ccb * cc1) / ab2;
  // u = (cca cc1 + ccb cc2)/ab2, v = (cca cc2 - ccb cc1)/ab2, ab2 = cca^2 + ccb^2
  const double cc1_b = (u_b * cca - v_b * ccb) / ab2;
  const double cc2_b = (u_b * ccb + v_b * cca) / ab2;
  const double cca_b = (u_b * (cc1 - 2.0 * cca * u) + v_b * (cc2 - 2.0 * cca * v)) / ab2;
  const double ccb_b = (u_b * (cc2 - 2.0 * ccb * u) - v_b * (cc1 + 2.0 * ccb * v)) / ab2;
  const double vrel_b = cca_b * s.cosw + ccb_b * sg * s.sinw + cc1_b * waterx + cc2_b * watery;
  const double Cb_b = cca_b;
  double uold_b = cc1_b * umassdti * s.brlx;
  double vold_b = cc2_b * umassdti * s.brlx;
  if (R > 0.0) {                          // vrel = A |(uocn - uold, vocn - vold)|
    uold_b -= vrel_b * A * du / R;
    vold_b -= vrel_b * A * dv / R;
  }
  if (S > 0.0) {                          // Cb = Tbu / (|(uold, vold)| + u0)
    const double t = -Cb_b * Tbu / (den * den) / S;
    uold_b += t * uold;
    vold_b += t * vold;
  }
  o.uold_b = uold_b;
  o.vold_b = vold_b;
  o.gx = uarear * cc1_b;
  o.gy = uarear * cc2_b;
  o.fx_b = cc1_b;
  o.fy_b = cc2_b;
  o.cw_b = vrel_b * aiX * s.rhow * R;
}

// ---------------------------------------------------------------------------
// stress (bench_v2b.F90:129-310)
// ---------------------------------------------------------------------------
struct StressAdj {
  double sig_b[12];   // adjoint of the stresses at the start of the subcycle
  double vel_b[8];    // adjoint contributions to u,v at cc, ee, se, ne
  double strength_b;  // d/d strength
};

// Reverse of the str1..8 combinations (bench_v2b.F90:217-310), which are
// linear in the updated stresses: accumulates M^T b into sb[12].
__host__ __device__ inline void combos_adjoint(double dxT, double dyT, double dxhy, double dyhx,
                                               const double *b, double *sb) {
  const double b1 = b[0], b2 = b[1], b3 = b[2], b4 = b[3], b5 = b[4], b6 = b[5], b7 = b[6], b8 = b[7];
  // str1..8 in terms of strp/strm/str12 and the csig terms
  const double strp_a = -b1 + b2, strm_a = -b1 + b2;      // dF/dx north pair
  const double strp_b = -b3 + b4, strm_b = -b3 + b4;      // dF/dx south pair
  const double strp_c = -b5 + b6, strm_c = b5 - b6;       // dF/dy east pair
  const double strp_d = -b7 + b8, strm_d = b7 - b8;       // dF/dy west pair
  const double s12ew = -b1 + b3, s12we = -b2 + b4, s12ns = -b5 + b7, s12sn = -b6 + b8;
  const double cpne = -dxhy * b1 - dyhx * b5, cmne = dxhy * b1 - dyhx * b5, c12ne = dyhx * b1 + dxhy * b5;
  const double cpnw = -dxhy * b2 - dyhx * b7, cmnw = dxhy * b2 - dyhx * b7, c12nw = dyhx * b2 + dxhy * b7;
  const double cpse = -dxhy * b3 - dyhx * b6, cmse = dxhy * b3 - dyhx * b6, c12se = dyhx * b3 + dxhy * b6;
  const double cpsw = -dxhy * b4 - dyhx * b8, cmsw = dxhy * b4 - dyhx * b8, c12sw = dyhx * b4 + dxhy * b8;
  // strp/strm = p25*d*(p333*A + p166*B)
  const double ay = p25 * dyT, ax = p25 * dxT;
  const double pn = ay * (p333 * strp_a + p166 * strp_b), ps = ay * (p166 * strp_a + p333 * strp_b);
  const double pe = ax * (p333 * strp_c + p166 * strp_d), pw = ax * (p166 * strp_c + p333 * strp_d);
  const double mn = ay * (p333 * strm_a + p166 * strm_b), ms = ay * (p166 * strm_a + p333 * strm_b);
  const double me = ax * (p333 * strm_c + p166 * strm_d), mw = ax * (p166 * strm_c + p333 * strm_d);
  // str12 = p5*d*(p333*A + p166*B)
  const double hx = p5 * dxT, hy = p5 * dyT;
  const double e12 = hx * (p333 * s12ew + p166 * s12we), w12 = hx * (p166 * s12ew + p333 * s12we);
  const double n12 = hy * (p333 * s12ns + p166 * s12sn), s12 = hy * (p166 * s12ns + p333 * s12sn);
  // csig p/m: csig_ne = p111*x1 + ssig2 + p027*x3, nw = p111*x2 + ssig1 + p027*x4,
  //           sw = p111*x3 + ssig2 + p027*x1, se = p111*x4 + ssig1 + p027*x2
  // ssig1 = (x1 + x3)*p055, ssig2 = (x2 + x4)*p055
  const double qp1 = (cpnw + cpse) * p055, qp2 = (cpne + cpsw) * p055;
  const double qm1 = (cmnw + cmse) * p055, qm2 = (cmne + cmsw) * p055;
  // csig12: ne = p222*x1 + ssig122 + p055*x3, nw = p222*x2 + ssig121 + p055*x4,
  //         sw = p222*x3 + ssig122 + p055*x1, se = p222*x4 + ssig121 + p055*x2
  // ssig121 = (x1 + x3)*p111, ssig122 = (x2 + x4)*p111
  const double q121 = (c12nw + c12se) * p111, q122 = (c12ne + c12sw) * p111;
  // p family: ssigpn = x1+x2, ps = x3+x4, pe = x1+x4, pw = x2+x3
  sb[0] += pn + pe + p111 * cpne + p027 * cpsw + qp1;
  sb[1] += pn + pw + p111 * cpnw + p027 * cpse + qp2;
  sb[2] += ps + pw + p111 * cpsw + p027 * cpne + qp1;
  sb[3] += ps + pe + p111 * cpse + p027 * cpnw + qp2;
  sb[4] += mn + me + p111 * cmne + p027 * cmsw + qm1;
  sb[5] += mn + mw + p111 * cmnw + p027 * cmse + qm2;
  sb[6] += ms + mw + p111 * cmsw + p027 * cmne + qm1;
  sb[7] += ms + me + p111 * cmse + p027 * cmnw + qm2;
  sb[8] += n12 + e12 + p222 * c12ne + p055 * c12sw + q121;
  sb[9] += n12 + w12 + p222 * c12nw + p055 * c12se + q122;
  sb[10] += s12 + w12 + p222 * c12sw + p055 * c12ne + q121;
  sb[11] += s12 + e12 + p222 * c12se + p055 * c12nw + q122;
}

// Reverse of one subcycle's stress at one T point.  Inputs: the forward
// inputs of stress_point, the adjoint of the updated stresses (sig_b_new)
// and of str1..8 (str_b).  Outputs the adjoint of the stresses before the
// update, the velocity adjoint contributions and d/d strength.
__host__ __device__ inline void stress_adjoint(const Scalars &s,
    double uvel_cc, double vvel_cc, double uvel_ee, double vvel_ee,
    double uvel_se, double vvel_se, double uvel_ne, double vvel_ne,
    double dxT, double dyT, double strength,
    double hte, double htn, double htem1, double htnm1,
    const double *sig_b_new, const double *str_b, StressAdj &o) {
  const double cxp = c1p5 * htn - p5 * htnm1;
  const double cyp = c1p5 * hte - p5 * htem1;
  const double cxm = -(c1p5 * htnm1 - p5 * htn);
  const double cym = -(c1p5 * htem1 - p5 * hte);
  const double DminTarea = s.deltaminEVP * (dxT * dyT);
  const double dxhy = p5 * (hte - htem1);
  const double dyhx = p5 * (htn - htnm1);

  // strain rates and Delta per corner: 0 ne, 1 nw, 2 sw, 3 se (stress index 1..4)
  double dv[4], tn[4], sh[4], De[4];
  dv[0] = cyp * uvel_cc - dyT * uvel_ee + cxp * vvel_cc - dxT * vvel_se;
  dv[1] = cym * uvel_ee + dyT * uvel_cc + cxp * vvel_ee - dxT * vvel_ne;
  dv[2] = cym * uvel_ne + dyT * uvel_se + cxm * vvel_ne + dxT * vvel_ee;
  dv[3] = cyp * uvel_se - dyT * uvel_ne + cxm * vvel_se + dxT * vvel_cc;
  tn[0] = -cym * uvel_cc - dyT * uvel_ee + cxm * vvel_cc + dxT * vvel_se;
  tn[1] = -cyp * uvel_ee + dyT * uvel_cc + cxm * vvel_ee + dxT * vvel_ne;
  tn[2] = -cyp * uvel_ne + dyT * uvel_se + cxp * vvel_ne - dxT * vvel_ee;
  tn[3] = -cym * uvel_se - dyT * uvel_ne + cxp * vvel_se - dxT * vvel_cc;
  sh[0] = -cym * vvel_cc - dyT * vvel_ee - cxm * uvel_cc - dxT * uvel_se;
  sh[1] = -cyp * vvel_ee + dyT * vvel_cc - cxm * uvel_ee - dxT * uvel_ne;
  sh[2] = -cyp * vvel_ne + dyT * vvel_se - cxp * uvel_ne + dxT * uvel_ee;
  sh[3] = -cym * vvel_se - dyT * vvel_ne - cxp * uvel_se + dxT * uvel_cc;
  for (int c = 0; c < 4; ++c) De[c] = sqrt(dv[c] * dv[c] + s.e_factor * (tn[c] * tn[c] + sh[c] * sh[c]));

  // adjoint of the updated stresses including the str1..8 path
  double sb[12];
  for (int k = 0; k < 12; ++k) sb[k] = sig_b_new[k];
  combos_adjoint(dxT, dyT, dxhy, dyhx, str_b, sb);

  const double keep = c1 - s.arlx1i * s.revp;
  double dv_b[4], tn_b[4], sh_b[4], strength_b = 0.0;
  for (int c = 0; c < 4; ++c) {
    // forward: tmpcalc, zetax2, etax2, rep_prs (visc_replpress)
    const double D = De[c];
    const double mx = (D > DminTarea) ? D : DminTarea;
    const double dsum = D + DminTarea;
    const double T = s.capping * (strength / mx) + (c1 - s.capping) * (strength / dsum);
    const double zeta = (c1 + s.Ktens) * T;
    const double eta = s.epp2i * zeta;
    // sp = (sp0*keep + arlx1i*(zeta*div - rep))*denom1, rep = (1-K)*T*D
    // sm = (sm0*keep + arlx1i*eta*ten)*denom1, s12 = (s120*keep + arlx1i*p5*eta*shr)*denom1
    const double P = sb[c] * s.denom1, M = sb[4 + c] * s.denom1, Q = sb[8 + c] * s.denom1;
    o.sig_b[c] = P * keep;
    o.sig_b[4 + c] = M * keep;
    o.sig_b[8 + c] = Q * keep;
    const double zeta_b0 = P * s.arlx1i * dv[c];
    const double rep_b = -P * s.arlx1i;
    const double eta_b = M * s.arlx1i * tn[c] + Q * s.arlx1i * p5 * sh[c];
    dv_b[c] = P * s.arlx1i * zeta;
    tn_b[c] = M * s.arlx1i * eta;
    sh_b[c] = Q * s.arlx1i * p5 * eta;
    const double zeta_b = zeta_b0 + s.epp2i * eta_b;
    const double T_b = (c1 + s.Ktens) * zeta_b + (c1 - s.Ktens) * D * rep_b;
    double D_b = (c1 - s.Ktens) * T * rep_b;
    strength_b += T_b * (s.capping / mx + (c1 - s.capping) / dsum);
    const double dT_dD = s.capping * ((D > DminTarea) ? -strength / (D * D) : 0.0) +
                         (c1 - s.capping) * (-strength / (dsum * dsum));
    D_b += T_b * dT_dD;
    if (D > 0.0) {
      dv_b[c] += D_b * dv[c] / D;
      tn_b[c] += D_b * s.e_factor * tn[c] / D;
      sh_b[c] += D_b * s.e_factor * sh[c] / D;
    }
  }
  o.strength_b = strength_b;

  // reverse of the strain rates (linear in the 8 velocities)
  double ucc = 0, vcc = 0, uee = 0, vee = 0, use = 0, vse = 0, une = 0, vne = 0;
  // divergence
  ucc += cyp * dv_b[0]; uee -= dyT * dv_b[0]; vcc += cxp * dv_b[0]; vse -= dxT * dv_b[0];
  uee += cym * dv_b[1]; ucc += dyT * dv_b[1]; vee += cxp * dv_b[1]; vne -= dxT * dv_b[1];
  une += cym * dv_b[2]; use += dyT * dv_b[2]; vne += cxm * dv_b[2]; vee += dxT * dv_b[2];
  use += cyp * dv_b[3]; une -= dyT * dv_b[3]; vse += cxm * dv_b[3]; vcc += dxT * dv_b[3];
  // tension
  ucc -= cym * tn_b[0]; uee -= dyT * tn_b[0]; vcc += cxm * tn_b[0]; vse += dxT * tn_b[0];
  uee -= cyp * tn_b[1]; ucc += dyT * tn_b[1]; vee += cxm * tn_b[1]; vne += dxT * tn_b[1];
  une -= cyp * tn_b[2]; use += dyT * tn_b[2]; vne += cxp * tn_b[2]; vee -= dxT * tn_b[2];
  use -= cym * tn_b[3]; une -= dyT * tn_b[3]; vse += cxp * tn_b[3]; vcc -= dxT * tn_b[3];
  // shear
  vcc -= cym * sh_b[0]; vee -= dyT * sh_b[0]; ucc -= cxm * sh_b[0]; use -= dxT * sh_b[0];
  vee -= cyp * sh_b[1]; vcc += dyT * sh_b[1]; uee -= cxm * sh_b[1]; une -= dxT * sh_b[1];
  vne -= cyp * sh_b[2]; vse += dyT * sh_b[2]; une -= cxp * sh_b[2]; uee += dxT * sh_b[2];
  vse -= cym * sh_b[3]; vne -= dyT * sh_b[3]; use -= cxp * sh_b[3]; ucc += dxT * sh_b[3];
  o.vel_b[0] = ucc; o.vel_b[1] = vcc;
  o.vel_b[2] = uee; o.vel_b[3] = vee;
  o.vel_b[4] = use; o.vel_b[5] = vse;
  o.vel_b[6] = une; o.vel_b[7] = vne;
}

}  // namespace evp1d
