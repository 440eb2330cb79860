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
// differences); rounding differs from the forward (quotients by one value
// are taken as products with its reciprocal), so no bit-exactness claim.
#pragma once

#include "evp1d_cache.hpp"
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

// The forward result (u, v) = (uvel, vvel) after the update is an input: with
// w = cc1 + i cc2 and a = cca + i ccb the update is u + i v = conj(a) w / |a|^2,
// so w = a (u + i v), i.e. cc1 = cca u - ccb v, cc2 = ccb u + cca v.  That
// gives cc1 and cc2 from the stored velocities alone, without the str1..8
// gathers they are summed from.
__host__ __device__ inline void stepu_adjoint(const Scalars &s,
    double uold, double vold, double Cw, double aiX, double uocn, double vocn,
    double waterx, double watery, double umassdti, double fm, double uarear, double Tbu,
    double u, double v, double u_b, double v_b, StepuAdj &o) {
  // forward recompute of the coefficients
  const double du = uocn - uold, dv = vocn - vold;
  const double R = sqrt(du * du + dv * dv);
  const double A = aiX * s.rhow * Cw;
  const double vrel = A * R;
  const double S = sqrt(uold * uold + vold * vold);
  const double den = S + s.u0;
  const double rden = c1 / den;
  const double Cb = Tbu * rden;
  const double sg = copysign(c1, fm);
  const double cca = (s.brlx + s.revp) * umassdti + vrel * s.cosw + Cb;
  const double ccb = fm + sg * vrel * s.sinw;
  const double ab2 = cca * cca + ccb * ccb;
  const double rab2 = c1 / ab2;
  const double cc1 = cca * u - ccb * v;
  const double cc2 = ccb * u + cca * v;
  // u = (cca cc1 + ccb cc2)/ab2, v = (cca cc2 - ccb cc1)/ab2, ab2 = cca^2 + ccb^2
  const double cc1_b = (u_b * cca - v_b * ccb) * rab2;
  const double cc2_b = (u_b * ccb + v_b * cca) * rab2;
  const double cca_b = (u_b * (cc1 - 2.0 * cca * u) + v_b * (cc2 - 2.0 * cca * v)) * rab2;
  const double ccb_b = (u_b * (cc2 - 2.0 * ccb * u) - v_b * (cc1 + 2.0 * ccb * v)) * rab2;
  const double vrel_b = cca_b * s.cosw + ccb_b * sg * s.sinw + cc1_b * waterx + cc2_b * watery;
  const double Cb_b = cca_b;
  double uold_b = cc1_b * umassdti * s.brlx;
  double vold_b = cc2_b * umassdti * s.brlx;
  if (R > 0.0) {                          // vrel = A |(uocn - uold, vocn - vold)|
    const double t = vrel_b * A / R;
    uold_b -= t * du;
    vold_b -= t * dv;
  }
  if (S > 0.0) {                          // Cb = Tbu / (|(uold, vold)| + u0)
    const double t = -Cb_b * Cb * rden / S;
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
// inputs of stress_point, the adjoint of the updated stresses (sig_b_new[c *
// ss], c = 0..11) and of str1..8 (str_b).  Writes the adjoint of the
// stresses before the update to sig_b_out[c * ss] (which may be sig_b_new:
// every element is read first), the velocity adjoint contributions at cc,
// ee, se, ne to vel_b[8], and returns d/d strength.  The four corners are
// reversed one after the other, each from its own strain rates, and feed the
// velocity accumulators as they go: few values stay live (no per-corner
// arrays), so the kernel keeps several waves per SIMD without spills.
// NT: the adjoint stresses read and written with the cache hints
// (evp1d_cache.hpp).
template <bool NT = false>
__host__ __device__ inline double stress_adjoint(const Scalars &s,
    double uvel_cc, double vvel_cc, double uvel_ee, double vvel_ee,
    double uvel_se, double vvel_se, double uvel_ne, double vvel_ne,
    double dxT, double dyT, double strength,
    double hte, double htn, double htem1, double htnm1,
    const double *sig_b_new, const double *str_b, double *sig_b_out, long ss, double *vel_b) {
  const double cxp = c1p5 * htn - p5 * htnm1;
  const double cyp = c1p5 * hte - p5 * htem1;
  const double cxm = -(c1p5 * htnm1 - p5 * htn);
  const double cym = -(c1p5 * htem1 - p5 * hte);
  const double DminTarea = s.deltaminEVP * (dxT * dyT);
  const double dxhy = p5 * (hte - htem1);
  const double dyhx = p5 * (htn - htnm1);

  // adjoint of the updated stresses including the str1..8 path
  double sb[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) sb[k] = in1<NT>(sig_b_new + k * ss);
  combos_adjoint(dxT, dyT, dxhy, dyhx, str_b, sb);

  const double keep = c1 - s.arlx1i * s.revp;
  const bool cap1 = s.capping == c1;
  double strength_b = 0.0;
  double ucc = 0, vcc = 0, uee = 0, vee = 0, use = 0, vse = 0, une = 0, vne = 0;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    // strain rates and Delta of corner c: 0 ne, 1 nw, 2 sw, 3 se (stress index 1..4)
    double dv, tn, sh;
    if (c == 0) {
      dv = cyp * uvel_cc - dyT * uvel_ee + cxp * vvel_cc - dxT * vvel_se;
      tn = -cym * uvel_cc - dyT * uvel_ee + cxm * vvel_cc + dxT * vvel_se;
      sh = -cym * vvel_cc - dyT * vvel_ee - cxm * uvel_cc - dxT * uvel_se;
    } else if (c == 1) {
      dv = cym * uvel_ee + dyT * uvel_cc + cxp * vvel_ee - dxT * vvel_ne;
      tn = -cyp * uvel_ee + dyT * uvel_cc + cxm * vvel_ee + dxT * vvel_ne;
      sh = -cyp * vvel_ee + dyT * vvel_cc - cxm * uvel_ee - dxT * uvel_ne;
    } else if (c == 2) {
      dv = cym * uvel_ne + dyT * uvel_se + cxm * vvel_ne + dxT * vvel_ee;
      tn = -cyp * uvel_ne + dyT * uvel_se + cxp * vvel_ne - dxT * vvel_ee;
      sh = -cyp * vvel_ne + dyT * vvel_se - cxp * uvel_ne + dxT * uvel_ee;
    } else {
      dv = cyp * uvel_se - dyT * uvel_ne + cxm * vvel_se + dxT * vvel_cc;
      tn = -cym * uvel_se - dyT * uvel_ne + cxp * vvel_se - dxT * vvel_cc;
      sh = -cym * vvel_se - dyT * vvel_ne - cxp * uvel_se + dxT * uvel_cc;
    }
    const double D = sqrt(dv * dv + s.e_factor * (tn * tn + sh * sh));
    // forward: tmpcalc, zetax2, etax2, rep_prs (visc_replpress), with
    // T_s = dT/dstrength, T_D = dT/dDelta and rD = 1/Delta (0 at Delta = 0).
    // capping = 1 (the default): T = strength / max(Delta, DminArea), so one
    // reciprocal gives T, T_s, T_D and, where Delta > DminArea, rD.
    const double mx = (D > DminTarea) ? D : DminTarea;
    double T, T_s, T_D, rD;
    if (cap1) {
      const double rmx = c1 / mx;
      T = strength * rmx;
      T_s = rmx;
      T_D = (D > DminTarea) ? -T * rmx : 0.0;
      rD = (D > DminTarea) ? rmx : (D > 0.0 ? c1 / D : 0.0);
    } else {
      const double dsum = D + DminTarea;
      T = s.capping * (strength / mx) + (c1 - s.capping) * (strength / dsum);
      T_s = s.capping / mx + (c1 - s.capping) / dsum;
      T_D = s.capping * ((D > DminTarea) ? -strength / (D * D) : 0.0) +
            (c1 - s.capping) * (-strength / (dsum * dsum));
      rD = D > 0.0 ? c1 / D : 0.0;
    }
    const double zeta = (c1 + s.Ktens) * T;
    const double eta = s.epp2i * zeta;
    // sp = (sp0*keep + arlx1i*(zeta*div - rep))*denom1, rep = (1-K)*T*D
    // sm = (sm0*keep + arlx1i*eta*ten)*denom1, s12 = (s120*keep + arlx1i*p5*eta*shr)*denom1
    const double P = sb[c] * s.denom1, M = sb[4 + c] * s.denom1, Q = sb[8 + c] * s.denom1;
    out1<NT>(sig_b_out + c * ss, P * keep);
    out1<NT>(sig_b_out + (4 + c) * ss, M * keep);
    out1<NT>(sig_b_out + (8 + c) * ss, Q * keep);
    const double zeta_b0 = P * s.arlx1i * dv;
    const double rep_b = -P * s.arlx1i;
    const double eta_b = M * s.arlx1i * tn + Q * s.arlx1i * p5 * sh;
    double dv_b = P * s.arlx1i * zeta;
    double tn_b = M * s.arlx1i * eta;
    double sh_b = Q * s.arlx1i * p5 * eta;
    const double zeta_b = zeta_b0 + s.epp2i * eta_b;
    const double T_b = (c1 + s.Ktens) * zeta_b + (c1 - s.Ktens) * D * rep_b;
    double D_b = (c1 - s.Ktens) * T * rep_b;
    strength_b += T_b * T_s;
    D_b += T_b * T_D;
    if (D > 0.0) {
      const double w = D_b * rD;
      dv_b += w * dv;
      tn_b += w * s.e_factor * tn;
      sh_b += w * s.e_factor * sh;
    }
    // reverse of this corner's strain rates (linear in the 8 velocities)
    if (c == 0) {
      ucc += cyp * dv_b; uee -= dyT * dv_b; vcc += cxp * dv_b; vse -= dxT * dv_b;
      ucc -= cym * tn_b; uee -= dyT * tn_b; vcc += cxm * tn_b; vse += dxT * tn_b;
      vcc -= cym * sh_b; vee -= dyT * sh_b; ucc -= cxm * sh_b; use -= dxT * sh_b;
    } else if (c == 1) {
      uee += cym * dv_b; ucc += dyT * dv_b; vee += cxp * dv_b; vne -= dxT * dv_b;
      uee -= cyp * tn_b; ucc += dyT * tn_b; vee += cxm * tn_b; vne += dxT * tn_b;
      vee -= cyp * sh_b; vcc += dyT * sh_b; uee -= cxm * sh_b; une -= dxT * sh_b;
    } else if (c == 2) {
      une += cym * dv_b; use += dyT * dv_b; vne += cxm * dv_b; vee += dxT * dv_b;
      une -= cyp * tn_b; use += dyT * tn_b; vne += cxp * tn_b; vee -= dxT * tn_b;
      vne -= cyp * sh_b; vse += dyT * sh_b; une -= cxp * sh_b; uee += dxT * sh_b;
    } else {
      use += cyp * dv_b; une -= dyT * dv_b; vse += cxm * dv_b; vcc += dxT * dv_b;
      use -= cym * tn_b; une -= dyT * tn_b; vse += cxp * tn_b; vcc -= dxT * tn_b;
      vse -= cym * sh_b; vne -= dyT * sh_b; use -= cxp * sh_b; ucc += dxT * sh_b;
    }
  }
  vel_b[0] = ucc; vel_b[1] = vcc;
  vel_b[2] = uee; vel_b[3] = vee;
  vel_b[4] = use; vel_b[5] = vse;
  vel_b[6] = une; vel_b[7] = vne;
  return strength_b;
}

}  // namespace evp1d
