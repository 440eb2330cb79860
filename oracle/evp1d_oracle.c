/*
 * evp1d_oracle.c - CPU restatement of the reference EVP1D kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the GPU library:
 * only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * load it.  No product code path calls it.
 *
 * It restates, statement by statement and with the same association, the
 * serial reference kernels in src/bench_v2b.F90:
 *   oracle_stress  <- stress         bench_v2b.F90:125-311
 *                     strain_rates   bench_v2b.F90:321-397
 *                     visc_replpress bench_v2b.F90:415-437
 *   oracle_stepu   <- stepu          bench_v2b.F90:505-547
 *   oracle_run     <- the subcycle loop of src/evp_v2b.F90:91-110
 *   oracle_set_evp_parameters <- set_evp_parameters, src/ice.F90:109-130
 * Build: -O2 -ffp-contract=off (no FMA, no reassociation, SSE2 scalar
 * double) - the analogue of the reference's f0 flag set.  Built with -fopenmp
 * the two loops split over threads; results do not change because every
 * point only writes its own slots.
 *
 * Arrays are 0-based in C; neighbour indices are the reference's 1-based
 * values and are used as x[idx-1].  Logical skip flags are int32 (nonzero =
 * skip).  The 12 stresses are passed as an array of 12 pointers in the order
 * stressp_1..4, stressm_1..4, stress12_1..4; str as 8 pointers str1..str8.
 *
 * Pinning: bit-identical to the flang-compiled reference v2b (oracle/Makefile
 * target `ref`) on the committed golden cases and the 400x300 workload (md5
 * of both output files), plus the closed-form checks in tests/test_oracle.py.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

/* ice_constants (src/ice.F90:25-68) */
static const double c0 = 0.0, c1 = 1.0, c2 = 2.0, c1p5 = 1.5, p5 = 0.5, p25 = 0.25;
#define P111 (1.0 / 9.0)
#define P166 (1.0 / 6.0)
#define P222 (2.0 / 9.0)
#define P333 (1.0 / 3.0)
#define P055 (P111 * 0.5)
#define P027 (P055 * 0.5)

/* scalar block, same order as evp1d_params in include/evp1d.h */
typedef struct {
  int32_t ndte, reserved;
  double capping, e_factor, epp2i, deltaminEVP;
  double arlx, arlx1i, brlx, denom1;
  double revp, Ktens, u0, cosw, sinw, rhow;
} oracle_params;

/* src/ice.F90:80-107 and 109-130 */
void oracle_set_evp_parameters(int32_t ndte, oracle_params *p) {
  const double elasticDamp = 0.36, e_plasticpot = c2, e_yieldcurve = c2;
  memset(p, 0, sizeof(*p));
  p->ndte = ndte;
  p->u0 = 5e-5;
  p->cosw = c1;
  p->sinw = c0;
  p->Ktens = c0;
  p->revp = c0;
  p->rhow = 1026.0;
  p->epp2i = c1 / pow(e_plasticpot, 2);
  p->capping = c1;
  p->e_factor = pow(e_yieldcurve, 2) / pow(e_plasticpot, 4);
  p->arlx = c2 * elasticDamp * (double)ndte;
  p->arlx1i = c1 / p->arlx;
  p->brlx = (double)ndte;
  p->denom1 = c1 / (c1 + p->arlx1i);
  p->deltaminEVP = 10e-9;
}

/* bench_v2b.F90:415-437 */
static void visc_replpress(const oracle_params *p, double strength, double DminArea, double Delta,
                           double *zetax2, double *etax2, double *rep_prs) {
  double tmpcalc;
  tmpcalc = p->capping * (strength / fmax(Delta, DminArea)) +
            (c1 - p->capping) * (strength / (Delta + DminArea));
  *zetax2 = (c1 + p->Ktens) * tmpcalc;
  *rep_prs = (c1 - p->Ktens) * tmpcalc * Delta;
  *etax2 = p->epp2i * *zetax2;
}

void oracle_stress(const oracle_params *p, int32_t lb, int32_t ub,
                   const int32_t *ee, const int32_t *ne, const int32_t *se,
                   const double *uvel, const double *vvel, const double *dxT, const double *dyT,
                   const int32_t *skipme, const double *strength,
                   const double *hte, const double *htn, const double *htem1, const double *htnm1,
                   double *const *sig, double *const *str) {
  double *stressp_1 = sig[0], *stressp_2 = sig[1], *stressp_3 = sig[2], *stressp_4 = sig[3];
  double *stressm_1 = sig[4], *stressm_2 = sig[5], *stressm_3 = sig[6], *stressm_4 = sig[7];
  double *stress12_1 = sig[8], *stress12_2 = sig[9], *stress12_3 = sig[10], *stress12_4 = sig[11];
  double *str1 = str[0], *str2 = str[1], *str3 = str[2], *str4 = str[3];
  double *str5 = str[4], *str6 = str[5], *str7 = str[6], *str8 = str[7];
  const double arlx1i = p->arlx1i, denom1 = p->denom1, revp = p->revp, e_factor = p->e_factor;
  int32_t iw;
#pragma omp parallel for schedule(static)
  for (iw = lb; iw <= ub; ++iw) {
    const int32_t w = iw - 1;
    double tmp_uvel_cc, tmp_vvel_cc, tmp_uvel_ee, tmp_vvel_se, tmp_vvel_ee, tmp_vvel_ne, tmp_uvel_ne,
        tmp_uvel_se, tmp_dxT, tmp_dyT, tmp_cxp, tmp_cyp, tmp_cxm, tmp_cym, tmp_strength, tmparea,
        tmp_DminTarea, tmp_dxhy, tmp_dyhx;
    double divune, divunw, divuse, divusw, tensionne, tensionnw, tensionse, tensionsw;
    double shearne, shearnw, shearse, shearsw, Deltane, Deltanw, Deltase, Deltasw;
    double zetax2ne, zetax2nw, zetax2se, zetax2sw, etax2ne, etax2nw, etax2se, etax2sw;
    double rep_prsne, rep_prsnw, rep_prsse, rep_prssw;
    double ssigpn, ssigps, ssigpe, ssigpw, ssigmn, ssigms, ssigme, ssigmw;
    double ssig12n, ssig12s, ssig12e, ssig12w, ssigp1, ssigp2, ssigm1, ssigm2, ssig121, ssig122;
    double csigpne, csigpnw, csigpse, csigpsw, csigmne, csigmnw, csigmse, csigmsw;
    double csig12ne, csig12nw, csig12se, csig12sw, str12ew, str12we, str12ns, str12sn;
    double strp_tmp, strm_tmp;
    if (skipme[w]) continue;
    tmp_uvel_cc = uvel[w];
    tmp_vvel_cc = vvel[w];
    tmp_uvel_ee = uvel[ee[w] - 1];
    tmp_vvel_se = vvel[se[w] - 1];
    tmp_vvel_ee = vvel[ee[w] - 1];
    tmp_vvel_ne = vvel[ne[w] - 1];
    tmp_uvel_ne = uvel[ne[w] - 1];
    tmp_uvel_se = uvel[se[w] - 1];
    tmp_dxT = dxT[w];
    tmp_dyT = dyT[w];
    tmp_cxp = c1p5 * htn[w] - p5 * htnm1[w];
    tmp_cyp = c1p5 * hte[w] - p5 * htem1[w];
    tmp_cxm = -(c1p5 * htnm1[w] - p5 * htn[w]);
    tmp_cym = -(c1p5 * htem1[w] - p5 * hte[w]);
    tmp_strength = strength[w];
    tmparea = dxT[w] * dyT[w];
    tmp_DminTarea = p->deltaminEVP * tmparea;
    tmp_dxhy = p5 * (hte[w] - htem1[w]);
    tmp_dyhx = p5 * (htn[w] - htnm1[w]);

    /* strain_rates (bench_v2b.F90:363-395); Fortran -a*b is -(a*b) */
    divune = tmp_cyp * tmp_uvel_cc - tmp_dyT * tmp_uvel_ee + tmp_cxp * tmp_vvel_cc - tmp_dxT * tmp_vvel_se;
    divunw = tmp_cym * tmp_uvel_ee + tmp_dyT * tmp_uvel_cc + tmp_cxp * tmp_vvel_ee - tmp_dxT * tmp_vvel_ne;
    divusw = tmp_cym * tmp_uvel_ne + tmp_dyT * tmp_uvel_se + tmp_cxm * tmp_vvel_ne + tmp_dxT * tmp_vvel_ee;
    divuse = tmp_cyp * tmp_uvel_se - tmp_dyT * tmp_uvel_ne + tmp_cxm * tmp_vvel_se + tmp_dxT * tmp_vvel_cc;
    tensionne = -(tmp_cym * tmp_uvel_cc) - tmp_dyT * tmp_uvel_ee + tmp_cxm * tmp_vvel_cc + tmp_dxT * tmp_vvel_se;
    tensionnw = -(tmp_cyp * tmp_uvel_ee) + tmp_dyT * tmp_uvel_cc + tmp_cxm * tmp_vvel_ee + tmp_dxT * tmp_vvel_ne;
    tensionsw = -(tmp_cyp * tmp_uvel_ne) + tmp_dyT * tmp_uvel_se + tmp_cxp * tmp_vvel_ne - tmp_dxT * tmp_vvel_ee;
    tensionse = -(tmp_cym * tmp_uvel_se) - tmp_dyT * tmp_uvel_ne + tmp_cxp * tmp_vvel_se - tmp_dxT * tmp_vvel_cc;
    shearne = -(tmp_cym * tmp_vvel_cc) - tmp_dyT * tmp_vvel_ee - tmp_cxm * tmp_uvel_cc - tmp_dxT * tmp_uvel_se;
    shearnw = -(tmp_cyp * tmp_vvel_ee) + tmp_dyT * tmp_vvel_cc - tmp_cxm * tmp_uvel_ee - tmp_dxT * tmp_uvel_ne;
    shearsw = -(tmp_cyp * tmp_vvel_ne) + tmp_dyT * tmp_vvel_se - tmp_cxp * tmp_uvel_ne + tmp_dxT * tmp_uvel_ee;
    shearse = -(tmp_cym * tmp_vvel_se) - tmp_dyT * tmp_vvel_ne - tmp_cxp * tmp_uvel_se + tmp_dxT * tmp_uvel_cc;
    Deltane = sqrt(divune * divune + e_factor * (tensionne * tensionne + shearne * shearne));
    Deltanw = sqrt(divunw * divunw + e_factor * (tensionnw * tensionnw + shearnw * shearnw));
    Deltasw = sqrt(divusw * divusw + e_factor * (tensionsw * tensionsw + shearsw * shearsw));
    Deltase = sqrt(divuse * divuse + e_factor * (tensionse * tensionse + shearse * shearse));

    visc_replpress(p, tmp_strength, tmp_DminTarea, Deltane, &zetax2ne, &etax2ne, &rep_prsne);
    visc_replpress(p, tmp_strength, tmp_DminTarea, Deltanw, &zetax2nw, &etax2nw, &rep_prsnw);
    visc_replpress(p, tmp_strength, tmp_DminTarea, Deltasw, &zetax2sw, &etax2sw, &rep_prssw);
    visc_replpress(p, tmp_strength, tmp_DminTarea, Deltase, &zetax2se, &etax2se, &rep_prsse);

    stressp_1[w] = (stressp_1[w] * (c1 - arlx1i * revp) + arlx1i * (zetax2ne * divune - rep_prsne)) * denom1;
    stressp_2[w] = (stressp_2[w] * (c1 - arlx1i * revp) + arlx1i * (zetax2nw * divunw - rep_prsnw)) * denom1;
    stressp_3[w] = (stressp_3[w] * (c1 - arlx1i * revp) + arlx1i * (zetax2sw * divusw - rep_prssw)) * denom1;
    stressp_4[w] = (stressp_4[w] * (c1 - arlx1i * revp) + arlx1i * (zetax2se * divuse - rep_prsse)) * denom1;
    stressm_1[w] = (stressm_1[w] * (c1 - arlx1i * revp) + arlx1i * etax2ne * tensionne) * denom1;
    stressm_2[w] = (stressm_2[w] * (c1 - arlx1i * revp) + arlx1i * etax2nw * tensionnw) * denom1;
    stressm_3[w] = (stressm_3[w] * (c1 - arlx1i * revp) + arlx1i * etax2sw * tensionsw) * denom1;
    stressm_4[w] = (stressm_4[w] * (c1 - arlx1i * revp) + arlx1i * etax2se * tensionse) * denom1;
    stress12_1[w] = (stress12_1[w] * (c1 - arlx1i * revp) + arlx1i * p5 * etax2ne * shearne) * denom1;
    stress12_2[w] = (stress12_2[w] * (c1 - arlx1i * revp) + arlx1i * p5 * etax2nw * shearnw) * denom1;
    stress12_3[w] = (stress12_3[w] * (c1 - arlx1i * revp) + arlx1i * p5 * etax2sw * shearsw) * denom1;
    stress12_4[w] = (stress12_4[w] * (c1 - arlx1i * revp) + arlx1i * p5 * etax2se * shearse) * denom1;

    ssigpn = stressp_1[w] + stressp_2[w];
    ssigps = stressp_3[w] + stressp_4[w];
    ssigpe = stressp_1[w] + stressp_4[w];
    ssigpw = stressp_2[w] + stressp_3[w];
    ssigp1 = (stressp_1[w] + stressp_3[w]) * P055;
    ssigp2 = (stressp_2[w] + stressp_4[w]) * P055;
    ssigmn = stressm_1[w] + stressm_2[w];
    ssigms = stressm_3[w] + stressm_4[w];
    ssigme = stressm_1[w] + stressm_4[w];
    ssigmw = stressm_2[w] + stressm_3[w];
    ssigm1 = (stressm_1[w] + stressm_3[w]) * P055;
    ssigm2 = (stressm_2[w] + stressm_4[w]) * P055;
    ssig12n = stress12_1[w] + stress12_2[w];
    ssig12s = stress12_3[w] + stress12_4[w];
    ssig12e = stress12_1[w] + stress12_4[w];
    ssig12w = stress12_2[w] + stress12_3[w];
    ssig121 = (stress12_1[w] + stress12_3[w]) * P111;
    ssig122 = (stress12_2[w] + stress12_4[w]) * P111;

    csigpne = P111 * stressp_1[w] + ssigp2 + P027 * stressp_3[w];
    csigpnw = P111 * stressp_2[w] + ssigp1 + P027 * stressp_4[w];
    csigpsw = P111 * stressp_3[w] + ssigp2 + P027 * stressp_1[w];
    csigpse = P111 * stressp_4[w] + ssigp1 + P027 * stressp_2[w];
    csigmne = P111 * stressm_1[w] + ssigm2 + P027 * stressm_3[w];
    csigmnw = P111 * stressm_2[w] + ssigm1 + P027 * stressm_4[w];
    csigmsw = P111 * stressm_3[w] + ssigm2 + P027 * stressm_1[w];
    csigmse = P111 * stressm_4[w] + ssigm1 + P027 * stressm_2[w];
    csig12ne = P222 * stress12_1[w] + ssig122 + P055 * stress12_3[w];
    csig12nw = P222 * stress12_2[w] + ssig121 + P055 * stress12_4[w];
    csig12sw = P222 * stress12_3[w] + ssig122 + P055 * stress12_1[w];
    csig12se = P222 * stress12_4[w] + ssig121 + P055 * stress12_2[w];

    str12ew = p5 * tmp_dxT * (P333 * ssig12e + P166 * ssig12w);
    str12we = p5 * tmp_dxT * (P333 * ssig12w + P166 * ssig12e);
    str12ns = p5 * tmp_dyT * (P333 * ssig12n + P166 * ssig12s);
    str12sn = p5 * tmp_dyT * (P333 * ssig12s + P166 * ssig12n);

    strp_tmp = p25 * tmp_dyT * (P333 * ssigpn + P166 * ssigps);
    strm_tmp = p25 * tmp_dyT * (P333 * ssigmn + P166 * ssigms);
    str1[w] = -strp_tmp - strm_tmp - str12ew + tmp_dxhy * (-csigpne + csigmne) + tmp_dyhx * csig12ne;
    str2[w] = strp_tmp + strm_tmp - str12we + tmp_dxhy * (-csigpnw + csigmnw) + tmp_dyhx * csig12nw;
    strp_tmp = p25 * tmp_dyT * (P333 * ssigps + P166 * ssigpn);
    strm_tmp = p25 * tmp_dyT * (P333 * ssigms + P166 * ssigmn);
    str3[w] = -strp_tmp - strm_tmp + str12ew + tmp_dxhy * (-csigpse + csigmse) + tmp_dyhx * csig12se;
    str4[w] = strp_tmp + strm_tmp + str12we + tmp_dxhy * (-csigpsw + csigmsw) + tmp_dyhx * csig12sw;
    strp_tmp = p25 * tmp_dxT * (P333 * ssigpe + P166 * ssigpw);
    strm_tmp = p25 * tmp_dxT * (P333 * ssigme + P166 * ssigmw);
    str5[w] = -strp_tmp + strm_tmp - str12ns - tmp_dyhx * (csigpne + csigmne) + tmp_dxhy * csig12ne;
    str6[w] = strp_tmp - strm_tmp - str12sn - tmp_dyhx * (csigpse + csigmse) + tmp_dxhy * csig12se;
    strp_tmp = p25 * tmp_dxT * (P333 * ssigpw + P166 * ssigpe);
    strm_tmp = p25 * tmp_dxT * (P333 * ssigmw + P166 * ssigme);
    str7[w] = -strp_tmp + strm_tmp + str12ns - tmp_dyhx * (csigpnw + csigmnw) + tmp_dxhy * csig12nw;
    str8[w] = strp_tmp - strm_tmp + str12sn - tmp_dyhx * (csigpsw + csigmsw) + tmp_dxhy * csig12sw;
  }
}

/* bench_v2b.F90:505-547.  Cw = cdn_ocn, aiX = aiu; stepu's rhow is its
 * local parameter 1026 (bench_v2b.F90:489-490), passed in p->rhow. */
void oracle_stepu(const oracle_params *p, int32_t lb, int32_t ub,
                  const double *Cw, const double *aiX, const double *uocn, const double *vocn,
                  const double *waterx, const double *watery, const double *forcex, const double *forcey,
                  const double *umassdti, const double *fm, const double *uarear,
                  double *strintx, double *strinty, const double *uvel_init, const double *vvel_init,
                  double *uvel, double *vvel, double *const *str,
                  const int32_t *nw, const int32_t *sw, const int32_t *sse, const int32_t *skipme,
                  const double *Tbu, double *Cb) {
  const double *str1 = str[0], *str2 = str[1], *str3 = str[2], *str4 = str[3];
  const double *str5 = str[4], *str6 = str[5], *str7 = str[6], *str8 = str[7];
  const double brlx = p->brlx, revp = p->revp, u0 = p->u0, cosw = p->cosw, sinw = p->sinw;
  const double rhow = p->rhow;
  int32_t iw;
#pragma omp parallel for schedule(static)
  for (iw = lb; iw <= ub; ++iw) {
    const int32_t w = iw - 1;
    double uold, vold, vrel, cca, ccb, ab2, cc1, cc2, taux, tauy;
    double tmp_str2_nw, tmp_str3_sse, tmp_str4_sw, tmp_str6_sse, tmp_str7_nw, tmp_str8_sw;
    if (skipme[w]) continue;
    uold = uvel[w];
    vold = vvel[w];
    vrel = aiX[w] * rhow * Cw[w] * sqrt((uocn[w] - uold) * (uocn[w] - uold) + (vocn[w] - vold) * (vocn[w] - vold));
    taux = vrel * waterx[w];
    tauy = vrel * watery[w];
    Cb[w] = Tbu[w] / (sqrt(uold * uold + vold * vold) + u0);
    cca = (brlx + revp) * umassdti[w] + vrel * cosw + Cb[w];
    ccb = fm[w] + copysign(c1, fm[w]) * vrel * sinw;
    ab2 = cca * cca + ccb * ccb;
    tmp_str2_nw = str2[nw[w] - 1];
    tmp_str3_sse = str3[sse[w] - 1];
    tmp_str4_sw = str4[sw[w] - 1];
    tmp_str6_sse = str6[sse[w] - 1];
    tmp_str7_nw = str7[nw[w] - 1];
    tmp_str8_sw = str8[sw[w] - 1];
    strintx[w] = uarear[w] * (str1[w] + tmp_str2_nw + tmp_str3_sse + tmp_str4_sw);
    strinty[w] = uarear[w] * (str5[w] + tmp_str6_sse + tmp_str7_nw + tmp_str8_sw);
    cc1 = strintx[w] + forcex[w] + taux + umassdti[w] * (brlx * uold + revp * uvel_init[w]);
    cc2 = strinty[w] + forcey[w] + tauy + umassdti[w] * (brlx * vold + revp * vvel_init[w]);
    uvel[w] = (cca * cc1 + ccb * cc2) / ab2;
    vvel[w] = (cca * cc2 - ccb * cc1) / ab2;
  }
}

/* 1D copy-list halo (src/vars.F90:100-143 in 1D form): ghost dst[k] takes
 * the value of slot src[k], or 0 when src[k] == 0.  1-based. */
void oracle_halo(int32_t n, const int32_t *dst, const int32_t *src, double *uvel, double *vvel) {
  int32_t k;
  for (k = 0; k < n; ++k) {
    uvel[dst[k] - 1] = src[k] > 0 ? uvel[src[k] - 1] : c0;
    vvel[dst[k] - 1] = src[k] > 0 ? vvel[src[k] - 1] : c0;
  }
}

/* Field bundle for oracle_run: pointers in the order of include/evp1d.h
 * enum evp1d_field (EVP1D_STRENGTH .. EVP1D_STR8). */
enum {
  F_STRENGTH = 0, F_UVEL, F_VVEL, F_DXT, F_DYT, F_UAREAR, F_CDN_OCN, F_AIU, F_UOCN, F_VOCN,
  F_WATERX, F_WATERY, F_FORCEX, F_FORCEY, F_UMASSDTI, F_FM, F_STRINTX, F_STRINTY, F_TBU,
  F_SIG0, F_HTE = F_SIG0 + 12, F_HTN, F_HTEM1, F_HTNM1, F_UVEL_INIT, F_VVEL_INIT, F_CB, F_STR1,
  F_NFIELDS = F_STR1 + 8
};

/* nsub subcycles of stress; stepu; halo over 1..na (src/evp_v2b.F90:91-110) */
void oracle_run(const oracle_params *p, int32_t na, int32_t nsub, double *const *f,
                const int32_t *const *nbr /* ee,ne,se,nw,sw,sse */, const int32_t *skipT,
                const int32_t *skipU, int32_t nhalo, const int32_t *hdst, const int32_t *hsrc) {
  int32_t i;
  for (i = 1; i <= nsub; ++i) {
    oracle_stress(p, 1, na, nbr[0], nbr[1], nbr[2], f[F_UVEL], f[F_VVEL], f[F_DXT], f[F_DYT], skipT,
                  f[F_STRENGTH], f[F_HTE], f[F_HTN], f[F_HTEM1], f[F_HTNM1], f + F_SIG0, f + F_STR1);
    oracle_stepu(p, 1, na, f[F_CDN_OCN], f[F_AIU], f[F_UOCN], f[F_VOCN], f[F_WATERX], f[F_WATERY],
                 f[F_FORCEX], f[F_FORCEY], f[F_UMASSDTI], f[F_FM], f[F_UAREAR], f[F_STRINTX],
                 f[F_STRINTY], f[F_UVEL_INIT], f[F_VVEL_INIT], f[F_UVEL], f[F_VVEL], f + F_STR1,
                 nbr[3], nbr[4], nbr[5], skipU, f[F_TBU], f[F_CB]);
    if (nhalo > 0) oracle_halo(nhalo, hdst, hsrc, f[F_UVEL], f[F_VVEL]);
  }
}
