/*
 * evp2d_oracle.c - CPU restatement of the reference's v0 2D block kernels.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the GPU 2D block path, like
 * evp1d_oracle.c for the 1D path); only tests/ load it.
 *
 * Statement by statement, with the reference's association, from
 * src/bench_v0.F90 and src/vars.F90:
 *   oracle_v0_stress <- stress          bench_v0.F90:62-331 (str zeroed, 135)
 *                       strain_rates    bench_v0.F90:342-423
 *                       visc_replpress  bench_v0.F90:441-465
 *   oracle_v0_stepu  <- stepu           bench_v0.F90:490-602
 *   oracle_v0_halo   <- ice_haloUpdate  vars.F90:100-143
 *   oracle_v0_run    <- the subcycle loop of evp_v0.F90:93-136
 *
 * Arrays are one block of (nx_block, ny_block) in Fortran order: cell (i, j),
 * 1-based, is element (i-1) + (j-1)*nx_block.  Cell lists indxti/indxtj are
 * the reference's 1-based values.  The 12 stresses are 12 pointers
 * (stressp_1..4, stressm_1..4, stress12_1..4); str is one scratch block of
 * 8 planes (str(:,:,1..8)).
 *
 * Pinning: bit-identical to the flang-compiled reference v0 (oracle/Makefile
 * target `ref`, _ref/evp_v0_f0) on the committed 2D golden cases (md5 of both
 * output files), and to the 1D restatement (evp1d_oracle.c) on the same grid
 * (tests/test_oracle_v0.py).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

static const double c0 = 0.0, c1 = 1.0, p5 = 0.5, p25 = 0.25;
#define P111 (1.0 / 9.0)
#define P166 (1.0 / 6.0)
#define P222 (2.0 / 9.0)
#define P333 (1.0 / 3.0)
#define P055 (P111 * 0.5)
#define P027 (P055 * 0.5)

/* same layout as oracle_params in evp1d_oracle.c and evp1d_params */
typedef struct {
  int32_t ndte, reserved;
  double capping, e_factor, epp2i, deltaminEVP;
  double arlx, arlx1i, brlx, denom1;
  double revp, Ktens, u0, cosw, sinw, rhow;
} oracle_params;

#define AT(a, i, j) (a)[(size_t)((i) - 1) + (size_t)((j) - 1) * (size_t)nxb]

/* bench_v0.F90:441-465 */
static void visc_replpress_v0(const oracle_params *p, double strength, double DminArea, double Delta,
                              double *zetax2, double *etax2, double *rep_prs) {
  double tmpcalc;
  tmpcalc = p->capping * (strength / fmax(Delta, DminArea)) +
            (c1 - p->capping) * (strength / (Delta + DminArea));
  *zetax2 = (c1 + p->Ktens) * tmpcalc;
  *rep_prs = (c1 - p->Ktens) * tmpcalc * Delta;
  *etax2 = p->epp2i * *zetax2;
}

void oracle_v0_stress(const oracle_params *p, int32_t nxb, int32_t nyb, int32_t icellt,
                      const int32_t *indxti, const int32_t *indxtj, const double *uvel, const double *vvel,
                      const double *dxT, const double *dyT, const double *dxhy, const double *dyhx,
                      const double *cxp, const double *cyp, const double *cxm, const double *cym,
                      const double *DminTarea, const double *strength, double *const *sig, double *str) {
  double *stressp_1 = sig[0], *stressp_2 = sig[1], *stressp_3 = sig[2], *stressp_4 = sig[3];
  double *stressm_1 = sig[4], *stressm_2 = sig[5], *stressm_3 = sig[6], *stressm_4 = sig[7];
  double *stress12_1 = sig[8], *stress12_2 = sig[9], *stress12_3 = sig[10], *stress12_4 = sig[11];
  const size_t plane = (size_t)nxb * (size_t)nyb;
  double *s1 = str, *s2 = str + plane, *s3 = str + 2 * plane, *s4 = str + 3 * plane;
  double *s5 = str + 4 * plane, *s6 = str + 5 * plane, *s7 = str + 6 * plane, *s8 = str + 7 * plane;
  const double arlx1i = p->arlx1i, denom1 = p->denom1, revp = p->revp, e_factor = p->e_factor;
  int32_t ij;
  memset(str, 0, 8 * plane * sizeof(double)); /* str(:,:,:) = c0 */
  for (ij = 0; ij < icellt; ++ij) {
    const int32_t i = indxti[ij], j = indxtj[ij];
    double divune, divunw, divuse, divusw, tensionne, tensionnw, tensionse, tensionsw;
    double shearne, shearnw, shearse, shearsw, Deltane, Deltanw, Deltase, Deltasw;
    double zetax2ne, zetax2nw, zetax2se, zetax2sw, etax2ne, etax2nw, etax2se, etax2sw;
    double rep_prsne, rep_prsnw, rep_prsse, rep_prssw;
    double ssigpn, ssigps, ssigpe, ssigpw, ssigmn, ssigms, ssigme, ssigmw;
    double ssig12n, ssig12s, ssig12e, ssig12w, ssigp1, ssigp2, ssigm1, ssigm2, ssig121, ssig122;
    double csigpne, csigpnw, csigpse, csigpsw, csigmne, csigmnw, csigmse, csigmsw;
    double csig12ne, csig12nw, csig12se, csig12sw, str12ew, str12we, str12ns, str12sn;
    double strp_tmp, strm_tmp;
    /* strain_rates (bench_v0.F90:386-419); Fortran -a*b is -(a*b) */
    divune = AT(cyp, i, j) * AT(uvel, i, j) - AT(dyT, i, j) * AT(uvel, i - 1, j) +
             AT(cxp, i, j) * AT(vvel, i, j) - AT(dxT, i, j) * AT(vvel, i, j - 1);
    divunw = AT(cym, i, j) * AT(uvel, i - 1, j) + AT(dyT, i, j) * AT(uvel, i, j) +
             AT(cxp, i, j) * AT(vvel, i - 1, j) - AT(dxT, i, j) * AT(vvel, i - 1, j - 1);
    divusw = AT(cym, i, j) * AT(uvel, i - 1, j - 1) + AT(dyT, i, j) * AT(uvel, i, j - 1) +
             AT(cxm, i, j) * AT(vvel, i - 1, j - 1) + AT(dxT, i, j) * AT(vvel, i - 1, j);
    divuse = AT(cyp, i, j) * AT(uvel, i, j - 1) - AT(dyT, i, j) * AT(uvel, i - 1, j - 1) +
             AT(cxm, i, j) * AT(vvel, i, j - 1) + AT(dxT, i, j) * AT(vvel, i, j);
    tensionne = -(AT(cym, i, j) * AT(uvel, i, j)) - AT(dyT, i, j) * AT(uvel, i - 1, j) +
                AT(cxm, i, j) * AT(vvel, i, j) + AT(dxT, i, j) * AT(vvel, i, j - 1);
    tensionnw = -(AT(cyp, i, j) * AT(uvel, i - 1, j)) + AT(dyT, i, j) * AT(uvel, i, j) +
                AT(cxm, i, j) * AT(vvel, i - 1, j) + AT(dxT, i, j) * AT(vvel, i - 1, j - 1);
    tensionsw = -(AT(cyp, i, j) * AT(uvel, i - 1, j - 1)) + AT(dyT, i, j) * AT(uvel, i, j - 1) +
                AT(cxp, i, j) * AT(vvel, i - 1, j - 1) - AT(dxT, i, j) * AT(vvel, i - 1, j);
    tensionse = -(AT(cym, i, j) * AT(uvel, i, j - 1)) - AT(dyT, i, j) * AT(uvel, i - 1, j - 1) +
                AT(cxp, i, j) * AT(vvel, i, j - 1) - AT(dxT, i, j) * AT(vvel, i, j);
    shearne = -(AT(cym, i, j) * AT(vvel, i, j)) - AT(dyT, i, j) * AT(vvel, i - 1, j) -
              AT(cxm, i, j) * AT(uvel, i, j) - AT(dxT, i, j) * AT(uvel, i, j - 1);
    shearnw = -(AT(cyp, i, j) * AT(vvel, i - 1, j)) + AT(dyT, i, j) * AT(vvel, i, j) -
              AT(cxm, i, j) * AT(uvel, i - 1, j) - AT(dxT, i, j) * AT(uvel, i - 1, j - 1);
    shearsw = -(AT(cyp, i, j) * AT(vvel, i - 1, j - 1)) + AT(dyT, i, j) * AT(vvel, i, j - 1) -
              AT(cxp, i, j) * AT(uvel, i - 1, j - 1) + AT(dxT, i, j) * AT(uvel, i - 1, j);
    shearse = -(AT(cym, i, j) * AT(vvel, i, j - 1)) - AT(dyT, i, j) * AT(vvel, i - 1, j - 1) -
              AT(cxp, i, j) * AT(uvel, i, j - 1) + AT(dxT, i, j) * AT(uvel, i, j);
    Deltane = sqrt(divune * divune + e_factor * (tensionne * tensionne + shearne * shearne));
    Deltanw = sqrt(divunw * divunw + e_factor * (tensionnw * tensionnw + shearnw * shearnw));
    Deltasw = sqrt(divusw * divusw + e_factor * (tensionsw * tensionsw + shearsw * shearsw));
    Deltase = sqrt(divuse * divuse + e_factor * (tensionse * tensionse + shearse * shearse));

    visc_replpress_v0(p, AT(strength, i, j), AT(DminTarea, i, j), Deltane, &zetax2ne, &etax2ne, &rep_prsne);
    visc_replpress_v0(p, AT(strength, i, j), AT(DminTarea, i, j), Deltanw, &zetax2nw, &etax2nw, &rep_prsnw);
    visc_replpress_v0(p, AT(strength, i, j), AT(DminTarea, i, j), Deltasw, &zetax2sw, &etax2sw, &rep_prssw);
    visc_replpress_v0(p, AT(strength, i, j), AT(DminTarea, i, j), Deltase, &zetax2se, &etax2se, &rep_prsse);

    AT(stressp_1, i, j) = (AT(stressp_1, i, j) * (c1 - arlx1i * revp) + arlx1i * (zetax2ne * divune - rep_prsne)) * denom1;
    AT(stressp_2, i, j) = (AT(stressp_2, i, j) * (c1 - arlx1i * revp) + arlx1i * (zetax2nw * divunw - rep_prsnw)) * denom1;
    AT(stressp_3, i, j) = (AT(stressp_3, i, j) * (c1 - arlx1i * revp) + arlx1i * (zetax2sw * divusw - rep_prssw)) * denom1;
    AT(stressp_4, i, j) = (AT(stressp_4, i, j) * (c1 - arlx1i * revp) + arlx1i * (zetax2se * divuse - rep_prsse)) * denom1;
    AT(stressm_1, i, j) = (AT(stressm_1, i, j) * (c1 - arlx1i * revp) + arlx1i * etax2ne * tensionne) * denom1;
    AT(stressm_2, i, j) = (AT(stressm_2, i, j) * (c1 - arlx1i * revp) + arlx1i * etax2nw * tensionnw) * denom1;
    AT(stressm_3, i, j) = (AT(stressm_3, i, j) * (c1 - arlx1i * revp) + arlx1i * etax2sw * tensionsw) * denom1;
    AT(stressm_4, i, j) = (AT(stressm_4, i, j) * (c1 - arlx1i * revp) + arlx1i * etax2se * tensionse) * denom1;
    AT(stress12_1, i, j) = (AT(stress12_1, i, j) * (c1 - arlx1i * revp) + arlx1i * p5 * etax2ne * shearne) * denom1;
    AT(stress12_2, i, j) = (AT(stress12_2, i, j) * (c1 - arlx1i * revp) + arlx1i * p5 * etax2nw * shearnw) * denom1;
    AT(stress12_3, i, j) = (AT(stress12_3, i, j) * (c1 - arlx1i * revp) + arlx1i * p5 * etax2sw * shearsw) * denom1;
    AT(stress12_4, i, j) = (AT(stress12_4, i, j) * (c1 - arlx1i * revp) + arlx1i * p5 * etax2se * shearse) * denom1;

    ssigpn = AT(stressp_1, i, j) + AT(stressp_2, i, j);
    ssigps = AT(stressp_3, i, j) + AT(stressp_4, i, j);
    ssigpe = AT(stressp_1, i, j) + AT(stressp_4, i, j);
    ssigpw = AT(stressp_2, i, j) + AT(stressp_3, i, j);
    ssigp1 = (AT(stressp_1, i, j) + AT(stressp_3, i, j)) * P055;
    ssigp2 = (AT(stressp_2, i, j) + AT(stressp_4, i, j)) * P055;
    ssigmn = AT(stressm_1, i, j) + AT(stressm_2, i, j);
    ssigms = AT(stressm_3, i, j) + AT(stressm_4, i, j);
    ssigme = AT(stressm_1, i, j) + AT(stressm_4, i, j);
    ssigmw = AT(stressm_2, i, j) + AT(stressm_3, i, j);
    ssigm1 = (AT(stressm_1, i, j) + AT(stressm_3, i, j)) * P055;
    ssigm2 = (AT(stressm_2, i, j) + AT(stressm_4, i, j)) * P055;
    ssig12n = AT(stress12_1, i, j) + AT(stress12_2, i, j);
    ssig12s = AT(stress12_3, i, j) + AT(stress12_4, i, j);
    ssig12e = AT(stress12_1, i, j) + AT(stress12_4, i, j);
    ssig12w = AT(stress12_2, i, j) + AT(stress12_3, i, j);
    ssig121 = (AT(stress12_1, i, j) + AT(stress12_3, i, j)) * P111;
    ssig122 = (AT(stress12_2, i, j) + AT(stress12_4, i, j)) * P111;

    csigpne = P111 * AT(stressp_1, i, j) + ssigp2 + P027 * AT(stressp_3, i, j);
    csigpnw = P111 * AT(stressp_2, i, j) + ssigp1 + P027 * AT(stressp_4, i, j);
    csigpsw = P111 * AT(stressp_3, i, j) + ssigp2 + P027 * AT(stressp_1, i, j);
    csigpse = P111 * AT(stressp_4, i, j) + ssigp1 + P027 * AT(stressp_2, i, j);
    csigmne = P111 * AT(stressm_1, i, j) + ssigm2 + P027 * AT(stressm_3, i, j);
    csigmnw = P111 * AT(stressm_2, i, j) + ssigm1 + P027 * AT(stressm_4, i, j);
    csigmsw = P111 * AT(stressm_3, i, j) + ssigm2 + P027 * AT(stressm_1, i, j);
    csigmse = P111 * AT(stressm_4, i, j) + ssigm1 + P027 * AT(stressm_2, i, j);
    csig12ne = P222 * AT(stress12_1, i, j) + ssig122 + P055 * AT(stress12_3, i, j);
    csig12nw = P222 * AT(stress12_2, i, j) + ssig121 + P055 * AT(stress12_4, i, j);
    csig12sw = P222 * AT(stress12_3, i, j) + ssig122 + P055 * AT(stress12_1, i, j);
    csig12se = P222 * AT(stress12_4, i, j) + ssig121 + P055 * AT(stress12_2, i, j);

    str12ew = p5 * AT(dxT, i, j) * (P333 * ssig12e + P166 * ssig12w);
    str12we = p5 * AT(dxT, i, j) * (P333 * ssig12w + P166 * ssig12e);
    str12ns = p5 * AT(dyT, i, j) * (P333 * ssig12n + P166 * ssig12s);
    str12sn = p5 * AT(dyT, i, j) * (P333 * ssig12s + P166 * ssig12n);

    strp_tmp = p25 * AT(dyT, i, j) * (P333 * ssigpn + P166 * ssigps);
    strm_tmp = p25 * AT(dyT, i, j) * (P333 * ssigmn + P166 * ssigms);
    AT(s1, i, j) = -strp_tmp - strm_tmp - str12ew + AT(dxhy, i, j) * (-csigpne + csigmne) + AT(dyhx, i, j) * csig12ne;
    AT(s2, i, j) = strp_tmp + strm_tmp - str12we + AT(dxhy, i, j) * (-csigpnw + csigmnw) + AT(dyhx, i, j) * csig12nw;
    strp_tmp = p25 * AT(dyT, i, j) * (P333 * ssigps + P166 * ssigpn);
    strm_tmp = p25 * AT(dyT, i, j) * (P333 * ssigms + P166 * ssigmn);
    AT(s3, i, j) = -strp_tmp - strm_tmp + str12ew + AT(dxhy, i, j) * (-csigpse + csigmse) + AT(dyhx, i, j) * csig12se;
    AT(s4, i, j) = strp_tmp + strm_tmp + str12we + AT(dxhy, i, j) * (-csigpsw + csigmsw) + AT(dyhx, i, j) * csig12sw;
    strp_tmp = p25 * AT(dxT, i, j) * (P333 * ssigpe + P166 * ssigpw);
    strm_tmp = p25 * AT(dxT, i, j) * (P333 * ssigme + P166 * ssigmw);
    AT(s5, i, j) = -strp_tmp + strm_tmp - str12ns - AT(dyhx, i, j) * (csigpne + csigmne) + AT(dxhy, i, j) * csig12ne;
    AT(s6, i, j) = strp_tmp - strm_tmp - str12sn - AT(dyhx, i, j) * (csigpse + csigmse) + AT(dxhy, i, j) * csig12se;
    strp_tmp = p25 * AT(dxT, i, j) * (P333 * ssigpw + P166 * ssigpe);
    strm_tmp = p25 * AT(dxT, i, j) * (P333 * ssigmw + P166 * ssigme);
    AT(s7, i, j) = -strp_tmp + strm_tmp + str12ns - AT(dyhx, i, j) * (csigpnw + csigmnw) + AT(dxhy, i, j) * csig12nw;
    AT(s8, i, j) = strp_tmp - strm_tmp + str12sn - AT(dyhx, i, j) * (csigpsw + csigmsw) + AT(dxhy, i, j) * csig12sw;
  }
}

/* bench_v0.F90:563-600.  Cw = cdn_ocn; rhow is stepu's local parameter
 * 1026 (bench_v0.F90:537-538), passed in p->rhow. */
void oracle_v0_stepu(const oracle_params *p, int32_t nxb, int32_t nyb, int32_t icellu,
                     const int32_t *indxui, const int32_t *indxuj, const double *Cw, const double *aiX,
                     const double *str, const double *uocn, const double *vocn, const double *waterx,
                     const double *watery, const double *forcex, const double *forcey, const double *umassdti,
                     const double *fm, const double *uarear, double *strintx, double *strinty, double *taubx,
                     double *tauby, const double *uvel_init, const double *vvel_init, double *uvel,
                     double *vvel, const double *Tbu) {
  const size_t plane = (size_t)nxb * (size_t)nyb;
  const double *s1 = str, *s2 = str + plane, *s3 = str + 2 * plane, *s4 = str + 3 * plane;
  const double *s5 = str + 4 * plane, *s6 = str + 5 * plane, *s7 = str + 6 * plane, *s8 = str + 7 * plane;
  const double brlx = p->brlx, revp = p->revp, u0 = p->u0, cosw = p->cosw, sinw = p->sinw, rhow = p->rhow;
  int32_t ij;
  (void)nyb;
  for (ij = 0; ij < icellu; ++ij) {
    const int32_t i = indxui[ij], j = indxuj[ij];
    double uold, vold, vrel, cca, ccb, ab2, cc1, cc2, taux, tauy, Cb;
    uold = AT(uvel, i, j);
    vold = AT(vvel, i, j);
    vrel = AT(aiX, i, j) * rhow * AT(Cw, i, j) *
           sqrt((AT(uocn, i, j) - uold) * (AT(uocn, i, j) - uold) + (AT(vocn, i, j) - vold) * (AT(vocn, i, j) - vold));
    taux = vrel * AT(waterx, i, j);
    tauy = vrel * AT(watery, i, j);
    Cb = AT(Tbu, i, j) / (sqrt(uold * uold + vold * vold) + u0);
    cca = (brlx + revp) * AT(umassdti, i, j) + vrel * cosw + Cb;
    ccb = AT(fm, i, j) + copysign(c1, AT(fm, i, j)) * vrel * sinw;
    ab2 = cca * cca + ccb * ccb;
    AT(strintx, i, j) = AT(uarear, i, j) * (AT(s1, i, j) + AT(s2, i + 1, j) + AT(s3, i, j + 1) + AT(s4, i + 1, j + 1));
    AT(strinty, i, j) = AT(uarear, i, j) * (AT(s5, i, j) + AT(s6, i, j + 1) + AT(s7, i + 1, j) + AT(s8, i + 1, j + 1));
    cc1 = AT(strintx, i, j) + AT(forcex, i, j) + taux + AT(umassdti, i, j) * (brlx * uold + revp * AT(uvel_init, i, j));
    cc2 = AT(strinty, i, j) + AT(forcey, i, j) + tauy + AT(umassdti, i, j) * (brlx * vold + revp * AT(vvel_init, i, j));
    AT(uvel, i, j) = (cca * cc1 + ccb * cc2) / ab2;
    AT(vvel, i, j) = (cca * cc2 - ccb * cc1) / ab2;
    AT(taubx, i, j) = -(AT(uvel, i, j) * Cb);
    AT(tauby, i, j) = -(AT(vvel, i, j) * Cb);
  }
}

/* vars.F90:100-143, called only when max_blocks > 1.  array: nblocks blocks
 * of (nxb, nyb).  lo/hi: 1-based interior bounds per block; src/dst: 3 ints
 * (i, j, block) per copy; block 0 in src means fill (0). */
void oracle_v0_halo(int32_t nxb, int32_t nyb, int32_t nblocks, const int32_t *ilo, const int32_t *jlo,
                    const int32_t *ihi, const int32_t *jhi, int32_t ncopies, const int32_t *src,
                    const int32_t *dst, double *array) {
  const size_t plane = (size_t)nxb * (size_t)nyb;
  int32_t b, i, j, m;
  for (b = 0; b < nblocks; ++b) {
    double *a = array + (size_t)b * plane;
    for (i = 1; i <= nxb; ++i) {
      AT(a, i, jlo[b] - 1) = c0;
      AT(a, i, jhi[b] + 1) = c0;
    }
    for (j = 1; j <= nyb; ++j) {
      AT(a, ilo[b] - 1, j) = c0;
      AT(a, ihi[b] + 1, j) = c0;
    }
  }
  for (m = 0; m < ncopies; ++m) {
    const int32_t is = src[3 * m], js = src[3 * m + 1], bs = src[3 * m + 2];
    const int32_t id = dst[3 * m], jd = dst[3 * m + 1], bd = dst[3 * m + 2];
    if (bs > 0) {
      if (bd > 0) AT(array + (size_t)(bd - 1) * plane, id, jd) = AT(array + (size_t)(bs - 1) * plane, is, js);
    } else if (bs == 0) {
      AT(array + (size_t)(bd - 1) * plane, id, jd) = c0;
    }
  }
}

/* Field bundle of oracle_v0_run: each a pointer to nblocks blocks. */
enum {
  V_STRENGTH = 0, V_UVEL, V_VVEL, V_DXT, V_DYT, V_DXHY, V_DYHX, V_CXP, V_CYP, V_CXM, V_CYM, V_DMIN,
  V_UAREAR, V_CDN, V_AIX, V_UOCN, V_VOCN, V_WATERX, V_WATERY, V_FORCEX, V_FORCEY, V_UMASSDTI, V_FM,
  V_STRINTX, V_STRINTY, V_TBU, V_SIG0, V_UVEL_INIT = V_SIG0 + 12, V_VVEL_INIT, V_TAUBX, V_TAUBY, V_NF
};

/* nsub subcycles of evp_v0.F90:93-136: per block stress then stepu (str is
 * one block of scratch, as strtmp is), then, with nblocks > 1 and a halo,
 * the halo update of uvel and vvel.  lists: nblocks x (nxb*nyb) each. */
void oracle_v0_run(const oracle_params *p, int32_t nxb, int32_t nyb, int32_t nblocks, int32_t nsub,
                   double *const *f, const int32_t *icellt, const int32_t *icellu, const int32_t *indxti,
                   const int32_t *indxtj, const int32_t *indxui, const int32_t *indxuj, double *str,
                   int32_t nhb, const int32_t *ilo, const int32_t *jlo, const int32_t *ihi, const int32_t *jhi,
                   int32_t ncopies, const int32_t *src, const int32_t *dst) {
  const size_t plane = (size_t)nxb * (size_t)nyb;
  int32_t k, b, c;
  for (k = 0; k < nsub; ++k) {
    for (b = 0; b < nblocks; ++b) {
      const size_t o = (size_t)b * plane;
      double *sig[12];
      for (c = 0; c < 12; ++c) sig[c] = f[V_SIG0 + c] + o;
      oracle_v0_stress(p, nxb, nyb, icellt[b], indxti + o, indxtj + o, f[V_UVEL] + o, f[V_VVEL] + o,
                       f[V_DXT] + o, f[V_DYT] + o, f[V_DXHY] + o, f[V_DYHX] + o, f[V_CXP] + o, f[V_CYP] + o,
                       f[V_CXM] + o, f[V_CYM] + o, f[V_DMIN] + o, f[V_STRENGTH] + o, sig, str);
      oracle_v0_stepu(p, nxb, nyb, icellu[b], indxui + o, indxuj + o, f[V_CDN] + o, f[V_AIX] + o, str,
                      f[V_UOCN] + o, f[V_VOCN] + o, f[V_WATERX] + o, f[V_WATERY] + o, f[V_FORCEX] + o,
                      f[V_FORCEY] + o, f[V_UMASSDTI] + o, f[V_FM] + o, f[V_UAREAR] + o, f[V_STRINTX] + o,
                      f[V_STRINTY] + o, f[V_TAUBX] + o, f[V_TAUBY] + o, f[V_UVEL_INIT] + o,
                      f[V_VVEL_INIT] + o, f[V_UVEL] + o, f[V_VVEL] + o, f[V_TBU] + o);
    }
    if (nblocks > 1 && nhb > 0) {
      oracle_v0_halo(nxb, nyb, nhb, ilo, jlo, ihi, jhi, ncopies, src, dst, f[V_UVEL]);
      oracle_v0_halo(nxb, nyb, nhb, ilo, jlo, ihi, jhi, ncopies, src, dst, f[V_VVEL]);
    }
  }
}
