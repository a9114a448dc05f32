/*
 * ORACLE -- test infrastructure only (see hybgen_oracle.h).
 *
 * Plain-C FP64 restatement of HYCOM-src hybgen.F90 (hybgen, hybgenaj,
 * hybgenbj, hybgencj and the PCM/PLM/PPM/WENO remappers), dpudpv.F90 and
 * the single-process xctilr halo (mod_xc_sm.h:1337-1429).  Statement order,
 * association and operation order follow the Fortran text so that, built
 * with -O2 -ffp-contract=off on x86-64, every output is bit-identical to the
 * reference compiled with flang (checked by tests/test_oracle_ref.py).
 * Arrays are indexed 1-based like the Fortran.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "hybgen_oracle.h"

/* ---- physical constants, mod_cb_arrays.F90:842-857 ------------------- */
static const double g_grav = 9.806;
static const double rhoref = 1000.0;
#define SVREF (1.0 / rhoref)
#define ONEM  (g_grav / SVREF)
#define TENM  (ONEM * 10.0)
#define ONEMM (ONEM * 0.001)
#define EPSIL 1.0e-11
#define QONEM (1.0 / ONEM)

/* ---- EOS_17T + EOS_SIG2 coefficients, stmt_fns.h:191-262 ------------- */
static const double c001 = 9.9984085444849347e+02;
static const double c002 = 7.3471625860981584e+00;
static const double c003 = -5.3211231792841769e-02;
static const double c004 = 3.6492439109814549e-04;
static const double c005 = 2.5880571023991390e+00;
static const double c006 = 6.7168282786692355e-03;
static const double c007 = 1.9203202055760151e-03;
static const double c008 = 1.0000000000000000e+00;
static const double c009 = 7.2815210113327091e-03;
static const double c010 = -4.4787265461983921e-05;
static const double c011 = 3.3851002965802430e-07;
static const double c012 = 1.3651202389758572e-10;
static const double c013 = 1.7632126669040377e-03;
static const double c014 = 8.8066583251206474e-06;
static const double c015 = 1.8832689434804897e-10;
static const double c016 = 5.7463776745432097e-06;
static const double c017 = 1.4716275472242334e-09;
static const double c018 = 1.1798263740430364e-02;
static const double c019 = 9.8920219266399117e-08;
static const double c020 = 4.6996642771754730e-06;
static const double c021 = 2.5862187075154352e-08;
static const double c022 = 3.2921414007960662e-12;
static const double c023 = 6.7103246285651894e-06;
static const double c024 = 2.4461698007024582e-17;
static const double c025 = 9.1534417604289062e-18;

static double C101, C103, C105, C108, C109, C111;
static int eos_ready = 0;

/* stmt_fns.h:251-262: folded coefficients at rpdb = pref*prs2pdb. */
static void eos_init(void) {
  if (eos_ready) return;
  const double pref = 2000.0e4, prs2pdb = 1.0e-4;
  const double rpdb = pref * prs2pdb;
  C101 = c001 + (c018 - c021 * rpdb) * rpdb;
  C103 = c003 + (c019 - c022 * rpdb) * rpdb;
  C105 = c005 + c020 * rpdb;
  C108 = c008 + c023 * rpdb;
  C109 = c009 - c025 * (rpdb * rpdb * rpdb);
  C111 = c011 - c024 * (rpdb * rpdb);
  eos_ready = 1;
}

static inline double fmax2(double a, double b) { return a > b ? a : b; }
static inline double fmin2(double a, double b) { return a < b ? a : b; }
static inline double fmin3(double a, double b, double c) { return fmin2(fmin2(a, b), c); }

/* sig(t,s), stmt_fns.h:505-511 */
static inline double sig17(double t, double s) {
  double sn = C101 + t * (c002 + t * (C103 + t * c004)) +
              s * (C105 - t * c006 + s * c007);
  double sd = C108 + t * (C109 + t * (c010 + t * (C111 + t * c012))) +
              s * (c013 - t * (c014 + t * t * c015) +
                   sqrt(fmax2(0.0, s)) * (c016 + t * t * c017));
  double sq = 1.0 / sd;
  return sn * sq - 1000.0;
}
/* tofsig/sofsig are the constant 99 under EOS_17T, stmt_fns.h:532-536 */
#define TOFSIG(r, s) (99.0)
#define SOFSIG(r, t) (99.0)

double oracle_sig(double t, double s) { eos_init(); return sig17(t, s); }

/* ---- column work arrays: 1-based, [field][k] ------------------------- */
#define KD (HYB_KMAX + 3)
typedef double colv[KD];

/* ---------------------------------------------------------------------- */
/* remappers                                                              */
/* ---------------------------------------------------------------------- */

/* hybgen_pcm_remap, hybgen.F90:2276-2391 */
static void pcm_remap(colv* si, const double* pi, const double* dpi,
                      colv* so, const double* po, int ki, int ko, int ks,
                      double thin) {
  double si_min[HYB_TRMAX + 5], si_max[HYB_TRMAX + 5];
  for (int i = 1; i <= ks; i++) {
    double mn = si[i][1], mx = si[i][1];
    for (int k = 2; k <= ki; k++) { mn = fmin2(mn, si[i][k]); mx = fmax2(mx, si[i][k]); }
    si_min[i] = mn; si_max[i] = mx;
  }
  double zx = pi[ki + 1];
  double zb = fmax2(po[1], pi[1]);
  int lb = 1;
  while (pi[lb + 1] < zb && lb < ki) lb++;
  for (int k = 1; k <= ko; k++) {
    double zt = zb;
    zb = fmin2(po[k + 1], zx);
    int lt = lb;
    while (pi[lb + 1] < zb && lb < ki) lb++;
    if (zb - zt <= thin || zt >= zx) {
      if (k != 1) { for (int i = 1; i <= ks; i++) so[i][k] = so[i][k - 1]; }
      else        { for (int i = 1; i <= ks; i++) so[i][k] = si[i][k]; }
    } else {
      if (lt != lb) {
        double dpt = pi[lt + 1] - zt;
        double dpb = zb - pi[lb];
        for (int i = 1; i <= ks; i++) {
          double o = si[i][(lt + lb) / 2];
          double sz = dpt * (si[i][lt] - o);
          for (int l = lt + 1; l <= lb - 1; l++) sz = sz + dpi[l] * (si[i][l] - o);
          sz = sz + dpb * (si[i][lb] - o);
          double v = o + sz / (zb - zt);
          v = fmax2(si_min[i], v);
          v = fmin2(si_max[i], v);
          so[i][k] = v;
        }
      } else {
        for (int i = 1; i <= ks; i++) so[i][k] = si[i][lt];
      }
    }
  }
}

/* hybgen_plm_coefs, hybgen.F90:2393-2458 */
static void plm_coefs(colv* si, const double* dpi, const int* lc, colv* ci,
                      int kk, int ks, double thin) {
  for (int i = 1; i <= ks; i++) { ci[i][1] = 0.0; ci[i][kk] = 0.0; }
  for (int k = 2; k <= kk - 1; k++) {
    if (lc[k] || dpi[k] <= thin) {
      for (int i = 1; i <= ks; i++) ci[i][k] = 0.0;
    } else {
      double qcen = dpi[k] / (dpi[k] + 0.5 * (dpi[k - 1] + dpi[k + 1]));
      for (int i = 1; i <= ks; i++) {
        double ztop = 2.0 * (si[i][k] - si[i][k - 1]);
        double zbot = 2.0 * (si[i][k + 1] - si[i][k]);
        double zcen = qcen * (si[i][k + 1] - si[i][k - 1]);
        if (ztop * zbot > 0.0)
          ci[i][k] = copysign(fmin3(fabs(zcen), fabs(zbot), fabs(ztop)), zbot);
        else
          ci[i][k] = 0.0;
      }
    }
  }
}

/* hybgen_plm_remap, hybgen.F90:2460-2585 */
static void plm_remap(colv* si, const double* pi, const double* dpi, colv* ci,
                      colv* so, const double* po, int ki, int ko, int ks,
                      double thin) {
  double si_min[HYB_TRMAX + 5], si_max[HYB_TRMAX + 5];
  for (int i = 1; i <= ks; i++) {
    double mn = si[i][1], mx = si[i][1];
    for (int k = 2; k <= ki; k++) { mn = fmin2(mn, si[i][k]); mx = fmax2(mx, si[i][k]); }
    si_min[i] = mn; si_max[i] = mx;
  }
  double zx = pi[ki + 1];
  double zb = fmax2(po[1], pi[1]);
  int lb = 1;
  while (pi[lb + 1] < zb && lb < ki) lb++;
  for (int k = 1; k <= ko; k++) {
    double zt = zb;
    zb = fmin2(po[k + 1], zx);
    int lt = lb;
    while (pi[lb + 1] < zb && lb < ki) lb++;
    if (zb - zt <= thin || zt >= zx) {
      if (k != 1) { for (int i = 1; i <= ks; i++) so[i][k] = so[i][k - 1]; }
      else        { for (int i = 1; i <= ks; i++) so[i][k] = si[i][k]; }
    } else {
      double xt = (zt - pi[lt]) / fmax2(dpi[lt], thin);
      double xb = (zb - pi[lb]) / fmax2(dpi[lb], thin);
      if (lt != lb) {
        double qt0 = (1.0 - xt);
        double qt1 = (1.0 - xt * xt) * 0.5;
        double qb0 = xb;
        double qb1 = xb * xb * 0.5;
        for (int i = 1; i <= ks; i++) {
          double o = si[i][(lt + lb) / 2];
          double c0 = si[i][lt] - o - 0.5 * ci[i][lt];
          double sz = dpi[lt] * (c0 * qt0 + ci[i][lt] * qt1);
          for (int l = lt + 1; l <= lb - 1; l++) sz = sz + dpi[l] * (si[i][l] - o);
          c0 = si[i][lb] - o - 0.5 * ci[i][lb];
          sz = sz + dpi[lb] * (c0 * qb0 + ci[i][lb] * qb1);
          double v = o + sz / (zb - zt);
          v = fmax2(si_min[i], v);
          v = fmin2(si_max[i], v);
          so[i][k] = v;
        }
      } else {
        double qt1 = (xb * xb - xt * xt - (xb - xt)) * 0.5;
        for (int i = 1; i <= ks; i++) {
          double sz = dpi[lt] * (ci[i][lt] * qt1);
          double v = si[i][lt] + sz / (zb - zt);
          v = fmax2(si_min[i], v);
          v = fmin2(si_max[i], v);
          so[i][k] = v;
        }
      }
    }
  }
}

/* hybgen_ppm_coefs, hybgen.F90:2587-2706.  ci[3*(i)+c] holds ci(:,i,c). */
static void ppm_coefs(colv* s, const double* dp, const int* lc, colv* ci,
                      int kk, int ks, double thin) {
  double as[KD], al[KD], ar[KD];
  double dpjp[KD], dp2jp[KD], dpj2p[KD], qdpjp[KD], qdp2jp[KD], qdpj2p[KD],
         dpq3[KD], qdp4[KD];
  for (int j = 1; j <= kk - 1; j++) {
    dpjp[j] = dp[j] + dp[j + 1];
    dp2jp[j] = dp[j] + dpjp[j];
    dpj2p[j] = dpjp[j] + dp[j + 1];
    qdpjp[j] = 1.0 / dpjp[j];
    qdp2jp[j] = 1.0 / dp2jp[j];
    qdpj2p[j] = 1.0 / dpj2p[j];
  }
  dpq3[2] = dp[2] / (dp[1] + dpjp[2]);
  for (int j = 3; j <= kk - 1; j++) {
    dpq3[j] = dp[j] / (dp[j - 1] + dpjp[j]);
    qdp4[j] = 1.0 / (dpjp[j - 2] + dpjp[j]);
  }
  for (int i = 1; i <= ks; i++) {
    as[1] = 0.0;
    for (int j = 2; j <= kk - 1; j++) {
      if (lc[j] || dp[j] <= thin) {
        as[j] = 0.0;
      } else {
        double slj = s[i][j] - s[i][j - 1];
        double srj = s[i][j + 1] - s[i][j];
        if (slj * srj > 0.0) {
          double scj = dpq3[j] * (dp2jp[j - 1] * srj * qdpjp[j] +
                                  dpj2p[j] * slj * qdpjp[j - 1]);
          as[j] = copysign(fmin3(fabs(2.0 * slj), fabs(scj), fabs(2.0 * srj)), scj);
        } else {
          as[j] = 0.0;
        }
      }
    }
    as[kk] = 0.0;
    al[1] = s[i][1];
    ar[1] = s[i][1];
    al[2] = s[i][1];
    for (int j = 3; j <= kk - 1; j++) {
      al[j] = s[i][j - 1] + dp[j - 1] * (s[i][j] - s[i][j - 1]) * qdpjp[j - 1] +
              qdp4[j] * (2.0 * dp[j] * dp[j - 1] * qdpjp[j - 1] * (s[i][j] - s[i][j - 1]) *
                             (dpjp[j - 2] * qdp2jp[j - 1] - dpjp[j] * qdpj2p[j - 1]) -
                         dp[j - 1] * as[j] * dpjp[j - 2] * qdp2jp[j - 1] +
                         dp[j] * as[j - 1] * dpjp[j] * qdpj2p[j - 1]);
      ar[j - 1] = al[j];
    }
    ar[kk - 1] = s[i][kk];
    al[kk] = s[i][kk];
    ar[kk] = s[i][kk];
    for (int j = 2; j <= kk - 1; j++) {
      if (lc[j] || dp[j] <= thin) {
        al[j] = s[i][j];
        ar[j] = s[i][j];
      } else if ((s[i][j + 1] - s[i][j]) * (s[i][j] - s[i][j - 1]) <= 0.0) {
        al[j] = s[i][j];
        ar[j] = s[i][j];
      } else {
        double da = ar[j] - al[j];
        double a6 = 6.0 * s[i][j] - 3.0 * (al[j] + ar[j]);
        if (da * a6 > da * da) {
          al[j] = 3.0 * s[i][j] - 2.0 * ar[j];
        } else if (da * a6 < -(da * da)) {
          ar[j] = 3.0 * s[i][j] - 2.0 * al[j];
        }
      }
    }
    for (int j = 1; j <= kk; j++) {
      if (al[j] != ar[j]) {
        ci[3 * i + 0][j] = al[j];
        ci[3 * i + 1][j] = ar[j] - al[j];
        ci[3 * i + 2][j] = 6.0 * s[i][j] - 3.0 * (al[j] + ar[j]);
      } else {
        ci[3 * i + 0][j] = al[j];
        ci[3 * i + 1][j] = 0.0;
        ci[3 * i + 2][j] = 0.0;
      }
    }
  }
}

/* hybgen_ppm_remap, hybgen.F90:2708-2843 */
static void ppm_remap(colv* si, const double* pi, const double* dpi, colv* ci,
                      colv* so, const double* po, int ki, int ko, int ks,
                      double thin) {
  double si_min[HYB_TRMAX + 5], si_max[HYB_TRMAX + 5];
  for (int i = 1; i <= ks; i++) {
    double mn = si[i][1], mx = si[i][1];
    for (int k = 2; k <= ki; k++) { mn = fmin2(mn, si[i][k]); mx = fmax2(mx, si[i][k]); }
    si_min[i] = mn; si_max[i] = mx;
  }
  double zx = pi[ki + 1];
  double zb = fmax2(po[1], pi[1]);
  int lb = 1;
  while (pi[lb + 1] < zb && lb < ki) lb++;
  for (int k = 1; k <= ko; k++) {
    double zt = zb;
    zb = fmin2(po[k + 1], zx);
    int lt = lb;
    while (pi[lb + 1] < zb && lb < ki) lb++;
    if (zb - zt <= thin || zt >= zx) {
      if (k != 1) { for (int i = 1; i <= ks; i++) so[i][k] = so[i][k - 1]; }
      else        { for (int i = 1; i <= ks; i++) so[i][k] = si[i][k]; }
    } else {
      double xt = (zt - pi[lt]) / fmax2(dpi[lt], thin);
      double xb = (zb - pi[lb]) / fmax2(dpi[lb], thin);
      if (lt != lb) {
        double qt0 = (1.0 - xt);
        double qt1 = (1.0 - xt * xt) * 0.5;
        double qt2 = (1.0 - xt * xt * xt) / 3.0;
        double qb0 = xb;
        double qb1 = xb * xb * 0.5;
        double qb2 = xb * xb * xb / 3.0;
        for (int i = 1; i <= ks; i++) {
          double o = si[i][(lt + lb) / 2];
          double sz = dpi[lt] * ((ci[3 * i][lt] - o) * qt0 +
                                 (ci[3 * i + 1][lt] + ci[3 * i + 2][lt]) * qt1 -
                                 ci[3 * i + 2][lt] * qt2);
          for (int l = lt + 1; l <= lb - 1; l++) sz = sz + dpi[l] * (si[i][l] - o);
          sz = sz + dpi[lb] * ((ci[3 * i][lb] - o) * qb0 +
                               (ci[3 * i + 1][lb] + ci[3 * i + 2][lb]) * qb1 -
                               ci[3 * i + 2][lb] * qb2);
          double v = o + sz / (zb - zt);
          v = fmax2(si_min[i], v);
          v = fmin2(si_max[i], v);
          so[i][k] = v;
        }
      } else {
        double qt0 = (xb - xt);
        double qt1 = (xb * xb - xt * xt) * 0.5;
        double qt2 = (xb * xb * xb - xt * xt * xt) / 3.0;
        for (int i = 1; i <= ks; i++) {
          double o = si[i][lt];
          double sz = dpi[lt] * ((ci[3 * i][lt] - o) * qt0 +
                                 (ci[3 * i + 1][lt] + ci[3 * i + 2][lt]) * qt1 -
                                 ci[3 * i + 2][lt] * qt2);
          double v = o + sz / (zb - zt);
          v = fmax2(si_min[i], v);
          v = fmin2(si_max[i], v);
          so[i][k] = v;
        }
      }
    }
  }
}

/* hybgen_weno_coefs, hybgen.F90:2845-2968.  ci[2*i+c] holds ci(:,i,c). */
static void weno_coefs(colv* s, const double* dp, const int* lc, colv* ci,
                       int kk, int ks, double thin) {
  const double dsmll = 1.0e-8;
  double qdpjm[KD], qdpjmjp[KD], dpjm2jp[KD];
  double zw1[KD], zw2[KD], zw3[KD];
  for (int j = 2; j <= kk - 1; j++) {
    qdpjm[j] = 1.0 / (dp[j - 1] + dp[j]);
    qdpjmjp[j] = 1.0 / (dp[j - 1] + dp[j] + dp[j + 1]);
    dpjm2jp[j] = dp[j - 1] + 2.0 * dp[j] + dp[j + 1];
  }
  qdpjm[kk] = 1.0 / (dp[kk - 1] + dp[kk]);
  for (int i = 1; i <= ks; i++) {
    double* c1 = ci[2 * i];
    double* c2 = ci[2 * i + 1];
    for (int j = 2; j <= kk; j++) zw3[j] = qdpjm[j] * (s[i][j] - s[i][j - 1]);
    c1[1] = s[i][1]; c2[1] = s[i][1]; zw1[1] = 0.0; zw2[1] = 0.0;
    for (int j = 2; j <= kk - 1; j++) {
      if (lc[j] || dp[j] <= thin) {
        c1[j] = s[i][j]; c2[j] = s[i][j]; zw1[j] = 0.0; zw2[j] = 0.0;
      } else {
        double q001 = dp[j] * zw3[j + 1];
        double q002 = dp[j] * zw3[j];
        if (q001 * q002 < 0.0) { q001 = 0.0; q002 = 0.0; }
        double q01 = dpjm2jp[j] * zw3[j + 1];
        double q02 = dpjm2jp[j] * zw3[j];
        if (fabs(q001) > fabs(q02)) q001 = q02;
        if (fabs(q002) > fabs(q01)) q002 = q01;
        double q = (q001 - q002) * qdpjmjp[j];
        q001 = q001 - q * dp[j + 1];
        q002 = q002 + q * dp[j - 1];
        c2[j] = s[i][j] + q001;
        c1[j] = s[i][j] - q002;
        double a = 2.0 * q001 - q002, b = 2.0 * q002 - q001;
        zw1[j] = a * a;
        zw2[j] = b * b;
      }
    }
    c1[kk] = s[i][kk]; c2[kk] = s[i][kk]; zw1[kk] = 0.0; zw2[kk] = 0.0;
    for (int j = 2; j <= kk; j++) {
      double q002 = fmax2(zw2[j - 1], dsmll);
      double q001 = fmax2(zw1[j], dsmll);
      zw3[j] = (q001 * c2[j - 1] + q002 * c1[j]) / (q001 + q002);
    }
    zw3[1] = 2.0 * s[i][1] - zw3[2];
    zw3[kk + 1] = 2.0 * s[i][kk] - zw3[kk];
    for (int j = 2; j <= kk - 1; j++) {
      if (!(lc[j] || dp[j] <= thin)) {
        double q01 = zw3[j + 1] - s[i][j];
        double q02 = s[i][j] - zw3[j];
        double q001 = 2.0 * q01;
        double q002 = 2.0 * q02;
        if (q01 * q02 < 0.0) { q01 = 0.0; q02 = 0.0; }
        else if (fabs(q01) > fabs(q002)) { q01 = q002; }
        else if (fabs(q02) > fabs(q001)) { q02 = q001; }
        c1[j] = s[i][j] - q02;
        c2[j] = s[i][j] + q01;
      }
    }
  }
}

/* hybgen_weno_remap, hybgen.F90:2970-3107 */
static void weno_remap(colv* si, const double* pi, const double* dpi, colv* ci,
                       colv* so, const double* po, int ki, int ko, int ks,
                       double thin) {
  double si_min[HYB_TRMAX + 5], si_max[HYB_TRMAX + 5];
  for (int i = 1; i <= ks; i++) {
    double mn = si[i][1], mx = si[i][1];
    for (int k = 2; k <= ki; k++) { mn = fmin2(mn, si[i][k]); mx = fmax2(mx, si[i][k]); }
    si_min[i] = mn; si_max[i] = mx;
  }
  double zx = pi[ki + 1];
  double zb = fmax2(po[1], pi[1]);
  int lb = 1;
  while (pi[lb + 1] < zb && lb < ki) lb++;
  for (int k = 1; k <= ko; k++) {
    double zt = zb;
    zb = fmin2(po[k + 1], zx);
    int lt = lb;
    while (pi[lb + 1] < zb && lb < ki) lb++;
    if (zb - zt <= thin || zt >= zx) {
      if (k != 1) { for (int i = 1; i <= ks; i++) so[i][k] = so[i][k - 1]; }
      else        { for (int i = 1; i <= ks; i++) so[i][k] = si[i][k]; }
    } else {
      double xt = (zt - pi[lt]) / fmax2(dpi[lt], thin);
      double xb = (zb - pi[lb]) / fmax2(dpi[lb], thin);
      if (lt != lb) {
        double dpt = pi[lt + 1] - zt;
        double dpb = zb - pi[lb];
        double qt1 = xt * (xt - 1.0);
        double qt2 = qt1 + xt;
        double qt0 = 1.0 - qt1 - qt2;
        double qb1 = (xb - 1.0) * (xb - 1.0);
        double qb2 = qb1 - 1.0 + xb;
        double qb0 = 1.0 - qb1 - qb2;
        for (int i = 1; i <= ks; i++) {
          double o = si[i][(lt + lb) / 2];
          double sz = dpt * (qt0 * (si[i][lt] - o) + qt1 * (ci[2 * i][lt] - o) +
                             qt2 * (ci[2 * i + 1][lt] - o));
          for (int l = lt + 1; l <= lb - 1; l++) sz = sz + dpi[l] * (si[i][l] - o);
          sz = sz + dpb * (qb0 * (si[i][lb] - o) + qb1 * (ci[2 * i][lb] - o) +
                           qb2 * (ci[2 * i + 1][lb] - o));
          double v = o + sz / (zb - zt);
          v = fmax2(si_min[i], v);
          v = fmin2(si_max[i], v);
          so[i][k] = v;
        }
      } else {
        double qt1 = xb * xb + xt * xt + xb * xt + 1.0 - 2.0 * (xb + xt);
        double qt2 = qt1 - 1.0 + (xb + xt);
        for (int i = 1; i <= ks; i++) {
          double o = si[i][lt];
          double sz = qt1 * (ci[2 * i][lt] - o) + qt2 * (ci[2 * i + 1][lt] - o);
          double v = o + sz;
          v = fmax2(si_min[i], v);
          v = fmin2(si_max[i], v);
          so[i][k] = v;
        }
      }
    }
  }
}

/* Dispatch as at hybgen.F90:1530-1548: PLM coefs take dprs, PPM/WENO coefs
 * take dpi; all remaps take dprs. */
static void remap_dispatch(int hybmap, colv* s1d, const double* pres,
                           const double* dprs, const double* dpi,
                           const int* lcm, colv* c1d, colv* f1d,
                           const double* prsf, int kk, int nums1d,
                           double dpthin) {
  if (hybmap == 0) {
    pcm_remap(s1d, pres, dprs, f1d, prsf, kk, kk, nums1d, dpthin);
  } else if (hybmap == 1) {
    plm_coefs(s1d, dprs, lcm, c1d, kk, nums1d, dpthin);
    plm_remap(s1d, pres, dprs, c1d, f1d, prsf, kk, kk, nums1d, dpthin);
  } else if (hybmap == 2) {
    ppm_coefs(s1d, dpi, lcm, c1d, kk, nums1d, dpthin);
    ppm_remap(s1d, pres, dprs, c1d, f1d, prsf, kk, kk, nums1d, dpthin);
  } else if (hybmap == 3) {
    weno_coefs(s1d, dpi, lcm, c1d, kk, nums1d, dpthin);
    weno_remap(s1d, pres, dprs, c1d, f1d, prsf, kk, kk, nums1d, dpthin);
  }
}

void oracle_remap_column(int method, int kk, const double* si,
                         const double* pi, const double* po, const int* lcm0,
                         double thin, double* so) {
  static __thread colv s[2], f[2], c[8];
  double pres[KD], prsf[KD], dprs[KD], dpi[KD];
  int lcm[KD] = {0};
  for (int k = 1; k <= kk; k++) {
    s[1][k] = si[k - 1];
    lcm[k] = lcm0 ? lcm0[k - 1] : 0;
  }
  for (int k = 1; k <= kk + 1; k++) { pres[k] = pi[k - 1]; prsf[k] = po[k - 1]; }
  for (int k = 1; k <= kk; k++) {
    dprs[k] = pres[k + 1] - pres[k];
    dpi[k] = fmax2(dprs[k], thin);
  }
  remap_dispatch(method, s, pres, dprs, dpi, lcm, c, f, prsf, kk, 1, thin);
  for (int k = 1; k <= kk; k++) so[k - 1] = f[1][k];
}

/* One column of stokes_vertical_j's remap (mod_stokes.F90:686-694): ks
 * fields of ki old layers onto ko new ones, PLM slopes on dpi, lcm off.
 * si[f*ki + k], pi[ki+1], po[ko+1], so[f*ko + k], 0-based. */
void oracle_plm_dpi_column(int ki, int ko, int ks, const double* si, const double* pi,
                           const double* po, double thin, double* so) {
  static __thread colv s[3], f[3], c[3];
  double pres[KD], prsf[KD], dprs[KD], dpi[KD];
  int lcm[KD] = {0};
  for (int i = 1; i <= ks; i++)
    for (int k = 1; k <= ki; k++) s[i][k] = si[(i - 1) * ki + k - 1];
  for (int k = 1; k <= ki + 1; k++) pres[k] = pi[k - 1];
  for (int k = 1; k <= ko + 1; k++) prsf[k] = po[k - 1];
  for (int k = 1; k <= ki; k++) {
    dprs[k] = pres[k + 1] - pres[k];
    dpi[k] = fmax2(dprs[k], thin);
  }
  plm_coefs(s, dpi, lcm, c, ki, ks, thin);
  plm_remap(s, pres, dprs, c, f, prsf, ki, ko, ks, thin);
  for (int i = 1; i <= ks; i++)
    for (int k = 1; k <= ko; k++) so[(i - 1) * ko + k - 1] = f[i][k];
}

/* ---------------------------------------------------------------------- */
/* trcupd.F90 tracer column remaps: pcmtrc, plmtrc, plmtrcx, plmtrcs       */
/* ---------------------------------------------------------------------- */
#define TRC_THIN 1.e-6 /* trcupd.F90:857, 960 */

/* plmtrcs, trcupd.F90:1088-1147: slopes of one field (1-based a, s) */
static void trc_plmtrcs(const double* rl, const double* rr, const double* a, double* s, int n) {
  s[1] = 0.0;
  for (int j = 2; j <= n - 1; j++) {
    double sl = rl[j] * (a[j] - a[j - 1]);
    double sr = rr[j] * (a[j + 1] - a[j]);
    if (sl * sr > 0.0) s[j] = copysign(fmin2(fabs(sl), fabs(sr)), sl);
    else s[j] = 0.0;
  }
  s[n] = 0.0;
  for (int j = 2; j <= n - 1; j++) {
    if (s[j] != 0.0) {
      double dl = -0.5 * (s[j] + s[j - 1]) + a[j] - a[j - 1];
      double dr = -0.5 * (s[j + 1] + s[j]) + a[j + 1] - a[j];
      double ds = copysign(fmin2(fabs(dl), fabs(dr)), dl);
      s[j] = s[j] + 2.0 * ds;
    }
  }
}

/* pcmtrc (method 0, trcupd.F90:808-909) or plmtrc (method 1, 911-1036)
 * of one column: si[f*ki + l-1], pi[ki+1], po[ko+1], so[f*ko + k-1].
 * Returns -1 where the reference stops ('bad lf'); that output is NaN (and
 * thin outputs below it copy it), later outputs are formed as usual.  A
 * thin first output takes si(1): the reference reads so(0) there, outside
 * the array. */
int oracle_trc_column(int method, int ki, int ks, const double* si, const double* pi,
                      const double* po, int ko, double* so) {
  double p[KD], pit[KD], ql[KD], qr[KD], a[KD], sis[KD], sok[HYB_TRMAX + 5];
  double smin[HYB_TRMAX + 5], smax[HYB_TRMAX + 5];
  static __thread double slopes[HYB_TRMAX + 5][KD];
  int rc = 0;
  for (int l = 1; l <= ki + 1; l++) p[l] = pi[l - 1];
  if (method == 1) {
    for (int i = 1; i <= ks; i++) {
      smin[i] = si[(i - 1) * ki];
      smax[i] = si[(i - 1) * ki];
      for (int l = 2; l <= ki; l++) {
        smin[i] = fmin2(smin[i], si[(i - 1) * ki + l - 1]);
        smax[i] = fmax2(smax[i], si[(i - 1) * ki + l - 1]);
      }
    }
    for (int l = 1; l <= ki; l++) pit[l] = fmax2(p[l + 1] - p[l], TRC_THIN);
    /* plmtrcx, trcupd.F90:1038-1086 */
    ql[1] = qr[1] = 0.0;
    for (int l = 2; l <= ki - 1; l++) {
      ql[l] = 2.0 * pit[l] / (pit[l - 1] + pit[l]);
      qr[l] = 2.0 * pit[l] / (pit[l] + pit[l + 1]);
    }
    ql[ki] = qr[ki] = 0.0;
    for (int i = 1; i <= ks; i++) {
      for (int l = 1; l <= ki; l++) a[l] = si[(i - 1) * ki + l - 1];
      trc_plmtrcs(ql, qr, a, sis, ki);
      for (int l = 1; l <= ki; l++) slopes[i][l] = sis[l];
    }
  }
  int lf = 1;
  double zb = po[0];
  for (int k = 1; k <= ko; k++) {
    double zt = zb;
    zb = po[k];
    if (zb - zt < TRC_THIN || zt >= p[ki + 1]) {
      for (int i = 1; i <= ks; i++)
        so[(i - 1) * ko + k - 1] = (k > 1) ? so[(i - 1) * ko + k - 2] : si[(i - 1) * ki];
      continue;
    }
    if (p[lf] > zt) {
      for (int i = 1; i <= ks; i++) so[(i - 1) * ko + k - 1] = NAN;
      rc = -1;
      continue;
    }
    for (int i = 1; i <= ks; i++) sok[i] = 0.0;
    for (int l = lf; l <= ki; l++) {
      if (p[l] > zb) {
        lf = l - 1;
        break;
      } else if (p[l] >= zt && p[l + 1] <= zb) {
        double q = fmax2(p[l + 1] - p[l], TRC_THIN) / (zb - zt);
        for (int i = 1; i <= ks; i++) sok[i] = sok[i] + q * si[(i - 1) * ki + l - 1];
      } else {
        double q = fmax2(fmin2(p[l + 1], zb) - fmax2(p[l], zt), TRC_THIN) / (zb - zt);
        if (method == 0) {
          for (int i = 1; i <= ks; i++) sok[i] = sok[i] + q * si[(i - 1) * ki + l - 1];
        } else {
          double zc = 0.5 * (fmin2(p[l + 1], zb) + fmax2(p[l], zt));
          double qc = (zc - p[l]) / pit[l] - 0.5;
          for (int i = 1; i <= ks; i++)
            sok[i] = sok[i] + q * (si[(i - 1) * ki + l - 1] + qc * slopes[i][l]);
        }
      }
    }
    for (int i = 1; i <= ks; i++) {
      double v = sok[i];
      if (method == 1) {
        v = fmax2(smin[i], v);
        v = fmin2(smax[i], v);
      }
      so[(i - 1) * ko + k - 1] = v;
    }
  }
  return rc;
}

/* The column remap at every masked point of a grid (initrc trcupd.F90:
 * 167-199, thermf.F90:770-806): si (NI,NJ,ki,ks), pi (NI,NJ,ki+1), po
 * (NI,NJ,ko+1), so (NI,NJ,ko,ks). */
int oracle_trc_remap(const hyb_trc_params* P, const hyb_trc_fields* F) {
  const int ki = P->ki, ko = P->ko, ks = P->ks, nb = P->nbdy;
  if (ki < 2 || ki > HYB_KMAX || ko < 1 || ko > HYB_KMAX || ks < 1 || ks > HYB_TRMAX)
    return HYB_EINVAL;
  const long NI = P->idm + 2 * nb, NIJ = NI * (long)(P->jdm + 2 * nb);
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (int j = 1; j <= P->jj; j++) {
    double si[HYB_TRMAX * KD], pi[KD], po[KD], so[HYB_TRMAX * KD];
    for (int i = 1; i <= P->ii; i++) {
      const long ij = (long)(i - 1 + nb) + NI * (long)(j - 1 + nb);
      if (F->mask[ij] == 0) continue;
      for (int f = 0; f < ks; f++)
        for (int l = 0; l < ki; l++) si[f * ki + l] = F->si[ij + NIJ * ((long)l + (long)ki * f)];
      for (int l = 0; l <= ki; l++) pi[l] = F->pi[ij + NIJ * l];
      for (int k = 0; k <= ko; k++) po[k] = F->po[ij + NIJ * k];
      if (oracle_trc_column(P->method, ki, ks, si, pi, po, ko, so)) bad = 1;
      for (int f = 0; f < ks; f++)
        for (int k = 0; k < ko; k++) F->so[ij + NIJ * ((long)k + (long)ko * f)] = so[f * ko + k];
    }
  }
  return bad ? HYB_EINVAL : 0;
}

/* stokes_vertical_j(n,j) for j = 1..jj (mod_stokes.F90:541-546, 650-705) */
int oracle_stokes_vertical(const hyb_stokes_params* P, const hyb_stokes_fields* F, int n) {
  const int nz = P->nsdzi, kk = P->kdm, nb = P->nbdy;
  if (nz < 2 || nz > HYB_KMAX || kk < 1 || kk > HYB_KMAX || n < 1 || n > 2) return HYB_EINVAL;
  const long NI = P->idm + 2 * nb, NIJ = NI * (long)(P->jdm + 2 * nb);
  const double dpthin = 0.001 * ONEMM;
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= P->jj; j++) {
    colv s[3], f[3], c[3];
    double pres[KD], dprs[KD], dpi[KD], prsf[KD];
    int lcm[KD] = {0};
    for (int i = 1; i <= P->ii; i++) {
      const long ij = (long)(i - 1 + nb) + NI * (long)(j - 1 + nb);
      if (F->ip[ij] == 0) continue;
      const double pbot = F->p[ij + NIJ * kk];
      pres[1] = 0.0;
      for (int k = 1; k <= nz; k++) {
        const long a = ij + NIJ * (long)(2 * (k - 1));
        s[1][k] = P->ws0 * F->usdz[a] + P->ws1 * F->usdz[a + NIJ];
        s[2][k] = P->ws0 * F->vsdz[a] + P->ws1 * F->vsdz[a + NIJ];
        pres[k + 1] = fmin2(F->sdzi[k - 1], pbot);
        dprs[k] = pres[k + 1] - pres[k];
        dpi[k] = fmax2(dprs[k], dpthin);
      }
      prsf[1] = 0.0;
      for (int k = 1; k <= kk; k++)
        prsf[k + 1] = fmin2(prsf[k] + F->dp[ij + NIJ * ((long)(k - 1) + (long)kk * (n - 1))], pbot);
      plm_coefs(s, dpi, lcm, c, nz, 2, dpthin);
      plm_remap(s, pres, dprs, c, f, prsf, nz, kk, 2, dpthin);
      for (int k = 1; k <= kk; k++) {
        F->usdp[ij + NIJ * (long)(k - 1)] = f[1][k];
        F->vsdp[ij + NIJ * (long)(k - 1)] = f[2][k];
      }
    }
  }
  return 0;
}

/* ---------------------------------------------------------------------- */
/* field access                                                            */
/* ---------------------------------------------------------------------- */
typedef struct {
  const hyb_params* P;
  const hyb_fields* F;
  long NI, NJ, NIJ;
  int K, nb;
  int32_t *rec_fixlay, *rec_kp; /* optional per-column decisions (idm*jdm, i fastest) */
} ctx_t;

#define IJ(i, j) ((long)((i) - 1 + c->nb) + c->NI * (long)((j) - 1 + c->nb))
#define L3(k) (c->NIJ * (long)((k) - 1))
#define L4(k, l) (c->NIJ * ((long)((k) - 1) + (long)c->K * ((l) - 1)))
#define P3(a, i, j, k) ((a)[IJ(i, j) + L3(k)])
#define F4(a, i, j, k, l) ((a)[IJ(i, j) + L4(k, l)])
#define TRC(i, j, k, l, t) \
  (c->F->tracer[IJ(i, j) + c->NIJ * ((long)((k) - 1) + (long)c->K * (((l) - 1) + 2L * ((t) - 1)))])
#define P2(a, i, j) ((a)[IJ(i, j)])

/* ---------------------------------------------------------------------- */
/* hybgenaj, hybgen.F90:308-1756                                           */
/* ---------------------------------------------------------------------- */
static void hybgenaj(const ctx_t* c, int m, int n, int j) {
  (void)m;
  const hyb_params* P = c->P;
  const hyb_fields* F = c->F;
  const int kk = c->K, nhybrd = P->nhybrd, ntracr = P->ntracr;
  double *p = F->p, *dp = F->dp, *temp = F->temp, *saln = F->saln, *th3d = F->th3d;
  const double* theta = F->theta;

  const double qqmn = -4.0, qqmx = 2.0;
  const double cusha = qqmn * qqmn * (qqmx - 1.0) / ((qqmx - qqmn) * (qqmx - qqmn));
  const double cushb = 1.0 / qqmn;
#define QQ(delp, dp0) (fmax2(qqmn, fmin2(qqmx, (delp) / (dp0))))
#define CUSHN(delp, dp0)                                                      \
  ((dp0) * (1.0 + cusha * ((1.0 - cushb * QQ(delp, dp0)) * (1.0 - cushb * QQ(delp, dp0)))) * \
   fmax2(1.0, (delp) / ((dp0) * qqmx)))

  static __thread colv s1d[HYB_TRMAX + 5], f1d[HYB_TRMAX + 5], c1d[3 * (HYB_TRMAX + 5)];
  int lcm[KD] = {0};
  double dpi[KD], dprs[KD], pres[KD], prsf[KD], qhrlx[KD], dp0ij[KD], dp0cum[KD];

  const double dpthin = 0.001 * ONEMM;
  const double thkbop = P->thkbot * ONEM;
  const double hybrlx = 1.0 / P->qhybrlx;
  const double rthin = P->hybthn;
  const int nums1d = ntracr + 2;
  const double* dp0k = P->dp0k;
  const double* ds0k = P->ds0k;
  const double* dx0k = P->dx0k;

  if (!P->isopcm) {
    for (int k = 1; k <= nhybrd; k++) lcm[k] = 0;
    for (int k = nhybrd + 1; k <= kk; k++) lcm[k] = 1;
  }

  for (int i = 1; i <= c->P->ii; i++) {
    if (P2(F->ip, i, j) == 0) continue;
    double q, qdep, qts, p_hat, p_hat0 = 0.0, p_hat2, p_hat3;
    int k, ka, kp, fixlay, fixall;

    /* terrain following, 399-477 */
    if (P->dpns == P->dsns) {
      qdep = 1.0;
    } else {
      qdep = fmax2(0.0, fmin2(1.0, (P2(F->depths, i, j) - P->dsns) / (P->dpns - P->dsns)));
    }
    if (qdep < 1.0) {
      P3(p, i, j, 1) = 0.0;
      dp0cum[1] = 0.0;
      qhrlx[1] = 1.0;
      dp0ij[1] = qdep * dp0k[0] + (1.0 - qdep) * ds0k[0];
      dp0cum[2] = dp0cum[1] + dp0ij[1];
      qhrlx[2] = 1.0;
      P3(p, i, j, 2) = P3(p, i, j, 1) + F4(dp, i, j, 1, n);
      for (k = 2; k <= kk; k++) {
        qhrlx[k + 1] = 1.0;
        dp0ij[k] = qdep * dp0k[k - 1] + (1.0 - qdep) * ds0k[k - 1];
        dp0cum[k + 1] = dp0cum[k] + dp0ij[k];
        P3(p, i, j, k + 1) = P3(p, i, j, k) + F4(dp, i, j, k, n);
      }
    } else {
      P3(p, i, j, 1) = 0.0;
      dp0cum[1] = 0.0;
      qhrlx[1] = 1.0;
      dp0ij[1] = dp0k[0];
      dp0cum[2] = dp0cum[1] + dp0ij[1];
      qhrlx[2] = 1.0;
      P3(p, i, j, 2) = P3(p, i, j, 1) + F4(dp, i, j, 1, n);
      for (k = 2; k <= kk; k++) {
        double d = dp0k[k - 1];
        if (d <= P->dp00i) {
          q = d;
          qts = 0.0;
        } else {
          q = fmax2(P->dp00i, d * d / fmax2(d, P3(p, i, j, k) - dp0cum[k]));
          qts = 1.0 - (q - P->dp00i) / (d - P->dp00i);
        }
        qhrlx[k + 1] = 1.0 / (1.0 + qts * (hybrlx - 1.0));
        dp0ij[k] = fmin2(q, d);
        dp0cum[k + 1] = dp0cum[k] + d;
        P3(p, i, j, k + 1) = P3(p, i, j, k) + F4(dp, i, j, k, n);
      }
    }

    /* always-fixed layers, 479-488 */
    fixlay = 1;
    for (k = 2; k <= nhybrd; k++) {
      if (dp0cum[k] >= P2(F->topiso, i, j)) break;
      qhrlx[k + 1] = 1.0;
      fixlay = fixlay + 1;
    }

    if (P->ldenfix) { /* 496-549 */
      int lf = 1;
      fixall = fixlay;
      for (k = fixall + 1; k <= nhybrd; k++) {
        double denfix = F4(th3d, i, j, nhybrd, n);
        for (int l = lf; l <= nhybrd; l++) {
          double zk = dp0cum[k + 1];
          if (P3(p, i, j, l) <= zk && P3(p, i, j, l + 1) >= zk) {
            double z0 = 0.5 * (P3(p, i, j, l) + P3(p, i, j, l + 1));
            if (zk <= z0) {
              if (l == 1) {
                denfix = F4(th3d, i, j, 1, n);
              } else {
                double zm = 0.5 * (P3(p, i, j, l - 1) + P3(p, i, j, l));
                q = (z0 - zk) / (z0 - zm);
                denfix = q * F4(th3d, i, j, l - 1, n) + (1.0 - q) * F4(th3d, i, j, l, n);
              }
            } else {
              if (P3(p, i, j, l + 1) == P3(p, i, j, nhybrd + 1)) {
                denfix = F4(th3d, i, j, nhybrd, n);
              } else {
                double zp = 0.5 * (P3(p, i, j, l + 1) + P3(p, i, j, l + 2));
                q = (zk - z0) / (zp - z0);
                denfix = q * F4(th3d, i, j, l + 1, n) + (1.0 - q) * F4(th3d, i, j, l, n);
              }
            }
            lf = l;
            break;
          }
        }
        if (P3(p, i, j, k + 1) > dp0cum[k + 1] + 0.1 * dp0ij[k] &&
            P3(theta, i, j, k) > denfix) {
          if (fixlay > fixall) {
            if (P3(p, i, j, k) > dp0cum[k]) fixlay = fixlay - 1;
          }
          break;
        }
        qhrlx[k] = 1.0;
        fixlay = fixlay + 1;
      }
    } else { /* 550-573 */
      fixall = fixlay;
      for (k = fixlay + 1; k <= nhybrd; k++) {
        if (P3(p, i, j, k + 1) > dp0cum[k + 1] + 0.1 * dp0ij[k]) {
          if (fixlay > fixall) {
            if (P3(p, i, j, k) > dp0cum[k]) fixlay = fixlay - 1;
          }
          break;
        }
        qhrlx[k] = 1.0;
        fixlay = fixlay + 1;
      }
    }

    /* deepest inflated layer, 590-598 */
    kp = 2;
    for (k = kk; k >= 3; k--) {
      if (P3(p, i, j, k + 1) - P3(p, i, j, k) >= dpthin) { kp = k; break; }
    }

    k = kp;
    ka = (k - 2 > 1) ? k - 2 : 1;

    if (k > fixlay + 1 && qdep == 1.0 &&
        P3(p, i, j, k) - P3(p, i, j, k - 1) >= dpthin &&
        P3(theta, i, j, k) - EPSIL > F4(th3d, i, j, k, n) &&
        F4(th3d, i, j, k - 1, n) > F4(th3d, i, j, k, n) &&
        F4(th3d, i, j, ka, n) > F4(th3d, i, j, k, n)) {
      /* entrain the whole layer into the one above, 609-686 */
      q = (P3(p, i, j, k + 1) - P3(p, i, j, k)) / (P3(p, i, j, k + 1) - P3(p, i, j, k - 1));
      F4(temp, i, j, k - 1, n) = F4(temp, i, j, k - 1, n) - q * (F4(temp, i, j, k - 1, n) - F4(temp, i, j, k, n));
      F4(saln, i, j, k - 1, n) = F4(saln, i, j, k - 1, n) - q * (F4(saln, i, j, k - 1, n) - F4(saln, i, j, k, n));
      F4(th3d, i, j, k - 1, n) = sig17(F4(temp, i, j, k - 1, n), F4(saln, i, j, k - 1, n)) - P->thbase;
      for (int t = 1; t <= ntracr; t++)
        TRC(i, j, k - 1, n, t) = TRC(i, j, k - 1, n, t) - q * (TRC(i, j, k - 1, n, t) - TRC(i, j, k, n, t));
      P3(p, i, j, k) = P3(p, i, j, k + 1);
      kp = k - 1;
    } else if (k > fixlay + 1 && qdep == 1.0 &&
               P3(p, i, j, k) - P3(p, i, j, k - 1) >= dpthin &&
               P3(theta, i, j, k) - EPSIL > F4(th3d, i, j, k, n) &&
               F4(th3d, i, j, k - 1, n) > F4(th3d, i, j, k, n)) {
      /* swap with the layer above, 687-811 */
      if (P3(p, i, j, k + 1) - P3(p, i, j, k) <= P3(p, i, j, k) - P3(p, i, j, k - 1)) {
        double st = F4(temp, i, j, k - 1, n), ss = F4(saln, i, j, k - 1, n), sh = F4(th3d, i, j, k - 1, n);
        q = (P3(p, i, j, k + 1) - P3(p, i, j, k)) / (P3(p, i, j, k) - P3(p, i, j, k - 1));
        F4(temp, i, j, k - 1, n) = F4(temp, i, j, k - 1, n) - q * (F4(temp, i, j, k - 1, n) - F4(temp, i, j, k, n));
        F4(saln, i, j, k - 1, n) = F4(saln, i, j, k - 1, n) - q * (F4(saln, i, j, k - 1, n) - F4(saln, i, j, k, n));
        F4(th3d, i, j, k - 1, n) = sig17(F4(temp, i, j, k - 1, n), F4(saln, i, j, k - 1, n)) - P->thbase;
        F4(temp, i, j, k, n) = st;
        F4(saln, i, j, k, n) = ss;
        F4(th3d, i, j, k, n) = sh;
        for (int t = 1; t <= ntracr; t++) {
          double sv = TRC(i, j, k - 1, n, t);
          TRC(i, j, k - 1, n, t) = TRC(i, j, k - 1, n, t) - q * (TRC(i, j, k - 1, n, t) - TRC(i, j, k, n, t));
          TRC(i, j, k, n, t) = sv;
        }
      } else {
        double st = F4(temp, i, j, k, n), ss = F4(saln, i, j, k, n), sh = F4(th3d, i, j, k, n);
        q = (P3(p, i, j, k) - P3(p, i, j, k - 1)) / (P3(p, i, j, k + 1) - P3(p, i, j, k));
        F4(temp, i, j, k, n) = F4(temp, i, j, k, n) + q * (F4(temp, i, j, k - 1, n) - F4(temp, i, j, k, n));
        F4(saln, i, j, k, n) = F4(saln, i, j, k, n) + q * (F4(saln, i, j, k - 1, n) - F4(saln, i, j, k, n));
        F4(th3d, i, j, k, n) = sig17(F4(temp, i, j, k, n), F4(saln, i, j, k, n)) - P->thbase;
        F4(temp, i, j, k - 1, n) = st;
        F4(saln, i, j, k - 1, n) = ss;
        F4(th3d, i, j, k - 1, n) = sh;
        for (int t = 1; t <= ntracr; t++) {
          double sv = TRC(i, j, k, n, t);
          TRC(i, j, k, n, t) = TRC(i, j, k, n, t) + q * (TRC(i, j, k - 1, n, t) - TRC(i, j, k, n, t));
          TRC(i, j, k - 1, n, t) = sv;
        }
      }
    }

    /* unmix a too-light deepest layer, 813-970 */
    k = kp;
    ka = (k - 2 > 1) ? k - 2 : 1;
    {
      double dpthck = dx0k[k - 2] * 0.95;
      if (k > fixlay + 1 && qdep == 1.0 &&
          P3(p, i, j, k) - P3(p, i, j, k - 1) >= dpthin &&
          P3(p, i, j, k) - P3(p, i, j, k - 1) < dpthck &&
          P3(theta, i, j, k) - EPSIL > F4(th3d, i, j, k, n) &&
          P3(theta, i, j, k - 1) < F4(th3d, i, j, k, n) &&
          fabs(P3(theta, i, j, k - 1) - F4(th3d, i, j, k - 1, n)) < P->hybiso &&
          (F4(th3d, i, j, k, n) - F4(th3d, i, j, k - 1, n)) >
              (P3(theta, i, j, k) - P3(theta, i, j, k - 1)) * 0.001) {
        ka = 1;
        for (int kt = k - 2; kt >= 2; kt--) {
          if (F4(th3d, i, j, k - 1, n) - F4(th3d, i, j, kt, n) >=
              P3(theta, i, j, k - 1) - P3(theta, i, j, k - 2)) {
            ka = kt;
            break;
          }
        }
        double delsm = fabs(F4(saln, i, j, ka, n) - F4(saln, i, j, k - 1, n));
        double dels = fabs(F4(saln, i, j, k - 1, n) - F4(saln, i, j, k, n));
        double deltm = fabs(F4(temp, i, j, ka, n) - F4(temp, i, j, k - 1, n));
        double delt = fabs(F4(temp, i, j, k - 1, n) - F4(temp, i, j, k, n));
        q = fmin3(F4(temp, i, j, ka, n), F4(temp, i, j, k - 1, n), F4(temp, i, j, k, n));
        if (q > 6.0) {
          deltm = fmin2(deltm, 6.0 * (P3(theta, i, j, k) - P3(theta, i, j, k - 1)));
        } else {
          deltm = fmin2(deltm, 10.0 * (P3(theta, i, j, k) - P3(theta, i, j, k - 1)));
        }
        delsm = fmin2(delsm, 1.3 * (P3(theta, i, j, k) - P3(theta, i, j, k - 1)));
        qts = 0.0;
        if (delt > EPSIL) qts = fmax2(qts, (fmin2(deltm, 2.0 * delt) - delt) / delt);
        if (dels > EPSIL) qts = fmax2(qts, (fmin2(delsm, 2.0 * dels) - dels) / dels);
        q = (P3(theta, i, j, k) - F4(th3d, i, j, k, n)) /
            (P3(theta, i, j, k) - F4(th3d, i, j, k - 1, n));
        q = fmin2(q, qts / (1.0 + qts));
        q = qhrlx[k] * q;
        p_hat = q * (P3(p, i, j, k + 1) - P3(p, i, j, k));
        P3(p, i, j, k) = P3(p, i, j, k) + p_hat;
        F4(temp, i, j, k, n) = F4(temp, i, j, k, n) + (q / (1.0 - q)) * (F4(temp, i, j, k, n) - F4(temp, i, j, k - 1, n));
        F4(saln, i, j, k, n) = F4(saln, i, j, k, n) + (q / (1.0 - q)) * (F4(saln, i, j, k, n) - F4(saln, i, j, k - 1, n));
        F4(th3d, i, j, k, n) = sig17(F4(temp, i, j, k, n), F4(saln, i, j, k, n)) - P->thbase;
        if (ntracr > 0 && p_hat != 0.0) {
          double qtr = p_hat / fmax2(p_hat, P3(p, i, j, k) - P3(p, i, j, k - 1));
          for (int t = 1; t <= ntracr; t++) {
            if (P->trcflg[t - 1] == 2) {
              TRC(i, j, k, n, t) = TRC(i, j, k, n, t) +
                                   (q / (1.0 - q)) * (TRC(i, j, k, n, t) - TRC(i, j, k - 1, n, t));
            } else {
              TRC(i, j, k - 1, n, t) = TRC(i, j, k - 1, n, t) +
                                       qtr * (TRC(i, j, k, n, t) - TRC(i, j, k - 1, n, t));
            }
          }
        }
      }
    }

    /* the column's decisions: fixed layers (479-573) and the deepest
       inflated layer after the overturn (590-686) */
    if (c->rec_fixlay) {
      long r = (long)(i - 1) + (long)P->idm * (j - 1);
      c->rec_fixlay[r] = fixlay;
      c->rec_kp[r] = kp;
    }

    /* massless or near-massless layers, 972-1009 */
    for (k = kp + 1; k <= kk; k++) {
      if (k <= nhybrd) {
        F4(th3d, i, j, k, n) = F4(th3d, i, j, k - 1, n);
        F4(saln, i, j, k, n) = F4(saln, i, j, k - 1, n);
        F4(temp, i, j, k, n) = F4(temp, i, j, k - 1, n);
      } else if (F4(th3d, i, j, k, n) != P3(theta, i, j, k)) {
        /* hybflg == 0 */
        F4(th3d, i, j, k, n) = fmax2(P3(theta, i, j, k), F4(th3d, i, j, k - 1, n));
        F4(saln, i, j, k, n) = F4(saln, i, j, k - 1, n);
        F4(temp, i, j, k, n) = TOFSIG(F4(th3d, i, j, k, n) + P->thbase, F4(saln, i, j, k, n));
        F4(saln, i, j, k, n) = SOFSIG(F4(th3d, i, j, k, n) + P->thbase, F4(temp, i, j, k, n));
      }
      for (int t = 1; t <= ntracr; t++) TRC(i, j, k, n, t) = TRC(i, j, k - 1, n, t);
    }

    /* old-grid snapshot, 1013-1046 */
    pres[1] = P3(p, i, j, 1);
    for (k = 1; k <= kk; k++) {
      s1d[1][k] = F4(temp, i, j, k, n);
      s1d[2][k] = F4(saln, i, j, k, n);
      for (int t = 1; t <= ntracr; t++) s1d[2 + t][k] = TRC(i, j, k, n, t);
      pres[k + 1] = P3(p, i, j, k + 1);
      dprs[k] = pres[k + 1] - pres[k];
      dpi[k] = fmax2(dprs[k], dpthin);
      if (P->isopcm) {
        if (k <= fixlay) {
          lcm[k] = 0;
        } else {
          lcm[k] = k > nhybrd || dprs[k] <= dpthin ||
                   fabs(F4(th3d, i, j, k, n) - P3(theta, i, j, k)) < P->hybiso;
        }
      }
    }

    /* layer 1 constant thickness, 1051-1063 */
    if (fixlay >= 1) {
      k = 1;
      p_hat = P3(p, i, j, k) + dp0ij[k];
      P3(p, i, j, k + 1) = p_hat;
      for (k = 2; k <= kk; k++) {
        if (P3(p, i, j, k + 1) >= p_hat) break;
        P3(p, i, j, k + 1) = p_hat;
      }
    }

    /* interface relocation, 1065-1494 */
    for (k = 2; k <= nhybrd; k++) {
      if (k <= fixlay) {
        if (k < kk) {
          P3(p, i, j, k + 1) = fmin2(dp0cum[k + 1], P3(p, i, j, kk + 1));
          if (k == fixlay) {
            for (ka = k + 2; ka <= kk; ka++) {
              if (P3(p, i, j, ka) >= P3(p, i, j, k + 1)) break;
              P3(p, i, j, ka) = P3(p, i, j, k + 1);
            }
          }
        }
      } else {
        double dp_top = P3(p, i, j, k) - P3(p, i, j, k - 1) + EPSIL;
        double dp_cen = P3(p, i, j, k + 1) - P3(p, i, j, k) + EPSIL;
        double dp_bot, dp_far, dp_inc;
        if (k == kk) {
          dp_bot = dp_cen;
          dp_far = dp_cen;
        } else if (k == kk - 1) {
          dp_bot = P3(p, i, j, k + 2) - P3(p, i, j, k + 1) + EPSIL;
          dp_far = dp_bot;
        } else {
          dp_bot = P3(p, i, j, k + 2) - P3(p, i, j, k + 1) + EPSIL;
          dp_far = P3(p, i, j, k + 3) - P3(p, i, j, k + 2) + EPSIL;
        }
        if (dp_cen < P->hybthk * fmin2(dp_top, dp_bot)) {
          /* thick-thin-thick, 1131-1159 */
          if (fmin2(dp_top, dp_bot) < 0.2 * fmax2(dp_top, dp_bot)) {
            q = dp_top / (dp_top + dp_bot);
          } else if (P3(theta, i, j, k) > F4(th3d, i, j, k - 1, n) &&
                     P3(theta, i, j, k) < F4(th3d, i, j, k + 1, n)) {
            q = (P3(theta, i, j, k) - F4(th3d, i, j, k + 1, n)) /
                (F4(th3d, i, j, k - 1, n) - F4(th3d, i, j, k + 1, n));
          } else {
            q = (P3(theta, i, j, k) - P3(theta, i, j, k + 1)) /
                (P3(theta, i, j, k - 1) - P3(theta, i, j, k + 1));
          }
          if (dp_top <= dp_bot) {
            dp_inc = (P->hybthk * dp_top - dp_cen) / (1.0 + q * P->hybthk);
          } else {
            dp_inc = (P->hybthk * dp_bot - dp_cen) / (1.0 + (1.0 - q) * P->hybthk);
          }
          p_hat = P3(p, i, j, k) - q * dp_inc;
          P3(p, i, j, k) = (1.0 - qhrlx[k]) * P3(p, i, j, k) + qhrlx[k] * p_hat;
          p_hat = P3(p, i, j, k + 1) + (1.0 - q) * dp_inc;
          P3(p, i, j, k + 1) = (1.0 - qhrlx[k]) * P3(p, i, j, k + 1) + qhrlx[k] * p_hat;
        } else if (dp_cen + dp_bot < P->hybthk * fmin2(dp_top, dp_far)) {
          /* thick-thin-thin-thick, 1178-1208 */
          if (fmin2(dp_top, dp_far) < 0.2 * fmax2(dp_top, dp_far)) {
            q = dp_top / (dp_top + dp_far);
          } else if (0.5 * (P3(theta, i, j, k) + P3(theta, i, j, k + 1)) > F4(th3d, i, j, k - 1, n) &&
                     0.5 * (P3(theta, i, j, k) + P3(theta, i, j, k + 1)) < F4(th3d, i, j, k + 2, n)) {
            q = (0.5 * (P3(theta, i, j, k) + P3(theta, i, j, k + 1)) - F4(th3d, i, j, k + 2, n)) /
                (F4(th3d, i, j, k - 1, n) - F4(th3d, i, j, k + 2, n));
          } else {
            q = (0.5 * (P3(theta, i, j, k) + P3(theta, i, j, k + 1)) - P3(theta, i, j, k + 2)) /
                (P3(theta, i, j, k - 1) - P3(theta, i, j, k + 2));
          }
          if (dp_top <= dp_far) {
            dp_inc = (P->hybthk * dp_top - (dp_cen + dp_bot)) / (1.0 + q * P->hybthk);
          } else {
            dp_inc = (P->hybthk * dp_far - (dp_cen + dp_bot)) / (1.0 + (1.0 - q) * P->hybthk);
          }
          p_hat = P3(p, i, j, k) - q * dp_inc;
          P3(p, i, j, k) = (1.0 - qhrlx[k]) * P3(p, i, j, k) + qhrlx[k] * p_hat;
          p_hat = P3(p, i, j, k + 2) + (1.0 - q) * dp_inc;
          P3(p, i, j, k + 2) = (1.0 - qhrlx[k]) * P3(p, i, j, k + 2) + qhrlx[k] * p_hat;
        } else if (F4(th3d, i, j, k, n) > P3(theta, i, j, k) + EPSIL && k > fixlay + 1) {
          /* too dense, 1237-1387 */
          int ltop;
          if (rthin * (P3(p, i, j, k) - P3(p, i, j, k - 1)) <= (P3(p, i, j, k + 1) - P3(p, i, j, k))) {
            ltop = 1;
          } else if (rthin * (P3(p, i, j, k + 1) - P3(p, i, j, k)) <= (P3(p, i, j, k) - P3(p, i, j, k - 1))) {
            ltop = 0;
          } else {
            ltop = P3(theta, i, j, k - 1) - F4(th3d, i, j, k - 1, n) >=
                   F4(th3d, i, j, k, n) - P3(theta, i, j, k);
          }
          if (F4(th3d, i, j, k - 1, n) >= P3(theta, i, j, k - 1) || !ltop ||
              P3(p, i, j, k) <= dp0cum[k] + ONEM) {
            if ((P3(theta, i, j, k) - F4(th3d, i, j, k - 1, n)) <= EPSIL) {
              /* p_hat0 keeps the value the tail of the previous k set
               * (1479); the blocking branch below reads it (1301-1330) */
              p_hat = P3(p, i, j, k - 1) + dp0ij[k - 1];
            } else {
              q = (P3(theta, i, j, k) - F4(th3d, i, j, k, n)) /
                  (P3(theta, i, j, k) - F4(th3d, i, j, k - 1, n));
              p_hat0 = P3(p, i, j, k) + q * (P3(p, i, j, k + 1) - P3(p, i, j, k));
              if (k == fixlay + 2) {
                p_hat = P3(p, i, j, k - 1) + fmax2(p_hat0 - P3(p, i, j, k - 1), dp0ij[k - 1]);
              } else {
                p_hat = P3(p, i, j, k - 1) + CUSHN(p_hat0 - P3(p, i, j, k - 1), dp0ij[k - 1]);
              }
            }
            p_hat = fmin2(p_hat, P3(p, i, j, kk + 1));
            if (k <= fixlay + 2) {
              /* do nothing */
            } else if (p_hat >= P3(p, i, j, k) &&
                       P3(p, i, j, k - 1) > dp0cum[k - 1] + TENM &&
                       (P3(p, i, j, kk + 1) - P3(p, i, j, k - 1) < thkbop ||
                        P3(p, i, j, k - 1) - P3(p, i, j, k - 2) > qqmx * dp0ij[k - 2])) {
              if (k == fixlay + 3) {
                p_hat2 = P3(p, i, j, k - 2) +
                         fmax2(P3(p, i, j, k - 1) - p_hat + p_hat0 - P3(p, i, j, k - 2), dp0ij[k - 2]);
              } else {
                p_hat2 = P3(p, i, j, k - 2) +
                         CUSHN(P3(p, i, j, k - 1) - p_hat + p_hat0 - P3(p, i, j, k - 2), dp0ij[k - 2]);
              }
              if (p_hat2 < P3(p, i, j, k - 1) - ONEMM) {
                P3(p, i, j, k - 1) = (1.0 - qhrlx[k - 1]) * P3(p, i, j, k - 1) +
                                     qhrlx[k - 1] * fmax2(p_hat2, 2.0 * P3(p, i, j, k - 1) - p_hat);
                p_hat = P3(p, i, j, k - 1) + CUSHN(p_hat0 - P3(p, i, j, k - 1), dp0ij[k - 1]);
              } else if (k <= fixlay + 3) {
                /* do nothing */
              } else if (P3(p, i, j, k - 2) > dp0cum[k - 2] + TENM &&
                         (P3(p, i, j, kk + 1) - P3(p, i, j, k - 2) < thkbop ||
                          P3(p, i, j, k - 2) - P3(p, i, j, k - 3) > qqmx * dp0ij[k - 3])) {
                if (k == fixlay + 4) {
                  p_hat3 = P3(p, i, j, k - 3) +
                           fmax2(P3(p, i, j, k - 2) - p_hat + p_hat0 - P3(p, i, j, k - 3), dp0ij[k - 3]);
                } else {
                  p_hat3 = P3(p, i, j, k - 3) +
                           CUSHN(P3(p, i, j, k - 2) - p_hat + p_hat0 - P3(p, i, j, k - 3), dp0ij[k - 3]);
                }
                if (p_hat3 < P3(p, i, j, k - 2) - ONEMM) {
                  P3(p, i, j, k - 2) = (1.0 - qhrlx[k - 2]) * P3(p, i, j, k - 2) +
                                       qhrlx[k - 2] * fmax2(p_hat3, 2.0 * P3(p, i, j, k - 2) - P3(p, i, j, k - 1));
                  p_hat2 = P3(p, i, j, k - 2) +
                           CUSHN(P3(p, i, j, k - 1) - p_hat + p_hat0 - P3(p, i, j, k - 2), dp0ij[k - 2]);
                  if (p_hat2 < P3(p, i, j, k - 1) - ONEMM) {
                    P3(p, i, j, k - 1) = (1.0 - qhrlx[k - 1]) * P3(p, i, j, k - 1) +
                                         qhrlx[k - 1] * fmax2(p_hat2, 2.0 * P3(p, i, j, k - 1) - p_hat);
                    p_hat = P3(p, i, j, k - 1) + CUSHN(p_hat0 - P3(p, i, j, k - 1), dp0ij[k - 1]);
                  }
                }
              }
            }
            if (p_hat < P3(p, i, j, k)) {
              P3(p, i, j, k) = (1.0 - qhrlx[k]) * P3(p, i, j, k) + qhrlx[k] * p_hat;
            }
          }
        } else if (F4(th3d, i, j, k, n) < P3(theta, i, j, k) - EPSIL) {
          /* too light, 1389-1475 */
          if (P3(p, i, j, k + 1) < P3(p, i, j, kk + 1)) {
            int ltop;
            if (rthin * (P3(p, i, j, k + 1) - P3(p, i, j, k)) <= (P3(p, i, j, k + 2) - P3(p, i, j, k + 1))) {
              ltop = 1;
            } else if (rthin * (P3(p, i, j, k + 2) - P3(p, i, j, k + 1)) <= (P3(p, i, j, k + 1) - P3(p, i, j, k))) {
              ltop = 0;
            } else {
              ltop = P3(theta, i, j, k) - F4(th3d, i, j, k, n) >=
                     F4(th3d, i, j, k + 1, n) - P3(theta, i, j, k + 1);
            }
            if (F4(th3d, i, j, k + 1, n) <= P3(theta, i, j, k + 1) || ltop ||
                P3(p, i, j, k + 1) <= dp0cum[k + 1] + ONEM) {
              if ((F4(th3d, i, j, k + 1, n) - P3(theta, i, j, k)) <= EPSIL) {
                p_hat = P3(p, i, j, k + 2);
              } else {
                q = (F4(th3d, i, j, k, n) - P3(theta, i, j, k)) /
                    (F4(th3d, i, j, k + 1, n) - P3(theta, i, j, k));
                p_hat = P3(p, i, j, k + 1) + q * (P3(p, i, j, k) - P3(p, i, j, k + 1));
              }
              int k3 = (k + 3 < kk + 1) ? k + 3 : kk + 1;
              if (P3(p, i, j, k3) < P3(p, i, j, kk + 1)) {
                if (P3(p, i, j, kk + 1) - P3(p, i, j, k) > dp0ij[k] + dp0ij[k + 1]) {
                  p_hat = P3(p, i, j, k + 2) - CUSHN(P3(p, i, j, k + 2) - p_hat, dp0ij[k + 1]);
                }
                p_hat = P3(p, i, j, k) + fmax2(p_hat - P3(p, i, j, k), dp0ij[k]);
                p_hat = fmin2(p_hat, fmax2(0.5 * (P3(p, i, j, k + 1) + P3(p, i, j, k + 2)),
                                           P3(p, i, j, k + 2) - dp0ij[k + 1]));
              } else {
                p_hat = fmin2(P3(p, i, j, k + 2), p_hat);
              }
              if (p_hat > P3(p, i, j, k + 1)) {
                P3(p, i, j, k + 1) = (1.0 - qhrlx[k + 1]) * P3(p, i, j, k + 1) + qhrlx[k + 1] * p_hat;
              }
            }
          }
        }
        /* layer above still too thin, 1479-1490 */
        p_hat0 = fmin2(P3(p, i, j, k - 1) + dp0ij[k - 1], P3(p, i, j, kk + 1));
        if (p_hat0 > P3(p, i, j, k)) {
          p_hat = (1.0 - qhrlx[k - 1]) * P3(p, i, j, k) + qhrlx[k - 1] * p_hat0;
          P3(p, i, j, k) = fmin2(p_hat, P3(p, i, j, k + 1));
        }
      }
    }

    /* maximum thickness dx0k, 1496-1509 */
    for (k = 2; k <= kk - 1; k++) {
      if ((P3(p, i, j, k + 1) - P3(p, i, j, k)) > dx0k[k - 1]) {
        P3(p, i, j, k + 1) = P3(p, i, j, k) + dx0k[k - 1];
        lcm[k] = 0;
        lcm[k - 1] = 0;
      }
    }

    /* rebuild dp, p; remap, 1523-1548 */
    prsf[1] = P3(p, i, j, 1);
    for (k = 1; k <= kk; k++) {
      F4(dp, i, j, k, n) = fmax2(P3(p, i, j, k + 1) - prsf[k], 0.0);
      prsf[k + 1] = prsf[k] + F4(dp, i, j, k, n);
      P3(p, i, j, k + 1) = prsf[k + 1];
    }
    remap_dispatch(P->hybmap, s1d, pres, dprs, dpi, lcm, c1d, f1d, prsf, kk, nums1d, dpthin);

    /* write back, 1549-1591 */
    for (k = 1; k <= kk; k++) {
      F4(temp, i, j, k, n) = f1d[1][k];
      F4(saln, i, j, k, n) = f1d[2][k];
      F4(th3d, i, j, k, n) = sig17(F4(temp, i, j, k, n), F4(saln, i, j, k, n)) - P->thbase;
      for (int t = 1; t <= ntracr; t++) TRC(i, j, k, n, t) = f1d[2 + t][k];
    }
  }
#undef QQ
#undef CUSHN
}

/* ---------------------------------------------------------------------- */
/* hybgencj, hybgen.F90:1909-2274 (Robert-Asselin branch)                   */
/* ---------------------------------------------------------------------- */
static void hybgencj(const ctx_t* c, int m, int n, int j) {
  const hyb_params* P = c->P;
  const hyb_fields* F = c->F;
  const int kk = c->K, nhybrd = P->nhybrd, ntracr = P->ntracr;
  double *p = F->p, *dp = F->dp, *dpo = F->dpo, *temp = F->temp, *saln = F->saln, *th3d = F->th3d;
  const double* theta = F->theta;
  static __thread colv s1d[HYB_TRMAX + 5], f1d[HYB_TRMAX + 5], c1d[3 * (HYB_TRMAX + 5)];
  int lcm[KD] = {0};
  double dpi[KD], dprs[KD], pres[KD], prsf[KD], dp0ij[KD], dp0cum[KD];
  const double dpthin = 0.001 * ONEMM;
  const int nums1d = ntracr + 2;

  if (!P->isopcm) {
    for (int k = 1; k <= nhybrd; k++) lcm[k] = 0;
    for (int k = nhybrd + 1; k <= kk; k++) lcm[k] = 1;
  }
  for (int i = 1; i <= P->ii; i++) {
    if (P2(F->ip, i, j) == 0) continue;
    double q;
    int k, fixlay;
    double qdep = fmax2(0.0, fmin2(1.0, (P2(F->depths, i, j) - P->dsns) / (P->dpns - P->dsns)));
    dp0cum[1] = 0.0;
    dp0ij[1] = qdep * P->dp0k[0] + (1.0 - qdep) * P->ds0k[0];
    dp0cum[2] = dp0cum[1] + dp0ij[1];
    for (k = 2; k <= kk; k++) {
      double d = P->dp0k[k - 1];
      if (d <= P->dp00i) {
        q = d;
      } else {
        q = fmax2(P->dp00i, d * d / fmax2(d, P3(p, i, j, k) - dp0cum[k]));
      }
      dp0ij[k] = fmin2(q, qdep * d + (1.0 - qdep) * P->ds0k[k - 1]);
      dp0cum[k + 1] = dp0cum[k] + dp0ij[k];
    }
    fixlay = 1;
    for (k = 2; k <= nhybrd; k++) {
      if (dp0cum[k] >= P2(F->topiso, i, j)) break;
      fixlay = fixlay + 1;
    }
    pres[1] = 0.0;
    prsf[1] = 0.0;
    for (k = 1; k <= kk; k++) {
      s1d[1][k] = F4(temp, i, j, k, m);
      s1d[2][k] = F4(saln, i, j, k, m);
      for (int t = 1; t <= ntracr; t++) s1d[2 + t][k] = TRC(i, j, k, m, t);
      dprs[k] = F4(dp, i, j, k, m);
      dpi[k] = fmax2(dprs[k], dpthin);
      pres[k + 1] = pres[k] + F4(dp, i, j, k, m);
      double dpold = F4(dpo, i, j, k, n);
      double dpmid = F4(dpo, i, j, k, m);
      double dpnew = F4(dp, i, j, k, n);
      q = 0.5 * P->ra2fac * (dpold + dpnew - 2.0 * dpmid);
      F4(dp, i, j, k, m) = dpmid + q;
      prsf[k + 1] = prsf[k] + F4(dp, i, j, k, m);
      P3(p, i, j, k + 1) = prsf[k + 1];
      if (P->isopcm) {
        if (k <= fixlay) {
          lcm[k] = 0;
        } else {
          lcm[k] = k > nhybrd || dprs[k] <= dpthin ||
                   fabs(F4(th3d, i, j, k, m) - P3(theta, i, j, k)) < P->hybiso;
        }
      }
    }
    remap_dispatch(P->hybmap, s1d, pres, dprs, dpi, lcm, c1d, f1d, prsf, kk, nums1d, dpthin);
    for (k = 1; k <= kk; k++) {
      F4(temp, i, j, k, m) = f1d[1][k];
      F4(saln, i, j, k, m) = f1d[2][k];
      F4(th3d, i, j, k, m) = sig17(F4(temp, i, j, k, m), F4(saln, i, j, k, m)) - P->thbase;
      for (int t = 1; t <= ntracr; t++) TRC(i, j, k, m, t) = f1d[2 + t][k];
    }
  }
}

/* ---------------------------------------------------------------------- */
/* hybgenbj, hybgen.F90:1758-1907                                          */
/* ---------------------------------------------------------------------- */
static void hybgenb_point(const ctx_t* c, int nl, int i, int j, double* pu,
                          const double* dpu, double* u, const double* ubavg) {
  const hyb_params* P = c->P;
  const int kk = c->K;
  colv s1d[2], f1d[2], c1d[8];
  int lcm[KD] = {0};
  double dpi[KD], dprs[KD], pres[KD], prsf[KD];
  const double dpthin = 0.001 * ONEMM;
  for (int k = 1; k <= kk; k++) lcm[k] = 0;
  pres[1] = P3(pu, i, j, 1);
  for (int k = 1; k <= kk; k++) {
    s1d[1][k] = F4(u, i, j, k, nl);
    pres[k + 1] = P3(pu, i, j, k + 1);
    dprs[k] = pres[k + 1] - pres[k];
    dpi[k] = fmax2(dprs[k], dpthin);
  }
  prsf[1] = P3(pu, i, j, 1);
  for (int k = 1; k <= kk; k++) {
    P3(pu, i, j, k + 1) = P3(pu, i, j, k) + F4(dpu, i, j, k, nl);
    prsf[k + 1] = P3(pu, i, j, k + 1);
  }
  if (P->hybmap == 0) {
    pcm_remap(s1d, pres, dprs, f1d, prsf, kk, kk, 1, dpthin);
  } else if (P->hybmap == 1 && P->hybiso > 2.0) {
    plm_coefs(s1d, dprs, lcm, c1d, kk, 1, dpthin);
    plm_remap(s1d, pres, dprs, c1d, f1d, prsf, kk, kk, 1, dpthin);
  } else {
    weno_coefs(s1d, dpi, lcm, c1d, kk, 1, dpthin);
    weno_remap(s1d, pres, dprs, c1d, f1d, prsf, kk, kk, 1, dpthin);
  }
  for (int k = 1; k <= kk; k++) {
    if (dpi[k] > dpthin || prsf[k] <= prsf[kk + 1] - ONEMM) {
      F4(u, i, j, k, nl) = f1d[1][k];
    } else {
      F4(u, i, j, k, nl) = -P3(ubavg, i, j, nl);
    }
  }
}

static void hybgenbj(const ctx_t* c, int nl, int j) {
  const hyb_fields* F = c->F;
  for (int i = 1; i <= c->P->ii; i++)
    if (P2(F->iu, i, j) != 0) hybgenb_point(c, nl, i, j, F->pu, F->dpu, F->u, F->ubavg);
  for (int i = 1; i <= c->P->ii; i++)
    if (P2(F->iv, i, j) != 0) hybgenb_point(c, nl, i, j, F->pv, F->dpv, F->v, F->vbavg);
}

/* old u/v interfaces, hybgen.F90:171-195 (and 246-270 at level m) */
static void old_puv_row(const ctx_t* c, int l, int j) {
  const hyb_fields* F = c->F;
  const int kk = c->K, ii = c->P->ii;
  for (int i = 1; i <= ii; i++) {
    if (P2(F->iu, i, j) != 0) { P3(F->pu, i, j, 1) = 0.0; P3(F->pu, i, j, 2) = F4(F->dpu, i, j, 1, l); }
    if (P2(F->iv, i, j) != 0) { P3(F->pv, i, j, 1) = 0.0; P3(F->pv, i, j, 2) = F4(F->dpv, i, j, 1, l); }
  }
  for (int k = 2; k <= kk; k++)
    for (int i = 1; i <= ii; i++) {
      if (P2(F->iu, i, j) != 0) P3(F->pu, i, j, k + 1) = P3(F->pu, i, j, k) + F4(F->dpu, i, j, k, l);
      if (P2(F->iv, i, j) != 0) P3(F->pv, i, j, k + 1) = P3(F->pv, i, j, k) + F4(F->dpv, i, j, k, l);
    }
}

/* xctilr(p(:,:,2),1,kk,1,1,halo_ps), mod_xc_sm.h:1337-1429 */
static void p_halo(const ctx_t* c) {
  const hyb_params* P = c->P;
  double* p = c->F->p;
  const int ii = P->ii, jj = P->jj, kk = c->K;
  for (int k = 2; k <= kk + 1; k++) {
    for (int i = 1; i <= ii; i++) {
      if (P->nreg <= 2) {
        P3(p, i, 0, k) = P->vland;
        P3(p, i, jj + 1, k) = P->vland;
      } else {
        P3(p, i, 0, k) = P3(p, i, jj, k);
        P3(p, i, jj + 1, k) = P3(p, i, 1, k);
      }
    }
  }
  for (int k = 2; k <= kk + 1; k++) {
    for (int j = 0; j <= jj + 1; j++) {
      if (P->nreg == 0 || P->nreg == 4) {
        P3(p, 0, j, k) = P->vland;
        P3(p, ii + 1, j, k) = P->vland;
      } else {
        P3(p, 0, j, k) = P3(p, ii, j, k);
        P3(p, ii + 1, j, k) = P3(p, 1, j, k);
      }
    }
  }
}

/* dpudpvj, dpudpv.F90:73-107 (margin 0) */
static void dpudpv_row(const ctx_t* c, int l, int j) {
  const hyb_fields* F = c->F;
  const int kk = c->K, ii = c->P->ii;
  const double* p = F->p;
  for (int k = 1; k <= kk; k++)
    for (int i = 1; i <= ii; i++) {
      if (P2(F->iu, i, j) != 0) {
        double du = P2(F->depthu, i, j);
        F4(F->dpu, i, j, k, l) = fmax2(0.0, fmin2(du, 0.5 * (P3(p, i, j, k + 1) + P3(p, i - 1, j, k + 1))) -
                                                 fmin2(du, 0.5 * (P3(p, i, j, k) + P3(p, i - 1, j, k))));
      }
      if (P2(F->iv, i, j) != 0) {
        double dv = P2(F->depthv, i, j);
        F4(F->dpv, i, j, k, l) = fmax2(0.0, fmin2(dv, 0.5 * (P3(p, i, j, k + 1) + P3(p, i, j - 1, k + 1))) -
                                                 fmin2(dv, 0.5 * (P3(p, i, j, k) + P3(p, i, j - 1, k))));
      }
    }
}

static int check(const hyb_params* P, const hyb_fields* F, int m, int n, int raflag) {
  if (!P || !F) return HYB_ENULL;
  if (P->kdm < 4 || P->kdm > HYB_KMAX || P->ntracr < 0 || P->ntracr > HYB_TRMAX) return HYB_EINVAL;
  if (P->hybflg != 0 || P->mxlmy) return HYB_EUNSUPPORTED;
  if (P->hybmap < 0 || P->hybmap > 3) return HYB_EINVAL;
  if (P->nhybrd < 1 || P->nhybrd > P->kdm) return HYB_EINVAL;
  if (m < 1 || m > 2 || n < 1 || n > 2 || (raflag && m == n)) return HYB_EINVAL;
  if (raflag && !F->dpo) return HYB_ENULL;
  return 0;
}

/* hybgen, hybgen.F90:14-306 */
static int run(const hyb_params* P, const hyb_fields* F, int m, int n,
               int raflag, int nthreads, int32_t* rec_fixlay, int32_t* rec_kp) {
  int rc = check(P, F, m, n, raflag);
  if (rc) return rc;
  eos_init();
  ctx_t cc, *c = &cc;
  c->P = P; c->F = F;
  c->rec_fixlay = rec_fixlay;
  c->rec_kp = rec_kp;
  c->NI = P->idm + 2 * P->nbdy;
  c->NJ = P->jdm + 2 * P->nbdy;
  c->NIJ = c->NI * c->NJ;
  c->K = P->kdm;
  c->nb = P->nbdy;
  const int kk = P->kdm, ii = P->ii, jj = P->jj;
  (void)nthreads;

  /* diagnostic tracers 801-803, before (44-87) */
  const int itr[3] = {P->itracr801, P->itracr802, P->itracr803};
  double* src[3] = {F->dp, F->temp, F->saln};
  for (int d = 0; d < 3; d++) {
    int t = itr[d];
    if (t == 0) continue;
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int j = 1; j <= jj; j++)
      for (int k = 1; k <= kk; k++)
        for (int i = 1; i <= ii; i++) TRC(i, j, k, n, t) = F4(src[d], i, j, k, n);
  }

#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int j = 1; j <= jj; j++) hybgenaj(c, m, n, j);

  /* diagnostic tracers 801-803, after (97-164) */
  for (int d = 0; d < 3; d++) {
    int t = itr[d];
    if (t == 0) continue;
    double q = (d == 0) ? QONEM * 86400.0 / P->baclin : 86400.0 / P->baclin;
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int j = 1; j <= jj; j++)
      for (int k = 1; k <= kk; k++)
        for (int i = 1; i <= ii; i++)
          if (P2(F->ip, i, j) != 0)
            TRC(i, j, k, n, t) = q * (F4(src[d], i, j, k, n) - TRC(i, j, k, n, t));
  }

#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int j = 1; j <= jj; j++) old_puv_row(c, n, j);
  p_halo(c);
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int j = 1; j <= jj; j++) dpudpv_row(c, n, j);
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int j = 1; j <= jj; j++) hybgenbj(c, n, j);

  if (raflag) {
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int j = 1; j <= jj; j++) hybgencj(c, m, n, j);
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int j = 1; j <= jj; j++) old_puv_row(c, m, j);
    p_halo(c);
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int j = 1; j <= jj; j++) dpudpv_row(c, m, j);
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int j = 1; j <= jj; j++) hybgenbj(c, m, j);
  }
  return 0;
}

int oracle_hybgen(const hyb_params* P, const hyb_fields* F, int m, int n, int raflag) {
  return run(P, F, m, n, raflag, 1, NULL, NULL);
}

int oracle_hybgen_omp(const hyb_params* P, const hyb_fields* F, int m, int n,
                      int raflag, int nthreads) {
  return run(P, F, m, n, raflag, nthreads < 1 ? 1 : nthreads, NULL, NULL);
}

int oracle_hybgen_decisions(const hyb_params* P, const hyb_fields* F, int m, int n,
                            int raflag, int nthreads, int32_t* fixlay, int32_t* kp) {
  if (!fixlay || !kp) return HYB_ENULL;
  return run(P, F, m, n, raflag, nthreads < 1 ? 1 : nthreads, fixlay, kp);
}
