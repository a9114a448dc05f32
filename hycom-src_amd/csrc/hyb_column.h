// hyb_column.h -- per-column hybgen work, FP64, for one lane.
//
// One water column (p-point) or one u/v point per lane; lanes run along i
// so that every k-level access of the Fortran-layout fields is a coalesced
// row.  The statement and operation order of every floating-point
// expression follows the reference text (cited per block) so that, compiled
// with -ffp-contract=off, the result is bit-identical to HYCOM's hybgen.
// The same code compiles for the host (HYB_HD), which lets the CPU suite
// check this exact logic without a GPU.
#pragma once

#include "hyb_math.h"

namespace hyb {

// Kernel-side copy of hyb_params (include/hycom_hybgen.h) plus derived
// scalars.  Passed by value as a kernel argument.
struct KParams {
  int32_t ii, jj, kk, nhybrd, ntracr, hybmap, isopcm, ldenfix, nreg;
  int32_t itr801, itr802, itr803;
  int32_t precision;  // HYB_PREC_FP64 or HYB_PREC_MIXED (remap arithmetic)
  double thbase, hybrlx, hybiso, hybthn, hybthk, thkbop, dp00i, dpns, dsns;
  double ra2fac, vland, diag_q801, diag_q80x;
  double dp0k[64], ds0k[64], dx0k[64];
  double dp0cum[65];  // dp0cum(k), k=1..kk+1, of columns with qdep==1
  uint64_t trc_t;  // bit t-1 set when trcflg(t)==2 (tracer follows T)
};

// Field view: raw pointers in the mod_cb_arrays layout.
struct View {
  double *dp, *dpo, *dpu, *dpv, *temp, *saln, *th3d, *u, *v, *tracer;
  double *p, *pu, *pv;
  const double *theta, *depths, *topiso, *depthu, *depthv, *ubavg, *vbavg;
  const int32_t *ip, *iu, *iv;
  int64_t NI, NIJ;  // row length and plane size including halo
  int32_t nb, K;
  int32_t *rec_fixlay, *rec_kp;  // optional per-column decision record
  int32_t rec_ld;                // leading dimension of the record (idm)
};

HYB_HD int64_t ij_of(const View& V, int i, int j) {
  return (int64_t)(i - 1 + V.nb) + V.NI * (int64_t)(j - 1 + V.nb);
}
// First i of a row's lane mapping with 128-byte-aligned wave rows: lane 0
// of the row's first block sits on a 128-byte boundary of the row (of
// every level when NI*NJ is a multiple of 16 and the array is 128-byte
// aligned), so the 64 columns of a wave are four whole cache lines instead
// of five partial ones.  Kernels mapped this way launch
// ceil((ii + kAlignPad) / width) blocks per row and skip i < 1 and i > ii.
constexpr int kAlignPad = 15;
HYB_HD int i_aligned0(const View& V, int j) {
  return 1 - (int)(((int64_t)V.nb + V.NI * (int64_t)(j - 1 + V.nb)) & 15);
}

// offsets of level k (1-based), time level l, tracer t
HYB_HD int64_t lev(const View& V, int k) { return V.NIJ * (int64_t)(k - 1); }
HYB_HD int64_t lev4(const View& V, int k, int l) {
  return V.NIJ * ((int64_t)(k - 1) + (int64_t)V.K * (l - 1));
}
HYB_HD int64_t levt(const View& V, int k, int l, int t) {
  return V.NIJ * ((int64_t)(k - 1) + (int64_t)V.K * ((l - 1) + 2 * (int64_t)(t - 1)));
}

HYB_HD double dmin3(double a, double b, double c) { return dmin(dmin(a, b), c); }
HYB_HD double dabs(double a) { return a < 0.0 ? -a : a; }
HYB_HD double dsign(double a, double b) {  // Fortran sign(a,b)
  double x = dabs(a);
  return signbit(b) ? -x : x;
}
HYB_HD bool lbit(uint64_t m, int k) { return (m >> (k - 1)) & 1ull; }
HYB_HD uint64_t setbit(uint64_t m, int k, bool v) {
  uint64_t b = 1ull << (k - 1);
  return v ? (m | b) : (m & ~b);
}

// ---------------------------------------------------------------------
// remap of one field on one column (hybgen.F90:2276-3107).
// s[1..kk] input values; pres[1..kk+1] old interfaces; dprs[k] =
// pres(k+1)-pres(k); dpi[k] = max(dprs,thin); lcm bit mask; prsf[1..kk+1]
// new interfaces.  Calls out(k, value) for k = 1..kk in order.
// c1,c2,c3 are caller scratch of length >= kk+1.
// ---------------------------------------------------------------------
template <int METHOD>
HYB_HD void coefs(int kk, const double* s, const double* dprs, const double* dpi,
                  uint64_t lcm, double thin, double* c1, double* c2, double* c3) {
  if (METHOD == 1) {  // hybgen_plm_coefs, 2393-2458 (takes dprs)
    c1[1] = 0.0;
    c1[kk] = 0.0;
    for (int k = 2; k <= kk - 1; k++) {
      if (lbit(lcm, k) || dprs[k] <= thin) {
        c1[k] = 0.0;
      } else {
        double qcen = dprs[k] / (dprs[k] + 0.5 * (dprs[k - 1] + dprs[k + 1]));
        double ztop = 2.0 * (s[k] - s[k - 1]);
        double zbot = 2.0 * (s[k + 1] - s[k]);
        double zcen = qcen * (s[k + 1] - s[k - 1]);
        if (ztop * zbot > 0.0)
          c1[k] = dsign(dmin3(dabs(zcen), dabs(zbot), dabs(ztop)), zbot);
        else
          c1[k] = 0.0;
      }
    }
  } else if (METHOD == 2) {  // hybgen_ppm_coefs, 2587-2706 (takes dpi)
    const double* dp = dpi;
    // as -> c3 scratch, al -> c1, ar -> c2 (overwritten by the coefficients)
    double* as = c3;
    double* al = c1;
    double* ar = c2;
    as[1] = 0.0;
    for (int j = 2; j <= kk - 1; j++) {
      if (lbit(lcm, j) || dp[j] <= thin) {
        as[j] = 0.0;
      } else {
        double slj = s[j] - s[j - 1];
        double srj = s[j + 1] - s[j];
        if (slj * srj > 0.0) {
          double dpjp_m = dp[j - 1] + dp[j];
          double dpjp_j = dp[j] + dp[j + 1];
          double dp2jp_m = dp[j - 1] + dpjp_m;
          double dpj2p_j = dpjp_j + dp[j + 1];
          double qdpjp_j = 1.0 / dpjp_j;
          double qdpjp_m = 1.0 / dpjp_m;
          double dpq3;
          if (j == 2) dpq3 = dp[2] / (dp[1] + dpjp_j);
          else        dpq3 = dp[j] / (dp[j - 1] + dpjp_j);
          double scj = dpq3 * (dp2jp_m * srj * qdpjp_j + dpj2p_j * slj * qdpjp_m);
          as[j] = dsign(dmin3(dabs(2.0 * slj), dabs(scj), dabs(2.0 * srj)), scj);
        } else {
          as[j] = 0.0;
        }
      }
    }
    as[kk] = 0.0;
    al[1] = s[1];
    ar[1] = s[1];
    al[2] = s[1];
    for (int j = 3; j <= kk - 1; j++) {
      double dpjp_m2 = dp[j - 2] + dp[j - 1];
      double dpjp_m = dp[j - 1] + dp[j];
      double dpjp_j = dp[j] + dp[j + 1];
      double dp2jp_m = dp[j - 1] + dpjp_m;
      double dpj2p_m = dpjp_m + dp[j];
      double qdpjp_m = 1.0 / dpjp_m;
      double qdp2jp_m = 1.0 / dp2jp_m;
      double qdpj2p_m = 1.0 / dpj2p_m;
      double qdp4 = 1.0 / (dpjp_m2 + dpjp_j);
      al[j] = s[j - 1] + dp[j - 1] * (s[j] - s[j - 1]) * qdpjp_m +
              qdp4 * (2.0 * dp[j] * dp[j - 1] * qdpjp_m * (s[j] - s[j - 1]) *
                          (dpjp_m2 * qdp2jp_m - dpjp_j * qdpj2p_m) -
                      dp[j - 1] * as[j] * dpjp_m2 * qdp2jp_m +
                      dp[j] * as[j - 1] * dpjp_j * qdpj2p_m);
      ar[j - 1] = al[j];
    }
    ar[kk - 1] = s[kk];
    al[kk] = s[kk];
    ar[kk] = s[kk];
    for (int j = 2; j <= kk - 1; j++) {
      if (lbit(lcm, j) || dp[j] <= thin) {
        al[j] = s[j];
        ar[j] = s[j];
      } else if ((s[j + 1] - s[j]) * (s[j] - s[j - 1]) <= 0.0) {
        al[j] = s[j];
        ar[j] = s[j];
      } else {
        double da = ar[j] - al[j];
        double a6 = 6.0 * s[j] - 3.0 * (al[j] + ar[j]);
        if (da * a6 > da * da) {
          al[j] = 3.0 * s[j] - 2.0 * ar[j];
        } else if (da * a6 < -(da * da)) {
          ar[j] = 3.0 * s[j] - 2.0 * al[j];
        }
      }
    }
    for (int j = 1; j <= kk; j++) {
      double a = al[j], b = ar[j];
      if (a != b) {
        c1[j] = a;
        c2[j] = b - a;
        c3[j] = 6.0 * s[j] - 3.0 * (a + b);
      } else {
        c1[j] = a;
        c2[j] = 0.0;
        c3[j] = 0.0;
      }
    }
  } else if (METHOD == 3) {  // hybgen_weno_coefs, 2845-2968 (takes dpi)
    const double* dp = dpi;
    const double dsmll = 1.0e-8;
    // zw(j,1) -> c3[j], zw(j,2) -> zw2[j] (tail of c3); zw(j,3) is
    // recomputed where it is read.
    double* zw2 = c3 + (kk + 2);
    c1[1] = s[1];
    c2[1] = s[1];
    c3[1] = 0.0;
    zw2[1] = 0.0;
    for (int j = 2; j <= kk - 1; j++) {
      if (lbit(lcm, j) || dp[j] <= thin) {
        c1[j] = s[j];
        c2[j] = s[j];
        c3[j] = 0.0;
        zw2[j] = 0.0;
      } else {
        double zw3_j = (1.0 / (dp[j - 1] + dp[j])) * (s[j] - s[j - 1]);
        double zw3_jp = (1.0 / (dp[j] + dp[j + 1])) * (s[j + 1] - s[j]);
        double q001 = dp[j] * zw3_jp;
        double q002 = dp[j] * zw3_j;
        if (q001 * q002 < 0.0) { q001 = 0.0; q002 = 0.0; }
        double dpjm2jp = dp[j - 1] + 2.0 * dp[j] + dp[j + 1];
        double q01 = dpjm2jp * zw3_jp;
        double q02 = dpjm2jp * zw3_j;
        if (dabs(q001) > dabs(q02)) q001 = q02;
        if (dabs(q002) > dabs(q01)) q002 = q01;
        double q = (q001 - q002) * (1.0 / (dp[j - 1] + dp[j] + dp[j + 1]));
        q001 = q001 - q * dp[j + 1];
        q002 = q002 + q * dp[j - 1];
        c2[j] = s[j] + q001;
        c1[j] = s[j] - q002;
        double a = 2.0 * q001 - q002, b = 2.0 * q002 - q001;
        c3[j] = a * a;
        zw2[j] = b * b;
      }
    }
    c1[kk] = s[kk];
    c2[kk] = s[kk];
    c3[kk] = 0.0;
    zw2[kk] = 0.0;
    // interface values zw(j,3) (second sweep, 2940-2944) rolled into the
    // final limiter sweep (2948-2965): layer j reads zw(j,3), zw(j+1,3),
    // both formed from the first-sweep values of c1, c2.
    double zj;
    {
      double q002 = dmax(zw2[1], dsmll);
      double q001 = dmax(c3[2], dsmll);
      zj = (q001 * c2[1] + q002 * c1[2]) / (q001 + q002);
    }
    for (int j = 2; j <= kk - 1; j++) {
      double q002n = dmax(zw2[j], dsmll);
      double q001n = dmax(c3[j + 1], dsmll);
      double zjp = (q001n * c2[j] + q002n * c1[j + 1]) / (q001n + q002n);
      if (!(lbit(lcm, j) || dp[j] <= thin)) {
        double q01 = zjp - s[j];
        double q02 = s[j] - zj;
        double q001 = 2.0 * q01;
        double q002 = 2.0 * q02;
        if (q01 * q02 < 0.0) { q01 = 0.0; q02 = 0.0; }
        else if (dabs(q01) > dabs(q002)) { q01 = q002; }
        else if (dabs(q02) > dabs(q001)) { q02 = q001; }
        c1[j] = s[j] - q02;
        c2[j] = s[j] + q01;
      }
      zj = zjp;
    }
  }
}

// The remap walk shared by the four methods.  OUT is a callable out(k, v).
template <int METHOD, class OUT>
HYB_HD void walk(int kk, const double* si, const double* pi, const double* dpi,
                 const double* c1, const double* c2, const double* c3,
                 const double* po, double thin, OUT out) {
  double si_min = si[1], si_max = si[1];
  for (int k = 2; k <= kk; k++) { si_min = dmin(si_min, si[k]); si_max = dmax(si_max, si[k]); }
  const double zx = pi[kk + 1];
  double zb = dmax(po[1], pi[1]);
  int lb = 1;
  while (pi[lb + 1] < zb && lb < kk) lb++;
  double prev = 0.0;
  for (int k = 1; k <= kk; k++) {
    double zt = zb;
    zb = dmin(po[k + 1], zx);
    int lt = lb;
    while (pi[lb + 1] < zb && lb < kk) lb++;
    double v;
    if (zb - zt <= thin || zt >= zx) {
      v = (k != 1) ? prev : si[k];
    } else if (METHOD == 0) {  // PCM, 2367-2387
      if (lt != lb) {
        double dpt = pi[lt + 1] - zt;
        double dpb = zb - pi[lb];
        double o = si[(lt + lb) / 2];
        double sz = dpt * (si[lt] - o);
        for (int l = lt + 1; l <= lb - 1; l++) sz = sz + dpi[l] * (si[l] - o);
        sz = sz + dpb * (si[lb] - o);
        v = o + sz / (zb - zt);
        v = dmax(si_min, v);
        v = dmin(si_max, v);
      } else {
        v = si[lt];
      }
    } else {
      double xt = (zt - pi[lt]) / dmax(dpi[lt], thin);
      double xb = (zb - pi[lb]) / dmax(dpi[lb], thin);
      if (METHOD == 1) {  // PLM, 2553-2581
        if (lt != lb) {
          double qt0 = (1.0 - xt);
          double qt1 = (1.0 - xt * xt) * 0.5;
          double qb0 = xb;
          double qb1 = xb * xb * 0.5;
          double o = si[(lt + lb) / 2];
          double c0 = si[lt] - o - 0.5 * c1[lt];
          double sz = dpi[lt] * (c0 * qt0 + c1[lt] * qt1);
          for (int l = lt + 1; l <= lb - 1; l++) sz = sz + dpi[l] * (si[l] - o);
          c0 = si[lb] - o - 0.5 * c1[lb];
          sz = sz + dpi[lb] * (c0 * qb0 + c1[lb] * qb1);
          v = o + sz / (zb - zt);
        } else {
          double qt1 = (xb * xb - xt * xt - (xb - xt)) * 0.5;
          double sz = dpi[lt] * (c1[lt] * qt1);
          v = si[lt] + sz / (zb - zt);
        }
      } else if (METHOD == 2) {  // PPM, 2799-2839
        if (lt != lb) {
          double qt0 = (1.0 - xt);
          double qt1 = (1.0 - xt * xt) * 0.5;
          double qt2 = (1.0 - xt * xt * xt) / 3.0;
          double qb0 = xb;
          double qb1 = xb * xb * 0.5;
          double qb2 = xb * xb * xb / 3.0;
          double o = si[(lt + lb) / 2];
          double sz = dpi[lt] * ((c1[lt] - o) * qt0 + (c2[lt] + c3[lt]) * qt1 - c3[lt] * qt2);
          for (int l = lt + 1; l <= lb - 1; l++) sz = sz + dpi[l] * (si[l] - o);
          sz = sz + dpi[lb] * ((c1[lb] - o) * qb0 + (c2[lb] + c3[lb]) * qb1 - c3[lb] * qb2);
          v = o + sz / (zb - zt);
        } else {
          double qt0 = (xb - xt);
          double qt1 = (xb * xb - xt * xt) * 0.5;
          double qt2 = (xb * xb * xb - xt * xt * xt) / 3.0;
          double o = si[lt];
          double sz = dpi[lt] * ((c1[lt] - o) * qt0 + (c2[lt] + c3[lt]) * qt1 - c3[lt] * qt2);
          v = o + sz / (zb - zt);
        }
      } else {  // WENO, 3066-3103
        if (lt != lb) {
          double dpt = pi[lt + 1] - zt;
          double dpb = zb - pi[lb];
          double qt1 = xt * (xt - 1.0);
          double qt2 = qt1 + xt;
          double qt0 = 1.0 - qt1 - qt2;
          double qb1 = (xb - 1.0) * (xb - 1.0);
          double qb2 = qb1 - 1.0 + xb;
          double qb0 = 1.0 - qb1 - qb2;
          double o = si[(lt + lb) / 2];
          double sz = dpt * (qt0 * (si[lt] - o) + qt1 * (c1[lt] - o) + qt2 * (c2[lt] - o));
          for (int l = lt + 1; l <= lb - 1; l++) sz = sz + dpi[l] * (si[l] - o);
          sz = sz + dpb * (qb0 * (si[lb] - o) + qb1 * (c1[lb] - o) + qb2 * (c2[lb] - o));
          v = o + sz / (zb - zt);
        } else {
          double qt1 = xb * xb + xt * xt + xb * xt + 1.0 - 2.0 * (xb + xt);
          double qt2 = qt1 - 1.0 + (xb + xt);
          double o = si[lt];
          double sz = qt1 * (c1[lt] - o) + qt2 * (c2[lt] - o);
          v = o + sz;
        }
      }
      v = dmax(si_min, v);
      v = dmin(si_max, v);
    }
    prev = v;
    out(k, v);
  }
}

// ---------------------------------------------------------------------
// Phase A: one p-column, hybgenaj (hybgen.F90:396-1753), hybflg==0.
// KC is the capacity of the per-lane arrays (>= kk).
// ---------------------------------------------------------------------
template <int KC>
HYB_HD void column_a(const KParams& P, const View& V, int i, int j, int n,
                     uint64_t& lcm_row) {
  const int kk = P.kk, nhybrd = P.nhybrd, ntracr = P.ntracr;
  const int64_t c0 = ij_of(V, i, j);
  double* temp = V.temp + c0 + lev4(V, 1, n);
  double* saln = V.saln + c0 + lev4(V, 1, n);
  double* th3d = V.th3d + c0 + lev4(V, 1, n);
  double* dpc = V.dp + c0 + lev4(V, 1, n);
  const double* theta = V.theta + c0;
  const int64_t S = V.NIJ;
  // field(k) of this column: x[(k-1)*S]
#define FK(x, k) (x)[(int64_t)((k) - 1) * S]
#define TRK(t, k) V.tracer[c0 + levt(V, k, n, t)]

  const double qqmn = -4.0, qqmx = 2.0;
  const double cusha = qqmn * qqmn * (qqmx - 1.0) / ((qqmx - qqmn) * (qqmx - qqmn));
  const double cushb = 1.0 / qqmn;
  auto cushn = [&](double delp, double dp0) -> double {
    double qq = dmax(qqmn, dmin(qqmx, delp / dp0));
    double a = 1.0 - cushb * qq;
    return dp0 * (1.0 + cusha * (a * a)) * dmax(1.0, delp / (dp0 * qqmx));
  };

  double p[KC + 2], dp0ij[KC + 1], dp0cum[KC + 2], qhrlx[KC + 2];
  double pres[KC + 2], dprs[KC + 1], dpi[KC + 1], prsf[KC + 2];
  const double dpthin = kDpthin;
  const double thkbop = P.thkbop;
  const double hybrlx = P.hybrlx;
  const double rthin = P.hybthn;
  double q, qdep, qts, p_hat, p_hat0 = 0.0, p_hat2, p_hat3;
  int k, ka, kp, fixlay, fixall;

  uint64_t lcm = lcm_row;

  // 399-477
  if (P.dpns == P.dsns) {
    qdep = 1.0;
  } else {
    qdep = dmax(0.0, dmin(1.0, (V.depths[c0] - P.dsns) / (P.dpns - P.dsns)));
  }
  if (qdep < 1.0) {
    p[1] = 0.0;
    dp0cum[1] = 0.0;
    qhrlx[1] = 1.0;
    dp0ij[1] = qdep * P.dp0k[0] + (1.0 - qdep) * P.ds0k[0];
    dp0cum[2] = dp0cum[1] + dp0ij[1];
    qhrlx[2] = 1.0;
    p[2] = p[1] + FK(dpc, 1);
    for (k = 2; k <= kk; k++) {
      qhrlx[k + 1] = 1.0;
      dp0ij[k] = qdep * P.dp0k[k - 1] + (1.0 - qdep) * P.ds0k[k - 1];
      dp0cum[k + 1] = dp0cum[k] + dp0ij[k];
      p[k + 1] = p[k] + FK(dpc, k);
    }
  } else {
    p[1] = 0.0;
    dp0cum[1] = 0.0;
    qhrlx[1] = 1.0;
    dp0ij[1] = P.dp0k[0];
    dp0cum[2] = dp0cum[1] + dp0ij[1];
    qhrlx[2] = 1.0;
    p[2] = p[1] + FK(dpc, 1);
    for (k = 2; k <= kk; k++) {
      double d = P.dp0k[k - 1];
      if (d <= P.dp00i) {
        q = d;
        qts = 0.0;
      } else {
        q = dmax(P.dp00i, d * d / dmax(d, p[k] - dp0cum[k]));
        qts = 1.0 - (q - P.dp00i) / (d - P.dp00i);
      }
      qhrlx[k + 1] = 1.0 / (1.0 + qts * (hybrlx - 1.0));
      dp0ij[k] = dmin(q, d);
      dp0cum[k + 1] = dp0cum[k] + d;
      p[k + 1] = p[k] + FK(dpc, k);
    }
  }

  // 479-488
  fixlay = 1;
  for (k = 2; k <= nhybrd; k++) {
    if (dp0cum[k] >= V.topiso[c0]) break;
    qhrlx[k + 1] = 1.0;
    fixlay = fixlay + 1;
  }

  if (P.ldenfix) {  // 496-549
    int lf = 1;
    fixall = fixlay;
    for (k = fixall + 1; k <= nhybrd; k++) {
      double denfix = FK(th3d, nhybrd);
      for (int l = lf; l <= nhybrd; l++) {
        double zk = dp0cum[k + 1];
        if (p[l] <= zk && p[l + 1] >= zk) {
          double z0 = 0.5 * (p[l] + p[l + 1]);
          if (zk <= z0) {
            if (l == 1) {
              denfix = FK(th3d, 1);
            } else {
              double zm = 0.5 * (p[l - 1] + p[l]);
              q = (z0 - zk) / (z0 - zm);
              denfix = q * FK(th3d, l - 1) + (1.0 - q) * FK(th3d, l);
            }
          } else {
            if (p[l + 1] == p[nhybrd + 1]) {
              denfix = FK(th3d, nhybrd);
            } else {
              double zp = 0.5 * (p[l + 1] + p[l + 2]);
              q = (zk - z0) / (zp - z0);
              denfix = q * FK(th3d, l + 1) + (1.0 - q) * FK(th3d, l);
            }
          }
          lf = l;
          break;
        }
      }
      if (p[k + 1] > dp0cum[k + 1] + 0.1 * dp0ij[k] && FK(theta, k) > denfix) {
        if (fixlay > fixall) {
          if (p[k] > dp0cum[k]) fixlay = fixlay - 1;
        }
        break;
      }
      qhrlx[k] = 1.0;
      fixlay = fixlay + 1;
    }
  } else {  // 550-573
    fixall = fixlay;
    for (k = fixlay + 1; k <= nhybrd; k++) {
      if (p[k + 1] > dp0cum[k + 1] + 0.1 * dp0ij[k]) {
        if (fixlay > fixall) {
          if (p[k] > dp0cum[k]) fixlay = fixlay - 1;
        }
        break;
      }
      qhrlx[k] = 1.0;
      fixlay = fixlay + 1;
    }
  }

  // 590-598
  kp = 2;
  for (k = kk; k >= 3; k--) {
    if (p[k + 1] - p[k] >= dpthin) { kp = k; break; }
  }

  k = kp;
  ka = (k - 2 > 1) ? k - 2 : 1;
  if (k > fixlay + 1 && qdep == 1.0 && p[k] - p[k - 1] >= dpthin &&
      FK(theta, k) - kEpsil > FK(th3d, k) && FK(th3d, k - 1) > FK(th3d, k) &&
      FK(th3d, ka) > FK(th3d, k)) {
    // 609-686
    q = (p[k + 1] - p[k]) / (p[k + 1] - p[k - 1]);
    FK(temp, k - 1) = FK(temp, k - 1) - q * (FK(temp, k - 1) - FK(temp, k));
    FK(saln, k - 1) = FK(saln, k - 1) - q * (FK(saln, k - 1) - FK(saln, k));
    FK(th3d, k - 1) = sig(FK(temp, k - 1), FK(saln, k - 1)) - P.thbase;
    for (int t = 1; t <= ntracr; t++)
      TRK(t, k - 1) = TRK(t, k - 1) - q * (TRK(t, k - 1) - TRK(t, k));
    p[k] = p[k + 1];
    kp = k - 1;
  } else if (k > fixlay + 1 && qdep == 1.0 && p[k] - p[k - 1] >= dpthin &&
             FK(theta, k) - kEpsil > FK(th3d, k) && FK(th3d, k - 1) > FK(th3d, k)) {
    // 687-811
    if (p[k + 1] - p[k] <= p[k] - p[k - 1]) {
      double st = FK(temp, k - 1), ss = FK(saln, k - 1), sh = FK(th3d, k - 1);
      q = (p[k + 1] - p[k]) / (p[k] - p[k - 1]);
      FK(temp, k - 1) = FK(temp, k - 1) - q * (FK(temp, k - 1) - FK(temp, k));
      FK(saln, k - 1) = FK(saln, k - 1) - q * (FK(saln, k - 1) - FK(saln, k));
      FK(th3d, k - 1) = sig(FK(temp, k - 1), FK(saln, k - 1)) - P.thbase;
      FK(temp, k) = st;
      FK(saln, k) = ss;
      FK(th3d, k) = sh;
      for (int t = 1; t <= ntracr; t++) {
        double sv = TRK(t, k - 1);
        TRK(t, k - 1) = TRK(t, k - 1) - q * (TRK(t, k - 1) - TRK(t, k));
        TRK(t, k) = sv;
      }
    } else {
      double st = FK(temp, k), ss = FK(saln, k), sh = FK(th3d, k);
      q = (p[k] - p[k - 1]) / (p[k + 1] - p[k]);
      FK(temp, k) = FK(temp, k) + q * (FK(temp, k - 1) - FK(temp, k));
      FK(saln, k) = FK(saln, k) + q * (FK(saln, k - 1) - FK(saln, k));
      FK(th3d, k) = sig(FK(temp, k), FK(saln, k)) - P.thbase;
      FK(temp, k - 1) = st;
      FK(saln, k - 1) = ss;
      FK(th3d, k - 1) = sh;
      for (int t = 1; t <= ntracr; t++) {
        double sv = TRK(t, k);
        TRK(t, k) = TRK(t, k) + q * (TRK(t, k - 1) - TRK(t, k));
        TRK(t, k - 1) = sv;
      }
    }
  }

  // 813-970
  k = kp;
  {
    double dpthck = P.dx0k[k - 2] * 0.95;
    if (k > fixlay + 1 && qdep == 1.0 && p[k] - p[k - 1] >= dpthin &&
        p[k] - p[k - 1] < dpthck && FK(theta, k) - kEpsil > FK(th3d, k) &&
        FK(theta, k - 1) < FK(th3d, k) &&
        dabs(FK(theta, k - 1) - FK(th3d, k - 1)) < P.hybiso &&
        (FK(th3d, k) - FK(th3d, k - 1)) > (FK(theta, k) - FK(theta, k - 1)) * 0.001) {
      ka = 1;
      for (int kt = k - 2; kt >= 2; kt--) {
        if (FK(th3d, k - 1) - FK(th3d, kt) >= FK(theta, k - 1) - FK(theta, k - 2)) {
          ka = kt;
          break;
        }
      }
      double delsm = dabs(FK(saln, ka) - FK(saln, k - 1));
      double dels = dabs(FK(saln, k - 1) - FK(saln, k));
      double deltm = dabs(FK(temp, ka) - FK(temp, k - 1));
      double delt = dabs(FK(temp, k - 1) - FK(temp, k));
      q = dmin3(FK(temp, ka), FK(temp, k - 1), FK(temp, k));
      if (q > 6.0) {
        deltm = dmin(deltm, 6.0 * (FK(theta, k) - FK(theta, k - 1)));
      } else {
        deltm = dmin(deltm, 10.0 * (FK(theta, k) - FK(theta, k - 1)));
      }
      delsm = dmin(delsm, 1.3 * (FK(theta, k) - FK(theta, k - 1)));
      qts = 0.0;
      if (delt > kEpsil) qts = dmax(qts, (dmin(deltm, 2.0 * delt) - delt) / delt);
      if (dels > kEpsil) qts = dmax(qts, (dmin(delsm, 2.0 * dels) - dels) / dels);
      q = (FK(theta, k) - FK(th3d, k)) / (FK(theta, k) - FK(th3d, k - 1));
      q = dmin(q, qts / (1.0 + qts));
      q = qhrlx[k] * q;
      p_hat = q * (p[k + 1] - p[k]);
      p[k] = p[k] + p_hat;
      FK(temp, k) = FK(temp, k) + (q / (1.0 - q)) * (FK(temp, k) - FK(temp, k - 1));
      FK(saln, k) = FK(saln, k) + (q / (1.0 - q)) * (FK(saln, k) - FK(saln, k - 1));
      FK(th3d, k) = sig(FK(temp, k), FK(saln, k)) - P.thbase;
      if (ntracr > 0 && p_hat != 0.0) {
        double qtr = p_hat / dmax(p_hat, p[k] - p[k - 1]);
        for (int t = 1; t <= ntracr; t++) {
          if ((P.trc_t >> (t - 1)) & 1ull) {
            TRK(t, k) = TRK(t, k) + (q / (1.0 - q)) * (TRK(t, k) - TRK(t, k - 1));
          } else {
            TRK(t, k - 1) = TRK(t, k - 1) + qtr * (TRK(t, k) - TRK(t, k - 1));
          }
        }
      }
    }
  }

  // 972-1009
  for (k = kp + 1; k <= kk; k++) {
    if (k <= nhybrd) {
      FK(th3d, k) = FK(th3d, k - 1);
      FK(saln, k) = FK(saln, k - 1);
      FK(temp, k) = FK(temp, k - 1);
    } else if (FK(th3d, k) != FK(theta, k)) {
      FK(th3d, k) = dmax(FK(theta, k), FK(th3d, k - 1));
      FK(saln, k) = FK(saln, k - 1);
      FK(temp, k) = 99.0;  // tofsig under EOS_17T, stmt_fns.h:532-533
      FK(saln, k) = 99.0;  // sofsig under EOS_17T, stmt_fns.h:535-536
    }
    for (int t = 1; t <= ntracr; t++) TRK(t, k) = TRK(t, k - 1);
  }

  // 1013-1046
  pres[1] = p[1];
  for (k = 1; k <= kk; k++) {
    pres[k + 1] = p[k + 1];
    dprs[k] = pres[k + 1] - pres[k];
    dpi[k] = dmax(dprs[k], dpthin);
    if (P.isopcm) {
      bool v;
      if (k <= fixlay) v = false;
      else v = k > nhybrd || dprs[k] <= dpthin || dabs(FK(th3d, k) - FK(theta, k)) < P.hybiso;
      lcm = setbit(lcm, k, v);
    }
  }

  // 1051-1063
  if (fixlay >= 1) {
    p_hat = p[1] + dp0ij[1];
    p[2] = p_hat;
    for (k = 2; k <= kk; k++) {
      if (p[k + 1] >= p_hat) break;
      p[k + 1] = p_hat;
    }
  }

  // 1065-1494
  for (k = 2; k <= nhybrd; k++) {
    if (k <= fixlay) {
      if (k < kk) {
        p[k + 1] = dmin(dp0cum[k + 1], p[kk + 1]);
        if (k == fixlay) {
          for (ka = k + 2; ka <= kk; ka++) {
            if (p[ka] >= p[k + 1]) break;
            p[ka] = p[k + 1];
          }
        }
      }
    } else {
      double dp_top = p[k] - p[k - 1] + kEpsil;
      double dp_cen = p[k + 1] - p[k] + kEpsil;
      double dp_bot, dp_far, dp_inc;
      if (k == kk) {
        dp_bot = dp_cen;
        dp_far = dp_cen;
      } else if (k == kk - 1) {
        dp_bot = p[k + 2] - p[k + 1] + kEpsil;
        dp_far = dp_bot;
      } else {
        dp_bot = p[k + 2] - p[k + 1] + kEpsil;
        dp_far = p[k + 3] - p[k + 2] + kEpsil;
      }
      if (dp_cen < P.hybthk * dmin(dp_top, dp_bot)) {
        if (dmin(dp_top, dp_bot) < 0.2 * dmax(dp_top, dp_bot)) {
          q = dp_top / (dp_top + dp_bot);
        } else if (FK(theta, k) > FK(th3d, k - 1) && FK(theta, k) < FK(th3d, k + 1)) {
          q = (FK(theta, k) - FK(th3d, k + 1)) / (FK(th3d, k - 1) - FK(th3d, k + 1));
        } else {
          q = (FK(theta, k) - FK(theta, k + 1)) / (FK(theta, k - 1) - FK(theta, k + 1));
        }
        if (dp_top <= dp_bot) {
          dp_inc = (P.hybthk * dp_top - dp_cen) / (1.0 + q * P.hybthk);
        } else {
          dp_inc = (P.hybthk * dp_bot - dp_cen) / (1.0 + (1.0 - q) * P.hybthk);
        }
        p_hat = p[k] - q * dp_inc;
        p[k] = (1.0 - qhrlx[k]) * p[k] + qhrlx[k] * p_hat;
        p_hat = p[k + 1] + (1.0 - q) * dp_inc;
        p[k + 1] = (1.0 - qhrlx[k]) * p[k + 1] + qhrlx[k] * p_hat;
      } else if (dp_cen + dp_bot < P.hybthk * dmin(dp_top, dp_far)) {
        double thm = 0.5 * (FK(theta, k) + FK(theta, k + 1));
        if (dmin(dp_top, dp_far) < 0.2 * dmax(dp_top, dp_far)) {
          q = dp_top / (dp_top + dp_far);
        } else if (thm > FK(th3d, k - 1) && thm < FK(th3d, k + 2)) {
          q = (thm - FK(th3d, k + 2)) / (FK(th3d, k - 1) - FK(th3d, k + 2));
        } else {
          q = (thm - FK(theta, k + 2)) / (FK(theta, k - 1) - FK(theta, k + 2));
        }
        if (dp_top <= dp_far) {
          dp_inc = (P.hybthk * dp_top - (dp_cen + dp_bot)) / (1.0 + q * P.hybthk);
        } else {
          dp_inc = (P.hybthk * dp_far - (dp_cen + dp_bot)) / (1.0 + (1.0 - q) * P.hybthk);
        }
        p_hat = p[k] - q * dp_inc;
        p[k] = (1.0 - qhrlx[k]) * p[k] + qhrlx[k] * p_hat;
        p_hat = p[k + 2] + (1.0 - q) * dp_inc;
        p[k + 2] = (1.0 - qhrlx[k]) * p[k + 2] + qhrlx[k] * p_hat;
      } else if (FK(th3d, k) > FK(theta, k) + kEpsil && k > fixlay + 1) {
        // too dense, 1237-1387
        bool ltop;
        if (rthin * (p[k] - p[k - 1]) <= (p[k + 1] - p[k])) {
          ltop = true;
        } else if (rthin * (p[k + 1] - p[k]) <= (p[k] - p[k - 1])) {
          ltop = false;
        } else {
          ltop = FK(theta, k - 1) - FK(th3d, k - 1) >= FK(th3d, k) - FK(theta, k);
        }
        if (FK(th3d, k - 1) >= FK(theta, k - 1) || !ltop || p[k] <= dp0cum[k] + kOnem) {
          if ((FK(theta, k) - FK(th3d, k - 1)) <= kEpsil) {
            // p_hat0 keeps the previous level's 1479 value: the blocking
            // branch below reads it (1301-1330)
            p_hat = p[k - 1] + dp0ij[k - 1];
          } else {
            q = (FK(theta, k) - FK(th3d, k)) / (FK(theta, k) - FK(th3d, k - 1));
            p_hat0 = p[k] + q * (p[k + 1] - p[k]);
            if (k == fixlay + 2) {
              p_hat = p[k - 1] + dmax(p_hat0 - p[k - 1], dp0ij[k - 1]);
            } else {
              p_hat = p[k - 1] + cushn(p_hat0 - p[k - 1], dp0ij[k - 1]);
            }
          }
          p_hat = dmin(p_hat, p[kk + 1]);
          if (k <= fixlay + 2) {
            // do nothing
          } else if (p_hat >= p[k] && p[k - 1] > dp0cum[k - 1] + kTenm &&
                     (p[kk + 1] - p[k - 1] < thkbop ||
                      p[k - 1] - p[k - 2] > qqmx * dp0ij[k - 2])) {
            if (k == fixlay + 3) {
              p_hat2 = p[k - 2] + dmax(p[k - 1] - p_hat + p_hat0 - p[k - 2], dp0ij[k - 2]);
            } else {
              p_hat2 = p[k - 2] + cushn(p[k - 1] - p_hat + p_hat0 - p[k - 2], dp0ij[k - 2]);
            }
            if (p_hat2 < p[k - 1] - kOnemm) {
              p[k - 1] = (1.0 - qhrlx[k - 1]) * p[k - 1] +
                         qhrlx[k - 1] * dmax(p_hat2, 2.0 * p[k - 1] - p_hat);
              p_hat = p[k - 1] + cushn(p_hat0 - p[k - 1], dp0ij[k - 1]);
            } else if (k <= fixlay + 3) {
              // do nothing
            } else if (p[k - 2] > dp0cum[k - 2] + kTenm &&
                       (p[kk + 1] - p[k - 2] < thkbop ||
                        p[k - 2] - p[k - 3] > qqmx * dp0ij[k - 3])) {
              if (k == fixlay + 4) {
                p_hat3 = p[k - 3] + dmax(p[k - 2] - p_hat + p_hat0 - p[k - 3], dp0ij[k - 3]);
              } else {
                p_hat3 = p[k - 3] + cushn(p[k - 2] - p_hat + p_hat0 - p[k - 3], dp0ij[k - 3]);
              }
              if (p_hat3 < p[k - 2] - kOnemm) {
                p[k - 2] = (1.0 - qhrlx[k - 2]) * p[k - 2] +
                           qhrlx[k - 2] * dmax(p_hat3, 2.0 * p[k - 2] - p[k - 1]);
                p_hat2 = p[k - 2] + cushn(p[k - 1] - p_hat + p_hat0 - p[k - 2], dp0ij[k - 2]);
                if (p_hat2 < p[k - 1] - kOnemm) {
                  p[k - 1] = (1.0 - qhrlx[k - 1]) * p[k - 1] +
                             qhrlx[k - 1] * dmax(p_hat2, 2.0 * p[k - 1] - p_hat);
                  p_hat = p[k - 1] + cushn(p_hat0 - p[k - 1], dp0ij[k - 1]);
                }
              }
            }
          }
          if (p_hat < p[k]) {
            p[k] = (1.0 - qhrlx[k]) * p[k] + qhrlx[k] * p_hat;
          }
        }
      } else if (FK(th3d, k) < FK(theta, k) - kEpsil) {
        // too light, 1389-1475
        if (p[k + 1] < p[kk + 1]) {
          bool ltop;
          if (rthin * (p[k + 1] - p[k]) <= (p[k + 2] - p[k + 1])) {
            ltop = true;
          } else if (rthin * (p[k + 2] - p[k + 1]) <= (p[k + 1] - p[k])) {
            ltop = false;
          } else {
            ltop = FK(theta, k) - FK(th3d, k) >= FK(th3d, k + 1) - FK(theta, k + 1);
          }
          if (FK(th3d, k + 1) <= FK(theta, k + 1) || ltop || p[k + 1] <= dp0cum[k + 1] + kOnem) {
            if ((FK(th3d, k + 1) - FK(theta, k)) <= kEpsil) {
              p_hat = p[k + 2];
            } else {
              q = (FK(th3d, k) - FK(theta, k)) / (FK(th3d, k + 1) - FK(theta, k));
              p_hat = p[k + 1] + q * (p[k] - p[k + 1]);
            }
            int k3 = (k + 3 < kk + 1) ? k + 3 : kk + 1;
            if (p[k3] < p[kk + 1]) {
              if (p[kk + 1] - p[k] > dp0ij[k] + dp0ij[k + 1]) {
                p_hat = p[k + 2] - cushn(p[k + 2] - p_hat, dp0ij[k + 1]);
              }
              p_hat = p[k] + dmax(p_hat - p[k], dp0ij[k]);
              p_hat = dmin(p_hat, dmax(0.5 * (p[k + 1] + p[k + 2]), p[k + 2] - dp0ij[k + 1]));
            } else {
              p_hat = dmin(p[k + 2], p_hat);
            }
            if (p_hat > p[k + 1]) {
              p[k + 1] = (1.0 - qhrlx[k + 1]) * p[k + 1] + qhrlx[k + 1] * p_hat;
            }
          }
        }
      }
      // 1479-1490
      p_hat0 = dmin(p[k - 1] + dp0ij[k - 1], p[kk + 1]);
      if (p_hat0 > p[k]) {
        p_hat = (1.0 - qhrlx[k - 1]) * p[k] + qhrlx[k - 1] * p_hat0;
        p[k] = dmin(p_hat, p[k + 1]);
      }
    }
  }

  // 1496-1509
  for (k = 2; k <= kk - 1; k++) {
    if ((p[k + 1] - p[k]) > P.dx0k[k - 1]) {
      p[k + 1] = p[k] + P.dx0k[k - 1];
      lcm = setbit(lcm, k, false);
      lcm = setbit(lcm, k - 1, false);
    }
  }

  // 1523-1529: new dp and p
  double* pg = V.p + c0;
  prsf[1] = p[1];
  pg[0] = p[1];
  for (k = 1; k <= kk; k++) {
    double d = dmax(p[k + 1] - prsf[k], 0.0);
    FK(dpc, k) = d;
    prsf[k + 1] = prsf[k] + d;
    pg[(int64_t)k * S] = prsf[k + 1];
  }
  if (V.rec_fixlay) {
    int64_t r = (int64_t)(i - 1) + (int64_t)V.rec_ld * (j - 1);
    V.rec_fixlay[r] = fixlay;
    V.rec_kp[r] = kp;
  }

  // 1530-1591: remap field by field.  T first (written back), then S with
  // th3d = sig(T,S) - thbase, then each tracer.
  double sfld[KC + 2], c1[KC + 2], c2[KC + 2], c3[2 * KC + 6];
  const int hybmap = P.hybmap;
  const double thbase = P.thbase;
  auto run_field = [&](auto load, auto store) {
    for (int kq = 1; kq <= kk; kq++) sfld[kq] = load(kq);
    switch (hybmap) {
      case 0:
        walk<0>(kk, sfld, pres, dprs, c1, c2, c3, prsf, dpthin, store);
        break;
      case 1:
        coefs<1>(kk, sfld, dprs, dpi, lcm, dpthin, c1, c2, c3);
        walk<1>(kk, sfld, pres, dprs, c1, c2, c3, prsf, dpthin, store);
        break;
      case 2:
        coefs<2>(kk, sfld, dprs, dpi, lcm, dpthin, c1, c2, c3);
        walk<2>(kk, sfld, pres, dprs, c1, c2, c3, prsf, dpthin, store);
        break;
      default:
        coefs<3>(kk, sfld, dprs, dpi, lcm, dpthin, c1, c2, c3);
        walk<3>(kk, sfld, pres, dprs, c1, c2, c3, prsf, dpthin, store);
        break;
    }
  };
  run_field([&](int kq) { return FK(temp, kq); },
            [&](int kq, double v) { FK(temp, kq) = v; });
  run_field([&](int kq) { return FK(saln, kq); },
            [&](int kq, double v) {
              FK(saln, kq) = v;
              FK(th3d, kq) = sig(FK(temp, kq), v) - thbase;
            });
  for (int t = 1; t <= ntracr; t++) {
    run_field([&](int kq) { return TRK(t, kq); },
              [&](int kq, double v) { TRK(t, kq) = v; });
  }
  lcm_row = lcm;
#undef FK
#undef TRK
}

// ---------------------------------------------------------------------
// Robert-Asselin branch: hybgencj (hybgen.F90:1962-2271), hybflg==0.
// Reads the level-n interfaces hybgenaj left in p, writes level-m fields
// and leaves the level-m interfaces in p.
// ---------------------------------------------------------------------
template <int KC>
HYB_HD void column_ra(const KParams& P, const View& V, int i, int j, int m,
                      int n, uint64_t lcm) {
  const int kk = P.kk, nhybrd = P.nhybrd, ntracr = P.ntracr;
  const int64_t c0 = ij_of(V, i, j);
  const int64_t S = V.NIJ;
  double* temp = V.temp + c0 + lev4(V, 1, m);
  double* saln = V.saln + c0 + lev4(V, 1, m);
  double* th3d = V.th3d + c0 + lev4(V, 1, m);
  double* dpm = V.dp + c0 + lev4(V, 1, m);
  const double* dpn = V.dp + c0 + lev4(V, 1, n);
  const double* dpon = V.dpo + c0 + lev4(V, 1, n);
  const double* dpom = V.dpo + c0 + lev4(V, 1, m);
  const double* theta = V.theta + c0;
  double* pg = V.p + c0;
#define FK(x, k) (x)[(int64_t)((k) - 1) * S]
#define TRM(t, k) V.tracer[c0 + levt(V, k, m, t)]
  double dp0ij[KC + 1], dp0cum[KC + 2];
  double pres[KC + 2], dprs[KC + 1], dpi[KC + 1], prsf[KC + 2];
  const double dpthin = kDpthin;
  double q;
  int k, fixlay;
  double qdep = dmax(0.0, dmin(1.0, (V.depths[c0] - P.dsns) / (P.dpns - P.dsns)));
  dp0cum[1] = 0.0;
  dp0ij[1] = qdep * P.dp0k[0] + (1.0 - qdep) * P.ds0k[0];
  dp0cum[2] = dp0cum[1] + dp0ij[1];
  for (k = 2; k <= kk; k++) {
    double d = P.dp0k[k - 1];
    if (d <= P.dp00i) {
      q = d;
    } else {
      q = dmax(P.dp00i, d * d / dmax(d, pg[(int64_t)(k - 1) * S] - dp0cum[k]));
    }
    dp0ij[k] = dmin(q, qdep * d + (1.0 - qdep) * P.ds0k[k - 1]);
    dp0cum[k + 1] = dp0cum[k] + dp0ij[k];
  }
  fixlay = 1;
  for (k = 2; k <= nhybrd; k++) {
    if (dp0cum[k] >= V.topiso[c0]) break;
    fixlay = fixlay + 1;
  }
  pres[1] = 0.0;
  prsf[1] = 0.0;
  for (k = 1; k <= kk; k++) {
    dprs[k] = FK(dpm, k);
    dpi[k] = dmax(dprs[k], dpthin);
    pres[k + 1] = pres[k] + FK(dpm, k);
    double dpold = FK(dpon, k);
    double dpmid = FK(dpom, k);
    double dpnew = FK(dpn, k);
    q = 0.5 * P.ra2fac * (dpold + dpnew - 2.0 * dpmid);
    double d = dpmid + q;
    FK(dpm, k) = d;
    prsf[k + 1] = prsf[k] + d;
    pg[(int64_t)k * S] = prsf[k + 1];
    if (P.isopcm) {
      bool v;
      if (k <= fixlay) v = false;
      else v = k > nhybrd || dprs[k] <= dpthin || dabs(FK(th3d, k) - FK(theta, k)) < P.hybiso;
      lcm = setbit(lcm, k, v);
    }
  }
  double sfld[KC + 2], c1[KC + 2], c2[KC + 2], c3[2 * KC + 6];
  const int hybmap = P.hybmap;
  const double thbase = P.thbase;
  auto run_field = [&](auto load, auto store) {
    for (int kq = 1; kq <= kk; kq++) sfld[kq] = load(kq);
    switch (hybmap) {
      case 0:
        walk<0>(kk, sfld, pres, dprs, c1, c2, c3, prsf, dpthin, store);
        break;
      case 1:
        coefs<1>(kk, sfld, dprs, dpi, lcm, dpthin, c1, c2, c3);
        walk<1>(kk, sfld, pres, dprs, c1, c2, c3, prsf, dpthin, store);
        break;
      case 2:
        coefs<2>(kk, sfld, dprs, dpi, lcm, dpthin, c1, c2, c3);
        walk<2>(kk, sfld, pres, dprs, c1, c2, c3, prsf, dpthin, store);
        break;
      default:
        coefs<3>(kk, sfld, dprs, dpi, lcm, dpthin, c1, c2, c3);
        walk<3>(kk, sfld, pres, dprs, c1, c2, c3, prsf, dpthin, store);
        break;
    }
  };
  run_field([&](int kq) { return FK(temp, kq); },
            [&](int kq, double v) { FK(temp, kq) = v; });
  run_field([&](int kq) { return FK(saln, kq); },
            [&](int kq, double v) {
              FK(saln, kq) = v;
              FK(th3d, kq) = sig(FK(temp, kq), v) - thbase;
            });
  for (int t = 1; t <= ntracr; t++) {
    run_field([&](int kq) { return TRM(t, kq); },
              [&](int kq, double v) { TRM(t, kq) = v; });
  }
#undef FK
#undef TRM
}

// ---------------------------------------------------------------------
// Phase C: one u- or v-point, hybgenbj (hybgen.F90:1791-1904) with the
// old-interface build (171-195) and dpudpv (dpudpv.F90:92-103) fused in:
// pres = running sum of the old dpu, the new dpu from the new p (after the
// halo update), prsf = running sum of the new dpu.
// ---------------------------------------------------------------------
template <int KC>
HYB_HD void point_c(const KParams& P, const View& V, int i, int j, int nl,
                    bool isv) {
  const int kk = P.kk;
  const int64_t c0 = ij_of(V, i, j);
  const int64_t S = V.NIJ;
  const int64_t cn = isv ? ij_of(V, i, j - 1) : ij_of(V, i - 1, j);
  double* dpu = (isv ? V.dpv : V.dpu) + c0 + lev4(V, 1, nl);
  double* uu = (isv ? V.v : V.u) + c0 + lev4(V, 1, nl);
  double* pu = (isv ? V.pv : V.pu) + c0;
  const double depthu = (isv ? V.depthv : V.depthu)[c0];
  const double ub = (isv ? V.vbavg : V.ubavg)[c0 + lev(V, nl)];
  const double* p0 = V.p + c0;
  const double* p1 = V.p + cn;
  const double dpthin = kDpthin;

  double pres[KC + 2], dprs[KC + 1], dpi[KC + 1], prsf[KC + 2];
  double sfld[KC + 2], c1[KC + 2], c2[KC + 2], c3[2 * KC + 6];
  // 171-195: old interfaces
  pres[1] = 0.0;
  for (int k = 1; k <= kk; k++) pres[k + 1] = pres[k] + dpu[(int64_t)(k - 1) * S];
  // dpudpv.F90:95-97 new thickness; hybgen.F90:1806-1810 new interfaces
  prsf[1] = 0.0;
  pu[0] = 0.0;
  for (int k = 1; k <= kk; k++) {
    double a = dmin(depthu, 0.5 * (p0[(int64_t)k * S] + p1[(int64_t)k * S]));
    double b = dmin(depthu, 0.5 * (p0[(int64_t)(k - 1) * S] + p1[(int64_t)(k - 1) * S]));
    double d = dmax(0.0, a - b);
    dpu[(int64_t)(k - 1) * S] = d;
    prsf[k + 1] = prsf[k] + d;
    pu[(int64_t)k * S] = prsf[k + 1];
  }
  for (int k = 1; k <= kk; k++) {
    sfld[k] = uu[(int64_t)(k - 1) * S];
    dprs[k] = pres[k + 1] - pres[k];
    dpi[k] = dmax(dprs[k], dpthin);
  }
  auto store = [&](int k, double v) {
    if (dpi[k] > dpthin || prsf[k] <= prsf[kk + 1] - kOnemm) {
      uu[(int64_t)(k - 1) * S] = v;
    } else {
      uu[(int64_t)(k - 1) * S] = -ub;
    }
  };
  if (P.hybmap == 0) {
    walk<0>(kk, sfld, pres, dprs, c1, c2, c3, prsf, dpthin, store);
  } else if (P.hybmap == 1 && P.hybiso > 2.0) {
    coefs<1>(kk, sfld, dprs, dpi, 0ull, dpthin, c1, c2, c3);
    walk<1>(kk, sfld, pres, dprs, c1, c2, c3, prsf, dpthin, store);
  } else {
    coefs<3>(kk, sfld, dprs, dpi, 0ull, dpthin, c1, c2, c3);
    walk<3>(kk, sfld, pres, dprs, c1, c2, c3, prsf, dpthin, store);
  }
}

// xctilr(p(:,:,2),1,kk,1,1,halo_ps) for single-tile domains,
// mod_xc_sm.h:1372-1424.  t enumerates the halo points of one level:
// [0, ii) j-rows 0 and jj+1 (pairs), then [ii, ii+jj+2) i-columns.
HYB_HD void p_halo_jrow(const KParams& P, const View& V, int i, int k) {
  double* p = V.p + lev(V, k);
  if (P.nreg <= 2) {
    p[ij_of(V, i, 0)] = P.vland;
    p[ij_of(V, i, P.jj + 1)] = P.vland;
  } else {
    p[ij_of(V, i, 0)] = p[ij_of(V, i, P.jj)];
    p[ij_of(V, i, P.jj + 1)] = p[ij_of(V, i, 1)];
  }
}
HYB_HD void p_halo_icol(const KParams& P, const View& V, int j, int k) {
  double* p = V.p + lev(V, k);
  if (P.nreg == 0 || P.nreg == 4) {
    p[ij_of(V, 0, j)] = P.vland;
    p[ij_of(V, P.ii + 1, j)] = P.vland;
  } else {
    p[ij_of(V, 0, j)] = p[ij_of(V, P.ii, j)];
    p[ij_of(V, P.ii + 1, j)] = p[ij_of(V, 1, j)];
  }
}

// Scratch shared by the streaming column kernels.  pres level k of column
// c0 is pres[c0 + (k-1)*NIJ]; the record words are indexed by c0.
struct Scratch {
  double* pres;       // (NIJ, K+1) old interfaces after overturn/unmix
  int32_t* kp;        // deepest inflated layer after overturn
  uint64_t* lcm;      // PCM-layer mask after the dx0k pass
  int32_t* uvf;       // (2, NIJ) u/v points left to the copy-first kernel
  double* ras;        // (NIJ, K, 3) outputs of the RA walks that cannot go in place
                      // (hybgen_raflag only; NULL: such columns go to column_ra)
};

// Row of per-lane storage in LDS (device) or a local array (host).
struct LaneRow {
  double* base;
  int stride;
  HYB_HD double& operator()(int k) const { return base[(int64_t)(k - 1) * stride]; }
};

}  // namespace hyb
