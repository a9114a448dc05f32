// hyb_fast.h -- register-resident FP64 hybgen kernels for a fixed layer count K.
//
// Same arithmetic as hyb_column.h (every expression keeps the reference's
// operation order), re-organised for gfx950:
//  * K is a template constant and every k-loop is unrolled, so the column's
//    interfaces live in VGPRs with compile-time indices;
//  * the column work is split in two kernels: regrid_fast (hybgenaj up to
//    the rebuild of dp/p, hybgen.F90:399-1529) and remap_fast (the
//    remap and write-back, 1530-1591); they communicate through the
//    outputs p and dp, the old interfaces `pres` and a two-word record;
//  * the remap walks the *input* layers in a static loop (the reference walks
//    the output layers): input values and reconstruction coefficients are
//    then compile-time indexed and streamed, and only the new interfaces
//    (an LDS row per lane) and the output stores carry a per-lane index.
//    Output k is still formed from exactly the same terms, in the same
//    order, as in the reference walk.
#pragma once

#include "hyb_column.h"

namespace hyb {

// Scratch shared by the two column kernels.  pres level k of column c0 is
// pres[c0 + (k-1)*NIJ]; the record words are indexed by c0.
struct Scratch {
  double* pres;       // (NIJ, K+1) old interfaces after overturn/unmix
  int32_t* kp;        // deepest inflated layer after overturn
  uint64_t* lcm;      // PCM-layer mask after the dx0k pass
  int32_t* uvf;       // (2, NIJ) u/v points left to the copy-first kernel
};

// Row of per-lane storage in LDS (device) or a local array (host).
struct LaneRow {
  double* base;
  int stride;
  HYB_HD double& operator()(int k) const { return base[(int64_t)(k - 1) * stride]; }
};

template <int N>
HYB_HD double sel(const double (&a)[N], int idx) {
  double r = a[0];
#pragma unroll
  for (int m = 1; m < N; m++)
    if (m == idx) r = a[m];
  return r;
}
template <int N>
HYB_HD void put(double (&a)[N], int idx, double v) {
#pragma unroll
  for (int m = 0; m < N; m++)
    if (m == idx) a[m] = v;
}

// ---------------------------------------------------------------------
// Kernel R body: hybgen.F90:399-1529 for one p-column.  Preconditions
// (checked on the host): hybflg==0, !ldenfix, nhybrd==kk==K.
// ---------------------------------------------------------------------
template <int K>
HYB_HD void regrid_fast(const KParams& P, const View& V, const Scratch& W, int i, int j,
                        int n) {
  const int64_t c0 = ij_of(V, i, j);
  const int64_t S = V.NIJ;
  const int ntracr = P.ntracr;
  double* temp = V.temp + c0 + lev4(V, 1, n);
  double* saln = V.saln + c0 + lev4(V, 1, n);
  double* th3d = V.th3d + c0 + lev4(V, 1, n);
  double* dpc = V.dp + c0 + lev4(V, 1, n);
  const double* theta = V.theta + c0;
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
  const double dpthin = kDpthin;
  const double hybrlx = P.hybrlx;
  const double rthin = P.hybthn;

  double p[K + 2], dp0ij[K + 1], dp0cum[K + 2], qhrlx[K + 2];
  double q, qdep, qts, p_hat, p_hat0, p_hat2, p_hat3;

  // 399-477
  if (P.dpns == P.dsns) {
    qdep = 1.0;
  } else {
    qdep = dmax(0.0, dmin(1.0, (V.depths[c0] - P.dsns) / (P.dpns - P.dsns)));
  }
  p[1] = 0.0;
  dp0cum[1] = 0.0;
  qhrlx[1] = 1.0;
  qhrlx[2] = 1.0;
  if (qdep < 1.0) {
    dp0ij[1] = qdep * P.dp0k[0] + (1.0 - qdep) * P.ds0k[0];
  } else {
    dp0ij[1] = P.dp0k[0];
  }
  dp0cum[2] = dp0cum[1] + dp0ij[1];
  p[2] = p[1] + FK(dpc, 1);
#pragma unroll
  for (int k = 2; k <= K; k++) {
    double d = P.dp0k[k - 1];
    if (qdep < 1.0) {
      qhrlx[k + 1] = 1.0;
      dp0ij[k] = qdep * d + (1.0 - qdep) * P.ds0k[k - 1];
      dp0cum[k + 1] = dp0cum[k] + dp0ij[k];
    } else {
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
    }
    p[k + 1] = p[k] + FK(dpc, k);
  }

  // 479-488 always-fixed layers
  int fixlay = 1;
  bool stop = false;
#pragma unroll
  for (int k = 2; k <= K; k++) {
    if (!stop) {
      if (dp0cum[k] >= V.topiso[c0]) {
        stop = true;
      } else {
        qhrlx[k + 1] = 1.0;
        fixlay = fixlay + 1;
      }
    }
  }
  // 550-573 sometimes-fixed layers
  {
    const int fixall = fixlay;
    stop = false;
#pragma unroll
    for (int k = 2; k <= K; k++) {
      if (!stop && k >= fixall + 1) {
        if (p[k + 1] > dp0cum[k + 1] + 0.1 * dp0ij[k]) {
          if (fixlay > fixall) {
            if (p[k] > dp0cum[k]) fixlay = fixlay - 1;
          }
          stop = true;
        } else {
          qhrlx[k] = 1.0;
          fixlay = fixlay + 1;
        }
      }
    }
  }

  // 590-598 deepest inflated layer
  int kp = 2;
  stop = false;
#pragma unroll
  for (int k = K; k >= 3; k--) {
    if (!stop && p[k + 1] - p[k] >= dpthin) {
      kp = k;
      stop = true;
    }
  }

  // 606-811 overturn of a too-light kp, on the fields in place
  {
    const int k = kp;
    const int ka = (k - 2 > 1) ? k - 2 : 1;
    const double pk = sel(p, k), pkm = sel(p, k - 1), pkp = sel(p, k + 1);
    const double thk = FK(theta, k);
    double tk = FK(th3d, k), tkm = FK(th3d, k - 1), tka = FK(th3d, ka);
    if (k > fixlay + 1 && qdep == 1.0 && pk - pkm >= dpthin && thk - kEpsil > tk && tkm > tk &&
        tka > tk) {
      q = (pkp - pk) / (pkp - pkm);
      FK(temp, k - 1) = FK(temp, k - 1) - q * (FK(temp, k - 1) - FK(temp, k));
      FK(saln, k - 1) = FK(saln, k - 1) - q * (FK(saln, k - 1) - FK(saln, k));
      FK(th3d, k - 1) = sig(FK(temp, k - 1), FK(saln, k - 1)) - P.thbase;
      for (int t = 1; t <= ntracr; t++)
        TRK(t, k - 1) = TRK(t, k - 1) - q * (TRK(t, k - 1) - TRK(t, k));
      put(p, k, pkp);
      kp = k - 1;
    } else if (k > fixlay + 1 && qdep == 1.0 && pk - pkm >= dpthin && thk - kEpsil > tk &&
               tkm > tk) {
      if (pkp - pk <= pk - pkm) {
        double st = FK(temp, k - 1), ss = FK(saln, k - 1), sh = FK(th3d, k - 1);
        q = (pkp - pk) / (pk - pkm);
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
        q = (pk - pkm) / (pkp - pk);
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
  }

  // 813-970 unmix a too-light deepest layer
  {
    const int k = kp;
    const double pk = sel(p, k), pkm = sel(p, k - 1), pkp = sel(p, k + 1);
    const double dpthck = P.dx0k[k - 2] * 0.95;
    const double thk = FK(theta, k), thkm = FK(theta, k - 1);
    const double tk = FK(th3d, k), tkm = FK(th3d, k - 1);
    if (k > fixlay + 1 && qdep == 1.0 && pk - pkm >= dpthin && pk - pkm < dpthck &&
        thk - kEpsil > tk && thkm < tk && dabs(thkm - tkm) < P.hybiso &&
        (tk - tkm) > (thk - thkm) * 0.001) {
      int ka = 1;
      for (int kt = k - 2; kt >= 2; kt--) {
        if (tkm - FK(th3d, kt) >= thkm - FK(theta, k - 2)) {
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
        deltm = dmin(deltm, 6.0 * (thk - thkm));
      } else {
        deltm = dmin(deltm, 10.0 * (thk - thkm));
      }
      delsm = dmin(delsm, 1.3 * (thk - thkm));
      qts = 0.0;
      if (delt > kEpsil) qts = dmax(qts, (dmin(deltm, 2.0 * delt) - delt) / delt);
      if (dels > kEpsil) qts = dmax(qts, (dmin(delsm, 2.0 * dels) - dels) / dels);
      q = (thk - tk) / (thk - tkm);
      q = dmin(q, qts / (1.0 + qts));
      q = sel(qhrlx, k) * q;
      p_hat = q * (pkp - pk);
      const double pknew = pk + p_hat;
      put(p, k, pknew);
      FK(temp, k) = FK(temp, k) + (q / (1.0 - q)) * (FK(temp, k) - FK(temp, k - 1));
      FK(saln, k) = FK(saln, k) + (q / (1.0 - q)) * (FK(saln, k) - FK(saln, k - 1));
      FK(th3d, k) = sig(FK(temp, k), FK(saln, k)) - P.thbase;
      if (ntracr > 0 && p_hat != 0.0) {
        double qtr = p_hat / dmax(p_hat, pknew - pkm);
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

  // 972-1009 massless layers take the values of layer kp (nhybrd==kk);
  // the remap applies this fill on the fly.  Here th3d(k>kp) = th3d(kp).
  const double th_kp = FK(th3d, kp);
  auto th3d_eff = [&](int k) -> double { return k > kp ? th_kp : FK(th3d, k); };

  // 1013-1046 old-grid snapshot and PCM mask
  uint64_t lcm = 0;
  {
    double* pres = W.pres + c0;
#pragma unroll
    for (int k = 1; k <= K + 1; k++) pres[(int64_t)(k - 1) * S] = p[k];
    if (P.isopcm) {
#pragma unroll
      for (int k = 1; k <= K; k++) {
        bool v;
        if (k <= fixlay) v = false;
        else v = (p[k + 1] - p[k]) <= dpthin || dabs(th3d_eff(k) - FK(theta, k)) < P.hybiso;
        if (v) lcm |= 1ull << (k - 1);
      }
    }
  }

  // 1051-1063
  if (fixlay >= 1) {
    p_hat = p[1] + dp0ij[1];
    p[2] = p_hat;
    stop = false;
#pragma unroll
    for (int k = 2; k <= K; k++) {
      if (!stop) {
        if (p[k + 1] >= p_hat) stop = true;
        else p[k + 1] = p_hat;
      }
    }
  }

  // 1065-1494 interface relocation (nhybrd == K)
#pragma unroll
  for (int k = 2; k <= K; k++) {
    if (k <= fixlay) {
      if (k < K) {
        p[k + 1] = dmin(dp0cum[k + 1], p[K + 1]);
        if (k == fixlay) {
          bool st2 = false;
#pragma unroll
          for (int ka = k + 2; ka <= K; ka++) {
            if (!st2) {
              if (p[ka] >= p[k + 1]) st2 = true;
              else p[ka] = p[k + 1];
            }
          }
        }
      }
    } else {
      const double thk = FK(theta, k);
      const double thm = FK(theta, k - 1);
      const double tk = th3d_eff(k);
      const double tkm = th3d_eff(k - 1);
      double dp_top = p[k] - p[k - 1] + kEpsil;
      double dp_cen = p[k + 1] - p[k] + kEpsil;
      double dp_bot, dp_far, dp_inc;
      if (k == K) {
        dp_bot = dp_cen;
        dp_far = dp_cen;
      } else if (k == K - 1) {
        dp_bot = p[k + 2] - p[k + 1] + kEpsil;
        dp_far = dp_bot;
      } else {
        dp_bot = p[k + 2] - p[k + 1] + kEpsil;
        dp_far = p[k + 3] - p[k + 2] + kEpsil;
      }
      if (k < K && dp_cen < P.hybthk * dmin(dp_top, dp_bot)) {
        const double thp = FK(theta, k + 1);
        const double tkp = th3d_eff(k + 1);
        if (dmin(dp_top, dp_bot) < 0.2 * dmax(dp_top, dp_bot)) {
          q = dp_top / (dp_top + dp_bot);
        } else if (thk > tkm && thk < tkp) {
          q = (thk - tkp) / (tkm - tkp);
        } else {
          q = (thk - thp) / (thm - thp);
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
      } else if (k < K - 1 && dp_cen + dp_bot < P.hybthk * dmin(dp_top, dp_far)) {
        const int k2 = (k + 2 <= K) ? k + 2 : K;
        const double thp = FK(theta, k + 1), thp2 = FK(theta, k2);
        const double tkp2 = th3d_eff(k2);
        const double thmid = 0.5 * (thk + thp);
        if (dmin(dp_top, dp_far) < 0.2 * dmax(dp_top, dp_far)) {
          q = dp_top / (dp_top + dp_far);
        } else if (thmid > tkm && thmid < tkp2) {
          q = (thmid - tkp2) / (tkm - tkp2);
        } else {
          q = (thmid - thp2) / (thm - thp2);
        }
        if (dp_top <= dp_far) {
          dp_inc = (P.hybthk * dp_top - (dp_cen + dp_bot)) / (1.0 + q * P.hybthk);
        } else {
          dp_inc = (P.hybthk * dp_far - (dp_cen + dp_bot)) / (1.0 + (1.0 - q) * P.hybthk);
        }
        const int kk2 = (k + 2 <= K + 1) ? k + 2 : K + 1;
        p_hat = p[k] - q * dp_inc;
        p[k] = (1.0 - qhrlx[k]) * p[k] + qhrlx[k] * p_hat;
        p_hat = p[kk2] + (1.0 - q) * dp_inc;
        p[kk2] = (1.0 - qhrlx[k]) * p[kk2] + qhrlx[k] * p_hat;
      } else if (tk > thk + kEpsil && k > fixlay + 1) {
        // too dense, 1237-1387
        bool ltop;
        if (rthin * (p[k] - p[k - 1]) <= (p[k + 1] - p[k])) {
          ltop = true;
        } else if (rthin * (p[k + 1] - p[k]) <= (p[k] - p[k - 1])) {
          ltop = false;
        } else {
          ltop = thm - tkm >= tk - thk;
        }
        if (tkm >= thm || !ltop || p[k] <= dp0cum[k] + kOnem) {
          p_hat0 = 0.0;
          if ((thk - tkm) <= kEpsil) {
            p_hat = p[k - 1] + dp0ij[k - 1];
          } else {
            q = (thk - tk) / (thk - tkm);
            p_hat0 = p[k] + q * (p[k + 1] - p[k]);
            if (k == fixlay + 2) {
              p_hat = p[k - 1] + dmax(p_hat0 - p[k - 1], dp0ij[k - 1]);
            } else {
              p_hat = p[k - 1] + cushn(p_hat0 - p[k - 1], dp0ij[k - 1]);
            }
          }
          p_hat = dmin(p_hat, p[K + 1]);
          const int km2 = (k - 2 >= 1) ? k - 2 : 1;
          const int km3 = (k - 3 >= 1) ? k - 3 : 1;
          if (k <= fixlay + 2) {
            // do nothing
          } else if (p_hat >= p[k] && p[k - 1] > dp0cum[k - 1] + kTenm &&
                     (p[K + 1] - p[k - 1] < P.thkbop ||
                      p[k - 1] - p[km2] > qqmx * dp0ij[km2])) {
            if (k == fixlay + 3) {
              p_hat2 = p[km2] + dmax(p[k - 1] - p_hat + p_hat0 - p[km2], dp0ij[km2]);
            } else {
              p_hat2 = p[km2] + cushn(p[k - 1] - p_hat + p_hat0 - p[km2], dp0ij[km2]);
            }
            if (p_hat2 < p[k - 1] - kOnemm) {
              p[k - 1] = (1.0 - qhrlx[k - 1]) * p[k - 1] +
                         qhrlx[k - 1] * dmax(p_hat2, 2.0 * p[k - 1] - p_hat);
              p_hat = p[k - 1] + cushn(p_hat0 - p[k - 1], dp0ij[k - 1]);
            } else if (k <= fixlay + 3) {
              // do nothing
            } else if (p[km2] > dp0cum[km2] + kTenm &&
                       (p[K + 1] - p[km2] < P.thkbop ||
                        p[km2] - p[km3] > qqmx * dp0ij[km3])) {
              if (k == fixlay + 4) {
                p_hat3 = p[km3] + dmax(p[km2] - p_hat + p_hat0 - p[km3], dp0ij[km3]);
              } else {
                p_hat3 = p[km3] + cushn(p[km2] - p_hat + p_hat0 - p[km3], dp0ij[km3]);
              }
              if (p_hat3 < p[km2] - kOnemm) {
                p[km2] = (1.0 - qhrlx[km2]) * p[km2] +
                         qhrlx[km2] * dmax(p_hat3, 2.0 * p[km2] - p[k - 1]);
                p_hat2 = p[km2] + cushn(p[k - 1] - p_hat + p_hat0 - p[km2], dp0ij[km2]);
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
      } else if (tk < thk - kEpsil) {
        // too light, 1389-1475
        if (k < K && p[k + 1] < p[K + 1]) {
          const int k2 = k + 2;  // <= K+1 since k < K
          const double thp = FK(theta, k + 1);
          const double tkp = th3d_eff(k + 1);
          bool ltop;
          if (rthin * (p[k + 1] - p[k]) <= (p[k2] - p[k + 1])) {
            ltop = true;
          } else if (rthin * (p[k2] - p[k + 1]) <= (p[k + 1] - p[k])) {
            ltop = false;
          } else {
            ltop = thk - tk >= tkp - thp;
          }
          if (tkp <= thp || ltop || p[k + 1] <= dp0cum[k + 1] + kOnem) {
            if ((tkp - thk) <= kEpsil) {
              p_hat = p[k2];
            } else {
              q = (tk - thk) / (tkp - thk);
              p_hat = p[k + 1] + q * (p[k] - p[k + 1]);
            }
            const int k3 = (k + 3 < K + 1) ? k + 3 : K + 1;
            if (p[k3] < p[K + 1]) {
              if (p[K + 1] - p[k] > dp0ij[k] + dp0ij[k + 1]) {
                p_hat = p[k2] - cushn(p[k2] - p_hat, dp0ij[k + 1]);
              }
              p_hat = p[k] + dmax(p_hat - p[k], dp0ij[k]);
              p_hat = dmin(p_hat, dmax(0.5 * (p[k + 1] + p[k2]), p[k2] - dp0ij[k + 1]));
            } else {
              p_hat = dmin(p[k2], p_hat);
            }
            if (p_hat > p[k + 1]) {
              p[k + 1] = (1.0 - qhrlx[k + 1]) * p[k + 1] + qhrlx[k + 1] * p_hat;
            }
          }
        }
      }
      // 1479-1490
      p_hat0 = dmin(p[k - 1] + dp0ij[k - 1], p[K + 1]);
      if (p_hat0 > p[k]) {
        p_hat = (1.0 - qhrlx[k - 1]) * p[k] + qhrlx[k - 1] * p_hat0;
        p[k] = dmin(p_hat, p[k + 1]);
      }
    }
  }

  // 1496-1509 maximum thickness
#pragma unroll
  for (int k = 2; k <= K - 1; k++) {
    if ((p[k + 1] - p[k]) > P.dx0k[k - 1]) {
      p[k + 1] = p[k] + P.dx0k[k - 1];
      lcm &= ~(3ull << (k - 2));  // lcm(k) = lcm(k-1) = .false.
    }
  }

  // 1523-1529 rebuild dp and p
  double* pg = V.p + c0;
  double prsf = p[1];
  pg[0] = p[1];
#pragma unroll
  for (int k = 1; k <= K; k++) {
    double d = dmax(p[k + 1] - prsf, 0.0);
    FK(dpc, k) = d;
    prsf = prsf + d;
    pg[(int64_t)k * S] = prsf;
  }
  W.kp[c0] = kp;
  W.lcm[c0] = lcm;
  if (V.rec_fixlay) {
    int64_t r = (int64_t)(i - 1) + (int64_t)V.rec_ld * (j - 1);
    V.rec_fixlay[r] = fixlay;
    V.rec_kp[r] = kp;
  }
#undef FK
#undef TRK
}

// ---------------------------------------------------------------------
// Remap of one field over the input layers (static l), METHOD as hybmap.
//  s[1..K]   input values (massless fill already applied)
//  pi[1..K+1] old interfaces; dprs/dpi derived per layer
//  po(k)     new interface k (dynamic index, LaneRow)
//  out(k,v)  output k (dynamic index)
// Reproduces hybgen_{pcm,plm,ppm,weno}_remap (hybgen.F90:2276-3107)
// together with the matching _coefs routine.
// ---------------------------------------------------------------------
template <int K>
HYB_HD int first_lb(const double (&pi)[K + 2], double zb) {
  // lb = min(K, 1 + #{l in 1..K-1 : pi(l+1) < zb}), pi non-decreasing
  int lb = 1;
#pragma unroll
  for (int l = 1; l <= K - 1; l++)
    if (pi[l + 1] < zb) lb = l + 1;
  return lb;
}

template <int K, int METHOD, class OUT>
HYB_HD void remap_l(const double (&s)[K + 2], const double (&pi)[K + 2], uint64_t lcm,
                    const LaneRow& po, OUT out) {
  const double thin = kDpthin;
  const double dsmll = 1.0e-8;
  double si_min = s[1], si_max = s[1];
#pragma unroll
  for (int k = 2; k <= K; k++) {
    si_min = dmin(si_min, s[k]);
    si_max = dmax(si_max, s[k]);
  }
  const double zx = pi[K + 1];
  auto dprs = [&](int l) -> double { return pi[l + 1] - pi[l]; };
  auto dpif = [&](int l) -> double { return dmax(pi[l + 1] - pi[l], thin); };

  // walk state
  double zb0 = dmax(po(1), pi[1]);
  int lt_c = first_lb<K>(pi, zb0);  // lt of output kc
  double zt_c = zb0;                 // zt of output kc
  int kc = 1;
  double prev = 0.0;
  bool accon = false, acc_thin = false;
  int lb_act = 0;
  double acc = 0.0, o = 0.0, zt_act = 0.0, zb_act = 0.0;

  // streaming coefficient state
  // PPM: as/al of the next layer and metrics
  double as_cur = 0.0, al_cur = 0.0;     // as(l), raw al(l)
  double qdpjp_prev = 0.0;                // qdpjp(l-1)
  // WENO: first-sweep values of layers l and l+1, interface value zw3(l)
  double wc1[3] = {0.0, 0.0, 0.0}, wc2[3] = {0.0, 0.0, 0.0};  // [0]=l, [1]=l+1
  double wz1[3] = {0.0, 0.0, 0.0}, wz2[3] = {0.0, 0.0, 0.0};
  double zw3_l = 0.0;

  auto lcmb = [&](int l) -> bool { return (lcm >> (l - 1)) & 1ull; };

  // WENO first sweep for layer j (hybgen.F90:2898-2938)
  auto weno_first = [&](int j, double& c1p, double& c2p, double& z1, double& z2) {
    if (j == 1 || j == K) {
      c1p = s[j];
      c2p = s[j];
      z1 = 0.0;
      z2 = 0.0;
    } else if (lcmb(j) || dpif(j) <= thin) {
      c1p = s[j];
      c2p = s[j];
      z1 = 0.0;
      z2 = 0.0;
    } else {
      const double dm = dpif(j - 1), d0 = dpif(j), dq = dpif(j + 1);
      double zw3_j = (1.0 / (dm + d0)) * (s[j] - s[j - 1]);
      double zw3_jp = (1.0 / (d0 + dq)) * (s[j + 1] - s[j]);
      double q001 = d0 * zw3_jp;
      double q002 = d0 * zw3_j;
      if (q001 * q002 < 0.0) { q001 = 0.0; q002 = 0.0; }
      double dpjm2jp = dm + 2.0 * d0 + dq;
      double q01 = dpjm2jp * zw3_jp;
      double q02 = dpjm2jp * zw3_j;
      if (dabs(q001) > dabs(q02)) q001 = q02;
      if (dabs(q002) > dabs(q01)) q002 = q01;
      double q = (q001 - q002) * (1.0 / (dm + d0 + dq));
      q001 = q001 - q * dq;
      q002 = q002 + q * dm;
      c2p = s[j] + q001;
      c1p = s[j] - q002;
      double a = 2.0 * q001 - q002, b = 2.0 * q002 - q001;
      z1 = a * a;
      z2 = b * b;
    }
  };

  if (METHOD == 2) {
    al_cur = s[1];  // al(1)
    as_cur = 0.0;   // as(1)
  }
  if (METHOD == 3) {
    weno_first(1, wc1[0], wc2[0], wz1[0], wz2[0]);
    weno_first(2, wc1[1], wc2[1], wz1[1], wz2[1]);
  }

#pragma unroll
  for (int l = 1; l <= K; l++) {
    // ---------------- coefficients of layer l ----------------
    double c1l = 0.0, c2l = 0.0, c3l = 0.0;
    if (METHOD == 1) {  // hybgen_plm_coefs on dprs
      if (l == 1 || l == K || lcmb(l) || dprs(l) <= thin) {
        c1l = 0.0;
      } else {
        double qcen = dprs(l) / (dprs(l) + 0.5 * (dprs(l - 1) + dprs(l + 1)));
        double ztop = 2.0 * (s[l] - s[l - 1]);
        double zbot = 2.0 * (s[l + 1] - s[l]);
        double zcen = qcen * (s[l + 1] - s[l - 1]);
        if (ztop * zbot > 0.0) c1l = dsign(dmin3(dabs(zcen), dabs(zbot), dabs(ztop)), zbot);
        else c1l = 0.0;
      }
    } else if (METHOD == 2) {  // hybgen_ppm_coefs on dpi
      // as(l+1) and raw al(l+1)
      double as_nxt = 0.0, al_nxt = 0.0;
      const int jn = l + 1;
      double qdpjp_l = 0.0;  // qdpjp(l) = 1/(dp(l)+dp(l+1))
      if (l <= K - 1) qdpjp_l = 1.0 / (dpif(l) + dpif(l + 1));
      if (jn >= 2 && jn <= K - 1) {
        if (!(lcmb(jn) || dpif(jn) <= thin)) {
          double slj = s[jn] - s[jn - 1];
          double srj = s[jn + 1] - s[jn];
          if (slj * srj > 0.0) {
            const double dm = dpif(jn - 1), d0 = dpif(jn), dq = dpif(jn + 1);
            double dpjp_m = dm + d0;
            double dpjp_j = d0 + dq;
            double dp2jp_m = dm + dpjp_m;
            double dpj2p_j = dpjp_j + dq;
            double qdpjp_j = 1.0 / dpjp_j;
            double dpq3 = d0 / (dm + dpjp_j);
            double scj = dpq3 * (dp2jp_m * srj * qdpjp_j + dpj2p_j * slj * qdpjp_l);
            as_nxt = dsign(dmin3(dabs(2.0 * slj), dabs(scj), dabs(2.0 * srj)), scj);
          }
        }
      }
      if (jn == 2) {
        al_nxt = s[1];
      } else if (jn >= 3 && jn <= K - 1) {
        const double dm2 = dpif(jn - 2), dm = dpif(jn - 1), d0 = dpif(jn), dq = dpif(jn + 1);
        double dpjp_m2 = dm2 + dm;
        double dpjp_m = dm + d0;
        double dpjp_j = d0 + dq;
        double dp2jp_m = dm + dpjp_m;
        double dpj2p_m = dpjp_m + d0;
        double qdpjp_m = qdpjp_l;  // qdpjp(jn-1) = qdpjp(l)
        double qdp2jp_m = 1.0 / dp2jp_m;
        double qdpj2p_m = 1.0 / dpj2p_m;
        double qdp4 = 1.0 / (dpjp_m2 + dpjp_j);
        al_nxt = s[jn - 1] + dm * (s[jn] - s[jn - 1]) * qdpjp_m +
                 qdp4 * (2.0 * d0 * dm * qdpjp_m * (s[jn] - s[jn - 1]) *
                             (dpjp_m2 * qdp2jp_m - dpjp_j * qdpj2p_m) -
                         dm * as_nxt * dpjp_m2 * qdp2jp_m + d0 * as_cur * dpjp_j * qdpj2p_m);
      } else if (jn == K) {
        al_nxt = s[K];
      }
      double al = al_cur, ar;
      if (l == 1) ar = s[1];
      else if (l <= K - 2) ar = al_nxt;
      else ar = s[K];
      if (l >= 2 && l <= K - 1) {
        if (lcmb(l) || dpif(l) <= thin) {
          al = s[l];
          ar = s[l];
        } else if ((s[l + 1] - s[l]) * (s[l] - s[l - 1]) <= 0.0) {
          al = s[l];
          ar = s[l];
        } else {
          double da = ar - al;
          double a6 = 6.0 * s[l] - 3.0 * (al + ar);
          if (da * a6 > da * da) {
            al = 3.0 * s[l] - 2.0 * ar;
          } else if (da * a6 < -(da * da)) {
            ar = 3.0 * s[l] - 2.0 * al;
          }
        }
      }
      if (al != ar) {
        c1l = al;
        c2l = ar - al;
        c3l = 6.0 * s[l] - 3.0 * (al + ar);
      } else {
        c1l = al;
        c2l = 0.0;
        c3l = 0.0;
      }
      as_cur = as_nxt;
      al_cur = al_nxt;
      qdpjp_prev = qdpjp_l;
    } else if (METHOD == 3) {  // hybgen_weno_coefs on dpi
      // first-sweep values of l+2 (if any) for zw3(l+1)
      if (l + 2 <= K) weno_first(l + 2, wc1[2], wc2[2], wz1[2], wz2[2]);
      double zw3_next = 0.0;
      if (l + 1 <= K) {
        double q002 = dmax(wz2[0], dsmll);
        double q001 = dmax(wz1[1], dsmll);
        zw3_next = (q001 * wc2[0] + q002 * wc1[1]) / (q001 + q002);
      }
      c1l = wc1[0];
      c2l = wc2[0];
      if (l >= 2 && l <= K - 1 && !(lcmb(l) || dpif(l) <= thin)) {
        double q01 = zw3_next - s[l];
        double q02 = s[l] - zw3_l;
        double q001 = 2.0 * q01;
        double q002 = 2.0 * q02;
        if (q01 * q02 < 0.0) { q01 = 0.0; q02 = 0.0; }
        else if (dabs(q01) > dabs(q002)) { q01 = q002; }
        else if (dabs(q02) > dabs(q001)) { q02 = q001; }
        c1l = s[l] - q02;
        c2l = s[l] + q01;
      }
      zw3_l = zw3_next;
      wc1[0] = wc1[1]; wc2[0] = wc2[1]; wz1[0] = wz1[1]; wz2[0] = wz2[1];
      wc1[1] = wc1[2]; wc2[1] = wc2[2]; wz1[1] = wz1[2]; wz2[1] = wz2[2];
    }
    (void)qdpjp_prev;

    const double pil = pi[l], pil1 = pi[l + 1];
    const double dprsl = pil1 - pil;
    // ---------------- an output in progress ----------------
    if (accon) {
      if (lb_act == l) {
        double v;
        if (acc_thin) {
          v = (kc != 1) ? prev : s[1];
        } else {
          double xb = (zb_act - pil) / dmax(dprsl, thin);
          if (METHOD == 0) {
            double dpb = zb_act - pil;
            acc = acc + dpb * (s[l] - o);
          } else if (METHOD == 1) {
            double qb0 = xb, qb1 = xb * xb * 0.5;
            double cc0 = s[l] - o - 0.5 * c1l;
            acc = acc + dprsl * (cc0 * qb0 + c1l * qb1);
          } else if (METHOD == 2) {
            double qb0 = xb, qb1 = xb * xb * 0.5, qb2 = xb * xb * xb / 3.0;
            acc = acc + dprsl * ((c1l - o) * qb0 + (c2l + c3l) * qb1 - c3l * qb2);
          } else {
            double dpb = zb_act - pil;
            double qb1 = (xb - 1.0) * (xb - 1.0);
            double qb2 = qb1 - 1.0 + xb;
            double qb0 = 1.0 - qb1 - qb2;
            acc = acc + dpb * (qb0 * (s[l] - o) + qb1 * (c1l - o) + qb2 * (c2l - o));
          }
          v = o + acc / (zb_act - zt_act);
          v = dmax(si_min, v);
          v = dmin(si_max, v);
        }
        out(kc, v);
        prev = v;
        lt_c = l;
        zt_c = zb_act;
        kc = kc + 1;
        accon = false;
      } else if (!acc_thin) {
        acc = acc + dprsl * (s[l] - o);
      }
    }
    // ---------------- outputs that start in layer l ----------------
    if (!accon && lt_c == l) {
      while (kc <= K) {
        const double zt = zt_c;
        const double zb = dmin(po(kc + 1), zx);
        const bool ends_here = (l == K) || (pil1 >= zb);
        const bool thn = (zb - zt <= thin) || (zt >= zx);
        if (ends_here) {
          double v;
          if (thn) {
            v = (kc != 1) ? prev : s[1];
          } else if (METHOD == 0) {
            v = s[l];
          } else {
            double xt = (zt - pil) / dmax(dprsl, thin);
            double xb = (zb - pil) / dmax(dprsl, thin);
            if (METHOD == 1) {
              double qt1 = (xb * xb - xt * xt - (xb - xt)) * 0.5;
              double sz = dprsl * (c1l * qt1);
              v = s[l] + sz / (zb - zt);
            } else if (METHOD == 2) {
              double qt0 = (xb - xt);
              double qt1 = (xb * xb - xt * xt) * 0.5;
              double qt2 = (xb * xb * xb - xt * xt * xt) / 3.0;
              double oo = s[l];
              double sz = dprsl * ((c1l - oo) * qt0 + (c2l + c3l) * qt1 - c3l * qt2);
              v = oo + sz / (zb - zt);
            } else {
              double qt1 = xb * xb + xt * xt + xb * xt + 1.0 - 2.0 * (xb + xt);
              double qt2 = qt1 - 1.0 + (xb + xt);
              double oo = s[l];
              double sz = qt1 * (c1l - oo) + qt2 * (c2l - oo);
              v = oo + sz;
            }
            v = dmax(si_min, v);
            v = dmin(si_max, v);
          }
          out(kc, v);
          prev = v;
          zt_c = zb;
          kc = kc + 1;
        } else {
          // output kc starts in layer l and ends in a deeper layer lb
          int lb = K;
          bool found = false;
          constexpr int kScan = 8;
#pragma unroll
          for (int d = l + 1; d <= K && d <= l + kScan; d++) {
            if (!found && (d == K || pi[d + 1] >= zb)) {
              lb = d;
              found = true;
            }
          }
          if (!found) {
#pragma unroll
            for (int d = l + kScan + 1; d <= K; d++) {
              if (!found && (d == K || pi[d + 1] >= zb)) {
                lb = d;
                found = true;
              }
            }
          }
          lb_act = lb;
          zt_act = zt;
          zb_act = zb;
          acc_thin = thn;
          accon = true;
          if (!thn) {
            const int mid = (l + lb) / 2;
            o = s[l];
            constexpr int kMid = 4;
#pragma unroll
            for (int d = l + 1; d <= K && d <= l + kMid; d++)
              if (d == mid) o = s[d];
            if (mid > l + kMid) {
#pragma unroll
              for (int d = l + kMid + 1; d <= K; d++)
                if (d == mid) o = s[d];
            }
            double xt = (zt - pil) / dmax(dprsl, thin);
            if (METHOD == 0) {
              double dpt = pil1 - zt;
              acc = dpt * (s[l] - o);
            } else if (METHOD == 1) {
              double qt0 = (1.0 - xt), qt1 = (1.0 - xt * xt) * 0.5;
              double cc0 = s[l] - o - 0.5 * c1l;
              acc = dprsl * (cc0 * qt0 + c1l * qt1);
            } else if (METHOD == 2) {
              double qt0 = (1.0 - xt);
              double qt1 = (1.0 - xt * xt) * 0.5;
              double qt2 = (1.0 - xt * xt * xt) / 3.0;
              acc = dprsl * ((c1l - o) * qt0 + (c2l + c3l) * qt1 - c3l * qt2);
            } else {
              double dpt = pil1 - zt;
              double qt1 = xt * (xt - 1.0);
              double qt2 = qt1 + xt;
              double qt0 = 1.0 - qt1 - qt2;
              acc = dpt * (qt0 * (s[l] - o) + qt1 * (c1l - o) + qt2 * (c2l - o));
            }
          }
          break;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------
// Kernel M body: remap of T, S and the tracers of one column and the
// th3d write-back (hybgen.F90:1530-1591).  prsf is the lane's LDS row.
// ---------------------------------------------------------------------
template <int K, int METHOD>
HYB_HD void remap_fast(const KParams& P, const View& V, const Scratch& W, int i, int j,
                       int n, const LaneRow& prsf) {
  const int64_t c0 = ij_of(V, i, j);
  const int64_t S = V.NIJ;
  double pi[K + 2];
  const double* pres = W.pres + c0;
#pragma unroll
  for (int k = 1; k <= K + 1; k++) pi[k] = pres[(int64_t)(k - 1) * S];
  pi[0] = 0.0;
  const double* pg = V.p + c0;
#pragma unroll
  for (int k = 1; k <= K + 1; k++) prsf(k) = pg[(int64_t)(k - 1) * S];
  const int kp = W.kp[c0];
  const uint64_t lcm = W.lcm[c0];
  double s[K + 2];
  s[0] = 0.0;
  s[K + 1] = 0.0;
  auto do_field = [&](double* f) {
#pragma unroll
    for (int k = 1; k <= K; k++) s[k] = f[(int64_t)(k - 1) * S];
    // massless fill, hybgen.F90:974-996 (k <= nhybrd)
    double skp = sel(s, kp);
#pragma unroll
    for (int k = 2; k <= K; k++)
      if (k > kp) s[k] = skp;
    remap_l<K, METHOD>(s, pi, lcm, prsf,
                       [&](int k, double v) { f[(int64_t)(k - 1) * S] = v; });
  };
  double* temp = V.temp + c0 + lev4(V, 1, n);
  double* saln = V.saln + c0 + lev4(V, 1, n);
  double* th3d = V.th3d + c0 + lev4(V, 1, n);
  do_field(temp);
  do_field(saln);
#pragma unroll
  for (int k = 1; k <= K; k++)
    th3d[(int64_t)(k - 1) * S] = sig(temp[(int64_t)(k - 1) * S], saln[(int64_t)(k - 1) * S]) - P.thbase;
  for (int t = 1; t <= P.ntracr; t++) do_field(V.tracer + c0 + levt(V, 1, n, t));
}

// ---------------------------------------------------------------------
// Kernel C body: one u- or v-point, as point_c in hyb_column.h.
// ---------------------------------------------------------------------
template <int K>
HYB_HD void uv_fast(const KParams& P, const View& V, int i, int j, int nl, bool isv,
                    const LaneRow& prsf) {
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
  double pi[K + 2], s[K + 2];
  pi[0] = 0.0;
  pi[1] = 0.0;
  s[0] = 0.0;
  s[K + 1] = 0.0;
#pragma unroll
  for (int k = 1; k <= K; k++) {
    pi[k + 1] = pi[k] + dpu[(int64_t)(k - 1) * S];
    s[k] = uu[(int64_t)(k - 1) * S];
  }
  double pf = 0.0;
  prsf(1) = 0.0;
  pu[0] = 0.0;
  double b = dmin(depthu, 0.5 * (p0[0] + p1[0]));
#pragma unroll
  for (int k = 1; k <= K; k++) {
    double a = dmin(depthu, 0.5 * (p0[(int64_t)k * S] + p1[(int64_t)k * S]));
    double d = dmax(0.0, a - b);
    b = a;
    dpu[(int64_t)(k - 1) * S] = d;
    pf = pf + d;
    pu[(int64_t)k * S] = pf;
    prsf(k + 1) = pf;
  }
  auto store = [&](int k, double v) { uu[(int64_t)(k - 1) * S] = v; };
  if (P.hybmap == 0) remap_l<K, 0>(s, pi, 0ull, prsf, store);
  else if (P.hybmap == 1 && P.hybiso > 2.0) remap_l<K, 1>(s, pi, 0ull, prsf, store);
  else remap_l<K, 3>(s, pi, 0ull, prsf, store);
  // thin near-bottom layers, hybgen.F90:1825-1833
  const double pbot = pf - kOnemm;
#pragma unroll
  for (int k = 1; k <= K; k++) {
    double dpi = dmax(pi[k + 1] - pi[k], dpthin);
    if (!(dpi > dpthin || prsf(k) <= pbot)) uu[(int64_t)(k - 1) * S] = -ub;
  }
}

}  // namespace hyb
