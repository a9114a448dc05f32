// hyb_stream.h -- FP64 hybgen column kernels with rolled k-loops and small
// register windows: no per-lane arrays, one code path for any kdm <= 64.
//
// Every floating-point expression keeps the reference's operands and
// operation order; only the loop structure differs from hybgen.F90.
//
// regrid_stream (hybgenaj, hybgen.F90:399-1529) makes three forward passes
// over a column:
//   1. interfaces p, fixed-layer count, relaxation prefix and kp
//      (399-598); p goes to the per-column scratch `pres`;
//   2. overturn / unmix of layer kp (606-970), in place;
//   3. interface relocation (1065-1494) on a 7-deep window p(k-3..k+3),
//      then, level by level as each p(k-2) becomes final, the maximum
//      thickness pass (1496-1509) and the rebuild of dp and p (1523-1529).
//   The two reference loops that write below the window -- the layer-1
//   thickness loop (1051-1063) and the fixed-layer order loop (1098-1107)
//   -- raise a monotone run of p to a floor, so they are applied as levels
//   enter the window.  dp0ij and qhrlx of level x are recomputed from the
//   initial p(x) when level x enters the window.
//
// remap_stream (1530-1591 with the hybgen_*_remap/_coefs routines) walks
// the old layers l = 1..kk once for up to kNFMax fields at a time (T and S
// first, with th3d), carrying each field's values s(l-1..l+2) and
// reconstruction state in registers.  The per-layer metrics (1/(dp+dp)
// etc.) and the per-output weights are shared by the fields.  Outputs are
// written in place as they complete; regrid_stream flags the columns where
// an output would overwrite a level the walk has not read yet, and those
// columns go to remap_fallback.  The clamp to the column's [min,max]
// (e.g. 2836-2837) is exact for an output inside the range of the levels
// read so far; other outputs are stored unclamped, marked, and clamped
// once the whole column has been read.
#pragma once

#include "hyb_column.h"
#include <type_traits>


namespace hyb {

constexpr int32_t kUnsafe = 1 << 16;  // kp word: remap this column in remap_fallback
constexpr int kNFMax = 6;             // fields per remap walk
constexpr int kKMax = 64;             // largest kdm

// minimum thickness of level k (hybgen.F90:456-466) from the initial
// interface pk = p(k) and dp0cum(k), qdep == 1; *qh = qhrlx(k+1).
HYB_HD double dp0ij_qh(const KParams& P, double d, double pk, double dp0cum_k, double* qh) {
  double q, qts;
  if (d <= P.dp00i) {
    q = d;
    qts = 0.0;
  } else {
    q = dmax(P.dp00i, d * d / dmax(d, pk - dp0cum_k));
    qts = 1.0 - (q - P.dp00i) / (d - P.dp00i);
  }
  *qh = 1.0 / (1.0 + qts * (P.hybrlx - 1.0));
  return dmin(q, d);
}
// dp0ij(k) when qdep < 1 (hybgen.F90:405-418)
HYB_HD double dp0ij_tf_of(const KParams& P, double qdep, int k) {
  return qdep * P.dp0k[k - 1] + (1.0 - qdep) * P.ds0k[k - 1];
}
HYB_HD double dp0ij_only(const KParams& P, double d, double pk, double dp0cum_k) {
  double q;
  if (d <= P.dp00i) {
    q = d;
  } else {
    q = dmax(P.dp00i, d * d / dmax(d, pk - dp0cum_k));
  }
  return dmin(q, d);
}

template <int W>
HYB_HD void shl(double (&a)[W], double v) {
#pragma unroll
  for (int q = 0; q < W - 1; q++) a[q] = a[q + 1];
  a[W - 1] = v;
}

// ---------------------------------------------------------------------
// regrid of one p-column (hybgenaj up to the remap)
// ---------------------------------------------------------------------
// AHEAD: the pass-3 window's next levels (pres, th3d, theta) are loaded
// AHEAD (0, 1 or 2) layer iterations before the slide takes them in (more
// registers)
template <int AHEAD = 0>
HYB_HDI void regrid_stream(const KParams& P, const View& V, const Scratch& W, int i, int j,
                           int n) {
  const int K = P.kk;
  const int64_t c0 = ij_of(V, i, j);
  const int64_t S = V.NIJ;
  const int ntracr = P.ntracr;
  double* temp = V.temp + c0 + lev4(V, 1, n);
  double* saln = V.saln + c0 + lev4(V, 1, n);
  double* th3d = V.th3d + c0 + lev4(V, 1, n);
  double* dpc = V.dp + c0 + lev4(V, 1, n);
  double* pres = W.pres + c0;
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
  const double rthin = P.hybthn;
  const double topiso = V.topiso[c0];

  // ------------------------------------------------ pass 1: 399-598
  double qdep;
  if (P.dpns == P.dsns) {
    qdep = 1.0;
  } else {
    qdep = dmax(0.0, dmin(1.0, (V.depths[c0] - P.dsns) / (P.dpns - P.dsns)));
  }
  const bool tf = qdep < 1.0;
  const double dp0ij1 = tf ? qdep * P.dp0k[0] + (1.0 - qdep) * P.ds0k[0] : P.dp0k[0];
  double pk = 0.0;               // p(k)
  double cumk = 0.0 + dp0ij1;    // dp0cum(k), k = 2
  double pk1 = pk + FK(dpc, 1);  // p(k+1), k = 1
  pres[0] = 0.0;
  pres[S] = pk1;
  // The remap's fast quotients need every old interface (after the
  // overturn) and every new one to be 0 or in [2^-60, 2^60], the old ones
  // in order; a column where that fails goes to the fallback remap.
  bool tame = (FK(dpc, 1) >= 0.0);
  tame &= tame_p(pk1);
  bool mono = FK(dpc, 1) >= 0.0;  // p(1..kk+1) in order
  int fixlay = 1, fixall = 1, phase = 1, qfix = 2;  // qhrlx(x) = 1 for x <= qfix
  int kp = 2;
  auto level = [&](int k, double dk) {
    pk = pk1;
    pk1 = pk + dk;
    pres[(int64_t)k * S] = pk1;
    tame &= (dk >= 0.0);
    tame &= tame_p(pk1);
    mono &= (dk >= 0.0);
    if (k >= 3 && pk1 - pk >= dpthin) kp = k;  // 590-598: the deepest such k
    double dtf = 0.0, cumk1;
    if (tf) {
      dtf = qdep * P.dp0k[k - 1] + (1.0 - qdep) * P.ds0k[k - 1];
      cumk1 = cumk + dtf;
    } else {
      cumk1 = P.dp0cum[k];
    }
    if (phase == 1) {  // 479-488
      if (cumk >= topiso) {
        phase = 2;
        fixall = fixlay;
      } else {
        fixlay = fixlay + 1;
        qfix = k + 1;
      }
    }
    if (phase == 2 && !P.ldenfix) {  // 550-573
      const double d0k = tf ? dtf : dp0ij_only(P, P.dp0k[k - 1], pk, cumk);
      if (pk1 > cumk1 + 0.1 * d0k) {
        if (fixlay > fixall) {
          if (pk > cumk) fixlay = fixlay - 1;
        }
        phase = 3;
      } else {
        if (k > qfix) qfix = k;
        fixlay = fixlay + 1;
      }
    }
    cumk = cumk1;
  };
  {
    // dp of levels 2..K in batches of 8, the next batch loaded before the
    // current one is scanned (indices clamped into the column; values
    // beyond it are not used)
    constexpr int B = 8;
    double cur[B];
#pragma unroll
    for (int q = 0; q < B; q++) cur[q] = FK(dpc, (2 + q <= K) ? 2 + q : K);
#pragma unroll 1
    for (int k0 = 2; k0 <= K; k0 += B) {
      double nxt[B];
#pragma unroll
      for (int q = 0; q < B; q++) nxt[q] = FK(dpc, (k0 + B + q <= K) ? k0 + B + q : K);
#pragma unroll
      for (int q = 0; q < B; q++)
        if (k0 + q <= K) level(k0 + q, cur[q]);
#pragma unroll
      for (int q = 0; q < B; q++) cur[q] = nxt[q];
    }
  }
  const double zx = pk1;  // p(kk+1); the overturn does not move it

  if (P.ldenfix) {
    // HYBGEN_DENFIX, 496-549: the fixed layers end where the target
    // interface dp0cum(k+1) is deeper than p(k+1) by a tenth of dp0ij(k)
    // and theta(k) is denser than th3d interpolated to dp0cum(k+1).  The
    // search for the layer l holding dp0cum(k+1) moves down from the last
    // one found (lf), over the interfaces pass 1 has just stored, before the
    // overturn moves any of them.
    auto pp = [&](int x) -> double { return pres[(int64_t)(x - 1) * S]; };  // p(x)
    const int fixall_d = fixlay;
    int lf = 1;
    double plf = pp(1), plf1 = pp(2);  // p(lf), p(lf+1)
    // With the interfaces in order, no l >= lf has p(l) <= zk <= p(l+1)
    // when zk lies above p(lf) or below the bottom p(kk+1), so that search
    // is skipped (it would find nothing); out of order, every search runs
    // as written.
    // dp0cum(k) and dp0cum(k+1) at k = 2, then advanced with the same
    // additions as 399-477 when qdep < 1
    double c_k = tf ? 0.0 + dp0ij1 : P.dp0cum[1];
    for (int k = 2; k <= fixall_d; k++) c_k = tf ? c_k + dp0ij_tf_of(P, qdep, k) : P.dp0cum[k];
    for (int k = fixall_d + 1; k <= K; k++) {
      const double c_k1 = tf ? c_k + dp0ij_tf_of(P, qdep, k) : P.dp0cum[k];  // dp0cum(k+1)
      const double pk_d = pp(k), pk1_d = pp(k + 1), thk_d = FK(theta, k);
      double denfix = FK(th3d, K);
      const double zk = c_k1;
      if (!(mono && (zk > zx || zk < plf))) {
        double pl = plf, pl1 = plf1;
        for (int l = lf; l <= K; l++) {
          if (l > lf) {
            pl = pl1;
            pl1 = pp(l + 1);
          }
          if (pl <= zk && pl1 >= zk) {
            const double z0 = 0.5 * (pl + pl1);
            if (zk <= z0) {
              if (l == 1) {
                denfix = FK(th3d, 1);
              } else {
                const double zm = 0.5 * (pp(l - 1) + pl);
                const double q = (z0 - zk) / (z0 - zm);
                denfix = q * FK(th3d, l - 1) + (1.0 - q) * FK(th3d, l);
              }
            } else {
              if (pl1 == zx) {  // p(l+1) == p(nhybrd+1), nhybrd == kk
                denfix = FK(th3d, K);
              } else {
                const double zp = 0.5 * (pl1 + pp(l + 2));
                const double q = (zk - z0) / (zp - z0);
                denfix = q * FK(th3d, l + 1) + (1.0 - q) * FK(th3d, l);
              }
            }
            lf = l;
            plf = pl;
            plf1 = pl1;
            break;
          }
        }
      }
      const double d0k = tf ? dp0ij_tf_of(P, qdep, k) : dp0ij_only(P, P.dp0k[k - 1], pk_d, c_k);
      if (pk1_d > c_k1 + 0.1 * d0k && thk_d > denfix) {
        if (fixlay > fixall_d) {
          if (pk_d > c_k) fixlay = fixlay - 1;
        }
        break;
      }
      if (k > qfix) qfix = k;  // qhrlx(k) = 1
      fixlay = fixlay + 1;
      c_k = c_k1;
    }
  }

  // ------------------------------------------------ pass 2: 606-970
  // The overturn moves p(kp0) or p(kp0-1) only; dp0ij and qhrlx are
  // functions of the initial p, so those two values are kept.
  const int kp0 = kp;
  const double pin_a = pres[(int64_t)(kp0 - 2) * S];  // p(kp0-1)
  const double pin_b = pres[(int64_t)(kp0 - 1) * S];  // p(kp0)
  auto p_init = [&](int x, double cur) -> double {
    return x == kp0 ? pin_b : (x == kp0 - 1 ? pin_a : cur);
  };
  auto qhrlx_of = [&](int x) -> double {  // final qhrlx(x)
    if (tf || x <= 2 || x <= qfix) return 1.0;
    double qh;
    (void)dp0ij_qh(P, P.dp0k[x - 2], p_init(x - 1, pres[(int64_t)(x - 2) * S]),
                   P.dp0cum[x - 2], &qh);
    return qh;
  };
  {
    const int k = kp;
    const int ka = (k - 2 > 1) ? k - 2 : 1;
    const double pkk = pres[(int64_t)(k - 1) * S], pkm = pres[(int64_t)(k - 2) * S],
                 pkp = pres[(int64_t)k * S];
    const double thk = FK(theta, k);
    const double tk = FK(th3d, k), tkm = FK(th3d, k - 1), tka = FK(th3d, ka);
    if (k > fixlay + 1 && qdep == 1.0 && pkk - pkm >= dpthin && thk - kEpsil > tk &&
        tkm > tk && tka > tk) {
      // 609-686 entrain layer k into k-1
      double q = (pkp - pkk) / (pkp - pkm);
      FK(temp, k - 1) = FK(temp, k - 1) - q * (FK(temp, k - 1) - FK(temp, k));
      FK(saln, k - 1) = FK(saln, k - 1) - q * (FK(saln, k - 1) - FK(saln, k));
      FK(th3d, k - 1) = sig(FK(temp, k - 1), FK(saln, k - 1)) - P.thbase;
      for (int t = 1; t <= ntracr; t++)
        TRK(t, k - 1) = TRK(t, k - 1) - q * (TRK(t, k - 1) - TRK(t, k));
      pres[(int64_t)(k - 1) * S] = pkp;
      kp = k - 1;
    } else if (k > fixlay + 1 && qdep == 1.0 && pkk - pkm >= dpthin && thk - kEpsil > tk &&
               tkm > tk) {
      // 687-811 swap with the layer above
      if (pkp - pkk <= pkk - pkm) {
        double st = FK(temp, k - 1), ss = FK(saln, k - 1), sh = FK(th3d, k - 1);
        double q = (pkp - pkk) / (pkk - pkm);
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
        double q = (pkk - pkm) / (pkp - pkk);
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
  {
    // 813-970 unmix a too-light deepest layer
    const int k = kp;
    const double pkk = pres[(int64_t)(k - 1) * S], pkm = pres[(int64_t)(k - 2) * S],
                 pkp = pres[(int64_t)k * S];
    const double dpthck = P.dx0k[k - 2] * 0.95;
    const double thk = FK(theta, k), thkm = FK(theta, k - 1);
    const double tk = FK(th3d, k), tkm = FK(th3d, k - 1);
    if (k > fixlay + 1 && qdep == 1.0 && pkk - pkm >= dpthin && pkk - pkm < dpthck &&
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
      double q = dmin3(FK(temp, ka), FK(temp, k - 1), FK(temp, k));
      if (q > 6.0) {
        deltm = dmin(deltm, 6.0 * (thk - thkm));
      } else {
        deltm = dmin(deltm, 10.0 * (thk - thkm));
      }
      delsm = dmin(delsm, 1.3 * (thk - thkm));
      double qts = 0.0;
      if (delt > kEpsil) qts = dmax(qts, (dmin(deltm, 2.0 * delt) - delt) / delt);
      if (dels > kEpsil) qts = dmax(qts, (dmin(delsm, 2.0 * dels) - dels) / dels);
      q = (thk - tk) / (thk - tkm);
      q = dmin(q, qts / (1.0 + qts));
      q = qhrlx_of(k) * q;
      const double p_hat = q * (pkp - pkk);
      const double pknew = pkk + p_hat;
      pres[(int64_t)(k - 1) * S] = pknew;
      tame &= tame_p(pknew);
      tame &= (pknew >= pkm);
      tame &= (pknew <= pkp);
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
  // 972-1009 (nhybrd == kk): levels below kp take level kp's values; the
  // remap applies this fill to T, S and the tracers as it loads them.
  const double th_kp = FK(th3d, kp);
  auto th3d_eff = [&](int x) -> double { return x > kp ? th_kp : FK(th3d, x); };

  // ------------------------------------------------ pass 3: 1036-1529
  const double floor1 = 0.0 + dp0ij1;              // 1053-1055: p(1)+dp0ij(1)
  // p(kk+1) after 1056-1062; the thick-thin branches below move it where
  // the deepest layer starts with a negative thickness
  double pbot = zx < floor1 ? floor1 : zx;
  double floor2 = 0.0;                             // 1098-1107, once set
  int floor2_from = K + 2;
  auto with_floor = [&](int x, double v) -> double {
    if (x == 2) v = floor1;
    else if (x >= 3 && v < floor1) v = floor1;
    if (x >= floor2_from && x <= K && v < floor2) v = floor2;
    return v;
  };
  auto raw_at = [&](int x) -> double {
    return (x >= 1 && x <= K + 1) ? pres[(int64_t)(x - 1) * S] : 0.0;
  };
  double tfc[4] = {0.0, 0.0, 0.0, 0.0};  // dp0cum(k-2..k+1) when qdep < 1
  auto dp0ij_tf = [&](int x) -> double {
    return qdep * P.dp0k[x - 1] + (1.0 - qdep) * P.ds0k[x - 1];
  };
  auto dp0cum_at = [&](int x, int k) -> double {
    return tf ? tfc[x - (k - 2)] : P.dp0cum[x - 1];
  };
  uint64_t lcm = 0;  // 1036-1045 from pres(x), pres(x+1), th3d(x), theta(x)
  auto lcm_set = [&](int x, double r0, double r1, double tx, double thx) {
    if (P.isopcm && x > fixlay && ((r1 - r0) <= dpthin || dabs(tx - thx) < P.hybiso))
      lcm |= 1ull << (x - 1);
  };
  // windows at k = 2
  double pw[7];          // p(k-3..k+3)
  double rw[3];          // pres(k+1..k+3) as left by the overturn
  double tw[4], thw[4];  // th3d (filled) and theta, levels k-1..k+2
  double dw[5];          // dp0ij(k-3..k+1)
  double qw[4];          // qhrlx(k-2..k+1)
  double qh_next = 1.0;  // qhrlx(k+2) before the fixed-layer override
#pragma unroll
  for (int q = 0; q < 7; q++) {
    const int x = q - 1;
    pw[q] = (x >= 1 && x <= K + 1) ? with_floor(x, raw_at(x)) : 0.0;
  }
  rw[0] = raw_at(3);
  rw[1] = raw_at(4);
  rw[2] = raw_at(5);
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int x = q + 1;
    tw[q] = (x <= K) ? th3d_eff(x) : 0.0;
    thw[q] = (x <= K) ? FK(theta, x) : 0.0;
  }
  {
    double r = raw_at(1);
#pragma unroll
    for (int x = 1; x <= 4; x++) {
      const double r1 = raw_at(x + 1);
      if (x <= K) lcm_set(x, r, r1, tw[x - 1], thw[x - 1]);
      r = r1;
    }
  }
  dw[0] = 0.0;
  dw[1] = 0.0;
  dw[2] = dp0ij1;
  qw[0] = 1.0;
  qw[1] = 1.0;
  qw[2] = 1.0;  // qhrlx(2)
  if (tf) {
    tfc[1] = 0.0;              // dp0cum(1)
    tfc[2] = tfc[1] + dp0ij1;  // dp0cum(2)
    dw[3] = K >= 2 ? dp0ij_tf(2) : 0.0;
    tfc[3] = tfc[2] + dw[3];   // dp0cum(3)
    dw[4] = K >= 3 ? dp0ij_tf(3) : 0.0;
    qw[3] = 1.0;
  } else {
    double qh3 = 1.0;
    dw[3] = K >= 2 ? dp0ij_qh(P, P.dp0k[1], p_init(2, raw_at(2)), P.dp0cum[1], &qh3) : 0.0;
    qw[3] = (3 <= qfix) ? 1.0 : qh3;
    dw[4] = K >= 3 ? dp0ij_qh(P, P.dp0k[2], p_init(3, raw_at(3)), P.dp0cum[2], &qh_next) : 0.0;
  }

  // 1496-1529 as each interface becomes final, plus the in-place check
  double pf = 0.0, prs = 0.0;
  double* pg = V.p + c0;
  bool unsafe = false;
  // pchk: pres(x-3) as left by the overturn, for the in-place check
  auto consume = [&](int x, double v, double pchk) {
    if (x == 1) {
      pf = v;
      prs = v;
      pg[0] = v;
      return;
    }
    const int jx = x - 1;
    if (jx >= 2 && jx <= K - 1 && (v - pf) > P.dx0k[jx - 1]) {
      v = pf + P.dx0k[jx - 1];
      lcm &= ~(3ull << (jx - 2));
    }
    const double d = dmax(v - prs, 0.0);
    FK(dpc, jx) = d;
    prs = prs + d;
    pg[(int64_t)jx * S] = prs;
    tame &= tame_p(prs);
    pf = v;
    // remap output jx ends in old layer lb <= jx-3, before the old layer
    // that reaches the bottom, only if pres(jx-2) >= zb = min(prsf(jx+1),zx)
    if (jx >= 4) {
      const double zb = dmin(prs, zx);
      if (zb < zx && pchk >= zb) unsafe = true;
    }
  };

  // with AHEAD: pres(k+4), th3d(k+3), theta(k+3) for the slide of
  // iteration k, loaded during iteration k-AHEAD (indices clamped into the
  // column; values beyond it are not used)
  auto cl1 = [&](int x) -> int64_t { return (int64_t)((x <= K + 1 ? x : K + 1) - 1) * S; };
  auto clk = [&](int x) -> int64_t { return (int64_t)((x <= K ? x : K) - 1) * S; };
  constexpr int AQ = AHEAD > 0 ? AHEAD : 1;
  double nr4[AQ], nth3[AQ], nta3[AQ];
#pragma unroll
  for (int q = 0; q < AQ; q++) {
    nr4[q] = nth3[q] = nta3[q] = 0.0;
    if (AHEAD > q) {
      nr4[q] = pres[cl1(6 + q)];
      nth3[q] = th3d[clk(5 + q)];
      nta3[q] = theta[clk(5 + q)];
    }
  }
  double q, p_hat, p_hat0 = 0.0, p_hat2, p_hat3;
#pragma unroll 1
  for (int k = 2; k <= K; k++) {
    // pres(k-5) for the in-place check of consume(k-2), issued early
    const double pchk = (k >= 7) ? pres[(int64_t)(k - 6) * S] : 0.0;
#define PW(x) pw[(x) - (k - 3)]
    const double dm1 = dw[2], dm2 = dw[1], dm3 = dw[0], d0 = dw[3], dp1 = dw[4];
    const double qm1 = qw[1], qm2 = qw[0], q0 = qw[2], qp1 = qw[3];
    if (k <= fixlay) {
      // 1092-1108
      if (k < K) {
        PW(k + 1) = dmin(dp0cum_at(k + 1, k), pbot);
        if (k == fixlay) {
          floor2 = PW(k + 1);
          floor2_from = k + 2;
          if (k + 2 <= K && PW(k + 2) < floor2) PW(k + 2) = floor2;
          if (k + 3 <= K && PW(k + 3) < floor2) PW(k + 3) = floor2;
        }
      }
    } else {
      const double thk = thw[1], thm = thw[0];
      const double tk = tw[1], tkm = tw[0];
      double dp_top = PW(k) - PW(k - 1) + kEpsil;
      double dp_cen = PW(k + 1) - PW(k) + kEpsil;
      double dp_bot, dp_far, dp_inc;
      if (k == K) {
        dp_bot = dp_cen;
        dp_far = dp_cen;
      } else if (k == K - 1) {
        dp_bot = PW(k + 2) - PW(k + 1) + kEpsil;
        dp_far = dp_bot;
      } else {
        dp_bot = PW(k + 2) - PW(k + 1) + kEpsil;
        dp_far = PW(k + 3) - PW(k + 2) + kEpsil;
      }
      // With thicknesses >= 0, hybthk <= 0.9 makes the first test false at
      // k == kk and the second false at k >= kk-1.  A deepest layer of
      // negative thickness (an interface under the bottom by a float32
      // rounding, as a start from a climatology gives) makes the second
      // true at kk-1 and the first true at kk; both then take the
      // thickness weighting (min < 0.2*max), so th3d(kk+1) stays unread,
      // and both move p(kk+1).  The second test stays off at kk: it can
      // pass there only with two layers of negative thickness, and would
      // move p(kk+2), an interface the column does not have.
      if (dp_cen < P.hybthk * dmin(dp_top, dp_bot)) {
        // 1131-1159 thick-thin-thick
        const double thp = thw[2], tkp = tw[2];
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
        p_hat = PW(k) - q * dp_inc;
        PW(k) = (1.0 - q0) * PW(k) + q0 * p_hat;
        p_hat = PW(k + 1) + (1.0 - q) * dp_inc;
        PW(k + 1) = (1.0 - q0) * PW(k + 1) + q0 * p_hat;
        if (k == K) pbot = PW(k + 1);
      } else if (k < K && dp_cen + dp_bot < P.hybthk * dmin(dp_top, dp_far)) {
        // 1178-1208 thick-thin-thin-thick
        const double thp = thw[2], thp2 = thw[3], tkp2 = tw[3];
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
        p_hat = PW(k) - q * dp_inc;
        PW(k) = (1.0 - q0) * PW(k) + q0 * p_hat;
        p_hat = PW(k + 2) + (1.0 - q) * dp_inc;
        PW(k + 2) = (1.0 - q0) * PW(k + 2) + q0 * p_hat;
        if (k == K - 1) pbot = PW(k + 2);
      } else if (tk > thk + kEpsil && k > fixlay + 1) {
        // 1237-1387 layer too dense: move the upper interface up
        bool ltop;
        if (rthin * (PW(k) - PW(k - 1)) <= (PW(k + 1) - PW(k))) {
          ltop = true;
        } else if (rthin * (PW(k + 1) - PW(k)) <= (PW(k) - PW(k - 1))) {
          ltop = false;
        } else {
          ltop = thm - tkm >= tk - thk;
        }
        if (tkm >= thm || !ltop || PW(k) <= dp0cum_at(k, k) + kOnem) {
          if ((thk - tkm) <= kEpsil) {
            // p_hat0 keeps the previous level's 1479 value: the blocking
            // branch below reads it (1301-1330)
            p_hat = PW(k - 1) + dm1;
          } else {
            q = (thk - tk) / (thk - tkm);
            p_hat0 = PW(k) + q * (PW(k + 1) - PW(k));
            if (k == fixlay + 2) {
              p_hat = PW(k - 1) + dmax(p_hat0 - PW(k - 1), dm1);
            } else {
              p_hat = PW(k - 1) + cushn(p_hat0 - PW(k - 1), dm1);
            }
          }
          p_hat = dmin(p_hat, pbot);
          if (k <= fixlay + 2) {
            // do nothing
          } else if (p_hat >= PW(k) && PW(k - 1) > dp0cum_at(k - 1, k) + kTenm &&
                     (pbot - PW(k - 1) < P.thkbop || PW(k - 1) - PW(k - 2) > qqmx * dm2)) {
            if (k == fixlay + 3) {
              p_hat2 = PW(k - 2) + dmax(PW(k - 1) - p_hat + p_hat0 - PW(k - 2), dm2);
            } else {
              p_hat2 = PW(k - 2) + cushn(PW(k - 1) - p_hat + p_hat0 - PW(k - 2), dm2);
            }
            if (p_hat2 < PW(k - 1) - kOnemm) {
              PW(k - 1) = (1.0 - qm1) * PW(k - 1) + qm1 * dmax(p_hat2, 2.0 * PW(k - 1) - p_hat);
              p_hat = PW(k - 1) + cushn(p_hat0 - PW(k - 1), dm1);
            } else if (k <= fixlay + 3) {
              // do nothing
            } else if (PW(k - 2) > dp0cum_at(k - 2, k) + kTenm &&
                       (pbot - PW(k - 2) < P.thkbop || PW(k - 2) - PW(k - 3) > qqmx * dm3)) {
              if (k == fixlay + 4) {
                p_hat3 = PW(k - 3) + dmax(PW(k - 2) - p_hat + p_hat0 - PW(k - 3), dm3);
              } else {
                p_hat3 = PW(k - 3) + cushn(PW(k - 2) - p_hat + p_hat0 - PW(k - 3), dm3);
              }
              if (p_hat3 < PW(k - 2) - kOnemm) {
                PW(k - 2) =
                    (1.0 - qm2) * PW(k - 2) + qm2 * dmax(p_hat3, 2.0 * PW(k - 2) - PW(k - 1));
                p_hat2 = PW(k - 2) + cushn(PW(k - 1) - p_hat + p_hat0 - PW(k - 2), dm2);
                if (p_hat2 < PW(k - 1) - kOnemm) {
                  PW(k - 1) =
                      (1.0 - qm1) * PW(k - 1) + qm1 * dmax(p_hat2, 2.0 * PW(k - 1) - p_hat);
                  p_hat = PW(k - 1) + cushn(p_hat0 - PW(k - 1), dm1);
                }
              }
            }
          }
          if (p_hat < PW(k)) {
            PW(k) = (1.0 - q0) * PW(k) + q0 * p_hat;
          }
        }
      } else if (tk < thk - kEpsil) {
        // 1389-1475 layer too light: move the lower interface down
        if (PW(k + 1) < pbot) {
          const double thp = thw[2], tkp = tw[2];
          bool ltop;
          if (rthin * (PW(k + 1) - PW(k)) <= (PW(k + 2) - PW(k + 1))) {
            ltop = true;
          } else if (rthin * (PW(k + 2) - PW(k + 1)) <= (PW(k + 1) - PW(k))) {
            ltop = false;
          } else {
            ltop = thk - tk >= tkp - thp;
          }
          if (tkp <= thp || ltop || PW(k + 1) <= dp0cum_at(k + 1, k) + kOnem) {
            if ((tkp - thk) <= kEpsil) {
              p_hat = PW(k + 2);
            } else {
              q = (tk - thk) / (tkp - thk);
              p_hat = PW(k + 1) + q * (PW(k) - PW(k + 1));
            }
            const double p3 = (k + 3 <= K + 1) ? PW(k + 3) : pbot;
            if (p3 < pbot) {
              if (pbot - PW(k) > d0 + dp1) {
                p_hat = PW(k + 2) - cushn(PW(k + 2) - p_hat, dp1);
              }
              p_hat = PW(k) + dmax(p_hat - PW(k), d0);
              p_hat = dmin(p_hat, dmax(0.5 * (PW(k + 1) + PW(k + 2)), PW(k + 2) - dp1));
            } else {
              p_hat = dmin(PW(k + 2), p_hat);
            }
            if (p_hat > PW(k + 1)) {
              PW(k + 1) = (1.0 - qp1) * PW(k + 1) + qp1 * p_hat;
            }
          }
        }
      }
      // 1479-1490 layer above still too thin
      p_hat0 = dmin(PW(k - 1) + dm1, pbot);
      if (p_hat0 > PW(k)) {
        p_hat = (1.0 - qm1) * PW(k) + qm1 * p_hat0;
        PW(k) = dmin(p_hat, PW(k + 1));
      }
    }
#undef PW
    // p(k-2) is final now
    if (k >= 3) consume(k - 2, pw[1], pchk);
    // slide: level k+4 enters pw, k+3 enters tw/thw/lcm, k+2 enters dw/qw
    double r4, th3, ta3;
    if (AHEAD) {
      r4 = (k + 4 <= K + 1) ? nr4[0] : 0.0;
      th3 = (k + 3 <= K) ? (k + 3 > kp ? th_kp : nth3[0]) : 0.0;
      ta3 = (k + 3 <= K) ? nta3[0] : 0.0;
#pragma unroll
      for (int q = 0; q < AQ - 1; q++) {
        nr4[q] = nr4[q + 1];
        nth3[q] = nth3[q + 1];
        nta3[q] = nta3[q + 1];
      }
      nr4[AQ - 1] = pres[cl1(k + 4 + AQ)];
      nth3[AQ - 1] = th3d[clk(k + 3 + AQ)];
      nta3[AQ - 1] = theta[clk(k + 3 + AQ)];
    } else {
      r4 = raw_at(k + 4);
      th3 = (k + 3 <= K) ? th3d_eff(k + 3) : 0.0;
      ta3 = (k + 3 <= K) ? FK(theta, k + 3) : 0.0;
    }
    shl(pw, (k + 4 <= K + 1) ? with_floor(k + 4, r4) : 0.0);
    if (k + 3 <= K) lcm_set(k + 3, rw[2], r4, th3, ta3);
    shl(tw, th3);
    shl(thw, ta3);
    if (tf) {
      const double dn = (k + 2 <= K) ? dp0ij_tf(k + 2) : 0.0;
      shl(tfc, tfc[3] + dw[4]);  // dp0cum(k+2) = dp0cum(k+1) + dp0ij(k+1)
      shl(dw, dn);
      shl(qw, 1.0);
    } else {
      shl(qw, (k + 2 <= qfix) ? 1.0 : qh_next);
      double qh = 1.0;
      const double dn =
          (k + 2 <= K) ? dp0ij_qh(P, P.dp0k[k + 1], p_init(k + 2, rw[1]), P.dp0cum[k + 1], &qh)
                       : 0.0;
      qh_next = qh;
      shl(dw, dn);
    }
    shl(rw, r4);
  }
  // levels kk-1, kk, kk+1 (pw holds p(kk-2..kk+4) now)
  consume(K - 1, pw[1], K - 1 >= 5 ? pres[(int64_t)(K - 5) * S] : 0.0);
  consume(K, pw[2], K >= 5 ? pres[(int64_t)(K - 4) * S] : 0.0);
  consume(K + 1, pw[3], K + 1 >= 5 ? pres[(int64_t)(K - 3) * S] : 0.0);
  W.kp[c0] = kp | ((unsafe || !tame) ? kUnsafe : 0);
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
// Sources of the old interfaces pi(x) for walk_core: pi(1), then pi(x+1)
// from pi(x) (the walk asks for them in increasing x, and scans ahead).
// ---------------------------------------------------------------------
// raw(x) is the memory word that gives pi(x+1) from pi(x) through comb(),
// so a walk can issue it early.
// kRawThk: the walk takes the old thicknesses dprs(x) from thk(x) as
// stored instead of as differences of the interfaces.
struct PiArray {  // p-columns: the overturned interfaces left by regrid_stream
  static constexpr bool kRawThk = false;
  const double* pres;
  int64_t S;
  HYB_HD double thk(int) const { return 0.0; }
  HYB_HD double first() const { return pres[0]; }
  HYB_HD double raw(int x) const { return pres[(int64_t)x * S]; }
  HYB_HD double comb(double, double r) const { return r; }
  HYB_HD double nextfrom(int x, double pix) const { return comb(pix, raw(x)); }
};
struct PiSum {  // u/v points: running sums of the old dpu (hybgen.F90:171-195)
  static constexpr bool kRawThk = false;
  const double* dpu;
  int64_t S;
  HYB_HD double thk(int) const { return 0.0; }
  HYB_HD double first() const { return 0.0; }
  HYB_HD double raw(int x) const { return dpu[(int64_t)(x - 1) * S]; }
  HYB_HD double comb(double pix, double r) const { return pix + r; }
  HYB_HD double nextfrom(int x, double pix) const { return comb(pix, raw(x)); }
};
// hybgencj (hybgen.F90:1909-2274): the old interfaces are running sums of
// the old dp(m), kept in the workspace, and the old thicknesses are dp(m)
// itself (hybgencj passes dprs(k) = dp(k,m), not pres(k+1)-pres(k)).
struct PiRa {
  static constexpr bool kRawThk = true;
  const double* pres;  // the running sums, level 1 of the column
  const double* dpm;   // dp(:,m), level 1 of the column
  int64_t S;
  HYB_HD double first() const { return pres[0]; }
  HYB_HD double raw(int x) const { return pres[(int64_t)x * S]; }
  HYB_HD double comb(double, double r) const { return r; }
  HYB_HD double nextfrom(int x, double pix) const { return comb(pix, raw(x)); }
  HYB_HD double thk(int x) const { return dpm[(int64_t)(x - 1) * S]; }
};

// New interfaces prsf(k): p-columns keep them in the lane's LDS column
// (LaneRow); u/v points keep the new thicknesses there and sum them as
// the walk asks for prsf(1), prsf(2), ... (hybgen.F90:1806-1810).
struct PoSum {
  LaneRow d;
  double prv = 0.0, cur = 0.0;  // prsf(k-1), prsf(k) of the last request k
  int k = 1;
  HYB_HD double operator()(int kq) {
    if (kq == 1) {
      prv = 0.0;
      cur = 0.0;
      k = 1;
      return cur;
    }
    while (k < kq) {
      prv = cur;
      cur = cur + d(k);
      k++;
    }
    return cur;
  }
};

// The same prsf(k) without the LDS column: the new thicknesses are formed
// again from p at the two neighbouring p-points as the walk asks for them
// (the operations of the first pass of uv_stream, so the same bits).  The
// p levels are loaded A levels ahead of the request, unconditionally (the
// index clamped into the column), so the loads are issued without a
// branch and have the work of A requests to arrive.
template <int A>
struct PoRecomputeT {
  const double* p0;
  const double* p1;
  int64_t S;
  double depthu;
  int k1;                                // kk+1, the deepest level
  double prv = 0.0, cur = 0.0, b = 0.0;  // prsf(k-1), prsf(k), min(depthu, p-mean at k)
  double dl = 0.0;                       // the new thickness of layer k-1 (cur = prv + dl)
  double q0[A], q1[A];                   // p at levels k+1..k+A of both points, loaded ahead
  int k = 0;                             // 0 until the first request
  HYB_HD PoRecomputeT(const double* a0, const double* a1, int64_t s, double du, int kk)
      : p0(a0), p1(a1), S(s), depthu(du), k1(kk + 1) {
#pragma unroll
    for (int q = 0; q < A; q++) q0[q] = q1[q] = 0.0;
  }
  HYB_HD double operator()(int kq) {
    if (kq == 1 || k == 0) {
      prv = 0.0;
      cur = 0.0;
      dl = 0.0;
      b = dmin(depthu, 0.5 * (p0[0] + p1[0]));
      k = 1;
#pragma unroll
      for (int q = 0; q < A; q++) {
        const int x = 1 + q < k1 ? 1 + q : k1 - 1;
        q0[q] = p0[(int64_t)x * S];
        q1[q] = p1[(int64_t)x * S];
      }
    }
    while (k < kq) {
      const double a = dmin(depthu, 0.5 * (q0[0] + q1[0]));
      const double d = dmax(0.0, a - b);
      b = a;
      prv = cur;
      cur = cur + d;
      dl = d;
      k++;
#pragma unroll
      for (int q = 0; q < A - 1; q++) {
        q0[q] = q0[q + 1];
        q1[q] = q1[q + 1];
      }
      // level k+A-1, clamped into the column (values beyond it unused)
      const int x = k + A - 1 < k1 ? k + A - 1 : k1 - 1;
      q0[A - 1] = p0[(int64_t)x * S];
      q1[A - 1] = p1[(int64_t)x * S];
    }
    return cur;
  }
};
using PoRecompute = PoRecomputeT<1>;
template <class PO>
struct is_recompute : std::false_type {};
template <int A>
struct is_recompute<PoRecomputeT<A>> : std::true_type {};

// Output sinks.  emit(kc, v, certain, thick): v[f] is output kc of field
// f before the clamp; certain means it lies inside the range of the input
// levels read so far, so the clamp against the whole column cannot change
// it.  fixup() clamps the others once the column range is known.
// NT: outputs written with non-temporal stores (streamed past the caches)
template <int NF, bool NT = false>
struct FieldSink {  // T, S, tracers; th3d from T and S (hybgen.F90:1549-1569)
  double* const* fp;
  int nf;
  double* th3d;
  double thbase;
  int64_t S;
  uint64_t fix = 0;
  HYB_HD void emit(int kc, const double* v, bool certain, bool) {
#pragma unroll
    for (int f = 0; f < NF; f++)
      if (f < nf) put(fp[f] + (int64_t)(kc - 1) * S, v[f]);
    if (!certain) {
      fix |= 1ull << (kc - 1);
    } else if constexpr (NF >= 2) {  // th3d needs T and S in the same walk
      if (th3d) put(th3d + (int64_t)(kc - 1) * S, sig(v[0], v[1]) - thbase);
    }
  }
  HYB_HD static void put(double* a, double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    if (NT) {
      __builtin_nontemporal_store(x, a);
      return;
    }
#endif
    *a = x;
  }
  // clamp to [smin, smax] the outputs outside [lo, hi] (lo = smin and
  // hi = smax except in the mixed form, see walk_core)
  HYB_HD void fixup(const double* lo, const double* hi, const double* smin,
                    const double* smax) {
    while (fix) {
      const int k = __builtin_ctzll(fix) + 1;
      fix &= fix - 1;
      double v[NF];
#pragma unroll
      for (int f = 0; f < NF; f++) {
        v[f] = 0.0;
        if (f < nf) {
          double x = fp[f][(int64_t)(k - 1) * S];
          x = (x < lo[f]) ? smin[f] : x;
          x = (x > hi[f]) ? smax[f] : x;
          fp[f][(int64_t)(k - 1) * S] = x;
          v[f] = x;
        }
      }
      if constexpr (NF >= 2)
        if (th3d) th3d[(int64_t)(k - 1) * S] = sig(v[0], v[1]) - thbase;
    }
  }
};
template <class PO>
struct UvSink {  // u or v with the thin-layer rule (hybgen.F90:1858-1865)
  double* uu;
  int64_t S;
  double ub;    // ubavg or vbavg at time level nl
  double ptop;  // prsf(kk+1)
  const PO* po;
  uint64_t fix = 0;
  // With PoRecompute, the new dpu of layer kc (dpudpv.F90:92-103) is
  // written as output kc is: the walk asks for prsf(kc+1) just before it
  // emits output kc, so the thickness that formed it is po->dl.  The walk
  // has read the old dpu of layer kc by then: outputs end no higher than
  // two layers above their index (the in-place test of uv_stream), and the
  // old interfaces are summed PD+3 layers ahead of the walk.
  double* dpn = nullptr;
  HYB_HD void emit(int kc, const double* v, bool certain, bool thick) {
    if (thick || po->prv <= ptop - kOnemm) {
      uu[(int64_t)(kc - 1) * S] = v[0];
      if (!certain) fix |= 1ull << (kc - 1);
    } else {
      uu[(int64_t)(kc - 1) * S] = -ub;
    }
    if constexpr (is_recompute<PO>::value) dpn[(int64_t)(kc - 1) * S] = po->dl;
  }
  HYB_HD void fixup(const double* lo, const double* hi, const double* smin,
                    const double* smax) {
    while (fix) {
      const int k = __builtin_ctzll(fix) + 1;
      fix &= fix - 1;
      double x = uu[(int64_t)(k - 1) * S];
      x = (x < lo[0]) ? smin[0] : x;
      x = (x > hi[0]) ? smax[0] : x;
      uu[(int64_t)(k - 1) * S] = x;
    }
  }
};

// ---------------------------------------------------------------------
// Reconstruction coefficients of old layer l for NF fields, in arithmetic
// type R (double: the reference's operations; float: the mixed form).
// Windows s0..s3 = s(l-1..l+2); dprs_* / dpi_* are the thicknesses of
// layers l-1, l, l+1, l+2 (dpi = max(dprs, thin)).  Carried state in st.
// ---------------------------------------------------------------------
template <class R, int NF>
struct CoefState {
  R asc[NF], alc[NF];                     // PPM: as(l), raw al(l)
  R wc1[NF], wc2[NF], wz1[NF], wz2[NF];   // WENO first sweep of layer l
  R zw3l[NF];                             // WENO interface value zw3(l)
  R qdpjp_l, wq_l;                        // PPM qdpjp(l); WENO 1/(dp(l)+dp(l+1))
};

template <class R>
HYB_HD R gmax(R a, R b) { return a > b ? a : b; }
template <class R>
HYB_HD R gmin(R a, R b) { return a < b ? a : b; }
template <class R>
HYB_HD R gmin3(R a, R b, R c) { return gmin(gmin(a, b), c); }
template <class R>
HYB_HD R gabs(R a) { return a < R(0) ? -a : a; }
template <class R>
HYB_HD R gsign(R a, R b) {  // Fortran sign(a,b)
  R x = gabs(a);
  return signbit(b) ? -x : x;
}

// FD (R = double only): the reciprocals and the quotients of thicknesses
// through rcp64/qrcp/qdiv_pos, bitwise the IEEE ones; every divisor here
// is a sum of thicknesses >= thin, which holds for columns whose
// interfaces are in order and bounded (the ones regrid_stream leaves to
// the walk).
template <int METHOD, int NF, class R, bool FD, class LCMB>
__attribute__((always_inline)) HYB_HD void coef_layer(
    int l, int K, int nf, const LCMB& lcmb, bool lc_l, R thin, R dsmll, R dprs_m, R dprs_0,
    R dprs_p, R dpi_m, R dpi_0, R dpi_p, R dpi_p2, const R* s0, const R* s1, const R* s2,
    const R* s3, CoefState<R, NF>& st, R* c1l, R* c2l, R* c3l) {
  auto RCP = [](R x) -> R {
    if constexpr (FD) return qrcp(rcp64(x));
    else return grcp(x);
  };
  auto DIV = [](R a, R b) -> R {
    if constexpr (FD) return qdiv_pos(a, rcp64(b));
    else return gdiv(a, b);
  };
  if (METHOD == 1) {
    // hybgen_plm_coefs (2393-2458) on dprs
    const bool flat = (l == 1 || l == K || lc_l || dprs_0 <= thin);
    R qcen = R(0);
    if (!flat) qcen = DIV(dprs_0, dprs_0 + R(0.5) * (dprs_m + dprs_p));
#pragma unroll
    for (int f = 0; f < NF; f++) {
      c1l[f] = R(0);
      c2l[f] = R(0);
      c3l[f] = R(0);
      if (f < nf && !flat) {
        R ztop = R(2) * (s1[f] - s0[f]);
        R zbot = R(2) * (s2[f] - s1[f]);
        R zcen = qcen * (s2[f] - s0[f]);
        if (ztop * zbot > R(0)) c1l[f] = gsign(gmin3(gabs(zcen), gabs(zbot), gabs(ztop)), zbot);
      }
    }
  } else if (METHOD == 2) {
    // hybgen_ppm_coefs (2587-2706) on dpi: as(l+1), al(l+1), then the
    // limited edge values of layer l
    const int jn = l + 1;
    const bool do_as = jn >= 2 && jn <= K - 1 && !(lcmb(jn) || dpi_p <= thin);
    const bool do_al = jn >= 3 && jn <= K - 1;
    R qdpjp_n = R(0), dpq3 = R(0), dp2jp_l = R(0), dpj2p_n = R(0);
    R dpjp_lm = R(0), dpjp_n = R(0), qdp2jp_l = R(0), qdpj2p_l = R(0), qdp4 = R(0);
    if (jn <= K - 1) {
      const R dpjp_l = dpi_0 + dpi_p;
      dpjp_n = dpi_p + dpi_p2;
      qdpjp_n = RCP(dpjp_n);  // qdpjp(l+1)
      dp2jp_l = dpi_0 + dpjp_l;
      dpj2p_n = dpjp_n + dpi_p2;
      dpq3 = DIV(dpi_p, dpi_0 + dpjp_n);
      if (do_al) {
        dpjp_lm = dpi_m + dpi_0;
        const R dpj2p_l = dpjp_l + dpi_p;
        qdp2jp_l = RCP(dp2jp_l);
        qdpj2p_l = RCP(dpj2p_l);
        qdp4 = RCP(dpjp_lm + dpjp_n);
      }
    }
    const bool lim = l >= 2 && l <= K - 1;
    const bool lim_flat = lim && (lc_l || dpi_0 <= thin);
#pragma unroll
    for (int f = 0; f < NF; f++) {
      c1l[f] = R(0);
      c2l[f] = R(0);
      c3l[f] = R(0);
      if (f < nf) {
        // the per-lane choices below are selects, not branches: the
        // values are those of the reference's taken branch
        R as_n = R(0), al_n = R(0);
        {
          const R slj = s2[f] - s1[f];
          const R srj = s3[f] - s2[f];
          const R scj = dpq3 * (dp2jp_l * srj * qdpjp_n + dpj2p_n * slj * st.qdpjp_l);
          const R asv = gsign(gmin3(gabs(R(2) * slj), gabs(scj), gabs(R(2) * srj)), scj);
          as_n = (do_as && slj * srj > R(0)) ? asv : R(0);
        }
        if (jn == K) {
          al_n = s2[f];  // al(kk) = s(kk)
        } else if (jn == 2) {
          al_n = s1[f];  // al(2) = s(1)
        } else if (do_al) {
          al_n = s1[f] + dpi_0 * (s2[f] - s1[f]) * st.qdpjp_l +
                 qdp4 * (R(2) * dpi_p * dpi_0 * st.qdpjp_l * (s2[f] - s1[f]) *
                             (dpjp_lm * qdp2jp_l - dpjp_n * qdpj2p_l) -
                         dpi_0 * as_n * dpjp_lm * qdp2jp_l + dpi_p * st.asc[f] * dpjp_n * qdpj2p_l);
        }
        R al = st.alc[f], ar;
        if (l == K - 1) ar = s2[f];  // ar(kk-1) = s(kk)
        else if (l == K) ar = s1[f];  // ar(kk) = s(kk)
        else if (l == 1) ar = s1[f];  // ar(1) = s(1)
        else ar = al_n;               // ar(l) = al(l+1)
        if (lim) {
          const bool flat = lim_flat || (s2[f] - s1[f]) * (s1[f] - s0[f]) <= R(0);
          const R da = ar - al;
          const R a6 = R(6) * s1[f] - R(3) * (al + ar);
          const bool peak_r = da * a6 > da * da;
          const bool peak_l = da * a6 < -(da * da);
          const R al_r = R(3) * s1[f] - R(2) * ar;
          const R ar_l = R(3) * s1[f] - R(2) * al;
          const R al2 = flat ? s1[f] : (peak_r ? al_r : al);
          const R ar2 = flat ? s1[f] : ((!peak_r && peak_l) ? ar_l : ar);
          al = al2;
          ar = ar2;
        }
        const bool pcm = !(al != ar);
        c1l[f] = al;
        c2l[f] = pcm ? R(0) : ar - al;
        c3l[f] = pcm ? R(0) : R(6) * s1[f] - R(3) * (al + ar);
        st.asc[f] = as_n;
        st.alc[f] = al_n;
      }
    }
    st.qdpjp_l = qdpjp_n;
  } else if (METHOD == 3) {
    // hybgen_weno_coefs (2845-2968) on dpi: first sweep of layer l+1,
    // interface value zw3(l+1), second sweep of layer l
    const int jn = l + 1;
    const bool first_full = jn >= 2 && jn <= K - 1 && !(lcmb(jn) || dpi_p <= thin);
    R wq_n = R(0), qmjp = R(0), dpjm2jp = R(0);
    if (jn <= K - 1) wq_n = RCP(dpi_p + dpi_p2);  // qdpjm(l+2)
    if (first_full) {
      qmjp = RCP(dpi_0 + dpi_p + dpi_p2);
      dpjm2jp = dpi_0 + R(2) * dpi_p + dpi_p2;
    }
    const bool second = l >= 2 && l <= K - 1 && !(lc_l || dpi_0 <= thin);
#pragma unroll
    for (int f = 0; f < NF; f++) {
      c1l[f] = R(0);
      c2l[f] = R(0);
      c3l[f] = R(0);
      if (f < nf) {
        R n1 = s2[f], n2 = s2[f], nz1 = R(0), nz2 = R(0);
        if (first_full) {
          const R zw3_j = st.wq_l * (s2[f] - s1[f]);
          const R zw3_jp = wq_n * (s3[f] - s2[f]);
          R q001 = dpi_p * zw3_jp;
          R q002 = dpi_p * zw3_j;
          if (q001 * q002 < R(0)) {
            q001 = R(0);
            q002 = R(0);
          }
          const R q01 = dpjm2jp * zw3_jp;
          const R q02 = dpjm2jp * zw3_j;
          if (gabs(q001) > gabs(q02)) q001 = q02;
          if (gabs(q002) > gabs(q01)) q002 = q01;
          const R qq = (q001 - q002) * qmjp;
          q001 = q001 - qq * dpi_p2;
          q002 = q002 + qq * dpi_0;
          n2 = s2[f] + q001;
          n1 = s2[f] - q002;
          const R a = R(2) * q001 - q002, b = R(2) * q002 - q001;
          nz1 = a * a;
          nz2 = b * b;
        }
        R zw3n = R(0);
        if (jn <= K) {
          const R q002 = gmax(st.wz2[f], dsmll);
          const R q001 = gmax(nz1, dsmll);
          zw3n = gdiv(q001 * st.wc2[f] + q002 * n1, q001 + q002);
        }
        c1l[f] = st.wc1[f];
        c2l[f] = st.wc2[f];
        if (second) {
          R q01 = zw3n - s1[f];
          R q02 = s1[f] - st.zw3l[f];
          const R q001 = R(2) * q01;
          const R q002 = R(2) * q02;
          if (q01 * q02 < R(0)) {
            q01 = R(0);
            q02 = R(0);
          } else if (gabs(q01) > gabs(q002)) {
            q01 = q002;
          } else if (gabs(q02) > gabs(q001)) {
            q02 = q001;
          }
          c1l[f] = s1[f] - q02;
          c2l[f] = s1[f] + q01;
        }
        st.zw3l[f] = zw3n;
        st.wc1[f] = n1;
        st.wc2[f] = n2;
        st.wz1[f] = nz1;
        st.wz2[f] = nz2;
      }
    }
    st.wq_l = wq_n;
  } else {
#pragma unroll
    for (int f = 0; f < NF; f++) c1l[f] = c2l[f] = c3l[f] = R(0);
  }

}

// ---------------------------------------------------------------------
// remap of up to NF fields of one column in one walk over the old layers
// (hybgen_{pcm,plm,ppm,weno}_{coefs,remap}).  fp[f] points at level 1 of
// field f (stride S); levels > kp take level kp's value (the massless
// fill); zx = pi(K+1).  PD levels of every field (and of the old
// interfaces) are loaded ahead of the window, to cover memory latency.
// K old layers, KO new ones (KO = 0: K).  CDPI: the PLM slopes are taken
// on dpi = max(dprs, thin) instead of dprs (mod_stokes.F90:683-684 passes
// dpi to hybgen_plm_coefs; hybgen.F90:1535 passes dprs).  fp[f][i] is a
// double, or a value that converts to one when the level enters the window
// (a blend of two stored words, so the loads stay outstanding until then).
// ---------------------------------------------------------------------
template <int METHOD, int NF, int PD, bool FAST = false, bool CDPI = false, bool FDIV = false,
          class PI, class PO, class SINK, class FP>
HYB_HDI void walk_core(int K, double zx, const PI& pi, PO& po, SINK& sink, const FP& fp,
                      int nf, int kp, uint64_t lcm, int64_t S, int KO = 0) {
  constexpr int PQ = PD > 0 ? PD : 1;
  const int KN = KO > 0 ? KO : K;
  using RawT = typename std::decay<decltype(fp[0][0])>::type;
  using CR = typename std::conditional<FAST, float, double>::type;
  const double thin = kDpthin;
  const double dsmll = 1.0e-8;
  // the FP64 quotients through rcp64/qdiv (hyb_math.h): the caller passes
  // only columns whose interfaces are in order and bounded, so every
  // divisor is in [thin, 2^63] and every thickness or depth difference is 0
  // or in [2^-112, 2^61]
  constexpr bool FD64 = FDIV && !FAST;
  auto lcmb = [&](int l) -> bool { return (lcm >> (l - 1)) & 1ull; };

  // per-field state
  double s0[NF], s1[NF], s2[NF], s3[NF];  // s(l-1), s(l), s(l+1), s(l+2)
  double smin[NF], smax[NF], skp[NF], sfirst[NF];
  double acc[NF], ov[NF], prev[NF];
  // FAST (the mixed form): each output's sum of piece integrals is
  // compensated (TwoSum: acc + accc), as the sums that set conservation
  double accc[NF];
  CoefState<CR, NF> cst;                       // reconstruction state
  RawT sq[NF][PQ];                             // s(l+3..l+2+PD), loaded ahead
  double prq[PQ];                              // raw(l+3..l+2+PD) of pi
  // the old thicknesses dprs(l-1..l+2) and dprs(l+3..l+2+PD) loaded ahead,
  // when the source gives them as stored (kRawThk)
  constexpr bool RAWT = PI::kRawThk;
  double tv0 = 0.0, tv1 = 0.0, tv2 = 0.0, tv3 = 0.0, tq[PQ];
  int xl = 4;                                  // next level to load
  // s(x) with the massless fill (972-996) and the running min/max
  // (e.g. 2727-2731: si_min/si_max over k = 1..kk in order)
  // fold(f, x, r): level x enters the column in order, r being its stored
  // value (ignored below kp); idempotent, so a level may be folded again
  auto fold = [&](int f, int x, double r) -> double {
    const double v = (x <= kp) ? r : skp[f];
    if (x == kp) skp[f] = v;
    if (x == 1) {
      smin[f] = v;
      smax[f] = v;
    } else {
      smin[f] = dmin(smin[f], v);
      smax[f] = dmax(smax[f], v);
    }
    return v;
  };
  auto load = [&](int f, int x) -> double {
    return fold(f, x, (x <= kp) ? static_cast<double>(fp[f][(int64_t)(x - 1) * S]) : 0.0);
  };
  // the stored value of level min(x, kp), loaded without a use: with
  // PD > 0 it is folded PD layer iterations later, so the load has that
  // long to arrive (its value is discarded below kp)
  auto raw = [&](int f, int x) -> RawT {
    const int xc = x < kp ? x : kp;
    return fp[f][(int64_t)(xc - 1) * S];
  };
#pragma unroll
  for (int f = 0; f < NF; f++) {
    s0[f] = s1[f] = s2[f] = s3[f] = 0.0;
    smin[f] = smax[f] = skp[f] = sfirst[f] = 0.0;
    acc[f] = ov[f] = prev[f] = accc[f] = 0.0;
    cst.asc[f] = cst.alc[f] = 0.0;
    cst.wc1[f] = cst.wc2[f] = cst.wz1[f] = cst.wz2[f] = cst.zw3l[f] = 0.0;
    if (f < nf) {
      s1[f] = load(f, 1);
      if (K >= 2) s2[f] = load(f, 2);
      if (K >= 3) s3[f] = load(f, 3);
      sfirst[f] = s1[f];
#pragma unroll
      for (int q = 0; q < PD; q++) sq[f][q] = (4 + q <= K) ? raw(f, 4 + q) : RawT{};
      if (METHOD == 2) cst.alc[f] = (CR)s1[f];  // al(1) = s(1)
      if (METHOD == 3) {                        // first sweep of layer 1
        cst.wc1[f] = (CR)s1[f];
        cst.wc2[f] = (CR)s1[f];
      }
    }
  }
  // interfaces pi(l-1..l+3) at l = 1
  double pv0 = 0.0, pv1 = pi.first(), pv2 = 0.0, pv3 = 0.0, pv4 = 0.0;
  if (K + 1 >= 2) pv2 = pi.nextfrom(1, pv1);
  if (K + 1 >= 3) pv3 = pi.nextfrom(2, pv2);
  if (K + 1 >= 4) pv4 = pi.nextfrom(3, pv3);
#pragma unroll
  for (int q = 0; q < PD; q++) prq[q] = (4 + q <= K) ? pi.raw(4 + q) : 0.0;
  if constexpr (RAWT) {
    tv1 = pi.thk(1);
    if (K >= 2) tv2 = pi.thk(2);
    if (K >= 3) tv3 = pi.thk(3);
#pragma unroll
    for (int q = 0; q < PQ; q++) tq[q] = (4 + q <= K) ? pi.thk(4 + q) : 0.0;
  }
  xl = 4 + PD;
  // layers thicker than thin, known up to l+2 (the u/v store rule)
  uint64_t thick = 0;
  if (dmax(pv2 - pv1, thin) > thin) thick |= 1ull;
  if (K >= 2 && dmax(pv3 - pv2, thin) > thin) thick |= 2ull;
  if (K >= 3 && dmax(pv4 - pv3, thin) > thin) thick |= 4ull;
  // carried metrics: PPM qdpjp(l); WENO 1/(dp(l)+dp(l+1))
  cst.qdpjp_l = 0.0;
  cst.wq_l = 0.0;
  if (K >= 2) {
    const CR d1 = gmax((CR)(RAWT ? tv1 : pv2 - pv1), (CR)thin),
             d2 = gmax((CR)(RAWT ? tv2 : pv3 - pv2), (CR)thin);
    if constexpr (FD64) {
      cst.qdpjp_l = qrcp(rcp64(d1 + d2));
      cst.wq_l = cst.qdpjp_l;
    } else {
      cst.qdpjp_l = grcp(d1 + d2);
      cst.wq_l = grcp(d1 + d2);
    }
  }

  // walk state (2729-2735)
  const double zb0 = dmax(po(1), pv1);
  int lt_c = 1;
  {
    double pc = pv1;
    while (lt_c < K) {
      const double pn = pi.nextfrom(lt_c, pc);
      if (!(pn < zb0)) break;
      lt_c++;
      pc = pn;
    }
  }
  double zt_c = zb0;
  int kc = 1;
  bool accon = false, acc_thin = false;
  int lb_act = 0;
  double zt_act = 0.0, zb_act = 0.0;

  // The clamp to the column range (e.g. 2836-2837).  The mixed form lets
  // values within a slack of 2^-20 of the range's magnitude stand (in the
  // final fixup): its float reconstruction can overshoot an edge that the
  // FP64 one meets exactly, and clamping those by rounding error would
  // break the conservation that the remap otherwise keeps.
  auto slack = [&](int f) -> double {
    return FAST ? 0x1p-20 * dmax(dabs(smin[f]), dabs(smax[f])) : 0.0;
  };
  // fold the levels below the window of layer l that are not folded yet:
  // those loaded ahead, then the rest (idempotent)
  auto fold_rest = [&](int l) {
#pragma unroll
    for (int q = 0; q < PD; q++) {  // levels loaded ahead, not folded yet
      const int x = l + 3 + q;
      if (x <= K && x < xl) {
#pragma unroll
        for (int f = 0; f < NF; f++)
          if (f < nf) (void)fold(f, x, sq[f][q]);
      }
    }
    for (int x = xl; x <= K; x++) {
#pragma unroll
      for (int f = 0; f < NF; f++)
        if (f < nf) (void)load(f, x);
    }
    xl = K + 1;
  };
  // output kc complete: v[f] before the clamp
  auto emit = [&](const double* v) {
    bool certain = true;
#pragma unroll
    for (int f = 0; f < NF; f++)
      if (f < nf) {
        prev[f] = v[f];
        if (!(v[f] >= smin[f] && v[f] <= smax[f])) certain = false;
      }
    sink.emit(kc, v, certain, (thick >> (kc - 1)) & 1ull);
    kc = kc + 1;
  };

#pragma unroll 1
  for (int l = 1; l <= K; l++) {
    const double dprs_m = RAWT ? tv0 : pv1 - pv0, dprs_0 = RAWT ? tv1 : pv2 - pv1,
                 dprs_p = RAWT ? tv2 : pv3 - pv2, dprs_p2 = RAWT ? tv3 : pv4 - pv3;
    const double dpi_m = dmax(dprs_m, thin), dpi_0 = dmax(dprs_0, thin),
                 dpi_p = dmax(dprs_p, thin), dpi_p2 = dmax(dprs_p2, thin);
    const bool lc_l = lcmb(l);
    // ------------------------------------------ coefficients of layer l
    double c1l[NF], c2l[NF], c3l[NF];
    if constexpr (!FAST) {
      coef_layer<METHOD, NF, double, FD64>(l, K, nf, lcmb, lc_l, thin, dsmll, CDPI ? dpi_m : dprs_m,
                                     CDPI ? dpi_0 : dprs_0, CDPI ? dpi_p : dprs_p, dpi_m, dpi_0,
                                     dpi_p, dpi_p2, s0, s1, s2, s3, cst, c1l, c2l, c3l);
    } else {
      // mixed mode: the reconstruction in float; the polynomial handed to
      // the FP64 piece integrals keeps the layer mean s(l) (PPM: c3 from
      // s(l) in FP64), so the remap still conserves to FP64 roundoff
      const float thf = (float)thin;
      const float fm = (float)dprs_m, f0 = (float)dprs_0, fq = (float)dprs_p,
                  fq2 = (float)dprs_p2;
      float sf0[NF], sf1[NF], sf2[NF], sf3[NF], g1[NF], g2[NF], g3[NF];
#pragma unroll
      for (int f = 0; f < NF; f++) {
        sf0[f] = (float)s0[f];
        sf1[f] = (float)s1[f];
        sf2[f] = (float)s2[f];
        sf3[f] = (float)s3[f];
      }
      coef_layer<METHOD, NF, float, false>(l, K, nf, lcmb, lc_l, thf, 1.0e-8f, fm, f0, fq,
                                    gmax(fm, thf), gmax(f0, thf), gmax(fq, thf), gmax(fq2, thf),
                                    sf0, sf1, sf2, sf3, cst, g1, g2, g3);
#pragma unroll
      for (int f = 0; f < NF; f++) {
        c1l[f] = g1[f];
        c2l[f] = g2[f];
        c3l[f] = g3[f];
        if (METHOD == 2) {
          if (g2[f] == 0.0f && g3[f] == 0.0f) {
            // flat: exactly s(l).  Edge values that coincide in float
            // without the limiter setting them (rare in FP64, not in
            // float for a nearly uniform field) are taken as flat too,
            // so every layer keeps its mean
            c1l[f] = s1[f];
          } else {
            c3l[f] = 6.0 * s1[f] - 3.0 * (c1l[f] + (c1l[f] + c2l[f]));
          }
        }
      }
    }

    const double pil = pv1, pil1 = pv2, dprsl = dprs_0;
    // x of a depth in layer l; the mixed mode multiplies by a reciprocal
    const double rdl = FAST ? frcp64(dmax(dprsl, thin)) : 0.0;
    // FD64: one reciprocal of the layer's thickness for every x of the layer
    const Rcp64 rl = FD64 ? rcp64(dmax(dprsl, thin)) : Rcp64{0.0, 0.0};
    auto xof = [&](double z) -> double {
      if constexpr (FD64) return qdiv(z - pil, rl);
      return FAST ? (z - pil) * rdl : (z - pil) / dmax(dprsl, thin);
    };
    auto third = [](double x) -> double {
      if constexpr (FD64) return div3(x);
      return FAST ? x * (1.0 / 3.0) : x / 3.0;
    };
    // hi + lo += x exactly (Knuth's TwoSum)
    auto tsum = [](double& hi, double& lo, double x) {
      const double t = hi + x;
      const double bp = t - hi;
      lo = lo + ((hi - (t - bp)) + (x - bp));
      hi = t;
    };
    // Every output still to come ends in this layer when it reaches the
    // bottom, and the walk stops here: fold the levels not read yet into
    // the column range now, before any of them is overwritten.
    if (l < K && pil1 >= zx) fold_rest(l);
    // ------------------------------------------ output in progress
    if (accon) {
      if (lb_act == l) {
        double v[NF];
        if (acc_thin) {
#pragma unroll
          for (int f = 0; f < NF; f++) v[f] = (kc != 1) ? prev[f] : sfirst[f];
        } else {
          const double xb = xof(zb_act);
          const double dpb = zb_act - pil;
          double qb0 = 0.0, qb1 = 0.0, qb2 = 0.0;
          if (METHOD == 1) {
            qb0 = xb;
            qb1 = xb * xb * 0.5;
          } else if (METHOD == 2) {
            qb0 = xb;
            qb1 = xb * xb * 0.5;
            qb2 = third(xb * xb * xb);
          } else if (METHOD == 3) {
            qb1 = (xb - 1.0) * (xb - 1.0);
            qb2 = qb1 - 1.0 + xb;
            qb0 = 1.0 - qb1 - qb2;
          }
          const double dz = zb_act - zt_act;
          const double rdz = FAST ? frcp64(dz) : 0.0;
          const Rcp64 rz = FD64 ? rcp64(dz) : Rcp64{0.0, 0.0};
#pragma unroll
          for (int f = 0; f < NF; f++) {
            v[f] = 0.0;
            if (f < nf) {
              const double o = ov[f];
              double t;  // the piece of layer l
              if (METHOD == 0) {
                t = dpb * (s1[f] - o);
              } else if (METHOD == 1) {
                const double cc0 = s1[f] - o - 0.5 * c1l[f];
                t = dprsl * (cc0 * qb0 + c1l[f] * qb1);
              } else if (METHOD == 2) {
                t = dprsl * ((c1l[f] - o) * qb0 + (c2l[f] + c3l[f]) * qb1 - c3l[f] * qb2);
              } else {
                t = dpb * (qb0 * (s1[f] - o) + qb1 * (c1l[f] - o) + qb2 * (c2l[f] - o));
              }
              if constexpr (FAST) {
                tsum(acc[f], accc[f], t);
                v[f] = o + (acc[f] + accc[f]) * rdz;
              } else {
                const double a = acc[f] + t;
                if constexpr (FD64) v[f] = o + qdiv_any(a, rz);
                else v[f] = o + a / dz;
              }
            }
          }
        }
        emit(v);
        lt_c = l;
        zt_c = zb_act;
        accon = false;
      } else if (!acc_thin) {
#pragma unroll
        for (int f = 0; f < NF; f++) {
          if (FAST && f < nf) tsum(acc[f], accc[f], dprsl * (s1[f] - ov[f]));
          else if (f < nf) acc[f] = acc[f] + dprsl * (s1[f] - ov[f]);
        }
      }
    }
    // ------------------------------------------ outputs starting in layer l
    if (!accon && lt_c == l) {
      while (kc <= KN) {
        const double zt = zt_c;
        const double zb = dmin(po(kc + 1), zx);
        const bool ends_here = (l == K) || (pil1 >= zb);
        const bool thn = (zb - zt <= thin) || (zt >= zx);
        if (ends_here) {
          double v[NF];
          if (thn) {
#pragma unroll
            for (int f = 0; f < NF; f++) v[f] = (kc != 1) ? prev[f] : sfirst[f];
          } else if (METHOD == 0) {
#pragma unroll
            for (int f = 0; f < NF; f++) v[f] = s1[f];
          } else {
            const double xt = xof(zt);
            const double xb = xof(zb);
            const double dz = zb - zt;
            const double rdz = FAST ? frcp64(dz) : 0.0;
            const Rcp64 rz = FD64 ? rcp64(dz) : Rcp64{0.0, 0.0};
            double qt0 = 0.0, qt1 = 0.0, qt2 = 0.0;
            if (METHOD == 1) {
              qt1 = (xb * xb - xt * xt - (xb - xt)) * 0.5;
            } else if (METHOD == 2) {
              qt0 = (xb - xt);
              qt1 = (xb * xb - xt * xt) * 0.5;
              qt2 = third(xb * xb * xb - xt * xt * xt);
            } else {
              qt1 = xb * xb + xt * xt + xb * xt + 1.0 - 2.0 * (xb + xt);
              qt2 = qt1 - 1.0 + (xb + xt);
            }
#pragma unroll
            for (int f = 0; f < NF; f++) {
              v[f] = 0.0;
              if (f < nf) {
                if (METHOD == 1) {
                  const double sz = dprsl * (c1l[f] * qt1);
                  if constexpr (FD64) v[f] = s1[f] + qdiv_any(sz, rz);
                  else v[f] = FAST ? s1[f] + sz * rdz : s1[f] + sz / dz;
                } else if (METHOD == 2) {
                  const double oo = s1[f];
                  const double sz =
                      dprsl * ((c1l[f] - oo) * qt0 + (c2l[f] + c3l[f]) * qt1 - c3l[f] * qt2);
                  if constexpr (FD64) v[f] = oo + qdiv_any(sz, rz);
                  else v[f] = FAST ? oo + sz * rdz : oo + sz / dz;
                } else {
                  const double oo = s1[f];
                  const double sz = qt1 * (c1l[f] - oo) + qt2 * (c2l[f] - oo);
                  v[f] = oo + sz;
                }
              }
            }
          }
          zt_c = zb;
          emit(v);
        } else {
          // output kc starts in layer l and ends in the first deeper layer
          // lb with pi(lb+1) >= zb (or kk)
          int lb;
          if (l + 1 == K || pv3 >= zb) {
            lb = l + 1;
          } else if (l + 2 == K || pv4 >= zb) {
            lb = l + 2;
          } else {
            lb = l + 3;
            double pc = pv4;  // pi(l+3)
            while (lb < K) {
              const double pn = pi.nextfrom(lb, pc);
              if (!(pn < zb)) break;
              lb++;
              pc = pn;
            }
          }
          lb_act = lb;
          zt_act = zt;
          zb_act = zb;
          acc_thin = thn;
          accon = true;
          if (!thn) {
            const int mid = (l + lb) / 2;
            const double xt = xof(zt);
            const double dpt = pil1 - zt;
            double qt0 = 0.0, qt1 = 0.0, qt2 = 0.0;
            if (METHOD == 1) {
              qt0 = (1.0 - xt);
              qt1 = (1.0 - xt * xt) * 0.5;
            } else if (METHOD == 2) {
              qt0 = (1.0 - xt);
              qt1 = (1.0 - xt * xt) * 0.5;
              qt2 = third(1.0 - xt * xt * xt);
            } else if (METHOD == 3) {
              qt1 = xt * (xt - 1.0);
              qt2 = qt1 + xt;
              qt0 = 1.0 - qt1 - qt2;
            }
#pragma unroll
            for (int f = 0; f < NF; f++) {
              if (f < nf) {
                // o = s(mid), with the fill; levels > l+2 are not written yet
                double o;
                if (mid == l) o = s1[f];
                else if (mid == l + 1) o = s2[f];
                else if (mid == l + 2) o = s3[f];
                else if (kp <= l + 2) o = skp[f];
                else o = static_cast<double>(fp[f][(int64_t)((mid < kp ? mid : kp) - 1) * S]);
                ov[f] = o;
                accc[f] = 0.0;
                if (METHOD == 0) {
                  acc[f] = dpt * (s1[f] - o);
                } else if (METHOD == 1) {
                  const double cc0 = s1[f] - o - 0.5 * c1l[f];
                  acc[f] = dprsl * (cc0 * qt0 + c1l[f] * qt1);
                } else if (METHOD == 2) {
                  acc[f] = dprsl * ((c1l[f] - o) * qt0 + (c2l[f] + c3l[f]) * qt1 - c3l[f] * qt2);
                } else {
                  acc[f] = dpt * (qt0 * (s1[f] - o) + qt1 * (c1l[f] - o) + qt2 * (c2l[f] - o));
                }
              }
            }
          }
          break;
        }
      }
    }
    if (kc > KN) {
      // every output is out; with fewer new layers than old ones the
      // levels below may not be read yet, and the range covers them all
      if (KO > 0) fold_rest(l);
      break;
    }
    // ------------------------------------------ slide to layer l+1
    // The slide's look-ahead loads issue at raised wave priority, ahead of
    // the arithmetic of the other waves on the SIMD (values unaffected).
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_s_setprio(2);
#endif
    const double top = pv4;
    pv0 = pv1;
    pv1 = pv2;
    pv2 = pv3;
    pv3 = top;
    if (PD == 0) {
      pv4 = (l + 4 <= K + 1) ? pi.nextfrom(l + 3, top) : 0.0;
    } else {
      pv4 = (l + 4 <= K + 1) ? pi.comb(top, prq[0]) : 0.0;
#pragma unroll
      for (int q = 0; q < PD - 1; q++) prq[q] = prq[q + 1];
      prq[PQ - 1] = (l + 3 + PD <= K) ? pi.raw(l + 3 + PD) : 0.0;
    }
    if (l + 3 <= K && dmax(pv4 - pv3, thin) > thin) thick |= 1ull << (l + 2);
    if constexpr (RAWT) {
      tv0 = tv1;
      tv1 = tv2;
      tv2 = tv3;
      tv3 = (l + 3 <= K) ? tq[0] : 0.0;
#pragma unroll
      for (int q = 0; q < PQ - 1; q++) tq[q] = tq[q + 1];
      tq[PQ - 1] = (l + 3 + PQ <= K) ? pi.thk(l + 3 + PQ) : 0.0;
    }
#pragma unroll
    for (int f = 0; f < NF; f++) {
      s0[f] = s1[f];
      s1[f] = s2[f];
      s2[f] = s3[f];
      if (PD == 0) {
        s3[f] = (f < nf && xl <= K) ? load(f, xl) : 0.0;
      } else {
        s3[f] = (f < nf && l + 3 <= K) ? fold(f, l + 3, sq[f][0]) : 0.0;
#pragma unroll
        for (int q = 0; q < PD - 1; q++) sq[f][q] = sq[f][q + 1];
        sq[f][PQ - 1] = (f < nf && xl <= K) ? raw(f, xl) : RawT{};
      }
    }
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_s_setprio(0);
#endif
    xl++;
  }
  double lo[NF], hi[NF];
#pragma unroll
  for (int f = 0; f < NF; f++) {
    lo[f] = FAST ? smin[f] - slack(f) : smin[f];
    hi[f] = FAST ? smax[f] + slack(f) : smax[f];
  }
  sink.fixup(lo, hi, smin, smax);
}

// ---------------------------------------------------------------------
// The p-column remap of T, S and the tracers: walks of NF fields at a
// time (T and S first, with th3d), then one walk over the remaining
// fields.  Every walk is instantiated for its exact field count, so no
// field of a walk is switched off at run time.
// ---------------------------------------------------------------------
// dst: where the outputs go (NULL: in place, over the fields read)
template <int METHOD, int NW, int PD, bool FAST, bool NT, class PO, class PI>
HYB_HDI void remap_walk_pi(const KParams& P, const View& V, int64_t c0, int n, int g0, int K,
                           double zx, const PI& pi, int kp, uint64_t lcm, PO& po, int KO,
                           double* const* dst = nullptr) {
  const int64_t S = V.NIJ;
  double* fp[NW];
#pragma unroll
  for (int f = 0; f < NW; f++) {
    const int g = g0 + f;
    if (g == 0) fp[f] = V.temp + c0 + lev4(V, 1, n);
    else if (g == 1) fp[f] = V.saln + c0 + lev4(V, 1, n);
    else fp[f] = V.tracer + c0 + levt(V, 1, n, g - 1);
  }
  double* th = (g0 == 0) ? V.th3d + c0 + lev4(V, 1, n) : nullptr;
  FieldSink<NW, NT> sink{dst ? dst : fp, NW, th, P.thbase, S};
  walk_core<METHOD, NW, PD, FAST, false, true>(K, zx, pi, po, sink, fp, NW, kp, lcm, S, KO);
}
template <int METHOD, int NW, int PD, bool FAST, bool NT, class PO>
HYB_HDI void remap_walk(const KParams& P, const View& V, int64_t c0, int n, int g0, int K,
                        double zx, const double* pres, int kp, uint64_t lcm, PO& po) {
  remap_walk_pi<METHOD, NW, PD, FAST, NT>(P, V, c0, n, g0, K, zx, PiArray{pres, V.NIJ}, kp, lcm,
                                          po, 0);
}

// the last walk, over r < NF fields (r = 0: none)
template <int METHOD, int R, int PD, bool FAST, bool NT, class PO>
HYB_HDI void remap_rest(const KParams& P, const View& V, int64_t c0, int n, int g0, int r, int K,
                        double zx, const double* pres, int kp, uint64_t lcm, PO& po) {
  if constexpr (R >= 1) {
    if (r == R) remap_walk<METHOD, R, PD, FAST, NT>(P, V, c0, n, g0, K, zx, pres, kp, lcm, po);
    else remap_rest<METHOD, R - 1, PD, FAST, NT>(P, V, c0, n, g0, r, K, zx, pres, kp, lcm, po);
  }
}

template <int METHOD, int NF, int PD, bool FAST, bool NT, class PO>
HYB_HDI void remap_fields(const KParams& P, const View& V, const Scratch& W, int i, int j, int n,
                          PO& po) {
  const int K = P.kk;
  const int64_t c0 = ij_of(V, i, j);
  const int64_t S = V.NIJ;
  const int kpw = W.kp[c0];
  if (kpw & kUnsafe) return;
  const int kp = kpw & 0xffff;
  const uint64_t lcm = W.lcm[c0];
  const double* pres = W.pres + c0;
  const double zx = pres[(int64_t)K * S];
  const int nfields = 2 + P.ntracr;
  int g0 = 0;
  for (; g0 + NF <= nfields; g0 += NF)
    remap_walk<METHOD, NF, PD, FAST, NT>(P, V, c0, n, g0, K, zx, pres, kp, lcm, po);
  remap_rest<METHOD, NF - 1, PD, FAST, NT>(P, V, c0, n, g0, nfields - g0, K, zx, pres, kp, lcm,
                                           po);
}

// Copies n levels of a column (level stride S) into a lane's LDS column
// (stride ld), 16 loads in flight before the first is stored: a column of
// up to 65 levels waits on memory at most five times instead of once per
// level or four.
// (ldw, when not 0, replaces ld as a 64-bit stride: a copy between two
// columns of global memory)
HYB_HDI void stage_column(double* row, int ld, const double* pg, int64_t S, int n,
                          int64_t ldw = 0) {
  const int64_t st = ldw ? ldw : (int64_t)ld;
  for (int k0 = 0; k0 < n; k0 += 16) {
    double t[16];
#pragma unroll
    for (int q = 0; q < 16; q++) t[q] = (k0 + q < n) ? pg[(int64_t)(k0 + q) * S] : 0.0;
#pragma unroll
    for (int q = 0; q < 16; q++)
      if (k0 + q < n) row[(int64_t)(k0 + q) * st] = t[q];
  }
}

// The new interfaces of a column as the walk reads them: levels 2..K
// staged in the lane's LDS column, levels 1 and K+1 kept in registers.
// K-1 doubles per lane: at K = 41 a 64-lane wave takes 20 KB of LDS, so a
// CU's 160 KB holds the 8 waves its registers allow (K+1 levels held 7).
HYB_HD int lds_levels(int K) { return K > 1 ? K - 1 : 1; }
struct StagedInterfaces {
  const double* base;  // level 2 of the lane's column
  int stride, k1;      // LDS stride; K+1
  double top, bot;     // levels 1 and K+1
  HYB_HD double operator()(int k) const {
    if (k == 1) return top;
    if (k >= k1) return bot;
    return base[(int64_t)(k - 2) * stride];
  }
};
HYB_HDI StagedInterfaces stage_interfaces(double* row, int ld, const double* pg, int64_t S,
                                          int K) {
  const double top = pg[0], bot = pg[(int64_t)K * S];
  stage_column(row, ld, pg + S, S, K - 1);
  return StagedInterfaces{row, ld, K + 1, top, bot};
}

// The new interfaces with only levels 2..ks staged in LDS (ks-1 doubles per
// lane) and the deeper ones read from p as the walk reaches them, the next
// one loaded ahead.  A walk's requests are non-decreasing by steps of one
// and start again at k = 1 for the next walk over the column; while a walk
// is still in the staged levels, the first deep level and the one after it
// are already on their way.  With ks = K every level is staged and p is not
// read again (StagedInterfaces).  The LDS saved lets a workgroup hold more
// columns, or a CU more waves, for the price of those reads.
struct SplitInterfaces {
  const double* base;  // level 2 of the lane's LDS column
  const double* pg;    // level 1 of the column in p
  int64_t S;
  int stride, k1, ks;  // LDS stride; K+1; the deepest staged level
  double top, bot;     // levels 1 and K+1
  int k = 0;           // level of cur (0: nothing loaded)
  double cur = 0.0, nxt = 0.0;
  HYB_HD double operator()(int kq) {
    if (kq == 1) return top;
    if (kq >= k1) return bot;
    if (kq <= ks) {
      if (ks + 1 < k1 && k != ks + 1) {  // a new walk: levels ks+1, ks+2 ahead
        k = ks + 1;
        cur = pg[(int64_t)ks * S];
        nxt = pg[(int64_t)(ks + 1) * S];
      }
      return base[(int64_t)(kq - 2) * stride];
    }
    if (k == 0 || kq < k) {  // (requests that skip the staged levels)
      k = kq;
      cur = pg[(int64_t)(kq - 1) * S];
      nxt = pg[(int64_t)kq * S];
    }
    while (k < kq) {  // kq <= K, so level k+1 <= K+1
      cur = nxt;
      k++;
      nxt = pg[(int64_t)k * S];
    }
    return cur;
  }
};
// ns levels staged (clamped to 1..K-1)
HYB_HDI SplitInterfaces stage_split(double* row, int ld, const double* pg, int64_t S, int K,
                                    int ns) {
  const int n = ns < K - 1 ? (ns > 1 ? ns : 1) : K - 1;
  const double top = pg[0], bot = pg[(int64_t)K * S];
  stage_column(row, ld, pg + S, S, n);
  SplitInterfaces po{row, pg, S, ld, K + 1, n + 1, top, bot};
  return po;
}

// the last RA walk, over r < NF fields (r = 0: none)
template <int METHOD, int NW, int PD, class PO>
HYB_HDI void ra_walk(const KParams& P, const View& V, const Scratch& W, int64_t c0, int m, int g0,
                     int K, double zx, const PiRa& pi, uint64_t lcm, PO& po, bool staged) {
  if (!staged) {
    remap_walk_pi<METHOD, NW, PD, false, false>(P, V, c0, m, g0, K, zx, pi, K, lcm, po, K);
    return;
  }
  // the outputs go to the workspace, then over the fields once the walk has
  // read every level
  const int64_t S = V.NIJ;
  double* dst[NW];
#pragma unroll
  for (int f = 0; f < NW; f++) dst[f] = W.ras + c0 + (int64_t)f * K * S;
  remap_walk_pi<METHOD, NW, PD, false, false>(P, V, c0, m, g0, K, zx, pi, K, lcm, po, K, dst);
#pragma unroll
  for (int f = 0; f < NW; f++) {
    const int g = g0 + f;
    double* fp = (g == 0) ? V.temp + c0 + lev4(V, 1, m)
               : (g == 1) ? V.saln + c0 + lev4(V, 1, m)
                          : V.tracer + c0 + levt(V, 1, m, g - 1);
    stage_column(fp, 0, dst[f], S, K, S);
  }
}

template <int METHOD, int R, int PD, class PO>
HYB_HDI void ra_rest(const KParams& P, const View& V, const Scratch& W, int64_t c0, int m,
                     int g0, int r, int K, double zx, const PiRa& pi, uint64_t lcm, PO& po,
                     bool staged) {
  if constexpr (R >= 1) {
    if (r == R)
      ra_walk<METHOD, R, PD>(P, V, W, c0, m, g0, K, zx, pi, lcm, po, staged);
    else
      ra_rest<METHOD, R - 1, PD>(P, V, W, c0, m, g0, r, K, zx, pi, lcm, po, staged);
  }
}

// ---------------------------------------------------------------------
// The Robert-Asselin branch of one p-column on the streaming design
// (hybgencj, hybgen.F90:1909-2274), after hybgenaj has left the level-n
// interfaces in p.  One pass over the column forms, without writing any
// field:
//   * dp0ij and dp0cum from the level-n p (1931-1960) and the fixed layers
//     from topiso (1962-1970);
//   * the old level-m interfaces, running sums of dp(m), into the
//     workspace;
//   * the filtered thicknesses dp(m)+0.5*ra2fac*(dpo(n)+dp(n)-2*dpo(m))
//     (2024-2032) and their running sums, the new interfaces, into the
//     lane's LDS column (levels 2..K; 1 and K+1 in registers);
//   * the PCM mask (2034-2046);
//   * the in-place test of the walks (regrid_stream's) and the range test
//     of the fast quotients.
// A column that fails the range test returns false untouched, for
// column_ra.  Otherwise T, S (with th3d) and the tracers of level m are
// remapped by walk_core, with the old thicknesses dp(m) as stored (PiRa), no
// massless fill, and the new grid's K layers (KO = K: its bottom need not
// be the old one, and the column range covers every old level): in place,
// or, where an output would overwrite a level the walk has not read (the
// filtered bottom above the old one with massless layers below, for one),
// into the workspace W.ras and then over the fields.  Then dp(m) is written
// and p takes the level-m interfaces (2032).
// ---------------------------------------------------------------------
template <int METHOD, int NF, int PD>
HYB_HDI bool ra_stream(const KParams& P, const View& V, const Scratch& W, int i, int j, int m,
                       int n, double* row, int ld) {
  const int K = P.kk;
  const int64_t c0 = ij_of(V, i, j);
  const int64_t S = V.NIJ;
  double* dpm = V.dp + c0 + lev4(V, 1, m);
  const double* dpn = V.dp + c0 + lev4(V, 1, n);
  const double* dpon = V.dpo + c0 + lev4(V, 1, n);
  const double* dpom = V.dpo + c0 + lev4(V, 1, m);
  const double* th3dm = V.th3d + c0 + lev4(V, 1, m);
  const double* theta = V.theta + c0;
  double* pg = V.p + c0;
  double* pres = W.pres + c0;
  const double dpthin = kDpthin;
  // 1931-1960
  const double qdep = dmax(0.0, dmin(1.0, (V.depths[c0] - P.dsns) / (P.dpns - P.dsns)));
  const double topiso = V.topiso[c0];
  double cum = 0.0 + (qdep * P.dp0k[0] + (1.0 - qdep) * P.ds0k[0]);  // dp0cum(2)
  bool fixed = true;  // k <= fixlay
  uint64_t lcm = 0;
  double pr = 0.0, ps = 0.0;  // pres(k), prsf(k)
  double q0 = 0.0, q1 = 0.0;  // pres(k-2), pres(k-1)
  double zlast = 0.0;         // the deepest old interface above the bottom
  bool tame = true, have = false;
  double cand = 0.0;
  pres[0] = 0.0;
  auto level = [&](int k, double dmk, double dold, double dmid, double dnew, double pnk,
                   double thk, double tak) {
    if (k >= 2) {
      const double d = P.dp0k[k - 1];
      double q;
      if (d <= P.dp00i) {
        q = d;
      } else {
        q = dmax(P.dp00i, d * d / dmax(d, pnk - cum));
      }
      const double dp0 = dmin(q, qdep * d + (1.0 - qdep) * P.ds0k[k - 1]);
      fixed = fixed && !(cum >= topiso);  // 1962-1970: the loop ends at dp0cum(k) >= topiso
      cum = cum + dp0;
    }
    const double prn = pr + dmk;
    const double qr = 0.5 * P.ra2fac * (dold + dnew - 2.0 * dmid);
    const double dk = dmid + qr;
    const double psn = ps + dk;
    if (P.isopcm) {
      const bool v = !fixed && (dmk <= dpthin || dabs(thk - tak) < P.hybiso);
      if (v) lcm |= 1ull << (k - 1);
    }
    tame &= (dmk >= 0.0);
    tame &= tame_p(prn);
    tame &= tame_p(psn);
    // output k ends in old layer <= k-3 only if pres(k-2) >= prsf(k+1)
    if (k >= 4 && q0 >= psn && (!have || psn < cand)) {
      cand = psn;
      have = true;
    }
    pres[(int64_t)k * S] = prn;
    if (prn > pr) zlast = pr;
    if (k <= K - 1) row[(int64_t)(k - 1) * ld] = psn;  // prsf(k+1) for k+1 in 2..K
    q0 = q1;
    q1 = pr;
    pr = prn;
    ps = psn;
  };
  {
    // batches of 8 levels, the next batch loaded before the current one is
    // used (indices clamped into the column; values beyond it unused);
    // p(k) of level n is read here, before any of p is rewritten
    constexpr int B = 8;
    double a0[B], a1[B], a2[B], a3[B], a4[B], a5[B], a6[B];
    auto load = [&](int k0, double* x0, double* x1, double* x2, double* x3, double* x4,
                    double* x5, double* x6) {
#pragma unroll
      for (int q = 0; q < B; q++) {
        const int k = k0 + q <= K ? k0 + q : K;
        x0[q] = dpm[(int64_t)(k - 1) * S];
        x1[q] = dpon[(int64_t)(k - 1) * S];
        x2[q] = dpom[(int64_t)(k - 1) * S];
        x3[q] = dpn[(int64_t)(k - 1) * S];
        x4[q] = pg[(int64_t)(k - 1) * S];
        x5[q] = th3dm[(int64_t)(k - 1) * S];
        x6[q] = theta[(int64_t)(k - 1) * S];
      }
    };
    load(1, a0, a1, a2, a3, a4, a5, a6);
#pragma unroll 1
    for (int k0 = 1; k0 <= K; k0 += B) {
      double b0[B], b1[B], b2[B], b3[B], b4[B], b5[B], b6[B];
      load(k0 + B <= K ? k0 + B : K, b0, b1, b2, b3, b4, b5, b6);
#pragma unroll
      for (int q = 0; q < B; q++)
        if (k0 + q <= K) level(k0 + q, a0[q], a1[q], a2[q], a3[q], a4[q], a5[q], a6[q]);
#pragma unroll
      for (int q = 0; q < B; q++) {
        a0[q] = b0[q];
        a1[q] = b1[q];
        a2[q] = b2[q];
        a3[q] = b3[q];
        a4[q] = b4[q];
        a5[q] = b5[q];
        a6[q] = b6[q];
      }
    }
  }
  const double zx = pr;  // pres(kk+1)
  // An output ending above zlast may overwrite a level the walk reads
  // later.  One ending below it ends in the deepest layer, which reaches the
  // bottom: the walk has folded every unread level into the column range at
  // the top of that layer (walk_core), every later output ends there too,
  // and nothing more is read.  (With the interfaces in order, tame.)
  const bool staged = have && cand <= zlast;
  if (!tame || (staged && !W.ras)) return false;
  StagedInterfaces po{row, ld, K + 1, 0.0, ps};
  PiRa pi{pres, dpm, S};
  const int nfields = 2 + P.ntracr;
  int g0 = 0;
  for (; g0 + NF <= nfields; g0 += NF)
    ra_walk<METHOD, NF, PD>(P, V, W, c0, m, g0, K, zx, pi, lcm, po, staged);
  ra_rest<METHOD, NF - 1, PD>(P, V, W, c0, m, g0, nfields - g0, K, zx, pi, lcm, po, staged);
  // 2024-2032: dp(m) and the level-m interfaces
  double sum = 0.0;
#pragma unroll 4
  for (int k = 1; k <= K; k++) {
    const double qr = 0.5 * P.ra2fac * (dpon[(int64_t)(k - 1) * S] + dpn[(int64_t)(k - 1) * S] -
                                        2.0 * dpom[(int64_t)(k - 1) * S]);
    const double dk = dpom[(int64_t)(k - 1) * S] + qr;
    dpm[(int64_t)(k - 1) * S] = dk;
    sum = sum + dk;
    pg[(int64_t)k * S] = sum;
  }
  return true;
}


// ---------------------------------------------------------------------
// Mixed-precision remap (HYB_PREC_MIXED; no reference counterpart): the
// walk of walk_core with FAST set.  The reconstruction (coef_layer) runs
// in float with hardware reciprocals; the piece integrals, the output sums
// and the interfaces stay FP64, and the quotients of the walk multiply by
// refined FP64 reciprocals.  Every old layer's polynomial keeps its mean
// s(l) exactly (the PPM c3 is formed from s(l) in FP64; PLM and WENO keep
// it by construction), so each layer hands out s(l)*dp(l) up to FP64
// roundoff and the column integrals are conserved as in the FP64 walk.
// The overturn/unmix, the interfaces and every decision of the regrid stay
// FP64 and bitwise the reference's.
// ---------------------------------------------------------------------
template <int METHOD, int NF, int PD, class PO>
HYB_HDI void remap_mixed(const KParams& P, const View& V, const Scratch& W, int i, int j, int n,
                         PO& po) {
  remap_fields<METHOD, NF, PD, true, false>(P, V, W, i, j, n, po);
}

template <int METHOD, int NF, int PD, class PO, bool NT = false>
HYB_HDI void remap_stream(const KParams& P, const View& V, const Scratch& W, int i, int j, int n,
                          PO& po) {
  remap_fields<METHOD, NF, PD, false, NT>(P, V, W, i, j, n, po);
}

// ---------------------------------------------------------------------
// One u- or v-point: the new thickness (dpudpv.F90:92-103), the old
// interfaces as running sums of the old dpu (hybgen.F90:171-195), the new
// ones as running sums of the new dpu (1806-1810), and the remap of u or v
// with the thin-layer rule (hybgenbj, 1758-1907).  dcol is the lane's LDS
// column for the new thicknesses.  Where an output of the in-place walk
// would overwrite a level not yet read (the test of regrid_stream), it
// writes pu only and returns false; point_c then redoes the point.
// ---------------------------------------------------------------------
template <int METHOD, int PD, bool MIXED = false, bool RECOMP = false>
HYB_HDI bool uv_stream(const KParams& P, const View& V, int i, int j, int nl, bool isv,
                      const LaneRow& dcol) {
  const int K = P.kk;
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
  double b = dmin(depthu, 0.5 * (p0[0] + p1[0]));
  double pr = 0.0;  // prsf(k+1)
  pu[0] = 0.0;
  double q0 = 0.0, q1 = 0.0, q2 = 0.0;  // pres(k-2), pres(k-1), pres(k)
  bool have = false;
  double cand = 0.0;
  // the walk's fast quotients need the old interfaces in order and both
  // sets of interfaces 0 or in [2^-60, 2^60]; other points go to point_c
  bool tame = true;
  // level k of the pass: p at both points (level k+1) and the old dpu(k)
  auto step = [&](int k, double pa, double pb, double dk) {
    const double a = dmin(depthu, 0.5 * (pa + pb));
    const double d = dmax(0.0, a - b);
    b = a;
    if (!RECOMP) dcol(k) = d;
    pr = pr + d;
    pu[(int64_t)k * S] = pr;
    // output k ends in old layer <= k-3 only if pres(k-2) >= prsf(k+1)
    if (k >= 4 && q0 >= pr && (!have || pr < cand)) {
      cand = pr;
      have = true;
    }
    const double q3 = q2 + dk;
    tame &= (dk >= 0.0);
    tame &= tame_p(q3);
    tame &= tame_p(pr);
    q0 = q1;
    q1 = q2;
    q2 = q3;
  };
  {
    // batches of 8 levels, the next batch's loads issued before the
    // current batch is summed (indices clamped into the column; values
    // beyond it are not used): a wave waits on memory once per batch with
    // the next batch already on its way
    constexpr int B = 8;
    double ca[B], cb[B], cd[B];
    auto load = [&](int k0, double* xa, double* xb, double* xd) {
#pragma unroll
      for (int q = 0; q < B; q++) {
        const int k = k0 + q <= K ? k0 + q : K;
        xa[q] = p0[(int64_t)k * S];
        xb[q] = p1[(int64_t)k * S];
        xd[q] = dpu[(int64_t)(k - 1) * S];
      }
    };
    load(1, ca, cb, cd);
#pragma unroll 1
    for (int k0 = 1; k0 <= K; k0 += B) {
      double na[B], nb[B], nd[B];
      load(k0 + B <= K ? k0 + B : K, na, nb, nd);
#pragma unroll
      for (int q = 0; q < B; q++)
        if (k0 + q <= K) step(k0 + q, ca[q], cb[q], cd[q]);
#pragma unroll
      for (int q = 0; q < B; q++) {
        ca[q] = na[q];
        cb[q] = nb[q];
        cd[q] = nd[q];
      }
    }
  }
  const double zx = q2;
  if ((have && cand < zx) || !tame) return false;
  PiSum pi{dpu, S};
  double* fp[1] = {uu};
  if (RECOMP) {
    PoRecompute po(p0, p1, S, depthu, K);
    UvSink<PoRecompute> sink{uu, S, ub, pr, &po, 0ull, dpu};
    walk_core<METHOD, 1, PD, MIXED, false, true>(K, zx, pi, po, sink, fp, 1, K, 0ull, S);
  } else {
    PoSum po{dcol};
    UvSink<PoSum> sink{uu, S, ub, pr, &po};
    walk_core<METHOD, 1, PD, MIXED, false, true>(K, zx, pi, po, sink, fp, 1, K, 0ull, S);
#pragma unroll 2
    for (int k = 1; k <= K; k++) dpu[(int64_t)(k - 1) * S] = dcol(k);
  }
  return true;
}

// remapper of hybgenbj (hybgen.F90:1819-1823): PCM, PLM when hybiso > 2,
// otherwise WENO
HYB_HD int uv_method(const KParams& P) {
  if (P.hybmap == 0) return 0;
  if (P.hybmap == 1 && P.hybiso > 2.0) return 1;
  return 3;
}

// ---------------------------------------------------------------------
// Remap of a column flagged kUnsafe: copies the column first (1013-1034),
// then the reference-shaped coefs/walk of hyb_column.h.
// ---------------------------------------------------------------------
HYB_HD void remap_fallback(const KParams& P, const View& V, const Scratch& W, int i, int j,
                           int n) {
  const int kk = P.kk;
  const int64_t c0 = ij_of(V, i, j);
  const int64_t S = V.NIJ;
  const int kpw = W.kp[c0];
  if (!(kpw & kUnsafe)) return;
  const int kp = kpw & 0xffff;
  const uint64_t lcm = W.lcm[c0];
  const double dpthin = kDpthin;
  double pres[kKMax + 2], dprs[kKMax + 2], dpi[kKMax + 2], prsf[kKMax + 2];
  double sfld[kKMax + 2], c1[kKMax + 2], c2[kKMax + 2], c3[2 * kKMax + 6];
  for (int k = 1; k <= kk + 1; k++) {
    pres[k] = W.pres[c0 + (int64_t)(k - 1) * S];
    prsf[k] = V.p[c0 + (int64_t)(k - 1) * S];
  }
  for (int k = 1; k <= kk; k++) {
    dprs[k] = pres[k + 1] - pres[k];
    dpi[k] = dmax(dprs[k], dpthin);
  }
  double* temp = V.temp + c0 + lev4(V, 1, n);
  double* saln = V.saln + c0 + lev4(V, 1, n);
  double* th3d = V.th3d + c0 + lev4(V, 1, n);
  auto run_field = [&](double* f, bool ts) {
    for (int k = 1; k <= kk; k++) sfld[k] = (k <= kp) ? f[(int64_t)(k - 1) * S] : sfld[kp];
    auto store = [&](int k, double v) {
      f[(int64_t)(k - 1) * S] = v;
      if (ts) th3d[(int64_t)(k - 1) * S] = sig(temp[(int64_t)(k - 1) * S], v) - P.thbase;
    };
    switch (P.hybmap) {
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
  run_field(temp, false);
  run_field(saln, true);
  for (int t = 1; t <= P.ntracr; t++) run_field(V.tracer + c0 + levt(V, 1, n, t), false);
}

}  // namespace hyb
