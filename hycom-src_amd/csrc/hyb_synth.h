// hyb_synth.h -- seeded synthetic ocean state for hybgen workloads.
//
// No reference counterpart: HYCOM reads its state from restart files.  The
// generator is a pure function of (seed, global i, global j, k) built from
// the exact helpers in hyb_math.h, so the host build and the gfx950 build
// produce identical bits for any window of the global grid.
//
// Column recipe (per global sea point):
//   depth   smooth random bathymetry: a sum of M plane waves, thresholded
//           to give the requested land fraction, mapped to
//           [depth_min, depth_max] with an optional shallow shelf band;
//   T(z)    Ts(lat) * exp(-z/800 m) + 2 C, Ts in [2,28];
//   S(z)    S_deep + (S_surf(lat) - S_deep) * exp(-z/300 m), S_surf in
//           [34, 36.5], plus N(0, ts_noise) per layer;
//   p       interface k+1 sits where sig(T(z),S(z)) crosses the midpoint
//           of the sigma-2 targets of layers k and k+1 (bisection), then
//           is perturbed by up to 0.42 of the smaller neighbouring gap;
//           lighter-than-surface layers vanish at the top, denser-than-
//           bottom layers vanish on the sea floor;
//   tracers trcflg==2 tracers are bit copies of T; others are positive
//           smooth profiles;
//   u, v    N(0, uv_std) on sea u/v points; ubavg/vbavg their dpu/dpv
//           weighted column means.
#pragma once

#include "hyb_column.h"
#include "hyb_math.h"

namespace hyb {

constexpr int kSynthModes = 12;

struct SynthDomain {
  uint64_t seed;
  int32_t itdm, jtdm, i0, j0;
  int32_t periodic_i, periodic_j;
  int32_t K, ntracr;
  int32_t extreme;           // config-4 style fronts and 1e-3 m layers
  int32_t pad_;
  double dmin, dmax, thr, hmax, shelf;
  double sig_top, sig_bot, uv_std, ts_noise, thbase;
  double amp[kSynthModes], kx[kSynthModes], ky[kSynthModes], ph[kSynthModes];
  int32_t trcflg[16];
  double topiso_pa;          // onem*isotop
};

constexpr double kTwoPi = 6.28318530717958647692;

// water depth (m) at global point (ig, jg), 0 on land.  Points outside the
// global grid wrap when periodic, else are land.
HYB_HD double synth_depth(const SynthDomain& D, long ig, long jg) {
  if (ig < 1 || ig > D.itdm) {
    if (!D.periodic_i) return 0.0;
    ig = ((ig - 1) % D.itdm + D.itdm) % D.itdm + 1;
  }
  if (jg < 1 || jg > D.jtdm) {
    if (!D.periodic_j) return 0.0;
    jg = ((jg - 1) % D.jtdm + D.jtdm) % D.jtdm + 1;
  }
  double x = ((double)ig - 0.5) / (double)D.itdm;
  double y = ((double)jg - 0.5) / (double)D.jtdm;
  double f = 0.0;
  for (int m = 0; m < kSynthModes; m++)
    f = f + D.amp[m] * xcos(kTwoPi * (D.kx[m] * x + D.ky[m] * y) + D.ph[m]);
  if (f <= D.thr) return 0.0;
  double h = (f - D.thr) / D.hmax;
  if (h > 1.0) h = 1.0;
  double shelf_depth = 200.0;
  if (shelf_depth > D.dmax) shelf_depth = D.dmax;
  double d;
  if (h < D.shelf) {
    d = D.dmin + (shelf_depth - D.dmin) * (h / D.shelf);
  } else {
    double q = D.shelf < 1.0 ? (h - D.shelf) / (1.0 - D.shelf) : 1.0;
    d = shelf_depth + (D.dmax - shelf_depth) * sqrt(q);
  }
  return d;
}

HYB_HD void synth_ts_surface(const SynthDomain& D, long ig, long jg,
                             double* Ts, double* Ss) {
  double x = ((double)ig - 0.5) / (double)D.itdm;
  double y = ((double)jg - 0.5) / (double)D.jtdm;
  double yy = 2.0 * y - 1.0;
  double a = 1.0 - yy * yy;
  *Ts = 2.0 + 26.0 * a * (0.85 + 0.15 * xcos(kTwoPi * x + 1.3));
  double q = ((yy < 0.0 ? -yy : yy) - 0.35) / 0.65;
  if (q < -1.0) q = -1.0;
  if (q > 1.0) q = 1.0;
  *Ss = 34.0 + 2.5 * (1.0 - q * q);
  if (D.extreme) {
    // 40 C / 40 psu fronts: alternate columns flip between two extremes
    if (((ig + jg) & 7) == 0) { *Ts = 38.0; *Ss = 40.0; }
    if (((ig + jg) & 7) == 4) { *Ts = -2.0; *Ss = 30.0; }
  }
}

HYB_HD double synth_T(double Ts, double z) { return Ts * xexp(-z / 800.0) + 2.0; }
HYB_HD double synth_S(double Ss, double z) { return 34.7 + (Ss - 34.7) * xexp(-z / 300.0); }

// Interface depths z[0..K] (m) of one sea column, z[0]=0, z[K]=depth.
HYB_HD void synth_interfaces(const SynthDomain& D, long ig, long jg,
                             double depth, double* z) {
  const int K = D.K;
  double Ts, Ss;
  synth_ts_surface(D, ig, jg, &Ts, &Ss);
  double sig0 = sig(synth_T(Ts, 0.0), synth_S(Ss, 0.0));
  double sigb = sig(synth_T(Ts, depth), synth_S(Ss, depth));
  z[0] = 0.0;
  for (int k = 1; k < K; k++) {
    double s1 = D.sig_top + (D.sig_bot - D.sig_top) * (double)(k - 1) / (double)(K - 1);
    double s2 = D.sig_top + (D.sig_bot - D.sig_top) * (double)(k) / (double)(K - 1);
    double tgt = 0.5 * (s1 + s2);
    double zk;
    if (sig0 >= tgt) {
      zk = 0.0;
    } else if (sigb <= tgt) {
      zk = depth;
    } else {
      double lo = 0.0, hi = depth;
      for (int it = 0; it < 36; it++) {
        double mid = 0.5 * (lo + hi);
        if (sig(synth_T(Ts, mid), synth_S(Ss, mid)) < tgt) lo = mid; else hi = mid;
      }
      zk = hi;
    }
    if (zk < z[k - 1]) zk = z[k - 1];
    z[k] = zk;
  }
  z[K] = depth;
  // perturb each interior interface by <= 0.42 of its smaller gap
  double prev = z[0];
  for (int k = 1; k < K; k++) {
    double gu = z[k] - z[k - 1], gd = z[k + 1] - z[k];
    double gap = gu < gd ? gu : gd;
    double d = 0.12 * nrm(D.seed + 3, (uint64_t)ig, (uint64_t)jg, (uint64_t)k) * gap;
    double zk = z[k] + d;
    if (zk < prev) zk = prev;
    if (zk > depth) zk = depth;
    z[k] = zk;
    prev = zk;
  }
  if (D.extreme) {
    // a stack of 1e-3 m layers just below the surface layer
    for (int k = 1; k < K && k < 6; k++) {
      double zk = z[1] + 0.001 * (double)(k - 1);
      if (zk > depth) zk = depth;
      if (zk > z[k]) z[k] = zk;
    }
    for (int k = 2; k <= K; k++) if (z[k] < z[k - 1]) z[k] = z[k - 1];
  }
}

// Layer k (1-based) properties from interface depths.
HYB_HD void synth_layer(const SynthDomain& D, long ig, long jg, int k,
                        const double* z, double* T, double* S, double* th) {
  double Ts, Ss;
  synth_ts_surface(D, ig, jg, &Ts, &Ss);
  double zm = 0.5 * (z[k - 1] + z[k]);
  *T = synth_T(Ts, zm);
  *S = synth_S(Ss, zm) + D.ts_noise * nrm(D.seed + 2, (uint64_t)ig, (uint64_t)jg, (uint64_t)k);
  *th = sig(*T, *S) - D.thbase;
}

HYB_HD double synth_tracer(const SynthDomain& D, long ig, long jg, int k,
                           int t, const double* z) {
  double zm = 0.5 * (z[k - 1] + z[k]);
  double x = ((double)ig - 0.5) / (double)D.itdm;
  double y = ((double)jg - 0.5) / (double)D.jtdm;
  double a = 1.0 + 0.5 * xcos(kTwoPi * (x + 0.37 * (double)t) + 3.0 * y);
  return a * xexp(-zm / (200.0 + 300.0 * (double)t));
}

// Pass 1 at window point (i,j), 0<=i<=ii+1, 0<=j<=jj+1: depths, ip, p; on
// interior points also the layer fields, theta and topiso.
HYB_HD void synth_point_pass1(const SynthDomain& D, const View& V, int ii, int jj,
                              int i, int j, int m, int n) {
  const int K = D.K;
  const long ig = (long)D.i0 + i, jg = (long)D.j0 + j;
  const int64_t c0 = ij_of(V, i, j);
  const int64_t S = V.NIJ;
  const bool interior = i >= 1 && i <= ii && j >= 1 && j <= jj;
  double depth = synth_depth(D, ig, jg);
  const_cast<double*>(V.depths)[c0] = depth;
  const_cast<int32_t*>(V.ip)[c0] = depth > 0.0 ? 1 : 0;
  double z[65];
  double pk = 0.0;
  if (depth > 0.0) synth_interfaces(D, ig, jg, depth, z);
  V.p[c0] = 0.0;
  for (int k = 1; k <= K; k++) {
    double d = 0.0;
    if (depth > 0.0) {
      double zt = z[k - 1] * kOnem;
      double zb = (k == K) ? depth * kOnem : z[k] * kOnem;
      d = zb - zt;
    }
    pk = pk + d;
    V.p[c0 + (int64_t)k * S] = pk;
    if (!interior) continue;
    const_cast<double*>(V.theta)[c0 + lev(V, k)] =
        D.sig_top + (D.sig_bot - D.sig_top) * (double)(k - 1) / (double)(K - 1) - D.thbase;
    double T = 0.0, Sa = 0.0, th = 0.0;
    if (depth > 0.0) synth_layer(D, ig, jg, k, z, &T, &Sa, &th);
    for (int l = 1; l <= 2; l++) {
      if (l != m && l != n) continue;
      int64_t a = c0 + lev4(V, k, l);
      V.dp[a] = d;
      if (V.dpo) V.dpo[a] = d;
      V.temp[a] = T;
      V.saln[a] = Sa;
      V.th3d[a] = th;
      for (int t = 1; t <= D.ntracr; t++) {
        double tr = 0.0;
        if (depth > 0.0) tr = (D.trcflg[t - 1] == 2) ? T : synth_tracer(D, ig, jg, k, t, z);
        V.tracer[c0 + levt(V, k, l, t)] = tr;
      }
    }
  }
  if (interior) const_cast<double*>(V.topiso)[c0] = D.topiso_pa;
}

// Pass 2 at interior point (i,j): u/v masks, bottom pressures, dpu/dpv from
// the pass-1 interfaces (dpudpv.F90:95-102), velocities and their means.
HYB_HD void synth_point_pass2(const SynthDomain& D, const View& V, int ii, int jj,
                              int i, int j, int m, int n) {
  (void)ii; (void)jj;
  const int K = D.K;
  const long ig = (long)D.i0 + i, jg = (long)D.j0 + j;
  const int64_t c0 = ij_of(V, i, j);
  const int64_t S = V.NIJ;
  const double pb = V.depths[c0] * kOnem;
  for (int isv = 0; isv <= 1; isv++) {
    const int64_t cn = isv ? ij_of(V, i, j - 1) : ij_of(V, i - 1, j);
    const double pbn = V.depths[cn] * kOnem;
    const double dep = dmin(pb, pbn);
    const bool sea = V.ip[c0] > 0 && V.ip[cn] > 0;
    const_cast<double*>(isv ? V.depthv : V.depthu)[c0] = dep;
    const_cast<int32_t*>(isv ? V.iv : V.iu)[c0] = sea ? 1 : 0;
    double* dpu = isv ? V.dpv : V.dpu;
    double* uu = isv ? V.v : V.u;
    double* pu = isv ? V.pv : V.pu;
    double sum = 0.0, wsum = 0.0;
    for (int k = 1; k <= K; k++) {
      double d = 0.0, vel = 0.0;
      if (sea) {
        double a = dmin(dep, 0.5 * (V.p[c0 + (int64_t)k * S] + V.p[cn + (int64_t)k * S]));
        double b = dmin(dep, 0.5 * (V.p[c0 + (int64_t)(k - 1) * S] + V.p[cn + (int64_t)(k - 1) * S]));
        d = dmax(0.0, a - b);
        vel = D.uv_std * nrm(D.seed + 5 + (uint64_t)isv, (uint64_t)ig, (uint64_t)jg, (uint64_t)k);
      }
      sum = sum + vel * d;
      wsum = wsum + d;
      for (int l = 1; l <= 2; l++) {
        if (l != m && l != n) continue;
        dpu[c0 + lev4(V, k, l)] = d;
        uu[c0 + lev4(V, k, l)] = vel;
      }
      pu[c0 + (int64_t)(k - 1) * S] = 0.0;
    }
    pu[c0 + (int64_t)K * S] = 0.0;
    double avg = wsum > 0.0 ? sum / wsum : 0.0;
    double* ub = const_cast<double*>(isv ? V.vbavg : V.ubavg);
    for (int l = 1; l <= 3; l++) ub[c0 + lev(V, l)] = avg;
  }
}

}  // namespace hyb
