// hyb_trc.h -- the tracer column remaps of trcupd.F90: pcmtrc (808-909)
// and plmtrc with its slopes plmtrcx/plmtrcs (911-1147), as initrc
// (trcupd.F90:167-199) and the tracer relaxation of thermf (thermf.F90:
// 770-806) call them per sea column.  Shared by the host checks and the
// gfx950 kernel: one column per lane, fields one after another.
//
// These are not hybgen's remappers: the overlap weights are
// max(overlap, thin)/(zb-zt) with thin = 1e-6, a layer's sum starts at
// zero, and each output scans the old layers from a carried start lf.
// A scan that does not end early leaves lf where it was, so the next
// output scans those layers again and each adds thin/(zb-zt) of its value;
// that is the reference's result, and it is kept.
#pragma once

#include "hyb_math.h"

namespace hyb {

struct KTrc {
  int32_t ii, jj, ki, ko, ks, method, nb;
  int64_t NI, NIJ;
  const double *si, *pi, *po;  // (NIJ,ki,ks), (NIJ,ki+1), (NIJ,ko+1)
  const int32_t* mask;
  double* so;                  // (NIJ,ko,ks)
  double* slope;               // (NIJ,ki,4) workspace: plmtrcx's slopes of up to 4 fields
  int32_t* bad;                // first column (c0) whose scan would stop, or INT32_MAX
};

HYB_HD int64_t trc_ij(const KTrc& Q, int i, int j) {
  return (int64_t)(i - 1 + Q.nb) + Q.NI * (int64_t)(j - 1 + Q.nb);
}

constexpr double kTrcThin = 1.e-6;  // trcupd.F90:857, 960

// plmtrcx/plmtrcs (1038-1147) for fields f0..f0+nf-1 of column c0: the
// slopes into Q.slope (zone-major, kTrcChunk per zone), each field's min
// and max (plmtrc 964-967) into mn, mx.  One pass: the raw slope of zone
// j+1 (Campbell eq. 15, 1121-1130) is formed one zone ahead, because eq. 19
// (1134-1143) changes zone j with zone j-1 already changed and zone j+1
// not yet.  The grid ratios ql, qr are shared by the fields.
constexpr int kTrcChunk = 4;  // fields remapped together

template <int NF>
HYB_HDI void trc_slopes(const KTrc& Q, int64_t c0, int f0, double* mn, double* mx) {
  const int64_t S = Q.NIJ;
  const int ki = Q.ki;
  const double* p = Q.pi + c0;
  const double* a[NF];
#pragma unroll
  for (int f = 0; f < NF; f++) a[f] = Q.si + c0 + S * (int64_t)ki * (f0 + f);
  auto A = [&](int f, int l) -> double { return a[f][(int64_t)(l - 1) * S]; };
  double* s = Q.slope + c0;
  auto SL = [&](int f, int l) -> double& { return s[S * ((int64_t)(l - 1) * kTrcChunk + f)]; };
  // windows a(j-1..j+2) of every field and pit(j-1..j+2), pit(l) =
  // max(pi(l+1)-pi(l), thin) (969-971); pn = pi(l+1) of the last pit formed
  double am[NF], a0[NF], ap[NF], aq[NF];
  double pn = p[0];
  auto next_pt = [&](int l) -> double {  // pit(l), l increasing by one from 1
    const double pl1 = p[(int64_t)l * S];
    const double t = dmax(pl1 - pn, kTrcThin);
    pn = pl1;
    return t;
  };
  const double tm0 = next_pt(1);
  const double t00 = (ki >= 2) ? next_pt(2) : 0.0;
  const double tp0 = (ki >= 3) ? next_pt(3) : 0.0;
  const double tq0 = (ki >= 4) ? next_pt(4) : 0.0;
  double tm = tm0, t0 = t00, tp = tp0, tq = tq0;
#pragma unroll
  for (int f = 0; f < NF; f++) {
    am[f] = A(f, 1);
    a0[f] = (ki >= 2) ? A(f, 2) : 0.0;
    ap[f] = (ki >= 3) ? A(f, 3) : 0.0;
    aq[f] = (ki >= 4) ? A(f, 4) : 0.0;
    double lo = am[f], hi = am[f];  // the range over all ki levels (any order: min/max are exact)
    if (ki >= 2) { lo = dmin(lo, a0[f]); hi = dmax(hi, a0[f]); }
    if (ki >= 3) { lo = dmin(lo, ap[f]); hi = dmax(hi, ap[f]); }
    if (ki >= 4) { lo = dmin(lo, aq[f]); hi = dmax(hi, aq[f]); }
    mn[f] = lo;
    mx[f] = hi;
    SL(f, 1) = 0.0;
  }
  // raw slope (Campbell eq. 15) of a zone from its window
  auto raw = [&](double xm, double x0, double xp, double sm_, double s0_, double sp_) -> double {
    const double rl = 2.0 * s0_ / (sm_ + s0_);
    const double rr = 2.0 * s0_ / (s0_ + sp_);
    const double sl = rl * (x0 - xm);
    const double sr = rr * (xp - x0);
    return (sl * sr > 0.0) ? copysign(dmin(fabs(sl), fabs(sr)), sl) : 0.0;
  };
  double sm[NF], rj[NF];
#pragma unroll
  for (int f = 0; f < NF; f++) {
    sm[f] = 0.0;                                                     // s(j-1), changed
    rj[f] = (ki >= 3) ? raw(am[f], a0[f], ap[f], tm, t0, tp) : 0.0;  // raw s(2)
  }
  for (int j = 2; j <= ki - 1; j++) {
    const bool has_n = j + 1 <= ki - 1;  // zone j+1 is an inner zone
#pragma unroll
    for (int f = 0; f < NF; f++) {
      const double rn = has_n ? raw(a0[f], ap[f], aq[f], t0, tp, tq) : 0.0;  // raw s(j+1)
      double sj = rj[f];
      if (sj != 0.0) {
        const double dl = -0.5 * (sj + sm[f]) + a0[f] - am[f];
        const double dr = -0.5 * (rn + sj) + ap[f] - a0[f];
        sj = sj + 2.0 * copysign(dmin(fabs(dl), fabs(dr)), dl);
      }
      SL(f, j) = sj;
      sm[f] = sj;
      rj[f] = rn;
    }
    // slide the windows by one zone
    const bool more = j + 3 <= ki;
    tm = t0;
    t0 = tp;
    tp = tq;
    tq = more ? next_pt(j + 3) : 0.0;
#pragma unroll
    for (int f = 0; f < NF; f++) {
      am[f] = a0[f];
      a0[f] = ap[f];
      ap[f] = aq[f];
      if (more) {
        aq[f] = A(f, j + 3);
        mn[f] = dmin(mn[f], aq[f]);
        mx[f] = dmax(mx[f], aq[f]);
      } else {
        aq[f] = 0.0;
      }
    }
  }
#pragma unroll
  for (int f = 0; f < NF; f++)
    if (ki >= 2) SL(f, ki) = 0.0;
}

// pcmtrc (METHOD 0) or plmtrc (METHOD 1) of fields f0..f0+NF-1 of column
// c0, in one scan of the old layers (the reference's loop over ks inside
// the loop over l).  Returns false where the reference stops ('bad lf',
// 866-869); the outputs from there on are NaN.
template <int METHOD, int NF>
HYB_HDI bool trc_remap_fields(const KTrc& Q, int64_t c0, int f0) {
  const int64_t S = Q.NIJ;
  const int ki = Q.ki, ko = Q.ko;
  const double thin = kTrcThin;
  const double* pi = Q.pi + c0;
  const double* po = Q.po + c0;
  const double* si[NF];
  double* so[NF];
#pragma unroll
  for (int f = 0; f < NF; f++) {
    si[f] = Q.si + c0 + S * (int64_t)ki * (f0 + f);
    so[f] = Q.so + c0 + S * (int64_t)ko * (f0 + f);
  }
  const double* sis = Q.slope + c0;
  auto PI = [&](int l) -> double { return pi[(int64_t)(l - 1) * S]; };
  double mn[NF], mx[NF], prev[NF], sok[NF];
  if (METHOD == 1) trc_slopes<NF>(Q, c0, f0, mn, mx);
#pragma unroll
  for (int f = 0; f < NF; f++) prev[f] = si[f][0];  // the reference reads so(0) here
  const double zx = PI(ki + 1);
  int lf = 1;
  double zb = po[0];
  bool ok = true;
  for (int k = 1; k <= ko; k++) {
    const double zt = zb;
    zb = po[(int64_t)k * S];
    if (zb - zt < thin || zt >= zx) {
      // thin or bottomed: the output above
    } else if (PI(lf) > zt) {
      ok = false;
#pragma unroll
      for (int f = 0; f < NF; f++) prev[f] = NAN;
    } else {
      const double dz = zb - zt;
#pragma unroll
      for (int f = 0; f < NF; f++) sok[f] = 0.0;
      // pi(l), pi(l+1) and layer l's words are in registers when
      // iteration l starts: each iteration loads layer l+1's while it sums
      // layer l (a break leaves those loads unused)
      double pl1 = PI(lf), pl2 = PI(lf + 1);
      double cv[NF], cs[NF];
#pragma unroll
      for (int f = 0; f < NF; f++) {
        cv[f] = si[f][(int64_t)(lf - 1) * S];
        cs[f] = (METHOD == 1) ? sis[S * ((int64_t)(lf - 1) * kTrcChunk + f)] : 0.0;
      }
      for (int l = lf; l <= ki; l++) {
        const double pl = pl1;
        if (pl > zb) {
          lf = l - 1;
          break;
        }
        pl1 = pl2;
        const bool more = l + 1 <= ki;
        double nv[NF], ns[NF];
        pl2 = more ? PI(l + 2) : 0.0;
#pragma unroll
        for (int f = 0; f < NF; f++) {
          nv[f] = more ? si[f][(int64_t)l * S] : 0.0;
          ns[f] = (METHOD == 1 && more) ? sis[S * ((int64_t)l * kTrcChunk + f)] : 0.0;
        }
        if (pl >= zt && pl1 <= zb) {
          const double q = dmax(pl1 - pl, thin) / dz;
#pragma unroll
          for (int f = 0; f < NF; f++) sok[f] = sok[f] + q * cv[f];
        } else {
          const double bot = dmin(pl1, zb), top = dmax(pl, zt);
          const double q = dmax(bot - top, thin) / dz;
          if (METHOD == 0) {
#pragma unroll
            for (int f = 0; f < NF; f++) sok[f] = sok[f] + q * cv[f];
          } else {
            const double zc = 0.5 * (bot + top);
            const double qc = (zc - pl) / dmax(pl1 - pl, thin) - 0.5;
#pragma unroll
            for (int f = 0; f < NF; f++) sok[f] = sok[f] + q * (cv[f] + qc * cs[f]);
          }
        }
#pragma unroll
        for (int f = 0; f < NF; f++) {
          cv[f] = nv[f];
          cs[f] = ns[f];
        }
      }
#pragma unroll
      for (int f = 0; f < NF; f++) {
        double v = sok[f];
        if (METHOD == 1) {
          v = dmax(mn[f], v);
          v = dmin(mx[f], v);
        }
        prev[f] = v;
      }
    }
#pragma unroll
    for (int f = 0; f < NF; f++) so[f][(int64_t)(k - 1) * S] = prev[f];
  }
  return ok;
}

template <int METHOD>
HYB_HDI bool trc_chunk(const KTrc& Q, int64_t c0, int f0, int nf) {
  switch (nf) {
    case 4: return trc_remap_fields<METHOD, 4>(Q, c0, f0);
    case 3: return trc_remap_fields<METHOD, 3>(Q, c0, f0);
    case 2: return trc_remap_fields<METHOD, 2>(Q, c0, f0);
    default: return trc_remap_fields<METHOD, 1>(Q, c0, f0);
  }
}

HYB_HDI bool trc_column(const KTrc& Q, int64_t c0) {
  bool ok = true;
  for (int f0 = 0; f0 < Q.ks; f0 += kTrcChunk) {
    const int nf = Q.ks - f0 < kTrcChunk ? Q.ks - f0 : kTrcChunk;
    if (Q.method == 0) ok = trc_chunk<0>(Q, c0, f0, nf) && ok;
    else ok = trc_chunk<1>(Q, c0, f0, nf) && ok;
  }
  return ok;
}

}  // namespace hyb
