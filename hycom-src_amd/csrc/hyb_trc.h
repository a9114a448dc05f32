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
  auto PT = [&](int l) -> double {  // pit(l) = max(pi(l+1)-pi(l), thin) (969-971)
    return dmax(p[(int64_t)l * S] - p[(int64_t)(l - 1) * S], kTrcThin);
  };
  double* s = Q.slope + c0;
  auto SL = [&](int f, int l) -> double& { return s[S * ((int64_t)(l - 1) * kTrcChunk + f)]; };
  // raw slopes of zone j for every field
  auto raw = [&](int j, double* r) {
#pragma unroll
    for (int f = 0; f < NF; f++) r[f] = 0.0;
    if (j <= 1 || j >= ki) return;
    const double pm = PT(j - 1), p0 = PT(j), pp = PT(j + 1);
    const double rl = 2.0 * p0 / (pm + p0);
    const double rr = 2.0 * p0 / (p0 + pp);
#pragma unroll
    for (int f = 0; f < NF; f++) {
      const double sl = rl * (A(f, j) - A(f, j - 1));
      const double sr = rr * (A(f, j + 1) - A(f, j));
      if (sl * sr > 0.0) r[f] = copysign(dmin(fabs(sl), fabs(sr)), sl);
    }
  };
#pragma unroll
  for (int f = 0; f < NF; f++) {
    double lo = A(f, 1), hi = lo;
    for (int l = 2; l <= ki; l++) {
      const double x = A(f, l);
      lo = dmin(lo, x);
      hi = dmax(hi, x);
    }
    mn[f] = lo;
    mx[f] = hi;
    SL(f, 1) = 0.0;
  }
  double sm[NF], rj[NF], rn[NF];
#pragma unroll
  for (int f = 0; f < NF; f++) sm[f] = 0.0;  // s(j-1), changed
  raw(2, rj);                                // raw s(j)
  for (int j = 2; j <= ki - 1; j++) {
    raw(j + 1, rn);                          // raw s(j+1)
#pragma unroll
    for (int f = 0; f < NF; f++) {
      double sj = rj[f];
      if (sj != 0.0) {
        const double dl = -0.5 * (sj + sm[f]) + A(f, j) - A(f, j - 1);
        const double dr = -0.5 * (rn[f] + sj) + A(f, j + 1) - A(f, j);
        sj = sj + 2.0 * copysign(dmin(fabs(dl), fabs(dr)), dl);
      }
      SL(f, j) = sj;
      sm[f] = sj;
      rj[f] = rn[f];
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
      for (int l = lf; l <= ki; l++) {
        const double pl = PI(l);
        if (pl > zb) {
          lf = l - 1;
          break;
        }
        const double pl1 = PI(l + 1);
        if (pl >= zt && pl1 <= zb) {
          const double q = dmax(pl1 - pl, thin) / dz;
#pragma unroll
          for (int f = 0; f < NF; f++) sok[f] = sok[f] + q * si[f][(int64_t)(l - 1) * S];
        } else {
          const double bot = dmin(pl1, zb), top = dmax(pl, zt);
          const double q = dmax(bot - top, thin) / dz;
          if (METHOD == 0) {
#pragma unroll
            for (int f = 0; f < NF; f++) sok[f] = sok[f] + q * si[f][(int64_t)(l - 1) * S];
          } else {
            const double zc = 0.5 * (bot + top);
            const double qc = (zc - pl) / dmax(pl1 - pl, thin) - 0.5;
#pragma unroll
            for (int f = 0; f < NF; f++)
              sok[f] = sok[f] + q * (si[f][(int64_t)(l - 1) * S] +
                                     qc * sis[S * ((int64_t)(l - 1) * kTrcChunk + f)]);
          }
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
