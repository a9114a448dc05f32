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
  int32_t* bad;                // first column (c0) whose scan would stop, or INT32_MAX
};

HYB_HD int64_t trc_ij(const KTrc& Q, int i, int j) {
  return (int64_t)(i - 1 + Q.nb) + Q.NI * (int64_t)(j - 1 + Q.nb);
}

constexpr double kTrcThin = 1.e-6;  // trcupd.F90:857, 960

constexpr int kTrcChunk = 4;  // fields remapped together

// plmtrcx/plmtrcs (1038-1147) for fields f0..f0+NF-1 of a column, formed
// zone by zone as the scan of plmtrc reaches them.  Campbell eq. 19
// (1134-1143) changes zone j with zone j-1 already changed and zone j+1
// not yet, so the slopes are a chain from the top: zone j needs the
// changed slope of j-1 (sm) and the raw slopes (eq. 15, 1121-1130) of j
// and j+1.  The chain holds the zone averages a(n-1..n+2) and interfaces
// pi(n-1..n+3) around the next zone n = head+1, the changed slope of zone
// head and the running min and max of the averages it has loaded
// (plmtrc 964-967: each field's range over all ki layers).  The scan asks
// for zones in increasing order and may go back only to the zone it holds;
// a scan that starts higher (the rescan after an output whose scan ran to
// the last layer) restarts the chain from zone 1.
template <int NF>
struct TrcChain {
  const double* a[NF];
  const double* p;
  int64_t S;
  int ki;
  int head;  // the zone whose slope sh holds
  int la;    // the deepest zone whose averages are in mn, mx (zones 1..la)
  double am[NF], a0[NF], ap[NF], aq[NF];
  double pm, p0, p1, p2, p3;
  double sm[NF], rj[NF];  // changed slope of zone head (0 at zone 1), raw slope of head+1
  double mn[NF], mx[NF];

  HYB_HDI double A(int f, int l) const { return a[f][(int64_t)(l - 1) * S]; }
  HYB_HDI double PI(int l) const { return p[(int64_t)(l - 1) * S]; }
  // raw slope (Campbell eq. 15) of a zone from its window
  HYB_HDI static double raw(double xm, double x0, double xp, double sm_, double s0_,
                            double sp_) {
    const double rl = 2.0 * s0_ / (sm_ + s0_);
    const double rr = 2.0 * s0_ / (s0_ + sp_);
    const double sl = rl * (x0 - xm);
    const double sr = rr * (xp - x0);
    return (sl * sr > 0.0) ? copysign(dmin(fabs(sl), fabs(sr)), sl) : 0.0;
  }
  // zone 1 (slope 0); the window around zone 2, indices clamped into the
  // column (values beyond it are not used and not ranged)
  HYB_HDI void start(bool first) {
    const int k1 = ki + 1;
    pm = PI(1);
    p0 = PI(2);
    p1 = PI(3 < k1 ? 3 : k1);
    p2 = PI(4 < k1 ? 4 : k1);
    p3 = PI(5 < k1 ? 5 : k1);
    const double tm = dmax(p0 - pm, kTrcThin);
    const double t0 = dmax(p1 - p0, kTrcThin);
    const double tp = dmax(p2 - p1, kTrcThin);
#pragma unroll
    for (int f = 0; f < NF; f++) {
      am[f] = A(f, 1);
      a0[f] = A(f, 2);
      ap[f] = A(f, 3 < ki ? 3 : ki);
      aq[f] = A(f, 4 < ki ? 4 : ki);
      if (first) {
        double lo = am[f], hi = am[f];
        lo = dmin(lo, a0[f]);
        hi = dmax(hi, a0[f]);
        if (ki >= 3) { lo = dmin(lo, ap[f]); hi = dmax(hi, ap[f]); }
        mn[f] = lo;
        mx[f] = hi;
      }
      sm[f] = 0.0;
      rj[f] = (ki >= 3) ? raw(am[f], a0[f], ap[f], tm, t0, tp) : 0.0;  // raw s(2)
    }
    head = 1;
    if (first) la = ki < 3 ? ki : 3;
  }
  HYB_HDI void init(const double* const* af, const double* pc, int64_t s, int k) {
#pragma unroll
    for (int f = 0; f < NF; f++) a[f] = af[f];
    p = pc;
    S = s;
    ki = k;
    start(true);
  }
  // the slope of zone head (s(ki) = 0)
  HYB_HDI double slope(int f) const { return head < ki ? sm[f] : 0.0; }
  // the slope of zone n = head+1, then the window one zone down
  HYB_HDI void advance() {
    const int n = head + 1;
    if (n <= ki - 1) {
      const bool has_n = n + 1 <= ki - 1;  // zone n+1 is an inner zone
      const double t0 = dmax(p1 - p0, kTrcThin), tp = dmax(p2 - p1, kTrcThin),
                   tq = dmax(p3 - p2, kTrcThin);
#pragma unroll
      for (int f = 0; f < NF; f++) {
        const double rn = has_n ? raw(a0[f], ap[f], aq[f], t0, tp, tq) : 0.0;  // raw s(n+1)
        double sj = rj[f];
        if (sj != 0.0) {
          const double dl = -0.5 * (sj + sm[f]) + a0[f] - am[f];
          const double dr = -0.5 * (rn + sj) + ap[f] - a0[f];
          sj = sj + 2.0 * copysign(dmin(fabs(dl), fabs(dr)), dl);
        }
        sm[f] = sj;
        rj[f] = rn;
      }
    }
    head = n;
    const int nl = n + 3;  // the zone entering the window
    const int lx = nl <= ki ? nl : ki;
    const int px = n + 4 <= ki + 1 ? n + 4 : ki + 1;
    pm = p0;
    p0 = p1;
    p1 = p2;
    p2 = p3;
    p3 = PI(px);
    // zone n+2 (loaded by the previous advance, used by the next one's
    // raw slope) joins the range; zone n+3 is still on its way
    const bool fresh = n + 2 <= ki && n + 2 > la;
#pragma unroll
    for (int f = 0; f < NF; f++) {
      am[f] = a0[f];
      a0[f] = ap[f];
      ap[f] = aq[f];
      aq[f] = A(f, lx);
      if (fresh) {
        mn[f] = dmin(mn[f], ap[f]);
        mx[f] = dmax(mx[f], ap[f]);
      }
    }
    if (fresh) la = n + 2;
  }
  // the range of every field over all ki zones (loads what the scans left)
  HYB_HDI void finish_range() {
    for (int l = la + 1; l <= ki; l++) {
#pragma unroll
      for (int f = 0; f < NF; f++) {
        const double v = A(f, l);
        mn[f] = dmin(mn[f], v);
        mx[f] = dmax(mx[f], v);
      }
    }
    la = ki;
  }
};

// pcmtrc (808-909) of fields f0..f0+NF-1 of column c0, in one scan of the
// old layers (the reference's loop over ks inside the loop over l).
// Returns false where the reference stops ('bad lf', 866-869); that
// output is NaN.
template <int NF>
HYB_HDI bool trc_pcm_fields(const KTrc& Q, int64_t c0, int f0) {
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
  auto PI = [&](int l) -> double { return pi[(int64_t)(l - 1) * S]; };
  double prev[NF], sok[NF];
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
      double cv[NF];
#pragma unroll
      for (int f = 0; f < NF; f++) cv[f] = si[f][(int64_t)(lf - 1) * S];
      for (int l = lf; l <= ki; l++) {
        const double pl = pl1;
        if (pl > zb) {
          lf = l - 1;
          break;
        }
        pl1 = pl2;
        const bool more = l + 1 <= ki;
        double nv[NF];
        pl2 = more ? PI(l + 2) : 0.0;
#pragma unroll
        for (int f = 0; f < NF; f++) nv[f] = more ? si[f][(int64_t)l * S] : 0.0;
        if (pl >= zt && pl1 <= zb) {
          const double q = dmax(pl1 - pl, thin) / dz;
#pragma unroll
          for (int f = 0; f < NF; f++) sok[f] = sok[f] + q * cv[f];
        } else {
          const double bot = dmin(pl1, zb), top = dmax(pl, zt);
          const double q = dmax(bot - top, thin) / dz;
#pragma unroll
          for (int f = 0; f < NF; f++) sok[f] = sok[f] + q * cv[f];
        }
#pragma unroll
        for (int f = 0; f < NF; f++) cv[f] = nv[f];
      }
#pragma unroll
      for (int f = 0; f < NF; f++) prev[f] = sok[f];
    }
#pragma unroll
    for (int f = 0; f < NF; f++) so[f][(int64_t)(k - 1) * S] = prev[f];
  }
  return ok;
}

// plmtrc (911-1037) of fields f0..f0+NF-1 of column c0, the slopes formed
// by a TrcChain as the scan reaches each layer (no slope array).  The
// clamp of each output to its field's range over all ki layers (1032-
// 1033) is applied at once where the output lies inside the range of the
// layers loaded so far (the full range can only be wider, so the clamp
// leaves it as it is); the other outputs are written unclamped, flagged,
// and clamped at the end with the full range.  Returns false where the
// reference stops ('bad lf', 989-992); that output is NaN.
template <int NF>
HYB_HDI bool trc_plm_fields(const KTrc& Q, int64_t c0, int f0) {
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
  TrcChain<NF> ch;
  ch.init(si, pi, S, ki);
  double prev[NF], sok[NF];
  uint64_t unc[NF];   // outputs written unclamped
  bool punc[NF];      // prev is unclamped
#pragma unroll
  for (int f = 0; f < NF; f++) {
    prev[f] = si[f][0];  // the reference reads so(0) here
    unc[f] = 0ull;
    punc[f] = false;
  }
  const double zx = pi[(int64_t)ki * S];
  int lf = 1;
  double zb = po[0];
  bool ok = true;
  for (int k = 1; k <= ko; k++) {
    const double zt = zb;
    zb = po[(int64_t)k * S];
    if (zb - zt < thin || zt >= zx) {
      // thin or bottomed: the output above
    } else if (pi[(int64_t)(lf - 1) * S] > zt) {
      ok = false;
#pragma unroll
      for (int f = 0; f < NF; f++) {
        prev[f] = NAN;
        punc[f] = false;
      }
    } else {
      const double dz = zb - zt;
#pragma unroll
      for (int f = 0; f < NF; f++) sok[f] = 0.0;
      if (ch.head > lf) ch.start(false);
      while (ch.head < lf) ch.advance();
      // layer l = ch.head: pi(l) = ch.pm, pi(l+1) = ch.p0, si(l) = ch.am,
      // its slope ch.slope()
      for (int l = lf;; l++) {
        const double pl = ch.pm;
        if (pl > zb) {
          lf = l - 1;
          break;
        }
        const double pl1 = ch.p0;
        if (pl >= zt && pl1 <= zb) {
          const double q = dmax(pl1 - pl, thin) / dz;
#pragma unroll
          for (int f = 0; f < NF; f++) sok[f] = sok[f] + q * ch.am[f];
        } else {
          const double bot = dmin(pl1, zb), top = dmax(pl, zt);
          const double q = dmax(bot - top, thin) / dz;
          const double zc = 0.5 * (bot + top);
          const double qc = (zc - pl) / dmax(pl1 - pl, thin) - 0.5;
#pragma unroll
          for (int f = 0; f < NF; f++) sok[f] = sok[f] + q * (ch.am[f] + qc * ch.slope(f));
        }
        if (l == ki) break;  // ran to the last layer: lf stays
        if (pl1 > zb) {      // layer l+1 starts below the output
          lf = l;
          break;
        }
        ch.advance();
      }
#pragma unroll
      for (int f = 0; f < NF; f++) {
        const double v = sok[f];
        prev[f] = v;
        punc[f] = !(v >= ch.mn[f] && v <= ch.mx[f]);
      }
    }
#pragma unroll
    for (int f = 0; f < NF; f++) {
      so[f][(int64_t)(k - 1) * S] = prev[f];
      if (punc[f]) unc[f] |= 1ull << (k - 1);
    }
  }
  bool any = false;
#pragma unroll
  for (int f = 0; f < NF; f++) any |= unc[f] != 0ull;
  if (any) {
    ch.finish_range();
#pragma unroll
    for (int f = 0; f < NF; f++) {
      uint64_t m = unc[f];
      while (m) {
        const int k = __builtin_ctzll(m) + 1;
        m &= m - 1;
        double v = so[f][(int64_t)(k - 1) * S];
        v = dmax(ch.mn[f], v);
        v = dmin(ch.mx[f], v);
        so[f][(int64_t)(k - 1) * S] = v;
      }
    }
  }
  return ok;
}

// the fields of column c0 in chunks of CHUNK, each chunk one scan
template <int METHOD, int CHUNK>
HYB_HDI bool trc_column_m(const KTrc& Q, int64_t c0) {
  bool ok = true;
  for (int f0 = 0; f0 < Q.ks; f0 += CHUNK) {
    const int nf = Q.ks - f0 < CHUNK ? Q.ks - f0 : CHUNK;
    bool r;
    if (METHOD == 0) {
      if (CHUNK >= 4 && nf == 4) r = trc_pcm_fields<(CHUNK >= 4 ? 4 : 1)>(Q, c0, f0);
      else if (CHUNK >= 3 && nf == 3) r = trc_pcm_fields<(CHUNK >= 3 ? 3 : 1)>(Q, c0, f0);
      else if (CHUNK >= 2 && nf == 2) r = trc_pcm_fields<(CHUNK >= 2 ? 2 : 1)>(Q, c0, f0);
      else r = trc_pcm_fields<1>(Q, c0, f0);
    } else {
      if (CHUNK >= 4 && nf == 4) r = trc_plm_fields<(CHUNK >= 4 ? 4 : 1)>(Q, c0, f0);
      else if (CHUNK >= 3 && nf == 3) r = trc_plm_fields<(CHUNK >= 3 ? 3 : 1)>(Q, c0, f0);
      else if (CHUNK >= 2 && nf == 2) r = trc_plm_fields<(CHUNK >= 2 ? 2 : 1)>(Q, c0, f0);
      else r = trc_plm_fields<1>(Q, c0, f0);
    }
    ok = r && ok;
  }
  return ok;
}

HYB_HDI bool trc_column(const KTrc& Q, int64_t c0) {
  return Q.method == 0 ? trc_column_m<0, kTrcChunk>(Q, c0) : trc_column_m<1, kTrcChunk>(Q, c0);
}

}  // namespace hyb
