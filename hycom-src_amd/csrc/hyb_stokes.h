// hyb_stokes.h -- Stokes drift onto the model layers (mod_stokes.F90:650-705,
// stokes_vertical_j), shared by the host checks and the gfx950 kernel.
// One sea p-column per lane: walk_core's PLM remap from the nsdzi fixed
// layers onto the kdm model layers, reading both snapshots of a level and
// blending them when the level enters the walk's window.
#pragma once

#include "hyb_stream.h"

namespace hyb {

struct KStokes {
  int32_t ii, jj, kk, nsdzi, nb, n;
  int64_t NI, NIJ;
  double ws0, ws1;
  double sdzi[kKMax];  // interface depths, Pa
  const double *usdz, *vsdz, *dp, *p;
  const int32_t* ip;
  double *usdp, *vsdp;
};

HYB_HD int64_t stokes_ij(const KStokes& Q, int i, int j) {
  return (int64_t)(i - 1 + Q.nb) + Q.NI * (int64_t)(j - 1 + Q.nb);
}

// Old interfaces pres(1) = 0, pres(k+1) = min(sdzi(k), p(kk+1))
// (mod_stokes.F90:675-681), for walk_core.
struct PiStokes {
  static constexpr bool kRawThk = false;
  const double* sdzi;
  double pbot;
  HYB_HD double first() const { return 0.0; }
  HYB_HD double raw(int x) const { return sdzi[x - 1]; }
  HYB_HD double comb(double, double r) const { return dmin(r, pbot); }
  HYB_HD double nextfrom(int x, double pix) const { return comb(pix, raw(x)); }
  HYB_HD double thk(int) const { return 0.0; }
};

// New interfaces prsf(1) = 0, prsf(k+1) = min(prsf(k) + dp(k), p(kk+1))
// (mod_stokes.F90:682-685), summed as the walk asks for them with the next
// thickness loaded ahead.
struct PoStokes {
  const double* dp;  // level 1 of dp(:,:,:,n) of the column
  int64_t S;
  double pbot;
  int k1;            // kk+1
  int k = 1;         // level of cur
  double cur = 0.0, nxt = 0.0;
  HYB_HD PoStokes(const double* d, int64_t s, double pb, int kk)
      : dp(d), S(s), pbot(pb), k1(kk + 1) {
    nxt = dp[0];
  }
  HYB_HD double operator()(int kq) {
    if (kq < k) {
      k = 1;
      cur = 0.0;
      nxt = dp[0];
    }
    while (k < kq) {
      cur = dmin(cur + nxt, pbot);
      k++;
      nxt = (k < k1) ? dp[(int64_t)(k - 1) * S] : 0.0;
    }
    return cur;
  }
};

// s1d(k) = ws0*usdz(i,j,1,k) + ws1*usdz(i,j,2,k) (mod_stokes.F90:676-679):
// both words are loaded when the walk asks for level k, and blended when
// the value is used
struct StokesBlend {
  double x, y, w0, w1;
  HYB_HD operator double() const { return w0 * x + w1 * y; }
};
struct StokesRow {  // one field of one column; level stride 2S, snapshot 2 at +S
  const double* a;
  int64_t S;
  double w0, w1;
  HYB_HD StokesBlend operator[](int64_t i) const {
    const int64_t o = 2 * i;  // i = (k-1)*S as walk_core forms it
    return StokesBlend{a[o], a[o + S], w0, w1};
  }
};
struct StokesFields {
  StokesRow row[2];
  HYB_HD const StokesRow& operator[](int f) const { return row[f]; }
};

// hybgen_plm_coefs on dpi and hybgen_plm_remap of the two blended fields
// (mod_stokes.F90:686-694), results to usdp, vsdp (695-698)
template <int PD>
HYB_HDI void stokes_column(const KStokes& Q, int64_t c0) {
  const int64_t S = Q.NIJ;
  const double pbot = Q.p[c0 + S * (int64_t)Q.kk];
  PiStokes pi{Q.sdzi, pbot};
  PoStokes po(Q.dp + c0 + S * (int64_t)Q.kk * (Q.n - 1), S, pbot, Q.kk);
  const StokesFields fp{{StokesRow{Q.usdz + c0, S, Q.ws0, Q.ws1},
                         StokesRow{Q.vsdz + c0, S, Q.ws0, Q.ws1}}};
  double* op[2] = {Q.usdp + c0, Q.vsdp + c0};
  FieldSink<2> sink{op, 2, nullptr, 0.0, S};
  const double zx = dmin(Q.sdzi[Q.nsdzi - 1], pbot);
  walk_core<1, 2, PD, false, true>(Q.nsdzi, zx, pi, po, sink, fp, 2, Q.nsdzi, 0ull, S, Q.kk);
}

}  // namespace hyb
