// hyb_tile.h -- the tiled p-column remap (hybgen.F90:1530-1569 with
// hybgen_ppm_coefs 2587-2706 and hybgen_ppm_remap 2708-2843).
//
// A workgroup of 256 threads remaps a tile of 16 neighbouring columns of
// one row: one 128-byte line of every level.  Each field passes through
// LDS once.  The work of the column walk is spread over the layers:
// thread (c, r) owns column c of the tile and layers k = r+1, r+17, r+33,
// ..., so every step below is one pass over the layers with no serial
// dependence between them:
//   geometry   per output k: zt, zb, the layers lt and lb that hold them
//              (a binary search of the old interfaces), the piece weights;
//              once per column, shared by all fields;
//   as(j)      slopes (Colella-Woodward eq. 1.8);
//   al(j)      first-guess edge values (eq. 1.6);
//   limit      monotone edge values of each layer (eq. 1.10);
//   outputs    the overlap integral of every output layer, clamped;
//   store      thin outputs copy the last thick one above; th3d from T, S.
// The operations of each value and their order are the reference's, so
// the result is bitwise the column walk's.
//
// The reference finds lb(k) by scanning from lb(k-1).  With the old
// interfaces pi(2..kk) non-decreasing and the output bottoms zb(k)
// non-decreasing (zb(0) = max(po(1), pi(1))), that scan stops at the
// first l with pi(l+1) >= zb(k), which a binary search finds directly.
// A column that breaks either order (or holds a NaN there) replays the
// scan instead.
//
// The same code runs on the host (TileTeamHost: the 256 threads one after
// another between barriers), so the CPU suite checks it bit for bit.
#pragma once

#include "hyb_stream.h"

namespace hyb {

constexpr int kTileW = 16;                // columns per tile
constexpr int kTileT = 256;               // threads per tile
constexpr int kTileR = kTileT / kTileW;   // layer slots

// LDS of a tile, in doubles with row stride kTileW, for kk = K
struct TileLayout {
  int K;
  HYB_HD int pi(int l) const { return (l - 1) * kTileW; }             // pres(l), l = 1..K+1
  // grid metrics of hybgen_ppm_coefs (2621-2633), once per tile
  HYB_HD int dp(int l) const { return (K + l) * kTileW; }             // dp(l) = max(dprs, thin)
  HYB_HD int qj(int l) const { return (2 * K + l) * kTileW; }         // qdpjp(l)
  HYB_HD int q2j(int l) const { return (3 * K + l) * kTileW; }        // qdp2jp(l)
  HYB_HD int qj2(int l) const { return (4 * K + l) * kTileW; }        // qdpj2p(l)
  HYB_HD int q3(int l) const { return (5 * K + l) * kTileW; }         // dpq3(l)
  HYB_HD int q4(int l) const { return (6 * K + l) * kTileW; }         // qdp4(l)
  HYB_HD int s(int l) const { return (7 * K + l) * kTileW; }          // s(l) with the fill
  HYB_HD int a(int l) const { return (8 * K + l) * kTileW; }          // al(l) limited
  HYB_HD int b(int l) const { return (9 * K + l) * kTileW; }          // raw al(l), then so(l)
  HYB_HD int c(int l) const { return (10 * K + l) * kTileW; }         // ar(l) limited
  HYB_HD int mn(int r) const { return (11 * K + 1 + r) * kTileW; }    // partial minima
  HYB_HD int mx(int r) const { return (11 * K + 1 + kTileR + r) * kTileW; }
  HYB_HD int lo() const { return (11 * K + 1 + 2 * kTileR) * kTileW; }  // si_min
  HYB_HD int hi() const { return lo() + kTileW; }                       // si_max
  HYB_HD int words() const { return hi() + kTileW; }
  // two uint64 per column follow: irregular-column flag, thick-output mask
  HYB_HD size_t bytes() const { return sizeof(double) * words() + 2 * sizeof(uint64_t) * kTileW; }
};

// registers of one thread: Q output layers (k = r + 1 + kTileR*q)
template <int Q>
struct TileThread {
  int c, r, kp;
  bool act;                 // the column is remapped here
  uint64_t lcm;
  int64_t c0;
  int lt[Q], lb[Q], src[Q];  // src: -1 thick output; else its source output (0: s(1))
  double zt[Q], zb[Q];
  double qt0[Q], qt1[Q], qt2[Q], qb0[Q], qb1[Q], qb2[Q], dz[Q];
  double pre[Q];            // the next field, loaded a field ahead
  double tv[Q];             // outputs of T, for th3d
};

HYB_HD void tile_or(uint64_t* a, uint64_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
  atomicOr((unsigned long long*)a, (unsigned long long)v);
#else
  *a |= v;
#endif
}

// the 256 threads run one after another between barriers (host checks)
template <class ST>
struct TileTeamHost {
  ST st[kTileT];
  template <class F>
  void each(F&& f) {
    for (int t = 0; t < kTileT; t++) f(st[t], t);
  }
  template <class F>
  bool each_or(F&& f) {
    bool any = false;
    for (int t = 0; t < kTileT; t++) any = f(st[t], t) || any;
    return any;
  }
};

// Level-1 pointer of field g (0 T, 1 S, 2.. tracers) of column c0
HYB_HDI double* tile_field(const View& V, int64_t c0, int n, int g) {
  if (g == 0) return V.temp + c0 + lev4(V, 1, n);
  if (g == 1) return V.saln + c0 + lev4(V, 1, n);
  return V.tracer + c0 + levt(V, 1, n, g - 1);
}

// PPM remap of the tile of row j whose first column is i0 (Q*kTileR >= kk+1)
template <int Q, class TEAM>
HYB_HDI void tile_remap_ppm(const KParams& P, const View& V, const Scratch& W, int n, int j,
                            int i0, double* lds, TEAM& team) {
  using ST = TileThread<Q>;
  const int K = P.kk;
  const int64_t S = V.NIJ;
  const TileLayout L{K};
  uint64_t* irr = (uint64_t*)(lds + L.words());
  uint64_t* ntm = irr + kTileW;
  const int nfields = 2 + P.ntracr;
  const double thin = kDpthin;

  // ---- columns, old interfaces, new interfaces of each output, field 0
  const bool any = team.each_or([&](ST& st, int t) -> bool {
    st.c = t % kTileW;
    st.r = t / kTileW;
    st.act = false;
    st.kp = 1;
    st.lcm = 0;
    st.c0 = 0;
    const int i = i0 + st.c;
    if (st.r == 0) {
      irr[st.c] = 0;
      ntm[st.c] = 0;
    }
    if (i < 1 || i > P.ii) return false;
    const int64_t c0 = ij_of(V, i, j);
    st.c0 = c0;
    if (V.ip[c0] == 0) return false;
    const int kpw = W.kp[c0];
    if (kpw & kUnsafe) return false;  // remapped by remap_fallback
    st.act = true;
    st.kp = kpw & 0xffff;
    st.lcm = W.lcm[c0];
    const double* f0 = tile_field(V, c0, n, 0);
#pragma unroll
    for (int q = 0; q < Q; q++) {
      const int k = st.r + 1 + kTileR * q;
      if (k <= K + 1) lds[L.pi(k) + st.c] = W.pres[c0 + (int64_t)(k - 1) * S];
      if (k <= K) {
        st.zt[q] = V.p[c0 + (int64_t)(k - 1) * S];  // po(k)
        st.zb[q] = V.p[c0 + (int64_t)k * S];        // po(k+1)
        st.pre[q] = f0[(int64_t)((k < st.kp ? k : st.kp) - 1) * S];
      }
    }
    return true;
  });
  if (!any) return;
  auto PI = [&](const ST& st, int l) -> double { return lds[L.pi(l) + st.c]; };
  auto DP = [&](const ST& st, int l) -> double { return dmax(PI(st, l + 1) - PI(st, l), thin); };

  // ---- output tops and bottoms (2739-2741), thin outputs (2747)
  team.each([&](ST& st, int) {
    if (!st.act) return;
    const double zx = PI(st, K + 1);
    bool bad = false;
    uint64_t thick = 0;
#pragma unroll
    for (int q = 0; q < Q; q++) {
      const int k = st.r + 1 + kTileR * q;
      if (k > K) break;
      const double zb = dmin(st.zb[q], zx);
      const double zt = (k == 1) ? dmax(st.zt[q], PI(st, 1)) : dmin(st.zt[q], zx);
      st.zt[q] = zt;
      st.zb[q] = zb;
      if (!(zb >= zt)) bad = true;
      if (k >= 2 && k <= K - 1 && !(PI(st, k + 1) >= PI(st, k))) bad = true;
      if (!((zb - zt <= thin) || (zt >= zx))) thick |= 1ull << (k - 1);
      // grid metrics of layer k (dpjp(k) = dp(k) + dp(k+1), ...)
      const double d0 = DP(st, k);
      lds[L.dp(k) + st.c] = d0;
      if (k <= K - 1) {
        const double dq = DP(st, k + 1);
        const double dpjp = d0 + dq;
        lds[L.qj(k) + st.c] = 1.0 / dpjp;
        lds[L.q2j(k) + st.c] = 1.0 / (d0 + dpjp);
        lds[L.qj2(k) + st.c] = 1.0 / (dpjp + dq);
        if (k >= 2) lds[L.q3(k) + st.c] = d0 / (DP(st, k - 1) + dpjp);
        if (k >= 3) lds[L.q4(k) + st.c] = 1.0 / ((DP(st, k - 2) + DP(st, k - 1)) + dpjp);
      }
    }
    if (bad) tile_or(&irr[st.c], 1ull);
    if (thick) tile_or(&ntm[st.c], thick);
  });

  // ---- lt, lb and the piece weights of every thick output (2742-2744,
  // 2766-2805); the source of every thin one
  team.each([&](ST& st, int) {
    if (!st.act) return;
    const bool irregular = irr[st.c] != 0;
    const uint64_t thick = ntm[st.c];
    const double zx = PI(st, K + 1);
    // 1 + #{l in 1..kk-1 : pi(l+1) < z}: pi(2..kk) is sorted here
    auto first_below = [&](double z) -> int {
      int lo = 0;
#pragma unroll
      for (int step = 32; step >= 1; step >>= 1)
        if (lo + step <= K - 1 && PI(st, lo + step + 1) < z) lo += step;
      return lo + 1;
    };
#pragma unroll
    for (int q = 0; q < Q; q++) {
      const int k = st.r + 1 + kTileR * q;
      if (k > K) break;
      if (!((thick >> (k - 1)) & 1ull)) {
        const uint64_t above = thick & ((1ull << (k - 1)) - 1);
        st.src[q] = above ? 64 - __builtin_clzll(above) : 0;
        continue;
      }
      st.src[q] = -1;
      const double zt = st.zt[q], zb = st.zb[q];
      int lt, lb;
      if (!irregular) {
        lt = first_below(zt);
        lb = first_below(zb);
      } else {  // the reference's scan from the surface (2729-2744)
        int l = 1;
        double z = dmax(V.p[st.c0], PI(st, 1));
        while (l < K && PI(st, l + 1) < z) l++;
        lt = l;
        for (int kk = 1; kk <= k; kk++) {
          z = dmin(V.p[st.c0 + (int64_t)kk * S], zx);
          lt = l;
          while (l < K && PI(st, l + 1) < z) l++;
        }
        lb = l;
      }
      st.lt[q] = lt;
      st.lb[q] = lb;
      const double xt = (zt - PI(st, lt)) / dmax(PI(st, lt + 1) - PI(st, lt), thin);
      const double xb = (zb - PI(st, lb)) / dmax(PI(st, lb + 1) - PI(st, lb), thin);
      if (lt != lb) {
        st.qt0[q] = (1.0 - xt);
        st.qt1[q] = (1.0 - xt * xt) * 0.5;
        st.qt2[q] = (1.0 - xt * xt * xt) / 3.0;
        st.qb0[q] = xb;
        st.qb1[q] = xb * xb * 0.5;
        st.qb2[q] = xb * xb * xb / 3.0;
      } else {
        st.qt0[q] = (xb - xt);
        st.qt1[q] = (xb * xb - xt * xt) * 0.5;
        st.qt2[q] = (xb * xb * xb - xt * xt * xt) / 3.0;
      }
      st.dz[q] = zb - zt;
    }
  });

  for (int g = 0; g < nfields; g++) {
    // ---- s(l) with the massless fill (972-996); partial min/max; the
    // next field's loads issued
    team.each([&](ST& st, int) {
      if (!st.act) return;
      double mn = 0.0, mx = 0.0;
      bool first = true;
#pragma unroll
      for (int q = 0; q < Q; q++) {
        const int k = st.r + 1 + kTileR * q;
        if (k > K) break;
        const double x = st.pre[q];
        lds[L.s(k) + st.c] = x;
        if (first) {
          mn = x;
          mx = x;
          first = false;
        } else {
          mn = dmin(mn, x);
          mx = dmax(mx, x);
        }
      }
      if (first) {  // no layer in this slot (kk < 16): neutral partials
        mn = HUGE_VAL;
        mx = -HUGE_VAL;
      }
      lds[L.mn(st.r) + st.c] = mn;
      lds[L.mx(st.r) + st.c] = mx;
      if (g + 1 < nfields) {
        const double* fn = tile_field(V, st.c0, n, g + 1);
#pragma unroll
        for (int q = 0; q < Q; q++) {
          const int k = st.r + 1 + kTileR * q;
          if (k > K) break;
          st.pre[q] = fn[(int64_t)((k < st.kp ? k : st.kp) - 1) * S];
        }
      }
    });
    auto SV = [&](const ST& st, int l) -> double { return lds[L.s(l) + st.c]; };
    auto lcmb = [&](const ST& st, int l) -> bool { return (st.lcm >> (l - 1)) & 1ull; };

    auto DPL = [&](const ST& st, int l) -> double { return lds[L.dp(l) + st.c]; };
    // slope as(j) (2646-2660), from the metrics of the tile
    auto slope = [&](const ST& st, int jl) -> double {
      double as = 0.0;
      if (jl >= 2 && jl <= K - 1 && !(lcmb(st, jl) || DPL(st, jl) <= thin)) {
        const double slj = SV(st, jl) - SV(st, jl - 1);
        const double srj = SV(st, jl + 1) - SV(st, jl);
        if (slj * srj > 0.0) {
          const double dm = DPL(st, jl - 1), d0 = DPL(st, jl), dq = DPL(st, jl + 1);
          const double dp2jp_m = dm + (dm + d0), dpj2p_j = (d0 + dq) + dq;
          const double scj = lds[L.q3(jl) + st.c] *
                             (dp2jp_m * srj * lds[L.qj(jl) + st.c] +
                              dpj2p_j * slj * lds[L.qj(jl - 1) + st.c]);
          as = gsign(gmin3(gabs(2.0 * slj), gabs(scj), gabs(2.0 * srj)), scj);
        }
      }
      return as;
    };

    // ---- si_min, si_max (2727-2731); first-guess edge values al(j)
    // (2662-2676) with the slopes as(j-1), as(j)
    team.each([&](ST& st, int) {
      if (!st.act) return;
      if (st.r == 0 || st.r == 1) {
        double x = lds[(st.r == 0 ? L.mn(0) : L.mx(0)) + st.c];
        for (int r = 1; r < kTileR; r++) {
          const double y = lds[(st.r == 0 ? L.mn(r) : L.mx(r)) + st.c];
          x = (st.r == 0) ? dmin(x, y) : dmax(x, y);
        }
        lds[(st.r == 0 ? L.lo() : L.hi()) + st.c] = x;
      }
#pragma unroll
      for (int q = 0; q < Q; q++) {
        const int jl = st.r + 1 + kTileR * q;
        if (jl > K) break;
        double al;
        if (jl == K) {
          al = SV(st, K);
        } else if (jl <= 2) {
          al = SV(st, 1);
        } else {
          const double d2 = DPL(st, jl - 2), dm = DPL(st, jl - 1), d0 = DPL(st, jl),
                       dq = DPL(st, jl + 1);
          const double dpjp_m = dm + d0, dpjp_j = d0 + dq, dpjp_m2 = d2 + dm;
          const double qdpjp_m = lds[L.qj(jl - 1) + st.c];
          const double qdp2jp_m = lds[L.q2j(jl - 1) + st.c], qdpj2p_m = lds[L.qj2(jl - 1) + st.c];
          const double qdp4 = lds[L.q4(jl) + st.c];
          const double s1 = SV(st, jl - 1), s2 = SV(st, jl);
          al = s1 + dm * (s2 - s1) * qdpjp_m +
               qdp4 * (2.0 * d0 * dm * qdpjp_m * (s2 - s1) *
                           (dpjp_m2 * qdp2jp_m - dpjp_j * qdpj2p_m) -
                       dm * slope(st, jl) * dpjp_m2 * qdp2jp_m +
                       d0 * slope(st, jl - 1) * dpjp_j * qdpj2p_m);
        }
        lds[L.b(jl) + st.c] = al;
      }
    });

    // ---- monotone edge values of each layer (2677-2693)
    team.each([&](ST& st, int) {
      if (!st.act) return;
#pragma unroll
      for (int q = 0; q < Q; q++) {
        const int l = st.r + 1 + kTileR * q;
        if (l > K) break;
        const double s1 = SV(st, l);
        double al = lds[L.b(l) + st.c], ar;
        if (l == K - 1 || l == K) ar = SV(st, K);
        else if (l == 1) ar = s1;
        else ar = lds[L.b(l + 1) + st.c];
        if (l >= 2 && l <= K - 1) {
          if (lcmb(st, l) || DPL(st, l) <= thin) {
            al = s1;
            ar = s1;
          } else if ((SV(st, l + 1) - s1) * (s1 - SV(st, l - 1)) <= 0.0) {
            al = s1;
            ar = s1;
          } else {
            const double da = ar - al;
            const double a6 = 6.0 * s1 - 3.0 * (al + ar);
            if (da * a6 > da * da) {
              al = 3.0 * s1 - 2.0 * ar;
            } else if (da * a6 < -(da * da)) {
              ar = 3.0 * s1 - 2.0 * al;
            }
          }
        }
        lds[L.a(l) + st.c] = al;
        lds[L.c(l) + st.c] = ar;
      }
    });

    // ---- thick outputs: overlap integrals (2755-2840), clamped
    team.each([&](ST& st, int) {
      if (!st.act) return;
      const double lo = lds[L.lo() + st.c], hi = lds[L.hi() + st.c];
      // piece of layer x: ci1*(w0) + (ci2+ci3)*w1 - ci3*w2 relative to o
      auto piece = [&](int x, double o, double w0, double w1, double w2) -> double {
        const double al = lds[L.a(x) + st.c], ar = lds[L.c(x) + st.c];
        double c2 = 0.0, c3 = 0.0;
        if (al != ar) {
          c2 = ar - al;
          c3 = 6.0 * SV(st, x) - 3.0 * (al + ar);
        }
        return (PI(st, x + 1) - PI(st, x)) * ((al - o) * w0 + (c2 + c3) * w1 - c3 * w2);
      };
#pragma unroll
      for (int q = 0; q < Q; q++) {
        const int k = st.r + 1 + kTileR * q;
        if (k > K) break;
        if (st.src[q] >= 0) continue;
        const int lt = st.lt[q], lb = st.lb[q];
        double v;
        if (lt == lb) {
          const double o = SV(st, lt);
          const double sz = piece(lt, o, st.qt0[q], st.qt1[q], st.qt2[q]);
          v = o + sz / st.dz[q];
        } else {
          const double o = SV(st, (lt + lb) / 2);
          double sz = piece(lt, o, st.qt0[q], st.qt1[q], st.qt2[q]);
          for (int l = lt + 1; l <= lb - 1; l++)
            sz = sz + (PI(st, l + 1) - PI(st, l)) * (SV(st, l) - o);
          sz = sz + piece(lb, o, st.qb0[q], st.qb1[q], st.qb2[q]);
          v = o + sz / st.dz[q];
        }
        v = (v < lo) ? lo : v;
        v = (v > hi) ? hi : v;
        lds[L.b(k) + st.c] = v;
      }
    });

    // ---- thin outputs take the output above (2747-2752); stores; th3d
    // (1549-1556)
    team.each([&](ST& st, int) {
      if (!st.act) return;
      double* fo = tile_field(V, st.c0, n, g);
      double* th = V.th3d + st.c0 + lev4(V, 1, n);
#pragma unroll
      for (int q = 0; q < Q; q++) {
        const int k = st.r + 1 + kTileR * q;
        if (k > K) break;
        const int sk = st.src[q] < 0 ? k : st.src[q];
        const double v = sk > 0 ? lds[L.b(sk) + st.c] : SV(st, 1);
        fo[(int64_t)(k - 1) * S] = v;
        if (g == 0) st.tv[q] = v;
        if (g == 1) th[(int64_t)(k - 1) * S] = sig(st.tv[q], v) - P.thbase;
      }
    });
  }
}

}  // namespace hyb
