// Host build of the device column code (hycom-src_amd/csrc/hyb_column.h),
// run in the same phase order as the kernel launcher, so the CPU suite can
// check the kernels' own arithmetic against the oracle without a GPU.
// Test infrastructure only: nothing in the library calls this.
#include "hyb_column.h"
#include "hyb_params_conv.h"

using namespace hyb;

static uint64_t row_lcm(const KParams& P) {
  uint64_t m = 0;
  if (!P.isopcm)
    for (int k = P.nhybrd + 1; k <= P.kk; k++) m |= 1ull << (k - 1);
  return m;
}

template <int KC>
static void run(const KParams& P, const View& V, int m, int n, int raflag) {
  const int ii = P.ii, jj = P.jj, kk = P.kk;
  const int itr[3] = {P.itr801, P.itr802, P.itr803};
  double* src[3] = {V.dp, V.temp, V.saln};
  for (int d = 0; d < 3; d++)
    if (itr[d] > 0)
      for (int j = 1; j <= jj; j++)
        for (int i = 1; i <= ii; i++)
          for (int k = 1; k <= kk; k++)
            V.tracer[ij_of(V, i, j) + levt(V, k, n, itr[d])] = src[d][ij_of(V, i, j) + lev4(V, k, n)];
#pragma omp parallel for schedule(dynamic, 1)
  for (int j = 1; j <= jj; j++)
    for (int i = 1; i <= ii; i++)
      if (V.ip[ij_of(V, i, j)]) {
        uint64_t l = row_lcm(P);
        column_a<KC>(P, V, i, j, n, l);
      }
  for (int d = 0; d < 3; d++)
    if (itr[d] > 0) {
      double q = d == 0 ? P.diag_q801 : P.diag_q80x;
      for (int j = 1; j <= jj; j++)
        for (int i = 1; i <= ii; i++)
          if (V.ip[ij_of(V, i, j)])
            for (int k = 1; k <= kk; k++) {
              int64_t a = ij_of(V, i, j) + levt(V, k, n, itr[d]);
              V.tracer[a] = q * (src[d][ij_of(V, i, j) + lev4(V, k, n)] - V.tracer[a]);
            }
    }
  for (int pass = 0; pass <= (raflag ? 1 : 0); pass++) {
    int nl = pass == 0 ? n : m;
    if (pass == 1) {
#pragma omp parallel for schedule(dynamic, 1)
      for (int j = 1; j <= jj; j++)
        for (int i = 1; i <= ii; i++)
          if (V.ip[ij_of(V, i, j)]) column_ra<KC>(P, V, i, j, m, n, row_lcm(P));
    }
    for (int k = 2; k <= kk + 1; k++)
      for (int i = 1; i <= ii; i++) p_halo_jrow(P, V, i, k);
    for (int k = 2; k <= kk + 1; k++)
      for (int j = 0; j <= jj + 1; j++) p_halo_icol(P, V, j, k);
#pragma omp parallel for schedule(dynamic, 1)
    for (int j = 1; j <= jj; j++)
      for (int isv = 0; isv <= 1; isv++)
        for (int i = 1; i <= ii; i++)
          if ((isv ? V.iv : V.iu)[ij_of(V, i, j)]) point_c<KC>(P, V, i, j, nl, isv != 0);
  }
}

extern "C" int hostcheck_hybgen(const hyb_params* P, const hyb_fields* F, int m, int n,
                                int raflag) {
  KParams K = kparams_of(P);
  View V = view_of(P, F);
  if (K.kk <= 24) run<24>(K, V, m, n, raflag);
  else if (K.kk <= 32) run<32>(K, V, m, n, raflag);
  else if (K.kk <= 48) run<48>(K, V, m, n, raflag);
  else run<64>(K, V, m, n, raflag);
  return 0;
}

// ---- streaming kernels (hyb_stream.h), same phase order ----
#include "hyb_stream.h"
#include <vector>

#include <memory>

static int64_t g_unsafe_cols = 0;
static int g_uv_recompute = 0;  // 1: the read-ahead regrid, remap and u/v variants
static int g_split_ns = 7;      // interface levels the split remap stages (the rest from p)

static int64_t g_uv_fallback = 0;
static int64_t g_ra_fallback = 0;

static void run_stream(const KParams& P, const View& V, int m, int n, int raflag) {
  const int ii = P.ii, jj = P.jj, kk = P.kk;
  std::vector<double> pres((size_t)V.NIJ * (kk + 1));
  std::vector<int32_t> kpv((size_t)V.NIJ);
  std::vector<uint64_t> lcmv((size_t)V.NIJ);
  std::vector<double> rasv(raflag ? (size_t)V.NIJ * kk * 3 : 0);
  Scratch W{pres.data(), kpv.data(), lcmv.data(), nullptr, raflag ? rasv.data() : nullptr};
  const int itr[3] = {P.itr801, P.itr802, P.itr803};
  double* src[3] = {V.dp, V.temp, V.saln};
  for (int d = 0; d < 3; d++)
    if (itr[d] > 0)
      for (int j = 1; j <= jj; j++)
        for (int i = 1; i <= ii; i++)
          for (int k = 1; k <= kk; k++)
            V.tracer[ij_of(V, i, j) + levt(V, k, n, itr[d])] = src[d][ij_of(V, i, j) + lev4(V, k, n)];
#pragma omp parallel for schedule(dynamic, 1)
  for (int j = 1; j <= jj; j++)
    for (int i = 1; i <= ii; i++)
      if (V.ip[ij_of(V, i, j)]) {
        if (g_uv_recompute) regrid_stream<1>(P, V, W, i, j, n);
        else regrid_stream(P, V, W, i, j, n);
      }
  int64_t unsafe = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : unsafe)
  for (int j = 1; j <= jj; j++) {
    double row[kKMax + 2];
    LaneRow po{row, 1};
    for (int i = 1; i <= ii; i++) {
      const int64_t c0 = ij_of(V, i, j);
      if (!V.ip[c0]) continue;
      if (W.kp[c0] & kUnsafe) {
        unsafe++;
        remap_fallback(P, V, W, i, j, n);
        continue;
      }
      for (int k = 1; k <= kk + 1; k++) row[k - 1] = V.p[c0 + (int64_t)(k - 1) * V.NIJ];
      if (P.precision == 1) {
        switch (P.hybmap) {
          case 0: remap_mixed<0, 3, 2>(P, V, W, i, j, n, po); break;
          case 1: remap_mixed<1, 3, 2>(P, V, W, i, j, n, po); break;
          case 2: remap_mixed<2, 3, 2>(P, V, W, i, j, n, po); break;
          default: remap_mixed<3, 3, 2>(P, V, W, i, j, n, po); break;
        }
        continue;
      }
      if (g_uv_recompute) {  // g_split_ns interface levels staged, the rest read from p
        double srow[kKMax + 2];
        SplitInterfaces pa = stage_split(srow, 1, V.p + c0, V.NIJ, kk, g_split_ns);
        switch (P.hybmap) {
          case 0: remap_stream<0, 2, 0>(P, V, W, i, j, n, pa); break;
          case 1: remap_stream<1, 2, 0>(P, V, W, i, j, n, pa); break;
          case 2: remap_stream<2, 3, 0>(P, V, W, i, j, n, pa); break;
          default: remap_stream<3, 2, 2>(P, V, W, i, j, n, pa); break;
        }
        continue;
      }
      switch (P.hybmap) {
        case 0: remap_stream<0, kNFMax, 0>(P, V, W, i, j, n, po); break;
        case 1: remap_stream<1, 3, 2>(P, V, W, i, j, n, po); break;
        case 2: remap_stream<2, kNFMax, 2>(P, V, W, i, j, n, po); break;
        default: remap_stream<3, 2, 4>(P, V, W, i, j, n, po); break;
      }
    }
  }
  g_unsafe_cols = unsafe;
  for (int d = 0; d < 3; d++)
    if (itr[d] > 0) {
      double q = d == 0 ? P.diag_q801 : P.diag_q80x;
      for (int j = 1; j <= jj; j++)
        for (int i = 1; i <= ii; i++)
          if (V.ip[ij_of(V, i, j)])
            for (int k = 1; k <= kk; k++) {
              int64_t a = ij_of(V, i, j) + levt(V, k, n, itr[d]);
              V.tracer[a] = q * (src[d][ij_of(V, i, j) + lev4(V, k, n)] - V.tracer[a]);
            }
    }
  for (int pass = 0; pass <= (raflag ? 1 : 0); pass++) {
    int nl = pass == 0 ? n : m;
    if (pass == 1) {
      int64_t ra_fb = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : ra_fb)
      for (int j = 1; j <= jj; j++) {
        double row[kKMax + 2];
        for (int i = 1; i <= ii; i++) {
          if (!V.ip[ij_of(V, i, j)]) continue;
          bool ok;
          switch (P.hybmap) {
            case 0: ok = ra_stream<0, 3, 2>(P, V, W, i, j, m, n, row, 1); break;
            case 1: ok = ra_stream<1, 3, 2>(P, V, W, i, j, m, n, row, 1); break;
            case 2: ok = ra_stream<2, 3, 2>(P, V, W, i, j, m, n, row, 1); break;
            default: ok = ra_stream<3, 3, 2>(P, V, W, i, j, m, n, row, 1); break;
          }
          if (!ok) {
            ra_fb++;
            column_ra<64>(P, V, i, j, m, n, row_lcm(P));
          }
        }
      }
      g_ra_fallback = ra_fb;
    }
    for (int k = 2; k <= kk + 1; k++)
      for (int i = 1; i <= ii; i++) p_halo_jrow(P, V, i, k);
    for (int k = 2; k <= kk + 1; k++)
      for (int j = 0; j <= jj + 1; j++) p_halo_icol(P, V, j, k);
    int64_t uv_fb = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : uv_fb)
    for (int j = 1; j <= jj; j++) {
      double col[kKMax + 2];
      LaneRow dcol{col, 1};
      for (int isv = 0; isv <= 1; isv++)
        for (int i = 1; i <= ii; i++) {
          if (!(isv ? V.iv : V.iu)[ij_of(V, i, j)]) continue;
          bool ok;
          if (P.precision == 1) {
            switch (uv_method(P)) {
              case 0: ok = uv_stream<0, 4, true, true>(P, V, i, j, nl, isv != 0, dcol); break;
              case 1: ok = uv_stream<1, 4, true, true>(P, V, i, j, nl, isv != 0, dcol); break;
              default: ok = uv_stream<3, 4, true, true>(P, V, i, j, nl, isv != 0, dcol); break;
            }
          } else if (g_uv_recompute) {
            switch (uv_method(P)) {
              case 0: ok = uv_stream<0, 4, false, true>(P, V, i, j, nl, isv != 0, dcol); break;
              case 1: ok = uv_stream<1, 0, false, true>(P, V, i, j, nl, isv != 0, dcol); break;
              default: ok = uv_stream<3, 4, false, true>(P, V, i, j, nl, isv != 0, dcol); break;
            }
          } else switch (uv_method(P)) {
            case 0: ok = uv_stream<0, 2>(P, V, i, j, nl, isv != 0, dcol); break;
            case 1: ok = uv_stream<1, 0>(P, V, i, j, nl, isv != 0, dcol); break;
            default: ok = uv_stream<3, 2>(P, V, i, j, nl, isv != 0, dcol); break;
          }
          if (!ok) {
            uv_fb++;
            point_c<64>(P, V, i, j, nl, isv != 0);
          }
        }
    }
    g_uv_fallback += uv_fb;
  }
}

extern "C" int hostcheck_hybgen_stream(const hyb_params* P, const hyb_fields* F, int m, int n,
                                       int raflag) {
  KParams K = kparams_of(P);
  View V = view_of(P, F);
  if (K.nhybrd != K.kk) return -2;
  g_uv_fallback = 0;
  run_stream(K, V, m, n, raflag);
  return 0;
}

// the streaming family with the LDS-free remap and u/v variants
extern "C" int hostcheck_hybgen_stream_rc(const hyb_params* P, const hyb_fields* F, int m,
                                          int n, int raflag) {
  g_uv_recompute = 1;
  const int rc = hostcheck_hybgen_stream(P, F, m, n, raflag);
  g_uv_recompute = 0;
  return rc;
}

// staged levels of the split remap in hostcheck_hybgen_stream_rc (>= 1)
extern "C" void hostcheck_set_split_levels(int ns) { g_split_ns = ns; }

extern "C" int64_t hostcheck_stream_unsafe_columns(void) { return g_unsafe_cols; }

extern "C" int64_t hostcheck_stream_uv_fallback_points(void) { return g_uv_fallback; }

extern "C" int64_t hostcheck_stream_ra_fallback_columns(void) { return g_ra_fallback; }

// ---- one column through the streaming walks (column-level checks) ----
template <bool MIXED, int M>
static void column_walk(int kk, const double* si, const double* pi, const double* po,
                        const int* lcm0, double* so) {
  double s[kKMax + 2], pres[kKMax + 2], prsf[kKMax + 2];
  uint64_t lcm = 0;
  for (int k = 0; k < kk; k++) {
    s[k] = si[k];
    if (lcm0 && lcm0[k]) lcm |= 1ull << k;
  }
  for (int k = 0; k <= kk; k++) {
    pres[k] = pi[k];
    prsf[k] = po[k];
  }
  double* fp[1] = {s};
  double* op[1] = {so};  // outputs to their own array: no in-place hazard
  PiArray pa{pres, 1};
  LaneRow pr{prsf, 1};
  FieldSink<1> sink{op, 1, nullptr, 0.0, 1};
  if (MIXED) walk_core<M, 1, 0, true>(kk, pres[kk], pa, pr, sink, fp, 1, kk, lcm, 1);
  else walk_core<M, 1, 0>(kk, pres[kk], pa, pr, sink, fp, 1, kk, lcm, 1);
}

extern "C" void hostcheck_remap_column(int mixed, int method, int kk, const double* si,
                                       const double* pi, const double* po, const int* lcm0,
                                       double* so) {
  switch (method + 4 * (mixed != 0)) {
    case 0: column_walk<false, 0>(kk, si, pi, po, lcm0, so); break;
    case 1: column_walk<false, 1>(kk, si, pi, po, lcm0, so); break;
    case 2: column_walk<false, 2>(kk, si, pi, po, lcm0, so); break;
    case 3: column_walk<false, 3>(kk, si, pi, po, lcm0, so); break;
    case 4: column_walk<true, 0>(kk, si, pi, po, lcm0, so); break;
    case 5: column_walk<true, 1>(kk, si, pi, po, lcm0, so); break;
    case 6: column_walk<true, 2>(kk, si, pi, po, lcm0, so); break;
    default: column_walk<true, 3>(kk, si, pi, po, lcm0, so); break;
  }
}


// ---- Stokes drift onto the model layers (hyb_stokes.h) ----
#include "hyb_stokes.h"

extern "C" int hostcheck_stokes(const hyb_stokes_params* P, const hyb_stokes_fields* F, int n) {
  KStokes Q{};
  Q.ii = P->ii;
  Q.jj = P->jj;
  Q.kk = P->kdm;
  Q.nsdzi = P->nsdzi;
  Q.nb = P->nbdy;
  Q.n = n;
  Q.NI = (int64_t)P->idm + 2 * P->nbdy;
  Q.NIJ = Q.NI * ((int64_t)P->jdm + 2 * P->nbdy);
  Q.ws0 = P->ws0;
  Q.ws1 = P->ws1;
  for (int k = 0; k < P->nsdzi; k++) Q.sdzi[k] = F->sdzi[k];
  Q.usdz = F->usdz;
  Q.vsdz = F->vsdz;
  Q.dp = F->dp;
  Q.p = F->p;
  Q.ip = F->ip;
  Q.usdp = F->usdp;
  Q.vsdp = F->vsdp;
#pragma omp parallel for schedule(dynamic, 1)
  for (int j = 1; j <= Q.jj; j++)
    for (int i = 1; i <= Q.ii; i++) {
      const int64_t c0 = stokes_ij(Q, i, j);
      if (!Q.ip[c0]) continue;
      stokes_column<2>(Q, c0);
    }
  return 0;
}

// ---- tracer column remaps of trcupd.F90 (hyb_trc.h) ----
#include "hyb_trc.h"

extern "C" int hostcheck_trc(const hyb_trc_params* P, const hyb_trc_fields* F) {
  KTrc Q{};
  Q.ii = P->ii;
  Q.jj = P->jj;
  Q.ki = P->ki;
  Q.ko = P->ko;
  Q.ks = P->ks;
  Q.method = P->method;
  Q.nb = P->nbdy;
  Q.NI = (int64_t)P->idm + 2 * P->nbdy;
  Q.NIJ = Q.NI * ((int64_t)P->jdm + 2 * P->nbdy);
  Q.si = F->si;
  Q.pi = F->pi;
  Q.po = F->po;
  Q.mask = F->mask;
  Q.so = F->so;
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
  for (int j = 1; j <= Q.jj; j++)
    for (int i = 1; i <= Q.ii; i++) {
      const int64_t c0 = trc_ij(Q, i, j);
      if (Q.mask[c0] && !trc_column(Q, c0)) bad = 1;
    }
  return bad ? HYB_EINVAL : 0;
}

// The quotient helpers of hyb_math.h on arrays, from a given first
// reciprocal y0 (so a test can stand in for the hardware's v_rcp_f64 with
// reciprocals of a chosen error).  which: 0 qdiv, 1 qdiv_pos, 2 qrcp,
// 3 qdiv_any, 4 div3 (b and y0 unused).
extern "C" void hostcheck_quot(int which, const double* a, const double* b, const double* y0,
                               double* out, int64_t n) {
  for (int64_t q = 0; q < n; q++) {
    if (which == 4) {
      out[q] = div3(a[q]);
      continue;
    }
    const Rcp64 r = rcp64_from(b[q], y0[q]);
    switch (which) {
      case 0: out[q] = qdiv(a[q], r); break;
      case 1: out[q] = qdiv_pos(a[q], r); break;
      case 2: out[q] = qrcp(r); break;
      default: out[q] = qdiv_any(a[q], r); break;
    }
  }
}
