// hyb_math.h -- constants and scalar functions shared by the host and the
// gfx950 device code.  Every function here is built with
// -ffp-contract=off on both sides, so host and device evaluate the same
// IEEE-754 double operations in the same order.
#pragma once

#include <stdint.h>
#include <math.h>

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#define HYB_HD __host__ __device__ inline
#else
#define HYB_HD inline
#endif
// the per-column bodies: inlined into every kernel that runs them, so a
// kernel never pays a device-function call (stack, argument registers,
// generic loads of the parameter block) for the hot path
#define HYB_HDI HYB_HD __attribute__((always_inline))

namespace hyb {

// mod_cb_arrays.F90:842-857.  onem is g/svref with svref = 1/rhoref,
// folded the way the reference folds it (one IEEE division each).
constexpr double kG      = 9.806;
constexpr double kRhoref = 1000.0;
constexpr double kSvref  = 1.0 / kRhoref;
constexpr double kOnem   = kG / kSvref;
constexpr double kTenm   = kOnem * 10.0;
constexpr double kOnemm  = kOnem * 0.001;
constexpr double kEpsil  = 1.0e-11;
constexpr double kQonem  = 1.0 / kOnem;
// hybgen.F90:374 / :1784
constexpr double kDpthin = 0.001 * kOnemm;

// stmt_fns.h:191-216, 251-262: 17-term sigma-2 rational function.
constexpr double c001 = 9.9984085444849347e+02;
constexpr double c002 = 7.3471625860981584e+00;
constexpr double c003 = -5.3211231792841769e-02;
constexpr double c004 = 3.6492439109814549e-04;
constexpr double c005 = 2.5880571023991390e+00;
constexpr double c006 = 6.7168282786692355e-03;
constexpr double c007 = 1.9203202055760151e-03;
constexpr double c008 = 1.0000000000000000e+00;
constexpr double c009 = 7.2815210113327091e-03;
constexpr double c010 = -4.4787265461983921e-05;
constexpr double c011 = 3.3851002965802430e-07;
constexpr double c012 = 1.3651202389758572e-10;
constexpr double c013 = 1.7632126669040377e-03;
constexpr double c014 = 8.8066583251206474e-06;
constexpr double c015 = 1.8832689434804897e-10;
constexpr double c016 = 5.7463776745432097e-06;
constexpr double c017 = 1.4716275472242334e-09;
constexpr double c018 = 1.1798263740430364e-02;
constexpr double c019 = 9.8920219266399117e-08;
constexpr double c020 = 4.6996642771754730e-06;
constexpr double c021 = 2.5862187075154352e-08;
constexpr double c022 = 3.2921414007960662e-12;
constexpr double c023 = 6.7103246285651894e-06;
constexpr double c024 = 2.4461698007024582e-17;
constexpr double c025 = 9.1534417604289062e-18;
constexpr double kRpdb = 2000.0e4 * 1.0e-4;                 // pref*prs2pdb
constexpr double c101 = c001 + (c018 - c021 * kRpdb) * kRpdb;
constexpr double c103 = c003 + (c019 - c022 * kRpdb) * kRpdb;
constexpr double c105 = c005 + c020 * kRpdb;
constexpr double c108 = c008 + c023 * kRpdb;
constexpr double c109 = c009 - c025 * (kRpdb * kRpdb * kRpdb);
constexpr double c111 = c011 - c024 * (kRpdb * kRpdb);

HYB_HD double dmax(double a, double b) { return a > b ? a : b; }
HYB_HD double dmin(double a, double b) { return a < b ? a : b; }

// Reciprocals of the mixed mode (HYB_PREC_MIXED; no reference
// counterpart).  On the device: the hardware reciprocal (v_rcp_f32 /
// v_rcp_f64, about 1 ulp), the double one refined by one Newton step; on
// the host the IEEE quotient.  The two agree to a few ulp, not bitwise.
// Arguments are normal and positive wherever they are used.
HYB_HD float frcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}
HYB_HD double frcp64(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double r = __builtin_amdgcn_rcp(x);
  const double e = fma(-x, r, 1.0);
  return fma(r, e, r);
#else
  return 1.0 / x;
#endif
}
// Quotients of the reconstruction in arithmetic type R: double keeps the
// reference's IEEE division, float (the mixed mode) multiplies by frcp.
HYB_HD double grcp(double x) { return 1.0 / x; }
HYB_HD float grcp(float x) { return frcp(x); }
HYB_HD double gdiv(double a, double b) { return a / b; }
HYB_HD float gdiv(float a, float b) { return a * frcp(b); }

// ---------------------------------------------------------------------
// FP64 quotients that are bitwise IEEE a/b on a bounded operand range.
//
// The method: a reciprocal y of b refined by two Newton-Raphson steps
// (e = 1 - b*y, y = y + y*e, each one fma), the quotient q0 = a*y, its
// exact remainder r = a - b*q0 (one fma) and Markstein's correction
// q = q0 + r*y (one fma; P. Markstein, "Computation of elementary functions
// on the IBM RISC System/6000 processor", IBM J. Res. Dev. 34(1), 1990).
// On gfx950 this is the core of the compiler's own correctly rounded
// division, which wraps it in v_div_scale (operands scaled by 2^+-128 when
// an exponent is extreme), v_div_fmas (the correction, rescaled) and
// v_div_fixup (zero, infinite and NaN operands, and the sign).  With
//   2^-8 <= |b| <= 2^100  and  a == 0 or 2^-918 <= |a| < 2^760
// (so the exponents differ by less than 768 and a/b is normal), no
// scaling applies: v_div_scale returns its operand and v_div_fmas is a
// plain fma, so the sequence below (fixup kept) computes the same bits as
// the compiler's a/b.  The reciprocal is formed once per divisor and
// shared by every quotient with that divisor, so a second quotient costs 4
// instructions instead of 11.
//
// On the host the same steps run from the IEEE reciprocal, so the host
// build of the kernels exercises the same correction and zero handling.
// ---------------------------------------------------------------------
struct Rcp64 {
  double b, y;  // divisor and its refined reciprocal
};
// two Newton-Raphson steps from a first reciprocal y0 of b
HYB_HD Rcp64 rcp64_from(double b, double y0) {
  double e = fma(-b, y0, 1.0);
  double y = fma(y0, e, y0);
  e = fma(-b, y, 1.0);
  y = fma(y, e, y);
  return Rcp64{b, y};
}
HYB_HD Rcp64 rcp64(double b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return rcp64_from(b, __builtin_amdgcn_rcp(b));  // v_rcp_f64
#else
  return rcp64_from(b, 1.0 / b);
#endif
}
// v_div_fixup_f64: special operands, then the sign of a/b on |q|
HYB_HD double div_fixup(double q, double b, double a) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_div_fixup(q, b, a);
#else
  if (isnan(a) || isnan(b) || a == 0.0 || b == 0.0 || isinf(a) || isinf(b)) return a / b;
  const double m = fabs(q);
  return (signbit(a) != signbit(b)) ? -m : m;
#endif
}
// a / r.b for a == 0 or 2^-918 <= |a| < 2^760 (and |b| in range)
HYB_HD double qdiv(double a, const Rcp64& r) {
  const double q0 = a * r.y;
  const double rr = fma(-r.b, q0, a);
  return div_fixup(fma(rr, r.y, q0), r.b, a);
}
// the same for a > 0 (no fixup needed: the quotient is positive and normal)
HYB_HD double qdiv_pos(double a, const Rcp64& r) {
  const double q0 = a * r.y;
  const double rr = fma(-r.b, q0, a);
  return fma(rr, r.y, q0);
}
// 1 / r.b
HYB_HD double qrcp(const Rcp64& r) {
  const double rr = fma(-r.b, r.y, 1.0);
  return fma(rr, r.y, r.y);
}
// a / r.b for any a: outside the range above, the IEEE division (a branch
// the wave skips when no lane takes it)
HYB_HD double qdiv_any(double a, const Rcp64& r) {
#if defined(__HIP_DEVICE_COMPILE__)
  const int e = __builtin_amdgcn_frexp_exp(a);  // 0 for 0, inf and NaN
#else
  int e = 0;
  if (isfinite(a) && a != 0.0) (void)frexp(a, &e);
#endif
  if (__builtin_expect((unsigned)(e + 917) > 1676u, 0)) return a / r.b;
  return qdiv(a, r);
}
// x / 3 for x == 0 or 2^-918 <= |x| < 2^760: Markstein's correction with
// y = RN(1/3) (the divisor is a constant, so its reciprocal is exact to
// half an ulp and the correction rounds correctly)
HYB_HD double div3(double x) {
  constexpr double y = 0x1.5555555555555p-2;
  const double q0 = x * y;
  const double rr = fma(-3.0, q0, x);
  return div_fixup(fma(rr, y, q0), 3.0, x);
}

// An interface pressure the fast quotients accept: 0 or in [2^-60, 2^60]
// (so a difference of two is 0 or in [2^-112, 2^61]); NaN is not.
HYB_HD bool tame_p(double v) { return (v == 0.0) | ((v >= 0x1p-60) & (v <= 0x1p60)); }

// sig(t,s) = sig_n*(1/sig_d) - 1000, stmt_fns.h:505-511
HYB_HD double sig(double t, double s) {
  double sn = c101 + t * (c002 + t * (c103 + t * c004)) +
              s * (c105 - t * c006 + s * c007);
  double sd = c108 + t * (c109 + t * (c010 + t * (c111 + t * c012))) +
              s * (c013 - t * (c014 + t * t * c015) +
                   sqrt(dmax(0.0, s)) * (c016 + t * t * c017));
  double sq = 1.0 / sd;
  return sn * sq - 1000.0;
}

// Single-precision form of the same rational function for the mixed mode.
HYB_HD float sigf(float t, float s) {
  float sn = (float)c101 + t * ((float)c002 + t * ((float)c103 + t * (float)c004)) +
             s * ((float)c105 - t * (float)c006 + s * (float)c007);
  float sd = (float)c108 + t * ((float)c109 + t * ((float)c010 + t * ((float)c111 + t * (float)c012))) +
             s * ((float)c013 - t * ((float)c014 + t * t * (float)c015) +
                  sqrtf(s > 0.0f ? s : 0.0f) * ((float)c016 + t * t * (float)c017));
  return sn / sd - 1000.0f;
}

// ---------------------------------------------------------------------
// Deterministic helpers for the synthetic-input generator.  They use only
// +,-,*,/, sqrt, floor and ldexp, which are exact/correctly rounded on
// both sides, so the host and device generators agree bit for bit.
// ---------------------------------------------------------------------
HYB_HD uint64_t mix64(uint64_t z) {  // splitmix64 finaliser
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
HYB_HD uint64_t hash4(uint64_t seed, uint64_t a, uint64_t b, uint64_t c) {
  uint64_t h = mix64(seed ^ 0x632BE59BD9B4E019ull);
  h = mix64(h ^ a);
  h = mix64(h ^ (b * 0x9E3779B97F4A7C15ull));
  h = mix64(h ^ (c * 0xD1B54A32D192ED03ull));
  return h;
}
// uniform in [0,1)
HYB_HD double u01(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }
// approximately standard normal: Irwin-Hall with 4 draws, rescaled
HYB_HD double nrm(uint64_t seed, uint64_t a, uint64_t b, uint64_t c) {
  uint64_t h = hash4(seed, a, b, c);
  double s = 0.0;
  for (int q = 0; q < 4; q++) { h = mix64(h + (uint64_t)q); s = s + u01(h); }
  return (s - 2.0) * 1.7320508075688772;
}

// exp(x) for |x| < 700: Cody-Waite reduction by ln2, degree-16 Taylor.
HYB_HD double xexp(double x) {
  const double ln2_hi = 6.93147180369123816490e-01;
  const double ln2_lo = 1.90821492927058770002e-10;
  const double inv_ln2 = 1.44269504088896338700e+00;
  double n = floor(x * inv_ln2 + 0.5);
  double r = (x - n * ln2_hi) - n * ln2_lo;
  double t = 1.0;
  for (int k = 16; k >= 1; k--) t = 1.0 + t * r / (double)k;
  return ldexp(t, (int)n);
}
// cos(x): reduction by 2*pi, degree-34 Taylor in r^2 on [-pi, pi].
HYB_HD double xcos(double x) {
  const double twopi_hi = 6.28318530717958623200e+00;
  const double twopi_lo = 2.44929359829470635445e-16;
  const double inv_twopi = 1.59154943091895345608e-01;
  double n = floor(x * inv_twopi + 0.5);
  double r = (x - n * twopi_hi) - n * twopi_lo;
  double r2 = r * r;
  double t = 1.0;
  for (int k = 34; k >= 2; k -= 2) t = 1.0 - t * r2 / ((double)k * (double)(k - 1));
  return t;
}

}  // namespace hyb
