// hyb_trc.cpp -- C ABI of the tracer column remaps (hycom_trc_remap*,
// include/hycom_hybgen.h; trcupd.F90:808-1147).
#include <hip/hip_runtime.h>
#include <limits.h>
#include <string.h>

#include <mutex>

#include "hyb_host.h"
#include "hyb_trc.h"

namespace hyb {
hipError_t launch_trc(const KTrc& Q, hipStream_t s);
}

using namespace hyb;

namespace {

int trc_check(const hyb_trc_params* P, const hyb_trc_fields* F) {
  if (!P || !F) return fail(HYB_EINVAL, "null parameter or field pack");
  if (P->idm < 1 || P->jdm < 1 || P->ii < 1 || P->jj < 1 || P->ii > P->idm || P->jj > P->jdm ||
      P->nbdy < 0)
    return fail(HYB_EINVAL, "extents idm=%d jdm=%d ii=%d jj=%d nbdy=%d", P->idm, P->jdm, P->ii,
                P->jj, P->nbdy);
  if (P->ki < 2 || P->ki > HYB_KMAX) return fail(HYB_EINVAL, "ki=%d outside 2..%d", P->ki, HYB_KMAX);
  if (P->ko < 1 || P->ko > HYB_KMAX) return fail(HYB_EINVAL, "ko=%d outside 1..%d", P->ko, HYB_KMAX);
  if (P->ks < 1 || P->ks > HYB_TRMAX)
    return fail(HYB_EINVAL, "ks=%d outside 1..%d", P->ks, HYB_TRMAX);
  if (P->method != 0 && P->method != 1) return fail(HYB_EINVAL, "method=%d: 0 or 1", P->method);
  if (P->reserved != 0) return fail(HYB_EINVAL, "reserved must be 0");
  if (!F->si || !F->pi || !F->po || !F->mask || !F->so)
    return fail(HYB_EINVAL, "a tracer remap field pointer is null");
  return 0;
}

int64_t nij_of(const hyb_trc_params* P) {
  return ((int64_t)P->idm + 2 * P->nbdy) * ((int64_t)P->jdm + 2 * P->nbdy);
}

// the first-bad-column word, kept
std::mutex g_ws_mu;
int32_t* g_bad = nullptr;

int enqueue(const hyb_trc_params* P, const hyb_trc_fields* F, hipStream_t s, int32_t** bad_out) {
  if (!g_bad) HIPCHK(hipMalloc(&g_bad, sizeof(int32_t)));
  int32_t* bad = g_bad;
  KTrc Q;
  memset(&Q, 0, sizeof Q);
  Q.ii = P->ii;
  Q.jj = P->jj;
  Q.ki = P->ki;
  Q.ko = P->ko;
  Q.ks = P->ks;
  Q.method = P->method;
  Q.nb = P->nbdy;
  Q.NI = (int64_t)P->idm + 2 * P->nbdy;
  Q.NIJ = nij_of(P);
  Q.si = F->si;
  Q.pi = F->pi;
  Q.po = F->po;
  Q.mask = F->mask;
  Q.so = F->so;
  Q.bad = bad;
  HIPCHK(launch_trc(Q, s));
  if (bad_out) *bad_out = bad;
  return 0;
}

struct TrcMirror {
  int64_t NIJ = 0;
  int ki = 0, ko = 0, ks = 0;
  double *si = nullptr, *pi = nullptr, *po = nullptr, *so = nullptr;
  int32_t* mask = nullptr;
  hipStream_t st = nullptr;
  void release() {
    void* a[] = {si, pi, po, so, mask};
    for (void* x : a)
      if (x) (void)hipFree(x);
    si = pi = po = so = nullptr;
    mask = nullptr;
    if (st) (void)hipStreamDestroy(st);
    st = nullptr;
    NIJ = 0;
  }
};
std::mutex g_tmirror_mu;
TrcMirror g_tmirror;

}  // namespace

extern "C" {

int hycom_trc_remap_device(const hyb_trc_params* P, const hyb_trc_fields* F, void* stream) {
  int rc = trc_check(P, F);
  if (rc) return rc;
  if ((rc = device_ready())) return rc;
  std::lock_guard<std::mutex> lk(g_ws_mu);
  return enqueue(P, F, (hipStream_t)stream, nullptr);
}

int hycom_trc_remap(const hyb_trc_params* P, const hyb_trc_fields* F) {
  int rc = trc_check(P, F);
  if (rc) return rc;
  if ((rc = device_ready())) return rc;
  std::lock_guard<std::mutex> lk(g_tmirror_mu);
  TrcMirror& M = g_tmirror;
  const int64_t NIJ = nij_of(P);
  const size_t nsi = (size_t)NIJ * P->ki * P->ks, nso = (size_t)NIJ * P->ko * P->ks;
  const size_t npi = (size_t)NIJ * (P->ki + 1), npo = (size_t)NIJ * (P->ko + 1);
  if (M.NIJ != NIJ || M.ki != P->ki || M.ko != P->ko || M.ks != P->ks) {
    M.release();
    HIPCHK(hipStreamCreate(&M.st));
    HIPCHK(hipMalloc(&M.si, nsi * sizeof(double)));
    HIPCHK(hipMalloc(&M.pi, npi * sizeof(double)));
    HIPCHK(hipMalloc(&M.po, npo * sizeof(double)));
    HIPCHK(hipMalloc(&M.so, nso * sizeof(double)));
    HIPCHK(hipMalloc(&M.mask, (size_t)NIJ * sizeof(int32_t)));
    M.NIJ = NIJ;
    M.ki = P->ki;
    M.ko = P->ko;
    M.ks = P->ks;
  }
  const hipMemcpyKind h2d = hipMemcpyHostToDevice;
  HIPCHK(hipMemcpyAsync(M.si, F->si, nsi * sizeof(double), h2d, M.st));
  HIPCHK(hipMemcpyAsync(M.pi, F->pi, npi * sizeof(double), h2d, M.st));
  HIPCHK(hipMemcpyAsync(M.po, F->po, npo * sizeof(double), h2d, M.st));
  HIPCHK(hipMemcpyAsync(M.so, F->so, nso * sizeof(double), h2d, M.st));
  HIPCHK(hipMemcpyAsync(M.mask, F->mask, (size_t)NIJ * sizeof(int32_t), h2d, M.st));
  hyb_trc_fields D{M.si, M.pi, M.po, M.mask, M.so};
  int32_t* bad = nullptr;
  int32_t first = INT_MAX;
  {
    std::lock_guard<std::mutex> lw(g_ws_mu);
    if ((rc = enqueue(P, &D, M.st, &bad))) return rc;
    HIPCHK(hipMemcpyAsync(&first, bad, sizeof first, hipMemcpyDeviceToHost, M.st));
    HIPCHK(hipMemcpyAsync(F->so, M.so, nso * sizeof(double), hipMemcpyDeviceToHost, M.st));
    HIPCHK(hipStreamSynchronize(M.st));
  }
  if (first != INT_MAX) {
    const int64_t NI = (int64_t)P->idm + 2 * P->nbdy;
    return fail(HYB_EINVAL, "po out of order at i=%d j=%d: the reference stops here ('bad lf')",
                (int)(first % NI) - P->nbdy + 1, (int)(first / NI) - P->nbdy + 1);
  }
  return 0;
}

}  // extern "C"
