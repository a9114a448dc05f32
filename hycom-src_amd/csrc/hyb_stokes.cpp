// hyb_stokes.cpp -- C ABI of the Stokes drift remap (hycom_stokes_vertical*,
// include/hycom_hybgen.h; mod_stokes.F90:541-546, 650-705).
#include <hip/hip_runtime.h>
#include <string.h>

#include <mutex>

#include "hyb_host.h"
#include "hyb_stokes.h"

namespace hyb {
hipError_t launch_stokes(const KStokes& Q, hipStream_t s);
}

using namespace hyb;

namespace {

int stokes_check(const hyb_stokes_params* P, const hyb_stokes_fields* F, int n) {
  if (!P || !F) return fail(HYB_EINVAL, "null parameter or field pack");
  if (P->idm < 1 || P->jdm < 1 || P->ii < 1 || P->jj < 1 || P->ii > P->idm || P->jj > P->jdm ||
      P->nbdy < 0)
    return fail(HYB_EINVAL, "extents idm=%d jdm=%d ii=%d jj=%d nbdy=%d", P->idm, P->jdm, P->ii,
                P->jj, P->nbdy);
  if (P->kdm < 1 || P->kdm > HYB_KMAX) return fail(HYB_EINVAL, "kdm=%d outside 1..%d", P->kdm, HYB_KMAX);
  if (P->nsdzi < 2 || P->nsdzi > HYB_KMAX)
    return fail(HYB_EINVAL, "nsdzi=%d outside 2..%d", P->nsdzi, HYB_KMAX);
  if (P->reserved != 0) return fail(HYB_EINVAL, "reserved must be 0");
  if (n < 1 || n > 2) return fail(HYB_EINVAL, "time level n=%d", n);
  if (!F->usdz || !F->vsdz || !F->sdzi || !F->dp || !F->p || !F->ip || !F->usdp || !F->vsdp)
    return fail(HYB_EINVAL, "a Stokes field pointer is null");
  // the reference stops on decreasing depths (mod_stokes.F90:346-356)
  for (int k = 2; k <= P->nsdzi; k++)
    if (!(F->sdzi[k - 1] >= F->sdzi[k - 2]))
      return fail(HYB_EINVAL, "sdzi(%d) < sdzi(%d): Stokes depths must be non-decreasing", k,
                  k - 1);
  return 0;
}

KStokes kstokes_of(const hyb_stokes_params* P, const hyb_stokes_fields* F, int n) {
  KStokes Q;
  memset(&Q, 0, sizeof Q);
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
  return Q;
}

// device mirrors of the host call, kept between calls of one shape
struct StokesMirror {
  int64_t NIJ = 0;
  int K = 0, NZ = 0;
  double *usdz = nullptr, *vsdz = nullptr, *dp = nullptr, *p = nullptr;
  double *usdp = nullptr, *vsdp = nullptr;
  int32_t* ip = nullptr;
  hipStream_t st = nullptr;
  void release() {
    void* a[] = {usdz, vsdz, dp, p, usdp, vsdp, ip};
    for (void* x : a)
      if (x) (void)hipFree(x);
    usdz = vsdz = dp = p = usdp = vsdp = nullptr;
    ip = nullptr;
    if (st) (void)hipStreamDestroy(st);
    st = nullptr;
    NIJ = 0;
  }
};
std::mutex g_smirror_mu;
StokesMirror g_smirror;

}  // namespace

extern "C" {

int hycom_stokes_vertical_device(const hyb_stokes_params* P, const hyb_stokes_fields* F, int n,
                                 void* stream) {
  int rc = stokes_check(P, F, n);
  if (rc) return rc;
  if ((rc = device_ready())) return rc;
  KStokes Q = kstokes_of(P, F, n);
  HIPCHK(launch_stokes(Q, (hipStream_t)stream));
  return 0;
}

int hycom_stokes_vertical(const hyb_stokes_params* P, const hyb_stokes_fields* F, int n) {
  int rc = stokes_check(P, F, n);
  if (rc) return rc;
  if ((rc = device_ready())) return rc;
  std::lock_guard<std::mutex> lk(g_smirror_mu);
  StokesMirror& M = g_smirror;
  const int64_t NIJ = ((int64_t)P->idm + 2 * P->nbdy) * ((int64_t)P->jdm + 2 * P->nbdy);
  const size_t nz = (size_t)NIJ * 2 * P->nsdzi, n4 = (size_t)NIJ * P->kdm * 2;
  const size_t n3p = (size_t)NIJ * (P->kdm + 1), n3 = (size_t)NIJ * P->kdm;
  if (M.NIJ != NIJ || M.K != P->kdm || M.NZ != P->nsdzi) {
    M.release();
    HIPCHK(hipStreamCreate(&M.st));
    HIPCHK(hipMalloc(&M.usdz, nz * sizeof(double)));
    HIPCHK(hipMalloc(&M.vsdz, nz * sizeof(double)));
    HIPCHK(hipMalloc(&M.dp, n4 * sizeof(double)));
    HIPCHK(hipMalloc(&M.p, n3p * sizeof(double)));
    HIPCHK(hipMalloc(&M.usdp, n3 * sizeof(double)));
    HIPCHK(hipMalloc(&M.vsdp, n3 * sizeof(double)));
    HIPCHK(hipMalloc(&M.ip, (size_t)NIJ * sizeof(int32_t)));
    M.NIJ = NIJ;
    M.K = P->kdm;
    M.NZ = P->nsdzi;
  }
  const hipMemcpyKind h2d = hipMemcpyHostToDevice, d2h = hipMemcpyDeviceToHost;
  HIPCHK(hipMemcpyAsync(M.usdz, F->usdz, nz * sizeof(double), h2d, M.st));
  HIPCHK(hipMemcpyAsync(M.vsdz, F->vsdz, nz * sizeof(double), h2d, M.st));
  HIPCHK(hipMemcpyAsync(M.dp, F->dp, n4 * sizeof(double), h2d, M.st));
  HIPCHK(hipMemcpyAsync(M.p, F->p, n3p * sizeof(double), h2d, M.st));
  HIPCHK(hipMemcpyAsync(M.usdp, F->usdp, n3 * sizeof(double), h2d, M.st));
  HIPCHK(hipMemcpyAsync(M.vsdp, F->vsdp, n3 * sizeof(double), h2d, M.st));
  HIPCHK(hipMemcpyAsync(M.ip, F->ip, (size_t)NIJ * sizeof(int32_t), h2d, M.st));
  hyb_stokes_fields D = *F;
  D.usdz = M.usdz;
  D.vsdz = M.vsdz;
  D.dp = M.dp;
  D.p = M.p;
  D.usdp = M.usdp;
  D.vsdp = M.vsdp;
  D.ip = M.ip;
  if ((rc = hycom_stokes_vertical_device(P, &D, n, M.st))) return rc;
  HIPCHK(hipMemcpyAsync(F->usdp, M.usdp, n3 * sizeof(double), d2h, M.st));
  HIPCHK(hipMemcpyAsync(F->vsdp, M.vsdp, n3 * sizeof(double), d2h, M.st));
  HIPCHK(hipStreamSynchronize(M.st));
  return 0;
}

}  // extern "C"
