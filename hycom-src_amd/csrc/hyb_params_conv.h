// hyb_params_conv.h -- hyb_params/hyb_fields (C ABI) to the kernel-side
// KParams/View structs.  Host-only, no HIP dependency.
#pragma once

#include <string.h>

#include "../../include/hycom_hybgen.h"
#include "hyb_column.h"

namespace hyb {

inline KParams kparams_of(const hyb_params* P) {
  KParams K;
  memset(&K, 0, sizeof K);
  K.ii = P->ii;
  K.jj = P->jj;
  K.kk = P->kdm;
  K.nhybrd = P->nhybrd;
  K.ntracr = P->ntracr;
  K.hybmap = P->hybmap;
  K.isopcm = P->isopcm != 0;
  K.ldenfix = P->ldenfix != 0;
  K.nreg = P->nreg;
  K.itr801 = P->itracr801;
  K.itr802 = P->itracr802;
  K.itr803 = P->itracr803;
  K.precision = P->precision;
  K.thbase = P->thbase;
  K.hybrlx = 1.0 / P->qhybrlx;             // hybgen.F90:376
  K.hybiso = P->hybiso;
  K.hybthn = P->hybthn;
  K.hybthk = P->hybthk;
  K.thkbop = P->thkbot * kOnem;            // hybgen.F90:375
  K.dp00i = P->dp00i;
  K.dpns = P->dpns;
  K.dsns = P->dsns;
  K.ra2fac = P->ra2fac;
  K.vland = P->vland;
  if (P->baclin > 0.0) {
    K.diag_q801 = kQonem * 86400.0 / P->baclin;  // hybgen.F90:100
    K.diag_q80x = 86400.0 / P->baclin;           // hybgen.F90:123,146
  }
  for (int k = 0; k < P->kdm; k++) {
    K.dp0k[k] = P->dp0k[k];
    K.ds0k[k] = P->ds0k[k];
    K.dx0k[k] = P->dx0k[k];
  }
  // hybgen.F90:436-452 with qdep==1: dp0cum(1)=0, dp0cum(2)=dp0ij(1)=dp0k(1),
  // dp0cum(k+1)=dp0cum(k)+dp0k(k); the same additions in the same order.
  K.dp0cum[0] = 0.0;
  if (P->kdm >= 1) K.dp0cum[1] = K.dp0cum[0] + P->dp0k[0];
  for (int k = 2; k <= P->kdm && k < 65; k++) K.dp0cum[k] = K.dp0cum[k - 1] + P->dp0k[k - 1];
  for (int t = 0; t < P->ntracr && t < 64; t++)
    if (P->trcflg[t] == 2) K.trc_t |= 1ull << t;
  return K;
}

inline View view_of(const hyb_params* P, const hyb_fields* F) {
  View V;
  memset(&V, 0, sizeof V);
  V.dp = F->dp; V.dpo = F->dpo; V.dpu = F->dpu; V.dpv = F->dpv;
  V.temp = F->temp; V.saln = F->saln; V.th3d = F->th3d;
  V.u = F->u; V.v = F->v; V.tracer = F->tracer;
  V.p = F->p; V.pu = F->pu; V.pv = F->pv;
  V.theta = F->theta; V.depths = F->depths; V.topiso = F->topiso;
  V.depthu = F->depthu; V.depthv = F->depthv;
  V.ubavg = F->ubavg; V.vbavg = F->vbavg;
  V.ip = F->ip; V.iu = F->iu; V.iv = F->iv;
  V.NI = P->idm + 2 * P->nbdy;
  V.NIJ = V.NI * (int64_t)(P->jdm + 2 * P->nbdy);
  V.nb = P->nbdy;
  V.K = P->kdm;
  V.rec_ld = P->idm;
  return V;
}

}  // namespace hyb
