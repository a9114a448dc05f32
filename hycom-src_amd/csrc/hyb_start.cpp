// hyb_start.cpp -- the grid masks and the derived fields of a model start
// from a restart, on one tile holding the whole domain (i0 = j0 = 0,
// ii = idm, jj = jdm):
//   hycom_grid_masks:  geopar.F90:244-256 and bigrid.F90:1-330 (land/sea
//                      masks of p, u and v points, the region type);
//   hycom_state_start: mod_hycom.F90:1684-1722 (dpthuv with no open
//                      boundaries, the dp halo, p and dpudpv of both time
//                      levels).
// Host code: it runs once per model start, beside the .a/.b reader.
#include <math.h>
#include <string.h>

#include "hyb_host.h"

using namespace hyb;

namespace {

constexpr double kHugel = 1.2676506002282294e30;  // 2^100 (mod_cb_arrays.F90:854)
enum HaloType { kHaloP = 1, kHaloU = 3, kHaloV = 4 };  // mod_xc.F90:41-44

// One (1-nbdy:idm+nbdy, 1-nbdy:jdm+nbdy) plane.
template <class T>
struct Plane {
  T* a;
  int ii, jj, nb;
  int64_t NI;
  T& operator()(int i, int j) const { return a[(i - 1 + nb) + NI * (int64_t)(j - 1 + nb)]; }
};

// xctilr(a, 1, 1, nbdy, nbdy, type) on one tile holding the whole domain.
// nreg 2 is the arctic build of the routine (mod_xc_sm.h:1172-1335): south
// closed, the north rows folded across the top row, periodic in i.
// nreg 0 and 1 are the other build (1337-1428): closed in latitude, closed
// (0) or periodic (1) in longitude.
template <class T>
void tile_halo(const Plane<T>& A, int type, T vland, int nreg) {
  const int ii = A.ii, jj = A.jj, nb = A.nb;
  if (nreg == 2) {
    for (int j = 1; j <= nb; j++)
      for (int i = 1; i <= ii; i++) A(i, 1 - j) = vland;
    for (int j = 1; j <= nb; j++)
      for (int i = 1; i <= ii; i++) {
        if (type == kHaloP) A(i, jj + j) = A(ii - (i - 1) % ii, jj - 1 - j);
        else if (type == kHaloU) A(i, jj + j) = A((ii - (i - 1)) % ii + 1, jj - 1 - j);
        else A(i, jj + j) = A(ii - (i - 1) % ii, jj - j);
      }
  } else {
    for (int j = 1; j <= nb; j++)
      for (int i = 1; i <= ii; i++) {
        A(i, 1 - j) = vland;
        A(i, jj + j) = vland;
      }
  }
  for (int j = 1 - nb; j <= jj + nb; j++)
    for (int i = 1; i <= nb; i++) {
      if (nreg == 0) {
        A(1 - i, j) = vland;
        A(ii + i, j) = vland;
      } else {
        A(1 - i, j) = A(ii + 1 - i, j);
        A(ii + i, j) = A(i, j);
      }
    }
}

// bigrid's boundary rule (parts I and II, bigrid.F90:123-152, 262-300): the
// halo beyond a closed boundary is land
template <class T>
void land_boundaries(const Plane<T>& A, bool lperiod, bool larctic) {
  const int ii = A.ii, jj = A.jj, nb = A.nb;
  for (int j = 1 - nb; j <= 0; j++)
    for (int i = 1 - nb; i <= ii + nb; i++) A(i, j) = T(0);
  if (!larctic)
    for (int j = jj + 1; j <= jj + nb; j++)
      for (int i = 1 - nb; i <= ii + nb; i++) A(i, j) = T(0);
  if (!lperiod)
    for (int j = 1 - nb; j <= jj + nb; j++) {
      for (int i = 1 - nb; i <= 0; i++) A(i, j) = T(0);
      for (int i = ii + 1; i <= ii + nb; i++) A(i, j) = T(0);
    }
}

}  // namespace

extern "C" {

int hycom_grid_masks(int32_t idm, int32_t jdm, int32_t nbdy, double* depths, int32_t* ip,
                     int32_t* iu, int32_t* iv, int32_t* nreg) {
  if (!depths || !ip || !iu || !iv || !nreg) return fail(HYB_EINVAL, "null argument");
  if (idm < 2 || jdm < 2 || nbdy < 1)
    return fail(HYB_EINVAL, "extents idm=%d jdm=%d nbdy=%d", idm, jdm, nbdy);
  const int64_t NI = (int64_t)idm + 2 * nbdy;
  const Plane<double> D{depths, idm, jdm, nbdy, NI};
  // data voids are land (geopar.F90:244-252)
  for (int j = 1; j <= jdm; j++)
    for (int i = 1; i <= idm; i++)
      if (D(i, j) > 0.5 * kHugel) D(i, j) = 0.0;
  // periodic in longitude if the last column has sea; across the arctic if
  // the last row has (bigrid.F90:26-46)
  double dmax_e = 0.0, dmax_n = 0.0;
  for (int j = 1; j <= jdm; j++) dmax_e = fmax(dmax_e, D(idm, j));
  for (int i = 1; i <= idm; i++) dmax_n = fmax(dmax_n, D(i, jdm));
  const bool lperiod = dmax_e > 0.0, larctic = dmax_n > 0.0;
  if (larctic && !lperiod)
    return fail(HYB_EINVAL, "arctic domain, but non-periodic: the reference stops (bigrid.F90:50-58)");
  const int reg = larctic ? 2 : (lperiod ? 1 : 0);
  tile_halo(D, kHaloP, 0.0, reg);  // vland is 0 until geopar sets it after bigrid
  land_boundaries(D, lperiod, larctic);
  // single-width inlets and one-point seas (bigrid.F90:154-190)
  for (int j = 1; j <= jdm; j++)
    for (int i = 1; i <= idm; i++)
      if (D(i, j) > 0.0 && D(i - 1, j) <= 0.0 && D(i + 1, j) <= 0.0 && D(i, j - 1) <= 0.0 &&
          D(i, j + 1) <= 0.0)
        return fail(HYB_EINVAL,
                    "dh(%d,%d) has 4 land neighbours: the bathymetry must be corrected "
                    "(bigrid.F90:160-190)", i, j);
  const Plane<int32_t> IP{ip, idm, jdm, nbdy, NI}, IU{iu, idm, jdm, nbdy, NI},
      IV{iv, idm, jdm, nbdy, NI};
  // p points where the water depth is positive; u and v points between two
  // p points (bigrid.F90:196-237), their halos as xctilr fills them
  for (int j = 1 - nbdy; j <= jdm + nbdy; j++)
    for (int i = 1 - nbdy; i <= idm + nbdy; i++) {
      IP(i, j) = D(i, j) > 0.0 ? 1 : 0;
      IU(i, j) = 0;
      IV(i, j) = 0;
    }
  for (int j = 1; j <= jdm; j++)
    for (int i = 1; i <= idm; i++) {
      IU(i, j) = (IP(i - 1, j) > 0 && IP(i, j) > 0) ? 1 : 0;
      IV(i, j) = (IP(i, j - 1) > 0 && IP(i, j) > 0) ? 1 : 0;
    }
  tile_halo(IU, kHaloU, 0, reg);
  tile_halo(IV, kHaloV, 0, reg);
  land_boundaries(IU, lperiod, larctic);
  land_boundaries(IV, lperiod, larctic);
  *nreg = reg;
  return 0;
}

int hycom_state_start(const hyb_params* P, const hyb_start_fields* F) {
  if (!P || !F) return fail(HYB_EINVAL, "null argument");
  if (!F->pbot || !F->dp || !F->ip || !F->iu || !F->iv || !F->depthu || !F->depthv || !F->p ||
      !F->dpu || !F->dpv)
    return fail(HYB_EINVAL, "a start field pointer is null");
  const int ii = P->idm, jj = P->jdm, nb = P->nbdy, kk = P->kdm;
  if (P->ii != ii || P->jj != jj)
    return fail(HYB_EINVAL, "a model start covers the whole domain: ii=%d jj=%d, idm=%d jdm=%d",
                P->ii, P->jj, ii, jj);
  if (ii < 2 || jj < 2 || nb < 1)
    return fail(HYB_EINVAL, "extents idm=%d jdm=%d nbdy=%d", ii, jj, nb);
  if (kk < 1 || kk > HYB_KMAX) return fail(HYB_EINVAL, "kdm=%d outside 1..%d", kk, HYB_KMAX);
  if (P->nreg < 0 || P->nreg > 2)
    return fail(HYB_EUNSUPPORTED, "nreg=%d: only 0 (closed), 1 (periodic) and 2 (arctic)", P->nreg);
  const int reg = P->nreg;
  const double vland = P->vland;
  const int64_t NI = (int64_t)ii + 2 * nb, NIJ = NI * ((int64_t)jj + 2 * nb);
  auto plane = [&](double* a) { return Plane<double>{a, ii, jj, nb, NI}; };
  auto cplane = [&](const int32_t* a) { return Plane<int32_t>{(int32_t*)a, ii, jj, nb, NI}; };
  const Plane<double> PB = plane(F->pbot), DU = plane(F->depthu), DV = plane(F->depthv);
  const Plane<int32_t> IP = cplane(F->ip), IU = cplane(F->iu), IV = cplane(F->iv);

  // geopar.F90:760-800: depthu, depthv, dpu, dpv and p(:,:,1) start at 0
  for (int64_t c = 0; c < NIJ; c++) {
    F->depthu[c] = 0.0;
    F->depthv[c] = 0.0;
    F->p[c] = 0.0;
  }
  memset(F->dpu, 0, sizeof(double) * (size_t)NIJ * kk * 2);
  memset(F->dpv, 0, sizeof(double) * (size_t)NIJ * kk * 2);

  // dpthuv (lbflag 0, no ports): bottom pressure at u and v points, one
  // point beyond the tile, then their halos (dpthuv.F90:48, 64-123)
  tile_halo(PB, kHaloP, vland, reg);
  for (int j = 0; j <= jj + 1; j++)
    for (int i = 0; i <= ii + 1; i++) {
      if (IU(i, j) == 1) DU(i, j) = fmin(PB(i, j), PB(i - 1, j));
      if (IV(i, j) == 1) DV(i, j) = fmin(PB(i, j), PB(i, j - 1));
    }
  tile_halo(DU, kHaloU, vland, reg);
  tile_halo(DV, kHaloV, vland, reg);

  // the halo of dp, both time levels (mod_hycom.F90:1686)
  for (int l = 0; l < 2 * kk; l++) tile_halo(plane(F->dp + NIJ * l), kHaloP, vland, reg);

  // each time level: p from dp at sea points, then dpudpv with p valid out
  // to nbdy and dpu, dpv out to nbdy-1 (mod_hycom.F90:1691-1722,
  // dpudpv.F90:85-100); p keeps the interfaces of level 2
  const int m = nb - 1;
  for (int nm = 1; nm <= 2; nm++) {
    const double* dp = F->dp + NIJ * (int64_t)kk * (nm - 1);
    for (int j = 1 - nb; j <= jj + nb; j++)
      for (int i = 1 - nb; i <= ii + nb; i++) {
        if (IP(i, j) == 0) continue;
        const int64_t c = (i - 1 + nb) + NI * (int64_t)(j - 1 + nb);
        F->p[c] = 0.0;
        for (int k = 1; k <= kk; k++) F->p[c + NIJ * k] = F->p[c + NIJ * (k - 1)] + dp[c + NIJ * (k - 1)];
      }
    double* dpu = F->dpu + NIJ * (int64_t)kk * (nm - 1);
    double* dpv = F->dpv + NIJ * (int64_t)kk * (nm - 1);
    for (int j = 1 - m; j <= jj + m; j++)
      for (int k = 1; k <= kk; k++)
        for (int i = 1 - m; i <= ii + m; i++) {
          const int64_t c = (i - 1 + nb) + NI * (int64_t)(j - 1 + nb);
          const double* pk = F->p + NIJ * (k - 1);
          const double* pk1 = F->p + NIJ * k;
          if (IU(i, j) != 0) {
            const double du = DU(i, j);
            dpu[c + NIJ * (k - 1)] =
                fmax(0.0, fmin(du, 0.5 * (pk1[c] + pk1[c - 1])) - fmin(du, 0.5 * (pk[c] + pk[c - 1])));
          }
          if (IV(i, j) != 0) {
            const double dv = DV(i, j);
            dpv[c + NIJ * (k - 1)] = fmax(
                0.0, fmin(dv, 0.5 * (pk1[c] + pk1[c - NI])) - fmin(dv, 0.5 * (pk[c] + pk[c - NI])));
          }
        }
  }
  return 0;
}

}  // extern "C"
