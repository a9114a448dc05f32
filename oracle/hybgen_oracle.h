/*
 * ORACLE -- test infrastructure, not product code.
 *
 * A plain-C, FP64, single-threaded-per-row restatement of HYCOM's hybgen
 * call path, used only by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  Nothing in the shipped
 * library links or calls it.  Each function cites the reference lines it
 * restates (paths relative to the HYCOM-src tree).
 *
 * Pinning: every function here is compared bit-for-bit against the
 * reference Fortran compiled from its own sources by oracle/ref/Makefile
 * (outputs in oracle/_ref/, see tests/test_oracle_ref.py), and against
 * committed golden vectors written by that build (tests/golden/).
 */
#ifndef HYBGEN_ORACLE_H
#define HYBGEN_ORACLE_H

#include "../include/hycom_hybgen.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Full hybgen(m,n,raflag) on host arrays, in place (hybgen.F90:14-306).
 * Returns 0, or HYB_EINVAL/HYB_EUNSUPPORTED for parameter packs the
 * restatement does not cover. */
int oracle_hybgen(const hyb_params* P, const hyb_fields* F,
                  int m, int n, int raflag);

/* Same, with the j loop split across nthreads OpenMP threads the way the
 * reference's !$OMP PARALLEL DO over j does (hybgen.F90:89-95). */
int oracle_hybgen_omp(const hyb_params* P, const hyb_fields* F,
                      int m, int n, int raflag, int nthreads);

/* oracle_hybgen_omp that also writes each sea p-column's decisions of
 * hybgenaj: fixlay, the number of fixed layers (hybgen.F90:479-573), and
 * kp, the deepest inflated layer after the overturn (590-686), at
 * (i-1) + idm*(j-1); land points are left as they are. */
int oracle_hybgen_decisions(const hyb_params* P, const hyb_fields* F, int m, int n,
                            int raflag, int nthreads, int32_t* fixlay, int32_t* kp);

/* sigma-2, 17-term EOS (stmt_fns.h:505-511). */
double oracle_sig(double t, double s);

/* pcmtrc (method 0, trcupd.F90:808-909) or plmtrc (method 1, 911-1147) of
 * one column: si[f*ki + l], pi[ki+1], po[ko+1], so[f*ko + k].  Returns -1
 * where the reference stops ('bad lf', its outputs NaN from there). */
int oracle_trc_column(int method, int ki, int ks, const double* si, const double* pi,
                      const double* po, int ko, double* so);

/* The same at every point of F->mask != 0 (initrc trcupd.F90:167-199,
 * thermf.F90:770-806); HYB_EINVAL if any column would stop. */
int oracle_trc_remap(const hyb_trc_params* P, const hyb_trc_fields* F);

/* stokes_vertical_j(n,j) for j = 1..jj (mod_stokes.F90:541-546, 650-705)
 * on host arrays: usdp, vsdp at sea points. */
int oracle_stokes_vertical(const hyb_stokes_params* P, const hyb_stokes_fields* F, int n);

/* One column of that remap: ks fields (<= 2) of ki old layers onto ko new
 * ones, PLM slopes on dpi = max(dprs, thin), no PCM layers.
 * si[f*ki + k], pi[ki+1], po[ko+1], so[f*ko + k]. */
void oracle_plm_dpi_column(int ki, int ko, int ks, const double* si, const double* pi,
                           const double* po, double thin, double* so);

/* Single-column remappers on caller arrays, 1 field (ks=1), for unit
 * tests: method 0 PCM, 1 PLM, 2 PPM, 3 WENO.  si[k], pi[k+1], po[k+1],
 * lcm[k] (0/1), so[k] are 0-based C arrays of the Fortran 1-based ones. */
void oracle_remap_column(int method, int kk, const double* si,
                         const double* pi, const double* po,
                         const int* lcm, double thin, double* so);

#ifdef __cplusplus
}
#endif
#endif
