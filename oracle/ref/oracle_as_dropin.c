/* Test infrastructure only: the oracle behind the drop-in's C entry points,
 * so that the Fortran shim (hycom-src_amd/fortran/hybgen_hip.F90) can run
 * inside the stand-alone model on the CPU with the oracle in place of the
 * GPU library (oracle/_ref/model/hycom_oracle).  Run against the stock model
 * it checks the oracle, and the shim's packing of the model's state, in the
 * model's own time loop. */
#include "hybgen_oracle.h"

static const char* g_msg = "";

int hycom_hybgen(const hyb_params* P, const hyb_fields* F, int m, int n, int raflag) {
  int rc = oracle_hybgen(P, F, m, n, raflag);
  g_msg = rc ? "the oracle rejected the call" : "";
  return rc;
}

const char* hycom_hybgen_last_error(void) { return g_msg; }
