// ref_harness.cpp -- C wrappers around the reference's own physics headers.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile directly against
// the reference headers where they lie (include/mhd/mhd_glm_common.hpp and
// include/mhd/mhd_common.hpp), with the reference's protocol flags
// (-O2, gamma 5/3, MHD_HLLD_TOLERANCE_VALUE=1e-4), into oracle/_ref/.
// Tests use it to pin the oracle restatement function by function; nothing
// in the product links it.  Built once with Real=double and once with
// -DUSE_FLOAT (Real=float); the symbol names are the same in both.
#define MHD_GAMMA_VALUE (5.0 / 3.0)
#include "mhd/mhd_glm_common.hpp"

#include <cstddef>

extern "C" {

int ref_glm_flux_batch(int n, int dir, const Real* UL, const Real* UR, Real ch,
                       int hlld, Real* F, int* status) {
  for (int s = 0; s < n; s++) {
    const Real* a = UL + 9 * (size_t)s;
    const Real* b = UR + 9 * (size_t)s;
    Real* f = F + 9 * (size_t)s;
    const int st = (dir == 0) ? mhdGlmFluxX(a, b, ch, hlld != 0, f)
                              : mhdGlmFluxY(a, b, ch, hlld != 0, f);
    if (status) status[s] = st;
  }
  return 0;
}

void ref_cons_to_prim_batch(int n, const Real* U, Real* W) {
  for (int s = 0; s < n; s++) {
    Real psi;
    mhdGlmConsToPrim(U + 9 * (size_t)s, W + 9 * (size_t)s, psi);
    W[9 * (size_t)s + 8] = psi;
  }
}

void ref_prim_to_cons_batch(int n, const Real* W, Real* U) {
  for (int s = 0; s < n; s++)
    mhdGlmPrimToCons(W + 9 * (size_t)s, W[9 * (size_t)s + 8], U + 9 * (size_t)s);
}

void ref_fast_speed_batch(int n, const Real* W, int dir, Real* out) {
  for (int s = 0; s < n; s++)
    out[s] = (dir == 0) ? mhdFastSpeedX(W + 9 * (size_t)s)
                        : mhdFastSpeedY(W + 9 * (size_t)s);
}

int ref_real_size(void) { return (int)sizeof(Real); }

}  // extern "C"
