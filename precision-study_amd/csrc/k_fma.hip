// k_fma.hip -- FP64 step kernels compiled with FMA contraction
// (-ffp-contract=fast), the default of the reference's nvcc build.  Checked
// against the contraction-off FP64 path to relative L2 <= 1e-12 per variable.
#define MHD_NS fma
#include "mhd_step.inc.hip"

namespace mhd {

const Launchers* launchers_fma() {
  static const Launchers fp64 = fma::make_launchers<fma::ExactPol<double>>(MODE_FP64_FMA);
  return &fp64;
}

}  // namespace mhd
