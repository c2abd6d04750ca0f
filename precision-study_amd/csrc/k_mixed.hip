// k_mixed.hip -- mixed-precision step kernels: FP32 reconstruction, HLLD and
// GLM (compiled with FMA contraction and the hardware's fast FP32 division
// and square root), FP64 pressure recovery where FP32 would cancel, FP64 (or
// FP32 + FP32 compensated) state and update accumulation, FP64 time and dt.
// Scored by relative L2 against the FP64 path, not bitwise.
#define MHD_NS mixed
#include "mhd_step.inc.hip"

namespace mhd {

const Launchers* launchers_mixed(int mode) {
  static const Launchers m = mixed::make_launchers<mixed::MixedPol>(MODE_MIXED);
  static const Launchers k = mixed::make_launchers<mixed::KahanPol>(MODE_MIXED_KAHAN);
  static const Launchers r = mixed::make_variant_launchers<mixed::MixVarPol<double, float, float>>(
      MODE_MIXED_RECON64);
  static const Launchers f = mixed::make_variant_launchers<mixed::MixVarPol<float, double, double>>(
      MODE_MIXED_FLUX64);
  static const Launchers n = mixed::make_variant_launchers<mixed::MixVarPol<float, double, float>>(
      MODE_MIXED_FLUX64_NARROW);
  switch (mode) {
    case MODE_MIXED_KAHAN: return &k;
    case MODE_MIXED_RECON64: return &r;
    case MODE_MIXED_FLUX64: return &f;
    case MODE_MIXED_FLUX64_NARROW: return &n;
    default: return &m;
  }
}

}  // namespace mhd
