// host_checks.cpp -- drives the host-side entry points of the C ABI that run
// without a device (initial conditions, argument validation, the autotuner's
// table, the selector's parameter layout), for a build of the library's host
// code under AddressSanitizer + UndefinedBehaviorSanitizer (`make sanitize`).
// Exits 0 when every call answered as documented.
#include "mhd_glm.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

namespace {
int failures = 0;
void expect(bool ok, const char* what) {
  if (!ok) {
    std::printf("FAILED: %s\n", what);
    failures++;
  }
}
mhd_cfg base(int nx, int ny, int precision) {
  mhd_cfg c;
  std::memset(&c, 0, sizeof(c));
  c.nx = nx;
  c.ny = ny;
  c.cfl = 0.25;
  c.alpha = 0.18;
  c.muscl = 1;
  c.hlld = 1;
  c.precision = precision;
  c.tile = -1;
  return c;
}
}  // namespace

int main() {
  // Initial conditions, every problem and both element types, padded planes.
  const double otv[5] = {1.1, 0.3, 1.2, 2.0, 4.5};
  const char* problems[] = {"ot", "kh", "khb"};
  for (int prec = 0; prec < 2; prec++) {
    const int nx = 33, ny = 17;
    const size_t ld = (size_t)nx * ny + 5;
    const size_t es = prec == MHD_PREC_FP32 ? 4 : 8;
    std::vector<unsigned char> buf(9 * ld * es);
    for (const char* p : problems)
      expect(mhd_problem_state(p, nx, ny, prec, buf.data(), ld) == 0, "problem_state");
    expect(mhd_problem_state_ex("otv", otv, 5, nx, ny, prec, buf.data(), ld) == 0, "otv");
    expect(mhd_problem_state_ex("otv", otv, 2, nx, ny, prec, buf.data(), ld) == 1, "otv nparams");
    expect(mhd_problem_state_ex("otv", nullptr, 5, nx, ny, prec, buf.data(), ld) == 1, "otv null");
    expect(mhd_problem_state("sod", nx, ny, prec, buf.data(), ld) == 1, "unknown problem");
    expect(mhd_problem_state("ot", nx, ny, prec, buf.data(), 10) == 1, "short ld");
    expect(mhd_problem_state(nullptr, nx, ny, prec, buf.data(), ld) == 1, "null problem");
    expect(mhd_problem_state("ot", nx, ny, prec, nullptr, ld) == 1, "null buffer");
  }
  expect(mhd_problem_state("ot", 8, 8, 42, nullptr, 64) == 1, "bad precision");

  // Configurations refused before any device call.
  mhd_ctx* ctx = nullptr;
  expect(mhd_create(nullptr, &ctx) == 1 && ctx == nullptr, "null cfg");
  mhd_cfg c = base(0, 8, MHD_PREC_FP64);
  expect(mhd_create(&c, &ctx) == 1, "nx < 1");
  c = base(8, 8, MHD_PREC_FP64);
  c.cfl = 1.5;
  expect(mhd_create(&c, &ctx) == 1, "cfl");
  c = base(8, 8, MHD_PREC_FP64);
  c.alpha = 0.0;
  expect(mhd_create(&c, &ctx) == 1, "alpha");
  c = base(8, 8, 99);
  expect(mhd_create(&c, &ctx) == 1, "precision");
  c = base(8, 8, MHD_PREC_REF_MIXED);
  c.muscl = 0;
  expect(mhd_create(&c, &ctx) == 1, "ref_mixed first order");
  c = base(8, 8, MHD_PREC_MIXED_FLUX64);
  c.tile = 0;
  expect(mhd_create(&c, &ctx) == 1, "variant tile");
  c = base(8, 8, MHD_PREC_FP64);
  c.fuse = 3;
  expect(mhd_create(&c, &ctx) == 1, "fuse");
  expect(mhd_create(&c, nullptr) == 1, "null out");
  mhd_ensemble* e = nullptr;
  const int nxs[2] = {8, 0}, nys[2] = {8, 8};
  c = base(2, 2, MHD_PREC_FP64);
  expect(mhd_ensemble_create(&c, 2, nxs, nys, &e) == 1 && e == nullptr, "ensemble size");
  c = base(2, 2, MHD_PREC_MIXED_RECON64);
  expect(mhd_ensemble_create(&c, 1, nxs, nys, &e) == 1, "ensemble variant");
  c = base(8, 8, MHD_PREC_FP64);
  expect(mhd_run_noalloc(&c, nullptr, 64, 0.5, 1, nullptr, nullptr) == 1, "noalloc null state");
  expect(mhd_run_noalloc(nullptr, nullptr, 64, 0.5, 1, nullptr, nullptr) == 1, "noalloc null");

  // Calls on no context.
  mhd_status st;
  expect(std::strlen(mhd_last_error(nullptr)) > 0, "last_error");
  expect(mhd_advance(nullptr, 0.5, 1, &st) == 1, "advance null");
  expect(mhd_get_status(nullptr, &st) == 1, "status null");
  expect(mhd_set_state(nullptr, nullptr, 0) == 1, "set_state null");
  expect(mhd_init_problem(nullptr, "ot") == 1, "init_problem null");
  expect(mhd_selector_configure(nullptr, nullptr, nullptr, nullptr, 0, 0, 0) == 1, "selector");
  expect(mhd_checkpoint_save(nullptr, "x") == 1, "checkpoint");
  mhd_destroy(nullptr);
  mhd_ensemble_destroy(nullptr);

  // The autotuner's table ($MHD_AUTOTUNE_FILE) and the network layout.
  for (int prec = 0; prec <= 5; prec++) {
    const int t = mhd_autotuned_tile(prec, 4096, 4096);
    const int f = mhd_autotuned_fuse(prec, 256, 256);
    expect(t >= 0 && t < 6 && (f == 1 || f == 2), "autotune");
  }
  expect(mhd_mlp_param_count() == 16 * 64 + 64 + 64 * 64 + 64 + 64 + 1, "mlp params");
  expect(mhd_glm_flux_batch(MHD_PREC_FP64, 0, 0, nullptr, nullptr, 1.0, 1, 1e-4, nullptr,
                            nullptr) == 0, "empty batch");
  expect(mhd_glm_flux_batch(MHD_PREC_FP64, 2, 1, nullptr, nullptr, 1.0, 1, 1e-4, nullptr,
                            nullptr) == 1, "batch dir");
  std::printf("host sanitize checks: %d failures\n", failures);
  return failures == 0 ? 0 : 1;
}
