// sanitize_checks.cpp -- TEST INFRASTRUCTURE: drives every entry point of the
// CPU oracle (oracle_capi.cpp, compiled into this program) on small grids and
// on special states, for a build under AddressSanitizer and
// UndefinedBehaviorSanitizer (`make -C oracle sanitize`).  Exits 0 when every
// call returned; the sanitizers abort on the first error they find.
#include "oracle_capi.cpp"

#include <cmath>
#include <cstdio>
#include <limits>
#include <random>
#include <vector>

namespace {

template <class R>
void special_states(std::vector<R>& a, std::vector<R>& b, int n) {
  std::mt19937 rng(7);
  std::uniform_real_distribution<double> u(-2.0, 2.0);
  a.assign((size_t)n * 9, (R)0);
  b.assign((size_t)n * 9, (R)0);
  for (int s = 0; s < n; s++)
    for (int k = 0; k < 9; k++) {
      a[s * 9 + k] = (R)u(rng);
      b[s * 9 + k] = (R)u(rng);
    }
  // rho and E positive in most states; zeros, NaN, infinities and huge and
  // tiny magnitudes in a few.
  for (int s = 0; s < n; s++) {
    a[s * 9 + 0] = std::fabs(a[s * 9 + 0]) + (R)0.1;
    b[s * 9 + 0] = std::fabs(b[s * 9 + 0]) + (R)0.1;
    a[s * 9 + 4] = std::fabs(a[s * 9 + 4]) + (R)5;
    b[s * 9 + 4] = std::fabs(b[s * 9 + 4]) + (R)5;
  }
  const R q = std::numeric_limits<R>::quiet_NaN(), inf = std::numeric_limits<R>::infinity();
  const R spec[] = {(R)0, -(R)0, q, inf, -inf, (R)1e30, (R)1e-30};
  for (int s = 0; s < 7 && s < n; s++)
    for (int k = 0; k < 9; k++) {
      a[s * 9 + k] = spec[(s + k) % 7];
      b[(n - 1 - s) * 9 + k] = spec[(s + 2 * k) % 7];
    }
}

template <class R>
void physics(bool f32) {
  const int n = 512;
  std::vector<R> a, b, f((size_t)n * 9), w((size_t)n * 9), o(n), pl((size_t)n * 9),
      pr((size_t)n * 9);
  std::vector<int> st(n);
  special_states(a, b, n);
  for (int dir = 0; dir < 2; dir++)
    for (int hlld = 0; hlld < 2; hlld++) {
      if (f32)
        orc_glm_flux_batch_f32(n, dir, (const float*)a.data(), (const float*)b.data(), 1.5f,
                               hlld, 1e-4, (float*)f.data(), st.data());
      else
        orc_glm_flux_batch_f64(n, dir, (const double*)a.data(), (const double*)b.data(), 1.5,
                               hlld, 1e-4, (double*)f.data(), st.data());
    }
  if (f32) {
    orc_cons_to_prim_batch_f32(n, (const float*)a.data(), (float*)w.data());
    orc_prim_to_cons_batch_f32(n, (const float*)w.data(), (float*)f.data());
    orc_fast_speed_batch_f32(n, (const float*)w.data(), 1, (float*)o.data());
    orc_reconstruct_batch_f32(n, (const float*)a.data(), (const float*)w.data(),
                              (const float*)b.data(), (float*)pl.data(), (float*)pr.data());
  } else {
    orc_cons_to_prim_batch_f64(n, (const double*)a.data(), (double*)w.data());
    orc_prim_to_cons_batch_f64(n, (const double*)w.data(), (double*)f.data());
    orc_fast_speed_batch_f64(n, (const double*)w.data(), 0, (double*)o.data());
    orc_reconstruct_batch_f64(n, (const double*)a.data(), (const double*)w.data(),
                              (const double*)b.data(), (double*)pl.data(), (double*)pr.data());
  }
}

int runs() {
  const double otv[5] = {1.067, 0.80, 1.76, 0.21, 0.53};
  const char* problems[] = {"ot", "kh", "khb", "otv"};
  for (const char* p : problems) {
    for (int prec = 0; prec < 2; prec++) {
      const int nx = 24, ny = 20;
      uint64_t cnt[4];
      int steps = 0, bad[2] = {-1, -1};
      if (prec == 0) {
        std::vector<double> u((size_t)nx * ny * 9), rhs(u.size());
        orc_init_params_f64(p, otv, nx, ny, u.data());
        orc_compute_rhs_f64(nx, ny, u.data(), 1.0, 0.18, 1, 1, 1e-4, rhs.data(), cnt);
        (void)orc_max_signal_speed_f64(nx, ny, u.data());
        for (int muscl = 0; muscl < 2; muscl++) {
          std::vector<double> v = u;
          double t = 0.0, ch = 0.0;
          orc_advance_f64(nx, ny, v.data(), 8, 0.5, 0.25, 0.18, muscl, 1, 1e-4, &t, &steps, cnt,
                          bad, &ch);
        }
        for (int mask = 0; mask < 32; mask += 7) {
          std::vector<double> v = u;
          double t = 0.0, ch = 0.0;
          orc_refmixed_advance_regions(nx, ny, v.data(), 4, 0.5, 0.25, 0.18, 1, 1e-4, &t, &steps,
                                       cnt, bad, &ch, mask);
        }
      } else {
        std::vector<float> u((size_t)nx * ny * 9), rhs(u.size());
        orc_init_params_f32(p, otv, nx, ny, u.data());
        orc_compute_rhs_f32(nx, ny, u.data(), 1.0f, 0.18f, 1, 0, 1e-4, rhs.data(), cnt);
        float t = 0.0f, ch = 0.0f;
        orc_advance_f32(nx, ny, u.data(), 8, 0.5f, 0.25f, 0.18f, 1, 1, 1e-4, &t, &steps, cnt,
                        bad, &ch);
      }
    }
  }
  // A state that fails (E far below the kinetic energy): the failure path.
  const int nx = 16, ny = 16;
  std::vector<double> u((size_t)nx * ny * 9);
  orc_init_f64("ot", nx, ny, u.data());
  u[(5 * nx + 7) * 9 + 4] = 1e-8;
  uint64_t cnt[4];
  int steps = 0, bad[2] = {-1, -1};
  double t = 0.0, ch = 0.0;
  const int code = orc_advance_f64(nx, ny, u.data(), 20, 0.5, 0.25, 0.18, 1, 1, 1e-4, &t, &steps,
                                   cnt, bad, &ch);
  return code == 3 ? 0 : 1;
}

}  // namespace

int main() {
  physics<double>(false);
  physics<float>(true);
  const int rc = runs();
  std::printf("oracle sanitize checks: %s\n", rc == 0 ? "ok" : "unexpected failure code");
  return rc;
}
