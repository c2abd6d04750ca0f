// mhd_api.cpp -- host side of the C ABI (include/mhd_glm.h).
//
// Compiled by the host C++ compiler (g++ -O2, no FMA) against the HIP
// runtime: owns device memory, the stream, the step-batching loop and the
// host-evaluated initial conditions; every kernel launch goes through the
// launchers of mhd_kernels.hip.
//
// Time loop (mhd2d_glm_cpu.cpp:513-570 / mhd2d_glm_gpu.cu:483-537): the
// reference synchronises with the host six times per step; here c_h, dt,
// t, the positivity verdicts and the fallback counters live on the device,
// steps are enqueued in batches (replayed as one HIP graph), and the host
// reads the 100-byte control block once per batch.
#include "mhd_glm.h"
#include "mhd_kernels.h"
#include "mhd_problems.hpp"

#include <hip/hip_runtime_api.h>

#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

using mhd::Ctrl;
using mhd::GridDesc;

struct mhd_ctx {
  mhd_cfg cfg;
  GridDesc grid;
  const mhd::Launchers* L = nullptr;
  hipStream_t stream = nullptr;
  void* u = nullptr;       // device state U   (9 planes)
  void* u1 = nullptr;      // device state U1  (9 planes)
  void* ctrl = nullptr;    // device Ctrl<R>
  void* ctrl_host = nullptr;  // pinned host copy of the control block
  size_t plane = 0;
  bool have_state = false;
  hipGraphExec_t graph = nullptr;  // `batch` steps, replayed
  int graph_steps = 0;
  std::string err;
};

namespace {

const char* kNoCtx = "mhd: no context";

// Element type of host state buffers: float for FP32, double otherwise.
size_t io_elem(int prec) { return prec == MHD_PREC_FP32 ? sizeof(float) : sizeof(double); }

const mhd::Launchers* launchers_for(int prec) {
  switch (prec) {
    case MHD_PREC_FP64: return mhd::launchers_exact(mhd::MODE_FP64);
    case MHD_PREC_FP32: return mhd::launchers_exact(mhd::MODE_FP32);
    case MHD_PREC_MIXED: return mhd::launchers_mixed(mhd::MODE_MIXED);
    case MHD_PREC_MIXED_KAHAN: return mhd::launchers_mixed(mhd::MODE_MIXED_KAHAN);
    case MHD_PREC_FP64_FMA: return mhd::launchers_fma();
    default: return nullptr;
  }
}

int fail(mhd_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int hip_fail(mhd_ctx* c, hipError_t e, const char* what) {
  std::string m = std::string("HIP error in ") + what + ": " + hipGetErrorString(e);
  return fail(c, 9, m);
}

#define HIP_OR_FAIL(ctx, call)                            \
  do {                                                    \
    hipError_t e_ = (call);                               \
    if (e_ != hipSuccess) return hip_fail(ctx, e_, #call); \
  } while (0)

template <class R>
void init_ctrl(Ctrl<R>* h, const mhd_cfg& cfg) {
  std::memset(h, 0, sizeof(*h));
  // dx = (x_max - x_min) / nx in the run's type (mhd2d_glm_cpu.cpp:465-466).
  h->dx = ((R)1.0 - (R)0.0) / (R)cfg.nx;
  h->dy = ((R)1.0 - (R)0.0) / (R)cfg.ny;
  h->cfl = (R)cfg.cfl;
  h->t = (R)0.0;
  h->t_end = (R)0.0;
  h->run = mhd::RUN_DONE;
  h->bad[0] = INT_MAX;
  h->bad[1] = INT_MAX;
  h->fail_cell = -1;
}

template <class R>
void fill_status(const mhd_ctx* c, const Ctrl<R>* h, mhd_status* st) {
  std::memset(st, 0, sizeof(*st));
  st->code = (h->run == mhd::RUN_HALTED) ? h->code : 0;
  st->done = (h->run == mhd::RUN_DONE) ? 1 : 0;
  st->steps = h->step;
  st->t = (double)h->t;
  st->dt = (double)(h->run == mhd::RUN_HALTED ? h->fail_dt : h->dt);
  st->ch_last = (double)(h->run == mhd::RUN_HALTED ? h->fail_ch : h->ch);
  const unsigned long long steps = (unsigned long long)h->step;
  if (c->cfg.hlld) {
    // The reference evaluates (nx+1)*ny x faces and nx*(ny+1) y faces per
    // stage (mhd2d_glm_cpu.cpp:253-271), two stages per step.
    st->x_evaluations = 2ull * steps * (unsigned long long)(c->cfg.nx + 1) * c->cfg.ny;
    st->y_evaluations = 2ull * steps * (unsigned long long)c->cfg.nx * (c->cfg.ny + 1);
    st->degenerate = h->degenerate;
    st->nonphysical = h->nonphysical;
  }
  st->bad_i = st->bad_j = -1;
  if (h->run == mhd::RUN_HALTED && (h->code == 3 || h->code == 4) && h->fail_cell >= 0) {
    st->bad_i = h->fail_cell % c->cfg.nx;
    st->bad_j = h->fail_cell / c->cfg.nx;
  }
  st->fail_step = h->fail_step;
}

int read_ctrl(mhd_ctx* c) {
  HIP_OR_FAIL(c, hipMemcpyAsync(c->ctrl_host, c->ctrl, c->L->ctrl_bytes,
                                hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
  return 0;
}

int status_of(mhd_ctx* c, mhd_status* st) {
  mhd_status tmp;
  if (c->L->time_is_float)
    fill_status(c, (const Ctrl<float>*)c->ctrl_host, &tmp);
  else
    fill_status(c, (const Ctrl<double>*)c->ctrl_host, &tmp);
  if (st) *st = tmp;
  return tmp.code;
}

hipError_t enqueue_step(mhd_ctx* c, hipEvent_t* ev = nullptr) {
  return c->L->step(c->grid, c->u, c->u1, c->ctrl, c->stream, ev);
}

hipError_t enqueue_target(mhd_ctx* c, double t_end) {
  return c->L->set_target(c->ctrl, t_end, c->stream);
}

// Builds (once) a graph of `batch` steps on the context stream.
int ensure_graph(mhd_ctx* c, int batch) {
  if (c->graph && c->graph_steps == batch) return 0;
  if (c->graph) {
    (void)hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
  }
  hipGraph_t g = nullptr;
  HIP_OR_FAIL(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  hipError_t le = hipSuccess;
  for (int s = 0; s < batch && le == hipSuccess; s++) le = enqueue_step(c);
  hipError_t ee = hipStreamEndCapture(c->stream, &g);
  if (le != hipSuccess) return hip_fail(c, le, "step launch during capture");
  if (ee != hipSuccess) return hip_fail(c, ee, "hipStreamEndCapture");
  hipError_t ie = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ie != hipSuccess) {
    c->graph = nullptr;
    return hip_fail(c, ie, "hipGraphInstantiate");
  }
  c->graph_steps = batch;
  return 0;
}

bool use_graphs() {
  const char* s = std::getenv("MHD_GRAPHS");
  return !(s && s[0] == '0');
}

template <class R>
void problem_soa(const char* problem, int nx, int ny, R* out, size_t ld) {
  const R dx = ((R)1.0 - (R)0.0) / (R)nx, dy = ((R)1.0 - (R)0.0) / (R)ny;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      R P[9], U[9];
      mhd::problem_primitive<R>(problem, i, j, dx, dy, P);
      mhd::host_prim_to_cons<R>(P, U);
      const size_t q = (size_t)j * nx + i;
      for (int k = 0; k < 9; k++) out[k * ld + q] = U[k];
    }
}

}  // namespace

extern "C" {

const char* mhd_last_error(const mhd_ctx* ctx) {
  return ctx ? ctx->err.c_str() : kNoCtx;
}

int mhd_create(const mhd_cfg* cfg, mhd_ctx** out) {
  if (!out) return 1;
  *out = nullptr;
  if (!cfg) return 1;
  if (cfg->nx < 2 || cfg->ny < 2) return 1;
  if (!(cfg->cfl > 0.0 && cfg->cfl < 1.0)) return 1;
  if (!(cfg->alpha > 0.0)) return 1;
  if (!launchers_for(cfg->precision)) return 1;
  if ((long long)cfg->nx * cfg->ny > (long long)INT_MAX) return 1;
  mhd_ctx* c = new mhd_ctx();
  c->cfg = *cfg;
  c->L = launchers_for(cfg->precision);
  if (!(c->cfg.lowbeta_ratio > 0.0)) c->cfg.lowbeta_ratio = 16.0;
  if (c->cfg.batch_steps <= 0) c->cfg.batch_steps = 32;
  if (c->cfg.hlld_tol <= 0.0) c->cfg.hlld_tol = 1.0e-4;
  if (c->cfg.tile < 0 || c->cfg.tile >= mhd::TILE_COUNT) c->cfg.tile = mhd::kDefaultTile;
  // Planes start on 512-byte boundaries.
  const size_t cells = (size_t)cfg->nx * cfg->ny;
  c->plane = (cells + 63) / 64 * 64;
  c->grid.nx = cfg->nx;
  c->grid.ny = cfg->ny;
  c->grid.plane = c->plane;
  c->grid.muscl = cfg->muscl ? 1 : 0;
  c->grid.hlld = cfg->hlld ? 1 : 0;
  c->grid.alpha = cfg->alpha;
  c->grid.tol = c->cfg.hlld_tol;
  c->grid.tile = c->cfg.tile;
  c->grid.lowbeta_ratio = c->cfg.lowbeta_ratio;
  const size_t bytes = (size_t)c->L->state_planes * c->plane * c->L->state_elem;
  hipError_t e;
  if ((e = hipStreamCreate(&c->stream)) != hipSuccess ||
      (e = hipMalloc(&c->u, bytes)) != hipSuccess ||
      (e = hipMalloc(&c->u1, bytes)) != hipSuccess ||
      (e = hipMalloc(&c->ctrl, c->L->ctrl_bytes)) != hipSuccess ||
      (e = hipHostMalloc(&c->ctrl_host, c->L->ctrl_bytes, hipHostMallocDefault)) !=
          hipSuccess) {
    std::fprintf(stderr, "mhd_create: %s\n", hipGetErrorString(e));
    mhd_destroy(c);
    return 9;
  }
  *out = c;
  return 0;
}

void mhd_destroy(mhd_ctx* c) {
  if (!c) return;
  if (c->graph) (void)hipGraphExecDestroy(c->graph);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->u) (void)hipFree(c->u);
  if (c->u1) (void)hipFree(c->u1);
  if (c->ctrl) (void)hipFree(c->ctrl);
  if (c->ctrl_host) (void)hipHostFree(c->ctrl_host);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int mhd_problem_state(const char* problem, int nx, int ny, int precision,
                      void* soa9, size_t ld) {
  if (!mhd::problem_known(problem) || nx < 1 || ny < 1 || !soa9) return 1;
  if (ld < (size_t)nx * ny) return 1;
  if (precision == MHD_PREC_FP32)
    problem_soa<float>(problem, nx, ny, (float*)soa9, ld);
  else if (launchers_for(precision))
    problem_soa<double>(problem, nx, ny, (double*)soa9, ld);
  else
    return 1;
  return 0;
}

int mhd_set_state(mhd_ctx* c, const void* soa9, size_t ld) {
  if (!c) return 1;
  if (!soa9 || ld < (size_t)c->cfg.nx * c->cfg.ny) return fail(c, 1, "mhd_set_state: bad buffer");
  const size_t cells = (size_t)c->cfg.nx * c->cfg.ny;
  const size_t es = c->L->state_elem;
  if (c->L->state_planes == 18) {
    // Compensated form: plane k holds hi = (float)x, plane 9+k lo = (float)(x - hi).
    std::vector<float> hl(18 * cells);
    const double* src = (const double*)soa9;
    for (int k = 0; k < 9; k++)
      for (size_t q = 0; q < cells; q++) {
        const double x = src[k * ld + q];
        const float hi = (float)x;
        hl[k * cells + q] = hi;
        hl[(9 + k) * cells + q] = (float)(x - (double)hi);
      }
    for (int k = 0; k < 18; k++)
      HIP_OR_FAIL(c, hipMemcpy((char*)c->u + k * c->plane * es, hl.data() + k * cells,
                               cells * es, hipMemcpyHostToDevice));
  } else {
    for (int k = 0; k < 9; k++)
      HIP_OR_FAIL(c, hipMemcpyAsync((char*)c->u + k * c->plane * es,
                                    (const char*)soa9 + k * ld * es, cells * es,
                                    hipMemcpyHostToDevice, c->stream));
  }
  if (c->L->time_is_float) {
    init_ctrl((Ctrl<float>*)c->ctrl_host, c->cfg);
  } else {
    init_ctrl((Ctrl<double>*)c->ctrl_host, c->cfg);
  }
  HIP_OR_FAIL(c, hipMemcpyAsync(c->ctrl, c->ctrl_host, c->L->ctrl_bytes,
                                hipMemcpyHostToDevice, c->stream));
  hipError_t e = c->L->speed(c->grid, c->u, c->ctrl, c->stream);
  if (e == hipSuccess) e = c->L->commit(c->ctrl, c->stream);
  if (e != hipSuccess) return hip_fail(c, e, "initial signal speed");
  if (int rc = read_ctrl(c)) return rc;
  c->have_state = true;
  return 0;
}

int mhd_init_problem(mhd_ctx* c, const char* problem) {
  if (!c) return 1;
  if (!mhd::problem_known(problem)) return fail(c, 1, "mhd_init_problem: unknown problem");
  const size_t cells = (size_t)c->cfg.nx * c->cfg.ny;
  std::vector<char> host(9 * cells * io_elem(c->cfg.precision));
  mhd_problem_state(problem, c->cfg.nx, c->cfg.ny, c->cfg.precision, host.data(), cells);
  return mhd_set_state(c, host.data(), cells);
}

int mhd_enqueue_steps(mhd_ctx* c, double t_end, int nsteps) {
  if (!c) return 1;
  if (!c->have_state) return fail(c, 1, "mhd_enqueue_steps: no state");
  hipError_t e = enqueue_target(c, t_end);
  if (e != hipSuccess) return hip_fail(c, e, "set target");
  const int batch = c->cfg.batch_steps;
  int left = nsteps;
  if (use_graphs() && left >= batch) {
    if (int rc = ensure_graph(c, batch)) return rc;
    while (left >= batch) {
      HIP_OR_FAIL(c, hipGraphLaunch(c->graph, c->stream));
      left -= batch;
    }
  }
  for (; left > 0; left--) {
    e = enqueue_step(c);
    if (e != hipSuccess) return hip_fail(c, e, "step launch");
  }
  return 0;
}

int mhd_advance(mhd_ctx* c, double t_end, long long max_steps, mhd_status* st) {
  if (!c) return 1;
  if (!c->have_state) return fail(c, 1, "mhd_advance: no state");
  hipError_t e = enqueue_target(c, t_end);
  if (e != hipSuccess) return hip_fail(c, e, "set target");
  if (int rc = read_ctrl(c)) return rc;
  long long left = max_steps < 0 ? LLONG_MAX : max_steps;
  const int batch = c->cfg.batch_steps;
  const bool graphs = use_graphs();
  mhd_status s;
  status_of(c, &s);
  while (left > 0 && s.code == 0 && !s.done) {
    const int n = (int)(left < batch ? left : batch);
    if (graphs && n == batch) {
      if (int rc = ensure_graph(c, batch)) return rc;
      HIP_OR_FAIL(c, hipGraphLaunch(c->graph, c->stream));
    } else {
      for (int q = 0; q < n; q++) {
        e = enqueue_step(c);
        if (e != hipSuccess) return hip_fail(c, e, "step launch");
      }
    }
    left -= n;
    if (int rc = read_ctrl(c)) return rc;
    status_of(c, &s);
  }
  if (st) *st = s;
  return s.code;
}

int mhd_get_status(mhd_ctx* c, mhd_status* st) {
  if (!c) return 1;
  if (int rc = read_ctrl(c)) return rc;
  return status_of(c, st);
}

int mhd_get_state(mhd_ctx* c, int which, void* soa9, size_t ld) {
  if (!c) return 1;
  if (!soa9 || ld < (size_t)c->cfg.nx * c->cfg.ny || (which != 0 && which != 1))
    return fail(c, 1, "mhd_get_state: bad arguments");
  const size_t es = c->L->state_elem;
  const size_t cells = (size_t)c->cfg.nx * c->cfg.ny;
  const char* src = (const char*)(which == 0 ? c->u : c->u1);
  if (c->L->state_planes == 18) {
    std::vector<float> hl(18 * cells);
    for (int k = 0; k < 18; k++)
      HIP_OR_FAIL(c, hipMemcpyAsync(hl.data() + k * cells, src + k * c->plane * es, cells * es,
                                    hipMemcpyDeviceToHost, c->stream));
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    double* out = (double*)soa9;
    for (int k = 0; k < 9; k++)
      for (size_t q = 0; q < cells; q++)
        out[k * ld + q] = (double)hl[k * cells + q] + (double)hl[(9 + k) * cells + q];
    return 0;
  }
  for (int k = 0; k < 9; k++)
    HIP_OR_FAIL(c, hipMemcpyAsync((char*)soa9 + k * ld * es, src + k * c->plane * es,
                                  cells * es, hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
  return 0;
}

int mhd_device_state(mhd_ctx* c, void** dev_ptr, size_t* ld) {
  if (!c || !dev_ptr || !ld) return 1;
  *dev_ptr = c->u;
  *ld = c->plane;
  return 0;
}

void* mhd_stream(mhd_ctx* c) { return c ? (void*)c->stream : nullptr; }

namespace {
const char kCkMagic[8] = {'M', 'H', 'D', 'G', 'L', 'M', 'C', 'K'};
struct CkHeader {
  char magic[8];
  int version, nx, ny, precision, planes, elem, ctrl_bytes, pad;
};
}  // namespace

int mhd_checkpoint_save(mhd_ctx* c, const char* path) {
  if (!c || !path) return 1;
  if (!c->have_state) return fail(c, 1, "mhd_checkpoint_save: no state");
  const size_t cells = (size_t)c->cfg.nx * c->cfg.ny;
  const size_t es = c->L->state_elem;
  std::vector<char> buf((size_t)c->L->state_planes * cells * es);
  if (int rc = read_ctrl(c)) return rc;
  for (int k = 0; k < c->L->state_planes; k++)
    HIP_OR_FAIL(c, hipMemcpyAsync(buf.data() + k * cells * es, (const char*)c->u + k * c->plane * es,
                                  cells * es, hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
  CkHeader h;
  std::memset(&h, 0, sizeof(h));
  std::memcpy(h.magic, kCkMagic, 8);
  h.version = 1;
  h.nx = c->cfg.nx;
  h.ny = c->cfg.ny;
  h.precision = c->cfg.precision;
  h.planes = c->L->state_planes;
  h.elem = (int)es;
  h.ctrl_bytes = (int)c->L->ctrl_bytes;
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(c, 1, std::string("mhd_checkpoint_save: cannot open ") + path);
  bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1 &&
            std::fwrite(c->ctrl_host, c->L->ctrl_bytes, 1, f) == 1 &&
            std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
  ok = (std::fclose(f) == 0) && ok;
  return ok ? 0 : fail(c, 1, "mhd_checkpoint_save: write failed");
}

int mhd_checkpoint_load(mhd_ctx* c, const char* path) {
  if (!c || !path) return 1;
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(c, 1, std::string("mhd_checkpoint_load: cannot open ") + path);
  CkHeader h;
  const size_t cells = (size_t)c->cfg.nx * c->cfg.ny;
  const size_t es = c->L->state_elem;
  bool ok = std::fread(&h, sizeof(h), 1, f) == 1 && std::memcmp(h.magic, kCkMagic, 8) == 0 &&
            h.version == 1 && h.nx == c->cfg.nx && h.ny == c->cfg.ny &&
            h.precision == c->cfg.precision && h.planes == c->L->state_planes &&
            h.elem == (int)es && h.ctrl_bytes == (int)c->L->ctrl_bytes;
  std::vector<char> buf;
  if (ok) {
    buf.resize((size_t)h.planes * cells * es);
    ok = std::fread(c->ctrl_host, c->L->ctrl_bytes, 1, f) == 1 &&
         std::fread(buf.data(), 1, buf.size(), f) == buf.size();
  }
  std::fclose(f);
  if (!ok) return fail(c, 1, "mhd_checkpoint_load: file does not match this context");
  for (int k = 0; k < h.planes; k++)
    HIP_OR_FAIL(c, hipMemcpy((char*)c->u + k * c->plane * es, buf.data() + k * cells * es,
                             cells * es, hipMemcpyHostToDevice));
  HIP_OR_FAIL(c, hipMemcpy(c->ctrl, c->ctrl_host, c->L->ctrl_bytes, hipMemcpyHostToDevice));
  c->have_state = true;
  return 0;
}

int mhd_time_kernels(mhd_ctx* c, double t_end, int nsteps, double* avg_ms) {
  if (!c || !avg_ms || nsteps <= 0) return 1;
  if (!c->have_state) return fail(c, 1, "mhd_time_kernels: no state");
  hipError_t e = enqueue_target(c, t_end);
  if (e != hipSuccess) return hip_fail(c, e, "set target");
  std::vector<hipEvent_t> ev((size_t)nsteps * 4, nullptr);
  for (auto& x : ev) HIP_OR_FAIL(c, hipEventCreate(&x));
  int rc = 0;
  for (int q = 0; q < nsteps && rc == 0; q++) {
    e = enqueue_step(c, &ev[(size_t)q * 4]);
    if (e != hipSuccess) rc = hip_fail(c, e, "timed step launch");
  }
  if (rc == 0) {
    e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) rc = hip_fail(c, e, "hipStreamSynchronize");
  }
  if (rc == 0) {
    double sum[3] = {0, 0, 0};
    for (int q = 0; q < nsteps; q++)
      for (int k = 0; k < 3; k++) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, ev[(size_t)q * 4 + k], ev[(size_t)q * 4 + k + 1]);
        sum[k] += ms;
      }
    for (int k = 0; k < 3; k++) avg_ms[k] = sum[k] / nsteps;
  }
  for (auto& x : ev)
    if (x) (void)hipEventDestroy(x);
  return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Operator-level batches.
// ---------------------------------------------------------------------------
namespace {

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

template <class R>
int flux_batch(int dir, int n, const void* UL, const void* UR, double ch, int hlld,
               double tol, void* F, int* status) {
  const size_t bytes = (size_t)n * 9 * sizeof(R);
  DevBuf a, b, f, s;
  if (hipMalloc(&a.p, bytes) != hipSuccess || hipMalloc(&b.p, bytes) != hipSuccess ||
      hipMalloc(&f.p, bytes) != hipSuccess || hipMalloc(&s.p, (size_t)n * sizeof(int)) != hipSuccess)
    return 9;
  if (hipMemcpy(a.p, UL, bytes, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b.p, UR, bytes, hipMemcpyHostToDevice) != hipSuccess)
    return 9;
  if (mhd::launch_glm_flux_batch<R>(n, dir, (const R*)a.p, (const R*)b.p, (R)ch, hlld, tol,
                                    (R*)f.p, (int*)s.p, nullptr) != hipSuccess)
    return 9;
  if (hipMemcpy(F, f.p, bytes, hipMemcpyDeviceToHost) != hipSuccess) return 9;
  if (status && hipMemcpy(status, s.p, (size_t)n * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
    return 9;
  return 0;
}

template <class R>
int c2p_batch(int n, const void* U, void* W) {
  const size_t bytes = (size_t)n * 9 * sizeof(R);
  DevBuf a, b;
  if (hipMalloc(&a.p, bytes) != hipSuccess || hipMalloc(&b.p, bytes) != hipSuccess) return 9;
  if (hipMemcpy(a.p, U, bytes, hipMemcpyHostToDevice) != hipSuccess) return 9;
  if (mhd::launch_cons_to_prim_batch<R>(n, (const R*)a.p, (R*)b.p, nullptr) != hipSuccess)
    return 9;
  if (hipMemcpy(W, b.p, bytes, hipMemcpyDeviceToHost) != hipSuccess) return 9;
  return 0;
}

template <class R>
int speed_batch(int dir, int n, const void* W, void* out) {
  DevBuf a, b;
  if (hipMalloc(&a.p, (size_t)n * 9 * sizeof(R)) != hipSuccess ||
      hipMalloc(&b.p, (size_t)n * sizeof(R)) != hipSuccess)
    return 9;
  if (hipMemcpy(a.p, W, (size_t)n * 9 * sizeof(R), hipMemcpyHostToDevice) != hipSuccess) return 9;
  if (mhd::launch_fast_speed_batch<R>(n, (const R*)a.p, dir, (R*)b.p, nullptr) != hipSuccess)
    return 9;
  if (hipMemcpy(out, b.p, (size_t)n * sizeof(R), hipMemcpyDeviceToHost) != hipSuccess) return 9;
  return 0;
}

}  // namespace

extern "C" {

int mhd_glm_flux_batch(int precision, int dir, int n, const void* UL, const void* UR,
                       double ch, int hlld, double hlld_tol, void* F, int* status) {
  if (n < 0 || (dir != 0 && dir != 1) || (n > 0 && (!UL || !UR || !F))) return 1;
  if (n == 0) return 0;
  if (precision == MHD_PREC_FP32)
    return flux_batch<float>(dir, n, UL, UR, ch, hlld, hlld_tol, F, status);
  if (precision == MHD_PREC_FP64)
    return flux_batch<double>(dir, n, UL, UR, ch, hlld, hlld_tol, F, status);
  return 1;
}

int mhd_cons_to_prim_batch(int precision, int n, const void* U, void* W) {
  if (n < 0 || (n > 0 && (!U || !W))) return 1;
  if (n == 0) return 0;
  if (precision == MHD_PREC_FP32) return c2p_batch<float>(n, U, W);
  if (precision == MHD_PREC_FP64) return c2p_batch<double>(n, U, W);
  return 1;
}

int mhd_fast_speed_batch(int precision, int dir, int n, const void* W, void* out) {
  if (n < 0 || (dir != 0 && dir != 1) || (n > 0 && (!W || !out))) return 1;
  if (n == 0) return 0;
  if (precision == MHD_PREC_FP32) return speed_batch<float>(dir, n, W, out);
  if (precision == MHD_PREC_FP64) return speed_batch<double>(dir, n, W, out);
  return 1;
}

}  // extern "C"
