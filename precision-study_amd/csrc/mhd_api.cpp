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
#include "mhd_select.h"

#include <hip/hip_runtime_api.h>

#include <dlfcn.h>

#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <cstddef>
#include <numeric>
#include <random>
#include <string>
#include <vector>

using mhd::Ctrl;
using mhd::GridDesc;

namespace {

// Rollback of the MIXED form (mhd_cfg.rollback_steps): device copies of one
// instance's state with host copies of its control block.  Slot 0 holds the
// state given at set_state (or a loaded checkpoint) for the whole run; slots
// 1 and 2 take turns holding snapshots of the run, every `every` steps.
struct Rollback {
  int every = 0;                // snapshot interval and window margin (0: off)
  void* slot[3] = {nullptr, nullptr, nullptr};
  std::vector<char> ctrl[3];
  int step[3] = {-1, -1, -1};   // step of the slot's state (-1: empty)
  int newest = 0;               // slot written last
  int rollbacks = 0;            // windows re-run in FP64
  int restarts = 0;             // restarts from slot 0 in FP64 throughout
  bool all_exact = false;       // the run is FP64 from slot 0 on
};

// More windows than this in one run restart it in FP64 throughout.
constexpr int kMaxRollbacks = 16;

}  // namespace

struct mhd_ctx {
  mhd_cfg cfg;
  GridDesc grid;
  const mhd::Launchers* L = nullptr;
  hipStream_t stream = nullptr;
  void* u = nullptr;       // device state U   (9 planes)
  void* u1 = nullptr;      // device state U1  (9 planes)
  // Two-stage steps (cfg.fuse 2): the state alternates between u and ualt
  // (Ctrl::cur); all_tiles lists every tile for the mixed form's
  // recomputation of a voted-out step.
  bool fused = false;
  void* ualt = nullptr;
  int* all_tiles = nullptr;
  void* ctrl = nullptr;    // device Ctrl<R>
  void* ctrl_host = nullptr;  // pinned host copy of the control block
  int* redo_lists = nullptr;  // MIXED FP64 fallback: 2 x ntiles tile indices
  size_t plane = 0;
  bool have_state = false;
  hipGraphExec_t graph = nullptr;  // `batch` steps, replayed
  int graph_steps = 0;
  std::string err;
  // Learned precision selector (MIXED form): device buffers, allocated on
  // first use.
  struct Selector {
    bool on = false;
    long long since = 0;            // selected steps enqueued since configure
    int ntiles = 0, tx = 16, ty = 15;
    float threshold = 0.5f, audit_frac = 0.01f, audit_budget = 1e-6f;
    float* params = nullptr;        // kParams FP32 weights
    float* mean = nullptr;          // NFEAT
    float* istd = nullptr;          // NFEAT
    float* feats = nullptr;         // ntiles x NIN
    float* err = nullptr;           // ntiles
    unsigned char* flags = nullptr;   // FP64 routing of the step
    unsigned char* sticky = nullptr;  // tiles forced to FP64 by an audit miss
    int* lists = nullptr;           // 5 x ntiles: the step's FP64 tiles, audit sample, mixed
                                    // tiles, the mixed launches' slots, each mixed tile's slot
    int* counts = nullptr;          // lengths of the three lists
    double* sumsq = nullptr;        // error scale (per-plane sum of squares)
    unsigned long long* counters = nullptr;  // audits, misses, FP64 tile-steps, steps
    void* u1x = nullptr;            // FP64 scratch stage output
    void* ctrl_scratch = nullptr;   // control-block snapshot for side evaluations
    // Concurrent schedule ($MHD_SELECTOR_CONCURRENT=1): the FP64 list
    // launch of the flagged and audit tiles runs on `side` beside the mixed
    // stage kernel (fork and join events, captured into the step's graph).
    hipStream_t side = nullptr;
    hipEvent_t fork[2] = {nullptr, nullptr}, join[2] = {nullptr, nullptr};
  } sel;
  Rollback rb;
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
    case MHD_PREC_REF_MIXED: return mhd::launchers_exact(mhd::MODE_REF_MIXED);
    case MHD_PREC_MIXED_RECON64: return mhd::launchers_mixed(mhd::MODE_MIXED_RECON64);
    case MHD_PREC_MIXED_FLUX64: return mhd::launchers_mixed(mhd::MODE_MIXED_FLUX64);
    case MHD_PREC_MIXED_FLUX64_NARROW:
      return mhd::launchers_mixed(mhd::MODE_MIXED_FLUX64_NARROW);
    default: return nullptr;
  }
}

// The allocation variants of the mixed form run single grids with MUSCL +
// HLLD on 16x15 tiles, one stage per launch.
bool is_variant(int prec) {
  return prec == MHD_PREC_MIXED_RECON64 || prec == MHD_PREC_MIXED_FLUX64 ||
         prec == MHD_PREC_MIXED_FLUX64_NARROW;
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
  h->rate_f32 = cfg.precision == MHD_PREC_REF_MIXED ? 1 : 0;
  // Test hook of the mixed form's rollback: $MHD_FAULT_STEP = s makes stage
  // 1 of step s (counted from the state set) halt with code 3 at cell 0,
  // as an FP64-confirmed failure would, unless a rollback window covers it.
  if (const char* f = std::getenv("MHD_FAULT_STEP")) {
    const int s = std::atoi(f);
    if (s >= 0 && s < INT_MAX) h->fail_inject = s + 1;
  }
}

template <class R>
void fill_status_of(int hlld, int nx, int ny, const Ctrl<R>* h, mhd_status* st) {
  std::memset(st, 0, sizeof(*st));
  st->code = (h->run == mhd::RUN_HALTED) ? h->code : 0;
  st->done = (h->run == mhd::RUN_DONE) ? 1 : 0;
  st->steps = h->step;
  st->t = (double)h->t;
  st->dt = (double)(h->run == mhd::RUN_HALTED ? h->fail_dt : h->dt);
  st->ch_last = (double)(h->run == mhd::RUN_HALTED ? h->fail_ch : h->ch);
  const unsigned long long steps = (unsigned long long)h->step;
  if (hlld) {
    // The reference evaluates (nx+1)*ny x faces and nx*(ny+1) y faces per
    // stage (mhd2d_glm_cpu.cpp:253-271), two stages per step.
    st->x_evaluations = 2ull * steps * (unsigned long long)(nx + 1) * ny;
    st->y_evaluations = 2ull * steps * (unsigned long long)nx * (ny + 1);
    st->degenerate = h->degenerate;
    st->nonphysical = h->nonphysical;
  }
  st->bad_i = st->bad_j = -1;
  if (h->run == mhd::RUN_HALTED && (h->code == 3 || h->code == 4) && h->fail_cell >= 0) {
    st->bad_i = h->fail_cell % nx;
    st->bad_j = h->fail_cell / nx;
  }
  st->fail_step = h->fail_step;
  st->fallback_tiles = h->redo_tiles;
  st->fused_redo_steps = h->fz_redo_steps;
  st->exact_steps = h->exact_steps;
}

template <class R>
void fill_status(const mhd_ctx* c, const Ctrl<R>* h, mhd_status* st) {
  fill_status_of(c->cfg.hlld, c->cfg.nx, c->cfg.ny, h, st);
}

// Host <-> device state transfer in the caller's SoA layout (element type
// float for FP32, double otherwise); the compensated form stores hi and lo
// FP32 planes on the device.
hipError_t upload_state(const mhd::Launchers* L, void* dev, size_t plane, size_t cells,
                        const void* soa9, size_t ld, hipStream_t s) {
  const size_t es = L->state_elem;
  hipError_t e = hipSuccess;
  if (L->state_planes == 18) {
    std::vector<float> hl(18 * cells);
    const double* src = (const double*)soa9;
    for (int k = 0; k < 9; k++)
      for (size_t q = 0; q < cells; q++) {
        const double x = src[k * ld + q];
        const float hi = (float)x;
        hl[k * cells + q] = hi;
        hl[(9 + k) * cells + q] = (float)(x - (double)hi);
      }
    for (int k = 0; k < 18 && e == hipSuccess; k++)
      e = hipMemcpy((char*)dev + k * plane * es, hl.data() + k * cells, cells * es,
                    hipMemcpyHostToDevice);
    return e;
  }
  for (int k = 0; k < 9 && e == hipSuccess; k++)
    e = hipMemcpyAsync((char*)dev + k * plane * es, (const char*)soa9 + k * ld * es, cells * es,
                       hipMemcpyHostToDevice, s);
  return e;
}

hipError_t download_state(const mhd::Launchers* L, const void* dev, size_t plane, size_t cells,
                          void* soa9, size_t ld, hipStream_t s) {
  const size_t es = L->state_elem;
  hipError_t e = hipSuccess;
  if (L->state_planes == 18) {
    std::vector<float> hl(18 * cells);
    for (int k = 0; k < 18 && e == hipSuccess; k++)
      e = hipMemcpyAsync(hl.data() + k * cells, (const char*)dev + k * plane * es, cells * es,
                         hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    double* out = (double*)soa9;
    for (int k = 0; k < 9; k++)
      for (size_t q = 0; q < cells; q++)
        out[k * ld + q] = (double)hl[k * cells + q] + (double)hl[(9 + k) * cells + q];
    return hipSuccess;
  }
  for (int k = 0; k < 9 && e == hipSuccess; k++)
    e = hipMemcpyAsync((char*)soa9 + k * ld * es, (const char*)dev + k * plane * es, cells * es,
                       hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e;
}

int read_ctrl(mhd_ctx* c) {
  HIP_OR_FAIL(c, hipMemcpyAsync(c->ctrl_host, c->ctrl, c->L->ctrl_bytes,
                                hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
  return 0;
}

// Two-stage steps leave the state in either buffer (Ctrl::cur).  Moves it
// back to `u` (one device copy) so that readers of the state and the
// single-stage paths find it there.
int settle_state(mhd_ctx* c) {
  if (!c->fused || !c->have_state) return 0;
  if (int rc = read_ctrl(c)) return rc;
  const size_t off =
      c->L->time_is_float ? offsetof(Ctrl<float>, cur) : offsetof(Ctrl<double>, cur);
  int* hcur = (int*)((char*)c->ctrl_host + off);
  if (*hcur == 0) return 0;
  const size_t bytes = (size_t)c->L->state_planes * c->plane * c->L->state_elem;
  HIP_OR_FAIL(c, hipMemcpyAsync(c->u, c->ualt, bytes, hipMemcpyDeviceToDevice, c->stream));
  *hcur = 0;
  HIP_OR_FAIL(c, hipMemcpyAsync((char*)c->ctrl + off, hcur, sizeof(int), hipMemcpyHostToDevice,
                                c->stream));
  HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
  return 0;
}

int status_of(mhd_ctx* c, mhd_status* st) {
  mhd_status tmp;
  if (c->L->time_is_float)
    fill_status(c, (const Ctrl<float>*)c->ctrl_host, &tmp);
  else
    fill_status(c, (const Ctrl<double>*)c->ctrl_host, &tmp);
  tmp.rollbacks = c->rb.rollbacks;
  tmp.restarts = c->rb.restarts;
  if (st) *st = tmp;
  return tmp.code;
}

// --- Rollback of the MIXED form ---------------------------------------------
bool rollback_enabled(const mhd_cfg& cfg) {
  return cfg.precision == MHD_PREC_MIXED && cfg.fallback == 0 && cfg.rollback_steps >= 0;
}

int rollback_interval(const mhd_cfg& cfg) {
  return cfg.rollback_steps > 0 ? cfg.rollback_steps : 256;
}

hipError_t rb_alloc(Rollback& rb, size_t bytes) {
  for (void*& p : rb.slot)
    if (!p)
      if (hipError_t e = hipMalloc(&p, bytes)) return e;
  return hipSuccess;
}

void rb_free(Rollback& rb) {
  for (void*& p : rb.slot)
    if (p) {
      (void)hipFree(p);
      p = nullptr;
    }
}

int ctrl_step(const void* ctrl_host) { return ((const Ctrl<double>*)ctrl_host)->step; }

// Starts the record of a run: slot 0 takes the given state (a device copy
// on the stream) and its control block as just read back.
hipError_t rb_start(Rollback& rb, const void* u, size_t bytes, const void* ctrl_host, size_t cb,
                    hipStream_t s) {
  rb.ctrl[0].assign((const char*)ctrl_host, (const char*)ctrl_host + cb);
  rb.step[0] = ctrl_step(ctrl_host);
  rb.step[1] = rb.step[2] = -1;
  rb.newest = 0;
  rb.rollbacks = rb.restarts = 0;
  rb.all_exact = false;
  return hipMemcpyAsync(rb.slot[0], u, bytes, hipMemcpyDeviceToDevice, s);
}

bool rb_snapshot_due(const Rollback& rb, const void* ctrl_host) {
  return ctrl_step(ctrl_host) - rb.step[rb.newest] >= rb.every;
}

// Snapshot of a running state into the rotating slot not written last.
hipError_t rb_snapshot(Rollback& rb, const void* u, size_t bytes, const void* ctrl_host,
                       size_t cb, hipStream_t s) {
  const int k = rb.newest == 1 ? 2 : 1;
  rb.ctrl[k].assign((const char*)ctrl_host, (const char*)ctrl_host + cb);
  rb.step[k] = ctrl_step(ctrl_host);
  rb.newest = k;
  return hipMemcpyAsync(rb.slot[k], u, bytes, hipMemcpyDeviceToDevice, s);
}

// How a run that halted with code 3/4 continues: the slot to restore and the
// step up to which every tile then runs in FP64.  A halt inside such a
// window (an FP64-confirmed failure of the drifted state), or the
// kMaxRollbacks-th one, restarts from slot 0 in FP64 throughout; a halt of
// that run stands (false), as it would in the FP64 form.  Otherwise the
// newest snapshot at least `every` steps before the failure is restored
// (slot 0 if none) and re-run in FP64 until `every` steps past it.
bool rb_plan(Rollback& rb, const void* ctrl_host, int& slot, int& exact_until) {
  if (rb.all_exact) return false;
  const Ctrl<double>* h = (const Ctrl<double>*)ctrl_host;
  const int fail = h->fail_step;
  slot = 0;
  if (fail < h->exact_until || rb.rollbacks >= kMaxRollbacks) {
    exact_until = INT_MAX;
    rb.all_exact = true;
    rb.restarts += 1;
    rb.step[1] = rb.step[2] = -1;
    rb.newest = 0;
    return true;
  }
  for (int k = 1; k <= 2; k++)
    if (rb.step[k] >= 0 && rb.step[k] <= fail - rb.every && rb.step[k] > rb.step[slot]) slot = k;
  exact_until = fail > INT_MAX - rb.every ? INT_MAX : fail + rb.every;
  rb.rollbacks += 1;
  // Snapshots later than the restored one belong to the abandoned trajectory.
  for (int k = 1; k <= 2; k++)
    if (k != slot && rb.step[k] > rb.step[slot]) rb.step[k] = -1;
  rb.newest = slot;
  return true;
}

// Restores `slot` into a state buffer and its device control block (through
// the pinned host copy `ctrl_host`), opens the FP64 window and re-arms the
// step for t_end.  The window's first c_h comes from the FP64 form's signal
// speed of the restored state (the snapshot's was the mixed form's), so a
// restart from the given state is the FP64 run bit for bit.
hipError_t rb_restore(const Rollback& rb, int slot, int exact_until, const GridDesc& g, void* u,
                      size_t bytes, void* ctrl, void* ctrl_host, size_t cb, double t_end,
                      hipStream_t s) {
  const mhd::Launchers* F64 = mhd::launchers_exact(mhd::MODE_FP64);
  hipError_t e = hipMemcpyAsync(u, rb.slot[slot], bytes, hipMemcpyDeviceToDevice, s);
  std::memcpy(ctrl_host, rb.ctrl[slot].data(), cb);
  ((Ctrl<double>*)ctrl_host)->exact_until = exact_until;
  if (e == hipSuccess) e = hipMemcpyAsync(ctrl, ctrl_host, cb, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = F64->speed(g, u, ctrl, s);
  if (e == hipSuccess) e = F64->commit(ctrl, s);
  if (e == hipSuccess) e = F64->set_target(ctrl, t_end, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e;
}

size_t state_bytes(const mhd_ctx* c) {
  return (size_t)c->L->state_planes * c->plane * c->L->state_elem;
}

// Steps between refreshes of the selector's tile features and flags (a
// tile's character changes slowly: information crosses less than one cell
// per step).
constexpr int kSelectorRefresh = 8;
static_assert(mhd::RUN_ACTIVE == 0, "sel_audit_pick treats run state 0 as active");

// One step with the learned selector: tile features -> MLP -> FP64 flags
// (every kSelectorRefresh steps); the mixed kernels take the unflagged tiles
// and the FP64 kernels the flagged ones; a random sample of the unflagged
// tiles is recomputed by the FP64 kernel into scratch and compared (an audit
// miss makes the tile FP64 from then on).
hipError_t enqueue_step_selected(mhd_ctx* c) {
  auto& S = c->sel;
  const mhd::Launchers* F64 = mhd::launchers_exact(mhd::MODE_FP64);
  const char* ctrl = (const char*)c->ctrl;
  hipError_t e = hipSuccess;
  if (S.since++ % kSelectorRefresh == 0) {
    e = mhd::sel_tile_features((const double*)c->u, c->plane, c->cfg.nx, c->cfg.ny, S.tx, S.ty,
                               ctrl + offsetof(Ctrl<double>, ch), S.feats, c->stream);
    if (e == hipSuccess)
      e = mhd::sel_mlp_infer(S.feats, S.ntiles, S.params, S.mean, S.istd, S.threshold, S.sticky,
                             nullptr, S.flags, c->stream);
    if (e == hipSuccess) e = mhd::sel_route(S.flags, S.ntiles, S.lists, S.counts, c->stream);
  }
  // FP64 tiles and mixed tiles as device lists (rebuilt with the flags), and
  // the step's audit sample (from the pick, which also snapshots the control
  // block for the audits).  The mixed stages take the mixed tiles from the
  // slots the routing dealt them (each XCD an equal part of the mixed list,
  // one load per block), so flagged tiles cost the mixed launch nothing
  // wherever they cluster; they only queue their fallback tiles here.  One
  // list launch per stage evaluates in FP64, over a small fixed grid: the
  // flagged tiles, the tiles the mixed stage queued for its fallback, and
  // (stage 1) the audit tiles into scratch.
  int* fp64_list = S.lists;
  int* audit_list = S.lists + S.ntiles;
  const int* fp64_count = S.counts;
  const int* audit_count = S.counts + 1;
  mhd::GridDesc gm = c->grid;
  gm.tiles = S.lists + 3 * S.ntiles;
  gm.redo = nullptr;
  auto redo_count = [&](int stage) {
    return (const int*)((char*)c->ctrl + offsetof(Ctrl<double>, redo_count) +
                        (stage - 1) * sizeof(int));
  };
  if (e == hipSuccess)
    e = mhd::sel_audit_pick(S.ntiles, (const int*)(ctrl + offsetof(Ctrl<double>, step)),
                            (const int*)(ctrl + offsetof(Ctrl<double>, run)), S.audit_frac, S.lists, S.counts, S.counters, c->ctrl, S.ctrl_scratch,
                            c->L->ctrl_bytes, c->stream);
  // Concurrent schedule: the flagged (and, in stage 1, audit) tiles on the
  // side stream beside the mixed stage kernel.  Their cells are disjoint
  // from the mixed list's, they read only what the previous stage left, and
  // they meet the mixed kernel only in the control block's order-independent
  // atomics (maximum speed, first bad cell, counters), so the step's results
  // are those of the serial schedule.  The fallback queue, which the mixed
  // kernel fills, runs after it on the main stream.
  const bool conc = S.side != nullptr;
  if (conc && e == hipSuccess) {
    e = hipEventRecord(S.fork[0], c->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(S.side, S.fork[0], 0);
    if (e == hipSuccess) {
      const mhd::ListJob jobs[2] = {{c->u, c->u, c->u1, c->ctrl, fp64_list, fp64_count},
                                    {c->u, c->u, S.u1x, S.ctrl_scratch, audit_list, audit_count}};
      e = F64->stage_listn(c->grid, 1, jobs, 2, S.side);
    }
    if (e == hipSuccess) e = hipEventRecord(S.join[0], S.side);
  }
  if (e == hipSuccess) e = c->L->stage(gm, 1, c->u, c->u, c->u1, c->ctrl, c->stream);
  if (e == hipSuccess && conc) {
    if (c->grid.redo_list[0]) {
      const mhd::ListJob jobs[1] = {
          {c->u, c->u, c->u1, c->ctrl, c->grid.redo_list[0], redo_count(1)}};
      e = F64->stage_listn(c->grid, 1, jobs, 1, c->stream);
    }
    if (e == hipSuccess) e = hipStreamWaitEvent(c->stream, S.join[0], 0);
  } else if (e == hipSuccess) {
    const mhd::ListJob jobs[3] = {
        {c->u, c->u, c->u1, c->ctrl, fp64_list, fp64_count},
        {c->u, c->u, S.u1x, S.ctrl_scratch, audit_list, audit_count},
        {c->u, c->u, c->u1, c->ctrl, c->grid.redo_list[0], redo_count(1)}};
    e = F64->stage_listn(c->grid, 1, jobs, c->grid.redo_list[0] ? 3 : 2, c->stream);
  }
  if (e == hipSuccess)
    e = mhd::sel_audit((const double*)c->u1, (const double*)S.u1x, c->plane, c->cfg.nx, c->cfg.ny,
                       S.tx, S.ty, S.sumsq, S.lists, S.counts, S.audit_budget, S.sticky, S.flags,
                       S.err, S.counters, c->stream);
  if (conc && e == hipSuccess) {
    e = hipEventRecord(S.fork[1], c->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(S.side, S.fork[1], 0);
    if (e == hipSuccess) {
      const mhd::ListJob jobs[1] = {{c->u1, c->u, c->u, c->ctrl, fp64_list, fp64_count}};
      e = F64->stage_listn(c->grid, 2, jobs, 1, S.side);
    }
    if (e == hipSuccess) e = hipEventRecord(S.join[1], S.side);
  }
  if (e == hipSuccess) e = c->L->stage(gm, 2, c->u1, c->u, c->u, c->ctrl, c->stream);
  if (e == hipSuccess && conc) {
    if (c->grid.redo_list[1]) {
      const mhd::ListJob jobs[1] = {
          {c->u1, c->u, c->u, c->ctrl, c->grid.redo_list[1], redo_count(2)}};
      e = F64->stage_listn(c->grid, 2, jobs, 1, c->stream);
    }
    if (e == hipSuccess) e = hipStreamWaitEvent(c->stream, S.join[1], 0);
  } else if (e == hipSuccess) {
    const mhd::ListJob jobs[2] = {
        {c->u1, c->u, c->u, c->ctrl, fp64_list, fp64_count},
        {c->u1, c->u, c->u, c->ctrl, c->grid.redo_list[1], redo_count(2)}};
    e = F64->stage_listn(c->grid, 2, jobs, c->grid.redo_list[1] ? 2 : 1, c->stream);
  }
  if (e == hipSuccess) e = c->L->finish(c->ctrl, c->stream);
  return e;
}

hipError_t enqueue_step(mhd_ctx* c, hipEvent_t* ev = nullptr) {
  if (c->sel.on && !ev) return enqueue_step_selected(c);
  if (c->fused)
    return c->L->step_fused(c->grid, c->u, c->ualt, c->u1, c->ctrl, c->all_tiles, c->stream, ev);
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

// Tile shape and stages per launch chosen by the autotuner (mhdglm.autotune)
// for a form and grid size.  The table is the text file `autotune.txt` next
// to the library's lib/ directory (or $MHD_AUTOTUNE_FILE): lines "form
// max_cells tile [fuse]", the first line with the form and max_cells >=
// cells wins; '#' starts a comment.  Without a match the built-in default
// tile and one stage per launch are used.
struct TuneRow {
  int form;
  long long max_cells;
  int tile;
  int fuse;  // stages per launch (1 or 2)
};

const std::vector<TuneRow>& tune_table() {
  static const std::vector<TuneRow> rows = [] {
    std::vector<TuneRow> out;
    std::string path;
    if (const char* env = std::getenv("MHD_AUTOTUNE_FILE")) {
      path = env;
    } else {
      Dl_info info;
      if (dladdr((void*)&tune_table, &info) && info.dli_fname) {
        std::string lib = info.dli_fname;
        const size_t slash = lib.rfind('/');
        const std::string dir = slash == std::string::npos ? "." : lib.substr(0, slash);
        path = dir + "/../autotune.txt";
      }
    }
    FILE* f = path.empty() ? nullptr : std::fopen(path.c_str(), "r");
    if (!f) return out;
    char line[256];
    while (std::fgets(line, sizeof(line), f)) {
      if (line[0] == '#') continue;
      TuneRow r;
      r.fuse = 1;
      const int n = std::sscanf(line, "%d %lld %d %d", &r.form, &r.max_cells, &r.tile, &r.fuse);
      if (n >= 3 && r.tile >= 0 && r.tile < mhd::TILE_COUNT && (r.fuse == 1 || r.fuse == 2))
        out.push_back(r);
    }
    std::fclose(f);
    return out;
  }();
  return rows;
}

const TuneRow* tuned_row(int form, long long cells) {
  for (const TuneRow& r : tune_table())
    if (r.form == form && cells <= r.max_cells) return &r;
  return nullptr;
}

int tuned_tile(int form, long long cells) {
  const TuneRow* r = tuned_row(form, cells);
  return r ? r->tile : mhd::kDefaultTile;
}

int tuned_fuse(int form, long long cells) {
  const TuneRow* r = tuned_row(form, cells);
  return r ? r->fuse : 1;
}

bool use_graphs() {
  const char* s = std::getenv("MHD_GRAPHS");
  return !(s && s[0] == '0');
}

template <class R>
void problem_soa(const char* problem, int nx, int ny, R* out, size_t ld,
                 const double* params = nullptr) {
  const R dx = ((R)1.0 - (R)0.0) / (R)nx, dy = ((R)1.0 - (R)0.0) / (R)ny;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      R P[9], U[9];
      mhd::problem_primitive<R>(problem, i, j, dx, dy, P, params);
      mhd::host_prim_to_cons<R>(P, U);
      const size_t q = (size_t)j * nx + i;
      for (int k = 0; k < 9; k++) out[k * ld + q] = U[k];
    }
}

// A checkpoint's control block as the device may take it: the run state a
// RunState, each first-bad-cell index INT_MAX or a cell of the grid; the
// per-step queue counts and two-stage flags cleared (a halted run's are not
// reset by the next target), the state in the context's own buffer (cur 0,
// as mhd_checkpoint_save leaves it) and the loading context's own test hook.
// Returns false for a block no save could have written.
template <class R>
bool sane_ctrl(Ctrl<R>* h, const Ctrl<R>* mine, size_t cells) {
  if (h->run != mhd::RUN_ACTIVE && h->run != mhd::RUN_DONE && h->run != mhd::RUN_HALTED)
    return false;
  for (int k = 0; k < 2; k++)
    if (h->bad[k] != INT_MAX && (h->bad[k] < 0 || (size_t)h->bad[k] >= cells)) return false;
  if (h->cur != 0 && h->cur != 1) return false;
  h->cur = 0;
  h->redo_count[0] = 0;
  h->redo_count[1] = 0;
  h->fz_fail = 0;
  h->fz_count = 0;
  h->fail_inject = mine->fail_inject;
  return true;
}

}  // namespace

extern "C" {

int mhd_autotuned_tile(int precision, int nx, int ny) {
  return tuned_tile(precision, (long long)nx * ny);
}

int mhd_autotuned_fuse(int precision, int nx, int ny) {
  return tuned_fuse(precision, (long long)nx * ny);
}

const char* mhd_last_error(const mhd_ctx* ctx) {
  return ctx ? ctx->err.c_str() : kNoCtx;
}

int mhd_create(const mhd_cfg* cfg, mhd_ctx** out) {
  if (!out) return 1;
  *out = nullptr;
  if (!cfg) return 1;
  if (cfg->nx < 1 || cfg->ny < 1) return 1;
  if (!(cfg->cfl > 0.0 && cfg->cfl < 1.0)) return 1;
  if (!(cfg->alpha > 0.0)) return 1;
  if (!launchers_for(cfg->precision)) return 1;
  if (cfg->precision == MHD_PREC_REF_MIXED && !cfg->muscl) return 1;
  if ((long long)cfg->nx * cfg->ny > (long long)INT_MAX) return 1;
  if (cfg->fuse < 0 || cfg->fuse > 2) return 1;
  if (is_variant(cfg->precision) &&
      (!cfg->muscl || !cfg->hlld || cfg->fuse == 2 ||
       (cfg->tile >= 0 && cfg->tile < mhd::TILE_COUNT && cfg->tile != mhd::TILE_16x15)))
    return 1;
  const long long ncells = (long long)cfg->nx * cfg->ny;
  int tile = cfg->tile, fuse = cfg->fuse;
  if (tile < 0 || tile >= mhd::TILE_COUNT) {
    tile = tuned_tile(cfg->precision, ncells);
    if (fuse == 0) fuse = tuned_fuse(cfg->precision, ncells);
  }
  if (fuse == 0) fuse = 1;
  // Only the MUSCL + HLLD scheme has every tile shape; the others run the
  // default tile (the selector's per-tile flags follow this shape).
  if (!(cfg->muscl && cfg->hlld)) tile = mhd::kDefaultTile;
  // Two stages per launch: MUSCL + HLLD on 16x15 tiles, not in the
  // fallback's check mode.  Asked for explicitly elsewhere, it is an error;
  // an autotuned choice falls back to one stage per launch.
  const bool can_fuse =
      cfg->muscl && cfg->hlld && tile == mhd::TILE_16x15 && cfg->fallback != 2;
  if (fuse == 2 && !can_fuse) {
    if (cfg->fuse == 2) return 1;
    fuse = 1;
  }
  mhd_ctx* c = new mhd_ctx();
  c->cfg = *cfg;
  c->L = launchers_for(cfg->precision);
  if (!(c->cfg.lowbeta_ratio > 0.0)) c->cfg.lowbeta_ratio = 16.0;
  if (c->cfg.batch_steps <= 0) c->cfg.batch_steps = 32;
  if (c->cfg.hlld_tol <= 0.0) c->cfg.hlld_tol = 1.0e-4;
  c->cfg.tile = tile;
  c->cfg.fuse = fuse;
  c->fused = fuse == 2;
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
  // FP64 fallback of the mixed form: one queue of tile indices per stage.
  if (cfg->precision == MHD_PREC_MIXED && cfg->fallback != 1) {
    int tx, ty;
    mhd::tile_dims(c->cfg.tile, tx, ty);
    const size_t ntiles = (size_t)((cfg->nx + tx - 1) / tx) * ((cfg->ny + ty - 1) / ty);
    if ((e = hipMalloc((void**)&c->redo_lists, 2 * ntiles * sizeof(int))) != hipSuccess) {
      std::fprintf(stderr, "mhd_create: %s\n", hipGetErrorString(e));
      mhd_destroy(c);
      return 9;
    }
    c->grid.redo_list[0] = c->redo_lists;
    c->grid.redo_list[1] = c->redo_lists + ntiles;
    c->grid.redo = mhd::launchers_exact(mhd::MODE_FP64)->stage_list;
    c->grid.redo_all = cfg->fallback == 2 ? 1 : 0;
  }
  // Rollback snapshots of the mixed form.
  if (rollback_enabled(c->cfg)) {
    c->rb.every = rollback_interval(c->cfg);
    if ((e = rb_alloc(c->rb, bytes)) != hipSuccess) {
      std::fprintf(stderr, "mhd_create: %s\n", hipGetErrorString(e));
      mhd_destroy(c);
      return 9;
    }
  }
  // Two-stage steps: the alternate state buffer, and (mixed form with the
  // fallback) the list of every tile for the recomputation of a step.
  if (c->fused) {
    e = hipMalloc(&c->ualt, bytes);
    if (e == hipSuccess && c->redo_lists) {
      int tx, ty;
      mhd::tile_dims(c->cfg.tile, tx, ty);
      const int ntiles = ((cfg->nx + tx - 1) / tx) * ((cfg->ny + ty - 1) / ty);
      std::vector<int> all(ntiles);
      for (int q = 0; q < ntiles; q++) all[q] = q;
      e = hipMalloc((void**)&c->all_tiles, (size_t)ntiles * sizeof(int));
      if (e == hipSuccess)
        e = hipMemcpy(c->all_tiles, all.data(), (size_t)ntiles * sizeof(int),
                      hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
      std::fprintf(stderr, "mhd_create: %s\n", hipGetErrorString(e));
      mhd_destroy(c);
      return 9;
    }
  }
  *out = c;
  return 0;
}

void mhd_destroy(mhd_ctx* c) {
  if (!c) return;
  if (c->graph) (void)hipGraphExecDestroy(c->graph);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  // Side work is joined into the main stream, except after a failed launch
  // between a fork and its join.
  if (c->sel.side) (void)hipStreamSynchronize(c->sel.side);
  for (void* p : {(void*)c->sel.params, (void*)c->sel.mean, (void*)c->sel.istd,
                  (void*)c->sel.feats, (void*)c->sel.err, (void*)c->sel.flags,
                  (void*)c->sel.sticky, (void*)c->sel.lists, (void*)c->sel.counts,
                  (void*)c->sel.sumsq,
                  (void*)c->sel.counters, c->sel.u1x, c->sel.ctrl_scratch})
    if (p) (void)hipFree(p);
  if (c->u) (void)hipFree(c->u);
  if (c->u1) (void)hipFree(c->u1);
  if (c->ualt) (void)hipFree(c->ualt);
  if (c->all_tiles) (void)hipFree(c->all_tiles);
  if (c->redo_lists) (void)hipFree(c->redo_lists);
  if (c->ctrl) (void)hipFree(c->ctrl);
  if (c->ctrl_host) (void)hipHostFree(c->ctrl_host);
  rb_free(c->rb);
  for (int k = 0; k < 2; k++) {
    if (c->sel.fork[k]) (void)hipEventDestroy(c->sel.fork[k]);
    if (c->sel.join[k]) (void)hipEventDestroy(c->sel.join[k]);
  }
  if (c->sel.side) (void)hipStreamDestroy(c->sel.side);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int mhd_problem_state_ex(const char* problem, const double* params, int nparams, int nx,
                         int ny, int precision, void* soa9, size_t ld) {
  if (!mhd::problem_known(problem) || nx < 1 || ny < 1 || !soa9) return 1;
  if (ld < (size_t)nx * ny) return 1;
  if (nparams != mhd::problem_nparams(problem) || (nparams > 0 && !params)) return 1;
  if (precision == MHD_PREC_FP32)
    problem_soa<float>(problem, nx, ny, (float*)soa9, ld, params);
  else if (launchers_for(precision))
    problem_soa<double>(problem, nx, ny, (double*)soa9, ld, params);
  else
    return 1;
  return 0;
}

int mhd_problem_state(const char* problem, int nx, int ny, int precision,
                      void* soa9, size_t ld) {
  return mhd_problem_state_ex(problem, nullptr, 0, nx, ny, precision, soa9, ld);
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
  if (c->rb.every > 0)
    HIP_OR_FAIL(c, rb_start(c->rb, c->u, state_bytes(c), c->ctrl_host, c->L->ctrl_bytes,
                            c->stream));
  c->have_state = true;
  return 0;
}

int mhd_init_problem(mhd_ctx* c, const char* problem) {
  if (!c) return 1;
  if (!mhd::problem_known(problem)) return fail(c, 1, "mhd_init_problem: unknown problem");
  const size_t cells = (size_t)c->cfg.nx * c->cfg.ny;
  if (mhd::problem_nparams(problem) != 0)
    return fail(c, 1, std::string("mhd_init_problem: problem '") + problem +
                          "' needs parameters (use mhd_problem_state_ex + mhd_set_state)");
  std::vector<char> host(9 * cells * io_elem(c->cfg.precision));
  if (int rc = mhd_problem_state(problem, c->cfg.nx, c->cfg.ny, c->cfg.precision, host.data(),
                                 cells))
    return fail(c, rc, "mhd_init_problem: initial condition failed");
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
  const int batch = c->cfg.batch_steps;
  const bool graphs = use_graphs();
  mhd_status s;
  status_of(c, &s);
  // Stop at max_steps more completed steps (steps a rollback re-runs count
  // once).
  const long long target = max_steps < 0 ? LLONG_MAX : (long long)s.steps + max_steps;
  while ((long long)s.steps < target && s.code == 0 && !s.done) {
    const long long left = target - s.steps;
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
    if (int rc = read_ctrl(c)) return rc;
    status_of(c, &s);
    if (c->rb.every <= 0) continue;
    int slot, until;
    if ((s.code == 3 || s.code == 4) && rb_plan(c->rb, c->ctrl_host, slot, until)) {
      HIP_OR_FAIL(c, rb_restore(c->rb, slot, until, c->grid, c->u, state_bytes(c), c->ctrl,
                                c->ctrl_host, c->L->ctrl_bytes, t_end, c->stream));
      if (int rc = read_ctrl(c)) return rc;
      status_of(c, &s);
    } else if (s.code == 0 && !s.done && rb_snapshot_due(c->rb, c->ctrl_host)) {
      if (int rc = settle_state(c)) return rc;
      HIP_OR_FAIL(c, rb_snapshot(c->rb, c->u, state_bytes(c), c->ctrl_host, c->L->ctrl_bytes,
                                 c->stream));
    }
  }
  if (int rc = settle_state(c)) return rc;
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
  if (which == 1 && c->fused)
    return fail(c, 1, "mhd_get_state: two-stage steps keep no U1");
  if (int rc = settle_state(c)) return rc;
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
  if (int rc = settle_state(c)) return rc;
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
  if (int rc = settle_state(c)) return rc;
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
  h.version = 2;
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
            h.version == 2 && h.nx == c->cfg.nx && h.ny == c->cfg.ny &&
            h.precision == c->cfg.precision && h.planes == c->L->state_planes &&
            h.elem == (int)es && h.ctrl_bytes == (int)c->L->ctrl_bytes;
  std::vector<char> buf, ctl;
  if (ok) {
    buf.resize((size_t)h.planes * cells * es);
    ctl.resize(c->L->ctrl_bytes);
    ok = std::fread(ctl.data(), ctl.size(), 1, f) == 1 &&
         std::fread(buf.data(), 1, buf.size(), f) == buf.size();
  }
  std::fclose(f);
  if (!ok) return fail(c, 1, "mhd_checkpoint_load: file does not match this context");
  ok = c->L->time_is_float
           ? sane_ctrl((Ctrl<float>*)ctl.data(), (const Ctrl<float>*)c->ctrl_host, cells)
           : sane_ctrl((Ctrl<double>*)ctl.data(), (const Ctrl<double>*)c->ctrl_host, cells);
  if (!ok) return fail(c, 1, "mhd_checkpoint_load: control block out of range");
  std::memcpy(c->ctrl_host, ctl.data(), ctl.size());
  for (int k = 0; k < h.planes; k++)
    HIP_OR_FAIL(c, hipMemcpy((char*)c->u + k * c->plane * es, buf.data() + k * cells * es,
                             cells * es, hipMemcpyHostToDevice));
  HIP_OR_FAIL(c, hipMemcpy(c->ctrl, c->ctrl_host, c->L->ctrl_bytes, hipMemcpyHostToDevice));
  if (c->rb.every > 0) {
    HIP_OR_FAIL(c, rb_start(c->rb, c->u, state_bytes(c), c->ctrl_host, c->L->ctrl_bytes,
                            c->stream));
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
  }
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
// No-allocation run: a caller's device state advanced with U1, the control
// block and the fallback lists in the fixed scratch array of k_noalloc.hip.
// ---------------------------------------------------------------------------
namespace {

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

int noalloc_reject(const char* why) {
  std::fprintf(stderr, "mhd_run_noalloc: %s\n", why);
  return 1;
}

int noalloc_hip(hipError_t e, const char* what) {
  std::fprintf(stderr, "mhd_run_noalloc: HIP error in %s: %s\n", what, hipGetErrorString(e));
  return 9;
}

}  // namespace

extern "C" int mhd_run_noalloc(const mhd_cfg* cfg_in, void* state, size_t ld, double t_end,
                               long long max_steps, void* stream, mhd_status* st) {
  if (!cfg_in || !state) return noalloc_reject("null configuration or state");
  mhd_cfg cfg = *cfg_in;
  if (cfg.nx < 1 || cfg.ny < 1 || (long long)cfg.nx * cfg.ny > (long long)INT_MAX)
    return noalloc_reject("grid must be at least 1 x 1 and at most INT_MAX cells");
  if (!(cfg.cfl > 0.0 && cfg.cfl < 1.0) || !(cfg.alpha > 0.0))
    return noalloc_reject("cfl must lie in (0,1) and alpha be positive");
  if (cfg.precision == MHD_PREC_MIXED_KAHAN || !launchers_for(cfg.precision))
    return noalloc_reject("unsupported precision (MIXED_KAHAN keeps two planes per variable)");
  if (cfg.precision == MHD_PREC_REF_MIXED && !cfg.muscl)
    return noalloc_reject("REF_MIXED runs MUSCL only");
  if (cfg.fuse == 2) return noalloc_reject("one stage per launch only (fuse 0 or 1)");
  if (is_variant(cfg.precision) &&
      (!cfg.muscl || !cfg.hlld ||
       (cfg.tile >= 0 && cfg.tile < mhd::TILE_COUNT && cfg.tile != mhd::TILE_16x15)))
    return noalloc_reject("the allocation variants run MUSCL + HLLD on 16x15 tiles");
  const mhd::Launchers* L = launchers_for(cfg.precision);
  const size_t cells = (size_t)cfg.nx * cfg.ny;
  if (ld < cells) return noalloc_reject("ld is smaller than nx * ny");
  if ((uintptr_t)state % L->state_elem != 0) return noalloc_reject("misaligned state pointer");
  if (!(cfg.lowbeta_ratio > 0.0)) cfg.lowbeta_ratio = 16.0;
  if (cfg.batch_steps <= 0) cfg.batch_steps = 32;
  if (cfg.hlld_tol <= 0.0) cfg.hlld_tol = 1.0e-4;
  int tile = cfg.tile;
  if (tile < 0 || tile >= mhd::TILE_COUNT) tile = tuned_tile(cfg.precision, (long long)cells);
  if (!(cfg.muscl && cfg.hlld)) tile = mhd::kDefaultTile;
  cfg.tile = tile;
  int tx, ty;
  mhd::tile_dims(tile, tx, ty);
  const size_t ntiles = (size_t)((cfg.nx + tx - 1) / tx) * ((cfg.ny + ty - 1) / ty);
  const bool lists = cfg.precision == MHD_PREC_MIXED && cfg.fallback != 1;
  // Scratch layout: U1 (planes ld apart, like the caller's state), the
  // control block, then the two fallback lists.
  const size_t u1_bytes = align256((size_t)L->state_planes * ld * L->state_elem);
  const size_t ctrl_bytes = align256(L->ctrl_bytes);
  const size_t list_bytes = lists ? align256(2 * ntiles * sizeof(int)) : 0;
  const size_t need = u1_bytes + ctrl_bytes + list_bytes;
  void* base = nullptr;
  size_t cap = 0;
  if (hipError_t e = mhd::noalloc_scratch(&base, &cap)) return noalloc_hip(e, "scratch lookup");
  if (need > cap) {
    char msg[160];
    std::snprintf(msg, sizeof(msg), "the grid needs %zu bytes of scratch; the maximum is %zu",
                  need, cap);
    return noalloc_reject(msg);
  }
  char* sc = (char*)base;
  void* u1 = sc;
  void* ctrl = sc + u1_bytes;
  int* redo = lists ? (int*)(sc + u1_bytes + ctrl_bytes) : nullptr;
  hipStream_t s = (hipStream_t)stream;

  mhd::GridDesc g;
  g.nx = cfg.nx;
  g.ny = cfg.ny;
  g.plane = ld;
  g.muscl = cfg.muscl ? 1 : 0;
  g.hlld = cfg.hlld ? 1 : 0;
  g.alpha = cfg.alpha;
  g.tol = cfg.hlld_tol;
  g.tile = tile;
  g.lowbeta_ratio = cfg.lowbeta_ratio;
  if (lists) {
    g.redo_list[0] = redo;
    g.redo_list[1] = redo + ntiles;
    g.redo = mhd::launchers_exact(mhd::MODE_FP64)->stage_list;
    g.redo_all = cfg.fallback == 2 ? 1 : 0;
  }
  // The control block starts from t = 0 with every counter and speed slot
  // cleared (the array keeps the previous call's contents).
  alignas(16) unsigned char host_ctrl[sizeof(Ctrl<double>) > sizeof(Ctrl<float>)
                                          ? sizeof(Ctrl<double>)
                                          : sizeof(Ctrl<float>)];
  auto read_back = [&]() -> hipError_t {
    hipError_t e = hipMemcpyAsync(host_ctrl, ctrl, L->ctrl_bytes, hipMemcpyDeviceToHost, s);
    return e == hipSuccess ? hipStreamSynchronize(s) : e;
  };
  auto status = [&](mhd_status* out) {
    if (L->time_is_float)
      fill_status_of(cfg.hlld, cfg.nx, cfg.ny, (const Ctrl<float>*)host_ctrl, out);
    else
      fill_status_of(cfg.hlld, cfg.nx, cfg.ny, (const Ctrl<double>*)host_ctrl, out);
  };
  if (L->time_is_float) init_ctrl((Ctrl<float>*)host_ctrl, cfg);
  else init_ctrl((Ctrl<double>*)host_ctrl, cfg);
  hipError_t e = hipMemcpyAsync(ctrl, host_ctrl, L->ctrl_bytes, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);  // host_ctrl is reused below
  if (e == hipSuccess) e = L->speed(g, state, ctrl, s);
  if (e == hipSuccess) e = L->commit(ctrl, s);
  if (e == hipSuccess) e = L->set_target(ctrl, t_end, s);
  if (e == hipSuccess) e = read_back();
  if (e != hipSuccess) return noalloc_hip(e, "start of the run");
  mhd_status cur;
  status(&cur);
  long long left = max_steps < 0 ? LLONG_MAX : max_steps;
  while (left > 0 && cur.code == 0 && !cur.done) {
    const int n = (int)(left < cfg.batch_steps ? left : cfg.batch_steps);
    for (int q = 0; q < n && e == hipSuccess; q++) e = L->step(g, state, u1, ctrl, s, nullptr);
    if (e == hipSuccess) e = read_back();
    if (e != hipSuccess) return noalloc_hip(e, "step");
    left -= n;
    status(&cur);
  }
  if (st) *st = cur;
  return cur.code;
}

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

// ---------------------------------------------------------------------------
// Ensembles.
// ---------------------------------------------------------------------------
struct mhd_ensemble {
  mhd_cfg cfg;
  const mhd::Launchers* L = nullptr;
  int n = 0;
  std::vector<int> nx, ny;
  std::vector<size_t> plane;
  std::vector<void*> u, u1;
  std::vector<char> have;
  void* ctrls = nullptr;       // device, n control blocks
  void* ctrls_host = nullptr;  // pinned copy
  void* desc = nullptr;        // device descriptors
  int* prefix = nullptr;       // device prefix of tile counts
  unsigned* queue = nullptr;   // device work-queue counters
  void* live = nullptr;        // device live list (rebuilt at every step)
  int* redo_items = nullptr;   // MIXED FP64 fallback: 2 x total work items
  std::vector<Rollback> rb;    // MIXED rollback, per instance
  mhd::EnsLaunch el;
  hipStream_t stream = nullptr;
  hipGraphExec_t graph = nullptr;
  int graph_steps = 0;
  std::string err;
};

namespace {

int efail(mhd_ensemble* e, int code, const std::string& msg) {
  if (e) e->err = msg;
  return code;
}

#define EHIP(e, call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return efail(e, 9, std::string("HIP error in ") + #call + ": " + hipGetErrorString(e_)); \
  } while (0)

mhd::GridDesc inst_grid(const mhd_ensemble* e, int i) {
  mhd::GridDesc g;
  g.nx = e->nx[i];
  g.ny = e->ny[i];
  g.plane = e->plane[i];
  g.muscl = e->cfg.muscl ? 1 : 0;
  g.hlld = e->cfg.hlld ? 1 : 0;
  g.alpha = e->cfg.alpha;
  g.tol = e->cfg.hlld_tol;
  g.lowbeta_ratio = e->cfg.lowbeta_ratio;
  g.tile = e->cfg.tile;
  return g;
}

void* ctrl_at(void* base, const mhd_ensemble* e, int i) {
  return (char*)base + (size_t)i * e->L->ctrl_bytes;
}

int ens_read_ctrls(mhd_ensemble* e) {
  EHIP(e, hipMemcpyAsync(e->ctrls_host, e->ctrls, (size_t)e->n * e->L->ctrl_bytes,
                         hipMemcpyDeviceToHost, e->stream));
  EHIP(e, hipStreamSynchronize(e->stream));
  return 0;
}

void ens_status(const mhd_ensemble* e, int i, mhd_status* st) {
  const void* h = ctrl_at(e->ctrls_host, e, i);
  if (e->L->time_is_float)
    fill_status_of(e->cfg.hlld, e->nx[i], e->ny[i], (const Ctrl<float>*)h, st);
  else
    fill_status_of(e->cfg.hlld, e->nx[i], e->ny[i], (const Ctrl<double>*)h, st);
  st->rollbacks = e->rb[i].rollbacks;
  st->restarts = e->rb[i].restarts;
}

size_t ens_state_bytes(const mhd_ensemble* e, int i) {
  return (size_t)e->L->state_planes * e->plane[i] * e->L->state_elem;
}

hipError_t ens_enqueue_step(mhd_ensemble* e, hipEvent_t* ev = nullptr) {
  return e->L->ens_step(e->el, e->stream, ev);
}

int ens_ensure_graph(mhd_ensemble* e, int batch) {
  if (e->graph && e->graph_steps == batch) return 0;
  if (e->graph) {
    (void)hipGraphExecDestroy(e->graph);
    e->graph = nullptr;
  }
  hipGraph_t g = nullptr;
  EHIP(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
  hipError_t le = hipSuccess;
  for (int q = 0; q < batch && le == hipSuccess; q++) le = ens_enqueue_step(e);
  hipError_t ee = hipStreamEndCapture(e->stream, &g);
  if (le != hipSuccess) return efail(e, 9, "ensemble step launch during capture");
  if (ee != hipSuccess) return efail(e, 9, "hipStreamEndCapture");
  hipError_t ie = hipGraphInstantiate(&e->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ie != hipSuccess) {
    e->graph = nullptr;
    return efail(e, 9, "hipGraphInstantiate");
  }
  e->graph_steps = batch;
  return 0;
}

}  // namespace

extern "C" {

const char* mhd_ensemble_last_error(const mhd_ensemble* e) {
  return e ? e->err.c_str() : kNoCtx;
}

void mhd_ensemble_destroy(mhd_ensemble* e) {
  if (!e) return;
  if (e->graph) (void)hipGraphExecDestroy(e->graph);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  for (void* p : e->u)
    if (p) (void)hipFree(p);
  for (void* p : e->u1)
    if (p) (void)hipFree(p);
  for (Rollback& r : e->rb) rb_free(r);
  if (e->ctrls) (void)hipFree(e->ctrls);
  if (e->ctrls_host) (void)hipHostFree(e->ctrls_host);
  if (e->desc) (void)hipFree(e->desc);
  if (e->prefix) (void)hipFree(e->prefix);
  if (e->queue) (void)hipFree(e->queue);
  if (e->live) (void)hipFree(e->live);
  if (e->redo_items) (void)hipFree(e->redo_items);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

int mhd_ensemble_create(const mhd_cfg* cfg, int n, const int* nx, const int* ny,
                        mhd_ensemble** out) {
  if (!out) return 1;
  *out = nullptr;
  if (!cfg || n < 1 || !nx || !ny) return 1;
  if (!(cfg->cfl > 0.0 && cfg->cfl < 1.0) || !(cfg->alpha > 0.0)) return 1;
  if (!launchers_for(cfg->precision) || is_variant(cfg->precision)) return 1;
  if (cfg->precision == MHD_PREC_REF_MIXED && !cfg->muscl) return 1;
  for (int i = 0; i < n; i++)
    if (nx[i] < 1 || ny[i] < 1 || (long long)nx[i] * ny[i] > (long long)INT_MAX) return 1;
  mhd_ensemble* e = new mhd_ensemble();
  e->cfg = *cfg;
  e->L = launchers_for(cfg->precision);
  e->n = n;
  if (e->cfg.batch_steps <= 0) e->cfg.batch_steps = 32;
  if (e->cfg.hlld_tol <= 0.0) e->cfg.hlld_tol = 1.0e-4;
  if (!(e->cfg.lowbeta_ratio > 0.0)) e->cfg.lowbeta_ratio = 16.0;
  if (e->cfg.tile < 0 || e->cfg.tile >= mhd::TILE_COUNT) {
    long long biggest = 0;
    for (int i = 0; i < n; i++) biggest = std::max(biggest, (long long)nx[i] * ny[i]);
    e->cfg.tile = tuned_tile(cfg->precision, biggest);
  }
  if (!(cfg->muscl && cfg->hlld)) e->cfg.tile = mhd::kDefaultTile;  // as single contexts
  e->nx.assign(nx, nx + n);
  e->ny.assign(ny, ny + n);
  e->u.assign(n, nullptr);
  e->u1.assign(n, nullptr);
  e->have.assign(n, 0);
  e->plane.resize(n);
  e->rb.resize(n);
  int rc = 0;
  auto hipok = [&](hipError_t x, const char* what) {
    if (x != hipSuccess && rc == 0) rc = efail(e, 9, std::string(what) + ": " + hipGetErrorString(x));
    return rc == 0;
  };
  hipok(hipStreamCreate(&e->stream), "hipStreamCreate");
  for (int i = 0; i < n && rc == 0; i++) {
    e->plane[i] = ((size_t)nx[i] * ny[i] + 63) / 64 * 64;
    const size_t bytes = (size_t)e->L->state_planes * e->plane[i] * e->L->state_elem;
    hipok(hipMalloc(&e->u[i], bytes), "hipMalloc");
    if (rc == 0) hipok(hipMalloc(&e->u1[i], bytes), "hipMalloc");
    if (rc == 0 && rollback_enabled(e->cfg)) {
      e->rb[i].every = rollback_interval(e->cfg);
      hipok(rb_alloc(e->rb[i], bytes), "hipMalloc");
    }
  }
  const size_t cb = (size_t)n * e->L->ctrl_bytes;
  if (rc == 0) hipok(hipMalloc(&e->ctrls, cb), "hipMalloc");
  if (rc == 0) hipok(hipHostMalloc(&e->ctrls_host, cb, hipHostMallocDefault), "hipHostMalloc");
  if (rc == 0) hipok(hipMalloc(&e->desc, (size_t)n * e->L->desc_bytes), "hipMalloc");
  if (rc == 0) hipok(hipMalloc((void**)&e->prefix, (size_t)(n + 1) * sizeof(int)), "hipMalloc");
  if (rc == 0)
    hipok(hipMalloc((void**)&e->queue, mhd::kEnsQueueWords * sizeof(unsigned)), "hipMalloc");
  if (rc == 0) hipok(hipMalloc(&e->live, mhd::ens_live_bytes(n)), "hipMalloc");
  if (rc == 0) hipok(hipMemset(e->live, 0, mhd::ens_live_bytes(n)), "hipMemset");
  if (rc == 0) {
    std::vector<char> hdesc((size_t)n * e->L->desc_bytes);
    std::vector<int> prefix(n + 1, 0);
    long long total = 0;
    for (int i = 0; i < n; i++) {
      prefix[i] = (int)total;
      total += e->L->ens_fill(hdesc.data(), i, e->u[i], e->u1[i], e->plane[i], nx[i], ny[i],
                              ctrl_at(e->ctrls, e, i), e->cfg.tile);
    }
    if (total > INT_MAX) rc = efail(e, 1, "ensemble too large");
    prefix[n] = (int)total;
    // Every instance starts as finished, so stray launches do nothing until
    // a state is set.
    for (int i = 0; i < n && rc == 0; i++) {
      mhd_cfg ci = e->cfg;
      ci.nx = nx[i];
      ci.ny = ny[i];
      if (e->L->time_is_float) init_ctrl((Ctrl<float>*)ctrl_at(e->ctrls_host, e, i), ci);
      else init_ctrl((Ctrl<double>*)ctrl_at(e->ctrls_host, e, i), ci);
    }
    const unsigned zeros[mhd::kEnsQueueWords] = {0};
    if (rc == 0) hipok(hipMemcpy(e->desc, hdesc.data(), hdesc.size(), hipMemcpyHostToDevice), "desc");
    if (rc == 0) hipok(hipMemcpy(e->prefix, prefix.data(), prefix.size() * sizeof(int),
                                 hipMemcpyHostToDevice), "prefix");
    if (rc == 0) hipok(hipMemcpy(e->queue, zeros, sizeof(zeros), hipMemcpyHostToDevice), "queue");
    if (rc == 0) hipok(hipMemcpy(e->ctrls, e->ctrls_host, cb, hipMemcpyHostToDevice), "ctrl");
    int dev = 0, cus = 256;
    if (rc == 0) hipok(hipGetDevice(&dev), "hipGetDevice");
    if (rc == 0) hipok(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev),
                       "hipDeviceGetAttribute");
    e->el.desc = e->desc;
    e->el.prefix = e->prefix;
    e->el.n = n;
    e->el.total = (int)total;
    e->el.queue = e->queue;
    e->el.ctrls = e->ctrls;
    e->el.tol = e->cfg.hlld_tol;
    e->el.alpha = e->cfg.alpha;
    e->el.lowbeta_ratio = e->cfg.lowbeta_ratio;
    e->el.muscl = e->cfg.muscl ? 1 : 0;
    e->el.hlld = e->cfg.hlld ? 1 : 0;
    e->el.tile = e->cfg.tile;
    e->el.grid_blocks = 0;
    if (const char* g = std::getenv("MHD_PERSISTENT_BLOCKS")) e->el.grid_blocks = std::atoi(g);
    e->el.live = e->live;
    e->el.kernel = mhd::ENS_FLAT;
    if (const char* k = std::getenv("MHD_ENSEMBLE_KERNEL"))
      if (std::strcmp(k, "queue") == 0) e->el.kernel = mhd::ENS_QUEUE;
    e->el.cus = cus;
    if (rc == 0 && cfg->precision == MHD_PREC_MIXED && cfg->fallback != 1) {
      hipok(hipMalloc((void**)&e->redo_items, 2 * (size_t)total * sizeof(int)), "hipMalloc");
      if (rc == 0) {
        e->el.redo_items[0] = e->redo_items;
        e->el.redo_items[1] = e->redo_items + total;
        e->el.redo = mhd::launchers_exact(mhd::MODE_FP64)->ens_redo;
        e->el.redo_all = cfg->fallback == 2 ? 1 : 0;
      }
    }
  }
  if (rc != 0) {
    std::fprintf(stderr, "mhd_ensemble_create: %s\n", e->err.c_str());
    mhd_ensemble_destroy(e);
    return rc;
  }
  *out = e;
  return 0;
}

int mhd_ensemble_set_state(mhd_ensemble* e, int i, const void* soa9, size_t ld) {
  if (!e || i < 0 || i >= e->n) return 1;
  const size_t cells = (size_t)e->nx[i] * e->ny[i];
  if (!soa9 || ld < cells) return efail(e, 1, "mhd_ensemble_set_state: bad buffer");
  EHIP(e, upload_state(e->L, e->u[i], e->plane[i], cells, soa9, ld, e->stream));
  mhd_cfg ci = e->cfg;
  ci.nx = e->nx[i];
  ci.ny = e->ny[i];
  void* h = ctrl_at(e->ctrls_host, e, i);
  EHIP(e, hipStreamSynchronize(e->stream));
  if (e->L->time_is_float) init_ctrl((Ctrl<float>*)h, ci);
  else init_ctrl((Ctrl<double>*)h, ci);
  void* d = ctrl_at(e->ctrls, e, i);
  EHIP(e, hipMemcpy(d, h, e->L->ctrl_bytes, hipMemcpyHostToDevice));
  const mhd::GridDesc g = inst_grid(e, i);
  EHIP(e, e->L->speed(g, e->u[i], d, e->stream));
  EHIP(e, e->L->commit(d, e->stream));
  if (e->rb[i].every > 0) {
    EHIP(e, hipMemcpyAsync(h, d, e->L->ctrl_bytes, hipMemcpyDeviceToHost, e->stream));
    EHIP(e, hipStreamSynchronize(e->stream));
    EHIP(e, rb_start(e->rb[i], e->u[i], ens_state_bytes(e, i), h, e->L->ctrl_bytes, e->stream));
  }
  EHIP(e, hipStreamSynchronize(e->stream));
  e->have[i] = 1;
  return 0;
}

int mhd_ensemble_init_problem(mhd_ensemble* e, int i, const char* problem, const double* params,
                              int nparams) {
  if (!e || i < 0 || i >= e->n) return 1;
  const size_t cells = (size_t)e->nx[i] * e->ny[i];
  std::vector<char> host(9 * cells * io_elem(e->cfg.precision));
  int rc = mhd_problem_state_ex(problem, params, nparams, e->nx[i], e->ny[i], e->cfg.precision,
                                host.data(), cells);
  if (rc) return efail(e, rc, "mhd_ensemble_init_problem: bad problem or parameters");
  return mhd_ensemble_set_state(e, i, host.data(), cells);
}

int mhd_ensemble_enqueue_steps(mhd_ensemble* e, double t_end, int nsteps) {
  if (!e) return 1;
  EHIP(e, e->L->ens_set_target(e->ctrls, e->n, t_end, e->queue, e->stream));
  const int batch = e->cfg.batch_steps;
  int left = nsteps;
  if (use_graphs() && left >= batch) {
    if (int rc = ens_ensure_graph(e, batch)) return rc;
    for (; left >= batch; left -= batch) EHIP(e, hipGraphLaunch(e->graph, e->stream));
  }
  for (; left > 0; left--) EHIP(e, ens_enqueue_step(e));
  return 0;
}

int mhd_ensemble_advance(mhd_ensemble* e, double t_end, long long max_steps, mhd_status* st) {
  if (!e) return 1;
  EHIP(e, e->L->ens_set_target(e->ctrls, e->n, t_end, e->queue, e->stream));
  if (int rc = ens_read_ctrls(e)) return rc;
  long long left = max_steps < 0 ? LLONG_MAX : max_steps;
  const int batch = e->cfg.batch_steps;
  const bool graphs = use_graphs();
  auto running = [&]() {
    for (int i = 0; i < e->n; i++) {
      mhd_status s;
      ens_status(e, i, &s);
      if (e->have[i] && s.code == 0 && !s.done) return true;
    }
    return false;
  };
  while (left > 0 && running()) {
    const int nb = (int)(left < batch ? left : batch);
    if (graphs && nb == batch) {
      if (int rc = ens_ensure_graph(e, batch)) return rc;
      EHIP(e, hipGraphLaunch(e->graph, e->stream));
    } else {
      for (int q = 0; q < nb; q++) EHIP(e, ens_enqueue_step(e));
    }
    left -= nb;
    if (int rc = ens_read_ctrls(e)) return rc;
    // Rollback of halted mixed instances; snapshots of the running ones.
    bool restored = false;
    for (int i = 0; i < e->n; i++) {
      Rollback& r = e->rb[i];
      if (r.every <= 0 || !e->have[i]) continue;
      mhd_status s;
      ens_status(e, i, &s);
      void* h = ctrl_at(e->ctrls_host, e, i);
      int slot, until;
      if ((s.code == 3 || s.code == 4) && rb_plan(r, h, slot, until)) {
        EHIP(e, rb_restore(r, slot, until, inst_grid(e, i), e->u[i], ens_state_bytes(e, i),
                           ctrl_at(e->ctrls, e, i), h, e->L->ctrl_bytes, t_end, e->stream));
        restored = true;
      } else if (s.code == 0 && !s.done && rb_snapshot_due(r, h)) {
        EHIP(e, rb_snapshot(r, e->u[i], ens_state_bytes(e, i), h, e->L->ctrl_bytes, e->stream));
      }
    }
    if (restored)
      if (int rc = ens_read_ctrls(e)) return rc;
  }
  int worst = 0;
  for (int i = 0; i < e->n; i++) {
    mhd_status s;
    ens_status(e, i, &s);
    if (st) st[i] = s;
    if (s.code > worst) worst = s.code;
  }
  return worst;
}

int mhd_ensemble_get_status(mhd_ensemble* e, int i, mhd_status* st) {
  if (!e || i < 0 || i >= e->n || !st) return 1;
  if (int rc = ens_read_ctrls(e)) return rc;
  ens_status(e, i, st);
  return st->code;
}

int mhd_ensemble_get_state(mhd_ensemble* e, int i, int which, void* soa9, size_t ld) {
  if (!e || i < 0 || i >= e->n || (which != 0 && which != 1)) return 1;
  const size_t cells = (size_t)e->nx[i] * e->ny[i];
  if (!soa9 || ld < cells) return efail(e, 1, "mhd_ensemble_get_state: bad buffer");
  EHIP(e, download_state(e->L, which == 0 ? e->u[i] : e->u1[i], e->plane[i], cells, soa9, ld,
                         e->stream));
  return 0;
}

int mhd_ensemble_time_kernels(mhd_ensemble* e, double t_end, int nsteps, double* avg_ms) {
  if (!e || !avg_ms || nsteps <= 0) return 1;
  EHIP(e, e->L->ens_set_target(e->ctrls, e->n, t_end, e->queue, e->stream));
  std::vector<hipEvent_t> ev((size_t)nsteps * 4, nullptr);
  for (auto& x : ev) EHIP(e, hipEventCreate(&x));
  int rc = 0;
  for (int q = 0; q < nsteps && rc == 0; q++) {
    hipError_t x = ens_enqueue_step(e, &ev[(size_t)q * 4]);
    if (x != hipSuccess) rc = efail(e, 9, "timed ensemble step");
  }
  if (rc == 0 && hipStreamSynchronize(e->stream) != hipSuccess) rc = efail(e, 9, "sync");
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
// Learned precision selector and its MLP.
// ---------------------------------------------------------------------------
namespace {

int sel_alloc(mhd_ctx* c) {
  auto& S = c->sel;
  if (S.params) return 0;
  int tx, ty;
  mhd::tile_dims(c->cfg.tile, tx, ty);
  S.tx = tx;
  S.ty = ty;
  S.ntiles = ((c->cfg.nx + tx - 1) / tx) * ((c->cfg.ny + ty - 1) / ty);
  const size_t nt = (size_t)S.ntiles;
  HIP_OR_FAIL(c, hipMalloc((void**)&S.params, mhd::sel::kParams * sizeof(float)));
  HIP_OR_FAIL(c, hipMalloc((void**)&S.mean, mhd::sel::NFEAT * sizeof(float)));
  HIP_OR_FAIL(c, hipMalloc((void**)&S.istd, mhd::sel::NFEAT * sizeof(float)));
  HIP_OR_FAIL(c, hipMalloc((void**)&S.feats, nt * mhd::sel::NIN * sizeof(float)));
  HIP_OR_FAIL(c, hipMalloc((void**)&S.err, nt * sizeof(float)));
  HIP_OR_FAIL(c, hipMalloc((void**)&S.flags, nt));
  HIP_OR_FAIL(c, hipMalloc((void**)&S.sticky, nt));
  HIP_OR_FAIL(c, hipMalloc((void**)&S.lists, 5 * nt * sizeof(int)));
  HIP_OR_FAIL(c, hipMalloc((void**)&S.counts, 3 * sizeof(int)));
  HIP_OR_FAIL(c, hipMalloc((void**)&S.sumsq, 9 * sizeof(double)));
  HIP_OR_FAIL(c, hipMalloc((void**)&S.counters, 4 * sizeof(unsigned long long)));
  HIP_OR_FAIL(c, hipMalloc(&S.u1x, 9 * c->plane * sizeof(double)));
  HIP_OR_FAIL(c, hipMalloc(&S.ctrl_scratch, c->L->ctrl_bytes));
  HIP_OR_FAIL(c, hipMemset(S.sticky, 0, nt));
  HIP_OR_FAIL(c, hipMemset(S.counters, 0, 4 * sizeof(unsigned long long)));
  const char* conc = std::getenv("MHD_SELECTOR_CONCURRENT");
  if (conc && std::atoi(conc) == 1) {
    HIP_OR_FAIL(c, hipStreamCreateWithFlags(&S.side, hipStreamNonBlocking));
    for (int k = 0; k < 2; k++) {
      HIP_OR_FAIL(c, hipEventCreateWithFlags(&S.fork[k], hipEventDisableTiming));
      HIP_OR_FAIL(c, hipEventCreateWithFlags(&S.join[k], hipEventDisableTiming));
    }
  }
  return 0;
}

// Copy of the control block, set up for a step from the current state
// (c_h and dt from the committed signal speed, no time limit), for the
// selector's stand-alone feature and label evaluations.
hipError_t prepared_snapshot(mhd_ctx* c) {
  hipError_t e = mhd::sel_copy_words(c->ctrl, c->sel.ctrl_scratch, c->L->ctrl_bytes, c->stream);
  if (e == hipSuccess) e = c->L->set_target(c->sel.ctrl_scratch, 1.0e300, c->stream);
  return e;
}

}  // namespace

extern "C" {

int mhd_selector_tiles(mhd_ctx* c) {
  if (!c || c->cfg.precision != MHD_PREC_MIXED) return -1;
  int tx, ty;
  mhd::tile_dims(c->cfg.tile, tx, ty);
  return ((c->cfg.nx + tx - 1) / tx) * ((c->cfg.ny + ty - 1) / ty);
}

int mhd_tile_features(mhd_ctx* c, float* feats) {
  if (!c || !feats) return 1;
  if (c->cfg.precision != MHD_PREC_MIXED) return fail(c, 1, "selector needs the MIXED form");
  if (!c->have_state) return fail(c, 1, "mhd_tile_features: no state");
  if (int rc = read_ctrl(c)) return rc;
  if (status_of(c, nullptr) != 0) return fail(c, 1, "mhd_tile_features: the run has halted");
  if (int rc = settle_state(c)) return rc;
  if (int rc = sel_alloc(c)) return rc;
  auto& S = c->sel;
  HIP_OR_FAIL(c, prepared_snapshot(c));
  HIP_OR_FAIL(c, mhd::sel_tile_features((const double*)c->u, c->plane, c->cfg.nx, c->cfg.ny, S.tx,
                                        S.ty, (const char*)S.ctrl_scratch + offsetof(Ctrl<double>, ch),
                                        S.feats, c->stream));
  HIP_OR_FAIL(c, hipMemcpyAsync(feats, S.feats, (size_t)S.ntiles * mhd::sel::NIN * sizeof(float),
                                hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
  return 0;
}

int mhd_pair_tile_errors(mhd_ctx* c, float* err) {
  if (!c || !err) return 1;
  if (c->cfg.precision != MHD_PREC_MIXED) return fail(c, 1, "selector needs the MIXED form");
  if (!c->have_state) return fail(c, 1, "mhd_pair_tile_errors: no state");
  if (int rc = read_ctrl(c)) return rc;
  if (status_of(c, nullptr) != 0) return fail(c, 1, "mhd_pair_tile_errors: the run has halted");
  if (int rc = settle_state(c)) return rc;
  if (int rc = sel_alloc(c)) return rc;
  auto& S = c->sel;
  const mhd::Launchers* F64 = mhd::launchers_exact(mhd::MODE_FP64);
  mhd::GridDesc g = c->grid;
  g.tiles = nullptr;
  g.redo_list[0] = g.redo_list[1] = nullptr;  // both updates of every tile
  // Both stage-1 updates of the current state, against a prepared snapshot
  // of the control block (the run's own counters and verdicts are
  // untouched).  The FP32-flux update goes to the context's stage buffer,
  // which the next step overwrites.
  HIP_OR_FAIL(c, prepared_snapshot(c));
  HIP_OR_FAIL(c, c->L->stage(g, 1, c->u, c->u, c->u1, S.ctrl_scratch, c->stream));
  HIP_OR_FAIL(c, prepared_snapshot(c));
  HIP_OR_FAIL(c, F64->stage(g, 1, c->u, c->u, S.u1x, S.ctrl_scratch, c->stream));
  double* sumsq = S.sumsq;
  HIP_OR_FAIL(c, mhd::sel_sumsq((const double*)S.u1x, c->plane, (long long)c->cfg.nx * c->cfg.ny,
                                sumsq, c->stream));
  HIP_OR_FAIL(c, mhd::sel_tile_errors((const double*)c->u1, (const double*)S.u1x, c->plane,
                                      c->cfg.nx, c->cfg.ny, S.tx, S.ty, sumsq, nullptr, nullptr,
                                      S.err, c->stream));
  HIP_OR_FAIL(c, hipMemcpyAsync(err, S.err, (size_t)S.ntiles * sizeof(float),
                                hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
  return 0;
}

int mhd_selector_configure(mhd_ctx* c, const float* params, const float* mean, const float* stdev,
                           float threshold, float audit_frac, float audit_budget) {
  if (!c) return 1;
  if (!params) {
    c->sel.on = false;
    if (c->graph) {
      (void)hipGraphExecDestroy(c->graph);
      c->graph = nullptr;
    }
    return 0;
  }
  if (c->cfg.precision != MHD_PREC_MIXED) return fail(c, 1, "selector needs the MIXED form");
  if (!c->have_state) return fail(c, 1, "mhd_selector_configure: set a state first");
  // The refresh every kSelectorRefresh steps is decided when a step is
  // enqueued, so a captured batch replays the refreshes it was captured
  // with: a batch that is a multiple of the refresh period keeps them every
  // kSelectorRefresh steps across replays and the steps enqueued one by one.
  if (c->cfg.batch_steps % kSelectorRefresh != 0)
    c->cfg.batch_steps += kSelectorRefresh - c->cfg.batch_steps % kSelectorRefresh;
  // The selector routes tiles of the single-stage step: two-stage steps
  // end for this context.
  if (int rc = settle_state(c)) return rc;
  c->fused = false;
  c->cfg.fuse = 1;
  if (int rc = sel_alloc(c)) return rc;
  auto& S = c->sel;
  float istd[mhd::sel::NFEAT], mu[mhd::sel::NFEAT];
  for (int k = 0; k < mhd::sel::NFEAT; k++) {
    mu[k] = mean ? mean[k] : 0.f;
    istd[k] = (stdev && stdev[k] > 0.f) ? 1.f / stdev[k] : 1.f;
  }
  HIP_OR_FAIL(c, hipMemcpy(S.params, params, mhd::sel::kParams * sizeof(float), hipMemcpyHostToDevice));
  HIP_OR_FAIL(c, hipMemcpy(S.mean, mu, sizeof(mu), hipMemcpyHostToDevice));
  HIP_OR_FAIL(c, hipMemcpy(S.istd, istd, sizeof(istd), hipMemcpyHostToDevice));
  HIP_OR_FAIL(c, hipMemset(S.sticky, 0, (size_t)S.ntiles));
  HIP_OR_FAIL(c, hipMemset(S.counters, 0, 4 * sizeof(unsigned long long)));
  // Error scale of the audits: the current state's per-plane RMS.
  HIP_OR_FAIL(c, mhd::sel_sumsq((const double*)c->u, c->plane, (long long)c->cfg.nx * c->cfg.ny,
                                S.sumsq, c->stream));
  HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
  S.threshold = threshold;
  S.audit_frac = audit_frac;
  S.audit_budget = audit_budget;
  S.since = 0;
  S.on = true;
  if (c->graph) {  // captured steps of the other path
    (void)hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
  }
  return 0;
}

int mhd_selector_stats(mhd_ctx* c, unsigned long long* out4) {
  if (!c || !out4) return 1;
  if (!c->sel.counters) {
    for (int k = 0; k < 4; k++) out4[k] = 0;
    return 0;
  }
  HIP_OR_FAIL(c, hipMemcpyAsync(out4, c->sel.counters, 4 * sizeof(unsigned long long),
                                hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
  return 0;
}

int mhd_mlp_param_count(void) { return mhd::sel::kParams; }

int mhd_mlp_forward(const float* params, const float* x, int n, float* prob) {
  if (!params || n < 0 || (n > 0 && (!x || !prob))) return 1;
  if (n == 0) return 0;
  DevBuf p, xb, pb;
  if (hipMalloc(&p.p, mhd::sel::kParams * sizeof(float)) != hipSuccess ||
      hipMalloc(&xb.p, (size_t)n * mhd::sel::NIN * sizeof(float)) != hipSuccess ||
      hipMalloc(&pb.p, (size_t)n * sizeof(float)) != hipSuccess)
    return 9;
  if (hipMemcpy(p.p, params, mhd::sel::kParams * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(xb.p, x, (size_t)n * mhd::sel::NIN * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    return 9;
  if (mhd::sel_mlp_infer((const float*)xb.p, n, (const float*)p.p, nullptr, nullptr, 2.f, nullptr,
                         (float*)pb.p, nullptr, nullptr) != hipSuccess)
    return 9;
  if (hipMemcpy(prob, pb.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) return 9;
  return 0;
}

int mhd_mlp_gradients(const float* params, const float* x, const float* y, const float* w, int n,
                      float* grad, float* loss) {
  if (!params || !x || !y || !w || !grad || n <= 0) return 1;
  DevBuf p, xb, yb, wb, ib, gb, lb, pb;
  const size_t np = mhd::sel::kParams;
  const int nparts = (n + 63) / 64;
  if (hipMalloc(&p.p, np * 4) != hipSuccess || hipMalloc(&xb.p, (size_t)n * mhd::sel::NIN * 4) != hipSuccess ||
      hipMalloc(&yb.p, (size_t)n * 4) != hipSuccess || hipMalloc(&wb.p, (size_t)n * 4) != hipSuccess ||
      hipMalloc(&ib.p, (size_t)n * 4) != hipSuccess || hipMalloc(&gb.p, np * 4) != hipSuccess ||
      hipMalloc(&lb.p, 4) != hipSuccess ||
      hipMalloc(&pb.p, (size_t)nparts * (np + 1) * 4) != hipSuccess)
    return 9;
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  double ws = 0;
  for (int s = 0; s < n; s++) ws += w[s];
  if (hipMemcpy(p.p, params, np * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(xb.p, x, (size_t)n * mhd::sel::NIN * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(yb.p, y, (size_t)n * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(wb.p, w, (size_t)n * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(ib.p, idx.data(), (size_t)n * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(gb.p, 0, np * 4) != hipSuccess || hipMemset(lb.p, 0, 4) != hipSuccess)
    return 9;
  if (mhd::sel_mlp_grad((const float*)xb.p, (const float*)yb.p, (const float*)wb.p, (const int*)ib.p,
                        0, n, (float)ws, (const float*)p.p, (float*)pb.p, nullptr) != hipSuccess ||
      mhd::sel_sum_partials((const float*)pb.p, nparts, (float*)gb.p, (float*)lb.p, nullptr) !=
          hipSuccess)
    return 9;
  if (hipMemcpy(grad, gb.p, np * 4, hipMemcpyDeviceToHost) != hipSuccess) return 9;
  if (loss && hipMemcpy(loss, lb.p, 4, hipMemcpyDeviceToHost) != hipSuccess) return 9;
  return 0;
}

int mhd_mlp_train(float* params, const float* x, const float* y, const float* w, int n,
                  int epochs, int batch, float lr, unsigned seed, float* epoch_loss) {
  if (!params || !x || !y || !w || n <= 0 || epochs <= 0 || batch <= 0) return 1;
  const size_t np = mhd::sel::kParams;
  DevBuf p, xb, yb, wb, ib, pb, mb, vb, lb;
  const int maxparts = (std::min(batch, n) + 63) / 64;
  if (hipMalloc(&p.p, np * 4) != hipSuccess || hipMalloc(&xb.p, (size_t)n * mhd::sel::NIN * 4) != hipSuccess ||
      hipMalloc(&yb.p, (size_t)n * 4) != hipSuccess || hipMalloc(&wb.p, (size_t)n * 4) != hipSuccess ||
      hipMalloc(&ib.p, (size_t)n * 4) != hipSuccess ||
      hipMalloc(&pb.p, (size_t)maxparts * (np + 1) * 4) != hipSuccess ||
      hipMalloc(&mb.p, np * 4) != hipSuccess || hipMalloc(&vb.p, np * 4) != hipSuccess ||
      hipMalloc(&lb.p, (size_t)epochs * 4) != hipSuccess)
    return 9;
  if (hipMemcpy(p.p, params, np * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(xb.p, x, (size_t)n * mhd::sel::NIN * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(yb.p, y, (size_t)n * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(wb.p, w, (size_t)n * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(mb.p, 0, np * 4) != hipSuccess || hipMemset(vb.p, 0, np * 4) != hipSuccess ||
      hipMemset(lb.p, 0, (size_t)epochs * 4) != hipSuccess)
    return 9;
  std::mt19937 rng(seed);
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  int step = 0;
  for (int ep = 0; ep < epochs; ep++) {
    std::shuffle(idx.begin(), idx.end(), rng);
    if (hipMemcpy(ib.p, idx.data(), (size_t)n * 4, hipMemcpyHostToDevice) != hipSuccess) return 9;
    for (int off = 0; off < n; off += batch) {
      const int cnt = std::min(batch, n - off);
      double ws = 0;
      for (int q = 0; q < cnt; q++) ws += w[idx[off + q]];
      // The epoch loss is the mean of the minibatch losses.  The gradient
      // rows of the minibatch's blocks are summed in block order, so
      // training is the same from run to run.
      if (mhd::sel_mlp_grad((const float*)xb.p, (const float*)yb.p, (const float*)wb.p,
                            (const int*)ib.p, off, cnt, (float)ws, (const float*)p.p, (float*)pb.p,
                            nullptr) != hipSuccess)
        return 9;
      if (mhd::sel_adam((float*)p.p, (const float*)pb.p, (cnt + 63) / 64, (float*)mb.p,
                        (float*)vb.p, ++step, lr, (float*)lb.p + ep, nullptr) != hipSuccess)
        return 9;
    }
  }
  if (hipDeviceSynchronize() != hipSuccess) return 9;
  if (hipMemcpy(params, p.p, np * 4, hipMemcpyDeviceToHost) != hipSuccess) return 9;
  if (epoch_loss) {
    if (hipMemcpy(epoch_loss, lb.p, (size_t)epochs * 4, hipMemcpyDeviceToHost) != hipSuccess) return 9;
    const int nb = (n + batch - 1) / batch;
    for (int ep = 0; ep < epochs; ep++) epoch_loss[ep] /= (float)nb;
  }
  return 0;
}

}  // extern "C"
