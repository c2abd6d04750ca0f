/* mhd_glm.h -- C ABI of the MI355X 2D ideal-MHD + mixed-GLM solver.
 *
 * The reference solver has no library boundary: its drivers are
 * executables (mhd2d_glm_cpu.cpp:417, mhd2d_glm_gpu.cu:353) whose loop calls
 * header-only physics (include/mhd/mhd_glm_common.hpp).  This ABI is the seam
 * between a host driver (CLI, initial conditions, .dat output) and the HIP
 * step, so the process contract of those executables (argv, OUTDIR, stdout,
 * exit codes) is kept by a thin host program on top of it, and the
 * reference's by-array operator signatures are exported as batched calls.
 *
 * Conventions: plain C types only; every int return is 0 on success or a
 * code of the reference executables (1 usage, 2 non-positive signal rate,
 * 3 non-physical stage-1 state, 4 non-physical state, 9 device error);
 * host buffers belong to the caller, device memory to the context; a
 * context owns one HIP stream and is used by one host thread at a time.
 *
 * State arrays exchanged with the host are structure-of-arrays: 9 planes
 * (rho, mx, my, mz, E, Bx, By, Bz, psi), plane k starting at element
 * k*ld, cell (i,j) of plane k at k*ld + j*nx + i.  Element type is float
 * for MHD_PREC_FP32 and double for every other form.
 */
#ifndef MHD_GLM_H
#define MHD_GLM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Arithmetic form of the step.
 *   FP64, FP32   the reference's two builds (base.hpp:27-31), FP contraction
 *                off: bitwise the reference CPU build of that precision.
 *   MIXED        FP32 reconstruction, HLLD and GLM; FP64 pressure recovery
 *                where FP32 would cancel (low beta / high Mach); FP64 state,
 *                update accumulation, dt and time.
 *   MIXED_KAHAN  as MIXED but the state is FP32 + FP32 compensation.
 *   FP64_FMA     FP64 with FMA contraction (relative L2 <= 1e-12 of FP64).
 *   REF_MIXED    the reference's stage-level mixed build
 *                (mhd2d_glm_gpu_mixed.cu:1-10): state, Riemann fluxes and
 *                SSPRK2 updates in FP64; recovery, CFL reductions, face
 *                reconstruction, rhs differences and psi damping in FP32.
 *                MUSCL only (first order keeps FP64 face states there).
 *   MIXED_RECON64, MIXED_FLUX64, MIXED_FLUX64_NARROW
 *                allocation variants of MIXED for locating its error: FP64
 *                recovery and reconstruction; FP64 GLM-HLLD flux differenced
 *                in FP64; FP64 flux narrowed to FP32 before the differences
 *                (as REF_MIXED narrows it).  Single contexts, MUSCL + HLLD,
 *                16x15 tiles, one stage per launch, no fallback or selector.
 * MIXED, MIXED_KAHAN, FP64_FMA and the variants have no reference
 * counterpart; MIXED has the opposite allocation to REF_MIXED. */
enum mhd_precision {
  MHD_PREC_FP64 = 0, MHD_PREC_FP32 = 1, MHD_PREC_MIXED = 2,
  MHD_PREC_MIXED_KAHAN = 3, MHD_PREC_FP64_FMA = 4, MHD_PREC_REF_MIXED = 5,
  MHD_PREC_MIXED_RECON64 = 6, MHD_PREC_MIXED_FLUX64 = 7, MHD_PREC_MIXED_FLUX64_NARROW = 8
};

/* Run configuration.  Mirrors the reference CLI
 * `nx ny t_end cfl order solver alpha [problem]` (mhd2d_glm_gpu.cu:354-374)
 * plus the compile-time switches of the reference build: precision
 * (base.hpp:27-31, -DUSE_FLOAT) and HLLD tolerance
 * (mhd_common.hpp:77-79, -DMHD_HLLD_TOLERANCE_VALUE). */
typedef struct mhd_cfg {
  int nx, ny;        /* grid cells, >= 2 each */
  double cfl;        /* in (0,1) */
  double alpha;      /* GLM damping parameter, > 0 */
  int muscl;         /* 1: minmod MUSCL ("muscl"), 0: first order ("first") */
  int hlld;          /* 1: HLLD with HLL fallback ("hlld"), 0: HLL ("hll") */
  int precision;     /* enum mhd_precision */
  double hlld_tol;   /* HLLD degeneracy tolerance (protocol value 1e-4) */
  int tile;          /* stage-kernel tile shape id (0 16x16, 1 32x8, 2 16x8,
                        3 16x15, 4 32x7, 5 16x7); -1 = the autotuned choice
                        for this form and size (autotune.txt), else 16x15 */
  int batch_steps;   /* steps launched per host check; <= 0 = default */
  double lowbeta_ratio; /* MIXED*: FP64 recovery where E > ratio * e_int;
                           <= 0 = default (16) */
  int fallback;      /* MIXED: a stage tile whose FP32-flux update fails the
                        positivity check is re-evaluated by the FP64 kernel
                        (only an FP64 failure halts the run).  0 = on
                        (default), 1 = off (the failure halts, as in the
                        other forms), 2 = every tile (check mode: the step
                        is then the FP64 step) */
  int fuse;          /* SSPRK2 stages per kernel launch (fusion depth).
                        1: one launch per stage (U1 through HBM); 2: one
                        launch per step, stage 1 recomputed on a 2-cell halo
                        ring of each tile (MUSCL + HLLD, tile 16x15, not with
                        fallback 2; the learned selector and ensembles run
                        one stage per launch).  Results as with 1: bitwise
                        in the exact forms, the same scheme in the mixed
                        ones.  0 = the autotuned choice when tile is -1,
                        else 1 */
  int rollback_steps; /* MIXED with the fallback on (fallback 0): a run that
                        halts anyway (an FP64-confirmed unusable cell after
                        the trajectory drifted from the FP64 one) is rolled
                        back to a device snapshot at least this many steps
                        before the failure (snapshots every this many steps;
                        the state given at set_state is always kept) and
                        re-run in FP64 until this many steps past the
                        failure, then resumes mixed.  A failure inside such
                        a window restarts the run from the given state in
                        FP64 throughout, which halts only where the FP64
                        form halts.  Costs three more state buffers.
                        0 = default (256), < 0 = off (halts as the other
                        forms do, the reference's contract) */
} mhd_cfg;

/* Progress and diagnostics of a context. */
typedef struct mhd_status {
  int code;                 /* 0 running/finished, 2/3/4 halted, 9 device */
  int done;                 /* 1 once t >= t_end */
  int steps;                /* completed SSPRK2 steps since set_state */
  double t;                 /* simulation time (value of the run's type) */
  double dt;                /* dt of the next step, or of the failed step */
  double ch_last;           /* c_h of the last step taken (or failed) */
  unsigned long long x_evaluations, y_evaluations; /* HLLD flux evaluations */
  unsigned long long degenerate, nonphysical;      /* HLLD->HLL fallbacks */
  int bad_i, bad_j;         /* first unusable cell on halt 3/4, else -1 */
  int fail_step;            /* step index of the failure */
  unsigned long long fallback_tiles; /* MIXED: stage tiles re-evaluated in
                                        FP64 by the fallback, whole run */
  unsigned long long fused_redo_steps; /* MIXED, fuse 2: steps in which a tile
                                          met an unusable cell, recomputed by
                                          the one-stage kernels (with their
                                          FP64 fallback) */
  unsigned long long exact_steps; /* MIXED rollback: steps of the run taken
                                     with every tile in FP64 */
  int rollbacks;                  /* MIXED rollback: windows re-run in FP64 */
  int restarts;                   /* MIXED rollback: restarts from the given
                                     state in FP64 throughout (0 or 1) */
} mhd_status;

typedef struct mhd_ctx mhd_ctx;

/* Creates a solver on the current HIP device.  Replaces the setup of
 * mhd2d_glm_gpu.cu:390-474 (grid, allocations).  Returns 0, 1 (invalid
 * configuration) or 9 (device error); *out is NULL on failure. */
int mhd_create(const mhd_cfg* cfg, mhd_ctx** out);
void mhd_destroy(mhd_ctx* ctx);
/* Last error message of the context (static text if ctx is NULL). */
const char* mhd_last_error(const mhd_ctx* ctx);

/* Initial condition of a named problem ("ot", "kh", "khb") evaluated on the
 * host exactly as the reference does (mhd2d_glm_cpu.cpp:486-505,
 * mhd2d_glm_gpu.cu:415-457) into a caller SoA buffer (no device needed). */
int mhd_problem_state(const char* problem, int nx, int ny, int precision,
                      void* soa9, size_t ld);

/* Same for problems with parameters: "otv" takes 5 (B amplitude, and the
 * phases added to the sine arguments of vx, vy, Bx, By); the others take 0.
 * No reference counterpart (the ensemble workload's initial conditions). */
int mhd_problem_state_ex(const char* problem, const double* params, int nparams,
                         int nx, int ny, int precision, void* soa9, size_t ld);

/* Uploads a state and restarts time at t=0 with cleared counters; the
 * first step's c_h and dt are computed on the device. */
int mhd_set_state(mhd_ctx* ctx, const void* soa9, size_t ld);
/* Same as mhd_set_state with mhd_problem_state's initial condition. */
int mhd_init_problem(mhd_ctx* ctx, const char* problem);

/* Advances until t >= t_end, a failure, or max_steps more steps (max_steps
 * < 0: no limit).  This is the time loop of mhd2d_glm_cpu.cpp:513-570.
 * Returns the status code (0 or 2/3/4/9) and fills *st if non-NULL. */
int mhd_advance(mhd_ctx* ctx, double t_end, long long max_steps,
                mhd_status* st);
int mhd_get_status(mhd_ctx* ctx, mhd_status* st);

/* Downloads the current state U (which = 0) or the stage-1 state U1
 * (which = 1, the state the reference prints on an exit-3 failure; not kept
 * by two-stage steps, cfg.fuse 2, which return 1 for it). */
int mhd_get_state(mhd_ctx* ctx, int which, void* soa9, size_t ld);

/* Device pointer of state plane 0 (planes are ctx-internal `ld` apart,
 * reported in *ld) for callers that keep data on the device.  Waits for the
 * enqueued steps.  With two-stage steps (cfg.fuse 2) the state moves between
 * two buffers as steps run: the pointer holds the current state until the
 * next step is enqueued (mhd_advance and this call move it back here). */
int mhd_device_state(mhd_ctx* ctx, void** dev_ptr, size_t* ld);
/* HIP stream of the context (as void*: hipStream_t). */
void* mhd_stream(mhd_ctx* ctx);
/* Enqueues max_steps steps on the context stream without any host
 * synchronisation or status read-back (for timing and graph capture). */
int mhd_enqueue_steps(mhd_ctx* ctx, double t_end, int nsteps);

/* Binary checkpoint of the whole solver state (no reference counterpart:
 * the reference writes only the final .dat, whose primitives cannot restore
 * U bitwise, mhd2d_glm_cpu.cpp:648-672).  The file holds a header (magic,
 * version, nx, ny, precision), the device control block (t, dt, c_h, step,
 * counters) and the raw state planes; loading it into a context of the same
 * nx, ny and precision continues the run bitwise.  A control block that no
 * save writes (a run state outside RunState, a first-bad-cell index outside
 * the grid) is refused; per-step queue counts in it are cleared.  Returns 0,
 * 1 (mismatch, out-of-range control block or I/O error) or 9. */
int mhd_checkpoint_save(mhd_ctx* ctx, const char* path);
int mhd_checkpoint_load(mhd_ctx* ctx, const char* path);

/* Tile shape a context of this form and size gets when cfg.tile is -1
 * (the autotuner's table, else the default).  No device needed. */
int mhd_autotuned_tile(int precision, int nx, int ny);
/* Stages per launch (1 or 2) a context of this form and size gets when
 * cfg.tile is -1 and cfg.fuse is 0.  No device needed.  No reference
 * counterpart (the reference launches one kernel per operation). */
int mhd_autotuned_fuse(int precision, int nx, int ny);

/* Measurement (no reference counterpart; the reference times its loop with
 * a host wall clock, mhd2d_glm_cpu.cpp:512,571-572): runs nsteps steps with
 * HIP events recorded on the context stream around each kernel and writes
 * the average duration in milliseconds of the stage-1, stage-2 and finish
 * kernels to avg_ms[0..2]. */
int mhd_time_kernels(mhd_ctx* ctx, double t_end, int nsteps, double* avg_ms);

/* Run without allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Equivalent to
 * mhd_create + mhd_set_state + mhd_advance + mhd_destroy on a state that is
 * already on the device: `state` (device memory, 9 planes `ld` elements
 * apart, element type as for mhd_set_state) starts at t = 0 and is advanced
 * in place until t >= t_end, a failure, or max_steps steps (< 0: no limit),
 * with the same kernels and results as mhd_advance without rollback.  The
 * stage-1 state, the control block and the MIXED form's fallback lists live
 * in a fixed scratch array of 24 MiB per device, reserved when the
 * library's code is loaded on that device and held until the process ends
 * (an FP64 grid of up to about 3.4e5 cells, e.g. 512 x 512); a larger grid
 * is rejected with 1 before any launch.  One stage per launch (cfg.fuse 2
 * is rejected), no selector, no graphs, no rollback snapshots (a MIXED run
 * that halts ends with its code 3/4, as mhd_advance does with
 * cfg.rollback_steps < 0); MIXED_KAHAN is rejected (its device state is two
 * FP32 planes per variable).  stream is a hipStream_t (NULL: the null
 * stream); the call returns once the run has stopped.  Every call on a
 * device shares that device's scratch array: calls that overlap on one
 * device race.  Returns 0, 1 (invalid arguments, message on stderr), 2/3/4
 * (halted) or 9, and fills *st if non-NULL. */
int mhd_run_noalloc(const mhd_cfg* cfg, void* state, size_t ld, double t_end,
                    long long max_steps, void* stream, mhd_status* st);

/* Operator-level calls: the reference's by-array physics functions, batched
 * over n states on the device.  Arrays are host AoS (state s, slot k at
 * s*9+k) of the precision's element type.
 *   mhd_glm_flux_batch  ~ mhdGlmFluxX/Y (mhd_glm_common.hpp:69-126);
 *                         dir 0 = x, 1 = y; status[s] = HLLD status
 *   mhd_cons_to_prim_batch ~ mhdGlmConsToPrim (mhd_glm_common.hpp:43-50)
 *   mhd_fast_speed_batch   ~ mhdFastSpeedX/Y (mhd_common.hpp:206-238) */
int mhd_glm_flux_batch(int precision, int dir, int n, const void* UL,
                       const void* UR, double ch, int hlld, double hlld_tol,
                       void* F, int* status);
int mhd_cons_to_prim_batch(int precision, int n, const void* U, void* W);
int mhd_fast_speed_batch(int precision, int dir, int n, const void* W,
                         void* out);

/* Learned precision selector (MIXED form; no reference counterpart).  Per
 * stage-kernel tile and step an MLP 12 -> 64 -> 64 -> 1 (BF16 MFMA, FP32
 * accumulate, sigmoid) scores 12 tile features (refreshed every 8 steps);
 * tiles scoring above the threshold are evaluated by the FP64 kernels, the
 * rest by the FP32-flux kernels.  A fraction audit_frac of the FP32 tiles
 * (a systematic sample, new each step) is recomputed in FP64 (audit); a
 * scaled difference above the budget makes the tile FP64 for the rest of
 * the run.  params holds mhd_mlp_param_count() FP32 weights in the
 * layout W1[16][64], b1[64], W2[64][64], b2[64], w3[64], b3; mean/stdev
 * standardise the 12 features.  params NULL turns the selector off. */
int mhd_selector_configure(mhd_ctx* ctx, const float* params, const float* mean,
                           const float* stdev, float threshold, float audit_frac,
                           float audit_budget);
/* out[0..3]: audited tiles, audit misses, FP64-routed tile-steps, steps --
 * counted over the steps the run actually took (the no-op tail of a batch of
 * steps after the run finished or halted counts nothing). */
int mhd_selector_stats(mhd_ctx* ctx, unsigned long long* out4);
/* Tiles of the selector's decomposition of this context (-1 if not MIXED). */
int mhd_selector_tiles(mhd_ctx* ctx);
/* Training data from the solver itself: the 12 raw features of every tile
 * (tiles x 16 floats, last 4 zero) of the current state, and every tile's
 * largest scaled difference between the FP32-flux and the FP64-flux stage-1
 * update of the current state (max over cells and the 8 MHD variables --
 * psi, which has no scale of its own, excluded -- of |a-b|/(|b|+rms_k),
 * rms_k the RMS of variable k of the FP64 update).  Audits use the same
 * measure with the RMS of the state at mhd_selector_configure.  Neither
 * call advances the run. */
int mhd_tile_features(mhd_ctx* ctx, float* feats);
int mhd_pair_tile_errors(mhd_ctx* ctx, float* err);
/* The MLP on the device: forward of n standardised samples (n x 16 floats),
 * gradients of the weighted binary cross-entropy, and Adam training
 * (epochs over shuffled minibatches; params updated in place). */
int mhd_mlp_param_count(void);
int mhd_mlp_forward(const float* params, const float* x, int n, float* prob);
int mhd_mlp_gradients(const float* params, const float* x, const float* y,
                      const float* w, int n, float* grad, float* loss);
int mhd_mlp_train(float* params, const float* x, const float* y, const float* w,
                  int n, int epochs, int batch, float lr, unsigned seed,
                  float* epoch_loss);

/* Ensembles (no reference counterpart: the reference runs one grid per
 * process).  n instances of sizes nx[i] x ny[i] share the configuration in
 * cfg (scheme, precision, tile) and advance together: each SSPRK2 stage is
 * one kernel over a device list of the running instances' (instance, tile)
 * work items, rebuilt on the device at the start of the stage
 * ($MHD_ENSEMBLE_KERNEL=queue: a persistent grid pulling the items from a
 * work queue), and each step ends with one control kernel; every instance keeps its own
 * dt, t, c_h, counters and halt state, exactly as a single context would
 * (an instance's result is bitwise that of mhd_advance on its own).  Calls
 * mirror the single-context ones with an instance index. */
typedef struct mhd_ensemble mhd_ensemble;
int mhd_ensemble_create(const mhd_cfg* cfg, int n, const int* nx, const int* ny,
                        mhd_ensemble** out);
void mhd_ensemble_destroy(mhd_ensemble* e);
const char* mhd_ensemble_last_error(const mhd_ensemble* e);
int mhd_ensemble_set_state(mhd_ensemble* e, int i, const void* soa9, size_t ld);
int mhd_ensemble_init_problem(mhd_ensemble* e, int i, const char* problem,
                              const double* params, int nparams);
/* Advances every instance until each reaches t_end or halts, or max_steps
 * steps (< 0: no limit); st (may be NULL) receives n statuses.  Returns the
 * largest instance status code. */
int mhd_ensemble_advance(mhd_ensemble* e, double t_end, long long max_steps,
                         mhd_status* st);
int mhd_ensemble_get_status(mhd_ensemble* e, int i, mhd_status* st);
int mhd_ensemble_get_state(mhd_ensemble* e, int i, int which, void* soa9, size_t ld);
int mhd_ensemble_enqueue_steps(mhd_ensemble* e, double t_end, int nsteps);
int mhd_ensemble_time_kernels(mhd_ensemble* e, double t_end, int nsteps, double* avg_ms);

#ifdef __cplusplus
}
#endif

#endif /* MHD_GLM_H */
