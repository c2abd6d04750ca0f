// mhd_kernels.h -- internal interface between the host library (mhd_api.cpp,
// compiled by the host C++ compiler) and the HIP kernels (k_exact.hip,
// k_fma.hip, k_mixed.hip).  Not part of the public C ABI (include/mhd_glm.h).
#ifndef MHD_KERNELS_H
#define MHD_KERNELS_H

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>

namespace mhd {

// Arithmetic forms of the step (values match enum mhd_precision).
enum Mode {
  MODE_FP64 = 0,         // FP64, contraction off: bitwise the reference CPU build
  MODE_FP32 = 1,         // FP32, contraction off: bitwise the reference FP32 build
  MODE_MIXED = 2,        // FP32 flux/reconstruction/GLM, FP64 state + update + time
  MODE_MIXED_KAHAN = 3,  // as MIXED with an FP32 + FP32 compensated state
  MODE_FP64_FMA = 4,     // FP64 with FMA contraction
  MODE_REF_MIXED = 5,    // the reference's stage-level mixed build (FP64 flux + update)
  // Allocation variants of MIXED (single grids, MUSCL + HLLD, 16x15 tiles):
  MODE_MIXED_RECON64 = 6,        // FP64 recovery and reconstruction
  MODE_MIXED_FLUX64 = 7,         // FP64 GLM-HLLD flux, differenced in FP64
  MODE_MIXED_FLUX64_NARROW = 8,  // FP64 flux narrowed to FP32 before differencing
  MODE_COUNT = 9
};
enum RunState { RUN_ACTIVE = 0, RUN_DONE = 1, RUN_HALTED = 2 };

// Device-resident step control for one solver instance.  T is the type of
// time, dt, c_h and the CFL arithmetic: float for FP32 (the reference keeps
// t and dt in Real), double otherwise.
template <class T>
struct Ctrl {
  T t, dt, ch, t_end, cfl, dx, dy;
  unsigned long long speed_cur;   // max signal speed of the current state (bits of T)
  unsigned long long speed_acc;   // max accumulated by stage 2 for the next state
  unsigned long long degenerate;  // HLLD->HLL fallbacks, by reason
  unsigned long long nonphysical;
  int valid_cur, valid_acc;       // some cell gave a non-NaN signal speed
  int step;                       // completed steps
  int run;                        // RunState
  int code;                       // exit code when halted (2, 3, 4)
  int bad[2];                     // first unusable cell (j*nx+i) per stage
  int fail_cell;                  // cell reported on halt
  int fail_step;
  T fail_t, fail_dt, fail_ch;
  // FP64 fallback of the mixed form: tiles queued for re-evaluation in this
  // step's stage 1 / stage 2 (reset when a step is prepared), and the total
  // over the run.
  int redo_count[2];
  unsigned long long redo_tiles;
  int rate_f32;  // CFL rate in float (the reference-mixed form)
  // Two-stage steps (fusion depth 2): the buffer that holds the current
  // state (0: the context's state buffer, 1: its alternate); the step's HLLD
  // counters per stage, added by the finish kernel once the step's outcome
  // is known; in the mixed form with the FP64 fallback, the flag of a step
  // in which some tile met an unusable cell, the tile count that the
  // unfused stage kernels then recompute the step over (else 0), and the
  // number of steps so recomputed.
  int cur;
  int fz_fail;
  int fz_count;
  int fz_pad;
  unsigned long long fz_degenerate[2], fz_nonphysical[2];
  unsigned long long fz_redo_steps;
  // Rollback window of the mixed form: while step < exact_until every tile
  // of the instance goes to the FP64 kernel through the fallback queue (set
  // by the host after it restored an earlier snapshot of a failed run);
  // exact_steps counts the completed steps so taken.
  int exact_until;
  // Injected failure (test hook of the rollback, $MHD_FAULT_STEP; 0: none):
  // with fail_inject = s + 1, stage 1 of step s of the mixed form counts
  // cell 0 as an FP64-confirmed unusable cell, outside rollback windows.
  int fail_inject;
  unsigned long long exact_steps;
  // Per-block partial maxima of the signal speed, spread over kSpeedSlots
  // addresses (block b uses slot b % kSpeedSlots) so that tens of thousands
  // of blocks do not serialise on one atomic; folded into speed_acc by the
  // next control kernel.
  static constexpr int kSpeedSlots = 64;
  unsigned long long speed_slot[kSpeedSlots];
};

// The XCD banding of a list of n items walked by a grid of at least n
// blocks (blocks are dealt to the 8 XCDs round-robin): block bid takes item
// xcd_item(bid, n), so XCD x takes the x-th eighth of the list; -1 for the
// blocks left over (the highest-numbered ones, so no XCD waits on another).
constexpr int xcd_item(int bid, int n) {
  const int q = n >> 3, r = n & 7, x = bid & 7, s = bid >> 3;
  if (s >= q + (x < r ? 1 : 0)) return -1;
  return x * q + (x < r ? x : r) + s;
}

// Launch parameters of one solver instance (one grid).
struct GridDesc {
  int nx, ny;
  size_t plane;          // elements per variable plane (>= nx*ny)
  int muscl, hlld;       // reconstruction order / Riemann solver
  double alpha;          // GLM damping parameter
  double tol;            // HLLD degeneracy tolerance
  double lowbeta_ratio;  // mixed forms: FP64 recovery where E > ratio * e_int
  int tile;              // tile-shape id
  // Per-tile precision routing (learned selector).  With `tiles` non-null,
  // block b of a stage launch evaluates tile tiles[b] (-1: none) instead of
  // the b-th tile of the whole grid: the routing kernel writes the mixed
  // tiles there in each XCD's block order (xcd_item), and an audit miss
  // strikes its tile out.
  const int* tiles = nullptr;
  // FP64 fallback of the mixed form.  With redo_list[stage-1] non-null, a
  // mixed stage tile with an unusable cell (the positivity check) writes
  // nothing and appends its index to that list (count in the control
  // block); `redo` (the FP64 launchers' stage_list) then re-evaluates the
  // listed tiles with the FP64 kernel, whose own check halts the run as the
  // FP64 form would.  redo_all queues every tile (check mode: the step is
  // then the FP64 step).
  int* redo_list[2] = {nullptr, nullptr};
  hipError_t (*redo)(const GridDesc&, int stage, const void* src, const void* uold, void* dst,
                     void* ctrl, const int* list, const int* count, hipStream_t) = nullptr;
  int redo_all = 0;
};

// Tile shapes of the stage kernels (cells TX x TY).  The face list of a
// tile has 2*TX*TY + TX + TY entries; 16x15, 32x7 and 16x7 fit it into whole
// waves evenly (511 / 487 / 247 faces over 256 / 256 / 128 lanes).
enum TileShape {
  TILE_16x16 = 0, TILE_32x8 = 1, TILE_16x8 = 2, TILE_16x15 = 3, TILE_32x7 = 4,
  TILE_16x7 = 5, TILE_COUNT = 6
};
constexpr int kDefaultTile = TILE_16x15;
inline void tile_dims(int tile, int& tx, int& ty) {
  static constexpr int dims[][2] = {{16, 16}, {32, 8}, {16, 8}, {16, 15}, {32, 7}, {16, 7}};
  static_assert(sizeof(dims) / sizeof(dims[0]) == TILE_COUNT, "one entry per tile shape");
  if (tile < 0 || tile >= TILE_COUNT) tile = kDefaultTile;
  tx = dims[tile][0];
  ty = dims[tile][1];
}

// How an ensemble's stage kernels take the live work items: one block per
// item, the grid covering every instance's tiles (blocks past the live
// items exit at once), or a persistent grid claiming items from a work
// queue.
enum EnsKernel { ENS_FLAT = 0, ENS_QUEUE = 1 };

// Device live list of an ensemble: a 64-byte header (live work items, live
// instances), the first live work item of each live instance (n ints,
// padded to 64 bytes), then one kEnsLiveEntryBytes entry per live instance
// (its descriptor, its first item in the fixed numbering over every
// instance, which the FP64 fallback queue uses, and the control-block values
// the stage reads).
constexpr size_t kEnsLiveHeaderBytes = 64;
constexpr size_t kEnsLiveEntryBytes = 128;
constexpr size_t ens_live_first_bytes(int n) { return ((size_t)n * sizeof(int) + 63) / 64 * 64; }
constexpr size_t ens_live_bytes(int n) {
  return kEnsLiveHeaderBytes + ens_live_first_bytes(n) + (size_t)n * kEnsLiveEntryBytes;
}

// Device description of an ensemble (type-erased; built by mhd_api.cpp).
struct EnsLaunch {
  const void* desc;     // device array of per-instance descriptors (desc_bytes each)
  const int* prefix;    // device: first work item of each instance, then the total
  int n, total;         // instances, work items (tiles over all instances)
  unsigned* queue;      // device: 16 work-queue counters
  void* ctrls;          // device: n control blocks, ctrl_bytes apart
  double tol, alpha, lowbeta_ratio;
  int muscl, hlld, tile;
  int grid_blocks;      // persistent grid; <= 0: occupancy x CUs
  int cus;              // compute units of the device
  // Live list (ens_live_bytes(n) bytes of device memory), rebuilt at the
  // start of every stage: the running instances' work items numbered from 0.
  void* live = nullptr;
  int kernel = ENS_FLAT;  // EnsKernel: how the stage kernels take the items
  // FP64 fallback of the mixed form (as GridDesc): queued work items (global
  // tile indices) per stage, counts in queue[16 + stage - 1], re-evaluated
  // by `redo` (the FP64 launchers' ens_redo).
  int* redo_items[2] = {nullptr, nullptr};
  int redo_all = 0;
  hipError_t (*redo)(const EnsLaunch&, int stage, hipStream_t) = nullptr;
};
// Work-queue counters of an ensemble: 8 chunks per stage, then the two
// fallback counts.
constexpr int kEnsQueueWords = 18;

// One segment of a list-driven stage launch: buffers, control block, and a
// device list of tile indices with its device count (list null: empty).
struct ListJob {
  const void* src;
  const void* uold;
  void* dst;
  void* ctrl;
  const int* list;
  const int* count;
  // With `cur` non-null the buffers follow the device word *cur: when it is
  // non-zero the *_alt buffers are used instead (two-stage steps, whose
  // state alternates between two buffers).
  const int* cur = nullptr;
  const void* src_alt = nullptr;
  const void* uold_alt = nullptr;
  void* dst_alt = nullptr;
};

// Launchers of one arithmetic form.  All launches go to `stream`; none
// synchronises.  State buffers hold `state_planes` planes of `plane`
// elements of `state_elem` bytes (SoA), cell (i,j) at j*nx+i.
struct Launchers {
  int mode;
  size_t ctrl_bytes;
  int time_is_float;
  size_t state_elem;
  int state_planes;
  hipError_t (*speed)(const GridDesc&, const void* state, void* ctrl, hipStream_t);
  hipError_t (*commit)(void* ctrl, hipStream_t);
  hipError_t (*set_target)(void* ctrl, double t_end, hipStream_t);
  // One step; with ev non-null an event is recorded before each of the three
  // kernels and after the last (ev[0..3]).
  hipError_t (*step)(const GridDesc&, void* u, void* u1, void* ctrl, hipStream_t,
                     hipEvent_t* ev);
  // One step as one two-stage kernel (MUSCL + HLLD, 16x15 tiles): the state
  // moves between u and ualt (Ctrl::cur names the current one).  With the
  // mixed form's FP64 fallback (g.redo_list set), a step in which a tile met
  // an unusable cell is recomputed by the unfused stage kernels over
  // all_tiles (a device list 0..ntiles-1), with u1 as their stage-1 output.
  // Events as for `step` (ev[1]..ev[2] brackets that recomputation).
  hipError_t (*step_fused)(const GridDesc&, void* u, void* ualt, void* u1, void* ctrl,
                           const int* all_tiles, hipStream_t, hipEvent_t* ev);
  // One stage (1 or 2) with explicit buffers: src (state L() is evaluated
  // on), uold (stage 2: U^n), dst; and the end-of-step kernel.
  hipError_t (*stage)(const GridDesc&, int stage, const void* src, const void* uold, void* dst,
                      void* ctrl, hipStream_t);
  hipError_t (*finish)(void* ctrl, hipStream_t);
  // One stage over the tiles list[0 .. *count) (device list and count, read
  // by the kernel), a fixed grid striding over the list (FP64 launchers
  // only; null elsewhere): the mixed form's fallback and the selector's
  // FP64 tiles.
  hipError_t (*stage_list)(const GridDesc&, int stage, const void* src, const void* uold,
                           void* dst, void* ctrl, const int* list, const int* count, hipStream_t);
  // The same over 1..3 segments in one launch (FP64 launchers only).
  hipError_t (*stage_listn)(const GridDesc&, int stage, const ListJob* jobs, int njobs,
                            hipStream_t);
  // Ensembles.
  size_t desc_bytes;
  // Fills descriptor i of a host array; returns the instance's tile count.
  int (*ens_fill)(void* host_desc, int i, void* u, void* u1, size_t plane, int nx, int ny,
                  void* ctrl, int tile);
  // One step of every running instance: the live-list kernel, the two stage
  // kernels (EnsLaunch::kernel), the FP64 fallback launches, the finish.
  hipError_t (*ens_step)(const EnsLaunch&, hipStream_t, hipEvent_t* ev);
  hipError_t (*ens_set_target)(void* ctrls, int n, double t_end, unsigned* queue, hipStream_t);
  // FP64 re-evaluation of an ensemble's queued items (FP64 launchers only).
  hipError_t (*ens_redo)(const EnsLaunch&, int stage, hipStream_t);
};

const Launchers* launchers_exact(int mode);  // MODE_FP64, MODE_FP32, MODE_REF_MIXED
const Launchers* launchers_fma();            // MODE_FP64_FMA
const Launchers* launchers_mixed(int mode);  // MODE_MIXED, MODE_MIXED_KAHAN, the variants

// Scratch of the no-allocation run (k_noalloc.hip): a fixed device array per
// device, reserved by the code object's loader.  24 MiB holds the stage-1
// state of an FP64 grid of up to 3.4e5 cells (512 x 512 with room to spare),
// its control block and the mixed form's fallback lists.
constexpr size_t kNoAllocScratchBytes = (size_t)24 << 20;
// Device address of the current device's scratch array (looked up once per
// device) and its size in bytes.
hipError_t noalloc_scratch(void** ptr, size_t* bytes);

// Operator-level batch kernels (exact arithmetic; device arrays, AoS 9 per
// state).  R is double or float.
template <class R>
hipError_t launch_glm_flux_batch(int n, int dir, const R* UL, const R* UR, R ch, int hlld,
                                 double tol, R* F, int* status, hipStream_t stream);
template <class R>
hipError_t launch_cons_to_prim_batch(int n, const R* U, R* W, hipStream_t stream);
template <class R>
hipError_t launch_fast_speed_batch(int n, const R* W, int dir, R* out, hipStream_t stream);

}  // namespace mhd

#endif  // MHD_KERNELS_H
