// mhd_kernels.h -- internal interface between the host library (mhd_api.cpp,
// compiled by the host C++ compiler) and the HIP kernels (mhd_kernels.hip).
// Not part of the public C ABI (include/mhd_glm.h).
#ifndef MHD_KERNELS_H
#define MHD_KERNELS_H

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>

namespace mhd {

enum Precision { PREC_FP64 = 0, PREC_FP32 = 1 };
enum RunState { RUN_ACTIVE = 0, RUN_DONE = 1, RUN_HALTED = 2 };

// Device-resident step control for one solver instance.  T is the type of
// time, dt, c_h and the CFL arithmetic: the arithmetic type of the run
// (float for FP32, double for FP64), as in the reference (Real t, dt).
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
};

// Launch parameters of one solver instance (one grid).
struct GridDesc {
  int nx, ny;
  size_t plane;      // elements per variable plane (>= nx*ny)
  int muscl, hlld;   // reconstruction order / Riemann solver
  double alpha;      // GLM damping parameter
  double tol;        // HLLD degeneracy tolerance
  int tile;          // tile-shape id (see mhd_kernels.hip)
};

// Tile shapes available to the stage kernels.
enum TileShape { TILE_16x16 = 0, TILE_32x8 = 1, TILE_16x8 = 2, TILE_COUNT = 3 };

// FP64 and FP32 entry points.  All launches go to `stream`; none synchronises.
// State buffers are 9 planes of `plane` elements (SoA), cell (i,j) at j*nx+i.
template <class R>
hipError_t launch_speed(const GridDesc& g, const R* state, Ctrl<R>* ctrl,
                        hipStream_t stream);
// Moves the speed accumulated by launch_speed into the current-state slot.
template <class R>
hipError_t launch_commit_speed(Ctrl<R>* ctrl, hipStream_t stream);
// Sets a new end time and re-prepares the coming step (no-op when halted).
template <class R>
hipError_t launch_set_target(Ctrl<R>* ctrl, double t_end, hipStream_t stream);
template <class R>
hipError_t launch_step(const GridDesc& g, R* u, R* u1, Ctrl<R>* ctrl,
                       hipStream_t stream);

// One step with an event recorded before each of its three kernels and
// after the last (ev[0..3]); for kernel timing.
template <class R>
hipError_t launch_step_timed(const GridDesc& g, R* u, R* u1, Ctrl<R>* ctrl,
                             hipStream_t stream, hipEvent_t ev[4]);

// Operator-level batch kernels (device arrays, AoS 9 per state).
template <class R>
hipError_t launch_glm_flux_batch(int n, int dir, const R* UL, const R* UR,
                                 R ch, int hlld, double tol, R* F, int* status,
                                 hipStream_t stream);
template <class R>
hipError_t launch_cons_to_prim_batch(int n, const R* U, R* W,
                                     hipStream_t stream);
template <class R>
hipError_t launch_fast_speed_batch(int n, const R* W, int dir, R* out,
                                   hipStream_t stream);

}  // namespace mhd

#endif  // MHD_KERNELS_H
