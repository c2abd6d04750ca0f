// mhd_select.h -- internal interface of the learned precision selector
// kernels (k_select.hip).  Not part of the public C ABI.
#ifndef MHD_SELECT_H
#define MHD_SELECT_H

#include <hip/hip_runtime_api.h>

#include <cstddef>

namespace mhd {
namespace sel {
constexpr int NFEAT = 12;   // features per tile
constexpr int NIN = 16;     // features padded to the MFMA K dimension
constexpr int H = 64;       // hidden width
constexpr int kParams = NIN * H + H + H * H + H + H + 1;  // FP32 master weights
}  // namespace sel

// feats: ntiles x NIN raw features of the FP64 state (c_h read from *ctrl_ch),
// tiles of tx x ty <= 256 cells (hipErrorInvalidValue otherwise).
hipError_t sel_tile_features(const double* u, size_t plane, int nx, int ny, int tx, int ty,
                             const void* ctrl_ch, float* feats, hipStream_t s);
// Per-plane sums of squares of an FP64 state (the error scale).
hipError_t sel_sumsq(const double* u, size_t plane, long long cells, double* sumsq9, hipStream_t s);
// err[t]: max over tile t's cells and the 8 MHD variables (psi excluded) of
// |a-b| / (|b| + s_k), s_k the grid RMS of rho, |momentum|, E or |B| (from
// the per-plane sums of squares sumsq9); every tile, or (list non-null) the
// tiles list[0 .. *count).
hipError_t sel_tile_errors(const double* a, const double* b, size_t plane, int nx, int ny, int tx,
                           int ty, const double* sumsq9, const int* list, const int* count,
                           float* err, hipStream_t s);
// prob/flags of n samples (raw features standardised with mean/istd when
// mean is non-null); flags[s] = prob > threshold || sticky[s].
hipError_t sel_mlp_infer(const float* raw, int n, const float* params, const float* mean,
                         const float* istd, float threshold, const unsigned char* sticky,
                         float* prob, unsigned char* flags, hipStream_t s);
// Gradient of the minibatch (samples idx[offset .. offset+count)) as one
// row of kParams + 1 floats (gradient, then the weighted loss) per block of
// 64 samples: partial must hold ceil(count / 64) rows.
hipError_t sel_mlp_grad(const float* X, const float* Y, const float* W, const int* idx, int offset,
                        int count, float wsum, const float* params, float* partial,
                        hipStream_t s);
// grad = sum of the nparts rows in row order; their loss is added to *loss.
hipError_t sel_sum_partials(const float* partial, int nparts, float* grad, float* loss,
                            hipStream_t s);
// Adam step with the gradient summed from the nparts rows in row order
// (the loss added to *loss when non-null).
hipError_t sel_adam(float* params, const float* partial, int nparts, float* m, float* v, int step,
                    float lr, float* loss, hipStream_t s);
// Tile lists of the current flags, each in tile order: lists[0 .. counts[0])
// the flagged (FP64) tiles, lists[2 ntiles .. 2 ntiles + counts[2]) the
// unflagged (mixed) tiles.  flags must be 16-byte aligned (hipErrorInvalidValue
// otherwise).
hipError_t sel_route(const unsigned char* flags, int ntiles, int* lists, int* counts,
                     hipStream_t s);
// The step's audit sample, a fraction `frac` of the mixed list taken at a
// stride of 1/frac from a phase hashed from the device step counter ->
// lists[ntiles .. ntiles + counts[1]); also copies the control block into
// ctrl_copy (the audits' snapshot).  Nothing is sampled or counted while
// *run (the control block's run state) is not 0 (active).
hipError_t sel_audit_pick(int ntiles, const int* step, const int* run, float frac, int* lists,
                          int* counts, unsigned long long* counters, const void* ctrl,
                          void* ctrl_copy, size_t ctrl_bytes, hipStream_t s);
// Errors of the audit list (a: FP32-flux, b: FP64 stage output) as
// sel_tile_errors; tiles above `budget` become sticky and flagged and join
// the FP64 list (lists[0 .. counts[0])).
hipError_t sel_audit(const double* a, const double* b, size_t plane, int nx, int ny, int tx, int ty,
                     const double* sumsq9, int* lists, int* counts, float budget,
                     unsigned char* sticky, unsigned char* flags, float* err,
                     unsigned long long* counters, hipStream_t s);
hipError_t sel_copy_words(const void* src, void* dst, size_t bytes, hipStream_t s);


}  // namespace mhd

#endif  // MHD_SELECT_H
