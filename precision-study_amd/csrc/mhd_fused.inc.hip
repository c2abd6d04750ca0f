// mhd_fused.inc.hip -- both SSPRK2 stages in one launch (fusion depth 2).
// Included by mhd_step.inc.hip inside namespace mhd::MHD_NS, after the
// single-stage tile pass whose pieces it shares (face_flux, advance_cell,
// the recovery and the checks), so each of the three translation units
// compiles it with its own floating point flags.
//
// One workgroup advances its TX x TY tile by a whole step
// (mhd2d_glm_cpu.cpp:526-566):
//   A1  U^n on the tile and a 4-cell halo (without the corners beyond 2
//       cells, which no stencil reaches) -> primitives in LDS; each thread
//       keeps the U^n of its stage-1 cells in registers for C1 and C2;
//   B1  the stage-1 faces of the tile and of a 2-cell halo ring around it:
//       the cells whose U1 the tile's stage-2 faces read;
//   C1  U1 of those cells (rhs, update, positivity check) -> primitives in
//       LDS, over A1's; the tile's own U1 stays in registers;
//   B2  the stage-2 faces of the tile;
//   C2  U^{n+1} of the tile -> the other state buffer; checks and speeds.
// U1 never leaves the chip: a cell-update moves U^n in and U^{n+1} out of
// HBM (144 B with an FP64 state) instead of the two launches' 360 B, for
// about a quarter more face evaluations (the ring's stage 1).  The ring's
// U1 is computed with the operations the owning tile uses, so in the exact
// forms the step is bitwise the two-launch step.  U^{n+1} cannot overwrite
// U^n, whose halo neighbouring tiles are still reading, so the state
// alternates between two buffers: Ctrl::cur names the current one, and
// k_finish_fused flips it when the step's result is the new buffer.
//
// Positivity failures keep the two-launch semantics.  Exact forms: the
// first unusable cell of stage 1 (code 3) or stage 2 (code 4) among the
// tiles' own cells, the state left as the two launches leave it (U^n on a
// stage-1 failure, the failed U^{n+1} on a stage-2 one), and the HLLD
// counters of stage 2 only when stage 1 passed.  Mixed form with the FP64
// fallback: a step in which any tile meets an unusable cell (its own or
// its ring's) is discarded and recomputed by the single-stage kernels,
// whose tiles fall back to FP64 one by one, in place in the current buffer.

template <int TX, int TY>
struct FusedGeom {
  static constexpr int NT = TX * TY;
  static constexpr int NW = (NT + 63) / 64, NTHR = NW * 64;
  // U^n box: tile + 4 cells on each side; box cell (x, y) is grid cell
  // (i0 + x - 4, j0 + y - 4), periodic.  The tile is x in [4, TX+4),
  // y in [4, TY+4).
  static constexpr int BX = TX + 8, BY = TY + 8, BN = BX * BY;
  // Stage-1 cells: the tile, and 2-cell arms on its rows (x in [2, 4) and
  // [TX+4, TX+6)) and on its columns (y in [2, 4) and [TY+4, TY+6)).
  static constexpr int NC1 = NT + 4 * TX + 4 * TY;
  // Stage-1 flux slots: x face (xf, y) between box cells xf-1 and xf of row
  // y, for y in [2, TY+6) and xf in [2, TX+7); y face (x, yf) likewise.
  // Only the slots a stage-1 cell reads are evaluated.
  static constexpr int XW = TX + 5, NXS = XW * (TY + 4);
  static constexpr int YW = TX + 4, NYS = YW * (TY + 5);
  static constexpr int NS1 = NXS + NYS;
  // Stage-1 faces as four groups: x faces of the tile's rows (xf in
  // [2, TX+7)), x faces of the 4 arm rows (xf in [4, TX+5)), y faces of the
  // tile's columns (yf in [2, TY+7)), y faces of the 4 arm columns (yf in
  // [4, TY+5)).
  static constexpr int G0 = TY * (TX + 5), G1 = 4 * (TX + 1);
  static constexpr int G2 = TX * (TY + 5), G3 = 4 * (TY + 1);
  static constexpr int NF1 = G0 + G1 + G2 + G3;
  // Stage 2: the single-stage geometry (tile + 2-cell box, its faces).
  static constexpr int HX = TX + 4, HN = HX * (TY + 4);
  static constexpr int NXF = (TX + 1) * TY, NF = NXF + TX * (TY + 1);
  __device__ static constexpr int xslot(int xf, int y) { return (y - 2) * XW + (xf - 2); }
  __device__ static constexpr int yslot(int x, int yf) { return NXS + (yf - 2) * YW + (x - 2); }
  // A stage-1 arm coordinate: k in [0, 4) -> 2, 3, n+4, n+5; the U^n
  // outer arm beyond it: 0, 1, n+6, n+7.
  __device__ static constexpr int arm(int k, int n) { return k < 2 ? 2 + k : n + 2 + k; }
  __device__ static constexpr int outer(int k, int n) { return k < 2 ? k : n + 4 + k; }
  // U^n cells outside the stage-1 cells: the outer arms and the 2x2
  // corners of the tile + 2 box.
  static constexpr int NOUT = 4 * TY + 4 * TX + 16;
  // Box cell of stage-1 cell e (tile cells first, row-major; then the arms
  // on the tile's rows, then those on its columns).
  __device__ static void c1_cell(int e, int& x, int& y) {
    if (e < NT) {
      const int r = e / TX;
      x = 4 + (e - r * TX);
      y = 4 + r;
    } else if (e < NT + 4 * TY) {
      const int m = e - NT;
      x = arm(m & 3, TX);
      y = 4 + (m >> 2);
    } else {
      const int m = e - NT - 4 * TY, k = m / TX;
      x = 4 + (m - k * TX);
      y = arm(k, TY);
    }
  }
  // Box cell of outer cell o.
  __device__ static void outer_cell(int o, int& x, int& y) {
    if (o < 4 * TY) {
      x = outer(o & 3, TX);
      y = 4 + (o >> 2);
    } else if (o < 4 * TY + 4 * TX) {
      const int m = o - 4 * TY, k = m / TX;
      x = 4 + (m - k * TX);
      y = outer(k, TY);
    } else {
      const int m = o - 4 * TY - 4 * TX;
      x = arm(m & 3, TX);
      y = arm(m >> 2, TY);
    }
  }
};

template <class RP, class RS, int TX, int TY>
struct FusedShared {
  using G = FusedGeom<TX, TY>;
  union {
    RP s1[9 * G::BN];  // A1-B1: primitives of U^n (box)
    RP s2[9 * G::HN];  // C1-B2: primitives of U1 (tile + 2 box)
  } p;
  union {
    RS s1[9 * G::NS1];  // B1-C1: stage-1 fluxes
    RS s2[9 * G::NF];   // B2-C2: stage-2 fluxes
  } f;
  unsigned long long sRed[5 * G::NW];
};
template <class Pol, int TX, int TY>
using FusedSharedOf = FusedShared<typename Pol::RP, typename Pol::RS, TX, TY>;

// Waves per SIMD the LDS footprint allows (160 KB per CU, 4 SIMDs), at
// most what the registers allow: an FP64 flux (about 250 registers) plus
// the U^n and U1 states this kernel holds across its face phases need more
// than the 256 of two waves, so FP64-flux forms run one.
template <class Pol, int TX, int TY>
struct FusedWaves {
  static constexpr int blocks = (int)(163840 / sizeof(FusedSharedOf<Pol, TX, TY>));
  static constexpr int w = blocks * FusedGeom<TX, TY>::NW / 4;
  static constexpr int cap = sizeof(typename Pol::R) == 8 ? 1 : StageWaves<Pol>::value;
  static constexpr int value = w < 1 ? 1 : (w > cap ? cap : w);
};

template <class Pol, bool MUSCL, bool HLLD, int TX, int TY>
__device__ __forceinline__ void step_tile_fused(const StageArgs<Pol>& a, int tile,
                                                FusedSharedOf<Pol, TX, TY>& sh) {
  using R = typename Pol::R;
  using RP = typename Pol::RP;
  using RS = typename Pol::RS;
  using T = typename Pol::T;
  using A = typename RhsType<Pol>::type;
  using G = FusedGeom<TX, TY>;
  constexpr int NT = G::NT, NW = G::NW, NTHR = G::NTHR, BX = G::BX, BN = G::BN;
  constexpr int HX = G::HX, HN = G::HN, NXF = G::NXF, NF = G::NF, NS1 = G::NS1;
  RP* sP1 = sh.p.s1;
  RP* sP2 = sh.p.s2;
  RS* sF1 = sh.f.s1;
  RS* sF2 = sh.f.s2;
  unsigned long long* sRed = sh.sRed;

  Ctrl<T>* ctrl = a.ctrl;
  const T ch_t = uni(ctrl->ch), dt = uni(ctrl->dt), dx = uni(ctrl->dx), dy = uni(ctrl->dy);
  const R ch = (R)ch_t;
  const Phys<R> c = make_phys<R>(a.tol);
  const Phys<RP> cp = make_phys<RP>(a.tol);
  const int nx = a.nx, ny = a.ny;
  const size_t ps = a.plane;
  const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
  const int i0 = tx * TX, j0 = ty * TY;
  const int tid = (int)threadIdx.x;
  const T damping = ch_t / smax((T)a.alpha, (T)1.0e-12);
  const A adx = (A)dx, ady = (A)dy, aidx = (A)a.inv_dx, aidy = (A)a.inv_dy;
  const bool pow2 = a.pow2 != 0;
  // Mixed form with the FP64 fallback: unusable cells vote the step out
  // instead of halting the run.
  const bool vote = Pol::kRedo && a.redo_list != nullptr;
  auto gidx = [&](int x, int y) -> size_t {
    return (size_t)wrap(j0 + y - 4, ny) * nx + wrap(i0 + x - 4, nx);
  };

  // A1: U^n -> primitives.  Thread tid loads stage-1 cells tid and
  // tid + NTHR, whose U^n it keeps in registers for C1 (and, for its tile
  // cell, C2), and outer cell tid; all loads are issued before any cell is
  // recovered (one memory round trip).
  static_assert(G::NC1 <= 2 * NTHR && G::NOUT <= NTHR, "phase A1 takes three cells per thread");
  CellIO<Pol> un[2];
  {
    bool have[3];
    int hh[3];
    CellIO<Pol> uo;
#pragma unroll
    for (int q = 0; q < 3; q++) {
      int x = 0, y = 0;
      if (q < 2) {
        const int e = tid + q * NTHR;
        have[q] = e < G::NC1;
        if (have[q]) G::c1_cell(e, x, y);
      } else {
        have[q] = tid < G::NOUT;
        if (have[q]) G::outer_cell(tid, x, y);
      }
      hh[q] = y * BX + x;
      if (have[q]) (q < 2 ? un[q] : uo).load(a.src, ps, gidx(x, y));
    }
#pragma unroll
    for (int q = 0; q < 3; q++) {
      if (!have[q]) continue;
      RP P[9];
      recover<MUSCL>(cp, q < 2 ? un[q] : uo, a.lowbeta_ratio, P);
#pragma unroll
      for (int k = 0; k < 9; k++) sP1[k * BN + hh[q]] = P[k];
    }
  }
  __syncthreads();

  // B1: stage-1 faces.  The HLLD counters count the faces the single-stage
  // kernel's tile counts (its x faces fc in [0, TX] of the tile's rows, y
  // faces fr in [0, TY] of its columns, each grid face once).
  unsigned d1 = 0, n1 = 0;
  for (int f = tid; f < G::NF1; f += NTHR) {
    bool ydir, owned = false;
    int base, stride, slot;
    if (f < G::G0) {
      const int row = f / (TX + 5), xf = 2 + (f - row * (TX + 5)), y = 4 + row;
      ydir = false;
      base = y * BX + xf - 2;
      stride = 1;
      slot = G::xslot(xf, y);
      const int fc = xf - 4;
      owned = fc >= 0 && fc <= TX && (j0 + row < ny) &&
              (fc >= 1 ? (i0 + fc - 1 < nx) : (i0 == 0));
    } else if (f < G::G0 + G::G1) {
      const int q = f - G::G0, k = q / (TX + 1), xf = 4 + (q - k * (TX + 1));
      const int y = G::arm(k, TY);
      ydir = false;
      base = y * BX + xf - 2;
      stride = 1;
      slot = G::xslot(xf, y);
    } else if (f < G::G0 + G::G1 + G::G2) {
      const int q = f - G::G0 - G::G1, k = q / TX, col = q - k * TX;
      const int x = 4 + col, yf = 2 + k;
      ydir = true;
      base = (yf - 2) * BX + x;
      stride = BX;
      slot = G::yslot(x, yf);
      const int fr = yf - 4;
      owned = fr >= 0 && fr <= TY && (i0 + col < nx) &&
              (fr >= 1 ? (j0 + fr - 1 < ny) : (j0 == 0));
    } else {
      const int q = f - G::G0 - G::G1 - G::G2, k = q >> 2;
      const int x = G::arm(q & 3, TX), yf = 4 + k;
      ydir = true;
      base = (yf - 2) * BX + x;
      stride = BX;
      slot = G::yslot(x, yf);
    }
    const int st = face_flux<Pol, MUSCL, HLLD, BN, NS1>(c, cp, ch, sP1, base, stride, ydir,
                                                         sF1, slot);
    if (HLLD && owned) {
      d1 += (st == HLLD_DEGENERATE);
      n1 += (st == HLLD_NONPHYSICAL);
    }
  }
  __syncthreads();

  // C1: U1 of the stage-1 cells the thread loaded in A1 (thread tid < NT
  // holds tile cell tid, the one it updates again in C2).
  CellIO<Pol> u1own;
  bool fail = false;
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const int e = tid + q * NTHR;
    if (e >= G::NC1) continue;
    int x, y;
    G::c1_cell(e, x, y);
    const int xl = G::xslot(x, y), yb = G::yslot(x, y);
    A rhs[9];
#pragma unroll
    for (int k = 0; k < 9; k++)
      rhs[k] = flux_div<A, RS>(sF1[k * NS1 + xl + 1], sF1[k * NS1 + xl],
                               sF1[k * NS1 + yb + G::YW], sF1[k * NS1 + yb], pow2, adx, ady,
                               aidx, aidy);
    const CellIO<Pol>& src = un[q];
    CellIO<Pol> nu;
    T out[9];
    advance_cell<1>(src, src, rhs, dt, damping, nu, out);
    const bool own = e < NT;
    const int gi = i0 + x - 4, gj = j0 + y - 4;
    const bool live = own && gi < nx && gj < ny;
    if constexpr (sizeof(R) == sizeof(T)) {
      // Exact forms: the reference's check on the tile's own cells (a ring
      // cell is checked by the tile that owns it).
      if (live) {
        const Phys<T> ct = make_phys<T>(a.tol);
        T P[9];
        if (!state_usable(ct, out, P)) atomicMin(&ctrl->bad[0], gj * nx + gi);
      }
    } else {
      double speed;
      bool valid;
      const bool unusable = !usable_and_speed_fp64(out, false, speed, valid);
      if (vote)
        fail = fail || unusable;
      else if (unusable && live)
        atomicMin(&ctrl->bad[0], gj * nx + gi);
    }
    if (own) u1own = nu;
    RP P[9];
    recover<MUSCL>(cp, nu, a.lowbeta_ratio, P);
    const int h2 = (y - 2) * HX + (x - 2);
#pragma unroll
    for (int k = 0; k < 9; k++) sP2[k * HN + h2] = P[k];
  }
  __syncthreads();

  // B2: stage-2 faces, as the single-stage kernel evaluates them.
  unsigned d2 = 0, n2 = 0;
  for (int f = tid; f < NF; f += NTHR) {
    const bool ydir = f >= NXF;
    int base, stride;
    bool owned;
    if (!ydir) {
      const int r = f / (TX + 1), fc = f - r * (TX + 1);
      base = (r + 2) * HX + fc;
      stride = 1;
      owned = (j0 + r < ny) && (fc >= 1 ? (i0 + fc - 1 < nx) : (i0 == 0));
    } else {
      const int g = f - NXF, fr = g / TX, cc = g - fr * TX;
      base = fr * HX + cc + 2;
      stride = HX;
      owned = (i0 + cc < nx) && (fr >= 1 ? (j0 + fr - 1 < ny) : (j0 == 0));
    }
    const int st = face_flux<Pol, MUSCL, HLLD, HN, NF>(c, cp, ch, sP2, base, stride, ydir, sF2, f);
    if (HLLD && owned) {
      d2 += (st == HLLD_DEGENERATE);
      n2 += (st == HLLD_NONPHYSICAL);
    }
  }
  __syncthreads();

  // C2: U^{n+1} of the tile's cells -> the other buffer.
  const int r = tid / TX, cc = tid - r * TX;
  const int gi = i0 + cc, gj = j0 + r;
  const bool live = tid < NT && gi < nx && gj < ny;
  unsigned long long sbits = 0;
  unsigned svalid = 0;
  if (live) {
    const size_t g = (size_t)gj * nx + gi;
    const int xl = r * (TX + 1) + cc, yb = NXF + r * TX + cc;
    A rhs[9];
#pragma unroll
    for (int k = 0; k < 9; k++)
      rhs[k] = flux_div<A, RS>(sF2[k * NF + xl + 1], sF2[k * NF + xl], sF2[k * NF + yb + TX],
                               sF2[k * NF + yb], pow2, adx, ady, aidx, aidy);
    CellIO<Pol> nu;
    T out[9];
    advance_cell<2>(u1own, un[0], rhs, dt, damping, nu, out);
    nu.store(a.dst, ps, g);
    if constexpr (sizeof(R) == sizeof(T)) {
      const Phys<T> ct = make_phys<T>(a.tol);
      T P[9];
      if (!state_usable(ct, out, P)) atomicMin(&ctrl->bad[1], (int)g);
      bool valid;
      T s;
      if constexpr (Pol::kRefMixed)
        s = refmix::cell_speed(out, valid);
      else
        s = cell_speed(ct, P, valid);
      sbits = speed_bits(s);
      svalid = valid ? 1u : 0u;
    } else {
      double speed;
      bool valid;
      const bool unusable = !usable_and_speed_fp64(out, true, speed, valid);
      if (vote)
        fail = fail || unusable;
      else if (unusable)
        atomicMin(&ctrl->bad[1], (int)g);
      sbits = speed_bits(speed);
      svalid = valid ? 1u : 0u;
    }
  }

  // Block reductions over one barrier: both stages' counters (staged for
  // the finish kernel), the speed maximum, the vote.
  const int lane = tid & 63, wid = tid >> 6;
  const unsigned wd1 = wave_sum_u32(d1), wn1 = wave_sum_u32(n1);
  const unsigned wd2 = wave_sum_u32(d2), wn2 = wave_sum_u32(n2);
  const unsigned long long wm = wave_max_u64(sbits);
  const unsigned wv = wave_sum_u32(svalid);
  const unsigned long long wf = __ballot(fail ? 1 : 0);
  if (lane == 0) {
    sRed[wid] = ((unsigned long long)wn1 << 32) | wd1;
    sRed[NW + wid] = ((unsigned long long)wn2 << 32) | wd2;
    sRed[2 * NW + wid] = wm;
    sRed[3 * NW + wid] = wv;
    sRed[4 * NW + wid] = wf;
  }
  __syncthreads();
  if (tid == 0) {
    unsigned long long dd1 = 0, nn1 = 0, dd2 = 0, nn2 = 0, m = 0, v = 0, any = 0;
    for (int w = 0; w < NW; w++) {
      dd1 += sRed[w] & 0xffffffffull;
      nn1 += sRed[w] >> 32;
      dd2 += sRed[NW + w] & 0xffffffffull;
      nn2 += sRed[NW + w] >> 32;
      m = u64max(m, sRed[2 * NW + w]);
      v += sRed[3 * NW + w];
      any |= sRed[4 * NW + w];
    }
    if (dd1) atomicAdd(&ctrl->fz_degenerate[0], dd1);
    if (nn1) atomicAdd(&ctrl->fz_nonphysical[0], nn1);
    if (dd2) atomicAdd(&ctrl->fz_degenerate[1], dd2);
    if (nn2) atomicAdd(&ctrl->fz_nonphysical[1], nn2);
    publish_speed(ctrl, m, v != 0);
    if (any) atomicOr(&ctrl->fz_fail, 1);
  }
}

// Buffers of a two-stage step: U^n is b[cur], U^{n+1} goes to b[1-cur].
template <class Pol>
struct FusedArgs {
  StageArgs<Pol> a;        // src = uold = b0, dst = b1
  CurBuffers<Pol> bufs;    // src_alt = uold_alt = b1, dst_alt = b0
};

template <class Pol, bool MUSCL, bool HLLD, int TX, int TY>
__global__ __launch_bounds__((FusedGeom<TX, TY>::NTHR))
__attribute__((amdgpu_waves_per_eu(FusedWaves<Pol, TX, TY>::value))) void k_step_fused(
    FusedArgs<Pol> fa) {
  __shared__ FusedSharedOf<Pol, TX, TY> sh;
  if (fa.a.ctrl->run != RUN_ACTIVE) return;
  if constexpr (Pol::kRedo) {
    // Rollback window: the step is voted out and recomputed by the
    // single-stage kernels, whose tiles then all go to the FP64 kernel.
    if (fa.a.redo_list && uni(fa.a.ctrl->step) < uni(fa.a.ctrl->exact_until)) {
      if (threadIdx.x == 0 && blockIdx.x == 0) atomicOr(&fa.a.ctrl->fz_fail, 1);
      return;
    }
  }
  StageArgs<Pol> a = fa.a;
  follow_cur(a, fa.bufs);
  step_tile_fused<Pol, MUSCL, HLLD, TX, TY>(a, xcd_tile((int)blockIdx.x, a.ntiles), sh);
}

// Mixed form with the FP64 fallback, after the two-stage kernel: when a tile
// voted the step out, its speed maxima are dropped and the single-stage
// kernels are given every tile (Ctrl::fz_count).
template <class T>
__global__ void k_fused_gate(Ctrl<T>* ctrl, int ntiles) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (ctrl->run != RUN_ACTIVE || ctrl->fz_fail == 0) return;
  for (int q = 0; q < Ctrl<T>::kSpeedSlots; q++) ctrl->speed_slot[q] = 0;
  ctrl->speed_acc = 0;
  ctrl->valid_acc = 0;
  ctrl->fz_count = ntiles;
}

// End of a two-stage step: the staged counters, the buffer flip, then the
// common end of step (finish_one).
template <class T>
__global__ void k_finish_fused(Ctrl<T>* ctrl) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (ctrl->run == RUN_ACTIVE) {
    if (ctrl->fz_fail) {
      // Recomputed in place by the single-stage kernels.
      ctrl->fz_redo_steps += 1;
    } else {
      ctrl->degenerate += ctrl->fz_degenerate[0];
      ctrl->nonphysical += ctrl->fz_nonphysical[0];
      if (ctrl->bad[0] == INT_MAX) {
        // Stage 2 counts as in the two-launch step, and the state (U^{n+1},
        // or the state that failed stage 2's check) is the other buffer.
        ctrl->degenerate += ctrl->fz_degenerate[1];
        ctrl->nonphysical += ctrl->fz_nonphysical[1];
        ctrl->cur ^= 1;
      }
    }
  }
  ctrl->fz_fail = 0;
  ctrl->fz_count = 0;
  ctrl->fz_degenerate[0] = ctrl->fz_degenerate[1] = 0;
  ctrl->fz_nonphysical[0] = ctrl->fz_nonphysical[1] = 0;
  finish_one(ctrl);
}

// Host launcher of a two-stage step (16x15 tiles, MUSCL + HLLD).
template <class Pol>
static hipError_t launch_step_fused(const GridDesc& g, void* u, void* ualt, void* u1, void* ctrl,
                                    const int* all_tiles, hipStream_t s, hipEvent_t* ev) {
  using S = typename Pol::S;
  using T = typename Pol::T;
  constexpr int TX = 16, TY = 15;
  if (!(g.muscl && g.hlld) || g.tile != TILE_16x15 || g.redo_all) return hipErrorInvalidValue;
  const bool vote = Pol::kRedo && g.redo_list[0] && g.redo_list[1];
  if (vote && !all_tiles) return hipErrorInvalidValue;
  const int* cur = (const int*)((char*)ctrl + offsetof(Ctrl<T>, cur));
  FusedArgs<Pol> fa;
  fa.a = stage_args<Pol, TX, TY>(g, ctrl);
  fa.a.tiles = nullptr;
  fa.a.src = fa.a.uold = (const S*)u;
  fa.a.dst = (S*)ualt;
  fa.a.redo_list = vote ? g.redo_list[0] : nullptr;
  fa.bufs.cur = cur;
  fa.bufs.src_alt = fa.bufs.uold_alt = (const S*)ualt;
  fa.bufs.dst_alt = (S*)u;
  if (ev) (void)hipEventRecord(ev[0], s);
  hipLaunchKernelGGL((k_step_fused<Pol, true, true, TX, TY>), dim3(fa.a.ntiles),
                     dim3(FusedGeom<TX, TY>::NTHR), 0, s, fa);
  if (ev) (void)hipEventRecord(ev[1], s);
  hipError_t e = hipGetLastError();
  if (vote && e == hipSuccess) {
    hipLaunchKernelGGL(k_fused_gate<T>, dim3(1), dim3(64), 0, s, (Ctrl<T>*)ctrl, fa.a.ntiles);
    const int* count = (const int*)((char*)ctrl + offsetof(Ctrl<T>, fz_count));
    const Launchers* F64 = launchers_exact(MODE_FP64);
    // Stage 1 over every tile, U^n -> u1, then the FP64 re-evaluation of
    // the tiles it queued; stage 2 in place in the current buffer.
    for (int stage = 1; stage <= 2 && e == hipSuccess; stage++) {
      ListJob j;
      if (stage == 1) {
        j = ListJob{u, u, u1, ctrl, all_tiles, count};
        j.src_alt = j.uold_alt = ualt;
        j.dst_alt = u1;
      } else {
        j = ListJob{u1, u, u, ctrl, all_tiles, count};
        j.src_alt = u1;
        j.uold_alt = ualt;
        j.dst_alt = ualt;
      }
      j.cur = cur;
      e = launch_list_t<Pol, true, true, TX, TY>(g, stage, &j, 1, s, g.redo_list[stage - 1]);
      if (e != hipSuccess) break;
      ListJob q = j;
      q.list = g.redo_list[stage - 1];
      q.count = redo_counter_of<T>(ctrl, stage);
      e = F64->stage_listn(g, stage, &q, 1, s);
    }
  }
  if (ev) (void)hipEventRecord(ev[2], s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_finish_fused<T>, dim3(1), dim3(64), 0, s, (Ctrl<T>*)ctrl);
    e = hipGetLastError();
  }
  if (ev) (void)hipEventRecord(ev[3], s);
  return e;
}
