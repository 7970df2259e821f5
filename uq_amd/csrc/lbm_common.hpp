// lbm_common.hpp — structures shared by the host runtime and the gfx950
// kernels of the D2Q13 thermal FD-LBM.
//
// Device layout of one instance (a grid of nx*ny cells):
//   * distributions: 26 planes (f0..f12, h0..h12), each ny rows of `pitch`
//     cells, x fastest — structure of arrays, so a wavefront reading one
//     direction of consecutive cells issues one coalesced row segment;
//   * cell type map: uint8 [ny][pitch] (0 fluid, 1 permanent, 2 solid);
//   * wall closure factor alpha: double [ny][pitch] (read at solid cells).
// Storage is FP64 or FP32 (mixed precision) per instance group. FP32
// storage holds every distribution as its deviation from the instance's
// reference equilibrium `ref` (one value per direction for f and for h),
// and each state buffer carries four FP64 planes with the cell's conserved
// moments (sum f, sum f e_x, sum f e_y, sum h): a reader restores
// f = dev + ref in FP64 and corrects it so that its moments equal the
// stored ones, so FP32 rounding never touches mass, momentum or energy.
#pragma once

#include <cstdint>

namespace uqlbm {

constexpr int kQ = 13;
constexpr int kPlanes = 26;

// Boundary flag bits (reference cpuLBM.h:47-49).
enum : int {
  kPeriodicX = 1,
  kPeriodicY = 2,
  kMirrorN = 4,
  kMirrorS = 8,
  kMirrorE = 16,
  kMirrorW = 32,
};

// Per-instance record read by every kernel. Plain data; lives in device
// memory as an array indexed by instance id.
struct InstDev {
  int nx, ny, pitch, flags;
  int flux_method;     // 0 NND, 1 WENO5
  int parity0;         // which state buffer holds the current state at launch step 0
  int step_base;       // steps taken before the current launch sequence
  long long plane;     // elements per plane (pitch*ny)
  void* state[2];      // ping-pong state buffers (26 planes each)
  void* f2;            // stage-2 distributions (26 planes)
  void* f3;            // stage-3 distributions (26 planes)
  void* dflux;         // net fluxes dF(f2) of directions 1..12, f then h (24 planes)
  const uint8_t* ctype;
  double* alpha;       // wall closure factor, [ny][pitch]
  int* fail;           // per-instance: first failing step (INT_MAX when none)
  double* mom[2];      // FP32 storage: conserved moments of state[k] (4 FP64
                       // planes each); null for FP64 storage
  double ref[26];      // FP32 storage: reference equilibrium f0..f12, h0..h12
                       // (zero for FP64 storage)
  double refsum[4];    // FP32 storage: conserved sums of ref (mass, x and y
                       // momentum, energy)
  double rho0, t0;     // FP32 storage: density and T/Tc of the reference equilibrium
  float eqref[8];      // FP32 storage: the base polynomials at rest at t0
  int eq32;            // FP32 storage near the reference: equilibria in FP32 (eq_dev), else FP64
                       // (f0, f1, f5, f9 per unit density; h0, h1, h5, h9)
  double i2c;          // 1/(2c): momentum correction per axis population pair
  // physics (reference setupVars, cpuLBM.cpp:360-450)
  double dt, dx, dy;
  double rdx, rdy;     // dt/dx, dt/dy
  double c;            // lattice speed sqrt(R*Tc)
  double sRTc;         // sqrt(R*Tc) as the equilibrium computes it
  double Tc, R, b, K, bR, RTc;
  double mu, Pr;
  // reciprocals and differences formed once on the host so that the
  // per-cell arithmetic multiplies instead of divides
  double isRTc, iTc, ibR;  // 1/sqrt(R*Tc), 1/Tc, 1/(b*R)
  double imu, Pr_m1;       // 1/mu, Pr - 1
  // viscosity law: 0 constant mu; 1 VHS mu_ref (T/T_ref)^vexp; 2 Sutherland
  // mu_ref (T/T_ref)^1.5 (T_ref + vS)/(T + vS)
  int visc_model;
  double mu_ref, T_ref_v, iT_ref_v, vexp, vS;
  // non-uniform grid: dt/dx per column [nx] and dt/dy per row [ny], and
  // dx per column and dy per row (the wall balance); null for a uniform grid
  // (rdx, rdy, dx, dy above)
  const double* rdxs;
  const double* rdys;
  const double* dxs;
  const double* dys;
  // residual tracking of a served request with a tolerance (the reference
  // driver's stop rule, main.cpp:54-110): res_hist[k - 2] receives the
  // residual of global step k for 2 <= k <= 20, res_hist[19] that of global
  // step res_last (the last step of the current launch sequence); null
  // when not tracked. ptile0, nptiles: the instance's run in the group's
  // pointwise tile list.
  double* res_hist;
  int res_last;
  int ptile0, nptiles;
  // the group's strip list holds this instance: the tile flux kernels skip
  // its tiles, except the moment-summing stage-3 flux of FP32 storage
  int strips;
};

constexpr int kResHist = 20;  // residual history slots per tracked instance

// The final combination of RK3 stage 3, new state = f3 + (dF2 - dF3)/2.
// FP64 storage: the stage-3 kernel writes f3 only and the stage-3 flux
// kernel forms the sum from f3 and the stored dF2. FP32 storage: the
// stage-3 kernel also writes G = f3 + dF2/2, formed in FP64 from the
// unrounded f3, and the flux kernel forms G - dF3/2: the FP64 form, from
// the rounded f3, raised the FP32 path's relative L2 against FP64 on 64^2
// Couette decks after 5000 steps 4-9 times (rho 5e-9 -> 4.7e-8, u 1.1e-7
// -> 4.4e-7), for 26 fewer values written per cell and step.
template <typename TS>
constexpr bool kStage3G = sizeof(TS) == 4;

// One tile of one instance: the unit of work of a block.
// Strip work items of the strip flux kernels (TileRef with tx = strip index,
// ty = first row, flags = kTileSolid | kStripDesc | row count << 8):
// kStripW own columns (with a 3-column halo on each side, 64 lanes) by up to
// kStripH rows; kStripDesc: the item takes its rows from the last to the
// first.
constexpr int kStripW = 58, kStripH = 32;

struct TileRef {
  int inst;
  int tx, ty;
  int flags;           // kTileSolid: the tile's halo region holds a solid cell;
                       // kTilePhysical: an own cell needs the reference's physical addressing
};

enum : int { kTileSolid = 1, kTilePhysical = 2, kStripDesc = 4 };

// Wall cell that needs a closure factor (has a non-solid axis neighbour).
struct WallRef {
  int inst;
  int x, y;
};

}  // namespace uqlbm
