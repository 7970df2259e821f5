// lbm_runtime.hpp — host-side objects behind the opaque handles of
// include/uqlbm.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <vector>

#include "../../include/uqlbm.h"
#include "lbm_common.hpp"
#include "lbm_launch.hpp"

namespace uqlbm {

// Instantiated HIP graphs of one launch group's step sequence: two steps
// (buffer parities 0 and 1, so that it replays back to back) and one step.
// Valid while the group's launch description is unchanged.
struct GraphCache {
  StepLaunch key{};
  int precision = -1;
  hipGraphExec_t two = nullptr, one = nullptr;
  GraphCache() = default;
  GraphCache(const GraphCache&) = delete;
  GraphCache& operator=(const GraphCache&) = delete;
  void clear() {
    if (two) (void)hipGraphExecDestroy(two);
    if (one) (void)hipGraphExecDestroy(one);
    two = one = nullptr;
    precision = -1;
  }
  ~GraphCache() { clear(); }
};

struct PropRow {
  double rho, p, T, ux, uy;
  int type;
};

// One solver instance: the deck, its derived constants and its device state.
struct Solver {
  lbm_params p{};
  int nx = 0, ny = 0, pitch = 0;
  int precision = LBM_PREC_FP64;
  std::vector<PropRow> props;
  double b = 0, K = 0, Tc = 0, c = 0, dt = 0;
  int nsteps = 0;

  std::vector<int> ids_host;        // [y][nx]
  std::vector<uint8_t> ctype_host;  // [y][nx]
  std::vector<char> phys_cells;     // [y][nx]: net fluxes need the reference addressing
  bool any_phys = false;
  std::vector<TileRef> tiles_shape[kFluxShapes];  // flux tiles per shape, built on first use
  std::vector<WallRef> walls_host;
  std::vector<TileRef> ptiles_host;
  std::vector<TileRef> strips_fast, strips_solid;  // strip items (none when a cell needs physical addressing)

  hipStream_t stream = nullptr;
  void* state[2] = {nullptr, nullptr};
  void* f2 = nullptr;
  void* f3 = nullptr;
  void* dflux = nullptr;
  double* mom[2] = {nullptr, nullptr};  // FP32 storage: conserved moments of state[k] (4 planes)
  double ref[kPlanes] = {};             // FP32 storage: reference equilibrium (deviation origin)
  double ref_rho = 0, ref_t = 0;        // its density and T/Tc
  bool eq32 = false;                    // FP32 storage near the reference: FP32 equilibria
  uint8_t* ctype = nullptr;
  double* alpha = nullptr;
  int* fail = nullptr;
  int* ids = nullptr;
  double* props_dev = nullptr;
  InstDev* inst_dev = nullptr;
  InstDev inst_host{};              // host image of *inst_dev (kept alive for async copies)
  TileRef* tiles_dev = nullptr;
  WallRef* walls_dev = nullptr;
  TileRef* ptiles_dev = nullptr;
  TileRef* strips_dev = nullptr;    // strips_fast then strips_solid
  double* macro_dev = nullptr;
  int* step_ctr = nullptr;          // device step counter of this solver's launches
  int flux_shape = 0;               // flux tile shape of tiles_dev
  int visc_model = 0;               // viscosity law (lbm_set_viscosity)
  double mu_ref = 0, T_ref_v = 0, visc_param = 0;
  std::vector<double> dxs, dys;     // non-uniform spacing (lbm_set_spacing), empty when uniform
  double* spacing_dev = nullptr;    // dt/dx[nx], dt/dy[ny], dx[nx], dy[ny]
  double* totals_dev = nullptr;     // conserved-totals reduction scratch (allocated on first use)
  double* resid_dev = nullptr;      // residual row sums, ny + 1 doubles (allocated on first use)
  double* res_hist = nullptr;       // served with a tolerance: per-step residual history (kResHist)
  int* fails_dev = nullptr;         // failure step gathered for the host (one int)
  std::unique_ptr<GraphCache> graphs{new GraphCache()};
  std::vector<std::pair<void*, size_t>> owned;  // buffers from solver_alloc, returned to its cache
  int ntiles = 0, nwalls = 0, nptiles = 0;
  int nphys = 0;  // flux tiles flagged kTilePhysical
  int nstrips = 0, nstrips_fast = 0;
  size_t device_bytes = 0;

  int cur = 0;    // state buffer holding the current state
  int step = 0;   // steps taken
  bool inited = false;
  bool failed = false;
  int batch_refs = 0;            // batches holding this solver
  bool destroy_pending = false;  // lbm_destroy called while batch_refs > 0

  ~Solver();
};

// A set of solvers advanced together by batched launches.
struct Batch {
  std::vector<Solver*> members;
  int precision = LBM_PREC_FP64;
  int rk_method = 1;
  hipStream_t stream = nullptr;
  InstDev* inst_dev = nullptr;
  InstDev inst_host{};              // host image of *inst_dev (kept alive for async copies)
  TileRef* tiles_dev = nullptr;
  WallRef* walls_dev = nullptr;
  TileRef* ptiles_dev = nullptr;
  TileRef* strips_dev = nullptr;
  int* step_ctr = nullptr;
  int* fails_dev = nullptr;         // members' failure steps gathered for one host copy
  int flux_shape = 0;
  std::unique_ptr<GraphCache> graphs{new GraphCache()};
  int ntiles = 0, nwalls = 0, nptiles = 0;
  int nphys = 0;  // flux tiles flagged kTilePhysical
  int eq_kinds = 1;  // equilibrium precisions among the members (StepLaunch::eq_kinds)
  int nstrips = 0, nstrips_fast = 0, tiles_rest = 0;
  Batch() = default;
  Batch(Batch&& o) noexcept { *this = std::move(o); }
  Batch& operator=(Batch&& o) noexcept {
    members = std::move(o.members);
    precision = o.precision;
    rk_method = o.rk_method;
    stream = o.stream;
    inst_dev = o.inst_dev;
    tiles_dev = o.tiles_dev;
    walls_dev = o.walls_dev;
    ptiles_dev = o.ptiles_dev;
    strips_dev = o.strips_dev;
    step_ctr = o.step_ctr;
    fails_dev = o.fails_dev;
    flux_shape = o.flux_shape;
    graphs = std::move(o.graphs);
    inst_host = o.inst_host;
    ntiles = o.ntiles;
    nwalls = o.nwalls;
    nptiles = o.nptiles;
    nphys = o.nphys;
    eq_kinds = o.eq_kinds;
    nstrips = o.nstrips;
    nstrips_fast = o.nstrips_fast;
    tiles_rest = o.tiles_rest;
    o.stream = nullptr;
    o.inst_dev = nullptr;
    o.tiles_dev = nullptr;
    o.walls_dev = nullptr;
    o.ptiles_dev = nullptr;
    o.strips_dev = nullptr;
    o.step_ctr = nullptr;
    o.fails_dev = nullptr;
    return *this;
  }
  ~Batch();
};

int fail_with(int status, const char* fmt, ...);
// A device buffer of `bytes` for solver s: a kept buffer of that size from a
// destroyed solver on this device when there is one (contents undefined),
// else hipMalloc; ~Solver returns it.
hipError_t solver_alloc(Solver& s, void** p, size_t bytes);
// Strip flux kernels in use (lbm_set_strips; default on), and whether a
// launch group with `nstrips` strip items takes them.
bool strips_enabled();
bool strips_for(int nstrips);
// Strip list of a launch group (member k's items with instance index k,
// items without solids first; returns how many those are) and the number
// of flux tiles of shape `shape` of the members not on strips.
int group_strips(const std::vector<Solver*>& members, int shape, std::vector<TileRef>& out, int* tiles_rest);
const std::vector<TileRef>& solver_tiles(Solver& s, int shape);
int default_flux_shape(int precision);
int batch_set_shape(Batch& b, int shape);
size_t solver_footprint(int nx, int ny, int precision);
int upload_inst(Solver& s);
int batch_finish(Batch& b, lbm_batch** out);
int batch_step(Batch& b, int nsteps);
int run_group(const std::vector<Solver*>& members, InstDev* table_dev, const StepLaunch& L, int precision,
              hipStream_t stream, int n, GraphCache* graphs, int* fails_dev, bool timed);

}  // namespace uqlbm
