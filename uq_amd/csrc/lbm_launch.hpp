// lbm_launch.hpp — host-side launch interface between the runtime and the
// kernels of lbm_kernels.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "lbm_common.hpp"

namespace uqlbm {

// Device pointers of one batched step: the instance table and the work lists.
struct StepLaunch {
  const InstDev* insts = nullptr;
  const TileRef* tiles = nullptr;
  int ntiles = 0;
  const WallRef* walls = nullptr;
  int nwalls = 0;
  const TileRef* ptiles = nullptr;  // 64 x 4 row tiles of the pointwise kernels
  int nptiles = 0;
  int rk_method = 1;                // RKMETHOD shared by every instance of the group
  int* step_ctr = nullptr;          // device step counter: step k of a sequence reads k + 1
  int flux_shape = 0;               // flux tile shape (flux_tile_dims); the tile list is built for it
  double* res_partial = nullptr;    // per pointwise tile residual partials (nptiles doubles) when
                                    // some instance tracks its residual; null otherwise
  int ninst = 0;                    // instances in the table (for the residual tracking)
  int nphys = 0;                    // tiles flagged kTilePhysical in `tiles` (k_phys runs when > 0)
  int eq_kinds = 1;                 // equilibrium precisions among the instances: bit 0 FP64,
                                    // bit 1 FP32 (deviation form; FP32 storage only)
  const TileRef* strips = nullptr;  // strip work items (instances with InstDev::strips set):
  int nstrips = 0;                  // items without a solid cell in their region first
  int nstrips_fast = 0;             // (the first nstrips_fast), then those with one
  int tiles_rest = 0;               // flux tiles of the instances not on strips
};

// Flux tiles of a list flagged kTilePhysical.
inline int count_physical(const std::vector<TileRef>& tiles) {
  int n = 0;
  for (const TileRef& t : tiles) n += (t.flags & kTilePhysical) ? 1 : 0;
  return n;
}

// Equilibrium precisions among a group's instances (StepLaunch::eq_kinds):
// FP64 storage always FP64; FP32 storage per instance.
template <typename It, typename Get>
inline int eq_kinds_of(int fp64, It begin, It end, Get solver_eq32) {
  if (fp64) return 1;
  int k = 0;
  for (It it = begin; it != end; ++it) k |= solver_eq32(*it) ? 2 : 1;
  return k ? k : 1;
}

inline bool same_launch(const StepLaunch& a, const StepLaunch& b) {
  return a.insts == b.insts && a.tiles == b.tiles && a.ntiles == b.ntiles && a.walls == b.walls &&
         a.nwalls == b.nwalls && a.ptiles == b.ptiles && a.nptiles == b.nptiles && a.rk_method == b.rk_method &&
         a.step_ctr == b.step_ctr && a.flux_shape == b.flux_shape && a.res_partial == b.res_partial && a.ninst == b.ninst &&
         a.nphys == b.nphys && a.eq_kinds == b.eq_kinds && a.strips == b.strips && a.nstrips == b.nstrips &&
         a.nstrips_fast == b.nstrips_fast && a.tiles_rest == b.tiles_rest;
}

// Optional per-kernel timing with HIP events recorded on the launch stream
// around each launch; elapsed times are collected after a synchronise.
class KernelTimer {
 public:
  static constexpr int kSlots = 8;
  void set_names(const std::vector<std::string>& names) { names_ = names; }
  void begin(int k, hipStream_t s) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
    (void)hipEventRecord(a, s);
    open_[k] = std::make_pair(a, b);
  }
  void end(int k, hipStream_t s) {
    auto ev = open_[k];
    (void)hipEventRecord(ev.second, s);
    pending_.push_back({k, ev});
  }
  // Collect elapsed times of all recorded pairs (call after synchronising).
  void collect() {
    for (auto& p : pending_) {
      float ms = 0;
      if (hipEventElapsedTime(&ms, p.second.first, p.second.second) == hipSuccess) {
        total_ms_[p.first] += ms;
        count_[p.first] += 1;
      }
      (void)hipEventDestroy(p.second.first);
      (void)hipEventDestroy(p.second.second);
    }
    pending_.clear();
  }
  void reset() {
    collect();
    for (int k = 0; k < kSlots; k++) { total_ms_[k] = 0; count_[k] = 0; }
  }
  double avg_ms(int k) const { return count_[k] ? total_ms_[k] / count_[k] : 0.0; }
  long long count(int k) const { return count_[k]; }
  const std::vector<std::string>& names() const { return names_; }

 private:
  std::vector<std::string> names_;
  std::pair<hipEvent_t, hipEvent_t> open_[kSlots] = {};
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending_;
  double total_ms_[kSlots] = {};
  long long count_[kSlots] = {};
};

// Flux tile shapes: 0 = 15 x 16 own cells (256 threads), 1 = 31 x 15 (512),
// 2 = 15 x 32 (512), 3 = 39 x 12 (512), 4 = 23 x 20 (512).
constexpr int kFluxShapes = 5;
void flux_tile_dims(int shape, int* w, int* h);
int tile_halo();
int dflux_planes();
constexpr int kPointW = 64, kPointH = 4;  // pointwise tile: 64 x 4 cells, 256 threads

// One step of every instance of L (buffer parity `par` relative to each
// instance's parity0), then the step counter advance.
template <typename TS, typename TC>
void launch_step(const StepLaunch& L, hipStream_t s, int par, KernelTimer* tm);

// Computes the persistent grid sizes of the step kernels (call before a
// stream capture).
void prepare_launch(int fp64);

template <typename TS>
void launch_init(const InstDev* insts, int inst, const int* ids, const double* props, long long cells,
                 hipStream_t s);

template <typename TS>
void launch_macro(const InstDev* insts, int inst, int par, double* out, long long cells, hipStream_t s);

// Conserved totals (mass, x and y momentum, total energy) of the fluid cells
// of one instance; `scratch` holds totals_scratch_doubles() doubles and the
// four results land in its last four.
template <typename TS>
void launch_totals(const InstDev* insts, int inst, int par, double* scratch, long long cells, hipStream_t s);
int totals_scratch_doubles();

// Copies *insts[k].fail into out[k] for the n members of a launch group.
void launch_gather_fail(const InstDev* insts, int n, int* out, hipStream_t s);

// Residual of the last step of one instance (RMS change of T over all
// cells between the other state buffer and state `par`): rows holds ny + 1
// doubles, per-row sums then the residual in rows[ny].
template <typename TS>
void launch_residual(const InstDev* insts, int inst, int par, double* rows, int nx, int ny, hipStream_t s);

}  // namespace uqlbm
