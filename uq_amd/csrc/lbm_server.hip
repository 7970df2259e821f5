// lbm_server.hip — continuous-batching request loop over the batched
// stepper: host-side queue, admission by device footprint and slots, device-
// resident state, batched rounds, retirement at each request's own step
// count. The reference steps a single cpuLBM from its main loop
// (main.cpp:21-24); this is the served form of that loop for many solves.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <vector>

#include "../../include/uqlbm.h"
#include "lbm_launch.hpp"
#include "lbm_runtime.hpp"

namespace uqlbm {

struct PendingReq {
  lbm_params params;
  int nx, ny;
  std::vector<int32_t> bnd;
  std::vector<double> dens, pres, temp, vx, vy;
  std::vector<int32_t> ctype;
  int nsteps;
  int64_t tag;
  double submit_s;
  size_t footprint;
};

struct ActiveReq {
  Solver* solver;
  int nsteps;              // step count; 0 with a tolerance: run until converged
  int64_t tag;
  double submit_s, admit_s;
  size_t footprint;
  double tol = 0;          // residual tolerance of the deck (0: none)
  double residual = 0;     // normalised residual at the last round end
  bool converged = false;
};

struct KeptMacro {
  int nx, ny;
  std::vector<double> rho, ux, uy, p, T;
};

struct Server {
  int precision = LBM_PREC_MIXED;
  int rk_method = 1;
  size_t budget = 0;
  int max_slots = 0;
  bool keep = false;
  std::deque<PendingReq> queue;
  std::vector<ActiveReq> active;
  std::vector<lbm_request_result> done;
  std::map<int64_t, KeptMacro> kept;
  size_t in_use = 0;
  long long lups = 0;
  hipStream_t stream = nullptr;
  InstDev* table_dev = nullptr;
  TileRef* tiles_dev = nullptr;
  WallRef* walls_dev = nullptr;
  TileRef* ptiles_dev = nullptr;
  TileRef* strips_dev = nullptr;
  size_t tiles_cap = 0, walls_cap = 0, ptiles_cap = 0, strips_cap = 0;
  int ntiles = 0, nwalls = 0, nptiles = 0;
  int nstrips = 0, nstrips_fast = 0, tiles_rest = 0;  // strip list of the active requests (StepLaunch)
  int nphys = 0;  // flux tiles flagged kTilePhysical
  int eq_kinds = 1;  // equilibrium precisions among the active requests (StepLaunch::eq_kinds)
  bool lists_dirty = true;
  int* step_ctr = nullptr;
  int* fails_dev = nullptr;  // failure steps of the active requests, gathered for one host copy
  double* res_partial = nullptr;  // residual partials per pointwise tile (ptiles_cap doubles)
  int flux_shape = 0;
  GraphCache graphs;  // step graphs of the current active set (recaptured when it changes)
  ~Server();
};

static double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

static size_t footprint_of(int nx, int ny, int precision) { return solver_footprint(nx, ny, precision); }

Server::~Server() {
  graphs.clear();
  for (auto& a : active) delete a.solver;
  for (void* p : {(void*)table_dev, (void*)tiles_dev, (void*)walls_dev, (void*)ptiles_dev, (void*)step_ctr,
                  (void*)fails_dev, (void*)res_partial, (void*)strips_dev})
    if (p) (void)hipFree(p);
  if (stream) (void)hipStreamDestroy(stream);
}

#define SV_TRY(expr)                                                                            \
  do {                                                                                          \
    hipError_t e_ = (expr);                                                                     \
    if (e_ != hipSuccess) return fail_with(LBM_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

static int rebuild_lists(Server& sv) {
  sv.flux_shape = default_flux_shape(sv.precision);
  std::vector<TileRef> tiles;
  std::vector<WallRef> walls;
  std::vector<TileRef> ptiles;
  for (size_t k = 0; k < sv.active.size(); k++) {
    for (TileRef t : solver_tiles(*sv.active[k].solver, sv.flux_shape)) {
      t.inst = (int)k;
      tiles.push_back(t);
    }
    for (TileRef t : sv.active[k].solver->ptiles_host) {
      t.inst = (int)k;
      ptiles.push_back(t);
    }
    for (WallRef w : sv.active[k].solver->walls_host) {
      w.inst = (int)k;
      walls.push_back(w);
    }
  }
  if (tiles.size() > sv.tiles_cap) {
    if (sv.tiles_dev) SV_TRY(hipFree(sv.tiles_dev));
    sv.tiles_cap = std::max(tiles.size(), sv.tiles_cap * 2);
    SV_TRY(hipMalloc((void**)&sv.tiles_dev, sv.tiles_cap * sizeof(TileRef)));
  }
  if (walls.size() > sv.walls_cap) {
    if (sv.walls_dev) SV_TRY(hipFree(sv.walls_dev));
    sv.walls_cap = std::max(walls.size(), sv.walls_cap * 2);
    SV_TRY(hipMalloc((void**)&sv.walls_dev, sv.walls_cap * sizeof(WallRef)));
  }
  std::vector<Solver*> members;
  members.reserve(sv.active.size());
  for (auto& a : sv.active) members.push_back(a.solver);
  std::vector<TileRef> strips;
  sv.nstrips_fast = group_strips(members, sv.flux_shape, strips, &sv.tiles_rest);
  sv.nstrips = (int)strips.size();
  if (strips.size() > sv.strips_cap) {
    if (sv.strips_dev) SV_TRY(hipFree(sv.strips_dev));
    sv.strips_cap = std::max(strips.size(), sv.strips_cap * 2);
    SV_TRY(hipMalloc((void**)&sv.strips_dev, sv.strips_cap * sizeof(TileRef)));
  }
  if (!strips.empty())
    SV_TRY(hipMemcpyAsync(sv.strips_dev, strips.data(), strips.size() * sizeof(TileRef), hipMemcpyHostToDevice,
                          sv.stream));
  if (ptiles.size() > sv.ptiles_cap) {
    if (sv.ptiles_dev) SV_TRY(hipFree(sv.ptiles_dev));
    if (sv.res_partial) SV_TRY(hipFree(sv.res_partial));
    sv.res_partial = nullptr;
    sv.ptiles_cap = std::max(ptiles.size(), sv.ptiles_cap * 2);
    SV_TRY(hipMalloc((void**)&sv.ptiles_dev, sv.ptiles_cap * sizeof(TileRef)));
    SV_TRY(hipMalloc((void**)&sv.res_partial, sv.ptiles_cap * sizeof(double)));
  }
  if (!ptiles.empty())
    SV_TRY(hipMemcpyAsync(sv.ptiles_dev, ptiles.data(), ptiles.size() * sizeof(TileRef), hipMemcpyHostToDevice,
                          sv.stream));
  if (!tiles.empty())
    SV_TRY(hipMemcpyAsync(sv.tiles_dev, tiles.data(), tiles.size() * sizeof(TileRef), hipMemcpyHostToDevice,
                          sv.stream));
  if (!walls.empty())
    SV_TRY(hipMemcpyAsync(sv.walls_dev, walls.data(), walls.size() * sizeof(WallRef), hipMemcpyHostToDevice,
                          sv.stream));
  SV_TRY(hipStreamSynchronize(sv.stream));
  sv.ntiles = (int)tiles.size();
  sv.nphys = count_physical(tiles);
  sv.eq_kinds = eq_kinds_of(sv.precision == LBM_PREC_FP64, sv.active.begin(), sv.active.end(),
                            [](const ActiveReq& a) { return a.solver->eq32; });
  sv.nwalls = (int)walls.size();
  sv.nptiles = (int)ptiles.size();
  sv.lists_dirty = false;
  return LBM_OK;
}

static int admit(Server& sv) {
  while (!sv.queue.empty() && (int)sv.active.size() < sv.max_slots) {
    PendingReq& r = sv.queue.front();
    if (sv.in_use + r.footprint > sv.budget && !sv.active.empty()) break;  // wait for memory
    lbm_props pr{(int32_t)r.ctype.size(), r.dens.data(), r.pres.data(), r.temp.data(),
                 r.vx.data(),             r.vy.data(),   r.ctype.data()};
    lbm_solver* h = nullptr;
    int st = lbm_create(&r.params, r.nx, r.ny, r.bnd.data(), &pr, sv.precision, &h);
    if (st == LBM_OK) st = lbm_init(h);
    const double t = now_s();
    if (st != LBM_OK) {
      if (h) lbm_destroy(h);
      lbm_request_result res{};
      res.tag = r.tag;
      res.status = st;
      res.submit_s = r.submit_s;
      res.admit_s = t;
      res.finish_s = t;
      res.cells = (int64_t)r.nx * r.ny;
      sv.done.push_back(res);
      sv.queue.pop_front();
      continue;
    }
    Solver* s = reinterpret_cast<Solver*>(h);
    ActiveReq a{s, r.nsteps, r.tag, r.submit_s, t, s->device_bytes};
    a.tol = r.params.tol > 0 ? r.params.tol : 0.0;
    if (a.tol > 0) {
      SV_TRY(solver_alloc(*s, (void**)&s->res_hist, kResHist * sizeof(double)));
      SV_TRY(hipMemset(s->res_hist, 0, kResHist * sizeof(double)));
    }
    sv.active.push_back(a);
    sv.in_use += s->device_bytes;
    sv.queue.pop_front();
    sv.lists_dirty = true;
  }
  return LBM_OK;
}

static int retire(Server& sv) {
  std::vector<ActiveReq> keep;
  for (auto& a : sv.active) {
    Solver* s = a.solver;
    const bool reached = a.nsteps > 0 && s->step >= a.nsteps;
    if (!reached && !a.converged && !s->failed) {
      keep.push_back(a);
      continue;
    }
    lbm_request_result res{};
    res.tag = a.tag;
    res.status = s->failed ? LBM_FAILED : LBM_OK;
    res.steps_done = s->step;
    res.submit_s = a.submit_s;
    res.admit_s = a.admit_s;
    res.finish_s = now_s();
    res.cells = (int64_t)s->nx * s->ny;
    res.residual = a.residual;
    if (sv.keep) {
      // the solver's own instance record still describes admission time
      int su = upload_inst(*s);
      if (su != LBM_OK) return su;
      KeptMacro km;
      km.nx = s->nx;
      km.ny = s->ny;
      const size_t n = (size_t)s->nx * s->ny;
      km.rho.resize(n);
      km.ux.resize(n);
      km.uy.resize(n);
      km.p.resize(n);
      km.T.resize(n);
      int st = lbm_get_macro(reinterpret_cast<lbm_solver*>(s), km.rho.data(), km.ux.data(), km.uy.data(),
                             km.p.data(), km.T.data());
      if (st == LBM_OK) sv.kept[a.tag] = std::move(km);
    }
    sv.done.push_back(res);
    sv.in_use -= a.footprint;
    delete s;
    sv.lists_dirty = true;
  }
  sv.active.swap(keep);
  return LBM_OK;
}

// Tolerance stop of the reference driver (main.cpp:54-110): the residual of
// steps 2..20 sets the normalisation (their maximum); from step 21 on the
// solve stops once residual / normalisation <= tol. A served request is
// tested at the end of each round (its last step's residual), so it stops
// at the first round end that meets the rule.
static int check_convergence(Server& sv) {
  std::vector<std::vector<double>> hist(sv.active.size());
  for (size_t k = 0; k < sv.active.size(); k++) {
    if (!(sv.active[k].tol > 0)) continue;
    hist[k].resize(kResHist);
    SV_TRY(hipMemcpyAsync(hist[k].data(), sv.active[k].solver->res_hist, kResHist * sizeof(double),
                          hipMemcpyDeviceToHost, sv.stream));
  }
  SV_TRY(hipStreamSynchronize(sv.stream));
  const int numRes = kResHist;
  for (size_t k = 0; k < sv.active.size(); k++) {
    ActiveReq& a = sv.active[k];
    if (!(a.tol > 0) || a.solver->failed || a.solver->step < numRes + 1) continue;
    double norm = 0;
    for (int i = 0; i < numRes - 1; i++) norm = std::max(norm, hist[k][i]);
    a.residual = norm > 0 ? hist[k][numRes - 1] / norm : hist[k][numRes - 1];
    if (a.residual <= a.tol) a.converged = true;
  }
  return LBM_OK;
}

}  // namespace uqlbm

using namespace uqlbm;

extern "C" int lbm_server_create(int32_t precision, int32_t rk_method, size_t hbm_budget_bytes, int32_t max_slots,
                                 lbm_server** out) {
  if (!out || max_slots < 1 || hbm_budget_bytes == 0) return fail_with(LBM_ERR_ARG, "bad server arguments");
  if (precision != LBM_PREC_FP64 && precision != LBM_PREC_MIXED)
    return fail_with(LBM_ERR_ARG, "unknown precision %d", precision);
  if (rk_method < 0 || rk_method > 3) return fail_with(LBM_ERR_ARG, "RKMETHOD %d unknown", rk_method);
  if (lbm_device_count() < 1) return fail_with(LBM_ERR_NOGPU, "no gfx950 device visible");
  std::unique_ptr<Server> sv(new Server());
  sv->precision = precision;
  sv->rk_method = rk_method;
  sv->budget = hbm_budget_bytes;
  sv->max_slots = max_slots;
  SV_TRY(hipStreamCreateWithFlags(&sv->stream, hipStreamNonBlocking));
  SV_TRY(hipMalloc((void**)&sv->table_dev, (size_t)max_slots * sizeof(InstDev)));
  SV_TRY(hipMalloc((void**)&sv->step_ctr, sizeof(int)));
  SV_TRY(hipMalloc((void**)&sv->fails_dev, (size_t)max_slots * sizeof(int)));
  *out = reinterpret_cast<lbm_server*>(sv.release());
  return LBM_OK;
}

extern "C" int lbm_server_submit(lbm_server* h, const lbm_params* params, int32_t nx, int32_t ny,
                                 const int32_t* bnd, const lbm_props* props, int32_t nsteps, int64_t tag) {
  if (!h || !params || !bnd || !props || nx < 4 || ny < 4 || nsteps < 0 || props->num_props < 1)
    return fail_with(LBM_ERR_ARG, "bad request");
  if (!props->density || !props->pressure || !props->temp || !props->velocity_x || !props->velocity_y ||
      !props->cell_type)
    return fail_with(LBM_ERR_ARG, "incomplete property table");
  if (nsteps == 0 && !(params->tol > 0))
    return fail_with(LBM_ERR_ARG, "a request needs a step count or a residual tolerance");
  Server& sv = *reinterpret_cast<Server*>(h);
  if (params->rk_method != sv.rk_method)
    return fail_with(LBM_ERR_ARG, "request RKMETHOD %d differs from the server's %d", params->rk_method, sv.rk_method);
  PendingReq r;
  r.params = *params;
  r.nx = nx;
  r.ny = ny;
  r.bnd.assign(bnd, bnd + (size_t)nx * ny);
  const int n = props->num_props;
  r.dens.assign(props->density, props->density + n);
  r.pres.assign(props->pressure, props->pressure + n);
  r.temp.assign(props->temp, props->temp + n);
  r.vx.assign(props->velocity_x, props->velocity_x + n);
  r.vy.assign(props->velocity_y, props->velocity_y + n);
  r.ctype.assign(props->cell_type, props->cell_type + n);
  r.nsteps = nsteps;
  r.tag = tag;
  r.submit_s = now_s();
  r.footprint = footprint_of(nx, ny, sv.precision);
  if (r.footprint > sv.budget) return fail_with(LBM_ERR_NOMEM, "request needs %zu bytes, budget %zu", r.footprint, sv.budget);
  sv.queue.push_back(std::move(r));
  return LBM_OK;
}

extern "C" int lbm_server_tick(lbm_server* h, int32_t tick_steps, int32_t* active_out, int32_t* queued_out) {
  if (!h || tick_steps < 1) return fail_with(LBM_ERR_ARG, "bad tick arguments");
  Server& sv = *reinterpret_cast<Server*>(h);
  int status = LBM_OK;
  int left = tick_steps;
  int st = admit(sv);
  if (st != LBM_OK) return st;
  st = retire(sv);  // zero-step requests
  if (st != LBM_OK) return st;
  while (left > 0 && !sv.active.empty()) {
    int round = left;
    for (auto& a : sv.active)
      if (a.nsteps > 0) round = std::min(round, a.nsteps - a.solver->step);
    if (round < 1) round = 1;
    bool track = false;
    for (auto& a : sv.active) track |= a.tol > 0;
    if (sv.lists_dirty) {
      st = rebuild_lists(sv);
      if (st != LBM_OK) return st;
    }
    std::vector<Solver*> members;
    members.reserve(sv.active.size());
    for (auto& a : sv.active) members.push_back(a.solver);
    StepLaunch L{sv.table_dev,  sv.tiles_dev, sv.ntiles,     sv.walls_dev, sv.nwalls,
                 sv.ptiles_dev, sv.nptiles,   sv.rk_method, sv.step_ctr,  sv.flux_shape};
    L.nphys = sv.nphys;
    L.eq_kinds = sv.eq_kinds;
    if (strips_for(sv.nstrips)) {
      L.strips = sv.strips_dev;
      L.nstrips = sv.nstrips;
      L.nstrips_fast = sv.nstrips_fast;
      L.tiles_rest = sv.tiles_rest;
    } else {
      L.tiles_rest = sv.ntiles;
    }
    if (track) {
      L.res_partial = sv.res_partial;
      L.ninst = (int)members.size();
    }
    long long cells = 0;
    for (Solver* s : members) cells += (long long)s->nx * s->ny;
    st = run_group(members, sv.table_dev, L, sv.precision, sv.stream, round, &sv.graphs, sv.fails_dev, true);
    if (st < 0) return st;
    if (st == LBM_FAILED) status = LBM_FAILED;
    sv.lups += cells * round;
    left -= round;
    if (track) {
      st = check_convergence(sv);
      if (st != LBM_OK) return st;
    }
    st = retire(sv);
    if (st != LBM_OK) return st;
    st = admit(sv);
    if (st != LBM_OK) return st;
  }
  if (active_out) *active_out = (int32_t)sv.active.size();
  if (queued_out) *queued_out = (int32_t)sv.queue.size();
  return status;
}

extern "C" int lbm_server_poll(lbm_server* h, lbm_request_result* out, int32_t max) {
  if (!h || (!out && max > 0) || max < 0) return fail_with(LBM_ERR_ARG, "bad poll arguments");
  Server& sv = *reinterpret_cast<Server*>(h);
  const int n = std::min<int>(max, (int)sv.done.size());
  for (int k = 0; k < n; k++) out[k] = sv.done[k];
  sv.done.erase(sv.done.begin(), sv.done.begin() + n);
  return n;
}

extern "C" int lbm_server_keep_results(lbm_server* h, int32_t enable) {
  if (!h) return fail_with(LBM_ERR_ARG, "null server");
  reinterpret_cast<Server*>(h)->keep = enable != 0;
  return LBM_OK;
}

extern "C" int lbm_server_result_macro(lbm_server* h, int64_t tag, double* rho, double* ux, double* uy, double* p,
                                       double* T) {
  if (!h) return fail_with(LBM_ERR_ARG, "null server");
  Server& sv = *reinterpret_cast<Server*>(h);
  auto it = sv.kept.find(tag);
  if (it == sv.kept.end()) return fail_with(LBM_ERR_ARG, "no kept result for tag %lld", (long long)tag);
  const KeptMacro& k = it->second;
  const size_t bytes = (size_t)k.nx * k.ny * sizeof(double);
  if (rho) std::memcpy(rho, k.rho.data(), bytes);
  if (ux) std::memcpy(ux, k.ux.data(), bytes);
  if (uy) std::memcpy(uy, k.uy.data(), bytes);
  if (p) std::memcpy(p, k.p.data(), bytes);
  if (T) std::memcpy(T, k.T.data(), bytes);
  return LBM_OK;
}

extern "C" int lbm_server_stats(lbm_server* h, int64_t* lups_done, size_t* hbm_in_use) {
  if (!h) return fail_with(LBM_ERR_ARG, "null server");
  Server& sv = *reinterpret_cast<Server*>(h);
  if (lups_done) *lups_done = sv.lups;
  if (hbm_in_use) *hbm_in_use = sv.in_use;
  return LBM_OK;
}

extern "C" void lbm_server_destroy(lbm_server* h) { delete reinterpret_cast<Server*>(h); }
