// gpuLBM.hpp — header-only C++ drop-in for the reference's cpuLBM class
// (reference CODE/C++/VS2010/simLB/simLB/cpuLBM.h:16-176), running the step
// on an MI355X through the C ABI of uqlbm.h.
//
// The public members keep the reference's names and meaning: the deck
// scalars, the geometry (Nx, Ny, bndArray, property table), the run-time
// counters (nsteps, step, tstep, residual, initRes, resNorm, stop, pause)
// and the macroscopic and distribution fields (rho_G, u_G, p_G, T_G, f_G,
// h_G) in the reference's [x][y] / [x][y][i] layouts. initSim() replaces
// refineInput + initSim (cpuLBM.cpp:283-462) and runOne(res) replaces
// runOne (:2284-2335), returning the failure flag. After each runOne the
// public fields are refreshed from the device unless syncFields is false
// (then call syncFromDevice() when they are needed).
#ifndef GPULBM_HPP
#define GPULBM_HPP

#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "uqlbm.h"

struct gpuDouble2 {
  double x, y;
};

class gpuLBM {
 public:
  // deck scalars (cpuLBM.h:38-57)
  int Nx = 0, Ny = 0;
  std::vector<int> bndArray;  // [x][y], already y-flipped
  int RKMETHOD = 1, FMETHOD = 1;
  double CFL = 0, dtau = 0;
  int PROCESSOR = 1;
  double tt_tref = 0, tt_time = 0;
  int steps = 0;
  double tol = 0;
  int periodicX = 0, periodicY = 0, mirrorNorth = 0, mirrorSouth = 0, mirrorEast = 0, mirrorWest = 0;
  double gamma = 1.4, R = 287.0, mu = 0, Pr = 0.71;
  double dx = 0, dy = 0, Lx = 0, Ly = 0;
  double rho_ref = 0, p_ref = 0, T_ref = 0, Tc_Tref = 1;
  int nPrintOut = 100;
  int precision = LBM_PREC_FP64;

  // property table (cpuLBM.h:101-108)
  int numProps = 0;
  std::vector<double> density, pressure, temp, velocityX, velocityY;
  std::vector<int> cellType;

  // switches and run-time values (cpuLBM.h:64-81)
  int pause = 0, stop = 0;
  int nsteps = 0, step = 0;
  double tstep = 0, residual = 1;
  double initRes[100] = {};
  double resNorm = 0;
  double dt = 0, b = 0, K = 0, Tc = 0;

  // fields (cpuLBM.h:90-96)
  std::vector<double> f_G, h_G, rho_G, p_G, T_G;
  std::vector<gpuDouble2> u_G;
  bool syncFields = true;

  gpuLBM() = default;
  gpuLBM(const gpuLBM&) = delete;
  gpuLBM& operator=(const gpuLBM&) = delete;
  ~gpuLBM() {
    if (solver_) lbm_destroy(solver_);
  }

  // refineInput + setupVars + initialiseFunctions on the device.
  void initSim() {
    lbm_params p{};
    p.rk_method = RKMETHOD;
    p.flux_method = FMETHOD;
    p.cfl = CFL;
    p.dtau = dtau;
    p.tt_tref = tt_tref;
    p.tt_time = tt_time;
    p.steps = steps;
    p.tol = tol;
    p.periodic_x = periodicX;
    p.periodic_y = periodicY;
    p.mirror_north = mirrorNorth;
    p.mirror_south = mirrorSouth;
    p.mirror_east = mirrorEast;
    p.mirror_west = mirrorWest;
    p.gamma = gamma;
    p.R = R;
    p.mu = mu;
    p.Pr = Pr;
    p.dx = dx;
    p.dy = dy;
    p.Lx = Lx;
    p.Ly = Ly;
    p.rho_ref = rho_ref;
    p.p_ref = p_ref;
    p.T_ref = T_ref;
    p.Tc_Tref = Tc_Tref;
    numProps = (int)cellType.size();
    lbm_props pr{numProps,         density.data(),   pressure.data(), temp.data(),
                 velocityX.data(), velocityY.data(), cellType.data()};
    if (solver_) lbm_destroy(solver_);
    solver_ = nullptr;
    check(lbm_create(&p, Nx, Ny, bndArray.data(), &pr, precision, &solver_), "lbm_create");
    check(lbm_init(solver_), "lbm_init");
    lbm_info info{};
    check(lbm_get_info(solver_, &info), "lbm_get_info");
    dt = info.dt;
    nsteps = info.nsteps;
    b = info.b;
    K = info.K;
    Tc = info.Tc;
    dx = info.dx;
    dy = info.dy;
    step = 0;
    tstep = 0;
    syncFromDevice();
  }

  // cpuLBM::updateStepCount (cpuLBM.cpp:771-775)
  void updateStepCount() {
    step += 1;
    tstep = step * dt;
  }

  // cpuLBM::runOne (cpuLBM.cpp:2284-2335): one step; returns 1 on failure.
  int runOne(int res = 0) {
    double r = 0;
    const int st = lbm_step(solver_, 1, res > 0, &r);
    if (st < 0) check(st, "lbm_step");
    if (res > 0) residual = r;
    if (syncFields) syncFromDevice();
    return st == LBM_FAILED ? 1 : 0;
  }

  void syncFromDevice() {
    const size_t n = (size_t)Nx * Ny;
    rho_G.resize(n);
    p_G.resize(n);
    T_G.resize(n);
    u_G.resize(n);
    f_G.resize(n * LBM_Q);
    h_G.resize(n * LBM_Q);
    std::vector<double> ux(n), uy(n);
    check(lbm_get_macro(solver_, rho_G.data(), ux.data(), uy.data(), p_G.data(), T_G.data()), "lbm_get_macro");
    for (size_t k = 0; k < n; k++) u_G[k] = {ux[k], uy[k]};
    check(lbm_get_dist(solver_, f_G.data(), h_G.data()), "lbm_get_dist");
  }

  lbm_solver* handle() { return solver_; }

 private:
  lbm_solver* solver_ = nullptr;
  static void check(int st, const char* where) {
    if (st < 0) throw std::runtime_error(std::string(where) + ": " + lbm_last_error());
  }
};

// The reference driver's step (main.cpp:32-112) for a gpuLBM: residual
// normalisation over steps 2..numRes, the failure stop (stop = 1), the end
// of iterations and the tolerance stop (stop = 2).
inline void gpuRunSimStep(gpuLBM& lbm, bool verbose = false) {
  const int numRes = 20;
  if (lbm.stop != 0) {
    lbm.pause = 1;
    return;
  }
  lbm.updateStepCount();
  if (lbm.tol != 0) {
    lbm.stop = lbm.runOne(1);
    if (lbm.step >= 2 && lbm.step <= numRes) {
      lbm.initRes[lbm.step - 2] = lbm.residual;
      lbm.residual = 1;
    } else if (lbm.step == numRes + 1) {
      for (int i = 0; i < numRes - 1; i++) lbm.resNorm = lbm.initRes[i] > lbm.resNorm ? lbm.initRes[i] : lbm.resNorm;
      lbm.residual = lbm.residual / lbm.resNorm;
    } else if (lbm.step >= numRes + 1) {
      lbm.residual = lbm.residual / lbm.resNorm;
    }
  } else {
    lbm.stop = lbm.runOne();
  }
  if (verbose) std::printf("iteration %d/%d\n", lbm.step, lbm.nsteps);
  if (lbm.step >= lbm.nsteps && lbm.nsteps != 0) lbm.stop = 2;
  if (lbm.tol != 0 && lbm.residual <= lbm.tol && lbm.step >= numRes + 1) lbm.stop = 2;
}

#endif  // GPULBM_HPP
