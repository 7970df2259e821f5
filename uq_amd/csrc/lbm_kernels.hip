// lbm_kernels.hip — gfx950 kernels of the D2Q13 thermal FD-LBM.
//
// One step of the reference's RKMETHOD = 1 stepper (IMEX-SSP2(3,2,2),
// reference RK3, cpuLBM.cpp:1738-2103) is four tile launches, each walking
// every tile of every instance of the group in one grid-stride (persistent)
// loop, plus the wall closure when the group has walls:
//
//   k_s12      pointwise: moments, stage-1 equilibrium, implicit stages 1-2
//              -> f2 (cpuLBM.cpp:1819-1924)
//  [k_alpha    wall cells: closure factor from the fluid neighbours' face
//              fluxes of f2 (wallFluxes, cpuLBM.cpp:1502-1590)]
//   k_flux<0>  stencil: per direction, f2 of tile + 3-cell halo in LDS,
//              x/y face fluxes (WENO5 or NND), net flux dF(f2) -> dF planes
//              (posFluxes + cFLUX, cpuLBM.cpp:1436-1500)
//   k_s3       pointwise: moments of f - dF(f2), stage-3 equilibrium,
//              implicit stage 3 -> f3 (cpuLBM.cpp:1926-1993)
//  [k_alpha    same for f3]
//   k_flux<1>  stencil: net flux dF(f3); the final combination, new state
//              f3 + (dF(f2) - dF(f3))/2 (cpuLBM.cpp:2009-2042)
//
// On launch groups large enough to fill the GPU the stencil stages run on
// the strip kernels of lbm_strip.hpp (one wavefront per 58-column strip,
// rows marched through registers) instead of the tile kernels below.
//
// The stencil kernels run one direction at a time through a runtime loop
// (one copy of the flux code, low register use) on tiles of one of five
// shapes (the autotuner picks one per launch group): every thread computes
// at most one x face and one y face per direction, so no wave waits on a
// straggler at the barriers. FP32 storage stages the next direction into a
// second LDS buffer with direct-to-LDS loads while the faces of the current
// one are computed; FP64 stages through registers.
#include <climits>
#include <type_traits>

#include "lbm_device.hpp"
#include "lbm_launch.hpp"

namespace uqlbm {

constexpr int kH = 3;                          // stencil halo (WENO5 reach)
constexpr int kNT = 256;                       // threads of the pointwise blocks
constexpr int kNDF = 24;                       // dF planes: f and h of directions 1..12

// region meta bits
enum : unsigned {
  kMetaType = 3u,
  kMetaMirEW = 4u,    // x position reached through an east/west mirror
  kMetaMirNS = 8u,    // y position reached through a north/south mirror
  kMetaClampX = 16u,  // x position clamped (zeroth-order extrapolation edge)
  kMetaClampY = 32u,
};

__constant__ int c_lat_x[13] = {0, 1, 0, -1, 0, 1, -1, -1, 1, 2, 0, -2, 0};
__constant__ int c_lat_y[13] = {0, 0, 1, 0, -1, 1, 1, -1, -1, 0, 2, 0, -2};
__constant__ int c_mir_ns[13] = {0, 1, 4, 3, 2, 8, 7, 6, 5, 9, 12, 11, 10};
__constant__ int c_mir_ew[13] = {0, 3, 2, 1, 4, 6, 5, 8, 7, 11, 10, 9, 12};

template <typename T>
__device__ __forceinline__ T* plane(void* base, long long plane_elems, int k) {
  return reinterpret_cast<T*>(base) + plane_elems * (long long)k;
}

// Reference index() along one axis (cpuLBM.cpp:1033-1113) for one position,
// then a clamp into the grid for positions no stencil reaches.
__device__ __forceinline__ int map_axis(int X, int n, int periodic, int mir_lo, int mir_hi, unsigned& mir,
                                        unsigned& clamped) {
  mir = 0;
  clamped = 0;
  if (X < 0) {
    if (periodic) X += n;
    else if (mir_lo) { X = -X - 1; mir = 1; }
    else { X = 0; clamped = 1; }
  } else if (X > n - 1) {
    if (periodic) X -= n;
    else if (mir_hi) { X = 2 * n - X - 1; mir = 1; }
    else { X = n - 1; clamped = 1; }
  }
  return X < 0 ? 0 : (X > n - 1 ? n - 1 : X);
}

// Grid-stride tile loop with an XCD-aware order: gfx950 hands consecutive
// workgroups to its 8 XCDs in turn, so block b is renumbered so that each
// XCD walks a contiguous run of tiles (neighbouring tiles share halos and
// that XCD's L2).
__device__ __forceinline__ int first_tile() {
  const int G = gridDim.x, b = blockIdx.x;
  if (G % 8 != 0) return b;
  return (b % 8) * (G / 8) + b / 8;
}

template <int... Is>
struct Seq {};
template <int N, int... Is>
struct MakeSeq : MakeSeq<N - 1, N - 1, Is...> {};
template <int... Is>
struct MakeSeq<0, Is...> {
  using type = Seq<Is...>;
};

// `step` counts the steps of the current launch sequence from 1; the global
// step index of an instance is in.step_base + step. A failure flagged during
// step s freezes the instance for steps > s only, so this test gives the
// same answer to every thread of a launch.
__device__ __forceinline__ bool frozen_by_failure(const InstDev& in, int step) {
  return *in.fail < in.step_base + step;
}

__device__ __forceinline__ void flag_failure(const InstDev& in, int step, double rho, double sumfeq, double tol) {
  // reference check |sum feq - rho| > 1e-6 (cpuLBM.cpp:825-843); a
  // non-finite density also fails (the reference misses NaN).
  if (!(fabs(sumfeq - rho) <= tol) || !isfinite(rho)) atomicMin(in.fail, in.step_base + step);
}

// Tolerance of the equilibrium mass check: the reference's 1e-6 for FP64
// equilibria; for the FP32 equilibria of FP32 storage (TE = float: the
// deviations from the reference equilibrium, whose absolute values sum to
// sum_abs) the rounding bound of the FP32 evaluation and sum is added, so
// that rounding alone never trips it.
template <typename TE>
__device__ __forceinline__ double eq_check_tol(double sum_abs) {
  return std::is_same<TE, double>::value ? 1e-6 : 1e-6 + 16.0 * 5.96e-8 * sum_abs;
}

// Global-memory pointers: loads and stores through them are global, not
// flat, operations (the pointers read out of the instance table are
// generic), so they count on the vector-memory counter only: a flat access
// also counts on the LDS/scalar counter, and every wait for a scalar load
// then waits for it too.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;

template <typename T>
__device__ __forceinline__ gptr<T> as_global(T* p) {
  return (gptr<T>)p;
}
template <typename T>
__device__ __forceinline__ gptr<T> as_global(gptr<T> p) {
  return p;
}

// ---------------------------------------------------------------------------
// Pointwise helpers
// ---------------------------------------------------------------------------
// The pointwise kernels read each input plane once: non-temporal loads keep
// them from displacing the lines the stencil kernels reuse (k_s12 and k_s3
// 2-2.5 % faster).
__device__ __forceinline__ void sums26(const InstDev& in, const double* f, const double* h, double* s) {
  double rho = 0, rux = 0, ruy = 0, sh = 0;
#pragma unroll
  for (int i = 0; i < kQ; i++) {
    rho += f[i];
    rux += f[i] * (lat_x(i) * in.c);
    ruy += f[i] * (lat_y(i) * in.c);
    sh += h[i];
  }
  s[0] = rho;
  s[1] = rux;
  s[2] = ruy;
  s[3] = sh;
}

__device__ __forceinline__ void moments26(const InstDev& in, const double* f, const double* h, Macro& m) {
  double s[4];
  sums26(in, f, h, s);
  finish_macro(in, s[0], s[1], s[2], s[3], m);
}

// The state of one cell and its conserved sums s (mass, x and y momentum,
// energy). FP64 storage: the stored values and their sums. FP32 storage:
// the stored deviations plus the reference equilibrium, corrected so that
// the sums equal the FP64 moments kept beside the state: the mass defect
// goes to the rest population f0, each momentum defect to the axis pair of
// its axis (+d/(2c) on f1 or f2, -d/(2c) on f3 or f4, which moves no mass),
// the energy defect to h0. The corrections are of the size of the FP32
// rounding; they keep that rounding out of the conserved quantities. The
// stored values stay in their storage type until a direction uses them
// (value<I>), which keeps the FP32 loads at half the registers.
template <typename TS>
struct CellState {
  TS f[kQ], h[kQ];
  double s[4];
  double cf[5], ch;  // FP32 storage: corrections of f0..f4 and h0
};

template <typename TS, typename PS, typename PM>
__device__ __forceinline__ void load_state(const InstDev& in, PS srcp, PM momp, long long o, CellState<TS>& c) {
  const long long P = in.plane;
  const gptr<const TS> src = as_global(srcp);
#pragma unroll
  for (int i = 0; i < kQ; i++) {
    c.f[i] = __builtin_nontemporal_load(&src[P * i + o]);
    c.h[i] = __builtin_nontemporal_load(&src[P * (kQ + i) + o]);
  }
  double t[4] = {0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < kQ; i++) {
    const double fv = double(c.f[i]);
    t[0] += fv;
    t[1] += fv * (lat_x(i) * in.c);
    t[2] += fv * (lat_y(i) * in.c);
    t[3] += double(c.h[i]);
  }
  if constexpr (sizeof(TS) == 4) {
#pragma unroll
    for (int k = 0; k < 4; k++) t[k] += in.refsum[k];
    const gptr<const double> mom = as_global(momp);
#pragma unroll
    for (int k = 0; k < 4; k++) c.s[k] = __builtin_nontemporal_load(&mom[P * k + o]);
    const double dx = (c.s[1] - t[1]) * in.i2c, dy = (c.s[2] - t[2]) * in.i2c;
    c.cf[0] = c.s[0] - t[0];
    c.cf[1] = dx;
    c.cf[3] = -dx;
    c.cf[2] = dy;
    c.cf[4] = -dy;
    c.ch = c.s[3] - t[3];
  } else {
    (void)momp;
#pragma unroll
    for (int k = 0; k < 4; k++) c.s[k] = t[k];
  }
}

// f_I and h_I of a loaded cell in FP64, in the stored form: FP32 storage
// works on deviations from the reference throughout the pointwise stages
// (every stage update is an affine combination of distribution values with
// weights summing to one, plus differences, so it maps deviations to
// deviations; only the equilibria need the reference taken off).
template <int I, typename TS>
__device__ __forceinline__ double fval(const CellState<TS>& c) {
  if constexpr (sizeof(TS) == 4) {
    double v = double(c.f[I]);
    if constexpr (I < 5) v += c.cf[I];
    return v;
  } else {
    return c.f[I];
  }
}
template <int I, typename TS>
__device__ __forceinline__ double hval(const CellState<TS>& c) {
  if constexpr (sizeof(TS) == 4) {
    double v = double(c.h[I]);
    if constexpr (I == 0) v += c.ch;
    return v;
  } else {
    return c.h[I];
  }
}

// Equilibrium of direction I in the stored form. TE is the type the
// equilibria are evaluated in: FP64 (the closed form; FP32 storage then
// takes the reference equilibrium off) or, for FP32 storage of decks near
// their reference equilibrium, FP32 in the deviation form (eq_dev), whose
// every term is of the deviation's size. With SUM, the stored-form value
// joins the failure-check sums (FP32 storage: sums of deviations, to which
// failure_mass adds the reference's mass).
template <int I, typename TS, bool SUM, typename TE, typename TC>
__device__ __forceinline__ void eq_stored(const InstDev& in, const EqIn<TC>& q, const EqDev& qd, TC& fe, TC& he,
                                          double& sumfe, double& sumabs) {
  if constexpr (sizeof(TE) == 4) {
    static_assert(sizeof(TS) == 4, "FP32 equilibria only for FP32 storage");
    float a, b;
    eq_dev<I>(qd, in.eqref, a, b);
    fe = TC(a);
    he = TC(b);
  } else {
    (void)qd;
    eq_dir<I>(q, fe, he);
    if constexpr (sizeof(TS) == 4) {
      fe -= TC(in.ref[I]);
      he -= TC(in.ref[kQ + I]);
    }
  }
  if constexpr (SUM) {
    sumfe += double(fe);
    sumabs += fabs(double(fe));
  }
}

// Sum of the equilibrium f of a cell from the failure-check sums.
template <typename TS>
__device__ __forceinline__ double failure_mass(const InstDev& in, double sumfe) {
  if constexpr (sizeof(TS) == 4) return sumfe + in.refsum[0];
  else return sumfe;
}

// Stored value of a full distribution value computed in FP64: the
// deviation from the reference for FP32 storage.
template <typename TS>
__device__ __forceinline__ TS to_store(const InstDev& in, int k, double v) {
  if constexpr (sizeof(TS) == 4) return TS(v - in.ref[k]);
  else return TS(v);
}

// The state buffer k of an instance and its moment planes.
template <typename TS>
__device__ __forceinline__ const TS* state_buf(const InstDev& in, int k) {
  return reinterpret_cast<const TS*>(in.state[k]);
}

// Stage-1/2 coefficients of one cell. Every divisor of the stage updates is
// applied as a reciprocal formed once per cell.
template <typename TC>
struct Rec12 {
  EqIn<TC> q;
  TC cU, cV;           // c*U, c*V: e_i . u = lat_x(i)*cU + lat_y(i)*cV
  TC af, ah, ahf, g;   // dt/(2 tf), dt/(2 th), dt/(2 thf), 1/thf
  TC iaf, iah;         // 1/(1 + af), 1/(1 + ah)
};

template <typename TC>
__device__ __forceinline__ Rec12<TC> make_rec12(const InstDev& in, const Macro& m) {
  Rec12<TC> r;
  r.q = eq_inputs<TC>(in, m);
  r.cU = TC(in.c) * TC(m.ux);
  r.cV = TC(in.c) * TC(m.uy);
  const TC hdt = TC(in.dt) * TC(0.5);
  r.af = hdt * TC(m.itf);
  r.ah = hdt * TC(m.ith);
  r.ahf = hdt * TC(m.ithf);
  r.g = TC(m.ithf);
  r.iaf = TC(1) / (TC(1) + r.af);
  r.iah = TC(1) / (TC(1) + r.ah);
  return r;
}

// f2, h2 of direction I (reference stages 1-2, cpuLBM.cpp:1853, 1862, 1905,
// 1920) and the stage-1 equilibria fe, he (feq2 = feq1).
template <int I, typename TC>
__device__ __forceinline__ void s12(const Rec12<TC>& r, TC hdt, TC fv, TC hv, TC fe, TC he, TC& f2, TC& h2) {
  const TC f1 = (fv + r.af * fe) * r.iaf;
  const TC eu = TC(lat_x(I)) * r.cU + TC(lat_y(I)) * r.cV;
  const TC h1 = (hv - r.ahf * eu * (fe - f1) + r.ah * he) * r.iah;
  f2 = (fv - r.af * (fe - f1) + r.af * fe) * r.iaf;
  h2 = (hv - hdt * ((eu * r.g) * (fe - f2) - (eu * r.g) * (fe - f1)) - r.ah * (he - h1) + r.ah * he) * r.iah;
}

// Own cell of a pointwise launch: 64 x 4 row tiles, so each wavefront reads
// one aligned 64-cell row segment of a plane.
struct PointCell {
  long long o;
  bool active;
};

__device__ __forceinline__ PointCell point_cell(const InstDev& in, const TileRef& tr) {
  const int x = tr.tx * kPointW + (int)threadIdx.x % kPointW;
  const int y = tr.ty * kPointH + (int)threadIdx.x / kPointW;
  PointCell c;
  c.active = x < in.nx && y < in.ny;
  c.o = c.active ? (long long)y * in.pitch + x : 0;
  return c;
}

// Whether the pointwise kernel instantiated for equilibrium type TE takes
// instance `in`: FP32 storage picks the equilibrium precision per instance
// (InstDev::eq32), and a launch group with both kinds launches each
// pointwise kernel once per kind.
template <typename TS, typename TE>
__device__ __forceinline__ bool eq_kernel_for(const InstDev& in) {
  if constexpr (sizeof(TS) == 4) return (in.eq32 != 0) == (sizeof(TE) == 4);
  else return true;
}

// ---------------------------------------------------------------------------
// k_s12: RK3 stages 1 and 2, pointwise -> f2.
// ---------------------------------------------------------------------------
template <typename TS, typename TC, typename TE, int... Is>
__device__ __forceinline__ void s12_cell(const InstDev& in, gptr<const TS> src, gptr<const double> mom, gptr<TS> dst,
                                         long long o, int step, Seq<Is...>) {
  const long long P = in.plane;
  CellState<TS> cs;
  load_state(in, src, mom, o, cs);
  Macro m;
  finish_macro(in, cs.s[0], cs.s[1], cs.s[2], cs.s[3], m);
  const Rec12<TC> r = make_rec12<TC>(in, m);
  const EqDev qd = eq_dev_inputs(in, m);
  const TC hdt = TC(in.dt) / TC(2);
  double sumfe = 0, sumabs = 0;
  auto one = [&](auto Ic) {
    constexpr int I = decltype(Ic)::value;
    TC fe, he, f2, h2;
    eq_stored<I, TS, true, TE>(in, r.q, qd, fe, he, sumfe, sumabs);
    s12<I>(r, hdt, TC(fval<I>(cs)), TC(hval<I>(cs)), fe, he, f2, h2);
    // streaming stores: f2 and, in stage 3, f3 are written once per step
    // and read back only by the next kernel, long after L2 has turned over.
    // The rest direction of f2 has no flux and no wall term, and stage 3
    // recomputes f2 from the state: nothing reads it.
    if constexpr (I != 0) {
      __builtin_nontemporal_store(TS(f2), &dst[P * I + o]);
      __builtin_nontemporal_store(TS(h2), &dst[P * (kQ + I) + o]);
    }
  };
  (one(std::integral_constant<int, Is>{}), ...);
  flag_failure(in, step, m.rho, failure_mass<TS>(in, sumfe), eq_check_tol<TE>(sumabs));
}

template <typename TS, typename TC, typename TE>
__global__ __launch_bounds__(kNT) void k_s12(const InstDev* __restrict__ insts, const TileRef* __restrict__ tiles,
                                             int ntiles, int par, const int* stepp) {
  const int step = *stepp;
  for (int t = first_tile(); t < ntiles; t += gridDim.x) {
    const TileRef tr = tiles[t];
    const InstDev& in = insts[tr.inst];
    if (frozen_by_failure(in, step)) continue;
    if (!eq_kernel_for<TS, TE>(in)) continue;
    const PointCell c = point_cell(in, tr);
    if (!c.active) continue;
    const int cur = in.parity0 ^ par;
    const gptr<const TS> src = as_global(state_buf<TS>(in, cur));
    const gptr<TS> dst = as_global(reinterpret_cast<TS*>(in.f2));
    if (as_global(in.ctype)[c.o] > 0) {
#pragma unroll
      for (int k = 0; k < kPlanes; k++) dst[in.plane * k + c.o] = src[in.plane * k + c.o];
      continue;
    }
    s12_cell<TS, TC, TE>(in, src, as_global(in.mom[cur]), dst, c.o, step, MakeSeq<kQ>::type{});
  }
}

// ---------------------------------------------------------------------------
// k_s3: RK3 stage 3 (cpuLBM.cpp:1926-1993), pointwise, one pass over the
// directions. With P = f - dF2 + af2 (feq2 - f2) and
// Q = (h - dH2) - dt/2 A2 + ah2 (heq2 - h2), A = (e.u/thf)(feq - f):
//   f3 = (P + af3 feq3)/(1 + af3),  h3 = (Q - dt/2 A3 + ah3 heq3)/(1 + ah3).
// The final combination (cpuLBM.cpp:2009-2040) is
//   f' = P + dF2/2 + dt/2 (feq3 - f3)/tf3 - dF3/2
//   h' = Q + dH2/2 + dt/2 (heq3 - h3)/th3 - dt/2 A3 - dH3/2
// (P + dF2/2 is f - dF2/2 + dt/2 (feq2 - f2)/tf2 since af2 = dt/(2 tf2)),
// and since f3 = P + af3 (feq3 - f3) and h3 = Q - dt/2 A3 + ah3 (heq3 - h3)
// with af3 = dt/(2 tf3), ah3 = dt/(2 th3), that is
//   f' = f3 + (dF2 - dF3)/2,  h' = h3 + (dH2 - dH3)/2:
// FP64 storage: the stage-3 flux kernel forms the new state from f3 and the
// stored dF2, and this kernel writes only f3 (its rest direction, which has
// no flux, straight into the new state). FP32 storage (kStage3G): this
// kernel writes f3 and G = f' + dF3/2, formed here in FP64 as above, and
// the flux kernel subtracts dF3/2.
// ---------------------------------------------------------------------------
template <typename TS, typename TC, typename TE, int... Is>
__device__ __forceinline__ void s3_cell(const InstDev& in, gptr<const TS> src, gptr<const double> mom,
                                        gptr<const TS> dfb, gptr<TS> f3b, gptr<TS> out, gptr<double> mom_out, long long o,
                                        int step, Seq<Is...>) {
  const long long P = in.plane;
  CellState<TS> cs;
  load_state(in, src, mom, o, cs);
  const double* s = cs.s;
  // stage-1 moments, and the stage-3 moments: those of f - dF(f2)
  // (macroProp3, cpuLBM.cpp:973-1032), formed as moments of f minus the
  // moments of the stored dF, summed here in a first pass over the
  // directions (the loads leave the lines in L2 for the second pass), so
  // that every direction can then be finished in one pass
  double drho = 0, drux = 0, druy = 0, dsh = 0;
#pragma unroll
  for (int I = 1; I < kQ; I++) {
    const double dFs = double(dfb[P * (I - 1) + o]);
    drho += dFs;
    drux += dFs * (lat_x(I) * in.c);
    druy += dFs * (lat_y(I) * in.c);
    dsh += double(dfb[P * (12 + I - 1) + o]);
  }
  Macro m1, m3;
  finish_macro(in, s[0], s[1], s[2], s[3], m1);
  finish_macro(in, s[0] - drho, s[1] - drux, s[2] - druy, s[3] - dsh, m3);
  if constexpr (sizeof(TS) == 4) {
    // conserved moments of the new state before its dF(f3)/2 term: the
    // relaxation terms move none, so they are those of f - dF(f2)/2; the
    // stage-3 flux kernel subtracts half those of dF(f3)
    mom_out[o] = s[0] - 0.5 * drho;
    mom_out[P + o] = s[1] - 0.5 * drux;
    mom_out[2 * P + o] = s[2] - 0.5 * druy;
    mom_out[3 * P + o] = s[3] - 0.5 * dsh;
  } else {
    (void)mom_out;
  }
  const Rec12<TC> r = make_rec12<TC>(in, m1);
  const EqDev qd1 = eq_dev_inputs(in, m1), qd3 = eq_dev_inputs(in, m3);
  const TC hdt = TC(in.dt) * TC(0.5);
  const EqIn<TC> q3 = eq_inputs<TC>(in, m3);
  const TC af3 = hdt * TC(m3.itf), ah3 = hdt * TC(m3.ith), ithf3 = TC(m3.ithf);
  const TC iaf3 = TC(1) / (TC(1) + af3), iah3 = TC(1) / (TC(1) + ah3);
  const TC cU3 = TC(in.c) * TC(m3.ux), cV3 = TC(in.c) * TC(m3.uy);
  double sumfe = 0, sumabs = 0;
  auto one = [&](auto Ic) {
    constexpr int I = decltype(Ic)::value;
    TC dF = 0, dH = 0;
    if constexpr (I != 0) {
      dF = TC(__builtin_nontemporal_load(&dfb[P * (I - 1) + o]));
      dH = TC(__builtin_nontemporal_load(&dfb[P * (12 + I - 1) + o]));
    }
    const TC fv = TC(fval<I>(cs)), hv = TC(hval<I>(cs));
    TC fe, he, f2, h2;
    eq_stored<I, TS, false, TE>(in, r.q, qd1, fe, he, sumfe, sumabs);
    s12<I>(r, hdt, fv, hv, fe, he, f2, h2);
    const TC D = fe - f2;
    const TC eu2 = TC(lat_x(I)) * r.cU + TC(lat_y(I)) * r.cV;
    const TC A2 = (eu2 * r.g) * D;
    const TC Pf = (fv - dF) + r.af * D;
    const TC Qh = ((hv - dH) - hdt * A2) + r.ah * (he - h2);
    TC fe3, he3;
    eq_stored<I, TS, true, TE>(in, q3, qd3, fe3, he3, sumfe, sumabs);
    const TC f3 = (Pf + af3 * fe3) * iaf3;
    const TC eu3 = TC(lat_x(I)) * cU3 + TC(lat_y(I)) * cV3;
    const TC A3 = (eu3 * ithf3) * (fe3 - f3);
    const TC h3 = ((Qh - hdt * A3) + ah3 * he3) * iah3;
    if constexpr (kStage3G<TS>) {
      const TC gf = (Pf + TC(0.5) * dF) + af3 * (fe3 - f3);
      const TC gh = ((Qh + TC(0.5) * dH) + ah3 * (he3 - h3)) - hdt * A3;
      if constexpr (I != 0) {  // the rest direction of f3 has no flux and no wall term: nothing reads it
        __builtin_nontemporal_store(TS(f3), &f3b[P * I + o]);
        __builtin_nontemporal_store(TS(h3), &f3b[P * (kQ + I) + o]);
      }
      __builtin_nontemporal_store(TS(gf), &out[P * I + o]);
      __builtin_nontemporal_store(TS(gh), &out[P * (kQ + I) + o]);
    } else {
      const gptr<TS> dst = I == 0 ? out : f3b;
      __builtin_nontemporal_store(TS(f3), &dst[P * I + o]);
      __builtin_nontemporal_store(TS(h3), &dst[P * (kQ + I) + o]);
    }
    // one direction at a time: keeps the scheduler from interleaving the
    // directions' temporaries (FP32 storage: 38 spilled registers otherwise)
    if constexpr (sizeof(TS) == 4) __builtin_amdgcn_sched_barrier(0);
  };
  (one(std::integral_constant<int, Is>{}), ...);
  flag_failure(in, step, m3.rho, failure_mass<TS>(in, sumfe), eq_check_tol<TE>(sumabs));
}

// Frozen cell (permanent or solid) of a pointwise stage: the stored values
// (and, for FP32 storage, the moment planes) carried over unchanged.
template <typename TS>
__device__ __forceinline__ void copy_frozen(const InstDev& in, gptr<const TS> src, gptr<TS> dst, int from, int to,
                                            long long o) {
#pragma unroll
  for (int k = 0; k < kPlanes; k++) dst[in.plane * k + o] = src[in.plane * k + o];
  if constexpr (sizeof(TS) == 4) {
    const gptr<const double> mf = as_global(in.mom[from]);
    const gptr<double> mt = as_global(in.mom[to]);
#pragma unroll
    for (int k = 0; k < 4; k++) mt[in.plane * k + o] = mf[in.plane * k + o];
  } else {
    (void)from;
    (void)to;
  }
}

// Two waves per SIMD (FP64 arithmetic for both storage types): the
// instantiation otherwise takes 262 registers (one wave per SIMD); capping
// it at 256 costs a few spills and hides twice the memory latency (4096^2
// FP64: 5.9 -> 4.2 ms per launch).
template <typename TS, typename TC, typename TE>
__global__ __launch_bounds__(kNT) __attribute__((amdgpu_waves_per_eu(2))) void k_s3(
    const InstDev* __restrict__ insts, const TileRef* __restrict__ tiles, int ntiles, int par, const int* stepp) {
  const int step = *stepp;
  for (int t = first_tile(); t < ntiles; t += gridDim.x) {
    const TileRef tr = tiles[t];
    const InstDev& in = insts[tr.inst];
    if (frozen_by_failure(in, step)) continue;
    if (!eq_kernel_for<TS, TE>(in)) continue;
    const PointCell c = point_cell(in, tr);
    if (!c.active) continue;
    const int cur = in.parity0 ^ par, nxt = 1 ^ cur;
    const gptr<const TS> src = as_global(state_buf<TS>(in, cur));
    const gptr<TS> out = as_global(reinterpret_cast<TS*>(in.state[nxt]));
    const gptr<TS> f3b = as_global(reinterpret_cast<TS*>(in.f3));
    if (as_global(in.ctype)[c.o] > 0) {
      // frozen cell: f3 = f2 = f, and the new state is f
#pragma unroll
      for (int k = 0; k < kPlanes; k++) f3b[in.plane * k + c.o] = src[in.plane * k + c.o];
      copy_frozen(in, src, out, cur, nxt, c.o);
      continue;
    }
    s3_cell<TS, TC, TE>(in, src, as_global(in.mom[cur]), as_global(reinterpret_cast<const TS*>(in.dflux)), f3b, out,
                        as_global(in.mom[nxt]), c.o, step, MakeSeq<kQ>::type{});
  }
}

// ---------------------------------------------------------------------------
// The one-flux-stage steppers (reference RK1 cpuLBM.cpp:1597-1737, EULER
// :2104-2193, EULER_EQ :2194-2283):
//   RK1:      k_s1 (f1 -> f2 buffer), k_flux<0>, k_fin<0>
//   EULER:    k_flux<2>, k_fin<2>
//   EULER_EQ: k_flux<2>, k_fin<3>
// ---------------------------------------------------------------------------
template <typename TC>
struct Rec1 {
  EqIn<TC> q;
  TC cU, cV;
  TC bf, bh, dt, ithf;  // dt/tf, dt/th, dt, 1/thf
  TC ibf, ibh;          // 1/(1 + bf), 1/(1 + bh)
};

template <typename TC>
__device__ __forceinline__ Rec1<TC> make_rec1(const InstDev& in, const Macro& m) {
  Rec1<TC> r;
  r.q = eq_inputs<TC>(in, m);
  r.cU = TC(in.c) * TC(m.ux);
  r.cV = TC(in.c) * TC(m.uy);
  r.dt = TC(in.dt);
  r.bf = r.dt * TC(m.itf);
  r.bh = r.dt * TC(m.ith);
  r.ithf = TC(m.ithf);
  r.ibf = TC(1) / (TC(1) + r.bf);
  r.ibh = TC(1) / (TC(1) + r.bh);
  return r;
}

// implicit stage of RK1 (cpuLBM.cpp:1650, 1660)
template <int I, typename TC>
__device__ __forceinline__ void s1(const Rec1<TC>& r, TC fv, TC hv, TC fe, TC he, TC& f1, TC& h1) {
  f1 = (fv + r.bf * fe) * r.ibf;
  const TC eu = TC(lat_x(I)) * r.cU + TC(lat_y(I)) * r.cV;
  h1 = (hv - r.dt * eu * ((fe - f1) * r.ithf) + r.bh * he) * r.ibh;
}

template <typename TS, typename TC, typename TE, int... Is>
__device__ __forceinline__ void s1_cell(const InstDev& in, gptr<const TS> src, gptr<const double> mom, gptr<TS> dst,
                                        long long o, int step, Seq<Is...>) {
  const long long P = in.plane;
  CellState<TS> cs;
  load_state(in, src, mom, o, cs);
  Macro m;
  finish_macro(in, cs.s[0], cs.s[1], cs.s[2], cs.s[3], m);
  const Rec1<TC> r = make_rec1<TC>(in, m);
  const EqDev qd = eq_dev_inputs(in, m);
  double sumfe = 0, sumabs = 0;
  auto one = [&](auto Ic) {
    constexpr int I = decltype(Ic)::value;
    TC fe, he, f1, h1;
    eq_stored<I, TS, true, TE>(in, r.q, qd, fe, he, sumfe, sumabs);
    s1<I>(r, TC(fval<I>(cs)), TC(hval<I>(cs)), fe, he, f1, h1);
    dst[P * I + o] = TS(f1);
    dst[P * (kQ + I) + o] = TS(h1);
  };
  (one(std::integral_constant<int, Is>{}), ...);
  flag_failure(in, step, m.rho, failure_mass<TS>(in, sumfe), eq_check_tol<TE>(sumabs));
}

template <typename TS, typename TC, typename TE>
__global__ __launch_bounds__(kNT) void k_s1(const InstDev* __restrict__ insts, const TileRef* __restrict__ tiles,
                                            int ntiles, int par, const int* stepp) {
  const int step = *stepp;
  for (int t = first_tile(); t < ntiles; t += gridDim.x) {
    const TileRef tr = tiles[t];
    const InstDev& in = insts[tr.inst];
    if (frozen_by_failure(in, step)) continue;
    if (!eq_kernel_for<TS, TE>(in)) continue;
    const PointCell c = point_cell(in, tr);
    if (!c.active) continue;
    const int cur = in.parity0 ^ par;
    const gptr<const TS> src = as_global(state_buf<TS>(in, cur));
    const gptr<TS> dst = as_global(reinterpret_cast<TS*>(in.f2));
    if (as_global(in.ctype)[c.o] > 0) {
#pragma unroll
      for (int k = 0; k < kPlanes; k++) dst[in.plane * k + c.o] = src[in.plane * k + c.o];
      continue;
    }
    s1_cell<TS, TC, TE>(in, src, as_global(in.mom[cur]), dst, c.o, step, MakeSeq<kQ>::type{});
  }
}

// Final update of the one-flux-stage steppers, METHOD 0 (RK1, cpuLBM.cpp:
// 1696, 1707), 2 (EULER, :2155, :2165), 3 (EULER_EQ, :2245, :2255).
template <int METHOD, typename TS, typename TC, typename TE, int... Is>
__device__ __forceinline__ void fin_cell(const InstDev& in, gptr<const TS> src, gptr<const double> mom,
                                         gptr<const TS> dfb, gptr<TS> out, gptr<double> mom_out, long long o, int step,
                                         Seq<Is...>) {
  const long long P = in.plane;
  CellState<TS> cs;
  load_state(in, src, mom, o, cs);
  Macro m;
  finish_macro(in, cs.s[0], cs.s[1], cs.s[2], cs.s[3], m);
  const Rec1<TC> r = make_rec1<TC>(in, m);
  const EqDev qd = eq_dev_inputs(in, m);
  double sumfe = 0, sumabs = 0;
  double t[4] = {0, 0, 0, 0};  // conserved sums of the new state (FP32 storage)
  auto one = [&](auto Ic) {
    constexpr int I = decltype(Ic)::value;
    TC dF = 0, dH = 0;
    if constexpr (I != 0) {
      dF = TC(__builtin_nontemporal_load(&dfb[P * (I - 1) + o]));
      dH = TC(__builtin_nontemporal_load(&dfb[P * (12 + I - 1) + o]));
    }
    const TC fv = TC(fval<I>(cs)), hv = TC(hval<I>(cs));
    const TC eu = TC(lat_x(I)) * r.cU + TC(lat_y(I)) * r.cV;
    TC nf, nh;
    if constexpr (METHOD == 0) {
      TC fe, he, f1, h1;
      eq_stored<I, TS, false, TE>(in, r.q, qd, fe, he, sumfe, sumabs);
      s1<I>(r, fv, hv, fe, he, f1, h1);
      nf = fv - dF + r.bf * (fe - f1);
      nh = hv - dH + r.bh * (he - h1) - ((r.dt * eu) * r.ithf) * (fe - f1);
    } else {
      TC fe, he;
      eq_stored<I, TS, true, TE>(in, r.q, qd, fe, he, sumfe, sumabs);
      if constexpr (METHOD == 2) {
        nf = fv - dF + r.bf * (fe - fv);
        nh = hv - dH + r.bh * (he - hv) - ((r.dt * eu) * r.ithf) * (fe - fv);
      } else {
        nf = fe - dF;
        nh = he - dH;
      }
    }
    if constexpr (sizeof(TS) == 4) {  // nf, nh are deviations: the reference's sums are added below
      t[0] += double(nf);
      t[1] += double(nf) * (lat_x(I) * in.c);
      t[2] += double(nf) * (lat_y(I) * in.c);
      t[3] += double(nh);
    }
    out[P * I + o] = TS(nf);
    out[P * (kQ + I) + o] = TS(nh);
  };
  (one(std::integral_constant<int, Is>{}), ...);
  if constexpr (sizeof(TS) == 4) {
#pragma unroll
    for (int k = 0; k < 4; k++) mom_out[P * k + o] = t[k] + in.refsum[k];
  } else {
    (void)mom_out;
  }
  if (METHOD != 0) flag_failure(in, step, m.rho, failure_mass<TS>(in, sumfe), eq_check_tol<TE>(sumabs));
}

template <typename TS, typename TC, typename TE, int METHOD>
__global__ __launch_bounds__(kNT) void k_fin(const InstDev* __restrict__ insts, const TileRef* __restrict__ tiles,
                                             int ntiles, int par, const int* stepp) {
  const int step = *stepp;
  for (int t = first_tile(); t < ntiles; t += gridDim.x) {
    const TileRef tr = tiles[t];
    const InstDev& in = insts[tr.inst];
    if (frozen_by_failure(in, step)) continue;
    if (!eq_kernel_for<TS, TE>(in)) continue;
    const PointCell c = point_cell(in, tr);
    if (!c.active) continue;
    const int cur = in.parity0 ^ par, nxt = 1 ^ cur;
    const gptr<const TS> src = as_global(state_buf<TS>(in, cur));
    const gptr<TS> out = as_global(reinterpret_cast<TS*>(in.state[nxt]));
    if (as_global(in.ctype)[c.o] > 0) {
      copy_frozen(in, src, out, cur, nxt, c.o);
      continue;
    }
    fin_cell<METHOD, TS, TC, TE>(in, src, as_global(in.mom[cur]), as_global(reinterpret_cast<const TS*>(in.dflux)),
                                 out, as_global(in.mom[nxt]), c.o, step, MakeSeq<kQ>::type{});
  }
}

// ---------------------------------------------------------------------------
// Stencil tiles: region (tile + 3-cell halo cross), staging and faces.
//
// A tile shape is TX x TY own cells worked by NT threads. Per direction the
// threads stage the direction's f, h over the tile and its halo into LDS
// (FP64 through registers: the own cells and the halo bands of the
// direction's nonzero axes; FP32 with direct-to-LDS loads of the whole
// region into one of two buffers, a direction ahead), compute the (TX+1) x
// TY x faces and TX x (TY+1) y faces (each face once, by the upwind owner's
// rule), then each own fluid cell's net flux. The shapes below keep the face and cell counts at or below one
// (or two) per thread so no wave waits on a straggler at the barriers; the
// autotuner picks one per launch group.
// ---------------------------------------------------------------------------
template <int TX_, int TY_, int NT_>
struct Shape {
  static constexpr int TX = TX_, TY = TY_, NT = NT_;
  static constexpr int RX = TX + 2 * kH, RY = TY + 2 * kH;
  static constexpr int Own = TX * TY;
  static constexpr int NXB = 2 * kH * TY, NYB = 2 * kH * TX;
  static constexpr int Slots = (Own + NXB + NYB + NT - 1) / NT;
  static constexpr int NXF = (TX + 1) * TY, NYF = TX * (TY + 1);
  static constexpr int XF = (NXF + NT - 1) / NT, YF = (NYF + NT - 1) / NT;
  static constexpr int Cells = (Own + NT - 1) / NT;
};
using Shape0 = Shape<15, 16, 256>;  // 240 own cells, 1.78 staged values per cell
using Shape1 = Shape<31, 15, 512>;  // 465 own cells, 1.59
using Shape2 = Shape<15, 32, 512>;  // 480 own cells, 1.59
using Shape3 = Shape<39, 12, 512>;  // 468 own cells, 1.65 (312-byte FP64 row segments)
using Shape4 = Shape<23, 20, 512>;  // 460 own cells, 1.56
static_assert(Shape0::XF == 1 && Shape0::YF == 1 && Shape1::XF == 1 && Shape1::YF == 1 &&
                  Shape2::XF == 1 && Shape2::YF == 1 && Shape3::XF == 1 && Shape3::YF == 1 &&
                  Shape4::XF == 1 && Shape4::YF == 1,
              "one x face and one y face per thread");

template <typename Sh>
struct Region {
  int off[Sh::RY][Sh::RX];
  unsigned char meta[Sh::RY][Sh::RX];
  double rdx[Sh::TX], rdy[Sh::TY];  // dt/dx of the tile's columns, dt/dy of its rows (non-uniform grids)
  float rdxf[Sh::TX], rdyf[Sh::TY];  // the same in FP32, for the FP32 stencil arithmetic
  float ref[kPlanes];               // FP32 storage: the instance's reference equilibrium
};

template <typename Sh>
__device__ void setup_region(Region<Sh>& R, const InstDev& in, int x0, int y0) {
  const int px_ = in.flags & kPeriodicX, py_ = in.flags & kPeriodicY;
  const int mw = in.flags & kMirrorW, me = in.flags & kMirrorE;
  const int ms = in.flags & kMirrorS, mn = in.flags & kMirrorN;
  for (int r = threadIdx.x; r < Sh::RX * Sh::RY; r += Sh::NT) {
    const int ly = r / Sh::RX - kH, lx = r % Sh::RX - kH;
    const bool inx = lx >= 0 && lx < Sh::TX, iny = ly >= 0 && ly < Sh::TY;
    if (!inx && !iny) {  // corner: never read
      R.off[ly + kH][lx + kH] = 0;
      R.meta[ly + kH][lx + kH] = 0;
      continue;
    }
    unsigned mx, cx, my, cy;
    const int X = map_axis(x0 + lx, in.nx, px_, mw, me, mx, cx);
    const int Y = map_axis(y0 + ly, in.ny, py_, ms, mn, my, cy);
    const long long o = (long long)Y * in.pitch + X;
    R.off[ly + kH][lx + kH] = (int)o;
    unsigned m = in.ctype[o] & kMetaType;
    m |= mx ? kMetaMirEW : 0u;
    m |= my ? kMetaMirNS : 0u;
    m |= cx ? kMetaClampX : 0u;
    m |= cy ? kMetaClampY : 0u;
    R.meta[ly + kH][lx + kH] = (unsigned char)m;
  }
  for (int k = threadIdx.x; k < kPlanes; k += Sh::NT) R.ref[k] = float(in.ref[k]);
  for (int k = threadIdx.x; k < Sh::TX + Sh::TY; k += Sh::NT) {  // dt/dx, dt/dy of the tile's columns and rows
    if (k < Sh::TX) {
      const double v = !in.rdxs ? in.rdx : (x0 + k < in.nx ? in.rdxs[x0 + k] : 0.0);
      R.rdx[k] = v;
      R.rdxf[k] = float(v);
    } else {
      const double v = !in.rdys ? in.rdy : (y0 + k - Sh::TX < in.ny ? in.rdys[y0 + k - Sh::TX] : 0.0);
      R.rdy[k - Sh::TX] = v;
      R.rdyf[k - Sh::TX] = float(v);
    }
  }
}

// An f value and its h partner, kept together in LDS so that one 8- or
// 16-byte access moves both.
template <typename TC>
struct alignas(2 * sizeof(TC)) FH {
  TC f, h;
};

// The staged f, h of one direction over the whole region, in region order
// (row-major RY x RX, halo corners included) and padded to whole wave loads.
// A wave load moves CH positions with LPP lanes per position, one dword per
// lane, straight from global memory into LDS.
template <typename Sh, typename TS>
struct StageBuf {
  static constexpr int LPP = int(sizeof(FH<TS>) / 4);
  static constexpr int CH = 64 / LPP;
  static constexpr int NCH = (Sh::RY * Sh::RX + CH - 1) / CH;
  FH<TS> v[NCH * CH];
};

template <typename Sh, typename TC, bool MOM = false>
struct FluxStage {
  FH<TC> fx[Sh::TY][Sh::TX + 1];  // x faces: face j is left of cell j
  FH<TC> fy[Sh::TY + 1][Sh::TX];  // y faces: face j is below cell j
  double mom[MOM ? 4 : 1][MOM ? Sh::Own : 1];  // per-cell sums of the stored dF, dH (MOM)
};

typedef __attribute__((address_space(1))) void global_void;
typedef __attribute__((address_space(3))) void lds_void;
constexpr int kWaitVmem = 0x0F70;  // s_waitcnt vmcnt(0): this wave's loads (also those into LDS) done

// Issue the loads of direction I's f, h over the region into B, without
// waiting for them: each wave takes every (NT/64)-th group of CH positions;
// the first LPP/2 lanes of a position load the f value of the
// mirror-mapped direction, the others its h value. Directions without a y
// component skip the groups that lie wholly in the top or bottom halo rows.
// A wave's loads have landed once it passes s_waitcnt vmcnt(0).
template <typename TS, typename Sh>
__device__ __forceinline__ void stage_dir(StageBuf<Sh, TS>& B, const Region<Sh>& R, const void* src,
                                          long long P, int I) {
  using SB = StageBuf<Sh, TS>;
  constexpr int NPos = Sh::RY * Sh::RX, NW = Sh::NT / 64;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int sub = lane / SB::LPP, q = lane % SB::LPP;
  const int comp = q / (SB::LPP / 2), word = q % (SB::LPP / 2);
  const bool ybands = c_lat_y[I] != 0;
  const int dew = c_mir_ew[I], dns = c_mir_ns[I], dboth = c_mir_ns[dew];
  const char* base = reinterpret_cast<const char*>(src);
  for (int c = w; c < SB::NCH; c += NW) {
    if (!ybands && (c * SB::CH + SB::CH <= kH * Sh::RX || c * SB::CH >= (Sh::RY - kH) * Sh::RX)) continue;
    const int p = c * SB::CH + sub, pc = p < NPos ? p : 0;  // padding lanes reload position 0
    const unsigned m = (&R.meta[0][0])[pc];
    const int d = (m & kMetaMirEW) ? ((m & kMetaMirNS) ? dboth : dew) : ((m & kMetaMirNS) ? dns : I);
    const long long e = (long long)(d + comp * kQ) * P + (&R.off[0][0])[pc];
    const char* g = base + e * (long long)sizeof(TS) + word * 4;
    __builtin_amdgcn_global_load_lds((global_void*)g, (lds_void*)&B.v[c * SB::CH], 4, 0, 0);
  }
}

// Synchronous staging through registers, for the kernels whose LDS cannot
// hold a second buffer: slot r = t + j NT of thread t is an own cell (r <
// Own), an x halo band position (the next NXB) or a y halo band position
// (the next NYB); a direction stages the own slots and the bands of its
// nonzero axes. Each slot keeps its LDS position, grid offset and meta byte
// for the whole tile.
template <typename Sh>
struct SlotSet {
  int lds[Sh::Slots], off[Sh::Slots], meta[Sh::Slots];
};

template <typename Sh>
__device__ __forceinline__ void setup_slots(SlotSet<Sh>& L, const Region<Sh>& R) {
#pragma unroll
  for (int j = 0; j < Sh::Slots; j++) {
    const int r = threadIdx.x + j * Sh::NT;
    int lx = 0, ly = 0, kind = 3;  // 0 own, 1 x band, 2 y band, 3 none
    if (r < Sh::Own) {
      lx = r % Sh::TX;
      ly = r / Sh::TX;
      kind = 0;
    } else if (r < Sh::Own + Sh::NXB) {
      const int k = r - Sh::Own, side = k / (kH * Sh::TY), kk = k % (kH * Sh::TY);
      ly = kk / kH;
      lx = side ? Sh::TX + kk % kH : -kH + kk % kH;
      kind = 1;
    } else if (r < Sh::Own + Sh::NXB + Sh::NYB) {
      const int k = r - Sh::Own - Sh::NXB, side = k / (kH * Sh::TX), kk = k % (kH * Sh::TX);
      lx = kk % Sh::TX;
      ly = side ? Sh::TY + kk / Sh::TX : -kH + kk / Sh::TX;
      kind = 2;
    }
    L.lds[j] = (ly + kH) * Sh::RX + (lx + kH);
    L.off[j] = kind < 3 ? R.off[ly + kH][lx + kH] : 0;
    L.meta[j] = kind < 3 ? (int)R.meta[ly + kH][lx + kH] | (kind << 8) : (3 << 8);
  }
}

// The staged values of one direction held in registers between their loads
// and their LDS writes, so that the loads of direction I+1 can be issued
// while direction I updates its cells.
template <typename Sh, typename TS>
struct SlotVals {
  TS f[Sh::Slots], h[Sh::Slots];
};

template <typename TS, typename Sh>
__device__ __forceinline__ void fetch_slots(SlotVals<Sh, TS>& V, const SlotSet<Sh>& L, const void* src, long long P,
                                            int I) {
  const int ex = c_lat_x[I], ey = c_lat_y[I];
  const int dew = c_mir_ew[I], dns = c_mir_ns[I], dboth = c_mir_ns[dew];
#pragma unroll
  for (int j = 0; j < Sh::Slots; j++) {
    const int kind = L.meta[j] >> 8;
    if (kind == 3 || (kind == 1 && !ex) || (kind == 2 && !ey)) continue;
    const unsigned m = (unsigned)L.meta[j];
    const int d = (m & kMetaMirEW) ? ((m & kMetaMirNS) ? dboth : dew) : ((m & kMetaMirNS) ? dns : I);
    const long long o = L.off[j];
    V.f[j] = as_global(plane<TS>(const_cast<void*>(src), P, d))[o];
    V.h[j] = as_global(plane<TS>(const_cast<void*>(src), P, kQ + d))[o];
  }
}

template <typename TS, typename Sh>
__device__ __forceinline__ void store_slots(StageBuf<Sh, TS>& B, const SlotVals<Sh, TS>& V, const SlotSet<Sh>& L,
                                            int I) {
  const int ex = c_lat_x[I], ey = c_lat_y[I];
#pragma unroll
  for (int j = 0; j < Sh::Slots; j++) {
    const int kind = L.meta[j] >> 8;
    if (kind == 3 || (kind == 1 && !ex) || (kind == 2 && !ey)) continue;
    B.v[L.lds[j]] = FH<TS>{V.f[j], V.h[j]};
  }
}

// Reference stencil solid handling along one line (cpuLBM.cpp:1242-1322):
// the first solid/fluid transition's solid point becomes 2 V(m1) - V(m2) of
// the two nearest fluid points, and the point beyond it the same value. The
// addressing equivalence with the reference is checked on the host per
// geometry (lbm_runtime.hip, validate_geometry).
// (The transition is located in the unrolled loop and the line updated in
// the same unrolled iteration, so every index is a constant and the stencil
// stays in registers.)
template <int L, typename TC>
__device__ __forceinline__ void extrapolate(const int* sld, const TC* V, TC* S) {
#pragma clang fp contract(off)
  bool found = false;
#pragma unroll
  for (int j = 0; j < L - 1; j++) {
    const int d = sld[j] - sld[j + 1];
    if (!found && d != 0) {
      found = true;
      if (d > 0) {  // solid point j, fluid beyond it at j+1, j+2
        const TC e = TC(2) * V[j + 1 < L ? j + 1 : L - 1] - V[j + 2 < L ? j + 2 : L - 1];
        S[j] = e;
        if (j - 1 >= 0) S[j - 1] = e;
      } else {      // solid point j+1, fluid before it at j, j-1
        const TC e = TC(2) * V[j] - V[j - 1 >= 0 ? j - 1 : 0];
        S[j + 1] = e;
        if (j + 2 < L) S[j + 2] = e;
      }
    }
  }
}

// Face flux (f and h) out of the owner at logical (lx, ly) along axis AX with
// stencil step sg and speed v = |e_axis|, from the staged direction V: the
// wall closure at solid owners, else WENO5/NND on the staged line. A tile
// whose region holds no solid cell (`solids` false) takes the line as
// staged, on a path of its own (no extrapolation state to carry).
template <int AX, typename Sh, typename TC>
__device__ __forceinline__ void face_flux(const FH<TC>* V, const Region<Sh>& R, const double* alpha, int method,
                                          bool solids, int lx, int ly, int sg, TC v, TC rf, TC rh, TC& vf,
                                          TC& vh) {
#pragma clang fp contract(off)  // the wall faces round as the strip kernels' do
  const int dx = AX == 0 ? sg : 0, dy = AX == 0 ? 0 : sg;
  if (!solids) {
    if (method == 1) {
      FH<TC> w[5];
#pragma unroll
      for (int k = 0; k < 5; k++) w[k] = V[(ly + (k - 2) * dy + kH) * Sh::RX + lx + (k - 2) * dx + kH];
      vf = weno5(v, w[0].f, w[1].f, w[2].f, w[3].f, w[4].f);
      vh = weno5(v, w[0].h, w[1].h, w[2].h, w[3].h, w[4].h);
    } else {
      FH<TC> w[3];
#pragma unroll
      for (int k = 0; k < 3; k++) w[k] = V[(ly + (k - 1) * dy + kH) * Sh::RX + lx + (k - 1) * dx + kH];
      vf = nnd(v, w[0].f, w[1].f, w[2].f);
      vh = nnd(v, w[0].h, w[1].h, w[2].h);
    }
    return;
  }
  const unsigned mo = R.meta[ly + kH][lx + kH];
  if ((mo & kMetaType) == 2) {
    // wall closure alpha |e| f_wall; staged deviations dev = f - ref give
    // the deviation of that face, |e| (alpha dev + (alpha - 1) ref), since
    // the |e| ref of every face cancels in the net flux (rf = rh = 0 for
    // FP64 storage)
    const TC a = TC(alpha[R.off[ly + kH][lx + kH]]);
    const FH<TC> w = V[(ly + kH) * Sh::RX + lx + kH];
    if constexpr (sizeof(TC) == 4) {
      vf = v * (a * w.f + (a - TC(1)) * rf);
      vh = v * (a * w.h + (a - TC(1)) * rh);
    } else {
      (void)rf;
      (void)rh;
      vf = a * v * w.f;
      vh = a * v * w.h;
    }
    return;
  }
  if (method == 1) {
    TC Vf[5], Vh[5];
    int sld[5], any = 0;
#pragma unroll
    for (int k = 0; k < 5; k++) {
      const int px = lx + (k - 2) * dx + kH, py = ly + (k - 2) * dy + kH;
      const FH<TC> w = V[py * Sh::RX + px];
      Vf[k] = w.f;
      Vh[k] = w.h;
      sld[k] = (R.meta[py][px] & kMetaType) == 2;
      any |= sld[k];
    }
    if (any) {
      TC Ef[5], Eh[5];
#pragma unroll
      for (int k = 0; k < 5; k++) { Ef[k] = Vf[k]; Eh[k] = Vh[k]; }
      extrapolate<5>(sld, Vf, Ef);
      extrapolate<5>(sld, Vh, Eh);
#pragma unroll
      for (int k = 0; k < 5; k++) { Vf[k] = Ef[k]; Vh[k] = Eh[k]; }
    }
    vf = weno5(v, Vf[0], Vf[1], Vf[2], Vf[3], Vf[4]);
    vh = weno5(v, Vh[0], Vh[1], Vh[2], Vh[3], Vh[4]);
  } else {
    TC Vf[3], Vh[3];
    int sld[3], any = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const int px = lx + (k - 1) * dx + kH, py = ly + (k - 1) * dy + kH;
      const FH<TC> w = V[py * Sh::RX + px];
      Vf[k] = w.f;
      Vh[k] = w.h;
      sld[k] = (R.meta[py][px] & kMetaType) == 2;
      any |= sld[k];
    }
    if (any) {
      TC Ef[3], Eh[3];
#pragma unroll
      for (int k = 0; k < 3; k++) { Ef[k] = Vf[k]; Eh[k] = Vh[k]; }
      extrapolate<3>(sld, Vf, Ef);
      extrapolate<3>(sld, Vh, Eh);
#pragma unroll
      for (int k = 0; k < 3; k++) { Vf[k] = Ef[k]; Vh[k] = Eh[k]; }
    }
    vf = nnd(v, Vf[0], Vf[1], Vf[2]);
    vh = nnd(v, Vh[0], Vh[1], Vh[2]);
  }
}

// ---------------------------------------------------------------------------
// k_flux<MODE>: net flux of every direction for the tile's cells.
//   MODE 0: source f2 (f1 for RK1); dF, dH -> the dF planes.
//   MODE 1: source f3; new state f3 + (dF2 - dF)/2, h3 + (dH2 - dH)/2
//           from the stored stage-2 net fluxes dF2, dH2 (FP32 storage:
//           from G, which stage 3 writes, G - dF/2, G - dH/2).
//   MODE 2: source the current state (EULER, EULER_EQ); dF, dH -> dF planes.
// MOM (MODE 1, FP32 storage): also the new state's conserved moments, those
// stage 3 wrote (of f - dF2/2) minus half those of each fluid cell's dF(f3)
// over the directions -> the moment planes of the new state buffer.
//
// FP32 storage: directions alternate between two staging buffers filled by
// direct-to-LDS loads; while the threads compute direction I's faces from
// one, the loads of direction I+1 stream into the other. FP64 storage: one
// buffer, written from registers; outside the moment kernel the loads of
// direction I+1 go out into registers once direction I's faces are done,
// so their latency hides behind direction I's net fluxes.
// ---------------------------------------------------------------------------


// ---------------------------------------------------------------------------
// The reference's physical stencil addressing (cpuLBM.cpp:1171-1326,
// 1436-1470, 1502-1590), for the wall closure and for the tiles whose cells
// the tile kernels' logical addressing would get wrong (next to solids one
// cell thick or near a mirror edge): there each own cell's net flux is
// formed from global memory, face by face, as the reference forms it.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int ref_index(const InstDev& in, int i, int pm, int d, int& mir) {
  unsigned m, cl;
  int r;
  if (d == 0)
    r = map_axis(i + pm, in.nx, in.flags & kPeriodicX, in.flags & kMirrorW, in.flags & kMirrorE, m, cl);
  else
    r = map_axis(i + pm, in.ny, in.flags & kPeriodicY, in.flags & kMirrorS, in.flags & kMirrorN, m, cl);
  mir = (int)m;
  return r;
}

// Value of plane k at offset o in FP64 (FP32 storage: deviation + reference).
template <typename TS>
__device__ __forceinline__ double full_val(const InstDev& in, const TS* src, int k, long long o) {
  if constexpr (sizeof(TS) == 4) return double(src[in.plane * k + o]) + in.ref[k];
  else return double(src[in.plane * k + o]);
}

// Face flux of distribution k0 + i (k0 = 0: f, kQ: h) out of fluid cell
// (x, y) along `axis`, with the reference's physical addressing: the
// stencil line through the mapped neighbours (direction remapped at a
// mirror), its first solid run extrapolated from the two nearest fluid
// points of the line in direction i; val(k, o) gives the FP64 value of
// plane k at offset o.
template <typename TC, typename Val>
__device__ TC ref_line_flux(const InstDev& in, const Val& val, int k0, int x, int y, int i, int axis) {
  const int e = axis == 0 ? c_lat_x[i] : c_lat_y[i];
  const int s = e > 0 ? 1 : (e < 0 ? -1 : 0);
  if (s == 0) return TC(0);
  const int L = in.flux_method == 1 ? 5 : 3;
  const int low = (L - 1) / 2;
  const int start = (axis == 0 ? x : y) - s * low;
  TC V[5];
  int sld[5], any = 0;
  for (int j = 0; j < L; j++) {
    int mir;
    const int p = ref_index(in, start, s * j, axis, mir);
    const int d = mir ? (axis == 0 ? c_mir_ew[i] : c_mir_ns[i]) : i;
    const long long o = axis == 0 ? (long long)y * in.pitch + p : (long long)p * in.pitch + x;
    V[j] = TC(val(k0 + d, o));
    sld[j] = in.ctype[o] == 2;
    any |= sld[j];
  }
  if (any) {
    int diff = 0, jj = 0, pos = 0, mir;
    for (int j = 0; j < L - 1; j++) {
      diff = sld[j] - sld[j + 1];
      if (diff > 0) { jj = j; pos = ref_index(in, start, s * j, axis, mir); break; }
      if (diff < 0) { jj = j + 1; pos = ref_index(in, start, s * (j + 1), axis, mir); break; }
    }
    const int m1 = ref_index(in, pos, s * diff, axis, mir);
    const int m2 = ref_index(in, pos, s * 2 * diff, axis, mir);
    const long long o1 = axis == 0 ? (long long)y * in.pitch + m1 : (long long)m1 * in.pitch + x;
    const long long o2 = axis == 0 ? (long long)y * in.pitch + m2 : (long long)m2 * in.pitch + x;
    V[jj] = TC(2) * TC(val(k0 + i, o1)) - TC(val(k0 + i, o2));
    const int nx = jj - diff;
    if (nx >= 0 && nx < L) V[nx] = V[jj];
  }
  const TC v = TC(double(e > 0 ? e : -e) * in.c);
  if (L == 5) return weno5(v, V[0], V[1], V[2], V[3], V[4]);
  return nnd(v, V[0], V[1], V[2]);
}

// Face fluxes (f, h) of direction i out of cell (x, y) along `axis`: the
// wall closure alpha |e| f at a wall owner, else the reference line.
template <typename Val>
__device__ void phys_face(const InstDev& in, const Val& val, int x, int y, int i, int axis, double& Ff,
                          double& Fh) {
  const long long o = (long long)y * in.pitch + x;
  if (in.ctype[o] == 2) {
    const int e = axis == 0 ? c_lat_x[i] : c_lat_y[i];
    const double av = in.alpha[o] * (double(e > 0 ? e : -e) * in.c);
    Ff = av * val(i, o);
    Fh = av * val(kQ + i, o);
    return;
  }
  Ff = ref_line_flux<double>(in, val, 0, x, y, i, axis);
  Fh = ref_line_flux<double>(in, val, kQ, x, y, i, axis);
}

// Net flux dF, dH of direction I out of cell (x, y) with the reference's
// addressing (cFLUX, cpuLBM.cpp:1436-1470): out-faces of the cell, in-faces
// of the upwind owner mapped through the boundary rules (the mirrored
// direction across a mirror edge; the cell itself across a clamped one).
// FP64 from the full values; combined as dir_net combines.
template <typename TS, typename TC>
__device__ void phys_net(const InstDev& in, const TS* src, int x, int y, int I, double rdx, double rdy, TC& dF,
                         TC& dH) {
  const auto val = [&](int k, long long o) { return full_val(in, src, k, o); };
  const int ex = c_lat_x[I], ey = c_lat_y[I];
  double oxf = 0, oxh = 0, ixf = 0, ixh = 0, oyf = 0, oyh = 0, iyf = 0, iyh = 0;
  if (ex) {
    phys_face(in, val, x, y, I, 0, oxf, oxh);
    unsigned m, cl;
    const int xu = map_axis(x - (ex > 0 ? 1 : -1), in.nx, in.flags & kPeriodicX, in.flags & kMirrorW,
                            in.flags & kMirrorE, m, cl);
    if (cl) {
      ixf = oxf;
      ixh = oxh;
    } else {
      phys_face(in, val, xu, y, m ? c_mir_ew[I] : I, 0, ixf, ixh);
    }
  }
  if (ey) {
    phys_face(in, val, x, y, I, 1, oyf, oyh);
    unsigned m, cl;
    const int yu = map_axis(y - (ey > 0 ? 1 : -1), in.ny, in.flags & kPeriodicY, in.flags & kMirrorS,
                            in.flags & kMirrorN, m, cl);
    if (cl) {
      iyf = oyf;
      iyh = oyh;
    } else {
      phys_face(in, val, x, yu, m ? c_mir_ns[I] : I, 1, iyf, iyh);
    }
  }
  double f, h;
  if (ex && ey) {
    f = (rdx * oxf + rdy * oyf) - (rdx * ixf + rdy * iyf);
    h = (rdx * oxh + rdy * oyh) - (rdx * ixh + rdy * iyh);
  } else if (ex) {
    f = rdx * oxf - rdx * ixf;
    h = rdx * oxh - rdx * ixh;
  } else {
    f = rdy * oyf - rdy * iyf;
    h = rdy * oyh - rdy * iyh;
  }
  dF = TC(f);
  dH = TC(h);
}

// Per-thread state of one tile pass. The instance fields the direction loop
// needs are copied here once per tile: read through the instance table
// inside the loop (after the loop's stores) they would be vector loads,
// whose wait would also wait for the staging loads in flight.
template <typename TS, typename TC, typename Sh>
struct TileWork {
  bool fluid[Sh::Cells];
  int oo[Sh::Cells];
  bool solids;
  int method;
  const void* src;
  gptr<TS> dst;
  gptr<const TS> dfl;  // MODE 1: the stored stage-2 net fluxes
  const double* alpha;
  long long P;
  double c;
};

// Direction I of a tile.
//   DB (two staging buffers): faces from cur, landed by the barrier on
//   entry, while the loads of direction I+1 (`next`) stream into nxt.
//   One buffer (cur and nxt the same): direction I's values go from
//   registers to cur on entry; they were loaded into PV while direction
//   I-1 updated its cells (with the moment sums: loaded here, on entry),
//   and PV takes the loads of direction I+1 once the faces are done.
template <typename TS, typename TC, int MODE, bool MOM, bool DB, typename Sh>
__device__ __forceinline__ void flux_dir(FluxStage<Sh, TC, MOM>& S, const Region<Sh>& R, const InstDev& in,
                                         const TileWork<TS, TC, Sh>& W, const SlotSet<Sh>& L, SlotVals<Sh, TS>& PV,
                                         StageBuf<Sh, TS>& cur, StageBuf<Sh, TS>& nxt, int I, bool next) {
  const int t = threadIdx.x;
  const long long P = W.P;
  // with no loads in flight across the faces the instance fields are read
  // where used, which keeps them out of the registers the faces need
  const double c = DB ? W.c : in.c;
  const int method = DB ? W.method : in.flux_method;
  const double* alpha = DB ? W.alpha : in.alpha;
  const int ex = c_lat_x[I], ey = c_lat_y[I];
  const int sx = ex > 0 ? 1 : (ex < 0 ? -1 : 0), sy = ey > 0 ? 1 : (ey < 0 ? -1 : 0);
  // reference equilibrium of direction I for the wall faces of FP32 storage
  // (a mirror image of I has the same reference: it depends on |e| only)
  const TC rf = sizeof(TS) == 4 ? TC(R.ref[I]) : TC(0), rh = sizeof(TS) == 4 ? TC(R.ref[kQ + I]) : TC(0);
  if constexpr (!DB) {  // cur was last read before the previous barrier
    if constexpr (MOM) {
      SlotVals<Sh, TS> V;
      fetch_slots<TS>(V, L, W.src, P, I);
      store_slots<TS>(cur, V, L, I);
    } else {
      store_slots<TS>(cur, PV, L, I);
    }
  }
  __syncthreads();  // cur complete; fx, fy (DB: nxt) free
  if (DB && next) stage_dir<TS>(nxt, R, W.src, P, I + 1);
  // faces: x face j of row r (owner column j - [sx > 0]); y face j of
  // column c (owner row j - [sy > 0])
  if (ex && t < Sh::NXF) {
    const int row = t / (Sh::TX + 1), j = t % (Sh::TX + 1);
    const TC v = TC(double(ex > 0 ? ex : -ex) * c);
    TC vf, vh;
    face_flux<0>(cur.v, R, alpha, method, W.solids, j - (sx > 0 ? 1 : 0), row, sx, v, rf, rh, vf, vh);
    S.fx[row][j] = FH<TC>{vf, vh};
  }
  if (ey && t < Sh::NYF) {
    const int col = t % Sh::TX, j = t / Sh::TX;
    const TC v = TC(double(ey > 0 ? ey : -ey) * c);
    TC vf, vh;
    face_flux<1>(cur.v, R, alpha, method, W.solids, col, j - (sy > 0 ? 1 : 0), sy, v, rf, rh, vf, vh);
    S.fy[j][col] = FH<TC>{vf, vh};
  }
  // MODE 1 (FP64 storage): the own cells' f3 of direction I, taken from cur
  // before the barrier after which (one buffer) the next direction may
  // overwrite it
  FH<TC> own[Sh::Cells];
  if constexpr (MODE == 1 && !kStage3G<TS>) {
#pragma unroll
    for (int q = 0; q < Sh::Cells; q++) {
      const int ci = t + q * Sh::NT;
      own[q] = W.fluid[q] ? cur.v[(ci / Sh::TX + kH) * Sh::RX + ci % Sh::TX + kH] : FH<TC>{TC(0), TC(0)};
    }
  }
  if constexpr (DB) __builtin_amdgcn_s_waitcnt(kWaitVmem);  // this wave's loads into nxt landed
  __syncthreads();
  // One buffer without the moment sums: the loads of direction I+1 go out
  // now and land in registers while the cells update (written to LDS when
  // I+1 starts; stage-3 flux 5-6 % faster, while the moment-summing kernel,
  // short of registers, loads at the start of each direction instead).
  // MODE 1 first issues the loads of the stored dF(f2) (FP32 storage: of
  // G), which it needs sooner.
  TS d2f[Sh::Cells], d2h[Sh::Cells];
  if constexpr (MODE == 1) {
    const gptr<const TS> b = kStage3G<TS> ? (gptr<const TS>)W.dst : W.dfl;
    const long long pf = kStage3G<TS> ? P * I : P * (I - 1), ph = kStage3G<TS> ? P * (kQ + I) : P * (12 + I - 1);
#pragma unroll
    for (int q = 0; q < Sh::Cells; q++) {
      d2f[q] = W.fluid[q] ? b[pf + W.oo[q]] : TS(0);
      d2h[q] = W.fluid[q] ? b[ph + W.oo[q]] : TS(0);
    }
  }
  if constexpr (!DB && !MOM) {
    if (next) fetch_slots<TS>(PV, L, W.src, P, I + 1);
  }
#pragma unroll
  for (int q = 0; q < Sh::Cells; q++) {
    // no contraction: the net flux and the state update round as the strip
    // kernels (and the oracle) round them, so both paths agree bit for bit
#pragma clang fp contract(off)
    if (!W.fluid[q]) continue;
    const int ci = t + q * Sh::NT, tx = ci % Sh::TX, ty = ci / Sh::TX;
    // reference cFLUX (cpuLBM.cpp:1436-1470): (dt/dx)(Fx_out - Fx_in) +
    // (dt/dy)(Fy_out - Fy_in), grouped out - in; an upwind owner across a
    // clamped edge is the cell itself. On a non-uniform grid dt/dx and
    // dt/dy are the cell's column and row values.
    TC rdx, rdy;
    if constexpr (sizeof(TC) == 8) {
      rdx = TC(R.rdx[tx]);
      rdy = TC(R.rdy[ty]);
    } else {
      rdx = TC(R.rdxf[tx]);
      rdy = TC(R.rdyf[ty]);
    }
    TC oxf = 0, oxh = 0, ixf = 0, ixh = 0, oyf = 0, oyh = 0, iyf = 0, iyh = 0;
    if (ex) {
      const int lo = sx > 0 ? -1 : 0;
      const FH<TC> out = S.fx[ty][tx - lo];
      oxf = out.f;
      oxh = out.h;
      const int ux = tx - sx;
      if (R.meta[ty + kH][ux + kH] & kMetaClampX) {
        ixf = oxf;
        ixh = oxh;
      } else {
        const FH<TC> in_ = S.fx[ty][ux - lo];
        ixf = in_.f;
        ixh = in_.h;
      }
    }
    if (ey) {
      const int lo = sy > 0 ? -1 : 0;
      const FH<TC> out = S.fy[ty - lo][tx];
      oyf = out.f;
      oyh = out.h;
      const int uy = ty - sy;
      if (R.meta[uy + kH][tx + kH] & kMetaClampY) {
        iyf = oyf;
        iyh = oyh;
      } else {
        const FH<TC> in_ = S.fy[uy - lo][tx];
        iyf = in_.f;
        iyh = in_.h;
      }
    }
    TC dF, dH;
    if (ex && ey) {
      dF = (rdx * oxf + rdy * oyf) - (rdx * ixf + rdy * iyf);
      dH = (rdx * oxh + rdy * oyh) - (rdx * ixh + rdy * iyh);
    } else if (ex) {
      dF = rdx * oxf - rdx * ixf;
      dH = rdx * oxh - rdx * ixh;
    } else {
      dF = rdy * oyf - rdy * iyf;
      dH = rdy * oyh - rdy * iyh;
    }
    const long long o = W.oo[q];
    if (MODE != 1) {
      W.dst[P * (I - 1) + o] = TS(dF);
      W.dst[P * (12 + I - 1) + o] = TS(dH);
      if constexpr (MOM) {
        const double dFs = double(TS(dF));
        S.mom[0][ci] += dFs;
        S.mom[1][ci] += dFs * (ex * c);
        S.mom[2][ci] += dFs * (ey * c);
        S.mom[3][ci] += double(TS(dH));
      }
    } else {
      // the new state f3 + (dF2 - dF3)/2 (FP32 storage: G - dF3/2), the
      // final combination of k_s3
      if constexpr (kStage3G<TS>) {
        W.dst[P * I + o] = TS(TC(d2f[q]) - TC(0.5) * dF);
        W.dst[P * (kQ + I) + o] = TS(TC(d2h[q]) - TC(0.5) * dH);
      } else {
        W.dst[P * I + o] = TS(own[q].f + TC(0.5) * (TC(d2f[q]) - dF));
        W.dst[P * (kQ + I) + o] = TS(own[q].h + TC(0.5) * (TC(d2h[q]) - dH));
      }
      if constexpr (MOM) {  // conserved moments of dF(f3), for the FP64 moment planes
        const double dFs = double(dF);
        S.mom[0][ci] += dFs;
        S.mom[1][ci] += dFs * (ex * c);
        S.mom[2][ci] += dFs * (ey * c);
        S.mom[3][ci] += double(dH);
      }
    }
  }
}

// Staging per storage type, as measured on the two workloads (4096^2
// Riemann, the 256-member UQ ensemble): FP32 two buffers with loads straight
// into LDS (stage-3 flux 3.85 -> 3.56 ms on the ensemble); FP64 one buffer
// through registers, which is as fast or faster there (4.75-4.77 against
// 4.87 ms) and leaves room for three 8-wave blocks per CU with the moment
// sums.
template <typename TS, bool MOM>
struct TwoBuffers {
  static constexpr bool value = sizeof(TS) != 8;
};

template <typename TS, typename TC, int MODE, bool MOM, typename Sh>
__device__ __forceinline__ void flux_tile(FluxStage<Sh, TC, MOM>& S, StageBuf<Sh, TS>& VA, StageBuf<Sh, TS>& VB,
                                          const Region<Sh>& R, const InstDev& in, const TileRef& tr, int par) {
  static_assert(std::is_same<TS, TC>::value, "LDS holds the stored values as computed on");
  static_assert((kQ - 1) % 2 == 0, "directions pair up over the two staging buffers");
  constexpr bool DB = TwoBuffers<TS, MOM>::value;
  const int t = threadIdx.x;
  TileWork<TS, TC, Sh> W;
  W.solids = tr.flags & kTileSolid;
  W.method = in.flux_method;
  W.alpha = in.alpha;
  W.c = in.c;
  W.P = in.plane;
  W.src = MODE == 0 ? in.f2 : (MODE == 1 ? in.f3 : in.state[in.parity0 ^ par]);
  W.dst = as_global(MODE == 1 ? reinterpret_cast<TS*>(in.state[1 ^ in.parity0 ^ par]) : reinterpret_cast<TS*>(in.dflux));
  W.dfl = as_global(reinterpret_cast<const TS*>(in.dflux));
  // own cells of this thread: cell ci = t + q NT at (ci % TX, ci / TX)
#pragma unroll
  for (int q = 0; q < Sh::Cells; q++) {
    const int ci = t + q * Sh::NT, tx = ci % Sh::TX, ty = ci / Sh::TX;
    const bool own = ci < Sh::Own && tr.tx * Sh::TX + tx < in.nx && tr.ty * Sh::TY + ty < in.ny;
    W.fluid[q] = own && (R.meta[ty + kH][tx + kH] & kMetaType) == 0;
    W.oo[q] = own ? R.off[ty + kH][tx + kH] : 0;
    if (MOM && ci < Sh::Own) {
#pragma unroll
      for (int k = 0; k < 4; k++) S.mom[k][ci] = 0.0;
    }
  }
  SlotSet<Sh> L;
  SlotVals<Sh, TS> PV;
  if constexpr (DB) {
    stage_dir<TS>(VA, R, W.src, W.P, 1);
    __builtin_amdgcn_s_waitcnt(kWaitVmem);
    for (int I = 1; I < kQ; I += 2) {
      flux_dir<TS, TC, MODE, MOM, true>(S, R, in, W, L, PV, VA, VB, I, true);
      flux_dir<TS, TC, MODE, MOM, true>(S, R, in, W, L, PV, VB, VA, I + 1, I + 2 < kQ);
    }
  } else {
    setup_slots(L, R);
    if constexpr (!MOM) fetch_slots<TS>(PV, L, W.src, W.P, 1);
    for (int I = 1; I < kQ; I++) flux_dir<TS, TC, MODE, MOM, false>(S, R, in, W, L, PV, VA, VA, I, I + 1 < kQ);
  }
  if constexpr (MOM) {
    // the new state's conserved moments: those written by stage 3 (of
    // f - dF(f2)/2) minus half those of dF(f3)
    static_assert(MODE == 1, "moment sums belong to the stage-3 flux of FP32 storage");
    const gptr<double> mo = as_global(in.mom[1 ^ in.parity0 ^ par]);
#pragma unroll
    for (int q = 0; q < Sh::Cells; q++) {
      if (!W.fluid[q]) continue;
      const int ci = t + q * Sh::NT;
      const long long o = W.oo[q];
#pragma unroll
      for (int k = 0; k < 4; k++) mo[k * W.P + o] -= 0.5 * S.mom[k][ci];
    }
  }
}

template <typename TS, typename TC, int MODE, bool MOM, typename Sh>
__device__ __forceinline__ void flux_tiles(const InstDev* insts, const TileRef* tiles, int ntiles, int par,
                                           int step, Region<Sh>& R, FluxStage<Sh, TC, MOM>& S,
                                           StageBuf<Sh, TS>& VA, StageBuf<Sh, TS>& VB) {
  for (int t = first_tile(); t < ntiles; t += gridDim.x) {
    const TileRef tr = tiles[t];
    const InstDev& in = insts[tr.inst];
    if (frozen_by_failure(in, step)) continue;
    if (tr.flags & kTilePhysical) continue;  // k_phys forms this tile's net fluxes
    if (!MOM && in.strips) continue;         // the strip kernels form this instance's net fluxes
    __syncthreads();
    setup_region(R, in, tr.tx * Sh::TX, tr.ty * Sh::TY);
    __syncthreads();
    flux_tile<TS, TC, MODE, MOM>(S, VA, VB, R, in, tr, par);
  }
}


// k_phys<MODE>: the flux kernel's work for the tiles flagged kTilePhysical,
// whose net fluxes the logical addressing of the staged tiles would get
// wrong: each thread takes an own fluid cell and forms its twelve net
// fluxes with the reference's addressing (phys_net), then writes what the
// flux kernel writes for the cell (MODE 0 and 2: the dF planes; MODE 1: the
// new state f3 + (dF2 - dF3)/2, FP32 storage G - dF3/2, with MOM the moment
// planes' update).
// Launched after the flux kernel, only for groups with such tiles; the flux
// kernel skips them.
template <typename TS, typename TC, int MODE, bool MOM, typename Sh>
__global__ __launch_bounds__(256) void k_phys(const InstDev* insts, const TileRef* tiles, int ntiles, int par,
                                              const int* stepp) {
  const int step = *stepp;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const TileRef tr = tiles[t];
    if (!(tr.flags & kTilePhysical)) continue;
    const InstDev& in = insts[tr.inst];
    if (frozen_by_failure(in, step)) continue;
    const long long P = in.plane;
    const int cur = in.parity0 ^ par, nxt = 1 ^ cur;
    const TS* src = reinterpret_cast<const TS*>(MODE == 0 ? in.f2 : (MODE == 1 ? in.f3 : in.state[cur]));
    TS* dst = reinterpret_cast<TS*>(MODE == 1 ? in.state[nxt] : in.dflux);
    for (int c = threadIdx.x; c < Sh::Own; c += blockDim.x) {
      const int x = tr.tx * Sh::TX + c % Sh::TX, y = tr.ty * Sh::TY + c / Sh::TX;
      if (x >= in.nx || y >= in.ny) continue;
      const long long o = (long long)y * in.pitch + x;
      if (in.ctype[o] != 0) continue;
      const double rdx = in.rdxs ? in.rdxs[x] : in.rdx, rdy = in.rdys ? in.rdys[y] : in.rdy;
      double m[4] = {0, 0, 0, 0};
      for (int I = 1; I < kQ; I++) {
        TC dF, dH;
        phys_net<TS, TC>(in, src, x, y, I, rdx, rdy, dF, dH);
        const double ex = c_lat_x[I] * in.c, ey = c_lat_y[I] * in.c;
        double vf, vh;
        if (MODE != 1) {
          dst[P * (I - 1) + o] = TS(dF);
          dst[P * (12 + I - 1) + o] = TS(dH);
          vf = double(TS(dF));
          vh = double(TS(dH));
        } else {
          if constexpr (kStage3G<TS>) {
            dst[P * I + o] = TS(TC(dst[P * I + o]) - TC(0.5) * dF);
            dst[P * (kQ + I) + o] = TS(TC(dst[P * (kQ + I) + o]) - TC(0.5) * dH);
          } else {
            const TS* d2 = reinterpret_cast<const TS*>(in.dflux);
            dst[P * I + o] = TS(TC(src[P * I + o]) + TC(0.5) * (TC(d2[P * (I - 1) + o]) - dF));
            dst[P * (kQ + I) + o] = TS(TC(src[P * (kQ + I) + o]) + TC(0.5) * (TC(d2[P * (12 + I - 1) + o]) - dH));
          }
          vf = double(dF);
          vh = double(dH);
        }
        m[0] += vf;
        m[1] += vf * ex;
        m[2] += vf * ey;
        m[3] += vh;
      }
      if constexpr (MOM) {
        static_assert(MODE == 1, "moment sums belong to the stage-3 flux of FP32 storage");
        for (int k = 0; k < 4; k++) in.mom[nxt][k * P + o] -= 0.5 * m[k];
      }
    }
  }
}

// FP32 storage: capped at 64 registers for 8 waves per SIMD (4 blocks per
// CU; with the stage-3 moment sums 3 spilled slots, mixed-ensemble stage-3
// flux 3.35 -> 3.18 ms). FP64: capped at 80 registers for 6 waves per SIMD.
// (The staging buffers are separate LDS variables so that the compiler sees
// that reads of one do not wait for the loads into the other; the second
// is dropped where it is unused.)
template <typename TS, typename TC, int MODE, bool MOM, typename Sh>
__global__ __launch_bounds__(Sh::NT) __attribute__((amdgpu_waves_per_eu(8))) void k_flux(
    const InstDev* __restrict__ insts, const TileRef* __restrict__ tiles, int ntiles, int par, const int* stepp) {
  __shared__ Region<Sh> R;
  __shared__ FluxStage<Sh, TC, MOM> S;
  __shared__ StageBuf<Sh, TS> VA;
  __shared__ StageBuf<Sh, TS> VB;
  flux_tiles<TS, TC, MODE, MOM>(insts, tiles, ntiles, par, *stepp, R, S, VA, VB);
}

template <int MODE, bool MOM, typename Sh>
__global__ __launch_bounds__(Sh::NT) __attribute__((amdgpu_waves_per_eu(6))) void k_flux64(
    const InstDev* __restrict__ insts, const TileRef* __restrict__ tiles, int ntiles, int par, const int* stepp) {
  __shared__ Region<Sh> R;
  __shared__ FluxStage<Sh, double, MOM> S;
  __shared__ StageBuf<Sh, double> VA;
  __shared__ StageBuf<Sh, double> VB;
  flux_tiles<double, double, MODE, MOM>(insts, tiles, ntiles, par, *stepp, R, S, VA, VB);
}

#include "lbm_strip.hpp"

// The flux kernel of a storage/compute type pair and tile shape.
template <typename TS, typename TC, int MODE, bool MOM, typename Sh>
struct FluxKernel {
  static constexpr auto fn = &k_flux<TS, TC, MODE, MOM, Sh>;
};
template <int MODE, bool MOM, typename Sh>
struct FluxKernel<double, double, MODE, MOM, Sh> {
  static constexpr auto fn = &k_flux64<MODE, MOM, Sh>;
};

// ---------------------------------------------------------------------------
// k_alpha: wall closure factor of each wall cell with a non-solid axis
// neighbour, from the neighbours' face fluxes of `which` (0: f2 or f1,
// 1: f3, 2: the current state),
// following the reference's physical stencil addressing (cpuLBM.cpp:1171-1326,
// 1502-1590).
// ---------------------------------------------------------------------------
__constant__ int c_wall_dx[4] = {-1, 1, 0, 0};
__constant__ int c_wall_dy[4] = {0, 0, -1, 1};
__constant__ int c_wall_in[4][4] = {{1, 5, 8, 9}, {3, 6, 7, 11}, {2, 5, 6, 10}, {4, 7, 8, 12}};
__constant__ int c_wall_out[4][4] = {{3, 6, 7, 11}, {1, 5, 8, 9}, {4, 8, 7, 12}, {2, 6, 5, 10}};

template <typename TS, typename TC>
__global__ void k_alpha(const InstDev* insts, const WallRef* walls, int nwalls, int which, int par,
                        const int* stepp) {
  const int step = *stepp;
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nwalls) return;
  const WallRef wr = walls[w];
  const InstDev& in = insts[wr.inst];
  if (frozen_by_failure(in, step)) return;
  const void* srcv = which == 0 ? in.f2 : (which == 1 ? in.f3 : in.state[in.parity0 ^ par]);
  const TS* src = reinterpret_cast<const TS*>(srcv);
  const auto fv = [&](int k, long long o) { return full_val(in, src, k, o); };
  const long long ow = (long long)wr.y * in.pitch + wr.x;
  double in_f = 0, out_f = 0;
  for (int j = 0; j < 4; j++) {
    int mx, my;
    const int xn = ref_index(in, wr.x, c_wall_dx[j], 0, my);
    const int yn = ref_index(in, wr.y, c_wall_dy[j], 1, mx);
    const long long on = (long long)yn * in.pitch + xn;
    if (in.ctype[on] >= 2) continue;
    const int axis = j < 2 ? 0 : 1;
    // the wall cell's own spacing: its axis neighbours share its row (x
    // faces) or column (y faces), so the face-length weights of a
    // non-uniform grid reduce to it
    const double dd = axis == 0 ? (in.dxs ? in.dxs[wr.x] : in.dx) : (in.dys ? in.dys[wr.y] : in.dy);
    for (int k = 0; k < 4; k++) {
      const int i = c_wall_in[j][k], io = c_wall_out[j][k];
      const double F = double(ref_line_flux<TC>(in, fv, 0, xn, yn, i, axis));
      in_f += F / dd;
      const double vv = fabs(double(axis == 0 ? c_lat_x[io] : c_lat_y[io]) * in.c);
      const double fw = full_val(in, src, io, ow);
      out_f += fw * vv / dd;
    }
  }
  in.alpha[ow] = in_f / out_f;
}

// ---------------------------------------------------------------------------
// Initialisation (reference initialiseFunctions, cpuLBM.cpp:846-875) and
// macroscopic output (macroPropShort, cpuLBM.cpp:877-914).
// ---------------------------------------------------------------------------
template <typename TS, int... Is>
__device__ __forceinline__ void init_cell(const InstDev& in, TS* dst, long long o, const double* pr, Seq<Is...>) {
  Macro m;
  m.rho = pr[0];
  m.ux = pr[1];
  m.uy = pr[2];
  m.T = pr[3];
  const EqIn<double> q = eq_inputs<double>(in, m);
  double t[4] = {0, 0, 0, 0};
  auto one = [&](auto Ic) {
    constexpr int I = decltype(Ic)::value;
    double fe, he;
    eq_dir<I>(q, fe, he);
    t[0] += fe;
    t[1] += fe * (lat_x(I) * in.c);
    t[2] += fe * (lat_y(I) * in.c);
    t[3] += he;
    dst[in.plane * I + o] = to_store<TS>(in, I, fe);
    dst[in.plane * (kQ + I) + o] = to_store<TS>(in, kQ + I, he);
  };
  (one(std::integral_constant<int, Is>{}), ...);
  if constexpr (sizeof(TS) == 4) {
#pragma unroll
    for (int k = 0; k < 4; k++) in.mom[in.parity0][in.plane * k + o] = t[k];
  }
}

// props: rows of {rho, ux, uy, T}; ids: [ny][pitch]
template <typename TS>
__global__ void k_init(const InstDev* insts, int inst, const int* ids, const double* props) {
  const InstDev& in = insts[inst];
  const long long n = (long long)in.ny * in.pitch;
  for (long long o = blockIdx.x * (long long)blockDim.x + threadIdx.x; o < n; o += (long long)gridDim.x * blockDim.x) {
    const int x = (int)(o % in.pitch);
    if (x >= in.nx) continue;
    TS* dst = reinterpret_cast<TS*>(in.state[in.parity0]);
    init_cell<TS>(in, dst, o, props + 4 * ids[o], MakeSeq<kQ>::type{});
  }
}

// out: 5 planes of double [ny][pitch]: rho, ux, uy, p, T
template <typename TS>
__global__ void k_macro(const InstDev* insts, int inst, int par, double* outp) {
  const InstDev& in = insts[inst];
  const long long n = (long long)in.ny * in.pitch;
  const TS* src = reinterpret_cast<const TS*>(in.state[in.parity0 ^ par]);
  for (long long o = blockIdx.x * (long long)blockDim.x + threadIdx.x; o < n; o += (long long)gridDim.x * blockDim.x) {
    const int x = (int)(o % in.pitch);
    if (x >= in.nx) continue;
    CellState<TS> cs;
    load_state(in, src, in.mom[in.parity0 ^ par], o, cs);
    Macro m;
    finish_macro(in, cs.s[0], cs.s[1], cs.s[2], cs.s[3], m);
    outp[o] = m.rho;
    outp[n + o] = m.ux;
    outp[2 * n + o] = m.uy;
    outp[3 * n + o] = m.p;
    outp[4 * n + o] = m.T;
  }
}

// ---------------------------------------------------------------------------
// Conserved totals of one instance over its fluid cells: sum f (mass),
// sum f e_x c, sum f e_y c (momentum), sum h (total energy), FP64. A
// deterministic two-pass reduction: every block reduces its cells with
// 64-lane wavefront shuffles and writes one partial per quantity; one block
// then sums the partials in a fixed order.
// ---------------------------------------------------------------------------
constexpr int kRedBlocks = 512;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Sums v over the block (256 threads = 4 wavefronts); the result is valid in
// thread 0.
__device__ __forceinline__ double block_sum(double v, double* lds4) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) lds4[w] = v;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0) r = (lds4[0] + lds4[1]) + (lds4[2] + lds4[3]);
  __syncthreads();
  return r;
}

template <typename TS>
__global__ __launch_bounds__(256) void k_totals(const InstDev* insts, int inst, int par, double* partials) {
  __shared__ double lds4[4];
  const InstDev& in = insts[inst];
  const long long n = (long long)in.ny * in.pitch;
  const TS* src = reinterpret_cast<const TS*>(in.state[in.parity0 ^ par]);
  double acc[4] = {0, 0, 0, 0};
  for (long long o = blockIdx.x * (long long)blockDim.x + threadIdx.x; o < n; o += (long long)gridDim.x * blockDim.x) {
    if ((int)(o % in.pitch) >= in.nx || in.ctype[o] != 0) continue;
    if constexpr (sizeof(TS) == 4) {  // the FP64 moment planes of the state
      const double* mo = in.mom[in.parity0 ^ par];
#pragma unroll
      for (int k = 0; k < 4; k++) acc[k] += mo[in.plane * k + o];
      continue;
    }
#pragma unroll
    for (int i = 0; i < kQ; i++) {
      const double f = double(src[in.plane * i + o]);
      acc[0] += f;
      acc[1] += f * (lat_x(i) * in.c);
      acc[2] += f * (lat_y(i) * in.c);
      acc[3] += double(src[in.plane * (kQ + i) + o]);
    }
  }
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const double b = block_sum(acc[k], lds4);
    if (threadIdx.x == 0) partials[k * kRedBlocks + blockIdx.x] = b;
  }
}

__global__ __launch_bounds__(256) void k_totals_final(const double* partials, int nparts, double* out) {
  __shared__ double lds4[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    double v = 0;
    for (int j = threadIdx.x; j < nparts; j += blockDim.x) v += partials[k * kRedBlocks + j];
    const double b = block_sum(v, lds4);
    if (threadIdx.x == 0) out[k] = b;
  }
}

template <typename TS>
void launch_totals(const InstDev* insts, int inst, int par, double* scratch, long long cells, hipStream_t s) {
  long long b = (cells + 255) / 256;
  const int blocks = (int)(b < kRedBlocks ? (b < 1 ? 1 : b) : kRedBlocks);
  k_totals<TS><<<blocks, 256, 0, s>>>(insts, inst, par, scratch);
  k_totals_final<<<1, 256, 0, s>>>(scratch, blocks, scratch + 4 * kRedBlocks);
}

int totals_scratch_doubles() { return 4 * kRedBlocks + 4; }

// ---------------------------------------------------------------------------
// Residual of the last step (reference runOne, cpuLBM.cpp:2322-2331): RMS
// over all nx*ny cells of the change of T between the previous state (the
// other state buffer, which a step leaves untouched) and the current one.
// Deterministic two-level FP64 reduction: one partial sum per grid row
// (rows[y], wavefront shuffles within the row's block), then one block adds
// the rows in a fixed order and writes sqrt(sum / (nx ny)) to rows[ny].
// ---------------------------------------------------------------------------
template <typename TS>
__device__ __forceinline__ double cell_T(const InstDev& in, int k, long long o) {
  CellState<TS> cs;
  load_state(in, state_buf<TS>(in, k), in.mom[k], o, cs);
  Macro m;
  finish_macro(in, cs.s[0], cs.s[1], cs.s[2], cs.s[3], m);
  return m.T;
}

template <typename TS>
__global__ __launch_bounds__(256) void k_resid_rows(const InstDev* insts, int inst, int par, double* rows) {
  __shared__ double lds4[4];
  const InstDev& in = insts[inst];
  const int now = in.parity0 ^ par, prev = 1 ^ now;
  for (int y = blockIdx.x; y < in.ny; y += gridDim.x) {
    double acc = 0;
    for (int x = threadIdx.x; x < in.nx; x += blockDim.x) {
      const long long o = (long long)y * in.pitch + x;
      const double d = cell_T<TS>(in, now, o) - cell_T<TS>(in, prev, o);
      acc += d * d;
    }
    const double b = block_sum(acc, lds4);
    if (threadIdx.x == 0) rows[y] = b;
  }
}

__global__ __launch_bounds__(256) void k_resid_final(double* rows, int ny, double cells) {
  __shared__ double lds4[4];
  double v = 0;
  for (int j = threadIdx.x; j < ny; j += blockDim.x) v += rows[j];
  const double b = block_sum(v, lds4);
  if (threadIdx.x == 0) rows[ny] = sqrt(b / cells);
}

// Per-step residual of the tracked instances of a launch group (served
// requests with a tolerance), inside the step sequence: after step `step`
// of the sequence (state parity par -> par ^ 1) each pointwise tile of an
// instance whose global step is in 2..20 or equals res_last writes its
// partial sum of (T_new - T_old)^2; one block per instance then adds its
// tiles in their fixed order.
template <typename TS>
__global__ __launch_bounds__(kNT) void k_track_tiles(const InstDev* insts, const TileRef* tiles, int ntiles, int par,
                                                     const int* stepp, double* partial) {
  __shared__ double lds4[4];
  const int step = *stepp;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const TileRef tr = tiles[t];
    const InstDev& in = insts[tr.inst];
    if (!in.res_hist) continue;
    const int gs = in.step_base + step;
    if (!((gs >= 2 && gs <= kResHist) || gs == in.res_last) || frozen_by_failure(in, step)) continue;
    const PointCell c = point_cell(in, tr);
    double d2 = 0;
    if (c.active) {
      const int prev = in.parity0 ^ par, now = 1 ^ prev;
      const double d = cell_T<TS>(in, now, c.o) - cell_T<TS>(in, prev, c.o);
      d2 = d * d;
    }
    const double b = block_sum(d2, lds4);
    if (threadIdx.x == 0) partial[t] = b;
  }
}

__global__ __launch_bounds__(256) void k_track_sum(const InstDev* insts, int ninst, const int* stepp,
                                                   const double* partial) {
  __shared__ double lds4[4];
  const int step = *stepp;
  for (int k = blockIdx.x; k < ninst; k += gridDim.x) {
    const InstDev& in = insts[k];
    if (!in.res_hist) continue;
    const int gs = in.step_base + step;
    if (!((gs >= 2 && gs <= kResHist) || gs == in.res_last) || frozen_by_failure(in, step)) continue;
    double v = 0;
    for (int j = threadIdx.x; j < in.nptiles; j += blockDim.x) v += partial[in.ptile0 + j];
    const double b = block_sum(v, lds4);
    if (threadIdx.x == 0) {
      const double r = sqrt(b / ((double)in.nx * (double)in.ny));
      if (gs >= 2 && gs <= kResHist) in.res_hist[gs - 2] = r;
      if (gs == in.res_last) in.res_hist[kResHist - 1] = r;
    }
  }
}

// Failure steps of the members of a launch group gathered into one array,
// so that the host reads them with a single copy.
__global__ void k_gather_fail(const InstDev* insts, int n, int* out) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) out[k] = *insts[k].fail;
}

void launch_gather_fail(const InstDev* insts, int n, int* out, hipStream_t s) {
  const int blocks = (n + 255) / 256 < 64 ? (n + 255) / 256 : 64;
  k_gather_fail<<<blocks, 256, 0, s>>>(insts, n, out);
}

template <typename TS>
void launch_residual(const InstDev* insts, int inst, int par, double* rows, int nx, int ny, hipStream_t s) {
  const int blocks = ny < 2048 ? ny : 2048;
  k_resid_rows<TS><<<blocks, 256, 0, s>>>(insts, inst, par, rows);
  k_resid_final<<<1, 256, 0, s>>>(rows, ny, (double)nx * (double)ny);
}

// ---------------------------------------------------------------------------
// Launch wrappers
// ---------------------------------------------------------------------------
void flux_tile_dims(int shape, int* w, int* h) {
  switch (shape) {
    case 1: *w = Shape1::TX; *h = Shape1::TY; break;
    case 2: *w = Shape2::TX; *h = Shape2::TY; break;
    case 3: *w = Shape3::TX; *h = Shape3::TY; break;
    case 4: *w = Shape4::TX; *h = Shape4::TY; break;
    default: *w = Shape0::TX; *h = Shape0::TY; break;
  }
}
int tile_halo() { return kH; }
int dflux_planes() { return kNDF; }

static int grid_for(const void* kernel, int threads, int max_per = 0) {
  int dev = 0, cus = 256, per = 4;
  (void)hipGetDevice(&dev);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) == hipSuccess) cus = prop.multiProcessorCount;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, 0) != hipSuccess || per < 1) per = 1;
  if (max_per > 0 && per > max_per) per = max_per;
  long long g = (long long)cus * per;
  g -= g % 8;
  return (int)(g < 8 ? 8 : g);
}

static int blocks_for(int ntiles, int g) { return ntiles <= g ? ntiles : g; }

// Persistent grid sizes of every step kernel of one storage/compute type and
// flux tile shape, computed once (before any stream capture).
// Compute type of the pointwise stage kernels and the wall closure: FP64
// for both storage types; the stage updates of FP32 storage are formed in
// FP64 and rounded once, on the store (stage updates in FP32, five or more
// roundings of the deviation per stored value, triple the error over 5000
// steps of the UQ decks). The equilibria of FP32 storage are evaluated in
// FP32 in the deviation form for instances that start near their reference
// equilibrium (InstDev::eq32), else in FP64: on a shock deck, whose
// deviations are of the order of the state itself, FP32 equilibria seed the
// scheme's unstable mode about as strongly as FP32 states do and the run
// ends earlier.
using TP = double;
using TW = double;

// The stage-3 flux kernel also sums the moments of dF(f3) for the FP64
// moment planes of FP32 storage.
template <typename TS>
constexpr bool kMom3 = sizeof(TS) == 4;

// Pointwise kernel grids are per equilibrium precision: [0] FP64, [1] FP32
// (FP32 storage only).
template <typename TS, typename TC, typename Sh>
struct StepGrids {
  int s12[2] = {0, 0}, s3[2] = {0, 0}, s1[2] = {0, 0}, fin0[2] = {0, 0}, fin2[2] = {0, 0}, fin3[2] = {0, 0};
  int flux3, flux0, fluxe;  // flux0 also serves the stage-2 flux of RK3
  template <typename TE>
  void pointwise(int k) {
    s12[k] = grid_for((const void*)k_s12<TS, TP, TE>, kNT);
    s3[k] = grid_for((const void*)k_s3<TS, TP, TE>, kNT);
    s1[k] = grid_for((const void*)k_s1<TS, TP, TE>, kNT);
    fin0[k] = grid_for((const void*)k_fin<TS, TP, TE, 0>, kNT);
    fin2[k] = grid_for((const void*)k_fin<TS, TP, TE, 2>, kNT);
    fin3[k] = grid_for((const void*)k_fin<TS, TP, TE, 3>, kNT);
  }
  static StepGrids make() {
    StepGrids g;
    g.template pointwise<double>(0);
    if constexpr (sizeof(TS) == 4) g.template pointwise<float>(1);
    g.flux3 = grid_for((const void*)FluxKernel<TS, TC, 1, kMom3<TS>, Sh>::fn, Sh::NT);
    g.flux0 = grid_for((const void*)FluxKernel<TS, TC, 0, false, Sh>::fn, Sh::NT);
    g.fluxe = grid_for((const void*)FluxKernel<TS, TC, 2, false, Sh>::fn, Sh::NT);
    return g;
  }
  static const StepGrids& get() {
    static const StepGrids g = make();
    return g;
  }
};

// The physical-addressing net fluxes of the flagged tiles, after the flux
// kernel of the same MODE.
template <typename TS, typename TC, int MODE, bool MOM, typename Sh>
static void launch_phys(const StepLaunch& L, hipStream_t s, int par) {
  if (L.nphys > 0)
    k_phys<TS, TC, MODE, MOM, Sh><<<L.ntiles < 1024 ? L.ntiles : 1024, 256, 0, s>>>(L.insts, L.tiles, L.ntiles, par,
                                                                                  L.step_ctr);
}

// Compute type of the strip kernels: FP64 for FP64 storage; for FP32
// storage the precision of the instance's equilibria (FP32 for instances
// near their reference equilibrium, FP64 for the others: on the 4096^2
// Riemann deck FP32 stencil arithmetic on strips ended the mixed run at
// step 1939, FP64 arithmetic completes the 2000 steps), each kind of
// instance taken by its own launch over the same item list.
template <typename TS, typename TC, int MODE>
struct StripGrids {
  int fast = 0, solid = 0;
  static const StripGrids& get() {
    static const StripGrids g = [] {
      StripGrids r;
      // FP64 arithmetic: three 4-wave blocks per CU (three waves per SIMD)
      // even where the registers would allow four, which run slower (4096^2
      // Riemann stage-2 flux 2.09 ms at four waves, 1.67 ms at three)
      const int per = sizeof(TC) == 8 ? 3 : 0;
      r.fast = grid_for((const void*)k_strip<TS, TC, MODE, false>, 256, per);
      r.solid = grid_for((const void*)k_strip<TS, TC, MODE, true>, 256, per);
      return r;
    }();
    return g;
  }
};

template <typename TS, typename TC, int MODE>
static void launch_strips(const StepLaunch& L, hipStream_t s, int par) {
  const StripGrids<TS, TC, MODE>& g = StripGrids<TS, TC, MODE>::get();
  const int nf = L.nstrips_fast, ns = L.nstrips - L.nstrips_fast;
  if (nf > 0)
    k_strip<TS, TC, MODE, false><<<blocks_for((nf + 3) / 4, g.fast), 256, 0, s>>>(L.insts, L.strips, nf, par,
                                                                                  L.step_ctr);
  if (ns > 0)
    k_strip<TS, TC, MODE, true><<<blocks_for((ns + 3) / 4, g.solid), 256, 0, s>>>(L.insts, L.strips + nf, ns, par,
                                                                                  L.step_ctr);
}

// One flux stage of MODE: the strip kernels for the instances on strips and
// the tile kernel for the others (every instance for the moment-summing
// stage-3 flux of FP32 storage, which has no strip form), then the
// physical-addressing net fluxes of flagged tiles.
template <typename TS, typename TC, int MODE, bool MOM, typename Sh>
static void launch_flux(const StepLaunch& L, hipStream_t s, int par, int grid_tiles) {
  const int* step = L.step_ctr;
  bool tiles = true;
  if constexpr (!MOM) {
    if (L.nstrips > 0) {
      if constexpr (sizeof(TS) == 4) {
        if (L.eq_kinds & 2) launch_strips<TS, float, MODE>(L, s, par);
        if (L.eq_kinds & 1) launch_strips<TS, double, MODE>(L, s, par);
      } else {
        launch_strips<TS, double, MODE>(L, s, par);
      }
    }
    tiles = L.tiles_rest > 0;
  }
  if (tiles)
    FluxKernel<TS, TC, MODE, MOM, Sh>::fn<<<blocks_for(L.ntiles, grid_tiles), Sh::NT, 0, s>>>(L.insts, L.tiles,
                                                                                             L.ntiles, par, step);
  launch_phys<TS, TC, MODE, MOM, Sh>(L, s, par);
}

template <typename TS, typename TC>
static void prepare_shapes() {
  (void)StepGrids<TS, TC, Shape0>::get();
  (void)StepGrids<TS, TC, Shape1>::get();
  (void)StepGrids<TS, TC, Shape2>::get();
  (void)StepGrids<TS, TC, Shape3>::get();
  (void)StepGrids<TS, TC, Shape4>::get();
  (void)StripGrids<TS, double, 0>::get();
  (void)StripGrids<TS, double, 2>::get();
  if constexpr (sizeof(TS) == 8) {
    (void)StripGrids<TS, double, 1>::get();
  } else {
    (void)StripGrids<TS, float, 0>::get();
    (void)StripGrids<TS, float, 2>::get();
  }
}

void prepare_launch(int fp64) {
  if (fp64) prepare_shapes<double, double>();
  else prepare_shapes<float, float>();
}

// Calls launch(TE{}, k) once per equilibrium precision present in the group
// (StepLaunch::eq_kinds; k indexes the StepGrids arrays).
template <typename TS, typename F>
static void per_eq_kind(const StepLaunch& L, F&& launch) {
  if constexpr (sizeof(TS) == 4) {
    if (L.eq_kinds & 2) launch(float{}, 1);
    if (L.eq_kinds & 1) launch(double{}, 0);
  } else {
    launch(double{}, 0);
  }
}

template <typename TS, typename TC, typename Sh>
static void launch_rk3(const StepLaunch& L, hipStream_t s, int par, KernelTimer* tm) {
  const int* step = L.step_ctr;
  const StepGrids<TS, TC, Sh>& G = StepGrids<TS, TC, Sh>::get();
  const int g1 = G.flux0, g3 = G.flux3;
  const int wb = (L.nwalls + 255) / 256;
  if (tm) tm->begin(0, s);
  per_eq_kind<TS>(L, [&](auto te, int k) {
    k_s12<TS, TP, decltype(te)><<<blocks_for(L.nptiles, G.s12[k]), kNT, 0, s>>>(L.insts, L.ptiles, L.nptiles, par,
                                                                                step);
  });
  if (tm) tm->end(0, s);
  if (L.nwalls) {
    if (tm) tm->begin(1, s);
    k_alpha<TS, TW><<<wb, 256, 0, s>>>(L.insts, L.walls, L.nwalls, 0, par, step);
    if (tm) tm->end(1, s);
  }
  if (tm) tm->begin(2, s);
  launch_flux<TS, TC, 0, false, Sh>(L, s, par, g1);
  if (tm) tm->end(2, s);
  if (tm) tm->begin(3, s);
  per_eq_kind<TS>(L, [&](auto te, int k) {
    k_s3<TS, TP, decltype(te)><<<blocks_for(L.nptiles, G.s3[k]), kNT, 0, s>>>(L.insts, L.ptiles, L.nptiles, par,
                                                                              step);
  });
  if (tm) tm->end(3, s);
  if (L.nwalls) {
    if (tm) tm->begin(1, s);
    k_alpha<TS, TW><<<wb, 256, 0, s>>>(L.insts, L.walls, L.nwalls, 1, par, step);
    if (tm) tm->end(1, s);
  }
  if (tm) tm->begin(4, s);
  launch_flux<TS, TC, 1, kMom3<TS>, Sh>(L, s, par, g3);
  if (tm) tm->end(4, s);
}

// RKMETHOD 0, 2, 3: one flux stage.
template <typename TS, typename TC, int METHOD, typename Sh>
static void launch_one_stage(const StepLaunch& L, hipStream_t s, int par, KernelTimer* tm) {
  const int* step = L.step_ctr;
  constexpr int FM = METHOD == 0 ? 0 : 2;
  const StepGrids<TS, TC, Sh>& G = StepGrids<TS, TC, Sh>::get();
  const int g1 = FM == 0 ? G.flux0 : G.fluxe;
  const int* g2 = METHOD == 0 ? G.fin0 : (METHOD == 2 ? G.fin2 : G.fin3);
  const int wb = (L.nwalls + 255) / 256;
  if constexpr (METHOD == 0) {
    if (tm) tm->begin(0, s);
    per_eq_kind<TS>(L, [&](auto te, int k) {
      k_s1<TS, TP, decltype(te)><<<blocks_for(L.nptiles, G.s1[k]), kNT, 0, s>>>(L.insts, L.ptiles, L.nptiles, par,
                                                                                step);
    });
    if (tm) tm->end(0, s);
  }
  if (L.nwalls) {
    if (tm) tm->begin(1, s);
    k_alpha<TS, TW><<<wb, 256, 0, s>>>(L.insts, L.walls, L.nwalls, METHOD == 0 ? 0 : 2, par, step);
    if (tm) tm->end(1, s);
  }
  if (tm) tm->begin(2, s);
  launch_flux<TS, TC, FM, false, Sh>(L, s, par, g1);
  if (tm) tm->end(2, s);
  if (tm) tm->begin(3, s);
  per_eq_kind<TS>(L, [&](auto te, int k) {
    k_fin<TS, TP, decltype(te), METHOD><<<blocks_for(L.nptiles, g2[k]), kNT, 0, s>>>(L.insts, L.ptiles, L.nptiles,
                                                                                     par, step);
  });
  if (tm) tm->end(3, s);
}

template <typename TS, typename TC, typename Sh>
static void launch_step_shape(const StepLaunch& L, hipStream_t s, int par, KernelTimer* tm) {
  switch (L.rk_method) {
    case 0: launch_one_stage<TS, TC, 0, Sh>(L, s, par, tm); break;
    case 2: launch_one_stage<TS, TC, 2, Sh>(L, s, par, tm); break;
    case 3: launch_one_stage<TS, TC, 3, Sh>(L, s, par, tm); break;
    default: launch_rk3<TS, TC, Sh>(L, s, par, tm); break;
  }
}

// Advances the step counter of a launch sequence: the kernels of step k read
// k + 1 from it, so one captured sequence replays for any step.
__global__ void k_tick(int* ctr) {
  if (threadIdx.x == 0) ctr[0] = ctr[0] + 1;
}

template <typename TS>
static void launch_track(const StepLaunch& L, hipStream_t s, int par, int ninst) {
  k_track_tiles<TS><<<L.nptiles < 8192 ? L.nptiles : 8192, kNT, 0, s>>>(L.insts, L.ptiles, L.nptiles, par,
                                                                        L.step_ctr, L.res_partial);
  k_track_sum<<<ninst < 1024 ? ninst : 1024, 256, 0, s>>>(L.insts, ninst, L.step_ctr, L.res_partial);
}

template <typename TS, typename TC>
void launch_step(const StepLaunch& L, hipStream_t s, int par, KernelTimer* tm) {
  switch (L.flux_shape) {
    case 1: launch_step_shape<TS, TC, Shape1>(L, s, par, tm); break;
    case 2: launch_step_shape<TS, TC, Shape2>(L, s, par, tm); break;
    case 3: launch_step_shape<TS, TC, Shape3>(L, s, par, tm); break;
    case 4: launch_step_shape<TS, TC, Shape4>(L, s, par, tm); break;
    default: launch_step_shape<TS, TC, Shape0>(L, s, par, tm); break;
  }
  if (L.res_partial && L.nptiles > 0) launch_track<TS>(L, s, par, L.ninst);
  k_tick<<<1, 64, 0, s>>>(L.step_ctr);
}

template <typename TS>
void launch_init(const InstDev* insts, int inst, const int* ids, const double* props, long long cells,
                 hipStream_t s) {
  const int blocks = (int)((cells + 255) / 256 < 4096 ? (cells + 255) / 256 : 4096);
  k_init<TS><<<blocks, 256, 0, s>>>(insts, inst, ids, props);
}

template <typename TS>
void launch_macro(const InstDev* insts, int inst, int par, double* out, long long cells, hipStream_t s) {
  const int blocks = (int)((cells + 255) / 256 < 4096 ? (cells + 255) / 256 : 4096);
  k_macro<TS><<<blocks, 256, 0, s>>>(insts, inst, par, out);
}

template void launch_step<double, double>(const StepLaunch&, hipStream_t, int, KernelTimer*);
template void launch_step<float, float>(const StepLaunch&, hipStream_t, int, KernelTimer*);
template void launch_init<double>(const InstDev*, int, const int*, const double*, long long, hipStream_t);
template void launch_init<float>(const InstDev*, int, const int*, const double*, long long, hipStream_t);
template void launch_macro<double>(const InstDev*, int, int, double*, long long, hipStream_t);
template void launch_macro<float>(const InstDev*, int, int, double*, long long, hipStream_t);
template void launch_totals<double>(const InstDev*, int, int, double*, long long, hipStream_t);
template void launch_totals<float>(const InstDev*, int, int, double*, long long, hipStream_t);
template void launch_residual<double>(const InstDev*, int, int, double*, int, int, hipStream_t);
template void launch_residual<float>(const InstDev*, int, int, double*, int, int, hipStream_t);

}  // namespace uqlbm
