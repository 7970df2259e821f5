// lbm_device.hpp — gfx950 device arithmetic of the D2Q13 thermal FD-LBM:
// lattice tables, equilibria, moments, WENO5 and NND face reconstructions.
//
// Compute type TC is double (FP64 path) or float (mixed path). Moments and
// the WENO weight normalisation are always evaluated in double.
#pragma once

#include <hip/hip_runtime.h>

#include "lbm_common.hpp"

namespace uqlbm {

// Lattice velocities in units of c (reference setupVars, cpuLBM.cpp:376-388).
__host__ __device__ constexpr int lat_x(int i) {
  return i == 1 ? 1 : i == 3 ? -1 : i == 5 ? 1 : i == 6 ? -1 : i == 7 ? -1 : i == 8 ? 1
       : i == 9 ? 2 : i == 11 ? -2 : 0;
}
__host__ __device__ constexpr int lat_y(int i) {
  return i == 2 ? 1 : i == 4 ? -1 : i == 5 ? 1 : i == 6 ? 1 : i == 7 ? -1 : i == 8 ? -1
       : i == 10 ? 2 : i == 12 ? -2 : 0;
}
__host__ __device__ constexpr int sgn_i(int a) { return a > 0 ? 1 : (a < 0 ? -1 : 0); }

// Direction remaps for reflections about the x axis (north/south mirror) and
// the y axis (east/west mirror) (reference mirIndex, cpuLBM.cpp:1147-1169).
__host__ __device__ constexpr int mirror_ns(int i) {
  return i == 2 ? 4 : i == 4 ? 2 : i == 5 ? 8 : i == 6 ? 7 : i == 7 ? 6 : i == 8 ? 5
       : i == 10 ? 12 : i == 12 ? 10 : i;
}
__host__ __device__ constexpr int mirror_ew(int i) {
  return i == 1 ? 3 : i == 3 ? 1 : i == 5 ? 6 : i == 6 ? 5 : i == 7 ? 8 : i == 8 ? 7
       : i == 9 ? 11 : i == 11 ? 9 : i;
}

// Equilibrium argument map: direction i evaluates base polynomial
// eq_base(i) at (a, b) where a = sa*(swap ? v : u), b = sb*(swap ? u : v)
// (reference equilibrium2D, cpuLBM.cpp:795-823).
__host__ __device__ constexpr int eq_base(int i) { return i == 0 ? 0 : i <= 4 ? 1 : i <= 8 ? 5 : 9; }
__host__ __device__ constexpr bool eq_swap(int i) { return i == 2 || i == 4 || i == 10 || i == 12; }
__host__ __device__ constexpr int eq_sa(int i) {
  return (i == 3 || i == 4 || i == 6 || i == 7 || i == 11 || i == 12) ? -1 : 1;
}
__host__ __device__ constexpr int eq_sb(int i) { return (i == 7 || i == 8) ? -1 : 1; }

// ---------------------------------------------------------------------------
// Equilibrium polynomials: the closed forms of the reference
// (inlineFunctions.cpp:10-61), evaluated with products instead of pow() and
// with the constant denominators applied as reciprocal products. Host and
// device: the host forms the FP32-storage reference equilibrium with them.
// ---------------------------------------------------------------------------
template <typename T>
__host__ __device__ __forceinline__ T poly_f0(T rho, T u, T v, T t) {
  const T u2 = u * u, v2 = v * v;
  return (rho * T(0.25)) * (T(4) + T(10) * (t * t) + u2 * u2 - T(5) * v2 + v2 * v2 +
                         T(10) * t * (T(-1) + u2 + v2) + u2 * (T(-5) + T(4) * v2));
}
template <typename T>
__host__ __device__ __forceinline__ T poly_f1(T rho, T a, T b, T t) {
  const T a2 = a * a, b2 = b * b;
  return (rho * T(1.0 / 6.0)) * (T(-6) * (t * t) - a * (T(1) + a) * (T(-4) + a2 + T(3) * b2) -
                         t * (T(-4) + T(6) * a + T(9) * a2 + T(3) * b2));
}
template <typename T>
__host__ __device__ __forceinline__ T poly_f5(T rho, T a, T b, T t) {
  return (rho * T(0.25)) * ((t + a + a * a) * (t + b + b * b));
}
template <typename T>
__host__ __device__ __forceinline__ T poly_f9(T rho, T a, T t) {
  return (rho * T(1.0 / 24.0)) * (T(3) * (t * t) + (T(-1) + a) * a * (T(1) + a) * (T(2) + a) +
                          t * (T(-1) + T(6) * a * (T(1) + a)));
}
template <typename T>
__host__ __device__ __forceinline__ T poly_h0(T K, T t, T u, T v) {
  const T u2 = u * u, v2 = v * v, u4 = u2 * u2, v4 = v2 * v2, t2 = t * t, t3 = t2 * t;
  return (T(10) * (T(16) + T(3) * K) * t3 +
          T(3) * t *
              (T(8) + T(4) * K - T(40) * u2 - T(5) * K * u2 + T(20) * u4 + K * u4 +
               (T(-5) * (T(8) + K) + T(4) * (T(15) + K) * u2) * v2 + (T(20) + K) * v4) +
          T(30) * t2 * (T(-4) + T(9) * u2 + T(9) * v2 + K * (T(-1) + u2 + v2)) +
          T(3) * (u2 + v2) * (T(4) + u4 - T(5) * v2 + v4 + u2 * (T(-5) + T(4) * v2))) *
         T(1.0 / 24.0);
}
template <typename T>
__host__ __device__ __forceinline__ T poly_h1(T K, T t, T a, T b) {
  const T a2 = a * a, a3 = a2 * a, b2 = b * b, t2 = t * t, t3 = t2 * t;
  return (T(-2) * (T(16) + T(3) * K) * t3 - a * (T(1) + a) * (a2 + b2) * (T(-4) + a2 + T(3) * b2) -
          t * (a * (T(-4) * (T(4) + K) - T(4) * (T(7) + K) * a + (T(14) + K) * a2 + (T(19) + K) * a3) +
               (T(-4) + T(3) * a * (T(10) + K + (T(14) + K) * a)) * b2 + T(3) * (b2 * b2)) -
          t2 * (T(-16) + T(6) * a * (T(6) + T(13) * a) + T(30) * b2 +
                K * (T(-4) + T(6) * a + T(9) * a2 + T(3) * b2))) *
         T(1.0 / 12.0);
}
template <typename T>
__host__ __device__ __forceinline__ T poly_h5(T K, T t, T a, T b) {
  const T a2 = a * a, b2 = b * b, t2 = t * t, t3 = t2 * t;
  return ((T(16) + T(3) * K) * t3 + T(3) * a * (T(1) + a) * b * (T(1) + b) * (a2 + b2) +
          T(3) * t2 * ((T(6) + K) * a + (T(9) + K) * a2 + b * (T(6) + K + (T(9) + K) * b)) +
          T(3) * t *
              (a2 * a + a2 * a2 + b2 * b * (T(1) + b) + a * b * (T(6) + K + (T(9) + K) * b) +
               a2 * b * (T(9) + K + (T(12) + K) * b))) *
         T(1.0 / 24.0);
}
template <typename T>
__host__ __device__ __forceinline__ T poly_h9(T K, T t, T a, T b) {
  const T a2 = a * a, b2 = b * b, t2 = t * t, t3 = t2 * t;
  return ((T(16) + T(3) * K) * t3 +
          t * a * (T(-8) + K * (T(-1) + a) * (T(1) + a) * (T(2) + a) + a * (T(-7) + T(2) * a * (T(11) + T(8) * a))) +
          t * (T(-1) + T(6) * a * (T(1) + a)) * b2 + (T(-1) + a) * a * (T(1) + a) * (T(2) + a) * (a2 + b2) +
          t2 * (T(-4) + T(36) * a + T(51) * a2 + K * (T(-1) + T(6) * a * (T(1) + a)) + T(3) * b2)) *
         T(1.0 / 48.0);
}

// Normalised equilibrium inputs of one cell.
template <typename T>
struct EqIn {
  T rho;    // density
  T u, v;   // velocity / sqrt(R*Tc)
  T t;      // temperature / Tc
  T K;      // internal degrees of freedom
  T scale;  // rho*R*Tc, the h scale
};

// f and h equilibrium of direction I (compile-time).
template <int I, typename T>
__host__ __device__ __forceinline__ void eq_dir(const EqIn<T>& q, T& fe, T& he) {
  constexpr int base = eq_base(I);
  const T a = T(eq_sa(I)) * (eq_swap(I) ? q.v : q.u);
  const T b = T(eq_sb(I)) * (eq_swap(I) ? q.u : q.v);
  if constexpr (base == 0) {
    fe = poly_f0(q.rho, q.u, q.v, q.t);
    he = q.scale * poly_h0(q.K, q.t, q.u, q.v);
  } else if constexpr (base == 1) {
    fe = poly_f1(q.rho, a, b, q.t);
    he = q.scale * poly_h1(q.K, q.t, a, b);
  } else if constexpr (base == 5) {
    fe = poly_f5(q.rho, a, b, q.t);
    he = q.scale * poly_h5(q.K, q.t, a, b);
  } else {
    fe = poly_f9(q.rho, a, q.t);
    he = q.scale * poly_h9(q.K, q.t, a, b);
  }
}

// ---------------------------------------------------------------------------
// Equilibria as deviations from the instance's reference equilibrium (the
// equilibrium at rest at density rho0 and t0 = T0/Tc), for FP32 storage:
//   feq - ref = rho (P_uv(u, v, t) + P(0, 0, t) - P(0, 0, t0)) + (rho - rho0) P(0, 0, t0)
// where P_uv collects the terms of the closed form that carry a velocity
// factor and the rest-state difference is factored by t - t0. Every term
// is of the size of the deviation, so FP32 evaluation errs relative to the
// deviation instead of to the full value. Same closed forms as above
// (reference inlineFunctions.cpp:10-61), regrouped.
// ---------------------------------------------------------------------------
// The kernels evaluate it in FP32 (EqDev); the FP64 instance checks the
// regrouping against the closed form on the host.
template <typename T>
struct EqDevT {
  T u, v, t, dt;      // velocity / sqrt(R Tc), T/Tc, T/Tc - t0
  T tt, t2s;          // t + t0, t^2 + t t0 + t0^2
  T rho, drho;        // density, density - rho0
  T K, scale, dscale;  // internal degrees of freedom, rho R Tc, (rho - rho0) R Tc
};
using EqDev = EqDevT<float>;

template <typename T>
__host__ __device__ __forceinline__ T pf0_uv(T u, T v, T t) {
  const T u2 = u * u, v2 = v * v;
  return T(0.25) * (u2 * u2 - T(5) * v2 + v2 * v2 + T(10) * t * (u2 + v2) + u2 * (-T(5) + T(4) * v2));
}
template <typename T>
__host__ __device__ __forceinline__ T pf1_uv(T a, T b, T t) {
  const T a2 = a * a, b2 = b * b;
  return (T(1) / T(6)) * (-a * (T(1) + a) * (-T(4) + a2 + T(3) * b2) - t * (T(6) * a + T(9) * a2 + T(3) * b2));
}
template <typename T>
__host__ __device__ __forceinline__ T pf5_uv(T a, T b, T t) {
  return T(0.25) * (t * (b + b * b) + (a + a * a) * (t + b + b * b));
}
template <typename T>
__host__ __device__ __forceinline__ T pf9_uv(T a, T t) {
  return (T(1) / T(24)) * ((-T(1) + a) * a * (T(1) + a) * (T(2) + a) + T(6) * t * a * (T(1) + a));
}
template <typename T>
__host__ __device__ __forceinline__ T ph0_uv(T K, T t, T u, T v) {
  const T u2 = u * u, v2 = v * v, u4 = u2 * u2, v4 = v2 * v2;
  return (T(3) * t * (-T(40) * u2 - T(5) * K * u2 + T(20) * u4 + K * u4 + (-T(5) * (T(8) + K) + T(4) * (T(15) + K) * u2) * v2 +
                     (T(20) + K) * v4) +
          T(30) * (t * t) * ((T(9) + K) * (u2 + v2)) +
          T(3) * (u2 + v2) * (T(4) + u4 - T(5) * v2 + v4 + u2 * (-T(5) + T(4) * v2))) *
         (T(1) / T(24));
}
template <typename T>
__host__ __device__ __forceinline__ T ph1_uv(T K, T t, T a, T b) {
  const T a2 = a * a, a3 = a2 * a, b2 = b * b;
  return (-a * (T(1) + a) * (a2 + b2) * (-T(4) + a2 + T(3) * b2) -
          t * (a * (-T(4) * (T(4) + K) - T(4) * (T(7) + K) * a + (T(14) + K) * a2 + (T(19) + K) * a3) +
               (-T(4) + T(3) * a * (T(10) + K + (T(14) + K) * a)) * b2 + T(3) * (b2 * b2)) -
          (t * t) * (T(6) * a * (T(6) + T(13) * a) + T(30) * b2 + K * (T(6) * a + T(9) * a2 + T(3) * b2))) *
         (T(1) / T(12));
}
template <typename T>
__host__ __device__ __forceinline__ T ph5_uv(T K, T t, T a, T b) {
  const T a2 = a * a, b2 = b * b;
  return (T(3) * a * (T(1) + a) * b * (T(1) + b) * (a2 + b2) +
          T(3) * (t * t) * ((T(6) + K) * a + (T(9) + K) * a2 + b * (T(6) + K + (T(9) + K) * b)) +
          T(3) * t *
              (a2 * a + a2 * a2 + b2 * b * (T(1) + b) + a * b * (T(6) + K + (T(9) + K) * b) +
               a2 * b * (T(9) + K + (T(12) + K) * b))) *
         (T(1) / T(24));
}
template <typename T>
__host__ __device__ __forceinline__ T ph9_uv(T K, T t, T a, T b) {
  const T a2 = a * a, b2 = b * b;
  return (t * a * (-T(8) + K * (-T(1) + a) * (T(1) + a) * (T(2) + a) + a * (-T(7) + T(2) * a * (T(11) + T(8) * a))) +
          t * (-T(1) + T(6) * a * (T(1) + a)) * b2 + (-T(1) + a) * a * (T(1) + a) * (T(2) + a) * (a2 + b2) +
          (t * t) * (T(36) * a + T(51) * a2 + T(6) * K * a * (T(1) + a) + T(3) * b2)) *
         (T(1) / T(48));
}

// f and h equilibrium deviations of direction I; eqref holds the base
// polynomials at rest at t0 (f per unit density, h per unit scale).
template <int I, typename T>
__host__ __device__ __forceinline__ void eq_dev(const EqDevT<T>& q, const T* eqref, T& fe, T& he) {
  constexpr int base = eq_base(I);
  const T a = T(eq_sa(I)) * (eq_swap(I) ? q.v : q.u);
  const T b = T(eq_sb(I)) * (eq_swap(I) ? q.u : q.v);
  const T K = q.K, t = q.t, dt = q.dt, tt = q.tt, t2s = q.t2s;
  const T c3 = T(16) + T(3) * K;
  if constexpr (base == 0) {
    fe = q.rho * (pf0_uv(q.u, q.v, t) + T(2.5) * dt * (tt - T(1))) + q.drho * eqref[0];
    he = q.scale * (ph0_uv(K, t, q.u, q.v) +
                    dt * (T(10) * c3 * t2s + T(3) * (T(8) + T(4) * K) - T(30) * (T(4) + K) * tt) * (T(1) / T(24))) +
         q.dscale * eqref[4];
  } else if constexpr (base == 1) {
    fe = q.rho * (pf1_uv(a, b, t) + dt * (-T(6) * tt + T(4)) * (T(1) / T(6))) + q.drho * eqref[1];
    he = q.scale * (ph1_uv(K, t, a, b) + dt * (-T(2) * c3 * t2s + (T(16) + T(4) * K) * tt) * (T(1) / T(12))) +
         q.dscale * eqref[5];
  } else if constexpr (base == 5) {
    fe = q.rho * (pf5_uv(a, b, t) + T(0.25) * dt * tt) + q.drho * eqref[2];
    he = q.scale * (ph5_uv(K, t, a, b) + dt * c3 * t2s * (T(1) / T(24))) + q.dscale * eqref[6];
  } else {
    fe = q.rho * (pf9_uv(a, t) + dt * (T(3) * tt - T(1)) * (T(1) / T(24))) + q.drho * eqref[3];
    he = q.scale * (ph9_uv(K, t, a, b) + dt * (c3 * t2s - (T(4) + K) * tt) * (T(1) / T(48))) + q.dscale * eqref[7];
  }
}

// Moments of one cell (reference macroProp, cpuLBM.cpp:916-970), FP64.
// The relaxation times enter every update as divisors; the cell keeps their
// reciprocals: 1/tf = p/mu, 1/th = Pr/tf, and since thf = th tf/(tf - th),
// 1/thf = 1/th - 1/tf = (Pr - 1)/tf (one rounding; Pr - 1 is exact for
// 0.5 <= Pr <= 2).
struct Macro {
  double rho, ux, uy, T, p;
  double itf, ith, ithf;  // 1/tf, 1/th, 1/thf
};

// Temperature-dependent viscosity of the OpenCL front end
// (PY/codeCL.py:987-990): VHS mu_ref (T/T_ref)^vexp or Sutherland
// mu_ref (T/T_ref)^1.5 (T_ref + S)/(T + S). Kept out of line so that the
// constant-viscosity path carries none of pow's registers.
__device__ __noinline__ static double viscosity_of(const InstDev& in, double T) {
  const double r = T * in.iT_ref_v;
  return in.visc_model == 1 ? in.mu_ref * pow(r, in.vexp)
                            : in.mu_ref * pow(r, 1.5) * ((in.T_ref_v + in.vS) / (T + in.vS));
}

__device__ __forceinline__ void finish_macro(const InstDev& in, double rho, double rux, double ruy,
                                             double sh, Macro& m) {
  const double irho = 1.0 / rho;
  m.rho = rho;
  m.ux = rux * irho;
  m.uy = ruy * irho;
  const double usq = m.ux * m.ux + m.uy * m.uy;
  m.T = (2.0 * (sh * irho) - usq) * in.ibR;
  m.p = rho * in.R * m.T;
  m.itf = in.visc_model == 0 ? m.p * in.imu : m.p / viscosity_of(in, m.T);
  m.ith = m.itf * in.Pr;
  m.ithf = m.itf * in.Pr_m1;
}

template <typename T = float>
__host__ __device__ __forceinline__ EqDevT<T> eq_dev_inputs(const InstDev& in, const Macro& m) {
  EqDevT<T> q;
  const double t = m.T * in.iTc;
  q.u = T(m.ux * in.isRTc);
  q.v = T(m.uy * in.isRTc);
  q.t = T(t);
  q.dt = T(t - in.t0);
  q.tt = T(t + in.t0);
  q.t2s = T(t * t + t * in.t0 + in.t0 * in.t0);
  q.rho = T(m.rho);
  q.drho = T(m.rho - in.rho0);
  q.K = T(in.K);
  q.scale = T(m.rho * in.RTc);
  q.dscale = T((m.rho - in.rho0) * in.RTc);
  return q;
}

template <typename T>
__host__ __device__ __forceinline__ EqIn<T> eq_inputs(const InstDev& in, const Macro& m) {
  EqIn<T> q;
  q.rho = T(m.rho);
  q.u = T(m.ux * in.isRTc);
  q.v = T(m.uy * in.isRTc);
  q.t = T(m.T * in.iTc);
  q.K = T(in.K);
  q.scale = T(m.rho * in.RTc);
  return q;
}

// ---------------------------------------------------------------------------
// Face reconstructions (reference WENO5 cpuLBM.cpp:1358-1403, NND
// :1328-1356, minmod :1130-1145). `v` is |e| along the axis (non-zero).
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T minmod(T a, T b) {
  if ((fabs(a) < fabs(b)) && (a * b > T(0))) return a;
  if ((fabs(b) < fabs(a)) && (a * b > T(0))) return b;
  if (a == b) return a;
  if (a * b <= T(0)) return T(0);
  return a + b;  // NaN input only
}

__device__ __forceinline__ float fused(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fused(double a, double b, double c) { return __builtin_fma(a, b, c); }

// The reconstructions are written with every fused multiply-add explicit and
// contraction off, so each inlined copy (the solid and solid-free face paths,
// every tile shape) rounds identically: a deck that is uniform along an axis
// stays bitwise uniform whichever path each tile takes.
template <typename T>
__device__ __forceinline__ T nnd(T v, T s0, T s1, T s2) {
#pragma clang fp contract(off)
  const T fi = v * s1, fp = v * s2, fm = v * s0;
  return fused(T(0.5), minmod(fp - fi, fi - fm), fi);
}

// Jiang-Shu WENO5 with d = (0.1, 0.6, 0.3), eps = 1e-6 on |e|*S. The
// smoothness indicators and candidate values are formed on the unscaled S
// and scaled afterwards (beta(vS) = v^2 beta(S), q(vS) = v q(S)), and the
// normalised weights w_k = a_k / sum(a) with a_k = d_k/(eps+b_k)^2 are
// formed as A_k / sum(A) with A_k = d_k * prod_{j!=k} (eps+b_j)^2 — the
// same ratio, one division instead of six — in FP64 for both precisions.
// The quotient num / sum(A) is formed by Newton-Raphson refinement of the
// hardware reciprocal (two steps r += r (1 - D r), then the residual
// correction q += r (num - D q), Markstein's scheme): every A_k >= 0.1e-24
// (each eps + b_j >= 1e-6), so sum(A) is a normal number far from both
// range limits, the range scaling of a general division is not needed, and
// the result is within one rounding of the quotient.
template <typename T>
__device__ __forceinline__ T weno5(T v, T s0, T s1, T s2, T s3, T s4) {
#pragma clang fp contract(off)
  const T c0 = fused(T(-2), s1, s0 + s2), d0 = fused(T(3), s2, fused(T(-4), s1, s0));
  const T c1 = fused(T(-2), s2, s1 + s3), d1 = s1 - s3;
  const T c2 = fused(T(-2), s3, s2 + s4), d2 = fused(T(3), s2, fused(T(-4), s3, s4));
  const T b0 = fused(T(13.0 / 12.0) * c0, c0, T(0.25) * (d0 * d0));
  const T b1 = fused(T(13.0 / 12.0) * c1, c1, T(0.25) * (d1 * d1));
  const T b2 = fused(T(13.0 / 12.0) * c2, c2, T(0.25) * (d2 * d2));
  const T r0 = fused(T(11.0 / 6.0), s2, fused(T(-7.0 / 6.0), s1, T(2.0 / 6.0) * s0));
  const T r1 = fused(T(2.0 / 6.0), s3, fused(T(5.0 / 6.0), s2, T(-1.0 / 6.0) * s1));
  const T r2 = fused(T(-1.0 / 6.0), s4, fused(T(5.0 / 6.0), s3, T(2.0 / 6.0) * s2));
  const double v2 = double(v) * double(v);
  const double e0 = fused(v2, double(b0), 1e-6), e1 = fused(v2, double(b1), 1e-6),
               e2 = fused(v2, double(b2), 1e-6);
  const double e00 = e0 * e0, e11 = e1 * e1, e22 = e2 * e2;
  const double A0 = 0.1 * (e11 * e22), A1 = 0.6 * (e00 * e22), A2 = 0.3 * (e00 * e11);
  const double num = fused(A2, double(r2), fused(A1, double(r1), A0 * double(r0)));
  const double den = (A0 + A1) + A2;
  double r = __builtin_amdgcn_rcp(den);
  r = fused(fused(-den, r, 1.0), r, r);
  r = fused(fused(-den, r, 1.0), r, r);
  double q = num * r;
  q = fused(fused(-den, q, num), r, q);
  return T(double(v) * q);
}

}  // namespace uqlbm
