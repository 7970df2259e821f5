// Host check of the deviation-form equilibria that FP32 storage evaluates
// (eq_dev, uq_amd/csrc/lbm_device.hpp) against the closed form (eq_dir, the
// reference's equilibrium2D, CPP/inlineFunctions.cpp:10-61): over random
// cells near and far from a reference state at rest,
//   * the FP64 instance of the regrouped form equals closed form minus
//     reference equilibrium to FP64 rounding (any dropped or altered term
//     shows as an error of the size of the deviation), and
//   * the FP32 instance errs relative to the deviation, not to the full
//     value.
// Prints the worst relative errors; exits 1 if either exceeds its bound.
#include <cmath>
#include <cstdio>
#include <random>
#include <utility>

#include "../uq_amd/csrc/lbm_device.hpp"

using namespace uqlbm;

template <typename F, int... Is>
static void each_dir(F&& f, std::integer_sequence<int, Is...>) {
  (f(std::integral_constant<int, Is>{}), ...);
}

int main() {
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  double worst64 = 0, worst32 = 0;
  for (int deck = 0; deck < 40; deck++) {
    // a deck: gas constants, lattice temperature and reference state
    InstDev in{};
    const double R = deck % 2 ? 287.0 : 1.0, gamma = 1.3 + 0.1 * (deck % 4);
    const double b = 2.0 / (gamma - 1.0);
    in.K = b - 2.0;
    const double T0 = deck % 2 ? 250.0 : 1.0;
    in.Tc = T0 * (1.0 + 1.5 * (deck % 3) / 2.0);
    in.iTc = 1.0 / in.Tc;
    in.RTc = R * in.Tc;
    in.isRTc = 1.0 / std::sqrt(in.RTc);
    in.rho0 = deck % 2 ? 0.9 : 1.0;
    in.t0 = T0 / in.Tc;
    // the reference equilibrium at rest and its base polynomials
    Macro m0{};
    m0.rho = in.rho0;
    m0.T = T0;
    const EqIn<double> q0 = eq_inputs<double>(in, m0);
    double ref_f[kQ], ref_h[kQ];
    each_dir([&](auto Ic) { eq_dir<decltype(Ic)::value>(q0, ref_f[Ic], ref_h[Ic]); },
             std::make_integer_sequence<int, kQ>{});
    const int bases[4] = {0, 1, 5, 9};
    double eqref64[8];
    float eqref32[8];
    for (int j = 0; j < 4; j++) {
      eqref64[j] = ref_f[bases[j]] / in.rho0;
      eqref64[4 + j] = ref_h[bases[j]] / (in.rho0 * in.RTc);
      eqref32[j] = float(eqref64[j]);
      eqref32[4 + j] = float(eqref64[4 + j]);
    }
    for (int cell = 0; cell < 2000; cell++) {
      // deviations from 1e-6 (near the reference) to 30 %
      const double amp = std::pow(10.0, -6.0 + 5.5 * (0.5 + 0.5 * U(rng)));
      Macro m{};
      m.rho = in.rho0 * (1.0 + amp * U(rng));
      m.T = T0 * (1.0 + amp * U(rng));
      m.ux = amp * U(rng) / in.isRTc;
      m.uy = amp * U(rng) / in.isRTc;
      const EqIn<double> q = eq_inputs<double>(in, m);
      const EqDevT<double> d64 = eq_dev_inputs<double>(in, m);
      const EqDevT<float> d32 = eq_dev_inputs<float>(in, m);
      double dev_scale = 0, e64 = 0, e32 = 0;
      each_dir(
          [&](auto Ic) {
            constexpr int I = decltype(Ic)::value;
            double fe, he, f64, h64;
            float f32, h32;
            eq_dir<I>(q, fe, he);
            eq_dev<I>(d64, eqref64, f64, h64);
            eq_dev<I>(d32, eqref32, f32, h32);
            const double df = fe - ref_f[I], dh = (he - ref_h[I]) / in.RTc;
            dev_scale = std::fmax(dev_scale, std::fmax(std::fabs(df), std::fabs(dh)));
            e64 = std::fmax(e64, std::fmax(std::fabs(f64 - df), std::fabs(h64 / in.RTc - dh)));
            e32 = std::fmax(e32, std::fmax(std::fabs(double(f32) - df), std::fabs(double(h32) / in.RTc - dh)));
          },
          std::make_integer_sequence<int, kQ>{});
      // FP64: relative to the full values (the closed form itself rounds at
      // that level); FP32: relative to the deviation
      worst64 = std::fmax(worst64, e64 / (1.0 + dev_scale));
      worst32 = std::fmax(worst32, e32 / dev_scale);
    }
  }
  std::printf("eq_dev FP64 worst |error| / (1 + deviation): %.3e\n", worst64);
  std::printf("eq_dev FP32 worst |error| / deviation: %.3e\n", worst32);
  return (worst64 <= 1e-13 && worst32 <= 1e-5) ? 0 : 1;
}
