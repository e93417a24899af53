// initializer.hpp -- workspace initializers for HIP states, in the reference's
// style (src/initializer.hpp): callables initializer(param, workspace) passed to
// the WorkspaceGeneric constructor.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <numbers>
#include <vector>

#include "hip_vector.hpp"
#include "workspace.hpp"

namespace cfl {

// phi = Gaussian random field with spectrum `kind` (CF_SPEC_*), white noise from
// std::mt19937(seed) (random_field.cpp:13-82), rescaled to rms (<= 0: not
// rescaled); dphi/dt = 0. The normals are drawn on the host, the FFTs and the
// spectrum multiply run on the device.
struct grf_initializer {
  uint32_t seed;
  int kind;
  double params[4];
  double rms;
  template <class Param, class Workspace>
  void operator()(const Param &param, Workspace &w) const {
    w.state.resize((std::size_t)(2 * param.N * param.N * param.N));
    check(cf_init_grf_state(w.fft_wrapper.plan(), seed, kind, params, param.L, rms, w.state.data(),
                            default_stream()),
          "grf_initializer");
  }
};

// homogeneous_field (initializer.hpp:210-224): phi = param.f, dphi/dt = param.dt_f.
inline auto homogeneous_field = [](const auto &param, auto &w) {
  const std::size_t n3 = (std::size_t)(param.N * param.N * param.N);
  std::vector<double> h(2 * n3);
  for (std::size_t i = 0; i < n3; ++i) {
    h[i] = param.f;
    h[n3 + i] = param.dt_f;
  }
  w.state.copy_from_host(h);
};

// plane_wave (initializer.hpp:247-269): phi = cos(2 pi c / N), dphi/dt = 0.
inline auto plane_wave = [](const auto &param, auto &w) {
  const long long N = param.N;
  std::vector<double> h((std::size_t)(2 * N * N * N), 0.0);
  for (long long a = 0; a < N; ++a)
    for (long long b = 0; b < N; ++b)
      for (long long c = 0; c < N; ++c) h[(std::size_t)(N * N * a + N * b + c)] = std::cos(2 * std::numbers::pi * c / N);
  w.state.copy_from_host(h);
};

// Ci(x) by the (12,12) Pade approximant the reference uses
// (special_function.hpp:17-26; coefficients of that approximant).
inline double ci_pade_12_12(double x) {
  using std::pow;
  const double num = 0.5772156649015329 - 0.24231497614160186 * pow(x, 2) + 0.007139183039136621 * pow(x, 4) -
                     0.00011466618094101764 * pow(x, 6) + 8.443734405201243e-7 * pow(x, 8) -
                     3.060472574705558e-9 * pow(x, 10) + 4.328624073851291e-12 * pow(x, 12);
  const double den = 1. + 0.013313955815300189 * pow(x, 2) + 0.00008836441800952094 * pow(x, 4) +
                     3.800404484365274e-7 * pow(x, 6) + 1.1376490214488613e-9 * pow(x, 8) +
                     2.297129602871981e-12 * pow(x, 10) + 2.510407760855278e-15 * pow(x, 12);
  return std::log(x) + num / den;
}

// Amplitude A_s of the scale-invariant comoving curvature spectrum that gives
// Psi the standard deviation Psi_std_dev at conformal time eta_i in a box L
// (initializer.hpp:156-167).
inline double comoving_curvature_amplitude(double L, double eta_i, double Psi_std_dev) {
  using std::pow;
  const double pi = std::numbers::pi, sq3 = std::sqrt(3);
  const double x = (4 * pi * eta_i) / (sq3 * L);
  return (-576 * pow(pi, 6) * pow(eta_i, 6) * pow(Psi_std_dev, 2)) /
         (-81 * pow(L, 4) * (pow(L, 2) + 2 * pow(pi, 2) * pow(eta_i, 2)) +
          L * (81 * pow(L, 5) - 54 * pow(L, 3) * pow(pi, 2) * pow(eta_i, 2) + 48 * L * pow(pi, 4) * pow(eta_i, 4)) *
              std::cos(x) +
          256 * pow(pi, 6) * pow(eta_i, 6) * ci_pade_12_12(x) +
          4 * sq3 * L * pi * eta_i * (27 * pow(L, 4) + 6 * pow(L, 2) * pow(pi, 2) * pow(eta_i, 2) - 16 * pow(pi, 4) * pow(eta_i, 4)) *
              std::sin(x));
}

// Psi_k / R_k at conformal time eta (initializer.hpp:177-179), tabulated over s.
inline std::vector<double> psi_kernel_table(long long N, double L, double eta) {
  const double pi = std::numbers::pi, sq3 = std::sqrt(3);
  std::vector<double> t((std::size_t)cf_spectrum_bins(N));
  for (std::size_t s = 0; s < t.size(); ++s) {
    const double k = std::sqrt((double)s) * (2 * pi / L);
    t[s] = k == 0.0 ? 0.0
                    : (6 * sq3 * (-((k * eta * std::cos((k * eta) / sq3)) / sq3) + std::sin((k * eta) / sq3))) /
                          (std::pow(k, 3) * std::pow(eta, 3));
  }
  return t;
}

// perturbed_grf_and_comoving_curvature_fft (initializer.hpp:150-206): the
// comoving curvature R (scale-invariant, amplitude from Psi_std_dev), its Psi,
// and a field whose variance is modulated by Psi (phi by exp(alpha_phi Psi),
// dphi/dt by exp(alpha_dphi Psi)); leaves R_fft in the workspace for the
// comoving-curvature equations. White noise: one std::mt19937(seed) stream,
// N^3 normals each for R, phi and dphi/dt in that order (the reference's static
// generator after set_generator_seed). FFTs and spectrum multiplies run on the
// device; R, Psi and both fields stay resident.
struct perturbed_grf_and_comoving_curvature_fft {
  uint32_t seed = 0;
  template <class Param, class Workspace>
  void operator()(const Param &param, Workspace &w) const {
    using std::pow;
    const long long N = param.N;
    const std::size_t n3 = (std::size_t)(N * N * N);
    cf_spectrum_plan *plan = w.fft_wrapper.plan();
    const double eta_i = w.cosmology.eta(param.t_start);
    const double A_s = comoving_curvature_amplitude(param.L, eta_i, param.Psi_std_dev);
    std::vector<double> nrm(3 * n3);
    check(cf_normals_host(seed, (int64_t)nrm.size(), nrm.data()), "normals");
    const double scale = 1.0 / (double)(N * N * N);  // random_field.cpp:44-46
    for (auto &v : nrm) v = scale * v;
    auto sqrt_table = [&](int kind, std::array<double, 4> prm) {
      std::vector<double> P((std::size_t)cf_spectrum_bins(N));
      check(cf_spectrum_table_host(kind, N, param.L, prm.data(), P.data()), "spectrum table");
      for (auto &v : P) v = std::sqrt(v);
      hip_vector<double> d;
      d.copy_from_host(P);
      return d;
    };
    hip_vector<double> wbuf(n3), R(n3), Psi(n3), aPsi(n3);
    copy_host_range(nrm.data(), nrm.data() + n3, wbuf);
    hip_vector<double> sP = sqrt_table(CF_SPEC_SCALE_INV, {A_s, 0, 0, 0});
    check(cf_grf(plan, wbuf.data(), sP.data(), nullptr, R.data(), default_stream()), "R");
    hip_vector<double> kern;
    kern.copy_from_host(psi_kernel_table(N, param.L, eta_i));
    check(cf_scaled_fourier_modes(plan, R.data(), kern.data(), Psi.data(), default_stream()), "Psi");
    // (3.17) of the reference's paper, with its factor 0.5 (initializer.hpp:184-188)
    const double v = param.k_ast / (param.a1 * param.m);
    const double den = 12 * pow(v, 4) + 50 * pow(v, 2) + 50;
    const double alpha_phi = 0.5 * (-3 * pow(4 * pow(v, 2) + 5, 2)) / den;
    const double alpha_dphi = 0.5 * (25 - 20 * pow(v, 2)) / den;
    w.state.resize(2 * n3);
    const double *src[1] = {Psi.data()};
    double a1[1] = {alpha_phi};
    check(cf_scale_sum(1, aPsi.data(), src, a1, (int64_t)n3, default_stream()), "alpha Psi");
    copy_host_range(nrm.data() + n3, nrm.data() + 2 * n3, wbuf);
    hip_vector<double> sPf = sqrt_table(CF_SPEC_POWER_CUTOFF, {param.varphi_std_dev, param.k_ast, 0.0, 0});
    check(cf_grf(plan, wbuf.data(), sPf.data(), aPsi.data(), w.state.data(), default_stream()), "phi");
    a1[0] = alpha_dphi;
    check(cf_scale_sum(1, aPsi.data(), src, a1, (int64_t)n3, default_stream()), "alpha Psi");
    copy_host_range(nrm.data() + 2 * n3, nrm.data() + 3 * n3, wbuf);
    hip_vector<double> sPd =
        sqrt_table(CF_SPEC_DERIV_POWER_CUTOFF, {param.varphi_std_dev, param.k_ast, 0.0, param.m});
    check(cf_grf(plan, wbuf.data(), sPd.data(), aPsi.data(), w.state.data() + n3, default_stream()), "dphi");
    w.R_fft = w.fft_wrapper.execute_d2z(R);
    check(cf_sync(default_stream()), "perturbed_grf_and_comoving_curvature_fft");
  }
};

}  // namespace cfl
