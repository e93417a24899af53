// initializer.hpp -- workspace initializers for HIP states, in the reference's
// style (src/initializer.hpp): callables initializer(param, workspace) passed to
// the WorkspaceGeneric constructor.
#pragma once

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

}  // namespace cfl
