// observer.hpp -- observers for HIP workspaces and the raw .dat writers.
//
// ConstIntervalObserver keeps the reference's cadence and file names
// (src/observer.hpp:64-126): at t >= t_last + t_interval, t == t_end or
// t == t_start it writes varphi_plus_spectrum_%d.dat (mode power spectrum),
// rho_spectrum_%d.dat (energy-density power spectrum) and, optionally,
// rho_slice_%d.dat / rho_axis_average_%d.dat, as raw little-endian float64
// (src/io.cpp:46-62). Spectra are computed and binned on the device; only the
// 3(N/2)^2+1 bins cross to the host.
//
// ConstStepObserver saves every K-th step (step-count based, so CPU and GPU
// runs save at identical steps regardless of time rounding).
//
// Both observers have an opt-in finite check (check_finite = true): at every save
// the state is scanned on the device and a NaN or infinity raises
// cfl::NonFiniteState, which ends integrate_const. The reference has no such check
// (it builds with -ffast-math).
#pragma once

#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "fft.hpp"
#include "hip_vector.hpp"

namespace cfl {

// write_vector_to_file / write_VectorXd_to_filename_template (src/io.cpp:30-62)
inline void write_vector_to_file(const std::vector<double> &v, const std::string &filename) {
  std::ofstream f(filename, std::ios::binary);
  if (f.is_open()) f.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(v.size() * sizeof(double)));
}

inline void write_vector_to_filename_template(const std::vector<double> &v, const std::string &fmt, int idx) {
  char name[512];
  std::snprintf(name, sizeof(name), fmt.c_str(), idx);
  write_vector_to_file(v, name);
}

inline std::vector<double> load_vector_from_file(const std::string &filename) {
  std::ifstream f(filename, std::ios::binary | std::ios::ate);
  std::vector<double> v;
  if (!f.is_open()) return v;
  const std::streamsize bytes = f.tellg();
  f.seekg(0);
  v.resize((std::size_t)bytes / sizeof(double));
  f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(v.size() * sizeof(double)));
  return v;
}

// Raised by the observers' finite check.
struct NonFiniteState : std::runtime_error {
  double t;
  long long count;
  NonFiniteState(double t_, long long count_)
      : std::runtime_error("state holds " + std::to_string(count_) + " NaN or infinite values at t = " +
                           std::to_string(t_)),
        t(t_), count(count_) {}
};

// Throws NonFiniteState if v holds a NaN or an infinity.
inline void require_finite(const hip_vector<double> &v, double t) {
  int64_t k = 0;
  check(cf_count_nonfinite(v.data(), (int64_t)v.size(), &k, default_stream()), "cf_count_nonfinite");
  if (k) throw NonFiniteState(t, (long long)k);
}

inline std::vector<double> to_host(const hip_vector<double> &d) {
  std::vector<double> h(d.size());
  copy_vector(h, d);
  return h;
}

// save_radial adds N/2 radial bins (bin j sums the s-bins with floor(sqrt(s)) = j)
// of each saved spectrum as varphi_plus_radial_%d.dat / rho_radial_%d.dat; the
// reference writes the s-bins only.
template <typename Equation, bool save_field_spectrum = true, bool save_density_spectrum = true,
          bool save_density = false, bool save_radial = false>
struct ConstIntervalObserver {
  typedef typename Equation::Workspace Workspace;
  typedef typename Workspace::State State;
  Workspace &workspace;
  int idx;
  std::string dir;
  double t_start, t_end, t_interval, t_last;
  bool check_finite = false;  // scan the state for NaN / inf at every save

  template <typename Param>
  ConstIntervalObserver(const std::string &dir_, const Param &param, Equation &eqn)
      : workspace(eqn.workspace), idx(0), dir(dir_), t_start(param.t_start), t_end(param.t_end),
        t_interval(param.t_interval), t_last(param.t_start) {}

  void operator()(const State &, double t) {
    if (t >= t_last + t_interval || t == t_end || t == t_start) {
      save(t);
      workspace.t_list.push_back(t);
      t_last = t;
      ++idx;
    }
  }

  void save(double t) {
    if (check_finite) require_finite(workspace.state, t);
    const long long N = workspace.N;
    const double L = workspace.L, m = workspace.m;
    const double a_t = workspace.cosmology.a(t);
    if constexpr (save_field_spectrum) {
      hip_vector<double> s = compute_mode_power_spectrum(N, L, m, a_t, workspace.state, workspace.fft_wrapper);
      write_vector_to_filename_template(to_host(s), dir + "varphi_plus_spectrum_%d.dat", idx);
      if constexpr (save_radial)
        write_vector_to_filename_template(to_host(compute_radial_spectrum(N, s, workspace.fft_wrapper)),
                                          dir + "varphi_plus_radial_%d.dat", idx);
    }
    if constexpr (save_density_spectrum || save_density) {
      hip_vector<double> rho = Equation::compute_energy_density(workspace, t);
      if constexpr (save_density) {
        // slice at a = 0 and the average over a (observer.hpp:110-119), both
        // taken on the device: 2 N^2 values cross PCIe instead of N^3. Taken
        // before the spectrum, whose transform consumes rho.
        const std::size_t n2 = (std::size_t)(N * N);
        std::vector<double> slice(n2), avg(n2);
        hip_vector<double> avg_d(n2);
        check(cf_axis_average(N, rho.data(), avg_d.data(), default_stream()), "cf_axis_average");
        check(cf_d2h(slice.data(), rho.data(), n2 * sizeof(double), default_stream()), "rho slice");
        check(cf_d2h(avg.data(), avg_d.data(), n2 * sizeof(double), default_stream()), "rho axis average");
        write_vector_to_filename_template(slice, dir + "rho_slice_%d.dat", idx);
        write_vector_to_filename_template(avg, dir + "rho_axis_average_%d.dat", idx);
      }
      if constexpr (save_density_spectrum) {
        hip_vector<double> s = compute_power_spectrum_consume(N, rho, workspace.fft_wrapper);
        write_vector_to_filename_template(to_host(s), dir + "rho_spectrum_%d.dat", idx);
        if constexpr (save_radial)
          write_vector_to_filename_template(to_host(compute_radial_spectrum(N, s, workspace.fft_wrapper)),
                                            dir + "rho_radial_%d.dat", idx);
      }
    }
  }
};

// Saves at steps 0, K, 2K, ... and at the last step (t_k = t0 + k*dt).
template <typename Equation, bool save_field_spectrum = true, bool save_density_spectrum = true>
struct ConstStepObserver {
  typedef typename Equation::Workspace Workspace;
  typedef typename Workspace::State State;
  ConstIntervalObserver<Equation, save_field_spectrum, save_density_spectrum, false> inner;
  double t0, dt;
  long long every, last_step;

  template <typename Param>
  ConstStepObserver(const std::string &dir, const Param &param, Equation &eqn, long long every_)
      : inner(dir, param, eqn), t0(param.t_start), dt(param.delta_t), every(every_),
        last_step((long long)((param.t_end - param.t_start) / param.delta_t + 0.5)) {}

  void operator()(const State &, double t) {
    const long long k = (long long)((t - t0) / dt + 0.5);
    if (k % every == 0 || k == last_step) {
      inner.save(t);
      inner.workspace.t_list.push_back(t);
      ++inner.idx;
    }
  }
};

}  // namespace cfl
