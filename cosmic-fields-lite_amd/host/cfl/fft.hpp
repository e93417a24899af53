// fft.hpp -- hipFFT wrapper with the reference's execute_* surface, and the
// spectrum functions the observer calls, for hip_vector<double>.
//
// hipfftWrapper mirrors cufftWrapperNoBatching (src/cuda_wrapper.cuh:113-130,
// cuda_wrapper.cu:202-255) and fftwWrapper (src/fftw_wrapper.hpp:20-70): D2Z and
// Z2D plans of an N^3 lattice sharing one work area; non-copyable.
// compute_power_spectrum / compute_mode_power_spectrum mirror the CUDA overloads
// (src/fdm3d_cuda.cuh:22-31) and return device vectors of 3(N/2)^2+1 bins,
// binned deterministically (no atomics).
#pragma once

#include <cstdint>
#include <stdexcept>

#include "hip_vector.hpp"

namespace cfl {

class hipfftWrapper {
 public:
  int N;
  explicit hipfftWrapper(int N_) : N(N_) { check(cf_spectrum_plan_create(N_, &plan_), "hipfftWrapper"); }
  ~hipfftWrapper() { cf_spectrum_plan_destroy(plan_); }
  hipfftWrapper(const hipfftWrapper &) = delete;
  hipfftWrapper &operator=(const hipfftWrapper &) = delete;
  hipfftWrapper(hipfftWrapper &&) = delete;
  hipfftWrapper &operator=(hipfftWrapper &&) = delete;

  cf_spectrum_plan *plan() const { return plan_; }
  std::size_t half_size() const { return (std::size_t)N * N * (N / 2 + 1) * 2; }

  // in: N^3 reals -> 2 N^2 (N/2+1) doubles (interleaved complex); in is preserved.
  hip_vector<double> execute_d2z(const hip_vector<double> &in) {
    hip_vector<double> out(half_size());
    check(cf_fft_d2z(plan_, in.data(), out.data(), default_stream()), "execute_d2z");
    return out;
  }
  // in: 2N^3 reals (phi | pi) -> two half spectra back to back (cuda_wrapper.cu:237-243).
  hip_vector<double> execute_batched_d2z(const hip_vector<double> &in) {
    hip_vector<double> out(2 * half_size());
    const std::size_t n3 = (std::size_t)N * N * N;
    check(cf_fft_d2z(plan_, in.data(), out.data(), default_stream()), "execute_batched_d2z");
    check(cf_fft_d2z(plan_, in.data() + n3, out.data() + half_size(), default_stream()), "execute_batched_d2z");
    return out;
  }
  // Unnormalised C2R; destroys in (as FFTW / cuFFT).
  hip_vector<double> execute_z2d(hip_vector<double> &in) {
    hip_vector<double> out((std::size_t)N * N * N);
    execute_z2d(in, out);
    return out;
  }
  void execute_z2d(hip_vector<double> &in, hip_vector<double> &out) {
    if (out.size() != (std::size_t)N * N * N) out.resize((std::size_t)N * N * N);
    check(cf_fft_z2d(plan_, in.data(), out.data(), default_stream()), "execute_z2d");
  }
  // In-place C2R (cuda_wrapper.cu:252-254): inout keeps its size 2 N^2 (N/2+1); the
  // N^3 results sit in the FFTW padded layout, site (a,b,c) at padded_idx(a,b,c).
  void execute_inplace_z2d(hip_vector<double> &inout) {
    if (inout.size() != half_size()) throw std::runtime_error("execute_inplace_z2d: size is not 2 N^2 (N/2+1)");
    check(cf_fft_z2d_inplace(plan_, inout.data(), default_stream()), "execute_inplace_z2d");
  }
  // PADDED_IDX_OF (fdm3d.hpp:24)
  std::size_t padded_idx(long long a, long long b, long long c) const {
    return (std::size_t)(2 * (N / 2 + 1)) * ((std::size_t)N * a + b) + c;
  }

 private:
  cf_spectrum_plan *plan_ = nullptr;
};

inline std::size_t spectrum_size(long long N) { return (std::size_t)cf_spectrum_bins(N); }

// compute_power_spectrum (fdm3d.cpp:7-34 / fdm3d_cuda.cu:161-176)
inline hip_vector<double> compute_power_spectrum(const long long N, const hip_vector<double> &f,
                                                 hipfftWrapper &fft) {
  hip_vector<double> out(spectrum_size(N));
  check(cf_power_spectrum(fft.plan(), f.data(), out.data(), default_stream()), "compute_power_spectrum");
  return out;
}

// The same bins from a field the caller is done with: f is the transform's input
// without the copy that preserves it, and its contents are undefined afterwards.
inline hip_vector<double> compute_power_spectrum_consume(const long long N, hip_vector<double> &f,
                                                         hipfftWrapper &fft) {
  hip_vector<double> out(spectrum_size(N));
  check(cf_power_spectrum_consume(fft.plan(), f.data(), out.data(), default_stream()),
        "compute_power_spectrum_consume");
  return out;
}

// compute_mode_power_spectrum (fdm3d.cpp:36-72 / fdm3d_cuda.cu:144-159)
inline hip_vector<double> compute_mode_power_spectrum(const long long N, const double L, const double m,
                                                      const double a_t, const hip_vector<double> &state,
                                                      hipfftWrapper &fft) {
  hip_vector<double> out(spectrum_size(N));
  check(cf_mode_power_spectrum(fft.plan(), state.data(), L, m, a_t, out.data(), default_stream()),
        "compute_mode_power_spectrum");
  return out;
}

// N/2 radial bins from the s-bins: bin j sums s with floor(sqrt(s)) = j.
inline hip_vector<double> compute_radial_spectrum(const long long N, const hip_vector<double> &sbins,
                                                  hipfftWrapper &fft) {
  hip_vector<double> out((std::size_t)(N / 2));
  check(cf_radial_bins(fft.plan(), sbins.data(), out.data(), default_stream()), "compute_radial_spectrum");
  return out;
}

}  // namespace cfl
