// ref_odeint_driver.cpp -- TEST INFRASTRUCTURE ONLY (built into oracle/_ref/).
//
// Runs the reference's own integrator headers, compiled from /root/reference:
//   external/boost/numeric/odeint (vendored Boost 1.84): integrate_const,
//     runge_kutta4_classic, vector_space_algebra for Eigen;
//   src/odeint_eigen/eigen_operations.hpp: the scale_sum ops the reference uses;
//   src/physics.hpp: StaticEOSCosmology.
// The right-hand side is the oracle's orc_rhs (the reference's equations.cpp
// cannot be compiled here: it includes fftw3.h, which this image lacks).
//
// Usage: odeint_ref <in_state.bin> <out_dir> N L m lambda a1 H1 t1 t0 t_end dt [threads]
//   in_state.bin: 2N^3 little-endian doubles. threads (default 1): threads of the
//   right-hand side's a-loop (the integrator's scale_sum passes stay on one). Writes out_dir/state.bin (final
//   state), out_dir/times.bin (every time the observer was called) and
//   out_dir/cosmo.bin (a, H, eta at 16 sample times t0 + k*(t_end-t0)/15).
//   Prints the number of steps on stdout and "elapsed_s <seconds>" of the
//   integrate_const call (std::chrono::steady_clock) on stderr.
#include <boost/numeric/odeint.hpp>
#include <boost/numeric/odeint/external/eigen/eigen.hpp>

#include <Eigen/Dense>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "odeint_eigen/eigen_operations.hpp"
#include "physics.hpp"

extern "C" void orc_rhs(int64_t N, const double *x, double *dxdt, double m, double lambda, double H,
                        double inv_ah2);
extern "C" void orc_set_threads(int n);

struct System {
  int64_t N;
  double L, m, lambda;
  StaticEOSCosmology cosmo;
  void operator()(const Eigen::VectorXd &x, Eigen::VectorXd &dxdt, const double t) {
    const double a_t = cosmo.a(t);
    const double H_t = cosmo.H(t);
    const double inv_ah_sqr = 1.0 / ((L / N) * (L / N)) / (a_t * a_t);  // equations.cpp:80
    orc_rhs(N, x.data(), dxdt.data(), m, lambda, H_t, inv_ah_sqr);
  }
};

struct Recorder {
  std::vector<double> *times;
  void operator()(const Eigen::VectorXd &, double t) { times->push_back(t); }
};

static bool write_all(const std::string &path, const double *p, size_t n) {
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) return false;
  bool ok = fwrite(p, sizeof(double), n, f) == n;
  fclose(f);
  return ok;
}

int main(int argc, char **argv) {
  using namespace boost::numeric::odeint;
  if (argc != 13 && argc != 14) {
    fprintf(stderr, "usage: %s in_state.bin out_dir N L m lambda a1 H1 t1 t0 t_end dt [threads]\n", argv[0]);
    return 2;
  }
  orc_set_threads(argc == 14 ? atoi(argv[13]) : 1);
  const std::string in = argv[1], out = argv[2];
  const int64_t N = atoll(argv[3]);
  const double L = atof(argv[4]), m = atof(argv[5]), lambda = atof(argv[6]);
  const double a1 = atof(argv[7]), H1 = atof(argv[8]), t1 = atof(argv[9]);
  const double t0 = atof(argv[10]), t_end = atof(argv[11]), dt = atof(argv[12]);
  const size_t n = 2 * (size_t)N * N * N;
  Eigen::VectorXd x(n);
  FILE *f = fopen(in.c_str(), "rb");
  if (!f || fread(x.data(), sizeof(double), n, f) != n) {
    fprintf(stderr, "cannot read %s\n", in.c_str());
    return 1;
  }
  fclose(f);
  System sys{N, L, m, lambda, (H1 == 0.0) ? StaticEOSCosmology() : StaticEOSCosmology(a1, H1, t1, 1.0)};
  std::vector<double> times;
  auto stepper = runge_kutta4_classic<Eigen::VectorXd, double, Eigen::VectorXd, double>();
  const auto c0 = std::chrono::steady_clock::now();
  size_t steps = integrate_const(stepper, sys, x, t0, t_end, dt, Recorder{&times});
  const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - c0).count();
  fprintf(stderr, "elapsed_s %.6f\n", el);
  std::vector<double> cz;
  for (int k = 0; k < 16; ++k) {
    const double t = t0 + k * (t_end - t0) / 15.0;
    cz.push_back(sys.cosmo.a(t));
    cz.push_back(sys.cosmo.H(t));
    cz.push_back(sys.cosmo.H1 != 0.0 ? sys.cosmo.eta(t) : 0.0);
  }
  if (!write_all(out + "/state.bin", x.data(), n) || !write_all(out + "/times.bin", times.data(), times.size()) ||
      !write_all(out + "/cosmo.bin", cz.data(), cz.size())) {
    fprintf(stderr, "cannot write outputs to %s\n", out.c_str());
    return 1;
  }
  printf("%zu\n", steps);
  return 0;
}
