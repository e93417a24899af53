// ref_odeint_hip_dropin.cpp -- TEST INFRASTRUCTURE ONLY (built into oracle/_ref/).
//
// The HIP equation of the product (cfl::HipKleinGordonEquationInFRW /
// cfl::HipLambdaEquationInFRW over libcfhip.so) driven by the reference's own
// vendored Boost.odeint (external/boost/numeric/odeint, Boost 1.84):
// integrate_const + runge_kutta4_classic<hip_vector<double>, double,
// hip_vector<double>, double>, exactly as src/main.cpp:153,183 drives
// CudaKleinGordonEquationInFRW. The only glue is cfl/odeint_boost_traits.hpp.
//
// usage: odeint_hip_dropin <in_state.bin> <out_dir> N L m lambda a1 H1 t1 t0 t_end dt
// Writes out_dir/state.bin and out_dir/times.bin (observer call times).
#include <boost/numeric/odeint.hpp>

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "cfl/cfl.hpp"
#include "cfl/odeint_boost_traits.hpp"

struct Param {
  long long int N;
  double L;
  double m;
  double lambda;
  double a1;
  double H1;
  double t1;
};

struct Recorder {
  std::vector<double> *times;
  void operator()(const cfl::hip_vector<double> &, double t) { times->push_back(t); }
};

template <class Equation>
static int run(const std::vector<double> &x0, const std::string &out, const Param &p, double t0, double t1,
               double dt) {
  using namespace boost::numeric::odeint;
  auto init = [&](const Param &, auto &w) { w.state.copy_from_host(x0); };
  typename Equation::Workspace workspace(p, init);
  Equation eqn(workspace);
  std::vector<double> times;
  auto stepper = runge_kutta4_classic<cfl::hip_vector<double>, double, cfl::hip_vector<double>, double>();
  std::size_t steps = integrate_const(stepper, eqn, workspace.state, t0, t1, dt, Recorder{&times});
  std::vector<double> x = cfl::to_host(workspace.state);
  cfl::write_vector_to_file(x, out + "/state.bin");
  cfl::write_vector_to_file(times, out + "/times.bin");
  std::printf("%zu\n", steps);
  return 0;
}

int main(int argc, char **argv) {
  if (argc != 13) {
    std::fprintf(stderr, "usage: %s in_state.bin out_dir N L m lambda a1 H1 t1 t0 t_end dt\n", argv[0]);
    return 2;
  }
  Param p{atoll(argv[3]), atof(argv[4]), atof(argv[5]), atof(argv[6]), atof(argv[7]), atof(argv[8]),
          atof(argv[9])};
  std::vector<double> x0 = cfl::load_vector_from_file(argv[1]);
  if (x0.size() != (std::size_t)(2 * p.N * p.N * p.N)) {
    std::fprintf(stderr, "bad input state\n");
    return 1;
  }
  try {
    if (p.lambda != 0.0) return run<cfl::HipLambdaEquationInFRW>(x0, argv[2], p, atof(argv[10]), atof(argv[11]), atof(argv[12]));
    return run<cfl::HipKleinGordonEquationInFRW>(x0, argv[2], p, atof(argv[10]), atof(argv[11]), atof(argv[12]));
  } catch (const std::exception &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
}
