// observer_finite_check.cpp -- the observers' opt-in finite check on a GPU
// workspace: a clean state saves normally; a NaN or an infinity planted in the
// state makes the next save raise cfl::NonFiniteState with the number of bad
// values, also from inside integrate_const.
//
// usage: observer_finite_check <out_dir/>      (exit 0 when every check holds)
#include <cmath>
#include <cstdio>
#include <limits>
#include <string>

#include "cfl/cfl.hpp"

struct RunParam {
  long long int N;
  double L;
  double m;
  double a1;
  double H1;
  double t1;
  double t_start;
  double t_end;
  double t_interval;
  double delta_t;
};

static void plant(cfl::hip_vector<double> &x, std::size_t i, double v) {
  cfl::check(cf_h2d(x.data() + i, &v, sizeof(double), cfl::default_stream()), "plant");
}

int main(int argc, char **argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: %s out_dir/\n", argv[0]);
    return 2;
  }
  const std::string dir = argv[1];
  RunParam p{16, 16.0, 1.0, 1.0, 0.5, 1.0, 1.0, 2.0, 0.5, 0.05};
  typedef cfl::HipKleinGordonEquationInFRW Equation;
  Equation::Workspace workspace(p, cfl::grf_initializer{3, CF_SPEC_GAUSS, {0.5, 0, 0, 0}, 0.5});
  Equation eqn(workspace);
  cfl::ConstIntervalObserver<Equation> obs(dir, p, eqn);
  obs.check_finite = true;
  int failures = 0;
  auto expect_throw = [&](long long count, const char *what) {
    try {
      obs(workspace.state, p.t_start);
      std::printf("FAIL %s: no exception\n", what);
      ++failures;
    } catch (const cfl::NonFiniteState &e) {
      const bool ok = e.count == count && e.t == p.t_start;
      std::printf("%s %s: %s\n", ok ? "ok" : "FAIL", what, e.what());
      failures += !ok;
    }
  };
  try {
    obs(workspace.state, p.t_start);  // clean state: saves, no exception
    std::printf("ok clean state saved (t_list size %zu)\n", workspace.t_list.size());
  } catch (const std::exception &e) {
    std::printf("FAIL clean state: %s\n", e.what());
    ++failures;
  }
  const std::size_t n = workspace.state.size();
  cfl::hip_vector<double> saved = workspace.state;
  plant(workspace.state, n - 1, std::numeric_limits<double>::infinity());
  expect_throw(1, "inf in the last pi value");
  plant(workspace.state, 0, std::numeric_limits<double>::quiet_NaN());
  expect_throw(2, "and NaN in the first phi value");
  plant(workspace.state, n / 2 + 7, -std::numeric_limits<double>::infinity());
  expect_throw(3, "and -inf in pi");
  // integrate_const: the exception from the save at t_start ends the run
  workspace.state = saved;
  plant(workspace.state, 5, std::numeric_limits<double>::quiet_NaN());
  workspace.t_list.clear();
  cfl::ConstIntervalObserver<Equation> obs2(dir, p, eqn);
  obs2.check_finite = true;
  try {
    cfl::odeint_lite::integrate_const(cfl::hip_runge_kutta4<>(), eqn, workspace.state, p.t_start, p.t_end,
                                      p.delta_t, obs2);
    std::printf("FAIL integrate_const: no exception\n");
    ++failures;
  } catch (const cfl::NonFiniteState &e) {
    std::printf("ok integrate_const stopped: %s\n", e.what());
  }
  return failures ? 1 : 0;
}
