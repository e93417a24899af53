// kg_run.cpp -- the reference's driver pattern (src/main.cpp:82-196) on the HIP
// layer: parameter struct -> workspace(param, initializer) -> equation ->
// ConstIntervalObserver -> integrate_const -> t_list.dat and the final state.
//
// usage: kg_run <out_dir/> <stepper> <equation> N L m lambda a1 H1 t1 t_start t_end t_interval dt
//               seed spectrum_kind spectrum_param rms
//   stepper:  classic (odeint_lite::runge_kutta4_classic over hip_operations)
//             fused   (cfl::hip_runge_kutta4: fused stage kernels)
//             f32c    (cfl::hip_runge_kutta4_f32c: FP32-compensated storage)
//             f32s    (cfl::hip_runge_kutta4_f32s: the same storage, float forces, tuned pace)
//   equation: kg | frw | lambda | sqrt (sqrt potential; the lambda argument is f_a)
// Writes state0.dat (initial state), the observer's spectra (s-bins and N/2 radial bins)
// and density slices, t_list.dat
// and state.dat; for frw in a radiation era with t1 = 1/(2 H1), also wkb_state.dat (the
// final state extrapolated by WKB to t_end + 10).
#include <cstdio>
#include <cstdlib>
#include <string>
#include <type_traits>

#include "cfl/cfl.hpp"

struct RunParam {
  long long int N;
  double L;
  double m;
  double lambda;
  double a1;
  double H1;
  double t1;
  double t_start;
  double t_end;
  double t_interval;
  double delta_t;
  double f_a;
};

template <class Equation>
static int run(const std::string &dir, const std::string &stepper, const RunParam &param,
               const cfl::grf_initializer &init) {
  typedef typename Equation::Workspace Workspace;
  Workspace workspace(param, init);
  cfl::write_vector_to_file(cfl::to_host(workspace.state), dir + "state0.dat");
  Equation eqn(workspace);
  cfl::ConstIntervalObserver<Equation, true, true, true, true> observer(dir, param, eqn);
  std::size_t steps = 0;
  if (stepper == "classic") {
    steps = cfl::odeint_lite::integrate_const(cfl::odeint_lite::runge_kutta4_classic<>(), eqn, workspace.state,
                                              param.t_start, param.t_end, param.delta_t, observer);
  } else if (stepper == "fused") {
    steps = cfl::odeint_lite::integrate_const(cfl::hip_runge_kutta4<>(), eqn, workspace.state, param.t_start,
                                              param.t_end, param.delta_t, observer);
  } else if constexpr (!requires { eqn.params(); }) {
    std::fprintf(stderr, "stepper %s is not available for this equation\n", stepper.c_str());
    return 2;
  } else if (stepper == "f32c" || stepper == "f32s") {
    // Compensated state; the observer sees the joined double state.
    cfl::compensated_state xs = cfl::to_compensated(workspace.state);
    auto obs = [&](const cfl::compensated_state &s, double t) {
      cfl::to_f64(s, workspace.state);
      observer(workspace.state, t);
    };
    if (stepper == "f32c") {
      steps = cfl::odeint_lite::integrate_const(cfl::hip_runge_kutta4_f32c(), eqn, xs, param.t_start, param.t_end,
                                                param.delta_t, obs);
    } else {
      // the pace of the float-force step for this lattice (timing only; the state is
      // left as it is)
      const cf_kg_params kp = eqn.params();
      cfl::hip_vector<unsigned char> work(cf_rk4_tb_work_bytes(kp.N, CF_PREC_F32S));
      int pace = 0;
      cfl::check(cf_rk4_tb_tune_pace(&kp, xs.hi.data(), xs.lo.data(), param.t_start, param.delta_t, 0, work.data(),
                                     cfl::default_stream(), &pace),
                 "cf_rk4_tb_tune_pace");
      steps = cfl::odeint_lite::integrate_const(cfl::hip_runge_kutta4_f32s(), eqn, xs, param.t_start, param.t_end,
                                                param.delta_t, obs);
    }
    cfl::to_f64(xs, workspace.state);
  } else {
    std::fprintf(stderr, "unknown stepper %s\n", stepper.c_str());
    return 2;
  }
  cfl::write_vector_to_file(workspace.t_list, dir + "t_list.dat");
  cfl::write_vector_to_file(cfl::to_host(workspace.state), dir + "state.dat");
  if constexpr (std::is_same_v<Equation, cfl::HipKleinGordonEquationInFRW>) {
    // WKB extrapolation of the final state to t_end + 10 (generate_wkb_solutions,
    // main.cpp:199-275), where the cosmology allows it (radiation era, t1 = 1/(2 H1))
    if (param.H1 > 0 && param.t1 == 1.0 / (2 * param.H1) && param.m > 0) {
      cfl::WKBSolutionForKleinGordonEquationInFRW wkb(workspace, param.t_end);
      cfl::write_vector_to_file(cfl::to_host(wkb.evaluate_at(param.t_end + 10.0)), dir + "wkb_state.dat");
    }
  }
  std::printf("steps %zu saves %zu\n", steps, workspace.t_list.size());
  return 0;
}

int main(int argc, char **argv) {
  if (argc != 19) {
    std::fprintf(stderr,
                 "usage: %s out_dir/ stepper equation N L m lambda a1 H1 t1 t_start t_end t_interval dt "
                 "seed spectrum_kind spectrum_param rms\n",
                 argv[0]);
    return 2;
  }
  const std::string dir = argv[1], stepper = argv[2], eq = argv[3];
  RunParam p{atoll(argv[4]), atof(argv[5]), atof(argv[6]), atof(argv[7]), atof(argv[8]), atof(argv[9]),
             atof(argv[10]), atof(argv[11]), atof(argv[12]), atof(argv[13]), atof(argv[14]), atof(argv[7])};
  cfl::grf_initializer init{(uint32_t)atol(argv[15]), atoi(argv[16]), {atof(argv[17]), 0, 0, 0}, atof(argv[18])};
  try {
    if (eq == "kg") return run<cfl::HipKleinGordonEquation>(dir, stepper, p, init);
    if (eq == "frw") return run<cfl::HipKleinGordonEquationInFRW>(dir, stepper, p, init);
    if (eq == "lambda") return run<cfl::HipLambdaEquationInFRW>(dir, stepper, p, init);
    if (eq == "sqrt") return run<cfl::HipSqrtPotentialEquationInFRW>(dir, stepper, p, init);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  std::fprintf(stderr, "unknown equation %s\n", eq.c_str());
  return 2;
}
