// comoving_run.cpp -- the reference's main run (src/main.cpp:82-196: a scalar field
// with a comoving-curvature metric perturbation) on the HIP layer: parameter
// struct -> workspace(param, perturbed_grf_and_comoving_curvature_fft) ->
// comoving-curvature equation -> ConstIntervalObserver (spectra + density
// slices) -> integrate_const -> t_list.dat and the final state.
//
// usage: comoving_run <out_dir/> <equation> <stepper> N L m k_ast varphi_std_dev Psi_std_dev
//                     a1 H1 t_end t_interval delta_t M
//   equation: comoving (HipComovingCurvatureEquationInFRW, the reference's
//             ComovingCurvatureEquationInFRW) | approx (HipApproximateComovingCurvatureEquationInFRW)
//   stepper:  classic (odeint_lite::runge_kutta4_classic) | fused (cfl::hip_runge_kutta4)
//   t1 = 1/(2 H1) and t_start = t1 (radiation era, main.cpp:111-113); the white
//   noise comes from std::mt19937(0) as after RandomNormal::set_generator_seed(0).
// Writes initial_R_spectrum.dat, initial_Psi_spectrum.dat (main.cpp:155-177),
// state0.dat, the observer's files, t_list.dat and state.dat.
#include <cstdio>
#include <cstdlib>
#include <string>

#include "cfl/cfl.hpp"

// The fields of MyParam (main.cpp:36-64).
struct MyParam {
  long long int N;
  double L;
  double m;
  double lambda;
  double f_a;
  double k_ast;
  double k_Psi;
  double varphi_std_dev;
  double Psi_std_dev;
  double a1;
  double H1;
  double t1;
  double t_start;
  double t_end;
  double t_interval;
  double delta_t;
  long long int M;
  double f;
  double delta_varphi_std_dev;
  double k_delta_varphi;
};

template <class Equation>
static int run(const std::string &dir, const std::string &stepper, const MyParam &param) {
  typedef typename Equation::Workspace Workspace;
  Workspace workspace(param, cfl::perturbed_grf_and_comoving_curvature_fft{0});
  cfl::write_vector_to_file(cfl::to_host(workspace.state), dir + "state0.dat");
  {
    // Spectra of R and of the initial Psi (main.cpp:155-177): R from R_fft by the
    // C2R transform / N^3, Psi from R by the scaled-mode transform.
    const long long N = param.N;
    const double eta_i = workspace.cosmology.eta(param.t_start);
    cfl::hip_vector<double> Rk = workspace.R_fft;
    cfl::hip_vector<double> R = workspace.fft_wrapper.execute_z2d(Rk);
    // times 1/N^3 (the reference divides by N^3: the same for powers of two,
    // within one rounding otherwise)
    const double *src[1] = {R.data()};
    const double inv_n3[1] = {1.0 / std::pow((double)N, 3)};
    cfl::check(cf_scale_sum(1, R.data(), src, inv_n3, (int64_t)R.size(), cfl::default_stream()), "R / N^3");
    cfl::hip_vector<double> kern;
    kern.copy_from_host(cfl::psi_kernel_table(N, param.L, eta_i));
    cfl::hip_vector<double> Psi(R.size());
    cfl::check(cf_scaled_fourier_modes(workspace.fft_wrapper.plan(), R.data(), kern.data(), Psi.data(),
                                       cfl::default_stream()),
               "Psi");
    cfl::write_vector_to_file(cfl::to_host(cfl::compute_power_spectrum(N, Psi, workspace.fft_wrapper)),
                              dir + "initial_Psi_spectrum.dat");
    cfl::write_vector_to_file(cfl::to_host(cfl::compute_power_spectrum(N, R, workspace.fft_wrapper)),
                              dir + "initial_R_spectrum.dat");
  }
  Equation eqn(workspace);
  cfl::ConstIntervalObserver<Equation, true, true, true> observer(dir, param, eqn);
  std::size_t steps = 0;
  if (stepper == "classic") {
    steps = cfl::odeint_lite::integrate_const(cfl::odeint_lite::runge_kutta4_classic<>(), eqn, workspace.state,
                                              param.t_start, param.t_end, param.delta_t, observer);
  } else if (stepper == "fused") {
    steps = cfl::odeint_lite::integrate_const(cfl::hip_runge_kutta4<>(), eqn, workspace.state, param.t_start,
                                              param.t_end, param.delta_t, observer);
  } else {
    std::fprintf(stderr, "unknown stepper %s\n", stepper.c_str());
    return 2;
  }
  cfl::write_vector_to_file(workspace.t_list, dir + "t_list.dat");
  cfl::write_vector_to_file(cfl::to_host(workspace.state), dir + "state.dat");
  std::printf("steps %zu saves %zu\n", steps, workspace.t_list.size());
  return 0;
}

int main(int argc, char **argv) {
  if (argc != 16) {
    std::fprintf(stderr,
                 "usage: %s out_dir/ equation stepper N L m k_ast varphi_std_dev Psi_std_dev a1 H1 t_end "
                 "t_interval delta_t M\n",
                 argv[0]);
    return 2;
  }
  const std::string dir = argv[1], eq = argv[2], stepper = argv[3];
  MyParam p{};
  p.N = atoll(argv[4]);
  p.L = atof(argv[5]);
  p.m = atof(argv[6]);
  p.k_ast = atof(argv[7]);
  p.varphi_std_dev = atof(argv[8]);
  p.Psi_std_dev = atof(argv[9]);
  p.a1 = atof(argv[10]);
  p.H1 = atof(argv[11]);
  p.t1 = 1.0 / (2 * p.H1);
  p.t_start = p.t1;
  p.t_end = atof(argv[12]);
  p.t_interval = atof(argv[13]);
  p.delta_t = atof(argv[14]);
  p.M = atoll(argv[15]);
  try {
    if (eq == "comoving") return run<cfl::HipComovingCurvatureEquationInFRW>(dir, stepper, p);
    if (eq == "approx") return run<cfl::HipApproximateComovingCurvatureEquationInFRW>(dir, stepper, p);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  std::fprintf(stderr, "unknown equation %s\n", eq.c_str());
  return 2;
}
