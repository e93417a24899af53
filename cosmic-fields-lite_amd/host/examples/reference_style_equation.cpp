// reference_style_equation.cpp -- an equation written against the workspace the
// way the reference writes CudaComovingCurvatureEquationInFRW::operator()
// (src/equations_cuda.cu:678-731): the Psi and dPsi/dt half spectra from
// workspace.R_fft, two in-place C2R transforms through
// workspace.fft_wrapper.execute_inplace_z2d, then the force with Psi and dPsi/dt
// read at PADDED_IDX_OF (Variable_Curvature_FRW_equation_kernel, :263-287; here
// evaluated on the host). Its dx/dt is compared with the library's
// HipComovingCurvatureEquationInFRW at several times.
//
// usage: reference_style_equation N L m a1 H1 Psi_std_dev
// Prints the largest |difference| / max|dx/dt| and exits 1 if it exceeds 1e-12.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <numbers>
#include <vector>

#include "cfl/cfl.hpp"

struct MyParam {
  long long int N;
  double L;
  double m;
  double k_ast;
  double varphi_std_dev;
  double Psi_std_dev;
  double a1;
  double H1;
  double t1;
  double t_start;
  double t_end;
  double t_interval;
  double delta_t;
};

struct RefStyleComovingEquation {
  typedef cfl::hip_vector<double> Vector;
  typedef Vector State;
  typedef cfl::WorkspaceGeneric<Vector> Workspace;
  Workspace &workspace;
  explicit RefStyleComovingEquation(Workspace &w) : workspace(w) {}

  void operator()(const State &x, State &dxdt, const double t) {
    const auto N = workspace.N;
    const auto L = workspace.L;
    const auto m = workspace.m;
    const auto a_t = workspace.cosmology.a(t);
    const auto H_t = workspace.cosmology.H(t);
    const auto eta_t = workspace.cosmology.eta(t);
    const double inv_ah_sqr = 1.0 / ((L / N) * (L / N)) / (a_t * a_t);
    const double factor = (2 * std::numbers::pi / L) * eta_t / std::sqrt(3);
    Vector Psi(workspace.R_fft.size());
    Vector dPsidt(workspace.R_fft.size());
    cfl::check(cf_comoving_psi_fft(N, workspace.R_fft.data(), Psi.data(), dPsidt.data(), factor, H_t,
                                   cfl::default_stream()),
               "Psi, dPsi/dt half spectra");
    // After the in-place C2R, site (a,b,c) of Psi and dPsidt is at PADDED_IDX_OF(N,a,b,c).
    workspace.fft_wrapper.execute_inplace_z2d(Psi);
    workspace.fft_wrapper.execute_inplace_z2d(dPsidt);
    const std::vector<double> hx = cfl::to_host(x), hP = cfl::to_host(Psi), hD = cfl::to_host(dPsidt);
    std::vector<double> out(hx.size());
    const long long N3 = N * N * N;
    auto idx = [&](long long a, long long b, long long c) { return N * N * a + N * b + c; };
    for (long long a = 0; a < N; ++a)
      for (long long b = 0; b < N; ++b)
        for (long long c = 0; c < N; ++c) {
          const long long i1 = idx(a, b, c), i2 = N3 + i1;
          const std::size_t i3 = workspace.fft_wrapper.padded_idx(a, b, c);
          out[i1] = hx[i2];
          out[i2] = (-3.0 * H_t + 4.0 * hD[i3]) * hx[i2] - std::exp(2 * hP[i3]) * m * m * hx[i1] +
                    std::exp(4 * hP[i3]) * inv_ah_sqr *
                        (-6.0 * hx[i1] + hx[idx((a + 1) % N, b, c)] + hx[idx((a + N - 1) % N, b, c)] +
                         hx[idx(a, (b + 1) % N, c)] + hx[idx(a, (b + N - 1) % N, c)] +
                         hx[idx(a, b, (c + 1) % N)] + hx[idx(a, b, (c + N - 1) % N)]);
        }
    dxdt.copy_from_host(out);
  }
};

int main(int argc, char **argv) {
  if (argc != 7) {
    std::fprintf(stderr, "usage: %s N L m a1 H1 Psi_std_dev\n", argv[0]);
    return 2;
  }
  MyParam p{};
  p.N = atoll(argv[1]);
  p.L = atof(argv[2]);
  p.m = atof(argv[3]);
  p.a1 = atof(argv[4]);
  p.H1 = atof(argv[5]);
  p.Psi_std_dev = atof(argv[6]);
  p.k_ast = 1.0;
  p.varphi_std_dev = 1.0;
  p.t1 = 1.0 / (2 * p.H1);
  p.t_start = p.t1;
  p.t_end = p.t1 + 1.0;
  p.t_interval = 1.0;
  p.delta_t = 0.1;
  try {
    typedef RefStyleComovingEquation::Workspace Workspace;
    Workspace workspace(p, cfl::perturbed_grf_and_comoving_curvature_fft{0});
    RefStyleComovingEquation ref(workspace);
    cfl::HipComovingCurvatureEquationInFRW hip(workspace);
    double worst = 0.0;
    for (double dtt : {0.0, 0.37, 2.5}) {
      const double t = p.t_start + dtt;
      cfl::hip_vector<double> d1(workspace.state.size()), d2(workspace.state.size());
      ref(workspace.state, d1, t);
      hip(workspace.state, d2, t);
      const std::vector<double> h1 = cfl::to_host(d1), h2 = cfl::to_host(d2);
      double dmax = 0.0, fmax = 0.0;
      for (std::size_t i = 0; i < h1.size(); ++i) {
        dmax = std::fmax(dmax, std::fabs(h1[i] - h2[i]));
        fmax = std::fmax(fmax, std::fabs(h2[i]));
      }
      std::printf("t=%.3f max|ref-style - library| / max|dx/dt| = %.3e\n", t, dmax / fmax);
      worst = std::fmax(worst, dmax / fmax);
    }
    return worst <= 1e-12 ? 0 : 1;
  } catch (const std::exception &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
}
