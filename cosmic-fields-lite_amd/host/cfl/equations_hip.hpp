// equations_hip.hpp -- HIP lattice equations with the reference's Equation
// interface (src/equations.hpp:24-30; src/equations_cuda.cuh:21-47):
//
//   typedef Vector / State / Workspace;  Workspace &workspace;
//   void operator()(const State &x, State &dxdt, const double t);   // odeint system
//   static Vector compute_energy_density(const Workspace &, const double t);
//
// so each is a drop-in beside CudaKleinGordonEquationInFRW. operator() is one
// HIP launch (cf_rhs_f64); the odeint ops are hip_operations. Each equation also
// offers rk4_step(x, t, dt): the whole classic RK4 step as four fused stage
// kernels (cf_rk4_step_f64), bit-identical to sys() + scale_sum2/scale_sum5; and
// rk4_step_into(x, out, t, dt): the same step with the same bits in one
// temporally blocked kernel that reads x once and writes out once.
// hip_runge_kutta4 (stepper.hpp) calls rk4_step_into.
#pragma once

#include <memory>

#include "hip_vector.hpp"
#include "workspace.hpp"

namespace cfl {

namespace detail {

struct KGCore {
  // Scalars of one RHS evaluation at time t (equations.cpp:75-80).
  static void coefs(const WorkspaceGeneric<hip_vector<double>> &w, double t, bool frw, double &H, double &inv) {
    if (frw) {
      const double a_t = w.cosmology.a(t);
      H = w.cosmology.H(t);
      inv = cf_inv_ah_sqr(w.N, w.L, a_t);
    } else {
      H = 0.0;
      inv = cf_inv_ah_sqr(w.N, w.L, 1.0);  // 1/((L/N)^2), equations.cpp:14
    }
  }
  static cf_kg_params params(const WorkspaceGeneric<hip_vector<double>> &w, double lambda, bool frw) {
    cf_kg_params p;
    p.N = w.N;
    p.L = w.L;
    p.m = w.m;
    p.lambda = lambda;
    p.cosmo = frw ? w.cosmology.c() : StaticEOSCosmology().c();
    return p;
  }
};

template <bool FRW, bool LAMBDA>
struct HipKG {
  typedef hip_vector<double> Vector;
  typedef Vector State;
  typedef WorkspaceGeneric<Vector> Workspace;
  Workspace &workspace;

  explicit HipKG(Workspace &workspace_) : workspace(workspace_) {}

  double lambda() const { return LAMBDA ? workspace.lambda : 0.0; }

  void operator()(const State &x, State &dxdt, const double t) {
    double H, inv;
    KGCore::coefs(workspace, t, FRW, H, inv);
    if (dxdt.size() != x.size()) dxdt.resize(x.size());
    check(cf_rhs_f64(workspace.N, x.data(), dxdt.data(), workspace.m, lambda(), H, inv, default_stream()),
          "Equation::operator()");
  }

  static Vector compute_energy_density(const Workspace &w, const double t) {
    double H, inv;
    KGCore::coefs(w, t, FRW, H, inv);
    Vector rho((std::size_t)(w.N * w.N * w.N));
    check(cf_energy_density_f64(w.N, w.state.data(), rho.data(), w.m, LAMBDA ? w.lambda : 0.0, inv,
                                default_stream()),
          "compute_energy_density");
    return rho;
  }

  // Parameters of the fused kernels (cf_rk4_step_f64 / cf_rk4_step_f32c).
  cf_kg_params params() const { return KGCore::params(workspace, lambda(), FRW); }

  // One fused classic RK4 step of x (in place).
  void rk4_step(State &x, double t, double dt) {
    const cf_kg_params p = KGCore::params(workspace, lambda(), FRW);
    const std::size_t wb = cf_rk4_work_bytes(p.N, CF_PREC_F64);
    if (work_->size() != wb) work_->resize(wb);
    check(cf_rk4_step_f64(&p, x.data(), t, dt, work_->data(), default_stream()), "rk4_step");
  }

  // One temporally blocked RK4 step from x into out (same bits as rk4_step).
  void rk4_step_into(const State &x, State &out, double t, double dt) {
    const cf_kg_params p = KGCore::params(workspace, lambda(), FRW);
    if (out.size() != x.size()) out.resize(x.size());
    check(cf_rk4_step_tb_f64(&p, x.data(), out.data(), t, dt, 0, default_stream()), "rk4_step_into");
  }

 private:
  // odeint passes the system by value on every step; copies share the scratch.
  std::shared_ptr<hip_vector<unsigned char>> work_ = std::make_shared<hip_vector<unsigned char>>();
};

// Owns a cf_variant handle (a snapshot of the workspace's Psi or R_fft).
class VariantHandle {
 public:
  VariantHandle(const WorkspaceGeneric<hip_vector<double>> &w, int kind) {
    cf_kg_params p = KGCore::params(w, 0.0, true);
    const double *field = nullptr;
    if (kind == CF_VAR_FIXED_CURVATURE) field = w.Psi.data();
    if (kind == CF_VAR_COMOVING_CURVATURE || kind == CF_VAR_APPROX_COMOVING_CURVATURE) field = w.R_fft.data();
    check(cf_variant_create(&p, kind, w.f_a, w.M, field, &v_, default_stream()), "cf_variant_create");
  }
  ~VariantHandle() { cf_variant_destroy(v_); }
  VariantHandle(const VariantHandle &) = delete;
  VariantHandle &operator=(const VariantHandle &) = delete;
  cf_variant get() const { return v_; }

 private:
  cf_variant v_ = nullptr;
};

// The curvature / potential equations (cf_variant_*). The workspace's Psi
// (fixed curvature) or R_fft (comoving curvature) is copied when the equation is
// constructed; call refresh() after changing it.
template <int KIND>
struct HipVariant {
  typedef hip_vector<double> Vector;
  typedef Vector State;
  typedef WorkspaceGeneric<Vector> Workspace;
  Workspace &workspace;

  explicit HipVariant(Workspace &workspace_)
      : workspace(workspace_), h_(std::make_shared<VariantHandle>(workspace_, KIND)) {}

  void refresh() { h_ = std::make_shared<VariantHandle>(workspace, KIND); }

  void operator()(const State &x, State &dxdt, const double t) {
    if (dxdt.size() != x.size()) dxdt.resize(x.size());
    check(cf_variant_rhs(h_->get(), x.data(), dxdt.data(), t, default_stream()), "Equation::operator()");
  }

  // The reference takes Workspace& for the curvature variants (equations_cuda.cuh:88).
  static Vector compute_energy_density(Workspace &w, const double t) {
    VariantHandle h(w, KIND);
    Vector rho((std::size_t)(w.N * w.N * w.N));
    check(cf_variant_energy_density(h.get(), w.state.data(), rho.data(), t, default_stream()),
          "compute_energy_density");
    return rho;
  }

  void rk4_step(State &x, double t, double dt) {
    const std::size_t wb = cf_rk4_work_bytes(workspace.N, CF_PREC_F64);
    if (work_->size() != wb) work_->resize(wb);
    check(cf_variant_rk4_step(h_->get(), x.data(), t, dt, work_->data(), default_stream()), "rk4_step");
  }

  // One temporally blocked RK4 step from x into out (same bits as rk4_step).
  void rk4_step_into(const State &x, State &out, double t, double dt) {
    if (out.size() != x.size()) out.resize(x.size());
    check(cf_variant_rk4_step_tb(h_->get(), x.data(), out.data(), t, dt, 0, default_stream()), "rk4_step_into");
  }

 private:
  std::shared_ptr<VariantHandle> h_;
  std::shared_ptr<hip_vector<unsigned char>> work_ = std::make_shared<hip_vector<unsigned char>>();
};

}  // namespace detail

// CudaSqrtPotentialEquationInFRW (equations_cuda.cu:83-106, 610-642): f_a from the workspace.
typedef detail::HipVariant<CF_VAR_SQRT_POTENTIAL> HipSqrtPotentialEquationInFRW;
// CudaFixedCurvatureEquationInFRW (equations_cuda.cu:114-137, 645-675): workspace.Psi.
typedef detail::HipVariant<CF_VAR_FIXED_CURVATURE> HipFixedCurvatureEquationInFRW;
// ComovingCurvatureEquationInFRW / CudaComovingCurvatureEquationInFRW
// (equations.cpp:141-322; equations_cuda.cu:678-761): workspace.R_fft.
typedef detail::HipVariant<CF_VAR_COMOVING_CURVATURE> HipComovingCurvatureEquationInFRW;
// CudaApproximateComovingCurvatureEquationInFRW (equations_cuda.cu:764-815):
// workspace.R_fft and workspace.M.
typedef detail::HipVariant<CF_VAR_APPROX_COMOVING_CURVATURE> HipApproximateComovingCurvatureEquationInFRW;

// KleinGordonEquation (equations.cpp:8-36): Minkowski, h = L/N.
typedef detail::HipKG<false, false> HipKleinGordonEquation;
// KleinGordonEquationInFRW / CudaKleinGordonEquationInFRW (equations.cpp:71-138,
// equations_cuda.cu:522-553).
typedef detail::HipKG<true, false> HipKleinGordonEquationInFRW;
// CudaLambdaEquationInFRW (equations_cuda.cu:575-607): + lambda phi^3 in the force,
// + lambda phi^4 / 4 in the energy density.
typedef detail::HipKG<true, true> HipLambdaEquationInFRW;

}  // namespace cfl
