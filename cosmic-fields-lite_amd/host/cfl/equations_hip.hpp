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
// kernels (cf_rk4_step_f64), bit-identical to sys() + scale_sum2/scale_sum5;
// hip_runge_kutta4 (stepper.hpp) calls it.
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

 private:
  // odeint passes the system by value on every step; copies share the scratch.
  std::shared_ptr<hip_vector<unsigned char>> work_ = std::make_shared<hip_vector<unsigned char>>();
};

}  // namespace detail

// KleinGordonEquation (equations.cpp:8-36): Minkowski, h = L/N.
typedef detail::HipKG<false, false> HipKleinGordonEquation;
// KleinGordonEquationInFRW / CudaKleinGordonEquationInFRW (equations.cpp:71-138,
// equations_cuda.cu:522-553).
typedef detail::HipKG<true, false> HipKleinGordonEquationInFRW;
// CudaLambdaEquationInFRW (equations_cuda.cu:575-607): + lambda phi^3 in the force,
// + lambda phi^4 / 4 in the energy density.
typedef detail::HipKG<true, true> HipLambdaEquationInFRW;

}  // namespace cfl
