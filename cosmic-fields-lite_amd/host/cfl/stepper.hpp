// stepper.hpp -- odeint-compatible steppers that run the fused RK4 kernels.
//
// hip_runge_kutta4 has the interface of an odeint explicit stepper (the shape
// of the reference's own custom stepper, src/midpoint.hpp:25-101): do_step(
// system, x, t, dt). If the system provides rk4_step_into(x, out, t, dt) (every
// HIP equation does) the step is one temporally blocked kernel from x into a
// spare state, and then x and the spare trade buffers (x.swap), so x's storage
// alternates between two allocations; else, if it provides rk4_step(x, t, dt),
// four fused stage kernels in place; otherwise the classic sys() + scale_sum
// sequence. All three give the same bits. Works with
// cfl::odeint_lite::integrate_const and, when Boost.odeint is present, with
// boost::numeric::odeint::integrate_const.
//
// hip_runge_kutta4_f32c steps a compensated_state: fields stored as float
// hi + lo, stage inputs and the RK4 increment stored as float, all arithmetic
// in double (the temporally blocked cf_rk4_step_tb_f32c, same bits as
// cf_rk4_step_f32c), with the same buffer swap. hip_runge_kutta4_f32s steps the
// same state with float forces and a double increment (cf_rk4_step_tb_f32s; tune
// its pace first with cf_rk4_tb_tune_pace for speed, the bits do not depend on it).
#pragma once

#if __has_include(<boost/numeric/odeint/stepper/stepper_categories.hpp>)
#include <boost/numeric/odeint/stepper/stepper_categories.hpp>
#define CFL_HAVE_ODEINT_CATEGORIES 1
#endif

#include <memory>

#include "equations_hip.hpp"
#include "odeint_hip.hpp"

namespace cfl {

template <class State = hip_vector<double>>
class hip_runge_kutta4 {
 public:
  typedef State state_type;
  typedef State deriv_type;
  typedef double value_type;
  typedef double time_type;
  typedef unsigned short order_type;
#ifdef CFL_HAVE_ODEINT_CATEGORIES
  typedef boost::numeric::odeint::stepper_tag stepper_category;
#endif
  static order_type order() { return 4; }

  template <class System>
  void do_step(System system, State &x, double t, double dt) {
    if constexpr (requires { system.rk4_step_into(x, x, t, dt); }) {
      // one kernel into the spare buffer, then the two buffers trade places
      system.rk4_step_into(x, *next_, t, dt);
      x.swap(*next_);
    } else if constexpr (requires { system.rk4_step(x, t, dt); }) {
      system.rk4_step(x, t, dt);
    } else {
      fallback_->do_step(system, x, t, dt);
    }
  }

 private:
  std::shared_ptr<State> next_ = std::make_shared<State>();
  std::shared_ptr<odeint_lite::runge_kutta4_classic<State>> fallback_ =
      std::make_shared<odeint_lite::runge_kutta4_classic<State>>();
};

// State of the FP32-compensated mode: the field value is hi + lo.
struct compensated_state {
  hip_vector<float> hi, lo;
  std::size_t size() const { return hi.size(); }
};

inline compensated_state to_compensated(const hip_vector<double> &x) {
  compensated_state s;
  s.hi.resize(x.size());
  s.lo.resize(x.size());
  check(cf_split_f64(x.data(), s.hi.data(), s.lo.data(), (int64_t)x.size(), default_stream()), "to_compensated");
  return s;
}

inline void to_f64(const compensated_state &s, hip_vector<double> &x) {
  if (x.size() != s.size()) x.resize(s.size());
  check(cf_join_f64(s.hi.data(), s.lo.data(), x.data(), (int64_t)s.size(), default_stream()), "to_f64");
}

template <bool FLOAT_FORCES>
class hip_runge_kutta4_compensated {
 public:
  typedef compensated_state state_type;
  typedef double value_type;
  typedef double time_type;
  typedef unsigned short order_type;
#ifdef CFL_HAVE_ODEINT_CATEGORIES
  typedef boost::numeric::odeint::stepper_tag stepper_category;
#endif
  static order_type order() { return 4; }

  template <class System>
  void do_step(System system, compensated_state &x, double t, double dt) {
    const cf_kg_params p = system.params();
    compensated_state &o = *next_;
    if (o.size() != x.size()) {
      o.hi.resize(x.size());
      o.lo.resize(x.size());
    }
    // the temporally blocked kernel, then a swap
    if constexpr (FLOAT_FORCES)
      check(cf_rk4_step_tb_f32s(&p, x.hi.data(), x.lo.data(), o.hi.data(), o.lo.data(), t, dt, 0, default_stream()),
            "hip_runge_kutta4_f32s");
    else
      check(cf_rk4_step_tb_f32c(&p, x.hi.data(), x.lo.data(), o.hi.data(), o.lo.data(), t, dt, 0, default_stream()),
            "hip_runge_kutta4_f32c");
    x.hi.swap(o.hi);
    x.lo.swap(o.lo);
  }

 private:
  std::shared_ptr<compensated_state> next_ = std::make_shared<compensated_state>();
};
typedef hip_runge_kutta4_compensated<false> hip_runge_kutta4_f32c;
typedef hip_runge_kutta4_compensated<true> hip_runge_kutta4_f32s;

}  // namespace cfl
