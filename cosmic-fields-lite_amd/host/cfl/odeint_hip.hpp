// odeint_hip.hpp -- odeint operations for cfl::hip_vector and a Boost-free
// integrator with odeint's semantics.
//
// hip_operations: the scale_sumK operations odeint's vector_space_algebra
// applies to whole state vectors -- the role of thrust_operations
// (src/odeint_thrust/thrust_operations.cuh:24-353). Each is one HIP launch
// (cf_scale_sum); v0 = a1*v1 + a2*v2 + ... evaluated left to right.
//
// cfl::odeint_lite: integrate_const (external/boost/numeric/odeint/integrate/
// detail/integrate_const.hpp:42-68 with less_eq_with_sign,
// util/detail/less_with_sign.hpp:46-53) and runge_kutta4_classic
// (stepper/runge_kutta4_classic.hpp:87-138 via explicit_stepper_base::do_step_v1,
// explicit_stepper_base.hpp:224-230), for builds without Boost.odeint.
//
// With Boost.odeint, include odeint_boost_traits.hpp instead and use the real
// integrate_const / runge_kutta4_classic.
#pragma once

#include <cstddef>
#include <limits>
#include <utility>

#include "hip_vector.hpp"

namespace cfl {

struct hip_operations {
  typedef hip_vector<double> State;

  template <int K>
  struct scale_sum_base {
    double a[K];
    template <class... Vs>
    void apply(State &v0, const Vs &...vs) const {
      static_assert(sizeof...(Vs) == K, "scale_sum arity");
      const double *p[K] = {vs.data()...};
      check(cf_scale_sum(K, v0.data(), p, a, (int64_t)v0.size(), default_stream()), "scale_sum");
    }
  };

#define CFL_SCALE_SUM(K)                                                            \
  template <class... Fac>                                                           \
  struct scale_sum##K : scale_sum_base<K> {                                         \
    template <class... A>                                                           \
    explicit scale_sum##K(A... alpha) : scale_sum_base<K>{{static_cast<double>(alpha)...}} {} \
    template <class... Vs>                                                          \
    void operator()(State &v0, const Vs &...vs) const {                             \
      this->apply(v0, vs...);                                                       \
    }                                                                               \
  };
  CFL_SCALE_SUM(1)
  CFL_SCALE_SUM(2)
  CFL_SCALE_SUM(3)
  CFL_SCALE_SUM(4)
  CFL_SCALE_SUM(5)
  CFL_SCALE_SUM(6)
  CFL_SCALE_SUM(7)
  CFL_SCALE_SUM(8)
  CFL_SCALE_SUM(9)
  CFL_SCALE_SUM(10)
  CFL_SCALE_SUM(11)
  CFL_SCALE_SUM(12)
  CFL_SCALE_SUM(13)
  CFL_SCALE_SUM(14)
#undef CFL_SCALE_SUM
};

// max |v| (vector_space_norm_inf, src/odeint_thrust/thrust_algebra.hpp:51-64)
inline double norm_inf(const hip_vector<double> &v) {
  double out = 0.0;
  check(cf_norm_inf(v.data(), (int64_t)v.size(), &out, default_stream()), "norm_inf");
  return out;
}

namespace odeint_lite {

inline bool less_eq_with_sign(double t1, double t2, double dt) {
  if (dt > 0) return t1 - t2 <= std::numeric_limits<double>::epsilon();
  return t2 - t1 <= std::numeric_limits<double>::epsilon();
}

// Classic RK4 over any State with hip_operations-style scale_sum2/scale_sum5.
template <class State = hip_vector<double>, class Ops = hip_operations>
class runge_kutta4_classic {
 public:
  typedef State state_type;
  typedef double value_type;
  typedef double time_type;
  static unsigned short order() { return 4; }

  template <class System>
  void do_step(System &sys, State &x, double t, double dt) {
    adjust(x);
    const double val1 = 1.0;
    const double dh = dt / 2.0;
    const double th = t + dh;
    sys(x, m_dxdt, t);
    typename Ops::template scale_sum2<double, double>(val1, dh)(m_x_tmp, x, m_dxdt);
    sys(m_x_tmp, m_dxt, th);
    typename Ops::template scale_sum2<double, double>(val1, dh)(m_x_tmp, x, m_dxt);
    sys(m_x_tmp, m_dxm, th);
    typename Ops::template scale_sum2<double, double>(val1, dt)(m_x_tmp, x, m_dxm);
    sys(m_x_tmp, m_dxh, t + dt);
    const double dt6 = dt / 6.0, dt3 = dt / 3.0;
    typename Ops::template scale_sum5<double, double, double, double, double>(1.0, dt6, dt3, dt3, dt6)(
        x, x, m_dxdt, m_dxt, m_dxm, m_dxh);
  }

 private:
  void adjust(const State &x) {
    for (State *s : {&m_dxdt, &m_dxt, &m_dxm, &m_dxh, &m_x_tmp})
      if (s->size() != x.size()) s->resize(x.size());
  }
  State m_dxdt, m_dxt, m_dxm, m_dxh, m_x_tmp;
};

// integrate_const with an explicit stepper: observer before every step and
// after the last; t_k = t0 + k*dt. Stepper, system and observer are taken by
// value, as odeint does. Returns the number of steps.
template <class Stepper, class System, class State, class Observer>
std::size_t integrate_const(Stepper stepper, System system, State &x, double t0, double t1, double dt,
                            Observer observer) {
  double time = t0;
  std::size_t step = 0;
  while (less_eq_with_sign(time + dt, t1, dt)) {
    observer(x, time);
    stepper.do_step(system, x, time, dt);
    ++step;
    time = t0 + static_cast<double>(step) * dt;
  }
  observer(x, time);
  return step;
}

struct null_observer {
  template <class S>
  void operator()(const S &, double) const {}
};

}  // namespace odeint_lite
}  // namespace cfl
