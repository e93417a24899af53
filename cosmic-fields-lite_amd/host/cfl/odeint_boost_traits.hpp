// odeint_boost_traits.hpp -- makes cfl::hip_vector<double> a Boost.odeint state.
//
// Include after (or instead of) <boost/numeric/odeint.hpp>. These are the
// specialisations the reference provides for thrust::device_vector in
// src/odeint_thrust/: algebra_dispatcher (thrust_algebra_dispatcher.hpp:42-47),
// operations_dispatcher (thrust_operations_dispatcher.hpp:41-45), is_resizeable /
// resize_impl / same_size_impl / copy_impl (thrust_resize.hpp:123-126) and
// vector_space_norm_inf (thrust_algebra.hpp:51-64). With them a HIP equation
// runs under the reference's own integrate_const and runge_kutta4_classic.
#pragma once

#include <boost/numeric/odeint.hpp>

#include "odeint_hip.hpp"

namespace boost {
namespace numeric {
namespace odeint {

template <>
struct algebra_dispatcher<cfl::hip_vector<double>> {
  typedef vector_space_algebra algebra_type;
};

template <>
struct operations_dispatcher<cfl::hip_vector<double>> {
  typedef cfl::hip_operations operations_type;
};

template <>
struct is_resizeable<cfl::hip_vector<double>> {
  typedef boost::true_type type;
  static const bool value = type::value;
};

template <>
struct resize_impl<cfl::hip_vector<double>, cfl::hip_vector<double>> {
  static void resize(cfl::hip_vector<double> &x, const cfl::hip_vector<double> &y) { x.resize(y.size()); }
};

template <>
struct same_size_impl<cfl::hip_vector<double>, cfl::hip_vector<double>> {
  static bool same_size(const cfl::hip_vector<double> &x, const cfl::hip_vector<double> &y) {
    return x.size() == y.size();
  }
};

template <>
struct copy_impl<cfl::hip_vector<double>, cfl::hip_vector<double>> {
  static void copy(const cfl::hip_vector<double> &from, cfl::hip_vector<double> &to) { to = from; }
};

template <>
struct vector_space_norm_inf<cfl::hip_vector<double>> {
  typedef double result_type;
  double operator()(const cfl::hip_vector<double> &v) const { return cfl::norm_inf(v); }
};

}  // namespace odeint
}  // namespace numeric
}  // namespace boost
