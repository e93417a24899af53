// physics.hpp -- StaticEOSCosmology with the reference's interface
// (src/physics.hpp:15-56): a(t), H(t), eta(t) for a constant equation of state.
// The formulas are evaluated by the library's host functions (cf_cosmo_*), the
// same code the fused stepper uses for its stage scalars.
#pragma once

#include "cfhip.h"

namespace cfl {

struct StaticEOSCosmology {
  double a1, H1, t1, p;

  StaticEOSCosmology(double a1_, double H1_, double t1_, double p_) : a1(a1_), H1(H1_), t1(t1_), p(p_) {}
  // radiation domination from a parameter struct with a1, H1, t1 (physics.hpp:25-27)
  template <typename T>
  explicit StaticEOSCosmology(const T &param) : a1(param.a1), H1(param.H1), t1(param.t1), p(1.0) {}
  // Minkowski (physics.hpp:30-31)
  StaticEOSCosmology() : a1(1.0), H1(0), t1(0), p(1.0) {}

  cf_cosmology c() const { return cf_cosmology{a1, H1, t1, p}; }
  double a(double t) const {
    cf_cosmology k = c();
    return cf_cosmo_a(&k, t);
  }
  double H(double t) const {
    cf_cosmology k = c();
    return cf_cosmo_H(&k, t);
  }
  double eta(double t) const {
    cf_cosmology k = c();
    return cf_cosmo_eta(&k, t);
  }
};

}  // namespace cfl
