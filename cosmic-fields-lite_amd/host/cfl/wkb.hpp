// wkb.hpp -- WKBSolutionForKleinGordonEquationInFRW (src/wkb.hpp, src/wkb.cpp:5-121)
// on the HIP layer: constructed from a workspace (its state at t_i), evaluate_at(t)
// returns the WKB-extrapolated state as a device vector. Radiation era only
// (cosmology p = 1, t1 = 1/(2 H1)), as the reference asserts.
#pragma once

#include "hip_vector.hpp"
#include "workspace.hpp"

namespace cfl {

class WKBSolutionForKleinGordonEquationInFRW {
 public:
  typedef hip_vector<double> Vector;
  typedef WorkspaceGeneric<Vector> Workspace;

  WKBSolutionForKleinGordonEquationInFRW(Workspace &workspace_, const double t_i_)
      : workspace(workspace_), t_i(t_i_) {
    cf_kg_params p;
    p.N = workspace.N;
    p.L = workspace.L;
    p.m = workspace.m;
    p.lambda = 0.0;
    p.cosmo = workspace.cosmology.c();
    check(cf_wkb_create(&p, workspace.state.data(), t_i, &w_, default_stream()), "cf_wkb_create");
  }
  ~WKBSolutionForKleinGordonEquationInFRW() { cf_wkb_destroy(w_); }
  WKBSolutionForKleinGordonEquationInFRW(const WKBSolutionForKleinGordonEquationInFRW &) = delete;
  WKBSolutionForKleinGordonEquationInFRW &operator=(const WKBSolutionForKleinGordonEquationInFRW &) = delete;

  Vector evaluate_at(const double t) {
    Vector out(workspace.state.size());
    check(cf_wkb_evaluate(w_, t, out.data(), default_stream()), "cf_wkb_evaluate");
    return out;
  }

  Workspace &workspace;
  double t_i;

 private:
  cf_wkb w_ = nullptr;
};

}  // namespace cfl
