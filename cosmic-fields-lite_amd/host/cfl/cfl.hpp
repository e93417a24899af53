// cfl.hpp -- umbrella header of the C++ host layer over libcfhip.so.
//
// The reference's user-facing API (README.md:15-61) with HIP types:
//   Workspace workspace(param, initializer);      // cfl::WorkspaceGeneric<hip_vector<double>>
//   Equation eqn(workspace);                      // cfl::HipKleinGordonEquationInFRW, ...
//   ConstIntervalObserver<Equation> observer(dir, param, eqn);
//   integrate_const(stepper, eqn, workspace.state, t_start, t_end, dt, observer);
// with stepper = cfl::hip_runge_kutta4<> (fused kernels) or any odeint
// runge_kutta4_classic over hip_operations.
#pragma once

#include "equations_hip.hpp"
#include "fft.hpp"
#include "hip_vector.hpp"
#include "initializer.hpp"
#include "observer.hpp"
#include "odeint_hip.hpp"
#include "physics.hpp"
#include "stepper.hpp"
#include "workspace.hpp"
#include "wkb.hpp"
