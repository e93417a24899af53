// workspace.hpp -- WorkspaceGeneric for HIP states, with the reference's fields
// and constructor protocol (src/workspace.hpp:20-94): a parameter struct is
// copied field by field when it has the matching members (C++20 concepts), then
// an initializer fills the state.
#pragma once

#include <cassert>
#include <concepts>
#include <memory>
#include <type_traits>
#include <vector>

#include "fft.hpp"
#include "hip_vector.hpp"
#include "physics.hpp"

namespace cfl {

template <typename P, typename T, typename M>
concept member_of_type = std::same_as<std::remove_cvref_t<M>, T>;

template <typename Param>
concept HasLatticeParams = requires(Param p) {
  requires member_of_type<Param, long long, decltype(p.N)>;
  requires member_of_type<Param, double, decltype(p.L)>;
};
template <typename Param>
concept HasMass = requires(Param p) { requires member_of_type<Param, double, decltype(p.m)>; };
template <typename Param>
concept HasLambda = requires(Param p) { requires member_of_type<Param, double, decltype(p.lambda)>; };
template <typename Param>
concept HasFa = requires(Param p) { requires member_of_type<Param, double, decltype(p.f_a)>; };
template <typename Param>
concept HasFRWParameters = requires(Param p) {
  requires member_of_type<Param, double, decltype(p.a1)>;
  requires member_of_type<Param, double, decltype(p.H1)>;
  requires member_of_type<Param, double, decltype(p.t1)>;
};
template <typename Param>
concept HasPsiApproximationParameters = requires(Param p) {
  requires member_of_type<Param, long long, decltype(p.M)>;
};

// fftWrapperDispatcher<Vector>::Generic (src/dispatcher.hpp:24-46)
template <typename Vector>
struct fftWrapperDispatcher;
template <>
struct fftWrapperDispatcher<hip_vector<double>> {
  typedef hipfftWrapper Generic;
};

template <typename Vector>
struct WorkspaceGeneric {
  typedef Vector State;
  long long int N;
  double L;
  double m{0.0};
  StaticEOSCosmology cosmology{};
  State state;
  double lambda{0};
  double f_a{1.0};
  Vector Psi;
  Vector dPsidt;
  Vector Psi_fft;
  Vector dPsidt_fft;
  Vector R_fft;
  std::vector<double> t_list;
  typename fftWrapperDispatcher<Vector>::Generic fft_wrapper;

  bool Psi_approximation_initialized{false};
  long long int M{0};
  std::unique_ptr<typename fftWrapperDispatcher<Vector>::Generic> fft_wrapper_M_ptr;
  Vector cutoff_R_fft;

  template <HasLatticeParams Param, typename Initializer>
  WorkspaceGeneric(const Param &param, Initializer &&initializer) : N(param.N), L(param.L), fft_wrapper((int)param.N) {
    if constexpr (HasFRWParameters<Param>) cosmology = StaticEOSCosmology(param);
    if constexpr (HasMass<Param>) m = param.m;
    if constexpr (HasLambda<Param>) lambda = param.lambda;
    if constexpr (HasFa<Param>) f_a = param.f_a;
    if constexpr (HasPsiApproximationParameters<Param>) {
      M = param.M;
      assert(N >= M);
    }
    initializer(param, *this);
  }
};

}  // namespace cfl
