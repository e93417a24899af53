// cf_march.hpp -- tile geometry and stage buffers shared by the marching
// stencil kernels (the KG stage kernels and the equation variants).
//
// Each thread owns one (b,c) column of a TC x TB tile and marches along a over a
// chunk of planes; blocks are numbered so that neighbouring tiles land on the
// same XCD when the block count allows.
#pragma once

#include "cf_common.hpp"

namespace cf {

constexpr int TC = 64;  // c extent of a tile (one wavefront)
constexpr int TB = 4;   // b extent of a tile (waves per block)

// Tile geometry shared by the marching kernels.
struct Tiling {
  int64_t N;
  int A;          // planes per a-chunk
  int n_ct, n_bt, n_ac;
  int total;      // n_ct * n_bt * n_ac
};

// Map the hardware block id to (c tile, b tile, a chunk). Blocks are handed to
// the 8 XCDs round robin; when the count allows, consecutive logical tiles are
// given to the same XCD so that the halo rows and planes shared by neighbouring
// tiles meet in one L2.
__device__ __forceinline__ void tile_of(const Tiling &t, int &ct, int &bt, int &ac) {
  int bid = blockIdx.x;
  if ((t.total & 7) == 0) bid = (bid & 7) * (t.total >> 3) + (bid >> 3);
  ct = bid % t.n_ct;
  int r = bid / t.n_ct;
  bt = r % t.n_bt;
  ac = r / t.n_bt;
}

static inline Tiling make_tiling(int64_t N, int A) {
  Tiling t;
  t.N = N;
  t.A = A;
  t.n_ct = (int)cdiv(N, TC);
  t.n_bt = (int)cdiv(N, TB);
  t.n_ac = (int)cdiv(N, A);
  t.total = t.n_ct * t.n_bt * t.n_ac;
  return t;
}

static inline int default_chunk(int64_t N) { return N >= 256 ? 32 : (N >= 64 ? 16 : (int)N); }

// Right-hand side of the pi equation at one site (KG_FRW_equation_kernel /
// Lambda_FRW_equation_kernel, equations_cuda.cu:15-61, in that operation order).
struct RhsCoef {
  double c_h;  // (-3.0) * H(t)
  double c_m;  // m * m
  double lam;  // lambda
  double inv;  // inv_ah_sqr(t)
};

__device__ __forceinline__ double rhs_pi(double phi, double pa1, double pam, double pb1, double pbm,
                                         double pc1, double pcm, double pi, const RhsCoef &k,
                                         bool has_lam) {
  double v = k.c_h * pi - k.c_m * phi;
  if (has_lam) v = v - ((k.lam * phi) * phi) * phi;
  double r = v + k.inv * ((((-6.0 * phi + pa1) + pam) + pb1) + pbm);
  return r + k.inv * (pc1 + pcm);
}

// Stage-specific inputs and outputs. T is the storage type (double, or float
// for the compensated mode); all arithmetic is double.
template <class T>
struct StageIO {
  const T *in;   // stencil input: the state (stage 1) or the previous stage output
  T *x;          // FP64: state (read s1-s3, written s4); F32C: hi (read; written s4)
  float *lo;     // F32C only: low part of the state (read and written in s4)
  T *yout;       // stage output y (s1-s3)
  T *acc;        // FP64: partial RK4 sum; F32C: increment D
  double cy;     // y = x + cy*k
  double ca;     // acc/D update coefficient
};

}  // namespace cf
