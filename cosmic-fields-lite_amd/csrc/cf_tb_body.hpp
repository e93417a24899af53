// cf_tb_body.hpp -- the temporally blocked RK4 step of one work unit (an output
// tile of one lattice over a chunk of a-planes), shared by the single-lattice
// kernel (cf_rk4_tb.hip) and the persistent ensemble kernel (cf_ensemble.hip).
//
// Geometry. A block is 32 x 32 threads over a 32 x 32 (b, c) region of the
// lattice: an output tile of T x T sites (T = 24, or N for small lattices) plus
// a halo of 4 sites on every side, one per stencil stage. Other region shapes
// RB x RC (rows x columns, RB RC <= 1024) give tiles of (RB - 8) x (RC - 8); with
// RB RC < 1024 the threads past the region idle through the unit. The block marches
// along a through a chunk of planes [a0, a1). Each thread owns one (b, c)
// column and keeps, in registers, the short history of planes every stage
// needs along a:
//
//   iteration j loads x at plane P = a0 - 4 + j, then computes
//     stage 1 at P-1 (needs x at P-2..P)       -> y1, acc
//     stage 2 at P-2 (needs y1 at P-3..P-1)    -> y2, acc
//     stage 3 at P-3 (needs y2 at P-4..P-2)    -> y3, acc
//     stage 4 at P-4 (needs y3 at P-5..P-3)    -> new state at P-4 (stored for j >= 8)
//
// so a chunk of A planes takes A + 8 iterations. The b+-1 neighbours of each
// stage input come from plane buffers in LDS (other wavefronts hold those rows),
// the c+-1 neighbours from the adjacent lanes of the same wavefront (DPP wave
// shifts) or, with LANE_NBR = false, from the same LDS buffers. The four
// stage inputs of an iteration are all known when it starts (the a+1 neighbour
// each stage needs from the current iteration is in registers), so they are
// written together and one barrier per iteration suffices. The region shrinks by
// one site per stage; only the inner T x T columns store.
//
// Arithmetic. Per site and stage the operations, their order and their
// rounding are those of the fused stage kernels (cf_stencil.hip): FP64 keeps
// acc = 1*x + dt/6 k1 + dt/3 k2 + dt/3 k3 and ends with acc + dt/6 k4; the
// compensated FP32 mode rounds y and the increment D to float after every
// stage and folds (hi + lo) + (D + dt/6 k4) in double. The results are
// therefore bit-identical to the four-kernel path. With SA (the single-precision
// stencil, compensated storage only) each stage's force is float arithmetic on the
// float stage inputs, y = fma(c, k, hi) in float, the increment D is kept unrounded
// in double (D = fma(dt/3, k, D)) and the new state is (hi + lo) + fma(dt/6, k4, D).
// That scheme has no four-kernel form; the oracle restates it (orc_rk4_step_f32s).
#pragma once

#include <type_traits>

#include "cf_common.hpp"

namespace cf {

constexpr int TBR = 32;  // region edge (threads per row, rows per block)
constexpr int TBH = 4;   // halo: one site per stencil stage

// Scalars of one step: per stage (-3H) and inv_ah_sqr at the stage time, and
// the step's constants.
struct TbStep {
  double c_h[4], inv[4];
  double c_m, lam, dh, dt, dt6, dt3;
  double m, fa2;  // equation variants: mass, f_a^2
};

// The scalars of the single-precision stencil (the compensated mode with float
// right-hand sides): the RHS coefficients and the stage-input step lengths of
// TbStep, rounded to float. The increment weights dt/6, dt/3 stay double.
struct TbStepF {
  float c_h[4], inv[4];
  float c_m, lam, dh, dt;
};

// Equation of the step body: the Klein-Gordon family (+ lambda phi^3), or the
// fixed-curvature / sqrt-potential variants in the operation order of their CUDA
// kernels (Fixed_Curvature_FRW_equation_kernel, equations_cuda.cu:114-137;
// Sqrt_Potential_FRW_equation_kernel, :83-106), as in csrc/cf_variants.hip.
enum : int { TB_KG = 0, TB_FIXED_CURVATURE = 1, TB_SQRT_POTENTIAL = 2, TB_COMOVING = 3, TB_APPROX_COMOVING = 4 };

// Comoving curvature: Psi | dPsi/dt ([Psi (G^3) | dPsi (G^3)]) at the step's three
// stage times (0: t, 1: t + dt/2 for stages 2 and 3, 2: t + dt); G = N (TB_COMOVING)
// or the coarse grid M (TB_APPROX_COMOVING, site (a,b,c) reads cell
// ((x + hs) % N) / sd per axis, sd = N / M, hs = sd / 2).
struct TbPsi {
  const double *f[3];
  int64_t G, sd, hs;
};

// Pacing. Neighbouring tiles read each other's edge rows and columns of every plane;
// those reads hit the L2 only while the two blocks march through the same planes at
// about the same time. Blocks that start together drift apart once an iteration's
// compute is short against its memory latency, and their halo reads then go to HBM.
// With PACE every wavefront of a block waits, before the barrier of iteration j, until
// pace * j ticks of the 100 MHz real-time counter have passed since the block began,
// so blocks cannot run ahead of the schedule (blocks that fall behind it catch up at
// full speed). Pacing changes timing only, never results.

// One work unit: lattice size, output tile origin and edges (T along b, Tc along
// c), plane chunk, and the planes NA that the state buffers hold along a (N for a
// periodic lattice; a slab of a larger lattice holds its planes plus ghost planes,
// and the unit's planes stay within it). Plane indices wrap at NA.
struct TbUnit {
  int64_t N;
  int64_t b0, c0, a0, a1;
  int T, Tc;
  int64_t NA;
  int pace;  // ticks per plane iteration (bodies built with PACE; see Pacing above)
};

__device__ __forceinline__ double tb_rhs(double phi, double pa1, double pam, double pb1, double pbm, double pc1,
                                         double pcm, double pi, double c_h, double c_m, double lam, double inv,
                                         bool has_lam) {
  double v = c_h * pi - c_m * phi;
  if (has_lam) v = v - ((lam * phi) * phi) * phi;
  double r = v + inv * ((((-6.0 * phi + pa1) + pam) + pb1) + pbm);
  return r + inv * (pc1 + pcm);
}

// The Klein-Gordon force in float (single-precision stencil): the seven-point sum
// in one chain (a, b, c neighbours; -6 phi + phi_a+ fused), then inv_ah_sqr times the
// sum fused onto (-3H) pi - m^2 phi (itself one fused multiply-add).
template <bool LAM>
__device__ __forceinline__ float tb_rhs_f32(float phi, float pa1, float pam, float pb1, float pbm, float pc1,
                                            float pcm, float pi, float c_h, float c_m, float lam, float inv) {
  float v = __builtin_fmaf(c_h, pi, -(c_m * phi));
  if constexpr (LAM) v = __builtin_fmaf(-((lam * phi) * phi), phi, v);
  const float s7 = ((((__builtin_fmaf(-6.0f, phi, pa1) + pam) + pb1) + pbm) + pc1) + pcm;
  return __builtin_fmaf(inv, s7, v);
}

template <int VAR, bool LAM>
__device__ __forceinline__ double tb_force(double phi, double pa1, double pam, double pb1, double pbm, double pc1,
                                           double pcm, double pi, double c_h, double inv, const TbStep &k, double ps,
                                           double dp = 0.0) {
  if constexpr (VAR == TB_KG) {
    return tb_rhs(phi, pa1, pam, pb1, pbm, pc1, pcm, pi, c_h, k.c_m, k.lam, inv, LAM);
  } else if constexpr (VAR == TB_FIXED_CURVATURE) {
    const double e2 = exp(2.0 * ps), e4 = exp(4.0 * ps);
    const double s7 = (((((-6.0 * phi + pa1) + pam) + pb1) + pbm) + pc1) + pcm;
    return (c_h * pi - ((e2 * k.m) * k.m) * phi) + (e4 * inv) * s7;
  } else if constexpr (VAR == TB_COMOVING) {
    const double e2 = exp(2.0 * ps), e4 = exp(4.0 * ps);
    const double ci = e4 * inv;
    const double v = ((c_h + 4.0 * dp) * pi - ((e2 * k.m) * k.m) * phi) +
                     ci * ((((-6.0 * phi + pa1) + pam) + pb1) + pbm);
    return v + ci * (pc1 + pcm);
  } else if constexpr (VAR == TB_APPROX_COMOVING) {
    const double s7 = (((((-6.0 * phi + pa1) + pam) + pb1) + pbm) + pc1) + pcm;
    return ((c_h + 4.0 * dp) * pi - ((((1.0 + 2.0 * ps) + (2.0 * ps) * ps) * k.m) * k.m) * phi) +
           (((1.0 + 4.0 * ps) + (8.0 * ps) * ps) * inv) * s7;
  } else {
    const double s7 = (((((-6.0 * phi + pa1) + pam) + pb1) + pbm) + pc1) + pcm;
    return (c_h * pi - (k.c_m * phi) / sqrt(1.0 + (phi * phi) / k.fa2)) + inv * s7;
  }
}

// Logical tile of hardware block b of a grid of `total` blocks. The dispatcher deals
// blocks to the 8 XCDs in turn (b -> XCD b % 8, its (b / 8)-th block); XCD x is given
// the contiguous logical range [x q + min(x, r), ...) of q or q + 1 tiles (q, r =
// total / 8, total % 8), so neighbouring tiles share one L2. A bijection for any total.
__device__ __forceinline__ int tb_xcd_order(int b, int total) {
  const int x = b & 7, i = b >> 3, q = total >> 3, r = total & 7;
  return x * q + (x < r ? x : r) + i;
}

// float rounding of the compensated mode (identity for FP64)
template <bool F32C>
__device__ __forceinline__ double rnd(double v) {
  if constexpr (F32C) return (double)(float)v;
  else return v;
}

// Buffer-resource access to one lattice plane (gfx9 raw buffer, stride 0,
// num_records = plane bytes; word 3 = 0x00020000 as for all gfx9 parts). An
// offset of TB_OOB is out of range: stores to it are dropped, loads return 0.
constexpr uint32_t TB_OOB = 0x80000000u;
typedef unsigned int tb_u2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t tb_plane(const void *base, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, bytes, 0x00020000);
}
__device__ __forceinline__ void tb_store(double v, __amdgpu_buffer_rsrc_t r, uint32_t off) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(tb_u2, v), r, (int)off, 0, 0);
}
__device__ __forceinline__ void tb_store(float v, __amdgpu_buffer_rsrc_t r, uint32_t off) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, (int)off, 0, 0);
}
__device__ __forceinline__ float tb_load_f32(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}

// The c+1 and c-1 neighbours of a stage input are the values of the lanes to the
// right and left in the same wavefront (a wavefront holds two rows of 32 columns),
// moved with DPP wave shifts instead of read back from LDS. The lanes at a row's
// ends receive a value from the other row (or 0); they are halo columns whose
// results are never used, exactly as with clamped LDS indices.
template <int CTRL>
__device__ __forceinline__ int tb_dpp(int v) {
  // bound_ctrl: a lane whose source is outside the wavefront reads 0 (no "old"
  // value to initialise first)
  return __builtin_amdgcn_mov_dpp(v, CTRL, 0xf, 0xf, true);
}
template <int CTRL, class V>
__device__ __forceinline__ V tb_lane_shift(V v) {
  if constexpr (sizeof(V) == 4) {
    return __builtin_bit_cast(V, tb_dpp<CTRL>(__builtin_bit_cast(int, v)));
  } else {
    int2 w = __builtin_bit_cast(int2, v);
    w.x = tb_dpp<CTRL>(w.x);
    w.y = tb_dpp<CTRL>(w.y);
    return __builtin_bit_cast(V, w);
  }
}
template <class V>
__device__ __forceinline__ V tb_lane_next(V v) {  // lane + 1 (DPP wave_shl:1)
  return tb_lane_shift<0x130>(v);
}
template <class V>
__device__ __forceinline__ V tb_lane_prev(V v) {  // lane - 1 (DPP wave_shr:1)
  return tb_lane_shift<0x138>(v);
}

// The new state of one iteration, stored two iterations later.
template <class T>
struct TbPending {
  T f, p;         // phi, pi (FP64) or their hi parts (F32C)
  float lf, lp;   // F32C: lo parts
  uint32_t off;   // lane offset in the plane, or TB_OOB
  int64_t plane;  // plane offset (times N^2)
};

// Two sets of four plane buffers: iteration j writes the four stage inputs
// (all known when it starts: they are the previous iteration's outputs) into set
// j&1, then one barrier. Set j&1 was last read in iteration j-2, which every
// thread finished before the barrier of iteration j-1. A caller running several
// units in a row must barrier between them.
template <class T>
using TbLds = T[2][4][TBR][TBR];

template <class T, bool F32C, bool LAM, bool DEFER = true, bool LANE_NBR = true, int VAR = TB_KG, bool RING = false,
          int RB = TBR, int RC = TBR, bool SA = false, bool PACE = false>
__device__ __forceinline__ void tb_unit(TbLds<T> &lds_all, const TbUnit &u, const TbStep &k,
                                        const T *__restrict__ xin, const float *__restrict__ lin,
                                        T *__restrict__ xout, float *__restrict__ lout,
                                        const double *__restrict__ psi = nullptr, const TbPsi cps = TbPsi{},
                                        const TbStepF kf32 = TbStepF{}) {
  static_assert(VAR == TB_KG || !F32C, "the equation variants are FP64");
  static_assert(!SA || (F32C && VAR == TB_KG), "the single-precision stencil is a compensated Klein-Gordon mode");
  static_assert(RB * RC <= TBR * TBR && RB > 2 * TBH && RC > 2 * TBH, "region shape");
  static_assert(!LANE_NBR || RC == TBR, "lane neighbours need rows of 32 lanes");
  // the plane buffers of this shape, in the storage of the 32 x 32 ones
  T(&lds)[2][4][RB][RC] = *reinterpret_cast<T(*)[2][4][RB][RC]>(&lds_all);
  const int tid = threadIdx.x;
  // threads past the region (RB RC < 1024) act as copies of its last row that
  // write no plane buffer and store nothing
  constexpr bool FULL = RB * RC == TBR * TBR;
  const bool live = FULL || tid < RB * RC;
  const int r = live ? tid / RC : RB - 1, cc = live ? tid % RC : RC - 1;
  static_assert(!F32C || sizeof(T) == sizeof(float), "the compensated mode stores float");
  // In the compensated mode every value carried from plane to plane is exactly
  // representable in float (the state hi, the rounded stage inputs and the rounded
  // increment), so the histories are held as float (half the registers) and
  // widened where they are used; the arithmetic stays double.
  typedef std::conditional_t<F32C, float, double> Hv;
  // Stage arithmetic (double; float with SA) and the increment's type (SA keeps the
  // increment D unrounded in double).
  typedef std::conditional_t<SA, float, double> Av;
  typedef std::conditional_t<SA, double, Hv> Dv;
  const int64_t N = u.N, NN = N * N, NA = u.NA, N3 = NN * NA;  // N3: offset of the pi planes
  const int plane_bytes = (int)(NN * (int64_t)sizeof(T)), plane_bytes_f = (int)(NN * (int64_t)sizeof(float));
  int64_t b = (u.b0 - TBH + r) % N;
  if (b < 0) b += N;
  int64_t c = (u.c0 - TBH + cc) % N;
  if (c < 0) c += N;
  const int64_t col = b * N + c;
  const bool out_ok = live && r >= TBH && r < TBH + u.T && u.b0 + (r - TBH) < N && cc >= TBH &&
                      cc < TBH + u.Tc && u.c0 + (cc - TBH) < N;
  const int rm = r > 0 ? r - 1 : 0, rp = r < RB - 1 ? r + 1 : RB - 1;
  const int cm = cc > 0 ? cc - 1 : 0, cp = cc < RC - 1 ? cc + 1 : RC - 1;
  // c+1 / c-1 neighbour of a stage input: from the adjacent lanes (LANE_NBR) or
  // from the LDS plane buffer (the same values; which is faster depends on the
  // kernel's mix of LDS and VALU work)
  auto cnext = [&](Hv v, const T(*pl)[RC]) -> Av {
    if constexpr (LANE_NBR) return (Av)tb_lane_next(v);
    else return (Av)pl[r][cp];
  };
  auto cprev = [&](Hv v, const T(*pl)[RC]) -> Av {
    if constexpr (LANE_NBR) return (Av)tb_lane_prev(v);
    else return (Av)pl[r][cm];
  };

  // plane offsets (times N^2) advanced by one plane per iteration, wrapping at N
  auto wrap_plane = [&](int64_t P) -> int64_t {
    int64_t q = P % NA;
    if (q < 0) q += NA;
    return q * NN;
  };
  auto next_plane = [&](int64_t off) -> int64_t {
    off += NN;
    return off == N3 ? 0 : off;
  };

  // Plane histories are arrays of TBU entries; the value d planes behind the
  // iteration's phase PH sits at [(PH - d) mod TBU]. Ring XF / XP holds x by plane (P
  // at [PH], P-d at [PH-d], the prefetched P+1, P+2 at [PH+1], [PH+2]); Y1F.. hold
  // the stage inputs made at P-1 by the iteration that made them; AF / AP the
  // partial sum (FP64) or increment D (F32C). The longest-lived value (x at P+2 ..
  // P-4) spans 7 iterations, hence TBU = 8. Two ways to advance one plane:
  //  RING: the loop is unrolled by TBU and PH = j mod TBU is a compile-time constant
  //        in each copy, so advancing renames registers (no moves; larger code);
  //  shift (default): PH = 0 throughout and every array rotates by one entry after
  //        each iteration (register moves), loop unrolled by two.
  // Which is faster depends on the kernel's register pressure and instruction mix:
  // the curvature variant gains from RING (it would otherwise spill), the plain
  // Klein-Gordon kernels do not.
  constexpr int TBU = 8;
  Hv XF[TBU] = {}, XP[TBU] = {};
  Hv Y1F[TBU] = {}, Y1P[TBU] = {}, Y2F[TBU] = {}, Y2P[TBU] = {}, Y3F[TBU] = {}, Y3P[TBU] = {};
  Dv AF[TBU] = {}, AP[TBU] = {};
  double PS[TBU] = {};  // fixed curvature: Psi rides with x (same ring, same planes)

  const int iters = (int)(u.a1 - u.a0) + 8;
  // Memory pipeline. Iteration j computes with x at plane P (loaded two iterations
  // earlier) and loads plane P+2. The new state of iteration j is stored at the
  // start of iteration j+2 (DEFER = false stores at the end of the iteration, for
  // callers with less register room), so every wait is for an operation issued at
  // least an iteration earlier. Stores go through buffer resources over one plane:
  // lanes that do not store get an offset past its end, which the hardware drops
  // (loads of it return 0), so no branch surrounds a memory operation. With RING
  // the loop runs a multiple of TBU iterations; those past `iters` store nothing.
  int64_t off_ld = wrap_plane(u.a0 - TBH);
  XF[0] = (Hv)xin[off_ld + col];
  XP[0] = (Hv)xin[N3 + off_ld + col];
  if constexpr (VAR == TB_FIXED_CURVATURE) PS[0] = psi[off_ld + col];
  off_ld = next_plane(off_ld);
  XF[1] = (Hv)xin[off_ld + col];
  XP[1] = (Hv)xin[N3 + off_ld + col];
  if constexpr (VAR == TB_FIXED_CURVATURE) PS[1] = psi[off_ld + col];
  int64_t off_st = wrap_plane(u.a0 - TBH - 4);  // plane P-4
  int64_t a_st = off_st / NN;                    // its index
  // comoving curvature: Psi and dPsi of stage s (plane P-s) from time slot `slot`
  int64_t cell_bc = 0;
  if constexpr (VAR == TB_APPROX_COMOVING) {
    const int64_t jb = ((b + cps.hs) % N) / cps.sd, jc = ((c + cps.hs) % N) / cps.sd;
    cell_bc = jb * cps.G + jc;
  }
  auto cpsi = [&](int s, int slot, double &ps, double &dp) {
    if constexpr (VAR == TB_COMOVING) {
      int64_t pl = off_st + (int64_t)(4 - s) * NN;
      if (pl >= N3) pl -= N3;
      ps = cps.f[slot][pl + col];
      dp = cps.f[slot][N3 + pl + col];
    } else if constexpr (VAR == TB_APPROX_COMOVING) {
      // 32-bit index arithmetic (N <= 2048)
      const int n = (int)N, sd = (int)cps.sd, hs = (int)cps.hs;
      int as = (int)a_st + (4 - s);
      if (as >= n) as -= n;
      int ah = as + hs;
      if (ah >= n) ah -= n;
      const int64_t j = (int64_t)(ah / sd) * (cps.G * cps.G) + cell_bc;
      ps = cps.f[slot][j];
      dp = cps.f[slot][cps.G * cps.G * cps.G + j];
    }
  };
  TbPending<T> Q[2] = {};
  Q[0].off = Q[1].off = TB_OOB;
  // F32C: the compensation term of output plane P-4 is prefetched the same way
  // (slot j&1, loaded two iterations ahead; lanes that will not store read 0).
  float LF[2] = {0.0f, 0.0f}, LP[2] = {0.0f, 0.0f};
  int64_t off_lo = off_st;
  auto lo_off = [&](int jj) -> uint32_t {
    return ((jj >= 8) && (jj < iters) && out_ok) ? (uint32_t)(col * (int64_t)sizeof(float)) : TB_OOB;
  };
  if constexpr (F32C) {
    LF[0] = tb_load_f32(tb_plane(lin + off_lo, plane_bytes_f), lo_off(0));
    LP[0] = tb_load_f32(tb_plane(lin + N3 + off_lo, plane_bytes_f), lo_off(0));
    off_lo = next_plane(off_lo);
    LF[1] = tb_load_f32(tb_plane(lin + off_lo, plane_bytes_f), lo_off(1));
    LP[1] = tb_load_f32(tb_plane(lin + N3 + off_lo, plane_bytes_f), lo_off(1));
  }

  auto flush = [&](const TbPending<T> &q) {
    tb_store(q.f, tb_plane(xout + q.plane, plane_bytes), q.off);
    tb_store(q.p, tb_plane(xout + N3 + q.plane, plane_bytes), q.off);
    if constexpr (F32C) {
      tb_store(q.lf, tb_plane(lout + q.plane, plane_bytes_f), q.off);
      tb_store(q.lp, tb_plane(lout + N3 + q.plane, plane_bytes_f), q.off);
    }
  };

  const uint64_t t_begin = PACE ? __builtin_amdgcn_s_memrealtime() : 0;
  auto iteration = [&](auto phase, auto slot, const int j) {
    constexpr int PH = decltype(phase)::value;
    constexpr int I0 = PH % TBU, I1 = (PH + TBU - 1) % TBU, I2 = (PH + TBU - 2) % TBU;
    constexpr int I3 = (PH + TBU - 3) % TBU, I4 = (PH + TBU - 4) % TBU, IN = (PH + 2) % TBU;
    constexpr int S = decltype(slot)::value;  // j & 1: LDS set, pending store, lo slot
    // RING: keep the scheduler within one iteration (interleaving copies of the
    // unrolled body lengthens live ranges past the register budget of 1024-thread
    // blocks)
    if constexpr (RING) __builtin_amdgcn_sched_barrier(0);
    T(*buf)[RB][RC] = lds[S];
    TbPending<T> &q = Q[S];
    double lo_f = 0.0, lo_p = 0.0;
    if constexpr (F32C) {
      lo_f = (double)LF[S];
      lo_p = (double)LP[S];
    }
    if constexpr (DEFER) flush(q);  // the new state of iteration j-2
    const uint32_t st_off = ((j >= 8) && (j < iters) && out_ok) ? (uint32_t)(col * (int64_t)sizeof(T)) : TB_OOB;
    if constexpr (F32C) {
      off_lo = next_plane(off_lo);  // plane P-2, the output plane of iteration j+2
      LF[S] = tb_load_f32(tb_plane(lin + off_lo, plane_bytes_f), lo_off(j + 2));
      LP[S] = tb_load_f32(tb_plane(lin + N3 + off_lo, plane_bytes_f), lo_off(j + 2));
    }
    off_ld = next_plane(off_ld);  // plane P+2
    XF[IN] = (Hv)xin[off_ld + col];
    XP[IN] = (Hv)xin[N3 + off_ld + col];
    if constexpr (VAR == TB_FIXED_CURVATURE) PS[IN] = psi[off_ld + col];

    if (live) {
      buf[0][r][cc] = (T)XF[I1];
      buf[1][r][cc] = (T)Y1F[I1];
      buf[2][r][cc] = (T)Y2F[I1];
      buf[3][r][cc] = (T)Y3F[I1];
    }
    if constexpr (PACE) {
      const uint64_t due = t_begin + (uint64_t)j * (uint64_t)u.pace;
      while (__builtin_amdgcn_s_memrealtime() < due) __builtin_amdgcn_s_sleep(1);
    }
    __syncthreads();

    // the stages' Psi: fixed curvature from its plane history, comoving from the
    // stage time's field at the stage's plane
    double ps1 = 0.0, dp1 = 0.0, ps2 = 0.0, dp2 = 0.0, ps3 = 0.0, dp3 = 0.0, ps4 = 0.0, dp4 = 0.0;
    if constexpr (VAR == TB_FIXED_CURVATURE) {
      ps1 = PS[I1];
      ps2 = PS[I2];
      ps3 = PS[I3];
      ps4 = PS[I4];
    } else if constexpr (VAR == TB_COMOVING || VAR == TB_APPROX_COMOVING) {
      cpsi(1, 0, ps1, dp1);
      cpsi(2, 1, ps2, dp2);
      cpsi(3, 1, ps3, dp3);
      cpsi(4, 2, ps4, dp4);
    }

    if constexpr (SA) {
      // Single-precision stencil: each stage's force in float on the float stage
      // inputs, y = hi + c k in float, the increment D = dt/6 k1 + dt/3 k2 + dt/3 k3
      // in double, and the new state (hi + lo) + (D + dt/6 k4) in double; every
      // product-then-sum is one fused multiply-add (17% fewer instructions per plane,
      // 97.3 against 95.0 GSite*step/s paced at N=512).
      float kf = XP[I1];
      float kp = tb_rhs_f32<LAM>(XF[I1], XF[I0], XF[I2], buf[0][rp][cc], buf[0][rm][cc], cnext(XF[I1], buf[0]),
                                 cprev(XF[I1], buf[0]), XP[I1], kf32.c_h[0], kf32.c_m, kf32.lam, kf32.inv[0]);
      Y1F[I0] = __builtin_fmaf(kf32.dh, kf, XF[I1]);
      Y1P[I0] = __builtin_fmaf(kf32.dh, kp, XP[I1]);
      AF[I0] = k.dt6 * (double)kf;
      AP[I0] = k.dt6 * (double)kp;

      kf = Y1P[I1];
      kp = tb_rhs_f32<LAM>(Y1F[I1], Y1F[I0], Y1F[I2], buf[1][rp][cc], buf[1][rm][cc], cnext(Y1F[I1], buf[1]),
                           cprev(Y1F[I1], buf[1]), Y1P[I1], kf32.c_h[1], kf32.c_m, kf32.lam, kf32.inv[1]);
      Y2F[I0] = __builtin_fmaf(kf32.dh, kf, XF[I2]);
      Y2P[I0] = __builtin_fmaf(kf32.dh, kp, XP[I2]);
      AF[I1] = __builtin_fma(k.dt3, (double)kf, AF[I1]);
      AP[I1] = __builtin_fma(k.dt3, (double)kp, AP[I1]);

      kf = Y2P[I1];
      kp = tb_rhs_f32<LAM>(Y2F[I1], Y2F[I0], Y2F[I2], buf[2][rp][cc], buf[2][rm][cc], cnext(Y2F[I1], buf[2]),
                           cprev(Y2F[I1], buf[2]), Y2P[I1], kf32.c_h[2], kf32.c_m, kf32.lam, kf32.inv[2]);
      Y3F[I0] = __builtin_fmaf(kf32.dt, kf, XF[I3]);
      Y3P[I0] = __builtin_fmaf(kf32.dt, kp, XP[I3]);
      AF[I2] = __builtin_fma(k.dt3, (double)kf, AF[I2]);
      AP[I2] = __builtin_fma(k.dt3, (double)kp, AP[I2]);

      kf = Y3P[I1];
      kp = tb_rhs_f32<LAM>(Y3F[I1], Y3F[I0], Y3F[I2], buf[3][rp][cc], buf[3][rm][cc], cnext(Y3F[I1], buf[3]),
                           cprev(Y3F[I1], buf[3]), Y3P[I1], kf32.c_h[3], kf32.c_m, kf32.lam, kf32.inv[3]);
      const double s0 = ((double)XF[I4] + lo_f) + __builtin_fma(k.dt6, (double)kf, AF[I3]);
      const double s1 = ((double)XP[I4] + lo_p) + __builtin_fma(k.dt6, (double)kp, AP[I3]);
      const float h0 = (float)s0, h1 = (float)s1;
      q.f = h0;
      q.p = h1;
      q.lf = (float)(s0 - (double)h0);
      q.lp = (float)(s1 - (double)h1);
    } else {
      // ---- stage 1 at P-1: k1 = f(x)
      double kf = (double)XP[I1];
      double kp = tb_force<VAR, LAM>((double)XF[I1], (double)XF[I0], (double)XF[I2], (double)buf[0][rp][cc],
                                    (double)buf[0][rm][cc], cnext(XF[I1], buf[0]), cprev(XF[I1], buf[0]),
                                    (double)XP[I1], k.c_h[0], k.inv[0], k, ps1, dp1);
      Y1F[I0] = (Hv)rnd<F32C>((double)XF[I1] + k.dh * kf);
      Y1P[I0] = (Hv)rnd<F32C>((double)XP[I1] + k.dh * kp);
      if constexpr (F32C) {
        AF[I0] = (Hv)(k.dt6 * kf);
        AP[I0] = (Hv)(k.dt6 * kp);
      } else {
        AF[I0] = XF[I1] + k.dt6 * kf;
        AP[I0] = XP[I1] + k.dt6 * kp;
      }

      // ---- stage 2 at P-2: k2 = f(y1)
      kf = (double)Y1P[I1];
      kp = tb_force<VAR, LAM>((double)Y1F[I1], (double)Y1F[I0], (double)Y1F[I2], (double)buf[1][rp][cc],
                              (double)buf[1][rm][cc], cnext(Y1F[I1], buf[1]), cprev(Y1F[I1], buf[1]),
                              (double)Y1P[I1], k.c_h[1], k.inv[1], k, ps2, dp2);
      Y2F[I0] = (Hv)rnd<F32C>((double)XF[I2] + k.dh * kf);
      Y2P[I0] = (Hv)rnd<F32C>((double)XP[I2] + k.dh * kp);
      AF[I1] = (Hv)((double)AF[I1] + k.dt3 * kf);
      AP[I1] = (Hv)((double)AP[I1] + k.dt3 * kp);

      // ---- stage 3 at P-3: k3 = f(y2)
      kf = (double)Y2P[I1];
      kp = tb_force<VAR, LAM>((double)Y2F[I1], (double)Y2F[I0], (double)Y2F[I2], (double)buf[2][rp][cc],
                              (double)buf[2][rm][cc], cnext(Y2F[I1], buf[2]), cprev(Y2F[I1], buf[2]),
                              (double)Y2P[I1], k.c_h[2], k.inv[2], k, ps3, dp3);
      Y3F[I0] = (Hv)rnd<F32C>((double)XF[I3] + k.dt * kf);
      Y3P[I0] = (Hv)rnd<F32C>((double)XP[I3] + k.dt * kp);
      AF[I2] = (Hv)((double)AF[I2] + k.dt3 * kf);
      AP[I2] = (Hv)((double)AP[I2] + k.dt3 * kp);

      // ---- stage 4 at P-4: k4 = f(y3); new state
      kf = (double)Y3P[I1];
      kp = tb_force<VAR, LAM>((double)Y3F[I1], (double)Y3F[I0], (double)Y3F[I2], (double)buf[3][rp][cc],
                              (double)buf[3][rm][cc], cnext(Y3F[I1], buf[3]), cprev(Y3F[I1], buf[3]),
                              (double)Y3P[I1], k.c_h[3], k.inv[3], k, ps4, dp4);
      if constexpr (F32C) {
        const double s0 = ((double)XF[I4] + lo_f) + ((double)AF[I3] + k.dt6 * kf);
        const double s1 = ((double)XP[I4] + lo_p) + ((double)AP[I3] + k.dt6 * kp);
        const float h0 = (float)s0, h1 = (float)s1;
        q.f = h0;
        q.p = h1;
        q.lf = (float)(s0 - (double)h0);
        q.lp = (float)(s1 - (double)h1);
      } else {
        q.f = AF[I3] + k.dt6 * kf;
        q.p = AP[I3] + k.dt6 * kp;
      }
    }
    q.off = st_off;
    q.plane = off_st;
    if constexpr (!DEFER) flush(q);
    off_st = next_plane(off_st);
    a_st = (a_st + 1 == NA) ? 0 : a_st + 1;
  };
  using std::integral_constant;
  if constexpr (RING) {
    for (int j = 0; j < iters; j += TBU) {
      iteration(integral_constant<int, 0>{}, integral_constant<int, 0>{}, j);
      iteration(integral_constant<int, 1>{}, integral_constant<int, 1>{}, j + 1);
      iteration(integral_constant<int, 2>{}, integral_constant<int, 0>{}, j + 2);
      iteration(integral_constant<int, 3>{}, integral_constant<int, 1>{}, j + 3);
      iteration(integral_constant<int, 4>{}, integral_constant<int, 0>{}, j + 4);
      iteration(integral_constant<int, 5>{}, integral_constant<int, 1>{}, j + 5);
      iteration(integral_constant<int, 6>{}, integral_constant<int, 0>{}, j + 6);
      iteration(integral_constant<int, 7>{}, integral_constant<int, 1>{}, j + 7);
    }
  } else {
    // new[i] = old[i + 1]: the entry for plane P+1 becomes the entry for P
    auto rotate = [](auto &a) {
      const auto t = a[0];
#pragma unroll
      for (int i = 0; i + 1 < TBU; ++i) a[i] = a[i + 1];
      a[TBU - 1] = t;
    };
    auto advance = [&]() {
      rotate(XF);
      rotate(XP);
      rotate(Y1F);
      rotate(Y1P);
      rotate(Y2F);
      rotate(Y2P);
      rotate(Y3F);
      rotate(Y3P);
      rotate(AF);
      rotate(AP);
      if constexpr (VAR == TB_FIXED_CURVATURE) rotate(PS);
    };
    int j = 0;
    for (; j + 1 < iters; j += 2) {
      iteration(integral_constant<int, 0>{}, integral_constant<int, 0>{}, j);
      advance();
      iteration(integral_constant<int, 0>{}, integral_constant<int, 1>{}, j + 1);
      advance();
    }
    if (j < iters) iteration(integral_constant<int, 0>{}, integral_constant<int, 0>{}, j);
  }
  if constexpr (DEFER) {
    flush(Q[0]);
    flush(Q[1]);
  }
}

// Output tile edge for an N^3 lattice (whole plane when N < 24).
inline int tb_tile(int64_t N) { return (N < TBR - 2 * TBH) ? (int)N : TBR - 2 * TBH; }

// Host: the step's scalars (runge_kutta4_classic.hpp:97-123 stage times; the
// RHS scalars as equations.cpp:75-80). inv_ah_sqr follows cf_inv_ah_sqr.
inline TbStep tb_step_host(const cf_kg_params *p, double t, double dt) {
  TbStep k;
  const double dh = dt / 2.0;
  const double th = t + dh;
  const double ts[4] = {t, th, th, t + dt};
  for (int s = 0; s < 4; ++s) {
    const double a_t = cf_cosmo_a(&p->cosmo, ts[s]);
    k.c_h[s] = (-3.0) * cf_cosmo_H(&p->cosmo, ts[s]);
    k.inv[s] = cf_inv_ah_sqr(p->N, p->L, a_t);
  }
  k.c_m = p->m * p->m;
  k.lam = p->lambda;
  k.m = p->m;
  k.fa2 = 0.0;
  k.dh = dh;
  k.dt = dt;
  k.dt6 = dt / 6.0;
  k.dt3 = dt / 3.0;
  return k;
}

}  // namespace cf
