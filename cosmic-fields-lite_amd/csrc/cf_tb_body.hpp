// cf_tb_body.hpp -- the temporally blocked RK4 step of one work unit (an output
// tile of one lattice over a chunk of a-planes), shared by the single-lattice
// kernel (cf_rk4_tb.hip) and the persistent ensemble kernel (cf_ensemble.hip).
//
// Geometry. A block is 32 x 32 threads over a 32 x 32 (b, c) region of the
// lattice: an output tile of T x T sites (T = 24, or N for small lattices) plus
// a halo of 4 sites on every side, one per stencil stage. The block marches
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
// therefore bit-identical to the four-kernel path.
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

// Equation of the step body: the Klein-Gordon family (+ lambda phi^3), or the
// fixed-curvature / sqrt-potential variants in the operation order of their CUDA
// kernels (Fixed_Curvature_FRW_equation_kernel, equations_cuda.cu:114-137;
// Sqrt_Potential_FRW_equation_kernel, :83-106), as in csrc/cf_variants.hip.
enum : int { TB_KG = 0, TB_FIXED_CURVATURE = 1, TB_SQRT_POTENTIAL = 2 };

// One work unit: lattice size, output tile origin and edge, plane chunk.
struct TbUnit {
  int64_t N;
  int64_t b0, c0, a0, a1;
  int T;
};

__device__ __forceinline__ double tb_rhs(double phi, double pa1, double pam, double pb1, double pbm, double pc1,
                                         double pcm, double pi, double c_h, double c_m, double lam, double inv,
                                         bool has_lam) {
  double v = c_h * pi - c_m * phi;
  if (has_lam) v = v - ((lam * phi) * phi) * phi;
  double r = v + inv * ((((-6.0 * phi + pa1) + pam) + pb1) + pbm);
  return r + inv * (pc1 + pcm);
}

template <int VAR, bool LAM>
__device__ __forceinline__ double tb_force(double phi, double pa1, double pam, double pb1, double pbm, double pc1,
                                           double pcm, double pi, double c_h, double inv, const TbStep &k, double ps) {
  if constexpr (VAR == TB_KG) {
    return tb_rhs(phi, pa1, pam, pb1, pbm, pc1, pcm, pi, c_h, k.c_m, k.lam, inv, LAM);
  } else if constexpr (VAR == TB_FIXED_CURVATURE) {
    const double e2 = exp(2.0 * ps), e4 = exp(4.0 * ps);
    const double s7 = (((((-6.0 * phi + pa1) + pam) + pb1) + pbm) + pc1) + pcm;
    return (c_h * pi - ((e2 * k.m) * k.m) * phi) + (e4 * inv) * s7;
  } else {
    const double s7 = (((((-6.0 * phi + pa1) + pam) + pb1) + pbm) + pc1) + pcm;
    return (c_h * pi - (k.c_m * phi) / sqrt(1.0 + (phi * phi) / k.fa2)) + inv * s7;
  }
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
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, false);
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

template <class T, bool F32C, bool LAM, bool DEFER = true, bool LANE_NBR = true, int VAR = TB_KG>
__device__ __forceinline__ void tb_unit(TbLds<T> &lds, const TbUnit &u, const TbStep &k, const T *__restrict__ xin,
                                        const float *__restrict__ lin, T *__restrict__ xout,
                                        float *__restrict__ lout, const double *__restrict__ psi = nullptr) {
  static_assert(VAR == TB_KG || !F32C, "the equation variants are FP64");
  const int tid = threadIdx.x;
  const int r = tid >> 5, cc = tid & 31;
  static_assert(!F32C || sizeof(T) == sizeof(float), "the compensated mode stores float");
  // In the compensated mode every value carried from plane to plane is exactly
  // representable in float (the state hi, the rounded stage inputs and the rounded
  // increment), so the histories are held as float (half the registers) and
  // widened where they are used; the arithmetic stays double.
  typedef std::conditional_t<F32C, float, double> Hv;
  const int64_t N = u.N, NN = N * N, N3 = NN * N;
  const int plane_bytes = (int)(NN * (int64_t)sizeof(T)), plane_bytes_f = (int)(NN * (int64_t)sizeof(float));
  int64_t b = (u.b0 - TBH + r) % N;
  if (b < 0) b += N;
  int64_t c = (u.c0 - TBH + cc) % N;
  if (c < 0) c += N;
  const int64_t col = b * N + c;
  const bool out_ok = r >= TBH && r < TBH + u.T && u.b0 + (r - TBH) < N && cc >= TBH && cc < TBH + u.T &&
                      u.c0 + (cc - TBH) < N;
  const int rm = r > 0 ? r - 1 : 0, rp = r < TBR - 1 ? r + 1 : TBR - 1;
  const int cm = cc > 0 ? cc - 1 : 0, cp = cc < TBR - 1 ? cc + 1 : TBR - 1;
  // c+1 / c-1 neighbour of a stage input: from the adjacent lanes (LANE_NBR) or
  // from the LDS plane buffer (the same values; which is faster depends on the
  // kernel's mix of LDS and VALU work)
  auto cnext = [&](Hv v, const T(*pl)[TBR]) -> double {
    if constexpr (LANE_NBR) return (double)tb_lane_next(v);
    else return (double)pl[r][cp];
  };
  auto cprev = [&](Hv v, const T(*pl)[TBR]) -> double {
    if constexpr (LANE_NBR) return (double)tb_lane_prev(v);
    else return (double)pl[r][cm];
  };

  // plane offsets (times N^2) advanced by one plane per iteration, wrapping at N
  auto wrap_plane = [&](int64_t P) -> int64_t {
    int64_t q = P % N;
    if (q < 0) q += N;
    return q * NN;
  };
  auto next_plane = [&](int64_t off) -> int64_t {
    off += NN;
    return off == N3 ? 0 : off;
  };

  Hv xf0 = 0, xf1 = 0, xf2 = 0, xf3 = 0, xf4 = 0;  // x phi at P, P-1, ..., P-4
  Hv xp0 = 0, xp1 = 0, xp2 = 0, xp3 = 0, xp4 = 0;  // x pi
  Hv y1f1 = 0, y1f2 = 0, y1p1 = 0;                  // y1 phi at P-2, P-3 (start of iteration); pi at P-2
  Hv y2f1 = 0, y2f2 = 0, y2p1 = 0;                  // y2 phi at P-3, P-4; pi at P-3
  Hv y3f1 = 0, y3f2 = 0, y3p1 = 0;                  // y3 phi at P-4, P-5; pi at P-4
  Hv af1 = 0, af2 = 0, af3 = 0, ap1 = 0, ap2 = 0, ap3 = 0;  // acc (FP64) / increment D (F32C) at P-2, P-3, P-4

  const int iters = (int)(u.a1 - u.a0) + 8;
  // Memory pipeline. Two prefetch slots are used alternately: iteration j takes
  // x at plane P from slot j&1 (loaded two iterations earlier) and refills it
  // with plane P+2. The new state of iteration j is stored at the start of
  // iteration j+2. The loop is unrolled by two, so slots and pending stores are
  // fixed registers and every wait is for an operation issued at least an
  // iteration earlier (DEFER = false stores at the end of the iteration, for
  // callers with less register room). Stores go through buffer resources over one plane: lanes
  // that do not store get an offset past its end, which the hardware drops
  // (loads of it return 0), so no branch surrounds a memory operation.
  int64_t off_ld = wrap_plane(u.a0 - TBH);
  Hv s0f = (Hv)xin[off_ld + col], s0p = (Hv)xin[N3 + off_ld + col];
  // fixed curvature: Psi rides with x (same slots, same planes)
  double s0q = 0.0, s1q = 0.0, ps0 = 0.0, ps1 = 0.0, ps2 = 0.0, ps3 = 0.0, ps4 = 0.0;
  if constexpr (VAR == TB_FIXED_CURVATURE) s0q = psi[off_ld + col];
  off_ld = next_plane(off_ld);
  Hv s1f = (Hv)xin[off_ld + col], s1p = (Hv)xin[N3 + off_ld + col];
  if constexpr (VAR == TB_FIXED_CURVATURE) s1q = psi[off_ld + col];
  int64_t off_st = wrap_plane(u.a0 - TBH - 4);  // plane P-4
  TbPending<T> q0{}, q1{};
  q0.off = q1.off = TB_OOB;
  // F32C: the compensation term of output plane P-4 is prefetched the same way
  // (slot j&1, loaded two iterations ahead; lanes that will not store read 0).
  float l0f = 0.0f, l0p = 0.0f, l1f = 0.0f, l1p = 0.0f;
  int64_t off_lo = off_st;
  auto lo_off = [&](int jj) -> uint32_t {
    return ((jj >= 8) && out_ok) ? (uint32_t)(col * (int64_t)sizeof(float)) : TB_OOB;
  };
  if constexpr (F32C) {
    l0f = tb_load_f32(tb_plane(lin + off_lo, plane_bytes_f), lo_off(0));
    l0p = tb_load_f32(tb_plane(lin + N3 + off_lo, plane_bytes_f), lo_off(0));
    off_lo = next_plane(off_lo);
    l1f = tb_load_f32(tb_plane(lin + off_lo, plane_bytes_f), lo_off(1));
    l1p = tb_load_f32(tb_plane(lin + N3 + off_lo, plane_bytes_f), lo_off(1));
  }

  auto flush = [&](const TbPending<T> &q) {
    tb_store(q.f, tb_plane(xout + q.plane, plane_bytes), q.off);
    tb_store(q.p, tb_plane(xout + N3 + q.plane, plane_bytes), q.off);
    if constexpr (F32C) {
      tb_store(q.lf, tb_plane(lout + q.plane, plane_bytes_f), q.off);
      tb_store(q.lp, tb_plane(lout + N3 + q.plane, plane_bytes_f), q.off);
    }
  };

  auto iteration = [&](const int j, Hv &slf, Hv &slp, double &slq, float &llf, float &llp, T(*buf)[TBR][TBR],
                       TbPending<T> &q) {
    xf0 = slf;
    xp0 = slp;
    if constexpr (VAR == TB_FIXED_CURVATURE) ps0 = slq;
    double lo_f = 0.0, lo_p = 0.0;
    if constexpr (F32C) {
      lo_f = (double)llf;
      lo_p = (double)llp;
    }
    if constexpr (DEFER) flush(q);  // the new state of iteration j-2
    const uint32_t st_off = ((j >= 8) && out_ok) ? (uint32_t)(col * (int64_t)sizeof(T)) : TB_OOB;
    if constexpr (F32C) {
      off_lo = next_plane(off_lo);  // plane P-2, the output plane of iteration j+2
      llf = tb_load_f32(tb_plane(lin + off_lo, plane_bytes_f), lo_off(j + 2));
      llp = tb_load_f32(tb_plane(lin + N3 + off_lo, plane_bytes_f), lo_off(j + 2));
    }
    off_ld = next_plane(off_ld);  // plane P+2
    slf = (Hv)xin[off_ld + col];
    slp = (Hv)xin[N3 + off_ld + col];
    if constexpr (VAR == TB_FIXED_CURVATURE) slq = psi[off_ld + col];

    buf[0][r][cc] = (T)xf1;
    buf[1][r][cc] = (T)y1f1;
    buf[2][r][cc] = (T)y2f1;
    buf[3][r][cc] = (T)y3f1;
    __syncthreads();

    // ---- stage 1 at P-1: k1 = f(x)
    double kf = (double)xp1;
    double kp = tb_force<VAR, LAM>((double)xf1, (double)xf0, (double)xf2, (double)buf[0][rp][cc],
                                  (double)buf[0][rm][cc], cnext(xf1, buf[0]), cprev(xf1, buf[0]),
                                  (double)xp1, k.c_h[0], k.inv[0], k, ps1);
    const Hv y1f0 = (Hv)rnd<F32C>((double)xf1 + k.dh * kf);
    const Hv y1p0 = (Hv)rnd<F32C>((double)xp1 + k.dh * kp);
    Hv af0, ap0;
    if constexpr (F32C) {
      af0 = (Hv)(k.dt6 * kf);
      ap0 = (Hv)(k.dt6 * kp);
    } else {
      af0 = xf1 + k.dt6 * kf;
      ap0 = xp1 + k.dt6 * kp;
    }

    // ---- stage 2 at P-2: k2 = f(y1)
    kf = (double)y1p1;
    kp = tb_force<VAR, LAM>((double)y1f1, (double)y1f0, (double)y1f2, (double)buf[1][rp][cc],
                                  (double)buf[1][rm][cc], cnext(y1f1, buf[1]), cprev(y1f1, buf[1]),
                                  (double)y1p1, k.c_h[1], k.inv[1], k, ps2);
    const Hv y2f0 = (Hv)rnd<F32C>((double)xf2 + k.dh * kf);
    const Hv y2p0 = (Hv)rnd<F32C>((double)xp2 + k.dh * kp);
    af1 = (Hv)((double)af1 + k.dt3 * kf);
    ap1 = (Hv)((double)ap1 + k.dt3 * kp);

    // ---- stage 3 at P-3: k3 = f(y2)
    kf = (double)y2p1;
    kp = tb_force<VAR, LAM>((double)y2f1, (double)y2f0, (double)y2f2, (double)buf[2][rp][cc],
                                  (double)buf[2][rm][cc], cnext(y2f1, buf[2]), cprev(y2f1, buf[2]),
                                  (double)y2p1, k.c_h[2], k.inv[2], k, ps3);
    const Hv y3f0 = (Hv)rnd<F32C>((double)xf3 + k.dt * kf);
    const Hv y3p0 = (Hv)rnd<F32C>((double)xp3 + k.dt * kp);
    af2 = (Hv)((double)af2 + k.dt3 * kf);
    ap2 = (Hv)((double)ap2 + k.dt3 * kp);

    // ---- stage 4 at P-4: k4 = f(y3); new state
    kf = (double)y3p1;
    kp = tb_force<VAR, LAM>((double)y3f1, (double)y3f0, (double)y3f2, (double)buf[3][rp][cc],
                                  (double)buf[3][rm][cc], cnext(y3f1, buf[3]), cprev(y3f1, buf[3]),
                                  (double)y3p1, k.c_h[3], k.inv[3], k, ps4);
    if constexpr (F32C) {
      const double s0 = ((double)xf4 + lo_f) + ((double)af3 + k.dt6 * kf);
      const double s1 = ((double)xp4 + lo_p) + ((double)ap3 + k.dt6 * kp);
      const float h0 = (float)s0, h1 = (float)s1;
      q.f = h0;
      q.p = h1;
      q.lf = (float)(s0 - (double)h0);
      q.lp = (float)(s1 - (double)h1);
    } else {
      q.f = af3 + k.dt6 * kf;
      q.p = ap3 + k.dt6 * kp;
    }
    q.off = st_off;
    q.plane = off_st;
    if constexpr (!DEFER) flush(q);
    off_st = next_plane(off_st);

    // ---- advance the histories by one plane
    xf4 = xf3; xf3 = xf2; xf2 = xf1; xf1 = xf0;
    ps4 = ps3; ps3 = ps2; ps2 = ps1; ps1 = ps0;
    xp4 = xp3; xp3 = xp2; xp2 = xp1; xp1 = xp0;
    y1f2 = y1f1; y1f1 = y1f0; y1p1 = y1p0;
    y2f2 = y2f1; y2f1 = y2f0; y2p1 = y2p0;
    y3f2 = y3f1; y3f1 = y3f0; y3p1 = y3p0;
    af3 = af2; af2 = af1; af1 = af0;
    ap3 = ap2; ap2 = ap1; ap1 = ap0;
  };
  int j = 0;
  for (; j + 1 < iters; j += 2) {
    iteration(j, s0f, s0p, s0q, l0f, l0p, lds[0], q0);
    iteration(j + 1, s1f, s1p, s1q, l1f, l1p, lds[1], q1);
  }
  if (j < iters) iteration(j, s0f, s0p, s0q, l0f, l0p, lds[0], q0);
  if constexpr (DEFER) {
    flush(q0);
    flush(q1);
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
