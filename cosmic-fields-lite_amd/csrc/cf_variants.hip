// cf_variants.hip -- the curvature and potential variants of the lattice
// equation on gfx950: fixed curvature, sqrt potential, comoving (variable)
// curvature and its coarse-grid approximation, each as an unfused right-hand
// side (the odeint system call), a fused FP64 RK4 step and an energy density.
//
// Right-hand sides (dphi = pi in every case), in the operation order of the
// reference expression each one restates:
//   FIXED     ((-3H)*pi - ((e^{2Psi}*m)*m)*phi) + (e^{4Psi}*inv)*S7
//             Fixed_Curvature_FRW_equation_kernel, equations_cuda.cu:114-137
//   SQRT      ((-3H)*pi - ((m*m)*phi)/sqrt(1 + (phi*phi)/(f_a*f_a))) + inv*S7
//             Sqrt_Potential_FRW_equation_kernel, equations_cuda.cu:83-106
//   COMOVING  (((-3H + 4*dPsi)*pi - ((e^{2Psi}*m)*m)*phi) + (e^{4Psi}*inv)*S5)
//             + (e^{4Psi}*inv)*(phi_c+ + phi_c-)
//             ComovingCurvatureEquationInFRW::operator(), equations.cpp:186-260
//   APPROX    ((-3H + 4*dPsi_M)*pi - ((((1 + 2P) + (2P)P)*m)*m)*phi)
//             + ((((1 + 4P) + (8P)P))*inv)*S7,  P = Psi_M at the coarse cell
//             Approximate_Variable_Curvature_FRW_equation_kernel, equations_cuda.cu:299-329
// with S7 = (((((-6phi + a+) + a-) + b+) + b-) + c+) + c- and S5 its first five
// terms. The comoving variants rebuild Psi and dPsi/dt from the stored comoving
// curvature spectrum at every distinct stage time:
//   Psi^(k) = R^(k) * (sin x - x cos x) * 2/x^3/G^3,
//   dPsi^(k) = R^(k) * H (3x cos x + (x^2-3) sin x) * 2/x^3/G^3,   x = |k| eta/sqrt(3)
// (equations.cpp:141-183), followed by one batched complex-to-real transform of
// both fields. A classic RK4 step needs them at t, t+dt/2 and t+dt only (stages
// 2 and 3 share a time), and the next step's first stage is usually this step's
// last, so a three-slot cache keyed by the exact stage time leaves two transforms
// per step; the reference computes two transforms per right-hand side. The
// temporally blocked step reads the three times' fields in one kernel.
//
// All arithmetic is FP64 and, as in the rest of the library, no multiply is
// fused into an add (-ffp-contract=off).
#include <hipfft/hipfft.h>

#include <cmath>

#include "cf_march.hpp"
#include "cf_tb_run.hpp"

namespace cf {

enum : int { V_FIXED = CF_VAR_FIXED_CURVATURE, V_SQRT = CF_VAR_SQRT_POTENTIAL,
             V_COMOVING = CF_VAR_COMOVING_CURVATURE, V_APPROX = CF_VAR_APPROX_COMOVING_CURVATURE };

struct VarCoef {
  double c_h;   // (-3.0) * H(t)
  double m;     // mass
  double c_m;   // m * m
  double inv;   // inv_ah_sqr(t)
  double fa2;   // f_a * f_a
};

struct VarFields {
  const double *psi;   // FIXED, COMOVING: N^3; APPROX: M^3
  const double *dpsi;  // COMOVING: N^3; APPROX: M^3
  int64_t M;           // APPROX: coarse grid size
  int64_t sd, hs;      // APPROX: N/M and (N/M)/2
};

template <int VAR>
__device__ __forceinline__ double var_rhs_pi(double phi, double pa1, double pam, double pb1, double pbm,
                                             double pc1, double pcm, double pi, const VarCoef &k,
                                             const VarFields &f, int64_t i, int64_t a, int64_t b, int64_t c,
                                             int64_t N) {
  if constexpr (VAR == V_FIXED) {
    const double ps = f.psi[i];
    const double e2 = exp(2.0 * ps), e4 = exp(4.0 * ps);
    const double s7 = (((((-6.0 * phi + pa1) + pam) + pb1) + pbm) + pc1) + pcm;
    return (k.c_h * pi - ((e2 * k.m) * k.m) * phi) + (e4 * k.inv) * s7;
  } else if constexpr (VAR == V_SQRT) {
    const double s7 = (((((-6.0 * phi + pa1) + pam) + pb1) + pbm) + pc1) + pcm;
    return (k.c_h * pi - (k.c_m * phi) / sqrt(1.0 + (phi * phi) / k.fa2)) + k.inv * s7;
  } else if constexpr (VAR == V_COMOVING) {
    const double ps = f.psi[i], dp = f.dpsi[i];
    const double e2 = exp(2.0 * ps), e4 = exp(4.0 * ps);
    const double ci = e4 * k.inv;
    const double v = ((k.c_h + 4.0 * dp) * pi - ((e2 * k.m) * k.m) * phi) +
                     ci * ((((-6.0 * phi + pa1) + pam) + pb1) + pbm);
    return v + ci * (pc1 + pcm);
  } else {
    const int64_t M = f.M;
    const int64_t ja = ((a + f.hs) % N) / f.sd, jb = ((b + f.hs) % N) / f.sd, jc = ((c + f.hs) % N) / f.sd;
    const int64_t j = (ja * M + jb) * M + jc;
    const double P = f.psi[j], dP = f.dpsi[j];
    const double s7 = (((((-6.0 * phi + pa1) + pam) + pb1) + pbm) + pc1) + pcm;
    return ((k.c_h + 4.0 * dP) * pi - ((((1.0 + 2.0 * P) + (2.0 * P) * P) * k.m) * k.m) * phi) +
           (((1.0 + 4.0 * P) + (8.0 * P) * P) * k.inv) * s7;
  }
}

// MODE 0: unfused right-hand side into io.yout; MODE 1: fused FP64 RK4 stage S
// (the same partial-sum scheme as the Klein-Gordon stage kernels).
template <int VAR, int MODE, int S>
__global__ __launch_bounds__(256) void k_var_march(Tiling tl, VarCoef vk, VarFields vf, StageIO<double> io) {
  int ct, bt, ac;
  tile_of(tl, ct, bt, ac);
  const int64_t N = tl.N, NN = N * N, N3 = NN * N;
  const int64_t c = (int64_t)ct * TC + threadIdx.x;
  const int64_t b = (int64_t)bt * TB + threadIdx.y;
  const int64_t a0 = (int64_t)ac * tl.A;
  if (c >= N || b >= N || a0 >= N) return;
  const int64_t a1 = (a0 + tl.A < N) ? a0 + tl.A : N;
  const int64_t cp = (c + 1 == N) ? 0 : c + 1, cm = (c == 0) ? N - 1 : c - 1;
  const int64_t bp = (b + 1 == N) ? 0 : b + 1, bm = (b == 0) ? N - 1 : b - 1;
  const double *__restrict__ in = io.in;
  const int64_t col = b * N + c;
  double prev = in[((a0 == 0) ? N - 1 : a0 - 1) * NN + col];
  double cur = in[a0 * NN + col];
  for (int64_t a = a0; a < a1; ++a) {
    const int64_t ap = (a + 1 == N) ? 0 : a + 1;
    const int64_t row = a * NN;
    const double next = in[ap * NN + col];
    const double pb1 = in[row + bp * N + c];
    const double pbm = in[row + bm * N + c];
    const double pc1 = in[row + b * N + cp];
    const double pcm = in[row + b * N + cm];
    const int64_t i = row + col;
    const double piv = in[N3 + i];
    const double kphi = piv;
    const double kpi = var_rhs_pi<VAR>(cur, next, prev, pb1, pbm, pc1, pcm, piv, vk, vf, i, a, b, c, N);
    if constexpr (MODE == 0) {
      io.yout[i] = kphi;
      io.yout[N3 + i] = kpi;
    } else if constexpr (S == 1) {
      io.yout[i] = cur + io.cy * kphi;
      io.yout[N3 + i] = piv + io.cy * kpi;
      io.acc[i] = cur + io.ca * kphi;
      io.acc[N3 + i] = piv + io.ca * kpi;
    } else if constexpr (S < 4) {
      const double x0 = io.x[i], x1 = io.x[N3 + i];
      io.yout[i] = x0 + io.cy * kphi;
      io.yout[N3 + i] = x1 + io.cy * kpi;
      io.acc[i] = io.acc[i] + io.ca * kphi;
      io.acc[N3 + i] = io.acc[N3 + i] + io.ca * kpi;
    } else {
      io.x[i] = io.acc[i] + io.ca * kphi;
      io.x[N3 + i] = io.acc[N3 + i] + io.ca * kpi;
    }
    prev = cur;
    cur = next;
  }
}

// Energy densities.
//   SQRT  ((0.5*pi)*pi + ((0.5/4)*inv)*((dx^2 + dy^2) + dz^2))
//         + (((m*m)*f_a)*f_a)*(sqrt(1 + (phi*phi)/(f_a*f_a)) - 1)
//         Sqrt_Potential_energy_density_kernel, equations_cuda.cu:422-444
//   FIXED (((e^{-2Psi}*0.5)*pi)*pi + ((e^{2Psi}*(0.5/4))*inv)*((dx^2 + dy^2) + dz^2))
//         + (((0.5*m)*m)*phi)*phi
//         Curvature_energy_density_kernel, equations_cuda.cu:452-473
//   COMOVING / APPROX (CPU order, equations.cpp:262-322, Psi on the full grid)
//         0.5*((e^{-2Psi}*pi^2 + (m*m)*phi^2) + ((e^{2Psi}*0.25)*inv)*(da^2 + db^2))
//         + (((e^{2Psi}*0.5)*0.25)*inv)*dc^2
struct VarRho {
  double m, c_m, inv, fa, fa2;
};

template <int VAR>
__global__ __launch_bounds__(256) void k_var_rho(Tiling tl, VarRho rc, const double *__restrict__ psi,
                                                 const double *__restrict__ x, double *__restrict__ rho) {
  int ct, bt, ac;
  tile_of(tl, ct, bt, ac);
  const int64_t N = tl.N, NN = N * N, N3 = NN * N;
  const int64_t c = (int64_t)ct * TC + threadIdx.x;
  const int64_t b = (int64_t)bt * TB + threadIdx.y;
  const int64_t a0 = (int64_t)ac * tl.A;
  if (c >= N || b >= N || a0 >= N) return;
  const int64_t a1 = (a0 + tl.A < N) ? a0 + tl.A : N;
  const int64_t cp = (c + 1 == N) ? 0 : c + 1, cm = (c == 0) ? N - 1 : c - 1;
  const int64_t bp = (b + 1 == N) ? 0 : b + 1, bm = (b == 0) ? N - 1 : b - 1;
  const int64_t col = b * N + c;
  double prev = x[((a0 == 0) ? N - 1 : a0 - 1) * NN + col];
  double cur = x[a0 * NN + col];
  for (int64_t a = a0; a < a1; ++a) {
    const int64_t ap = (a + 1 == N) ? 0 : a + 1;
    const int64_t row = a * NN;
    const int64_t i = row + col;
    const double next = x[ap * NN + col];
    const double da = next - prev;
    const double db = x[row + bp * N + c] - x[row + bm * N + c];
    const double dc = x[row + b * N + cp] - x[row + b * N + cm];
    const double piv = x[N3 + i];
    double r;
    if constexpr (VAR == V_SQRT) {
      r = ((0.5 * piv) * piv + ((0.5 / 4.0) * rc.inv) * ((da * da + db * db) + dc * dc)) +
          ((rc.c_m * rc.fa) * rc.fa) * (sqrt(1.0 + (cur * cur) / rc.fa2) - 1.0);
    } else if constexpr (VAR == V_FIXED) {
      const double ps = psi[i];
      r = (((exp(-2.0 * ps) * 0.5) * piv) * piv +
           ((exp(2.0 * ps) * (0.5 / 4.0)) * rc.inv) * ((da * da + db * db) + dc * dc)) +
          (((0.5 * rc.m) * rc.m) * cur) * cur;
    } else {
      const double ps = psi[i];
      const double e2 = exp(2.0 * ps);
      r = 0.5 * ((exp(-2.0 * ps) * (piv * piv) + rc.c_m * (cur * cur)) + ((e2 * 0.25) * rc.inv) * (da * da + db * db));
      r = r + (((e2 * 0.5) * 0.25) * rc.inv) * (dc * dc);
    }
    rho[i] = r;
    prev = cur;
    cur = next;
  }
}

// Transfer from the comoving curvature spectrum to Psi^ and dPsi^/dt over the
// half spectrum of a G^3 grid (equations.cpp:141-183; the c index runs over
// 0..G/2 as in Comoving_Curvature_Psi_dPsidt_fft_kernel, equations_cuda.cu:210-253).
__global__ void k_comoving_transfer(int64_t G, const double *__restrict__ R, double *__restrict__ psi_k,
                                    double *__restrict__ dpsi_k, double factor, double H) {
  const int64_t Hc = G / 2 + 1, n = G * G * Hc;
  const double g3 = (double)(G * G * G);
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = e % Hc, b = (e / Hc) % G, a = e / (Hc * G);
    const int64_t as = (a <= G / 2) ? a : G - a, bs = (b <= G / 2) ? b : G - b;
    const int64_t s = as * as + bs * bs + c * c;
    const double rr = R[2 * e], ri = R[2 * e + 1];
    if (s == 0) {
      psi_k[2 * e] = 0.0;
      psi_k[2 * e + 1] = 0.0;
      dpsi_k[2 * e] = 0.0;
      dpsi_k[2 * e + 1] = 0.0;
    } else {
      const double w = sqrt((double)s) * factor;
      double sv, cv;
      sincos(w, &sv, &cv);
      const double cf = 2.0 / ((w * w) * w) / g3;
      const double tp = (sv - w * cv) * cf;
      const double td = (H * ((3.0 * w) * cv + ((w * w) - 3.0) * sv)) * cf;
      psi_k[2 * e] = rr * tp;
      psi_k[2 * e + 1] = ri * tp;
      dpsi_k[2 * e] = rr * td;
      dpsi_k[2 * e + 1] = ri * td;
    }
  }
}

// Half spectrum of N^3 -> half spectrum of M^3 (compute_cutoff_fouriers,
// fdm3d.cpp:105-130): low |k| modes kept, divided by (N/M)^3.
__global__ void k_cutoff(int64_t N, int64_t M, const double *__restrict__ in, double *__restrict__ out,
                         double scale3) {
  const int64_t HM = M / 2 + 1, HN = N / 2 + 1, n = M * M * HM;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = e % HM, b = (e / HM) % M, a = e / (HM * M);
    const int64_t aN = (a < M / 2) ? a : N - (M - a), bN = (b < M / 2) ? b : N - (M - b);
    const int64_t f = (aN * N + bN) * HN + c;
    out[2 * e] = in[2 * f] / scale3;
    out[2 * e + 1] = in[2 * f + 1] / scale3;
  }
}

static unsigned grid_for(int64_t n) {
  int64_t g = cdiv(n, 256);
  return (unsigned)(g > 65536 ? 65536 : (g < 1 ? 1 : g));
}

// Psi and dPsi/dt on one grid, rebuilt from the stored spectrum R at a stage
// time, with a two-slot cache keyed by that time.
struct PsiGrid {
  int64_t G = 0;
  double *R = nullptr;     // 2*G*G*(G/2+1): comoving curvature, half spectrum
  double *spec = nullptr;  // 2 half spectra: Psi^, dPsi^
  // Psi | dPsi (2*G^3 each) at up to three times, least recently used replaced: a
  // temporally blocked step needs t, t + dt/2 and t + dt at once, and the next step
  // starts at the previous t + dt (two new times per step either way).
  static constexpr int NS = 3;
  double *slot[NS] = {nullptr, nullptr, nullptr};
  double slot_t[NS] = {0.0, 0.0, 0.0};
  bool valid[NS] = {false, false, false};
  uint64_t used[NS] = {0, 0, 0};
  uint64_t clock = 0;
  hipfftHandle plan = 0;
  bool have_plan = false;

  size_t half() const { return 2 * (size_t)G * G * (G / 2 + 1); }
  size_t cube() const { return (size_t)G * G * G; }

  cf_status init(int64_t g) {
    G = g;
    CF_REQUIRE(G <= 1024, "comoving curvature variants support grids up to 1024^3 (got %lld)", (long long)G);
    CF_HIP(hipMalloc(&R, half() * sizeof(double)));
    CF_HIP(hipMalloc(&spec, 2 * half() * sizeof(double)));
    for (int s = 0; s < NS; ++s) CF_HIP(hipMalloc(&slot[s], 2 * cube() * sizeof(double)));
    int n[3] = {(int)G, (int)G, (int)G};
    hipfftResult r = hipfftPlanMany(&plan, 3, n, nullptr, 1, (int)(G * G * (G / 2 + 1)), nullptr, 1, (int)cube(),
                                    HIPFFT_Z2D, 2);
    if (r != HIPFFT_SUCCESS) {
      set_error("hipfftPlanMany(Z2D, G=%lld, batch 2) failed: %d", (long long)G, (int)r);
      return CF_E_FFT;
    }
    have_plan = true;
    return CF_OK;
  }
  void release() {
    if (have_plan) hipfftDestroy(plan);
    have_plan = false;
    (void)hipFree(R);
    (void)hipFree(spec);
    for (int s = 0; s < NS; ++s) {
      (void)hipFree(slot[s]);
      slot[s] = nullptr;
    }
    R = spec = nullptr;
  }

  // Psi(t) | dPsi(t) as 2*G^3 reals, computed on stream st unless cached.
  cf_status get(double t, double factor, double H, hipStream_t st, const double **out) {
    for (int q = 0; q < NS; ++q)
      if (valid[q] && slot_t[q] == t) {
        used[q] = ++clock;
        *out = slot[q];
        return CF_OK;
      }
    int s = 0;
    for (int q = 1; q < NS; ++q)
      if (used[q] < used[s]) s = q;
    valid[s] = false;
    hipLaunchKernelGGL(k_comoving_transfer, dim3(grid_for((int64_t)(half() / 2))), dim3(256), 0, st, G, R, spec,
                       spec + half(), factor, H);
    CF_LAUNCHED();
    if (hipfftSetStream(plan, st) != HIPFFT_SUCCESS) {
      set_error("hipfftSetStream failed");
      return CF_E_FFT;
    }
    hipfftResult r = hipfftExecZ2D(plan, (hipfftDoubleComplex *)spec, slot[s]);
    if (r != HIPFFT_SUCCESS) {
      set_error("hipfftExecZ2D (G=%lld, batch 2) failed: %d", (long long)G, (int)r);
      return CF_E_FFT;
    }
    slot_t[s] = t;
    valid[s] = true;
    used[s] = ++clock;
    *out = slot[s];
    return CF_OK;
  }
};

}  // namespace cf

using namespace cf;

struct cf_variant_s {
  int kind = 0;
  cf_kg_params p{};
  double f_a = 1.0;
  int64_t M = 0;
  double *psi_fixed = nullptr;  // FIXED: N^3
  PsiGrid grid;                 // COMOVING: N; APPROX: M (cut-off spectrum)
  PsiGrid full;                 // APPROX only: N, for the energy density
  bool full_ready = false;
};

namespace cf {

static cf_status check_kg(const cf_kg_params *p) {
  CF_REQUIRE(p, "null cf_kg_params");
  CF_REQUIRE(valid_N(p->N), "unsupported lattice size N=%lld (even, 4..2048)", (long long)p->N);
  CF_REQUIRE(p->L > 0.0, "box size L must be positive");
  return CF_OK;
}

// factor = (2 pi / L) eta(t) / sqrt(3) (equations.cpp:196).
static double comoving_factor(const cf_kg_params &p, double t) {
  const double pi = 3.141592653589793;
  return (2 * pi / p.L) * cf_cosmo_eta(&p.cosmo, t) / sqrt(3.0);
}

static VarCoef var_coef(const cf_variant_s *v, double t) {
  VarCoef k;
  const double a_t = cf_cosmo_a(&v->p.cosmo, t);
  const double H_t = cf_cosmo_H(&v->p.cosmo, t);
  k.c_h = (-3.0) * H_t;
  k.m = v->p.m;
  k.c_m = v->p.m * v->p.m;
  k.inv = cf_inv_ah_sqr(v->p.N, v->p.L, a_t);
  k.fa2 = v->f_a * v->f_a;
  return k;
}

// Fields the right-hand side at time t reads (computed on st when needed).
static cf_status var_fields(cf_variant_s *v, double t, hipStream_t st, VarFields *f) {
  *f = VarFields{};
  if (v->kind == V_FIXED) {
    f->psi = v->psi_fixed;
  } else if (v->kind == V_COMOVING || v->kind == V_APPROX) {
    const double *pf = nullptr;
    CF_TRY(v->grid.get(t, comoving_factor(v->p, t), cf_cosmo_H(&v->p.cosmo, t), st, &pf));
    f->psi = pf;
    f->dpsi = pf + v->grid.cube();
    if (v->kind == V_APPROX) {
      f->M = v->M;
      f->sd = v->p.N / v->M;
      f->hs = f->sd / 2;
    }
  }
  return CF_OK;
}

template <int MODE, int S>
static cf_status launch_var(const cf_variant_s *v, const Tiling &tl, const VarCoef &k, const VarFields &f,
                            const StageIO<double> &io, hipStream_t st) {
  dim3 blk(TC, TB), grd(tl.total);
  switch (v->kind) {
    case V_FIXED: hipLaunchKernelGGL((k_var_march<V_FIXED, MODE, S>), grd, blk, 0, st, tl, k, f, io); break;
    case V_SQRT: hipLaunchKernelGGL((k_var_march<V_SQRT, MODE, S>), grd, blk, 0, st, tl, k, f, io); break;
    case V_COMOVING:
      hipLaunchKernelGGL((k_var_march<V_COMOVING, MODE, S>), grd, blk, 0, st, tl, k, f, io);
      break;
    default: hipLaunchKernelGGL((k_var_march<V_APPROX, MODE, S>), grd, blk, 0, st, tl, k, f, io); break;
  }
  CF_LAUNCHED();
  return CF_OK;
}

}  // namespace cf

extern "C" {

cf_status cf_cutoff_fouriers(int64_t N, int64_t M, const double *in, double *out, cf_stream stream) {
  CF_REQUIRE(valid_N(N) && valid_N(M) && M <= N, "cf_cutoff_fouriers: need even 4 <= M <= N <= 2048 (N=%lld M=%lld)",
             (long long)N, (long long)M);
  CF_REQUIRE(in && out && (const void *)in != (const void *)out, "cf_cutoff_fouriers: null or aliased buffers");
  const double sd = static_cast<double>(N) / static_cast<double>(M);
  const int64_t n = M * M * (M / 2 + 1);
  hipLaunchKernelGGL(k_cutoff, dim3(grid_for(n)), dim3(256), 0, S(stream), N, M, in, out, pow(sd, 3));
  CF_LAUNCHED();
  return CF_OK;
}

cf_status cf_comoving_psi_fft(int64_t G, const double *R_fft, double *psi_fft, double *dpsi_fft, double factor,
                              double H, cf_stream stream) {
  CF_REQUIRE(valid_N(G), "cf_comoving_psi_fft: unsupported grid size %lld", (long long)G);
  CF_REQUIRE(R_fft && psi_fft && dpsi_fft, "cf_comoving_psi_fft: null buffer");
  const int64_t n = G * G * (G / 2 + 1);
  hipLaunchKernelGGL(k_comoving_transfer, dim3(grid_for(n)), dim3(256), 0, S(stream), G, R_fft, psi_fft, dpsi_fft,
                     factor, H);
  CF_LAUNCHED();
  return CF_OK;
}

cf_status cf_variant_create(const cf_kg_params *p, int kind, double f_a, int64_t M, const double *field,
                            cf_variant *out, cf_stream stream) {
  CF_REQUIRE(out, "cf_variant_create: null output handle");
  *out = nullptr;
  CF_TRY(check_kg(p));
  CF_REQUIRE(kind >= CF_VAR_FIXED_CURVATURE && kind <= CF_VAR_APPROX_COMOVING_CURVATURE,
             "cf_variant_create: unknown variant kind %d", kind);
  const int64_t N = p->N;
  if (kind == V_SQRT) CF_REQUIRE(f_a > 0.0, "cf_variant_create: sqrt potential needs f_a > 0");
  if (kind != V_SQRT) CF_REQUIRE(field, "cf_variant_create: this variant needs a field (Psi or the R spectrum)");
  if (kind == V_APPROX)
    CF_REQUIRE(valid_N(M) && M <= N && N % M == 0,
               "cf_variant_create: coarse grid M=%lld must be even, >= 4 and divide N=%lld", (long long)M,
               (long long)N);
  cf_variant_s *v = new cf_variant_s();
  v->kind = kind;
  v->p = *p;
  v->f_a = f_a;
  v->M = (kind == V_APPROX) ? M : 0;
  hipStream_t st = S(stream);
  cf_status rc = CF_OK;
  const size_t N3 = (size_t)N * N * N, halfN = 2 * (size_t)N * N * (N / 2 + 1);
  if (kind == V_FIXED) {
    if (hipMalloc(&v->psi_fixed, N3 * sizeof(double)) != hipSuccess ||
        hipMemcpyAsync(v->psi_fixed, field, N3 * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess) {
      set_error("cf_variant_create: allocation or copy of Psi failed");
      rc = CF_E_HIP;
    }
  } else if (kind == V_COMOVING) {
    rc = v->grid.init(N);
    if (rc == CF_OK && hipMemcpyAsync(v->grid.R, field, halfN * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess) {
      set_error("cf_variant_create: copy of the R spectrum failed");
      rc = CF_E_HIP;
    }
  } else if (kind == V_APPROX) {
    rc = v->grid.init(M);
    if (rc == CF_OK) rc = v->full.init(N);
    if (rc == CF_OK && hipMemcpyAsync(v->full.R, field, halfN * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess) {
      set_error("cf_variant_create: copy of the R spectrum failed");
      rc = CF_E_HIP;
    }
    if (rc == CF_OK) rc = cf_cutoff_fouriers(N, M, v->full.R, v->grid.R, stream);
    v->full_ready = (rc == CF_OK);
  }
  if (rc == CF_OK && hipStreamSynchronize(st) != hipSuccess) {
    set_error("cf_variant_create: stream synchronisation failed");
    rc = CF_E_HIP;
  }
  if (rc != CF_OK) {
    (void)hipFree(v->psi_fixed);
    v->grid.release();
    v->full.release();
    delete v;
    return rc;
  }
  *out = v;
  return CF_OK;
}

void cf_variant_destroy(cf_variant v) {
  if (!v) return;
  (void)hipDeviceSynchronize();
  (void)hipFree(v->psi_fixed);
  v->grid.release();
  v->full.release();
  delete v;
}

cf_status cf_variant_rhs(cf_variant v, const double *x, double *dxdt, double t, cf_stream stream) {
  CF_REQUIRE(v, "cf_variant_rhs: null handle");
  CF_REQUIRE(x && dxdt && (const void *)x != (const void *)dxdt, "cf_variant_rhs: null or aliased buffers");
  hipStream_t st = S(stream);
  VarFields f;
  CF_TRY(var_fields(v, t, st, &f));
  StageIO<double> io{};
  io.in = x;
  io.yout = dxdt;
  return launch_var<0, 0>(v, make_tiling(v->p.N, default_chunk(v->p.N)), var_coef(v, t), f, io, st);
}

cf_status cf_variant_rk4_step(cf_variant v, double *x, double t, double dt, void *work, cf_stream stream) {
  CF_REQUIRE(v, "cf_variant_rk4_step: null handle");
  CF_REQUIRE(x && work, "cf_variant_rk4_step: null buffer");
  const int64_t n = 2 * v->p.N * v->p.N * v->p.N;
  double *yA = (double *)work, *yB = yA + n, *acc = yB + n;
  const double dh = dt / 2.0, dt6 = dt / 6.0, dt3 = dt / 3.0;
  const double th = t + dh;
  const double ts[4] = {t, th, th, t + dt};
  Tiling tl = make_tiling(v->p.N, default_chunk(v->p.N));
  hipStream_t st = S(stream);
  StageIO<double> io{};
  io.x = x;
  io.acc = acc;
  VarFields f;
  CF_TRY(var_fields(v, ts[0], st, &f));
  io.in = x; io.yout = yA; io.cy = dh; io.ca = dt6;
  CF_TRY((launch_var<1, 1>(v, tl, var_coef(v, ts[0]), f, io, st)));
  CF_TRY(var_fields(v, ts[1], st, &f));
  io.in = yA; io.yout = yB; io.cy = dh; io.ca = dt3;
  CF_TRY((launch_var<1, 2>(v, tl, var_coef(v, ts[1]), f, io, st)));
  CF_TRY(var_fields(v, ts[2], st, &f));
  io.in = yB; io.yout = yA; io.cy = dt; io.ca = dt3;
  CF_TRY((launch_var<1, 3>(v, tl, var_coef(v, ts[2]), f, io, st)));
  CF_TRY(var_fields(v, ts[3], st, &f));
  io.in = yA; io.yout = nullptr; io.ca = dt6;
  CF_TRY((launch_var<1, 4>(v, tl, var_coef(v, ts[3]), f, io, st)));
  return CF_OK;
}

cf_status cf_variant_rk4_run(cf_variant v, double *x, double t0, int64_t k0, int64_t n_steps, double dt, void *work,
                             cf_stream stream) {
  CF_REQUIRE(n_steps >= 0, "cf_variant_rk4_run: negative step count");
  for (int64_t k = 0; k < n_steps; ++k) {
    const double t = t0 + static_cast<double>(k0 + k) * dt;
    CF_TRY(cf_variant_rk4_step(v, x, t, dt, work, stream));
  }
  return CF_OK;
}

// One temporally blocked step of any variant from x_in into x_out (the comoving
// ones with Psi | dPsi at t, t + dt/2, t + dt from the cache, the same fields and
// times as the stage kernels use). The kinds equal the TB_* ids (CF_VAR_* 1-4).
static cf_status var_tb_step(cf_variant_s *v, const double *x_in, double *x_out, double t, double dt, int chunk,
                             hipStream_t st) {
  if (v->kind == V_FIXED || v->kind == V_SQRT)
    return tb_step_variant(v->kind, &v->p, v->f_a, v->psi_fixed, x_in, x_out, t, dt, chunk, st, nullptr);
  const double dh = dt / 2.0, th = t + dh;
  const double ts[3] = {t, th, t + dt};
  TbPsi cp{};
  for (int q = 0; q < 3; ++q) {
    const double *pf = nullptr;
    CF_TRY(v->grid.get(ts[q], comoving_factor(v->p, ts[q]), cf_cosmo_H(&v->p.cosmo, ts[q]), st, &pf));
    cp.f[q] = pf;
  }
  cp.G = v->grid.G;
  cp.sd = v->kind == V_APPROX ? v->p.N / v->M : 1;
  cp.hs = cp.sd / 2;
  return tb_step_variant(v->kind, &v->p, v->f_a, nullptr, x_in, x_out, t, dt, chunk, st, &cp);
}

cf_status cf_variant_rk4_run_tb(cf_variant v, double *x, double t0, int64_t k0, int64_t n_steps, double dt, int chunk,
                                void *work, cf_stream stream) {
  CF_REQUIRE(v, "cf_variant_rk4_run_tb: null handle");
  CF_REQUIRE(n_steps >= 0, "cf_variant_rk4_run_tb: negative step count");
  const size_t bytes = 2 * (size_t)v->p.N * v->p.N * v->p.N * sizeof(double);
  const char *xa = (const char *)x, *wa = (const char *)work;
  CF_REQUIRE(x && work && (xa + bytes <= wa || wa + bytes <= xa), "cf_variant_rk4_run_tb: null or overlapping buffer");
  hipStream_t st = S(stream);
  double *bufs[2] = {x, (double *)work};
  for (int64_t s = 0; s < n_steps; ++s) {
    const double t = t0 + static_cast<double>(k0 + s) * dt;
    CF_TRY(var_tb_step(v, bufs[s & 1], bufs[(s + 1) & 1], t, dt, chunk, st));
  }
  if (n_steps & 1) CF_HIP(hipMemcpyAsync(x, work, bytes, hipMemcpyDeviceToDevice, st));
  return CF_OK;
}

cf_status cf_variant_rk4_step_tb(cf_variant v, const double *x_in, double *x_out, double t, double dt, int chunk,
                                 cf_stream stream) {
  CF_REQUIRE(v, "cf_variant_rk4_step_tb: null handle");
  return var_tb_step(v, x_in, x_out, t, dt, chunk, S(stream));
}

cf_status cf_variant_psi(cf_variant v, double t, double *psi, double *dpsi, cf_stream stream) {
  CF_REQUIRE(v, "cf_variant_psi: null handle");
  CF_REQUIRE(v->kind == V_COMOVING || v->kind == V_APPROX, "cf_variant_psi: only the comoving variants rebuild Psi");
  CF_REQUIRE(psi && dpsi, "cf_variant_psi: null buffer");
  hipStream_t st = S(stream);
  VarFields f;
  CF_TRY(var_fields(v, t, st, &f));
  const size_t bytes = v->grid.cube() * sizeof(double);
  CF_HIP(hipMemcpyAsync(psi, f.psi, bytes, hipMemcpyDeviceToDevice, st));
  CF_HIP(hipMemcpyAsync(dpsi, f.dpsi, bytes, hipMemcpyDeviceToDevice, st));
  return CF_OK;
}

cf_status cf_variant_energy_density(cf_variant v, const double *x, double *rho, double t, cf_stream stream) {
  CF_REQUIRE(v, "cf_variant_energy_density: null handle");
  CF_REQUIRE(x && rho && (const void *)x != (const void *)rho, "cf_variant_energy_density: null or aliased buffers");
  hipStream_t st = S(stream);
  const double a_t = cf_cosmo_a(&v->p.cosmo, t);
  VarRho rc;
  rc.m = v->p.m;
  rc.c_m = v->p.m * v->p.m;
  rc.inv = cf_inv_ah_sqr(v->p.N, v->p.L, a_t);
  rc.fa = v->f_a;
  rc.fa2 = v->f_a * v->f_a;
  const double *psi = nullptr;
  if (v->kind == V_FIXED) {
    psi = v->psi_fixed;
  } else if (v->kind != V_SQRT) {
    // Psi on the full N grid (the approximate variant's energy density is the
    // comoving one, equations_cuda.cu:812-815).
    PsiGrid &g = (v->kind == V_COMOVING) ? v->grid : v->full;
    CF_TRY(g.get(t, comoving_factor(v->p, t), cf_cosmo_H(&v->p.cosmo, t), st, &psi));
  }
  Tiling tl = make_tiling(v->p.N, default_chunk(v->p.N));
  dim3 blk(TC, TB), grd(tl.total);
  switch (v->kind) {
    case V_FIXED: hipLaunchKernelGGL(k_var_rho<V_FIXED>, grd, blk, 0, st, tl, rc, psi, x, rho); break;
    case V_SQRT: hipLaunchKernelGGL(k_var_rho<V_SQRT>, grd, blk, 0, st, tl, rc, psi, x, rho); break;
    default: hipLaunchKernelGGL(k_var_rho<V_COMOVING>, grd, blk, 0, st, tl, rc, psi, x, rho); break;
  }
  CF_LAUNCHED();
  return CF_OK;
}

}  // extern "C"
