// cf_stencil.hip -- the periodic 7-point Klein-Gordon right-hand side, the four
// fused RK4 stage kernels (FP64 and FP32-compensated storage) and the energy
// density, all on gfx950.
//
// Arithmetic order. Every kernel evaluates the reference CPU expressions in their
// written order (the library is compiled with -ffp-contract=off, so no product is
// fused into an add):
//   dpi = ((-3H)*pi - (m*m)*phi [- ((lambda*phi)*phi)*phi])
//         + inv*((((-6*phi + phi_a+) + phi_a-) + phi_b+) + phi_b-)      equations.cpp:86-93
//   dpi = dpi + inv*(phi_c+ + phi_c-)                                    equations.cpp:94-100
// and the RK4 combination of runge_kutta4_classic.hpp:103-129 with scale_sum2 /
// scale_sum5 (1*x + dt/6 k1 + dt/3 k2 + dt/3 k3 + dt/6 k4, left to right). The
// fused FP64 path keeps the partial sum acc = 1*x + dt/6 k1 (+ dt/3 k2 + dt/3 k3)
// in its own array, so its result is bit-identical to the unfused sequence.
//
// Data layout in HBM: one state array [phi (N^3) | pi (N^3)], site (a,b,c) at
// N*N*a + N*b + c. Each thread owns one (b,c) column and marches along a over a
// chunk of planes, keeping phi(a-1), phi(a), phi(a+1) in registers; the b and c
// neighbours are read through the L1/L2 caches (they are the centre loads of the
// neighbouring lanes and rows of the same tile).
#include <cmath>

#include "cf_march.hpp"

namespace cf {

// MODE 0: unfused RHS (dxdt -> yout); MODE 1: FP64 fused stage; MODE 2: F32C fused stage.
template <class T, int MODE, int S, bool LAM>
__global__ __launch_bounds__(256) void k_march(Tiling tl, RhsCoef rk, StageIO<T> io) {
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
  const T *__restrict__ in = io.in;

  const int64_t col = b * N + c;
  double prev = (double)in[((a0 == 0) ? N - 1 : a0 - 1) * NN + col];
  double cur = (double)in[a0 * NN + col];
  for (int64_t a = a0; a < a1; ++a) {
    const int64_t ap = (a + 1 == N) ? 0 : a + 1;
    const int64_t row = a * NN;
    const double next = (double)in[ap * NN + col];
    const double pb1 = (double)in[row + bp * N + c];
    const double pbm = (double)in[row + bm * N + c];
    const double pc1 = (double)in[row + b * N + cp];
    const double pcm = (double)in[row + b * N + cm];
    const int64_t i = row + col;
    const double piv = (double)in[N3 + i];
    const double kphi = piv;
    const double kpi = rhs_pi(cur, next, prev, pb1, pbm, pc1, pcm, piv, rk, LAM);
    if constexpr (MODE == 0) {
      io.yout[i] = (T)kphi;
      io.yout[N3 + i] = (T)kpi;
    } else if constexpr (MODE == 1) {
      // FP64: acc holds 1*x + dt6*k1 (+ dt3*k2 + dt3*k3); y = 1*x + cy*k.
      if constexpr (S == 1) {
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
    } else {
      // F32C: D is the float increment, y = f32(hi + cy*k); stage 4 folds the
      // increment into hi + lo in double and re-splits.
      if constexpr (S == 1) {
        io.yout[i] = (float)(cur + io.cy * kphi);
        io.yout[N3 + i] = (float)(piv + io.cy * kpi);
        io.acc[i] = (float)(io.ca * kphi);
        io.acc[N3 + i] = (float)(io.ca * kpi);
      } else if constexpr (S < 4) {
        const double h0 = (double)io.x[i], h1 = (double)io.x[N3 + i];
        io.yout[i] = (float)(h0 + io.cy * kphi);
        io.yout[N3 + i] = (float)(h1 + io.cy * kpi);
        io.acc[i] = (float)((double)io.acc[i] + io.ca * kphi);
        io.acc[N3 + i] = (float)((double)io.acc[N3 + i] + io.ca * kpi);
      } else {
        const double s0 = ((double)io.x[i] + (double)io.lo[i]) + ((double)io.acc[i] + io.ca * kphi);
        const double s1 = ((double)io.x[N3 + i] + (double)io.lo[N3 + i]) +
                          ((double)io.acc[N3 + i] + io.ca * kpi);
        const float h0 = (float)s0, h1 = (float)s1;
        io.x[i] = h0;
        io.x[N3 + i] = h1;
        io.lo[i] = (float)(s0 - (double)h0);
        io.lo[N3 + i] = (float)(s1 - (double)h1);
      }
    }
    prev = cur;
    cur = next;
  }
}

// Energy density (equations.cpp:106-138 order; lambda term as
// energy_density_kernel, equations_cuda.cu:332-354).
struct RhoCoef {
  double c_m, c_q, c_e, c_l;  // m*m, 0.25*inv, (0.5*0.25)*inv, 0.25*lambda
};

template <bool LAM>
__global__ __launch_bounds__(256) void k_energy_density(Tiling tl, RhoCoef rc, const double *__restrict__ x,
                                                        double *__restrict__ rho) {
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
    const double next = x[ap * NN + col];
    const double da = next - prev;
    const double db = x[row + bp * N + c] - x[row + bm * N + c];
    const double dc = x[row + b * N + cp] - x[row + b * N + cm];
    const double piv = x[N3 + row + col];
    double r = 0.5 * ((piv * piv + rc.c_m * (cur * cur)) + rc.c_q * (da * da + db * db));
    r = r + rc.c_e * (dc * dc);
    if (LAM) r = r + (((rc.c_l * cur) * cur) * cur) * cur;
    rho[row + col] = r;
    prev = cur;
    cur = next;
  }
}

// rho.reshaped(N*N, N).rowwise().mean() (observer.hpp:114): for each (b, c) the
// sum over a in increasing a, then one division by N.
__global__ void k_axis_average(int64_t N, const double *__restrict__ rho, double *__restrict__ avg) {
  const int64_t n2 = N * N;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += (int64_t)gridDim.x * blockDim.x) {
    double s = rho[i];
    for (int64_t a = 1; a < N; ++a) s = s + rho[a * n2 + i];
    avg[i] = s / (double)N;
  }
}

__global__ void k_split(const double *__restrict__ x, float *__restrict__ hi, float *__restrict__ lo,
                        int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double v = x[i];
    const float h = (float)v;
    hi[i] = h;
    lo[i] = (float)(v - (double)h);
  }
}

__global__ void k_join(const float *__restrict__ hi, const float *__restrict__ lo, double *__restrict__ x,
                       int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] = (double)hi[i] + (double)lo[i];
}

static RhsCoef rhs_coef(double m, double lambda, double H, double inv) {
  RhsCoef k;
  k.c_h = (-3.0) * H;
  k.c_m = m * m;
  k.lam = lambda;
  k.inv = inv;
  return k;
}

template <class T, int MODE, int S>
static cf_status launch_march(const Tiling &tl, const RhsCoef &rk, const StageIO<T> &io,
                              hipStream_t st) {
  dim3 blk(TC, TB), grd(tl.total);
  if (rk.lam != 0.0)
    hipLaunchKernelGGL((k_march<T, MODE, S, true>), grd, blk, 0, st, tl, rk, io);
  else
    hipLaunchKernelGGL((k_march<T, MODE, S, false>), grd, blk, 0, st, tl, rk, io);
  CF_LAUNCHED();
  return CF_OK;
}

static cf_status check_params(const cf_kg_params *p) {
  CF_REQUIRE(p, "null cf_kg_params");
  CF_REQUIRE(valid_N(p->N), "unsupported lattice size N=%lld (even, 4..2048)", (long long)p->N);
  CF_REQUIRE(p->L > 0.0, "box size L must be positive");
  return CF_OK;
}

// Stage times and RHS coefficients of one RK4 step at t (runge_kutta4_classic.hpp:97-123).
static void stage_coefs(const cf_kg_params *p, double t, double dt, RhsCoef rk[4]) {
  const double dh = dt / 2.0;
  const double th = t + dh;
  const double ts[4] = {t, th, th, t + dt};
  for (int s = 0; s < 4; ++s) {
    const double a_t = cf_cosmo_a(&p->cosmo, ts[s]);
    const double H_t = cf_cosmo_H(&p->cosmo, ts[s]);
    rk[s] = rhs_coef(p->m, p->lambda, H_t, cf_inv_ah_sqr(p->N, p->L, a_t));
  }
}

}  // namespace cf

using namespace cf;

extern "C" {

cf_status cf_rhs_f64(int64_t N, const double *x, double *dxdt, double m, double lambda, double H,
                     double inv_ah2, cf_stream stream) {
  CF_REQUIRE(valid_N(N), "cf_rhs_f64: unsupported N=%lld", (long long)N);
  CF_REQUIRE(x && dxdt && x != dxdt, "cf_rhs_f64: null or aliased buffers");
  Tiling tl = make_tiling(N, default_chunk(N));
  StageIO<double> io{};
  io.in = x;
  io.yout = dxdt;
  return launch_march<double, 0, 0>(tl, rhs_coef(m, lambda, H, inv_ah2), io, S(stream));
}

size_t cf_rk4_work_bytes(int64_t N, int precision) {
  if (!valid_N(N)) return 0;
  const size_t n = 2 * (size_t)N * N * N;
  return precision == CF_PREC_F32C ? 3 * n * sizeof(float) : 3 * n * sizeof(double);
}

cf_status cf_rk4_step_f64(const cf_kg_params *p, double *x, double t, double dt, void *work,
                          cf_stream stream) {
  CF_TRY(check_params(p));
  CF_REQUIRE(x && work, "cf_rk4_step_f64: null buffer");
  const int64_t n = 2 * p->N * p->N * p->N;
  double *yA = (double *)work, *yB = yA + n, *acc = yB + n;
  RhsCoef rk[4];
  stage_coefs(p, t, dt, rk);
  const double dh = dt / 2.0, dt6 = dt / 6.0, dt3 = dt / 3.0;
  Tiling tl = make_tiling(p->N, default_chunk(p->N));
  hipStream_t st = S(stream);
  StageIO<double> io{};
  io.x = x;
  io.acc = acc;
  // k1 = f(x, t); y1 = x + dh*k1; acc = 1*x + dt6*k1
  io.in = x; io.yout = yA; io.cy = dh; io.ca = dt6;
  CF_TRY((launch_march<double, 1, 1>(tl, rk[0], io, st)));
  // k2 = f(y1, t+dh); y2 = x + dh*k2; acc += dt3*k2
  io.in = yA; io.yout = yB; io.cy = dh; io.ca = dt3;
  CF_TRY((launch_march<double, 1, 2>(tl, rk[1], io, st)));
  // k3 = f(y2, t+dh); y3 = x + dt*k3; acc += dt3*k3
  io.in = yB; io.yout = yA; io.cy = dt; io.ca = dt3;
  CF_TRY((launch_march<double, 1, 3>(tl, rk[2], io, st)));
  // k4 = f(y3, t+dt); x = acc + dt6*k4
  io.in = yA; io.yout = nullptr; io.ca = dt6;
  CF_TRY((launch_march<double, 1, 4>(tl, rk[3], io, st)));
  return CF_OK;
}

cf_status cf_rk4_run_f64(const cf_kg_params *p, double *x, double t0, int64_t k0, int64_t n_steps,
                         double dt, void *work, cf_stream stream) {
  CF_REQUIRE(n_steps >= 0, "cf_rk4_run_f64: negative step count");
  for (int64_t k = 0; k < n_steps; ++k) {
    const double t = t0 + static_cast<double>(k0 + k) * dt;
    CF_TRY(cf_rk4_step_f64(p, x, t, dt, work, stream));
  }
  return CF_OK;
}

cf_status cf_rk4_step_f32c(const cf_kg_params *p, float *hi, float *lo, double t, double dt, void *work,
                           cf_stream stream) {
  CF_TRY(check_params(p));
  CF_REQUIRE(hi && lo && work && hi != lo, "cf_rk4_step_f32c: null or aliased buffer");
  const int64_t n = 2 * p->N * p->N * p->N;
  float *yA = (float *)work, *yB = yA + n, *D = yB + n;
  RhsCoef rk[4];
  stage_coefs(p, t, dt, rk);
  const double dh = dt / 2.0, dt6 = dt / 6.0, dt3 = dt / 3.0;
  Tiling tl = make_tiling(p->N, default_chunk(p->N));
  hipStream_t st = S(stream);
  StageIO<float> io{};
  io.x = hi;
  io.lo = lo;
  io.acc = D;
  io.in = hi; io.yout = yA; io.cy = dh; io.ca = dt6;
  CF_TRY((launch_march<float, 2, 1>(tl, rk[0], io, st)));
  io.in = yA; io.yout = yB; io.cy = dh; io.ca = dt3;
  CF_TRY((launch_march<float, 2, 2>(tl, rk[1], io, st)));
  io.in = yB; io.yout = yA; io.cy = dt; io.ca = dt3;
  CF_TRY((launch_march<float, 2, 3>(tl, rk[2], io, st)));
  io.in = yA; io.yout = nullptr; io.ca = dt6;
  CF_TRY((launch_march<float, 2, 4>(tl, rk[3], io, st)));
  return CF_OK;
}

cf_status cf_rk4_run_f32c(const cf_kg_params *p, float *hi, float *lo, double t0, int64_t k0,
                          int64_t n_steps, double dt, void *work, cf_stream stream) {
  CF_REQUIRE(n_steps >= 0, "cf_rk4_run_f32c: negative step count");
  for (int64_t k = 0; k < n_steps; ++k) {
    const double t = t0 + static_cast<double>(k0 + k) * dt;
    CF_TRY(cf_rk4_step_f32c(p, hi, lo, t, dt, work, stream));
  }
  return CF_OK;
}

cf_status cf_split_f64(const double *x, float *hi, float *lo, int64_t n, cf_stream stream) {
  CF_REQUIRE(x && hi && lo && n >= 0, "cf_split_f64: bad arguments");
  if (n == 0) return CF_OK;
  int64_t g = cdiv(n, 256);
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(k_split, dim3((unsigned)g), dim3(256), 0, S(stream), x, hi, lo, n);
  CF_LAUNCHED();
  return CF_OK;
}

cf_status cf_join_f64(const float *hi, const float *lo, double *x, int64_t n, cf_stream stream) {
  CF_REQUIRE(x && hi && lo && n >= 0, "cf_join_f64: bad arguments");
  if (n == 0) return CF_OK;
  int64_t g = cdiv(n, 256);
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(k_join, dim3((unsigned)g), dim3(256), 0, S(stream), hi, lo, x, n);
  CF_LAUNCHED();
  return CF_OK;
}

cf_status cf_axis_average(int64_t N, const double *rho, double *avg, cf_stream stream) {
  CF_REQUIRE(valid_N(N), "cf_axis_average: unsupported N=%lld", (long long)N);
  CF_REQUIRE(rho && avg && (const void *)rho != (const void *)avg, "cf_axis_average: null or aliased buffers");
  int64_t g = cdiv(N * N, 256);
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(k_axis_average, dim3((unsigned)g), dim3(256), 0, S(stream), N, rho, avg);
  CF_LAUNCHED();
  return CF_OK;
}

cf_status cf_energy_density_f64(int64_t N, const double *x, double *rho, double m, double lambda,
                                double inv_ah2, cf_stream stream) {
  CF_REQUIRE(valid_N(N), "cf_energy_density_f64: unsupported N=%lld", (long long)N);
  CF_REQUIRE(x && rho, "cf_energy_density_f64: null buffer");
  RhoCoef rc;
  rc.c_m = m * m;
  rc.c_q = 0.25 * inv_ah2;
  rc.c_e = 0.5 * 0.25 * inv_ah2;
  rc.c_l = 0.25 * lambda;
  Tiling tl = make_tiling(N, default_chunk(N));
  dim3 blk(TC, TB), grd(tl.total);
  if (lambda != 0.0)
    hipLaunchKernelGGL(k_energy_density<true>, grd, blk, 0, S(stream), tl, rc, x, rho);
  else
    hipLaunchKernelGGL(k_energy_density<false>, grd, blk, 0, S(stream), tl, rc, x, rho);
  CF_LAUNCHED();
  return CF_OK;
}

}  // extern "C"
