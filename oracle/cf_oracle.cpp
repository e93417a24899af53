// cf_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the Cosmic-Fields-Lite lattice-evolution path, used as the
// parity checker for the HIP library. Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it; nothing in the product links it.
//
// Every function follows the reference file:line cited beside it, in the same
// floating-point operation order (the build uses -ffp-contract=off so that the
// written order is the evaluated order). Where the reference has no CPU
// counterpart (the lambda*phi^4 term), the formula of the reference's CUDA
// kernel is used and the citation says so.
//
// Each function's parity pins (golden vectors, the reference's own integrator
// build, closed forms and planted mistakes) are the tests/test_oracle_*.py files.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

typedef int64_t i64;

// fdm3d.hpp:16  IDX_OF(N,i,j,k) = N*N*i + N*j + k   (c fastest)
static inline i64 IDX(i64 N, i64 a, i64 b, i64 c) { return N * N * a + N * b + c; }

static int g_threads = 1;

extern "C" void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }

template <class F>
static void parallel_for(i64 n, F f) {
  int nt = g_threads;
  if (nt <= 1 || n < 2) {
    for (i64 i = 0; i < n; ++i) f(i);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t) {
    th.emplace_back([=]() {
      for (i64 i = t; i < n; i += nt) f(i);
    });
  }
  for (auto &x : th) x.join();
}

// ---------------------------------------------------------------------------
// Cosmology: physics.hpp:15-56 (StaticEOSCosmology)
// ---------------------------------------------------------------------------
struct OrcCosmo {
  double a1, H1, t1, p;
};

// physics.hpp:34-36
extern "C" double orc_cosmo_a(const OrcCosmo *c, double t) {
  return c->a1 * pow(1 + (1 + 1 / c->p) * c->H1 * (t - c->t1), c->p / (1 + c->p));
}
// physics.hpp:39-41
extern "C" double orc_cosmo_H(const OrcCosmo *c, double t) {
  return c->H1 * pow(1 + (1 + 1 / c->p) * c->H1 * (t - c->t1), -1);
}
// physics.hpp:52-55
extern "C" double orc_cosmo_eta(const OrcCosmo *c, double t) {
  return (c->p / (c->a1 * c->H1)) * pow(1 + (1 + 1 / c->p) * c->H1 * (t - c->t1), 1 / (1 + c->p));
}

// equations.cpp:80 inv_ah_sqr = 1.0 / ((L / N) * (L / N)) / (a_t * a_t)
extern "C" double orc_inv_ah_sqr(i64 N, double L, double a_t) {
  return 1.0 / ((L / N) * (L / N)) / (a_t * a_t);
}

// ---------------------------------------------------------------------------
// Right-hand side
// ---------------------------------------------------------------------------
// KleinGordonEquationInFRW::operator() -- equations.cpp:71-104.
//   dphi = pi
//   dpi  = ((-3H)*pi - (m*m)*phi) + inv_ah2*((((-6phi + phi_a+) + phi_a-) + phi_b+) + phi_b-)
//   dpi += inv_ah2*(phi_c+ + phi_c-)          (second pass over c, :94-100)
// With lambda != 0 the term -((lambda*phi)*phi)*phi is subtracted before the
// stencil, as in Lambda_FRW_equation_kernel (equations_cuda.cu:38-61); the
// reference has no CPU lambda equation.
// KleinGordonEquation (equations.cpp:8-36) is this function with H=0, a=1:
// (-0*pi - m^2 phi) == (-(m*m))*phi bit for bit.
extern "C" void orc_rhs(i64 N, const double *x, double *dxdt, double m, double lambda,
                        double H, double inv_ah2) {
  const i64 N3 = N * N * N;
  const double c_h = (-3.0) * H;
  const double c_m = m * m;
  parallel_for(N, [&](i64 a) {
    for (i64 b = 0; b < N; ++b) {
      const double *p0 = x + IDX(N, a, b, 0);
      const double *pa1 = x + IDX(N, (a + 1) % N, b, 0);
      const double *pam = x + IDX(N, (a + N - 1) % N, b, 0);
      const double *pb1 = x + IDX(N, a, (b + 1) % N, 0);
      const double *pbm = x + IDX(N, a, (b + N - 1) % N, 0);
      const double *pi = x + N3 + IDX(N, a, b, 0);
      double *dphi = dxdt + IDX(N, a, b, 0);
      double *dpi = dxdt + N3 + IDX(N, a, b, 0);
      for (i64 c = 0; c < N; ++c) dphi[c] = pi[c];
      if (lambda != 0.0) {
        for (i64 c = 0; c < N; ++c) {
          double v = (c_h * pi[c] - c_m * p0[c]) - ((lambda * p0[c]) * p0[c]) * p0[c];
          dpi[c] = v + inv_ah2 * ((((-6.0 * p0[c] + pa1[c]) + pam[c]) + pb1[c]) + pbm[c]);
        }
      } else {
        for (i64 c = 0; c < N; ++c) {
          double v = c_h * pi[c] - c_m * p0[c];
          dpi[c] = v + inv_ah2 * ((((-6.0 * p0[c] + pa1[c]) + pam[c]) + pb1[c]) + pbm[c]);
        }
      }
      for (i64 c = 1; c < N - 1; ++c) dpi[c] = dpi[c] + inv_ah2 * (p0[c + 1] + p0[c - 1]);
      dpi[0] = dpi[0] + inv_ah2 * (p0[N - 1] + p0[1]);
      dpi[N - 1] = dpi[N - 1] + inv_ah2 * (p0[N - 2] + p0[0]);
    }
  });
}

// Energy density -- KleinGordonEquationInFRW::compute_energy_density,
// equations.cpp:106-138:
//   rho  = 0.5*((pi^2 + (m*m)*phi^2) + (0.25*inv)*((phi_a+ - phi_a-)^2 + (phi_b+ - phi_b-)^2))
//   rho += ((0.5*0.25)*inv)*(phi_c+ - phi_c-)^2
// lambda != 0 adds ((((0.25*lambda)*phi)*phi)*phi)*phi (energy_density_kernel,
// equations_cuda.cu:332-354; no CPU counterpart in the reference).
extern "C" void orc_energy_density(i64 N, const double *x, double *rho, double m,
                                   double lambda, double inv_ah2) {
  const i64 N3 = N * N * N;
  const double c_m = m * m;
  const double c_q = 0.25 * inv_ah2;
  const double c_e = 0.5 * 0.25 * inv_ah2;
  const double c_l = 0.25 * lambda;
  parallel_for(N, [&](i64 a) {
    for (i64 b = 0; b < N; ++b) {
      const double *p0 = x + IDX(N, a, b, 0);
      const double *pa1 = x + IDX(N, (a + 1) % N, b, 0);
      const double *pam = x + IDX(N, (a + N - 1) % N, b, 0);
      const double *pb1 = x + IDX(N, a, (b + 1) % N, 0);
      const double *pbm = x + IDX(N, a, (b + N - 1) % N, 0);
      const double *pi = x + N3 + IDX(N, a, b, 0);
      double *r = rho + IDX(N, a, b, 0);
      for (i64 c = 0; c < N; ++c) {
        double da = pa1[c] - pam[c];
        double db = pb1[c] - pbm[c];
        r[c] = 0.5 * ((pi[c] * pi[c] + c_m * (p0[c] * p0[c])) + c_q * (da * da + db * db));
      }
      for (i64 c = 1; c < N - 1; ++c) {
        double dc = p0[c + 1] - p0[c - 1];
        r[c] = r[c] + c_e * (dc * dc);
      }
      {
        double d0 = p0[1] - p0[N - 1];
        r[0] = r[0] + c_e * (d0 * d0);
        double d1 = p0[0] - p0[N - 2];
        r[N - 1] = r[N - 1] + c_e * (d1 * d1);
      }
      if (lambda != 0.0)
        for (i64 c = 0; c < N; ++c) r[c] = r[c] + (((c_l * p0[c]) * p0[c]) * p0[c]) * p0[c];
    }
  });
}

// ---------------------------------------------------------------------------
// Classic RK4 step: runge_kutta4_classic::do_step_impl
// (external/boost/numeric/odeint/stepper/runge_kutta4_classic.hpp:87-138) via
// explicit_stepper_base::do_step_v1 (explicit_stepper_base.hpp:224-230), with
// eigen_operations::scale_sum2 / scale_sum5 (eigen_operations.hpp:33,87):
//   k1 = f(x,t); y = 1*x + dh*k1; k2 = f(y,t+dh); y = 1*x + dh*k2; k3 = f(y,t+dh);
//   y = 1*x + dt*k3; k4 = f(y,t+dt);
//   x = (((1*x + dt6*k1) + dt3*k2) + dt3*k3) + dt6*k4
// ---------------------------------------------------------------------------
struct OrcEq {
  i64 N;
  double L, m, lambda;
  OrcCosmo cosmo;
};

static void eq_rhs(const OrcEq *e, const double *x, double *dxdt, double t) {
  double a_t = orc_cosmo_a(&e->cosmo, t);
  double H_t = orc_cosmo_H(&e->cosmo, t);
  double inv = orc_inv_ah_sqr(e->N, e->L, a_t);
  orc_rhs(e->N, x, dxdt, e->m, e->lambda, H_t, inv);
}

// work must hold 5 * 2N^3 doubles (dxdt, dxt, dxm, dxh, x_tmp).
extern "C" void orc_rk4_step(const OrcEq *e, double *x, double t, double dt, double *work) {
  const i64 n = 2 * e->N * e->N * e->N;
  double *k1 = work, *k2 = work + n, *k3 = work + 2 * n, *k4 = work + 3 * n, *y = work + 4 * n;
  const double val1 = 1.0;
  const double dh = dt / 2.0;
  const double th = t + dh;
  eq_rhs(e, x, k1, t);
  for (i64 i = 0; i < n; ++i) y[i] = val1 * x[i] + dh * k1[i];
  eq_rhs(e, y, k2, th);
  for (i64 i = 0; i < n; ++i) y[i] = val1 * x[i] + dh * k2[i];
  eq_rhs(e, y, k3, th);
  for (i64 i = 0; i < n; ++i) y[i] = val1 * x[i] + dt * k3[i];
  eq_rhs(e, y, k4, t + dt);
  const double dt6 = dt / 6.0, dt3 = dt / 3.0;
  for (i64 i = 0; i < n; ++i)
    x[i] = (((1.0 * x[i] + dt6 * k1[i]) + dt3 * k2[i]) + dt3 * k3[i]) + dt6 * k4[i];
}

// integrate_const with an explicit stepper (detail/integrate_const.hpp:42-68):
// time_k = t0 + k*dt; loop while less_eq_with_sign(time+dt, t_end, dt)
// (util/detail/less_with_sign.hpp:46-53). This entry only runs the steps (no
// observer); the observer cadence is restated in cf_oracle.py. Returns the
// number of steps taken.
extern "C" int orc_rk4_run(const OrcEq *e, double *x, double t0, double dt, int n_steps,
                           double *work) {
  for (int k = 0; k < n_steps; ++k) {
    double t = t0 + static_cast<double>(k) * dt;
    orc_rk4_step(e, x, t, dt, work);
  }
  return n_steps;
}

// FP32-compensated RK4 step, the arithmetic the HIP kernels define for the
// single-precision storage mode (no reference counterpart; restated here so the
// GPU path can be checked bit for bit). State = hi + lo, both float. Stage
// inputs and the increment D are float; every product and sum is double.
//   s1: k1 = f(hi);  D = f32(dt6*k1);            y = f32(hi + dh*k1)
//   s2: k2 = f(y);   D = f32(D + dt3*k2);        y = f32(hi + dh*k2)
//   s3: k3 = f(y);   D = f32(D + dt3*k3);        y = f32(hi + dt*k3)
//   s4: k4 = f(y);   s = (hi + lo) + (D + dt6*k4); hi = f32(s); lo = f32(s - hi)
// f() is orc_rhs evaluated in double on the widened float inputs.
extern "C" void orc_rk4_step_f32c(const OrcEq *e, float *hi, float *lo, double t, double dt,
                                  double *work /* 2 * 2N^3 doubles */, float *fwork /* 2 * 2N^3 floats */) {
  const i64 n = 2 * e->N * e->N * e->N;
  double *xd = work, *k = work + n;
  float *D = fwork, *y = fwork + n;
  const double dh = dt / 2.0, th = t + dh, dt6 = dt / 6.0, dt3 = dt / 3.0;
  for (i64 i = 0; i < n; ++i) xd[i] = (double)hi[i];
  eq_rhs(e, xd, k, t);
  for (i64 i = 0; i < n; ++i) {
    D[i] = (float)(dt6 * k[i]);
    y[i] = (float)((double)hi[i] + dh * k[i]);
  }
  for (int s = 2; s <= 3; ++s) {
    for (i64 i = 0; i < n; ++i) xd[i] = (double)y[i];
    eq_rhs(e, xd, k, th);
    const double cy = (s == 2) ? dh : dt;
    for (i64 i = 0; i < n; ++i) {
      D[i] = (float)((double)D[i] + dt3 * k[i]);
      y[i] = (float)((double)hi[i] + cy * k[i]);
    }
  }
  for (i64 i = 0; i < n; ++i) xd[i] = (double)y[i];
  eq_rhs(e, xd, k, t + dt);
  for (i64 i = 0; i < n; ++i) {
    double s = ((double)hi[i] + (double)lo[i]) + ((double)D[i] + dt6 * k[i]);
    float h = (float)s;
    hi[i] = h;
    lo[i] = (float)(s - (double)h);
  }
}

// The compensated storage with the single-precision stencil (no reference
// counterpart; the arithmetic csrc/cf_tb_body.hpp defines for it, restated so the GPU
// path can be checked bit for bit). State = hi + lo, both float. Each stage's force is
// float arithmetic on float inputs with the coefficients rounded to float (fma: one
// rounding):
//   v = fma((-3H)f, pi, -((m*m)f * phi));  v = fma(-(((lambda f) * phi) * phi), phi, v)
//   s7 = ((((fma(-6, phi, phi_a+) + phi_a-) + phi_b+) + phi_b-) + phi_c+) + phi_c-
//   dpi = fma((inv_ah_sqr)f, s7, v);  dphi = pi
// the stage inputs y = fma(c, k, hi) in float (c = (dt/2)f, (dt/2)f, (dt)f), the
// increment D = dt/6 k1, then D = fma(dt/3, k, D) twice, in double, and
//   s = (hi + lo) + fma(dt/6, k4, D);  hi = f32(s);  lo = f32(s - hi)   (double).
static void rhs_f32(const OrcEq *e, const float *y, float *k, double t) {
  const i64 N = e->N, N3 = N * N * N;
  const double a_t = orc_cosmo_a(&e->cosmo, t);
  const float c_h = (float)((-3.0) * orc_cosmo_H(&e->cosmo, t));
  const float inv = (float)orc_inv_ah_sqr(N, e->L, a_t);
  const float c_m = (float)(e->m * e->m), lam = (float)e->lambda;
  const bool has_lam = e->lambda != 0.0;
  parallel_for(N, [&](i64 a) {
    for (i64 b = 0; b < N; ++b) {
      const float *p0 = y + IDX(N, a, b, 0);
      const float *pa1 = y + IDX(N, (a + 1) % N, b, 0);
      const float *pam = y + IDX(N, (a + N - 1) % N, b, 0);
      const float *pb1 = y + IDX(N, a, (b + 1) % N, 0);
      const float *pbm = y + IDX(N, a, (b + N - 1) % N, 0);
      const float *pi = y + N3 + IDX(N, a, b, 0);
      float *dphi = k + IDX(N, a, b, 0);
      float *dpi = k + N3 + IDX(N, a, b, 0);
      for (i64 c = 0; c < N; ++c) {
        const float phi = p0[c];
        const float pc1 = p0[(c + 1) % N], pcm = p0[(c + N - 1) % N];
        float v = std::fmaf(c_h, pi[c], -(c_m * phi));
        if (has_lam) v = std::fmaf(-((lam * phi) * phi), phi, v);
        const float s7 = ((((std::fmaf(-6.0f, phi, pa1[c]) + pam[c]) + pb1[c]) + pbm[c]) + pc1) + pcm;
        dphi[c] = pi[c];
        dpi[c] = std::fmaf(inv, s7, v);
      }
    }
  });
}

extern "C" void orc_rk4_step_f32s(const OrcEq *e, float *hi, float *lo, double t, double dt,
                                  double *work /* 2N^3 doubles */, float *fwork /* 2 * 2N^3 floats */) {
  const i64 n = 2 * e->N * e->N * e->N;
  double *D = work;
  float *y = fwork, *k = fwork + n;
  const double dh = dt / 2.0, th = t + dh, dt6 = dt / 6.0, dt3 = dt / 3.0;
  const float cy[3] = {(float)dh, (float)dh, (float)dt};
  const double ts[4] = {t, th, th, t + dt};
  rhs_f32(e, hi, k, ts[0]);
  for (i64 i = 0; i < n; ++i) {
    D[i] = dt6 * (double)k[i];
    y[i] = std::fmaf(cy[0], k[i], hi[i]);
  }
  for (int s = 1; s <= 2; ++s) {
    rhs_f32(e, y, k, ts[s]);
    for (i64 i = 0; i < n; ++i) {
      D[i] = std::fma(dt3, (double)k[i], D[i]);
      y[i] = std::fmaf(cy[s], k[i], hi[i]);
    }
  }
  rhs_f32(e, y, k, ts[3]);
  for (i64 i = 0; i < n; ++i) {
    const double s = ((double)hi[i] + (double)lo[i]) + std::fma(dt6, (double)k[i], D[i]);
    const float h = (float)s;
    hi[i] = h;
    lo[i] = (float)(s - (double)h);
  }
}

// ---------------------------------------------------------------------------
// FFT: a mixed-radix Stockham FFT written for the oracle (FFTW is absent from
// this image). Conventions follow FFTW's r2c/c2r 3-D transforms as used by
// fftwWrapper (fftw_wrapper.cpp:4-76): forward exp(-2 pi i k.x/N), unnormalised;
// the inverse is unnormalised; the half-spectrum is N x N x (N/2+1), interleaved.
// ---------------------------------------------------------------------------
struct cd {
  double re, im;
};
static inline cd cadd(cd a, cd b) { return {a.re + b.re, a.im + b.im}; }
static inline cd csub(cd a, cd b) { return {a.re - b.re, a.im - b.im}; }
static inline cd cmul(cd a, cd b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }

struct Fft1 {
  int n = 0, sign = -1;
  std::vector<int> radix;
  std::vector<std::vector<cd>> tw;  // tw[stage][p*r + k] = w_{len}^{p*k}
  std::vector<std::vector<cd>> om;  // om[stage][j*r + k] = omega_r^{j*k}

  Fft1(int n_, int sign_) : n(n_), sign(sign_) {
    int m = n;
    while (m > 1) {
      int r = 0;
      if (m % 4 == 0) r = 4;
      else if (m % 2 == 0) r = 2;
      else if (m % 3 == 0) r = 3;
      else if (m % 5 == 0) r = 5;
      else {
        for (int f = 7; f <= m; f += 2)
          if (m % f == 0) { r = f; break; }
      }
      radix.push_back(r);
      m /= r;
    }
    int len = n;
    for (int r : radix) {
      int mm = len / r;
      std::vector<cd> t((size_t)mm * r);
      for (int p = 0; p < mm; ++p)
        for (int k = 0; k < r; ++k) {
          long double ang = (long double)sign * 2.0L * 3.14159265358979323846264338327950288L *
                            (long double)(((long long)p * k) % len) / (long double)len;
          t[(size_t)p * r + k] = {(double)cosl(ang), (double)sinl(ang)};
        }
      tw.push_back(t);
      std::vector<cd> o((size_t)r * r);
      for (int j = 0; j < r; ++j)
        for (int k = 0; k < r; ++k) {
          long double ang = (long double)sign * 2.0L * 3.14159265358979323846264338327950288L *
                            (long double)((j * k) % r) / (long double)r;
          o[(size_t)j * r + k] = {(double)cosl(ang), (double)sinl(ang)};
        }
      om.push_back(o);
      len = mm;
    }
  }

  // In-place transform of x (length n) using scratch y (length n).
  void exec(cd *x, cd *y) const {
    int len = n, s = 1;
    cd *src = x, *dst = y;
    cd a[64], b[64];
    for (size_t st = 0; st < radix.size(); ++st) {
      const int r = radix[st];
      const int mm = len / r;
      const cd *T = tw[st].data();
      const cd *O = om[st].data();
      for (int p = 0; p < mm; ++p) {
        for (int q = 0; q < s; ++q) {
          for (int j = 0; j < r; ++j) a[j] = src[q + s * (p + j * mm)];
          if (r == 2) {
            b[0] = cadd(a[0], a[1]);
            b[1] = csub(a[0], a[1]);
          } else if (r == 4) {
            cd t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
            cd t2 = cadd(a[1], a[3]), t3 = csub(a[1], a[3]);
            // multiply t3 by sign*i
            cd t3r = (sign < 0) ? cd{t3.im, -t3.re} : cd{-t3.im, t3.re};
            b[0] = cadd(t0, t2);
            b[2] = csub(t0, t2);
            b[1] = cadd(t1, t3r);
            b[3] = csub(t1, t3r);
          } else {
            for (int k = 0; k < r; ++k) {
              cd acc = a[0];
              for (int j = 1; j < r; ++j) acc = cadd(acc, cmul(a[j], O[(size_t)j * r + k]));
              b[k] = acc;
            }
          }
          dst[q + s * (r * p)] = b[0];
          for (int k = 1; k < r; ++k) dst[q + s * (r * p + k)] = cmul(b[k], T[(size_t)p * r + k]);
        }
      }
      std::swap(src, dst);
      len = mm;
      s *= r;
    }
    if (src != x) std::memcpy(x, src, sizeof(cd) * n);
  }
};

// Apply a 1-D transform along one axis of a complex array, batching B lines.
// data: base pointer; count lines with given line start offsets; stride between
// consecutive elements of a line; lines grouped so that `group` consecutive
// lines have starts at unit offsets (for cache-friendly gathers).
static void fft_axis(cd *data, const Fft1 &f, i64 n_outer, i64 outer_stride, i64 n_inner,
                     i64 elem_stride) {
  // lines: for o in [0,n_outer), for i in [0,n_inner): start = o*outer_stride + i
  const int n = f.n;
  const int B = 16;
  parallel_for(n_outer, [&](i64 o) {
    std::vector<cd> buf((size_t)B * n), scr(n);
    for (i64 i0 = 0; i0 < n_inner; i0 += B) {
      int nb = (int)std::min<i64>(B, n_inner - i0);
      cd *base = data + o * outer_stride + i0;
      for (int j = 0; j < n; ++j)
        for (int l = 0; l < nb; ++l) buf[(size_t)l * n + j] = base[(i64)j * elem_stride + l];
      for (int l = 0; l < nb; ++l) f.exec(&buf[(size_t)l * n], scr.data());
      for (int j = 0; j < n; ++j)
        for (int l = 0; l < nb; ++l) base[(i64)j * elem_stride + l] = buf[(size_t)l * n + j];
    }
  });
}

// r2c: in N^3 reals -> out 2*N*N*(N/2+1) doubles (fftw_execute_dft_r2c).
extern "C" void orc_fft_r2c(i64 N, const double *in, double *out) {
  const i64 H = N / 2 + 1;
  cd *X = reinterpret_cast<cd *>(out);
  Fft1 f((int)N, -1);
  parallel_for(N, [&](i64 a) {
    std::vector<cd> line(N), scr(N);
    for (i64 b = 0; b < N; ++b) {
      const double *src = in + IDX(N, a, b, 0);
      for (i64 c = 0; c < N; ++c) line[c] = {src[c], 0.0};
      f.exec(line.data(), scr.data());
      cd *dst = X + (a * N + b) * H;
      for (i64 c = 0; c < H; ++c) dst[c] = line[c];
    }
  });
  // along b: for each a, lines over c in [0,H), elements stride H
  fft_axis(X, f, N, N * H, H, H);
  // along a: one outer block, lines over (b,c) in [0, N*H), elements stride N*H
  fft_axis(X, f, 1, 0, N * H, N * H);
}

// c2r: in 2*N*N*(N/2+1) doubles -> out N^3 reals, unnormalised
// (fftw_execute_dft_c2r). The input is not modified (FFTW may destroy it).
extern "C" void orc_fft_c2r(i64 N, const double *in, double *out) {
  const i64 H = N / 2 + 1;
  std::vector<cd> X((size_t)(N * N * H));
  std::memcpy(X.data(), in, sizeof(cd) * X.size());
  Fft1 f((int)N, +1);
  fft_axis(X.data(), f, 1, 0, N * H, N * H);
  fft_axis(X.data(), f, N, N * H, H, H);
  parallel_for(N, [&](i64 a) {
    std::vector<cd> line(N), scr(N);
    for (i64 b = 0; b < N; ++b) {
      const cd *src = X.data() + (a * N + b) * H;
      for (i64 c = 0; c < H; ++c) line[c] = src[c];
      for (i64 c = H; c < N; ++c) line[c] = {src[N - c].re, -src[N - c].im};
      f.exec(line.data(), scr.data());
      double *dst = out + IDX(N, a, b, 0);
      for (i64 c = 0; c < N; ++c) dst[c] = line[c].re;
    }
  });
}

// Direct O(N^6) DFT in long double, for pinning orc_fft_r2c on tiny N.
extern "C" void orc_dft_r2c_direct(i64 N, const double *in, double *out) {
  const i64 H = N / 2 + 1;
  const long double tp = 2.0L * 3.14159265358979323846264338327950288L;
  for (i64 ka = 0; ka < N; ++ka)
    for (i64 kb = 0; kb < N; ++kb)
      for (i64 kc = 0; kc < H; ++kc) {
        long double sr = 0, si = 0;
        for (i64 a = 0; a < N; ++a)
          for (i64 b = 0; b < N; ++b)
            for (i64 c = 0; c < N; ++c) {
              long double ph = -tp * (long double)(((ka * a + kb * b + kc * c) % N)) / (long double)N;
              long double v = in[IDX(N, a, b, c)];
              sr += v * cosl(ph);
              si += v * sinl(ph);
            }
        i64 o = ((ka * N + kb) * H + kc) * 2;
        out[o] = (double)sr;
        out[o + 1] = (double)si;
      }
}

// ---------------------------------------------------------------------------
// Spectra
// ---------------------------------------------------------------------------
// compute_power_spectrum -- fdm3d.cpp:7-34. out has 3*(N/2)^2+1 entries.
extern "C" void orc_bin_power(i64 N, const double *fk, double *spectrum) {
  const i64 nb = 3 * (N / 2) * (N / 2) + 1;
  for (i64 s = 0; s < nb; ++s) spectrum[s] = 0.0;
  for (i64 a = 0; a < N; ++a)
    for (i64 b = 0; b < N; ++b)
      for (i64 c = 0; c < N; ++c) {
        i64 as = (a <= N / 2) ? a : (N - a);
        i64 bs = (b <= N / 2) ? b : (N - b);
        i64 cs = (c <= N / 2) ? c : (N - c);
        i64 s = as * as + bs * bs + cs * cs;
        i64 idx = N * (N / 2 + 1) * a + (N / 2 + 1) * b + cs;
        double re = fk[2 * idx + 0], im = fk[2 * idx + 1];
        spectrum[s] += re * re + im * im;
      }
}

extern "C" void orc_power_spectrum(i64 N, const double *f, double *spectrum) {
  std::vector<double> fk((size_t)(2 * N * N * (N / 2 + 1)));
  orc_fft_r2c(N, f, fk.data());
  orc_bin_power(N, fk.data(), spectrum);
}

// compute_mode_power_spectrum -- fdm3d.cpp:36-72.
extern "C" void orc_bin_mode_power(i64 N, double L, double m, double a_t, const double *phik,
                                   const double *dphik, double *spectrum) {
  const double pi = 3.14159265358979323846;
  const i64 nb = 3 * (N / 2) * (N / 2) + 1;
  for (i64 s = 0; s < nb; ++s) spectrum[s] = 0.0;
  for (i64 a = 0; a < N; ++a)
    for (i64 b = 0; b < N; ++b)
      for (i64 c = 0; c < N; ++c) {
        i64 as = (a <= N / 2) ? a : (N - a);
        i64 bs = (b <= N / 2) ? b : (N - b);
        i64 cs = (c <= N / 2) ? c : (N - c);
        i64 s = as * as + bs * bs + cs * cs;
        i64 idx = N * (N / 2 + 1) * a + (N / 2 + 1) * b + cs;
        double pr = phik[2 * idx + 0], pim = phik[2 * idx + 1];
        double dr = dphik[2 * idx + 0], dim = dphik[2 * idx + 1];
        spectrum[s] += pr * pr + pim * pim +
                       (dr * dr + dim * dim) /
                           (m * m + (2 * pi / L) * (2 * pi / L) * s / (a_t * a_t));
      }
}

extern "C" void orc_mode_power_spectrum(i64 N, double L, double m, double a_t, const double *state,
                                        double *spectrum) {
  const i64 N3 = N * N * N;
  std::vector<double> pk((size_t)(2 * N * N * (N / 2 + 1))), dk(pk.size());
  orc_fft_r2c(N, state, pk.data());
  orc_fft_r2c(N, state + N3, dk.data());
  orc_bin_mode_power(N, L, m, a_t, pk.data(), dk.data(), spectrum);
}

// ---------------------------------------------------------------------------
// Gaussian random fields -- random_field.cpp:13-89
// ---------------------------------------------------------------------------
// Draw n standard normals from std::mt19937(seed) + std::normal_distribution<double>
// (RandomNormal, random_field.cpp:13-32; a freshly seeded generator and a fresh
// distribution, i.e. no cached polar value).
extern "C" void orc_normals(uint32_t seed, i64 n, double *out) {
  std::mt19937 mt{seed};
  std::normal_distribution<double> dist{0.0, 1.0};
  for (i64 i = 0; i < n; ++i) out[i] = dist(mt);
}

// generate_inhomogeneous_gaussian_random_field (random_field.cpp:36-82), with
// the normals passed in (so several fields can share one generator stream) and
// sqrt(P(|k|)) passed as a table over s = |n|^2 (|k| = sqrt(s)*2*pi/L depends on
// s only). Psi may be null (generate_gaussian_random_field, :85-89).
extern "C" void orc_grf_from_normals(i64 N, const double *normals, const double *sqrtP_by_s,
                                     const double *Psi, double *out) {
  const i64 N3 = N * N * N;
  std::vector<double> phi((size_t)N3);
  for (i64 i = 0; i < N3; ++i) phi[i] = (1.0 / (double)(N * N * N)) * normals[i];
  if (Psi)
    for (i64 i = 0; i < N3; ++i) phi[i] = phi[i] * std::exp(Psi[i]);
  const i64 H = N / 2 + 1;
  std::vector<double> fk((size_t)(2 * N * N * H));
  orc_fft_r2c(N, phi.data(), fk.data());
  for (i64 a = 0; a < N; ++a)
    for (i64 b = 0; b < N; ++b)
      for (i64 c = 0; c <= N / 2; ++c) {
        i64 as = (a <= N / 2) ? a : (a - N);
        i64 bs = (b <= N / 2) ? b : (b - N);
        i64 cs = (c <= N / 2) ? c : (c - N);
        i64 s = as * as + bs * bs + cs * cs;
        double sp = sqrtP_by_s[s];
        i64 o = 2 * (N * H * a + H * b + c);
        fk[o] *= sp;
        fk[o + 1] *= sp;
      }
  orc_fft_c2r(N, fk.data(), out);
}

// Spectrum tables P(|k|) over s for the spectra the reference and the
// workloads use; |k| = sqrt(s)*2*pi/L as random_field.cpp:67.
//   kind 0: exp(-k^2/(2 sigma^2)), P(0)=0          (params: sigma)
//   kind 1: k^-3 exp(-k^2/2), P(0)=0               (no params)
//   kind 2: scale_invariant_spectrum_3d            (random_field.cpp:234-240; params: A_s)
//   kind 3: power_law_with_cutoff_given_amplitude_3d (:212-221; params: sigma, k_ast, alpha)
//   kind 4: to_deriv_spectrum(m, kind 3)           (:248-250; params: sigma, k_ast, alpha, m)
extern "C" void orc_spectrum_table(int kind, i64 N, double L, const double *prm, double *P_by_s) {
  const double pi = 3.14159265358979323846;
  const i64 nb = 3 * (N / 2) * (N / 2) + 1;
  double P_k_ast = 0.0, k_ast = 0.0, alpha = 0.0, P_k0 = 0.0;
  if (kind == 3 || kind == 4) {
    double sigma = prm[0];
    k_ast = prm[1];
    alpha = prm[2];
    P_k_ast = alpha == -3.0
                  ? 2 * pi * pi * sigma * sigma / (k_ast * k_ast * k_ast * log(k_ast * L / (2 * pi)))
                  : (3 + alpha) * 2 * pi * pi * sigma * sigma * (L * L * L * pow(k_ast * L, alpha)) /
                        (pow(k_ast * L, alpha + 3) - pow(2 * pi, alpha + 3));
    P_k_ast *= (N / L) * (N / L) * (N / L);
  }
  if (kind == 2) {
    double k0 = 1.0;
    P_k0 = prm[0] / ((k0 * k0 * k0) / (2 * pi * pi));
    P_k0 *= (N / L) * (N / L) * (N / L);
  }
  for (i64 s = 0; s < nb; ++s) {
    double k = sqrt((double)s) * 2 * pi / L;
    double P = 0.0;
    switch (kind) {
      case 0: P = (k == 0.0) ? 0.0 : exp(-(k * k) / (2 * prm[0] * prm[0])); break;
      case 1: P = (k == 0.0) ? 0.0 : pow(k, -3) * exp(-(k * k) / 2); break;
      case 2: P = (k == 0.0) ? 0.0 : (P_k0 * pow(k / 1.0, -3)); break;
      case 3: P = (k == 0.0) ? 0.0 : (P_k_ast * (k <= k_ast ? pow(k / k_ast, alpha) : 0.0)); break;
      case 4: {
        double m = prm[3];
        double Pf = (k == 0.0) ? 0.0 : (P_k_ast * (k <= k_ast ? pow(k / k_ast, alpha) : 0.0));
        P = (k * k + m * m) * Pf;
        break;
      }
      default: P = 0.0;
    }
    P_by_s[s] = P;
  }
}

// Rescale x (n entries) to a target rms, summing squares in index order.
extern "C" double orc_rescale_rms(i64 n, double *x, double target) {
  double acc = 0.0;
  for (i64 i = 0; i < n; ++i) acc += x[i] * x[i];
  double rms = std::sqrt(acc / (double)n);
  double f = target / rms;
  for (i64 i = 0; i < n; ++i) x[i] = x[i] * f;
  return rms;
}

// compute_field_with_scaled_fourier_modes -- fdm3d.cpp:76-102, with the kernel
// passed as a table over s (kernel(sqrt(s)*(2*pi/L)) depends on s only).
extern "C" void orc_scaled_fourier_modes(i64 N, const double *f, const double *kernel_by_s,
                                         double *out) {
  const i64 H = N / 2 + 1;
  std::vector<double> fk((size_t)(2 * N * N * H));
  orc_fft_r2c(N, f, fk.data());
  for (i64 a = 0; a < N; ++a)
    for (i64 b = 0; b < N; ++b)
      for (i64 c = 0; c <= N / 2; ++c) {
        i64 as = (a <= N / 2) ? a : (N - a);
        i64 bs = (b <= N / 2) ? b : (N - b);
        i64 cs = (c <= N / 2) ? c : (N - c);
        i64 s = as * as + bs * bs + cs * cs;
        i64 idx = N * H * a + H * b + cs;
        double scale = kernel_by_s[s] / (N * N * N);
        fk[2 * idx + 0] *= scale;
        fk[2 * idx + 1] *= scale;
      }
  orc_fft_c2r(N, fk.data(), out);
}

// compute_Comoving_Curvature_Psi_dPsidt_fft -- equations.cpp:141-183
extern "C" void orc_comoving_psi_fft(i64 N, const double *R_fft, double *Psi_fft, double *dPsidt_fft,
                                     double factor, double H_t) {
  for (i64 a = 0; a < N; ++a)
    for (i64 b = 0; b < N; ++b)
      for (i64 c = 0; c <= N / 2; ++c) {
        i64 as = (a <= N / 2) ? a : (N - a);
        i64 bs = (b <= N / 2) ? b : (N - b);
        i64 cs = (c <= N / 2) ? c : (N - c);
        i64 s = as * as + bs * bs + cs * cs;
        i64 idx = N * (N / 2 + 1) * a + (N / 2 + 1) * b + cs;
        i64 o = 2 * idx;
        if (s == 0) {
          Psi_fft[o] = Psi_fft[o + 1] = 0;
          if (dPsidt_fft) dPsidt_fft[o] = dPsidt_fft[o + 1] = 0;
        } else {
          double w = sqrt(static_cast<double>(s)) * factor;
          double sv = sin(w), cv = cos(w);
          double cf = 2.0 / (w * w * w) / (N * N * N);
          double tp = (sv - w * cv) * cf;
          Psi_fft[o] = R_fft[o] * tp;
          Psi_fft[o + 1] = R_fft[o + 1] * tp;
          if (dPsidt_fft) {
            double td = H_t * (3 * w * cv + (w * w - 3.0) * sv) * cf;
            dPsidt_fft[o] = R_fft[o] * td;
            dPsidt_fft[o + 1] = R_fft[o + 1] * td;
          }
        }
      }
}

// ComovingCurvatureEquationInFRW::compute_energy_density -- equations.cpp:262-322
// (Psi given in real space, N^3).
extern "C" void orc_curvature_energy_density(i64 N, const double *x, const double *Psi, double *rho,
                                             double m, double inv_ah2) {
  const i64 N3 = N * N * N;
  parallel_for(N, [&](i64 a) {
    for (i64 b = 0; b < N; ++b) {
      const double *p0 = x + IDX(N, a, b, 0);
      const double *pa1 = x + IDX(N, (a + 1) % N, b, 0);
      const double *pam = x + IDX(N, (a + N - 1) % N, b, 0);
      const double *pb1 = x + IDX(N, a, (b + 1) % N, 0);
      const double *pbm = x + IDX(N, a, (b + N - 1) % N, 0);
      const double *pi = x + N3 + IDX(N, a, b, 0);
      const double *ps = Psi + IDX(N, a, b, 0);
      double *r = rho + IDX(N, a, b, 0);
      for (i64 c = 0; c < N; ++c) {
        double da = pa1[c] - pam[c], db = pb1[c] - pbm[c];
        r[c] = 0.5 * ((std::exp(-2 * ps[c]) * (pi[c] * pi[c]) + m * m * (p0[c] * p0[c])) +
                      std::exp(2 * ps[c]) * 0.25 * inv_ah2 * (da * da + db * db));
      }
      for (i64 c = 1; c < N - 1; ++c) {
        double dc = p0[c + 1] - p0[c - 1];
        r[c] += std::exp(2 * ps[c]) * 0.5 * 0.25 * inv_ah2 * (dc * dc);
      }
      double d0 = p0[1] - p0[N - 1];
      r[0] += std::exp(2 * ps[0]) * 0.5 * 0.25 * inv_ah2 * (d0 * d0);
      double d1 = p0[0] - p0[N - 2];
      r[N - 1] += std::exp(2 * ps[N - 1]) * 0.5 * 0.25 * inv_ah2 * (d1 * d1);
    }
  });
}

// Ci_pade_approximant_12_12 -- special_function.hpp:17-26
extern "C" double orc_ci_pade(double x) {
  return log(x) + (0.5772156649015329 - 0.24231497614160186 * pow(x, 2) +
                   0.007139183039136621 * pow(x, 4) - 0.00011466618094101764 * pow(x, 6) +
                   8.443734405201243e-7 * pow(x, 8) - 3.060472574705558e-9 * pow(x, 10) +
                   4.328624073851291e-12 * pow(x, 12)) /
                      (1. + 0.013313955815300189 * pow(x, 2) + 0.00008836441800952094 * pow(x, 4) +
                       3.800404484365274e-7 * pow(x, 6) + 1.1376490214488613e-9 * pow(x, 8) +
                       2.297129602871981e-12 * pow(x, 10) + 2.510407760855278e-15 * pow(x, 12));
}

// ---------------------------------------------------------------------------
// Curvature / potential variants
// ---------------------------------------------------------------------------
// The 7-point sum of the CUDA kernels, all six neighbours in one expression:
// ((((((-6 phi + a+) + a-) + b+) + b-) + c+) + c-)   (equations_cuda.cu:28-34)
static inline double stencil7(i64 N, const double *x, i64 a, i64 b, i64 c) {
  return (((((-6.0 * x[IDX(N, a, b, c)] + x[IDX(N, (a + 1) % N, b, c)]) + x[IDX(N, (a + N - 1) % N, b, c)]) +
            x[IDX(N, a, (b + 1) % N, c)]) +
           x[IDX(N, a, (b + N - 1) % N, c)]) +
          x[IDX(N, a, b, (c + 1) % N)]) +
         x[IDX(N, a, b, (c + N - 1) % N)];
}

// Fixed_Curvature_FRW_equation_kernel (equations_cuda.cu:114-137); no CPU counterpart.
extern "C" void orc_rhs_fixed_curvature(i64 N, const double *x, double *dxdt, const double *Psi, double m,
                                        double H, double inv) {
  const i64 N3 = N * N * N;
  parallel_for(N, [&](i64 a) {
    for (i64 b = 0; b < N; ++b)
      for (i64 c = 0; c < N; ++c) {
        const i64 i = IDX(N, a, b, c);
        dxdt[i] = x[N3 + i];
        dxdt[N3 + i] = ((-3.0) * H * x[N3 + i] - std::exp(2 * Psi[i]) * m * m * x[i]) +
                       std::exp(4 * Psi[i]) * inv * stencil7(N, x, a, b, c);
      }
  });
}

// Sqrt_Potential_FRW_equation_kernel (equations_cuda.cu:83-106): V = m^2 f_a^2 (sqrt(1 + phi^2/f_a^2) - 1).
extern "C" void orc_rhs_sqrt_potential(i64 N, const double *x, double *dxdt, double m, double f_a, double H,
                                       double inv) {
  const i64 N3 = N * N * N;
  parallel_for(N, [&](i64 a) {
    for (i64 b = 0; b < N; ++b)
      for (i64 c = 0; c < N; ++c) {
        const i64 i = IDX(N, a, b, c);
        dxdt[i] = x[N3 + i];
        dxdt[N3 + i] = ((-3.0) * H * x[N3 + i] - m * m * x[i] / std::sqrt(1.0 + x[i] * x[i] / (f_a * f_a))) +
                       inv * stencil7(N, x, a, b, c);
      }
  });
}

// Sqrt_Potential_energy_density_kernel (equations_cuda.cu:422-444).
extern "C" void orc_energy_density_sqrt(i64 N, const double *x, double *rho, double m, double f_a, double inv) {
  const i64 N3 = N * N * N;
  parallel_for(N, [&](i64 a) {
    for (i64 b = 0; b < N; ++b)
      for (i64 c = 0; c < N; ++c) {
        const i64 i = IDX(N, a, b, c);
        const double f = x[i];
        const double dx = x[IDX(N, (a + 1) % N, b, c)] - x[IDX(N, (a + N - 1) % N, b, c)];
        const double dy = x[IDX(N, a, (b + 1) % N, c)] - x[IDX(N, a, (b + N - 1) % N, c)];
        const double dz = x[IDX(N, a, b, (c + 1) % N)] - x[IDX(N, a, b, (c + N - 1) % N)];
        rho[i] = 0.5 * x[N3 + i] * x[N3 + i] + (0.5 / 4.0) * inv * (dx * dx + dy * dy + dz * dz) +
                 m * m * f_a * f_a * (std::sqrt(1.0 + f * f / (f_a * f_a)) - 1.0);
      }
  });
}

// ComovingCurvatureEquationInFRW::operator() right-hand side given Psi, dPsi/dt
// in real space (equations.cpp:357-380): CPU order, c-neighbours added in a
// second pass.
extern "C" void orc_rhs_variable_curvature(i64 N, const double *x, double *dxdt, const double *Psi,
                                           const double *dPsidt, double m, double H, double inv) {
  const i64 N3 = N * N * N;
  parallel_for(N, [&](i64 a) {
    for (i64 b = 0; b < N; ++b) {
      const i64 r0 = IDX(N, a, b, 0);
      const double *p0 = x + r0, *pi = x + N3 + r0, *ps = Psi + r0, *pd = dPsidt + r0;
      const double *pa1 = x + IDX(N, (a + 1) % N, b, 0), *pam = x + IDX(N, (a + N - 1) % N, b, 0);
      const double *pb1 = x + IDX(N, a, (b + 1) % N, 0), *pbm = x + IDX(N, a, (b + N - 1) % N, 0);
      double *d0 = dxdt + r0, *d1 = dxdt + N3 + r0;
      for (i64 c = 0; c < N; ++c) d0[c] = pi[c];
      for (i64 c = 0; c < N; ++c)
        d1[c] = ((-3.0 * H + 4.0 * pd[c]) * pi[c] - std::exp(2 * ps[c]) * m * m * p0[c]) +
                std::exp(4 * ps[c]) * inv * ((((-6.0 * p0[c] + pa1[c]) + pam[c]) + pb1[c]) + pbm[c]);
      for (i64 c = 1; c < N - 1; ++c) d1[c] = d1[c] + std::exp(4 * ps[c]) * inv * (p0[c + 1] + p0[c - 1]);
      d1[0] = d1[0] + std::exp(4 * ps[0]) * inv * (p0[N - 1] + p0[1]);
      d1[N - 1] = d1[N - 1] + std::exp(4 * ps[N - 1]) * inv * (p0[N - 2] + p0[0]);
    }
  });
}

// Approximate_Variable_Curvature_FRW_equation_kernel (equations_cuda.cu:299-329):
// Psi, dPsi/dt on a coarse M^3 grid (here unpadded, IDX_OF(M,...)), exponentials
// replaced by their second-order Taylor polynomials.
extern "C" void orc_rhs_approx_curvature(i64 N, i64 M, const double *x, double *dxdt, const double *PsiM,
                                         const double *dPsidtM, double m, double H, double inv) {
  const i64 N3 = N * N * N;
  const i64 sd = N / M, hs = sd / 2;
  parallel_for(N, [&](i64 a) {
    for (i64 b = 0; b < N; ++b)
      for (i64 c = 0; c < N; ++c) {
        const i64 i = IDX(N, a, b, c);
        const i64 j = IDX(M, ((a + hs) % N) / sd, ((b + hs) % N) / sd, ((c + hs) % N) / sd);
        const double P = PsiM[j];
        dxdt[i] = x[N3 + i];
        dxdt[N3 + i] = ((-3.0 * H + 4.0 * dPsidtM[j]) * x[N3 + i] - (1 + 2.0 * P + 2.0 * P * P) * m * m * x[i]) +
                       (1 + 4.0 * P + 8.0 * P * P) * inv * stencil7(N, x, a, b, c);
      }
  });
}

// compute_cutoff_fouriers (fdm3d.cpp:105-130): half spectrum of N^3 -> of M^3.
extern "C" void orc_cutoff_fouriers(i64 N, i64 M, const double *fft, double *out) {
  const double scale_diff = static_cast<double>(N) / static_cast<double>(M);
  for (i64 a = 0; a < M; ++a)
    for (i64 b = 0; b < M; ++b)
      for (i64 c = 0; c <= M / 2; ++c) {
        const i64 aN = (a < M / 2) ? a : (N - (M - a));
        const i64 bN = (b < M / 2) ? b : (N - (M - b));
        const i64 iM = M * (M / 2 + 1) * a + (M / 2 + 1) * b + c;
        const i64 iN = N * (N / 2 + 1) * aN + (N / 2 + 1) * bN + c;
        out[2 * iM] = fft[2 * iN] / pow(scale_diff, 3);
        out[2 * iM + 1] = fft[2 * iN + 1] / pow(scale_diff, 3);
      }
}

// Curvature_energy_density_kernel (equations_cuda.cu:452-473), the fixed-curvature
// energy density (CudaFixedCurvatureEquationInFRW::compute_energy_density, :660-675);
// no CPU counterpart.
extern "C" void orc_curvature_energy_density_cuda(i64 N, const double *x, const double *Psi, double *rho,
                                                  double m, double inv) {
  const i64 N3 = N * N * N;
  parallel_for(N, [&](i64 a) {
    for (i64 b = 0; b < N; ++b)
      for (i64 c = 0; c < N; ++c) {
        const i64 i = IDX(N, a, b, c);
        const double f = x[i];
        const double dx = x[IDX(N, (a + 1) % N, b, c)] - x[IDX(N, (a + N - 1) % N, b, c)];
        const double dy = x[IDX(N, a, (b + 1) % N, c)] - x[IDX(N, a, (b + N - 1) % N, c)];
        const double dz = x[IDX(N, a, b, (c + 1) % N)] - x[IDX(N, a, b, (c + N - 1) % N)];
        rho[i] = std::exp(-2 * Psi[i]) * 0.5 * x[N3 + i] * x[N3 + i] +
                 std::exp(2 * Psi[i]) * (0.5 / 4.0) * inv * (dx * dx + dy * dy + dz * dz) + 0.5 * m * m * f * f;
      }
  });
}

// ---------------------------------------------------------------------------
// WKB extrapolation: WKBSolutionForKleinGordonEquationInFRW (wkb.cpp:5-121),
// radiation era (p = 1, t1 = 1/(2 H1)).
// ---------------------------------------------------------------------------
// Per s = |n|^2: phase(s) = integral of omega_eff from t_i to t_f in closed form
// (wkb.cpp:40-79), omega_eff_i = sqrt(m^2 + k^2/(2 a1^2 H1 t_i) + 3/(16 t_i^2))
// (wkb.cpp:96); out: cos(phase), sin(phase), sqrt(omega_eff_i), phase.
extern "C" void orc_wkb_tables(i64 N, double L, double m, double a1, double H1, double t_i, double t_f,
                               double *cos_t, double *sin_t, double *sqw_t, double *phase_t) {
  const i64 nb = 3 * (N / 2) * (N / 2) + 1;
  const double pi = 3.141592653589793;
  for (i64 s = 0; s < nb; ++s) {
    const double k_sqr = s * std::pow(2 * pi / L, 2);
    auto Q = [&](double t) {
      return std::sqrt(H1 * (8 * k_sqr * t + std::pow(a1, 2) * H1 * (3 + 16 * std::pow(m, 2) * std::pow(t, 2))));
    };
    auto B = [&](double t) { return (3 * std::pow(a1, 2) * H1 + 4 * k_sqr * t) / (std::sqrt(3) * a1 * Q(t)); };
    auto D = [&](double t) { return k_sqr + a1 * m * (4 * a1 * H1 * m * t + Q(t)); };
    const double phase = Q(t_f) / (4. * a1 * H1) - Q(t_i) / (4. * a1 * H1) +
                         (std::sqrt(3) * std::log((1 - B(t_f)) / (1 - B(t_i)))) / 8. -
                         (std::sqrt(3) * std::log((1 + B(t_f)) / (1 + B(t_i)))) / 8. +
                         (k_sqr * std::log(D(t_f) / D(t_i))) / (4. * std::pow(a1, 2) * H1 * m);
    const double omega_eff_i = std::sqrt(m * m + k_sqr / (2 * a1 * a1 * H1 * t_i) + 3 / (16 * t_i * t_i));
    // wkb.cpp:100-101 computes cos(phase) and sin(phase); g++ -O3 merges the pair
    // into one sincos call, so the restatement calls sincos
    double sv, cv;
    ::sincos(phase, &sv, &cv);
    cos_t[s] = cv;
    sin_t[s] = sv;
    sqw_t[s] = std::sqrt(omega_eff_i);
    if (phase_t) phase_t[s] = phase;
  }
}

// Constructor (wkb.cpp:5-20): the two half spectra of the state, scaled by a^1.5,
// the second combined with 1.5 H times the (scaled) first.
extern "C" void orc_wkb_init(i64 N, const double *state, double a_t, double H_t, double *ffts) {
  const i64 fs = 2 * N * N * (N / 2 + 1);
  orc_fft_r2c(N, state, ffts);
  orc_fft_r2c(N, state + N * N * N, ffts + fs);
  const double a15 = pow(a_t, 1.5);
  for (i64 i = 0; i < fs; ++i) ffts[i] *= a15;
  for (i64 i = 0; i < fs; ++i) ffts[fs + i] = a15 * ffts[fs + i] + 1.5 * H_t * ffts[i];
}

// evaluate_at (wkb.cpp:24-121) given the tables of orc_wkb_tables and a(t), H(t).
extern "C" void orc_wkb_evaluate(i64 N, const double *ffts, const double *cos_t, const double *sin_t,
                                 const double *sqw_t, double a_t, double H_t, double *out) {
  const i64 H = N / 2 + 1, fs = 2 * N * N * H, N3 = N * N * N;
  std::vector<double> np((size_t)fs), nd((size_t)fs);
  for (i64 a = 0; a < N; ++a)
    for (i64 b = 0; b < N; ++b)
      for (i64 c = 0; c <= N / 2; ++c) {
        const i64 as = (a <= N / 2) ? a : (N - a), bs = (b <= N / 2) ? b : (N - b);
        const i64 s = as * as + bs * bs + c * c;
        const i64 o = 2 * (N * H * a + H * b + c);
        const double cv = cos_t[s], sv = sin_t[s], sw = sqw_t[s];
        for (int q = 0; q < 2; ++q) {
          np[o + q] = ffts[o + q] * cv + ffts[fs + o + q] * sv / sw;
          nd[o + q] = ffts[fs + o + q] * cv - ffts[o + q] * sv * sw;
        }
      }
  std::vector<double> zp((size_t)N3), zd((size_t)N3);
  orc_fft_c2r(N, np.data(), zp.data());
  orc_fft_c2r(N, nd.data(), zd.data());
  const double a15 = pow(a_t, 1.5);
  for (i64 i = 0; i < N3; ++i) {
    const double p = zp[i] / (N * N * N), d = zd[i] / (N * N * N);
    out[i] = p / a15;
    out[N3 + i] = d / a15 - 1.5 * H_t * p / a15;
  }
}
