// cf_wkb.hip -- WKB extrapolation of a Klein-Gordon field in a radiation-era
// FRW background (WKBSolutionForKleinGordonEquationInFRW, wkb.cpp:5-121): from
// the state at t_i, every Fourier mode is advanced analytically to any later t
// by its closed-form phase.
//
//   create:   phi^, pi^ = R2C(phi), R2C(pi);  phi^ *= a^1.5;  pi^ = a^1.5 pi^ + 1.5 H phi^
//   evaluate: per mode (s = |n|^2), with c = cos(phase), s_ = sin(phase), w = sqrt(omega_eff(t_i)):
//               phi^' = phi^ c + pi^ s_ / w,   pi^' = pi^ c - phi^ s_ w
//             phi = C2R(phi^')/N^3 / a^1.5,  pi = C2R(pi^')/N^3 / a^1.5 - 1.5 H (C2R(phi^')/N^3) / a^1.5
//
// The phase and omega depend on s only, so they are tabulated on the host (in
// the reference's operation order) and the device does two element-wise passes
// and two C2R transforms; the mode loop that the reference runs on one CPU core
// becomes one kernel.
#include <hipfft/hipfft.h>

#include <cmath>
#include <vector>

#include "cf_common.hpp"

namespace cf {

// Per-entry passes over the half spectrum (N*N*(N/2+1) complex values).
__global__ void k_wkb_init(double2 *__restrict__ ph, double2 *__restrict__ pd, int64_t n, double a15, double c15h) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    double2 p = ph[e], d = pd[e];
    p.x = p.x * a15;
    p.y = p.y * a15;
    d.x = a15 * d.x + c15h * p.x;
    d.y = a15 * d.y + c15h * p.y;
    ph[e] = p;
    pd[e] = d;
  }
}

__global__ void k_wkb_rotate(const double2 *__restrict__ ph, const double2 *__restrict__ pd,
                             double2 *__restrict__ oph, double2 *__restrict__ opd, const double *__restrict__ cs,
                             const double *__restrict__ sn, const double *__restrict__ sw, int64_t N) {
  const int64_t H = N / 2 + 1, n = N * N * H;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = e % H, b = (e / H) % N, a = e / (H * N);
    const int64_t as = (a <= N / 2) ? a : N - a, bs = (b <= N / 2) ? b : N - b;
    const int64_t s = as * as + bs * bs + c * c;
    const double cv = cs[s], sv = sn[s], w = sw[s];
    const double2 p = ph[e], d = pd[e];
    double2 q, r;
    q.x = p.x * cv + d.x * sv / w;
    q.y = p.y * cv + d.y * sv / w;
    r.x = d.x * cv - p.x * sv * w;
    r.y = d.y * cv - p.y * sv * w;
    oph[e] = q;
    opd[e] = r;
  }
}

__global__ void k_wkb_finish(double *__restrict__ x, int64_t N3, double n3, double a15, double c15h) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N3; i += (int64_t)gridDim.x * blockDim.x) {
    const double p = x[i] / n3, d = x[N3 + i] / n3;
    x[i] = p / a15;
    x[N3 + i] = d / a15 - c15h * p / a15;
  }
}

static unsigned wkb_grid(int64_t n) {
  int64_t g = cdiv(n, 256);
  return (unsigned)(g > 65536 ? 65536 : (g < 1 ? 1 : g));
}

}  // namespace cf

using namespace cf;

struct cf_wkb_s {
  int64_t N = 0;
  double L = 0, m = 0, t_i = 0;
  cf_cosmology cosmo{};
  double *ffts = nullptr;   // 2 half spectra: scaled phi^ | pi^
  double *rot = nullptr;    // 2 half spectra: rotated (destroyed by the C2R)
  double *tables = nullptr; // 3 tables over s: cos, sin, sqrt(omega_i)
  hipfftHandle d2z = 0, z2d = 0;
  bool have_d2z = false, have_z2d = false;
  std::vector<double> host_tables;
};

extern "C" {

cf_status cf_wkb_tables_host(int64_t N, double L, double m, double a1, double H1, double t_i, double t_f,
                             double *cos_t, double *sin_t, double *sqw_t, double *phase_t) {
  CF_REQUIRE(valid_N(N) && L > 0 && m > 0 && a1 > 0 && H1 > 0, "cf_wkb_tables_host: bad arguments");
  CF_REQUIRE(cos_t && sin_t && sqw_t, "cf_wkb_tables_host: null table");
  const int64_t nb = 3 * (N / 2) * (N / 2) + 1;
  const double pi = 3.14159265358979323846;
  using std::log;
  using std::pow;
  using std::sqrt;
  for (int64_t s = 0; s < nb; ++s) {
    // compute_phase_integral (wkb.cpp:40-79), term by term in the reference's order
    const double k_sqr = s * pow(2 * pi / L, 2);
    auto Q = [&](double t) { return sqrt(H1 * (8 * k_sqr * t + pow(a1, 2) * H1 * (3 + 16 * pow(m, 2) * pow(t, 2)))); };
    auto B = [&](double t) { return (3 * pow(a1, 2) * H1 + 4 * k_sqr * t) / (sqrt(3) * a1 * Q(t)); };
    auto D = [&](double t) { return k_sqr + a1 * m * (4 * a1 * H1 * m * t + Q(t)); };
    const double phase = Q(t_f) / (4. * a1 * H1) - Q(t_i) / (4. * a1 * H1) +
                         (sqrt(3) * log((1 - B(t_f)) / (1 - B(t_i)))) / 8. -
                         (sqrt(3) * log((1 + B(t_f)) / (1 + B(t_i)))) / 8. +
                         (k_sqr * log(D(t_f) / D(t_i))) / (4. * pow(a1, 2) * H1 * m);
    // omega_eff at t_i (wkb.cpp:96)
    const double omega_eff_i = sqrt(m * m + k_sqr / (2 * a1 * a1 * H1 * t_i) + 3 / (16 * t_i * t_i));
    // cos and sin of one argument: g++ -O3 (the reference's compiler) turns the
    // pair into one sincos call, whose last bit can differ from separate calls
    double sv, cv;
    ::sincos(phase, &sv, &cv);
    cos_t[s] = cv;
    sin_t[s] = sv;
    sqw_t[s] = sqrt(omega_eff_i);
    if (phase_t) phase_t[s] = phase;
  }
  return CF_OK;
}

cf_status cf_wkb_create(const cf_kg_params *p, const double *state, double t_i, cf_wkb *out, cf_stream stream) {
  CF_REQUIRE(out, "cf_wkb_create: null output handle");
  *out = nullptr;
  CF_REQUIRE(p && state, "cf_wkb_create: null argument");
  CF_REQUIRE(valid_N(p->N) && p->N <= 1024, "cf_wkb_create: unsupported N=%lld (even, 4..1024)", (long long)p->N);
  CF_REQUIRE(p->m > 0 && p->L > 0, "cf_wkb_create: need m > 0 and L > 0");
  // wkb.cpp:7-8: radiation era only
  CF_REQUIRE(p->cosmo.p == 1.0 && p->cosmo.H1 > 0 && p->cosmo.t1 == 1.0 / (2 * p->cosmo.H1),
             "cf_wkb_create: the WKB solution needs a radiation-era cosmology with t1 = 1/(2 H1)");
  const int64_t N = p->N, N3 = N * N * N, half = N * N * (N / 2 + 1);
  cf_wkb_s *w = new cf_wkb_s();
  w->N = N;
  w->L = p->L;
  w->m = p->m;
  w->t_i = t_i;
  w->cosmo = p->cosmo;
  hipStream_t st = S(stream);
  cf_status rc = CF_OK;
  const int n[3] = {(int)N, (int)N, (int)N};
  const int64_t nb = 3 * (N / 2) * (N / 2) + 1;
  if (hipMalloc(&w->ffts, 4 * half * sizeof(double)) != hipSuccess ||
      hipMalloc(&w->rot, 4 * half * sizeof(double)) != hipSuccess ||
      hipMalloc(&w->tables, 3 * nb * sizeof(double)) != hipSuccess) {
    set_error("cf_wkb_create: device allocation failed");
    rc = CF_E_NOMEM;
  }
  // execute_batched_d2z (fftw_wrapper.cpp:70-76): both halves of the state
  if (rc == CF_OK && hipfftPlanMany(&w->d2z, 3, const_cast<int *>(n), nullptr, 1, (int)N3, nullptr, 1, (int)half,
                                    HIPFFT_D2Z, 2) == HIPFFT_SUCCESS)
    w->have_d2z = true;
  else if (rc == CF_OK) {
    set_error("cf_wkb_create: hipfftPlanMany(D2Z) failed");
    rc = CF_E_FFT;
  }
  if (rc == CF_OK && hipfftPlanMany(&w->z2d, 3, const_cast<int *>(n), nullptr, 1, (int)half, nullptr, 1, (int)N3,
                                    HIPFFT_Z2D, 2) == HIPFFT_SUCCESS)
    w->have_z2d = true;
  else if (rc == CF_OK) {
    set_error("cf_wkb_create: hipfftPlanMany(Z2D) failed");
    rc = CF_E_FFT;
  }
  if (rc == CF_OK) {
    // the transform reads a copy of the state (the caller's state is preserved)
    double *tmp = w->rot;  // 4*half >= 2*N3 doubles
    if (hipMemcpyAsync(tmp, state, 2 * N3 * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess ||
        hipfftSetStream(w->d2z, st) != HIPFFT_SUCCESS ||
        hipfftExecD2Z(w->d2z, tmp, (hipfftDoubleComplex *)w->ffts) != HIPFFT_SUCCESS) {
      set_error("cf_wkb_create: forward transform failed");
      rc = CF_E_FFT;
    }
  }
  if (rc == CF_OK) {
    const double a15 = std::pow(cf_cosmo_a(&p->cosmo, t_i), 1.5), c15h = 1.5 * cf_cosmo_H(&p->cosmo, t_i);
    hipLaunchKernelGGL(k_wkb_init, dim3(wkb_grid(half)), dim3(256), 0, st, (double2 *)w->ffts,
                       (double2 *)w->ffts + half, half, a15, c15h);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
      set_error("cf_wkb_create: scaling kernel failed");
      rc = CF_E_HIP;
    }
  }
  if (rc != CF_OK) {
    cf_wkb_destroy(w);
    return rc;
  }
  *out = w;
  return CF_OK;
}

void cf_wkb_destroy(cf_wkb w) {
  if (!w) return;
  (void)hipDeviceSynchronize();
  if (w->have_d2z) hipfftDestroy(w->d2z);
  if (w->have_z2d) hipfftDestroy(w->z2d);
  (void)hipFree(w->ffts);
  (void)hipFree(w->rot);
  (void)hipFree(w->tables);
  delete w;
}

cf_status cf_wkb_evaluate(cf_wkb w, double t, double *state_out, cf_stream stream) {
  CF_REQUIRE(w && state_out, "cf_wkb_evaluate: null argument");
  const int64_t N = w->N, N3 = N * N * N, half = N * N * (N / 2 + 1), nb = 3 * (N / 2) * (N / 2) + 1;
  hipStream_t st = S(stream);
  w->host_tables.resize(3 * nb);
  double *h = w->host_tables.data();
  CF_TRY(cf_wkb_tables_host(N, w->L, w->m, w->cosmo.a1, w->cosmo.H1, w->t_i, t, h, h + nb, h + 2 * nb, nullptr));
  // blocking upload: the host table buffer is reused by the next call
  CF_HIP(hipMemcpyAsync(w->tables, h, 3 * nb * sizeof(double), hipMemcpyHostToDevice, st));
  CF_HIP(hipStreamSynchronize(st));
  hipLaunchKernelGGL(k_wkb_rotate, dim3(wkb_grid(half)), dim3(256), 0, st, (const double2 *)w->ffts,
                     (const double2 *)w->ffts + half, (double2 *)w->rot, (double2 *)w->rot + half, w->tables,
                     w->tables + nb, w->tables + 2 * nb, N);
  CF_LAUNCHED();
  if (hipfftSetStream(w->z2d, st) != HIPFFT_SUCCESS ||
      hipfftExecZ2D(w->z2d, (hipfftDoubleComplex *)w->rot, state_out) != HIPFFT_SUCCESS) {
    set_error("cf_wkb_evaluate: inverse transform failed");
    return CF_E_FFT;
  }
  const double a15 = std::pow(cf_cosmo_a(&w->cosmo, t), 1.5), c15h = 1.5 * cf_cosmo_H(&w->cosmo, t);
  hipLaunchKernelGGL(k_wkb_finish, dim3(wkb_grid(N3)), dim3(256), 0, st, state_out, N3, (double)(N * N * N), a15,
                     c15h);
  CF_LAUNCHED();
  return CF_OK;
}

}  // extern "C"
