// cf_spectrum.hip -- hipFFT wrapper, deterministic spectrum binning, Gaussian
// random fields and the fixed-order rms rescale.
//
// Binning without atomics. The half spectrum has N*N*(N/2+1) entries; entry
// (a,b,c) belongs to bin s = a'^2 + b'^2 + c^2 with x' = min(x, N-x), and counts
// twice unless c is 0 or N/2 (the full-cube loop of fdm3d.cpp:11-28 visits the
// mirrored column too). A row (a, b) of the half spectrum has the row class
// r = a'^2 + b'^2, so s = r + c^2. At plan creation the host sorts the N^2 rows by
// r (4 B per row). Step A sums, for each class r, its whole rows (each read
// contiguously, in sorted order) into Q[c][r]; step B forms P(s) as the sum over
// ascending c of Q[c][s - c^2]. Every sum has a fixed order and no atomics are
// used, so two runs give identical bits.
#include <hipfft/hipfft.h>

#include <cmath>
#include <cstring>
#include <random>
#include <type_traits>
#include <vector>

#include "cf_common.hpp"

struct cf_spectrum_plan {
  int64_t N = 0, H = 0, nb = 0, nhalf = 0;
  hipfftHandle d2z = 0, z2d = 0;
  bool have_d2z = false, have_z2d = false;
  void *work = nullptr;
  double *rbuf = nullptr;        // N^3 real scratch
  double *cbuf0 = nullptr;       // half-spectrum scratch (2*nhalf doubles)
  double *cbuf1 = nullptr;       // second half-spectrum scratch
  int64_t nr = 0;                // row classes r = as^2 + bs^2 <= 2 (N/2)^2
  int32_t *roff = nullptr;       // nr+1 offsets into rows
  int32_t *rows = nullptr;       // the N^2 rows (a*N + b) of the half spectrum sorted by r
  double *red = nullptr;         // reduction scratch (RED_BLOCKS + 1 doubles)
  double *bins = nullptr;        // one spectrum (nb doubles) for callers inside the library
};

namespace cf {

constexpr int RED_BLOCKS = 1024;

#define CF_FFT(expr)                                                        \
  do {                                                                      \
    hipfftResult r_ = (expr);                                               \
    if (r_ != HIPFFT_SUCCESS) {                                             \
      ::cf::set_error("%s failed: hipfft status %d (%s:%d)", #expr, (int)r_, \
                      __FILE__, __LINE__);                                  \
      return CF_E_FFT;                                                      \
    }                                                                       \
  } while (0)

// Binning (fdm3d.cpp:11-28 sums |f_k|^2 over the full k-cube by s = |n|^2; the half
// spectrum counts interior c columns twice). With r = as^2 + bs^2 of a row (a, b),
// s = r + c^2, so
//   step A: Q[c][r] = sum over the rows of class r (in sorted order) of w(c)|f|^2 --
//           whole c-rows are read, coalesced; RB consecutive r of one block are
//           summed into an LDS tile and written out transposed;
//   step B: P(s) = sum_{c ascending} Q[c][s - c^2], coalesced over consecutive s.
// Fixed summation orders, no atomics: bit-reproducible.
constexpr int BIN_LDS = 4096;  // doubles of the step-A tile (32 KB: several blocks per CU)
constexpr int BIN_CQ = 9;      // 64-column chunks of a row: H = N/2 + 1 <= 513

// CQ: 64-column chunks per row (CQ * 64 >= H). Each wavefront keeps bin_group rows
// of its class in flight (their indices read together, then every column load of
// the group issued before the group is summed, in row order: the same bits as a
// row-by-row sum); the group holds at most ~96 VGPRs of loaded values.
constexpr int bin_group(bool mode, int cq) {
  const int g = 24 / (cq * (mode ? 2 : 1));
  return g < 1 ? 1 : (g > 4 ? 4 : g);
}

template <bool MODE, int CQ>
__global__ __launch_bounds__(256) void k_bin_rows(const double2 *__restrict__ fk, const double2 *__restrict__ dk,
                                                  const int32_t *__restrict__ roff, const int32_t *__restrict__ rows,
                                                  int64_t H, int64_t nr, int lrb, double m2, double kir2, double a2,
                                                  double *__restrict__ Q) {
  __shared__ double tile[BIN_LDS];
  const int RB = 1 << lrb;
  const int64_t r0 = (int64_t)blockIdx.x * RB;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int rl = wave; rl < RB; rl += 4) {  // each wavefront its own row classes
    const int64_t r = r0 + rl;
    if (r >= nr) break;
    const int32_t beg = roff[r], end = roff[r + 1];
    // lane holds columns c = lane + 64 q
    double acc[CQ];
    double denom[CQ];
#pragma unroll
    for (int q = 0; q < CQ; ++q) {
      acc[q] = 0.0;
      const int64_t c = lane + 64 * q;
      // fdm3d.cpp:64: m*m + (2*pi/L)*(2*pi/L)*s/(a_t*a_t)
      denom[q] = MODE ? m2 + (kir2 * (double)(r + c * c)) / a2 : 1.0;
    }
    auto add = [&](int q, double2 u, double2 w) {
      const int64_t c = lane + 64 * q;
      double p = u.x * u.x + u.y * u.y;
      if constexpr (MODE) p = p + (w.x * w.x + w.y * w.y) / denom[q];
      acc[q] += (c == 0 || c == H - 1) ? p : 2.0 * p;
    };
    const double2 zero = make_double2(0.0, 0.0);
    constexpr int BIN_G = bin_group(MODE, CQ);
    int32_t j = beg;
    for (; j + BIN_G <= end; j += BIN_G) {
      int64_t base[BIN_G];
#pragma unroll
      for (int g = 0; g < BIN_G; ++g) base[g] = (int64_t)rows[j + g] * H;
      double2 u[BIN_G][CQ], w[BIN_G][CQ];
#pragma unroll
      for (int g = 0; g < BIN_G; ++g)
#pragma unroll
        for (int q = 0; q < CQ; ++q) {
          const int64_t c = lane + 64 * q;
          u[g][q] = c < H ? fk[base[g] + c] : zero;
          if constexpr (MODE) w[g][q] = c < H ? dk[base[g] + c] : zero;
          else w[g][q] = zero;
        }
#pragma unroll
      for (int g = 0; g < BIN_G; ++g)
#pragma unroll
        for (int q = 0; q < CQ; ++q)
          if (lane + 64 * q < H) add(q, u[g][q], w[g][q]);
    }
    for (; j < end; ++j) {
      const int64_t base = (int64_t)rows[j] * H;
#pragma unroll
      for (int q = 0; q < CQ; ++q) {
        const int64_t c = lane + 64 * q;
        if (c < H) add(q, fk[base + c], MODE ? dk[base + c] : zero);
      }
    }
#pragma unroll
    for (int q = 0; q < CQ; ++q) {
      const int64_t c = lane + 64 * q;
      if (c < H) tile[rl * H + c] = acc[q];
    }
  }
  __syncthreads();
  const int64_t nrb = (r0 + RB <= nr) ? RB : nr - r0;
  for (int64_t q = threadIdx.x; q < (int64_t)RB * H; q += blockDim.x) {
    const int64_t c = q >> lrb, rl = q & (RB - 1);
    if (rl < nrb) Q[c * nr + r0 + rl] = tile[rl * H + c];
  }
}

__global__ __launch_bounds__(256) void k_bin_cols(const double *__restrict__ Q, int64_t H, int64_t nr, int64_t nb,
                                                  double *__restrict__ bins) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nb) return;
  double acc = 0.0;
  for (int64_t c = 0; c < H && c * c <= s; ++c) {
    const int64_t r = s - c * c;
    if (r < nr) acc += Q[c * nr + r];
  }
  bins[s] = acc;
}

// radial[j] = sum_{s : isqrt(s) = j} sbins[s], j < N/2, summed in s order.
__global__ void k_radial(const double *__restrict__ sb, int64_t nb, int64_t nr, double *__restrict__ rad) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nr) return;
  int64_t lo = j * j, hi = (j + 1) * (j + 1);
  if (hi > nb) hi = nb;
  double acc = 0.0;
  for (int64_t s = lo; s < hi; ++s) acc += sb[s];
  rad[j] = acc;
}

// Real field (a, b, c) at N^2 a + N b + c -> the FFTW padded layout of an in-place
// real transform, rows of 2 (N/2 + 1) doubles (PADDED_IDX_OF, fdm3d.hpp:24). The
// padding entries are left as they are.
__global__ void k_pad_rows(const double *__restrict__ f, double *__restrict__ out, int64_t N, int64_t total) {
  const int64_t W = 2 * (N / 2 + 1);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = i / N, c = i - row * N;
    out[row * W + c] = f[i];
  }
}

// Multiply the half spectrum by sqrt(P) per bin (random_field.cpp:60-73).
__global__ void k_grf_scale(double2 *__restrict__ fk, const double *__restrict__ sqrtP, int64_t N,
                            int64_t H) {
  const int64_t total = N * N * H;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = e % H;
    const int64_t r = e / H;
    const int64_t b = r % N, a = r / N;
    const int64_t as = (a <= N / 2) ? a : a - N;
    const int64_t bs = (b <= N / 2) ? b : b - N;
    const int64_t s = as * as + bs * bs + c * c;
    const double sp = sqrtP[s];
    double2 v = fk[e];
    v.x *= sp;
    v.y *= sp;
    fk[e] = v;
  }
}

// Multiply the half spectrum by kernel(s)/N^3 (compute_field_with_scaled_fourier_modes,
// fdm3d.cpp:84-97; the kernel is tabulated over s = |n|^2).
__global__ void k_scale_modes(double2 *__restrict__ fk, const double *__restrict__ kern, int64_t N, int64_t H) {
  const int64_t total = N * N * H;
  const double n3 = (double)(N * N * N);
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = e % H;
    const int64_t r = e / H;
    const int64_t b = r % N, a = r / N;
    const int64_t as = (a <= N / 2) ? a : N - a;
    const int64_t bs = (b <= N / 2) ? b : N - b;
    const double scale = kern[as * as + bs * bs + c * c] / n3;
    double2 v = fk[e];
    v.x *= scale;
    v.y *= scale;
    fk[e] = v;
  }
}

__global__ void k_mul_exp(double *__restrict__ w, const double *__restrict__ psi, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    w[i] = w[i] * exp(psi[i]);
}

// Fixed-order sum of squares: RED_BLOCKS blocks of 256 threads, grid-stride
// partials, block tree, then one block over the partials. The grid is fixed,
// so the order of additions is the same on every run and every device.
__global__ __launch_bounds__(256) void k_sumsq_partial(const double *__restrict__ x, int64_t n,
                                                       double *__restrict__ part) {
  __shared__ double sm[256];
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)RED_BLOCKS * 256)
    acc += x[i] * x[i];
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

__global__ __launch_bounds__(256) void k_rms_factor(double *__restrict__ part, int64_t n, double target) {
  __shared__ double sm[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < RED_BLOCKS; i += 256) acc += part[i];
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[RED_BLOCKS] = target / sqrt(sm[0] / (double)n);
}

__global__ void k_scale_by(double *__restrict__ x, int64_t n, const double *__restrict__ f) {
  const double s = *f;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] = x[i] * s;
}

static unsigned grid_cap(int64_t n, int block) {
  int64_t g = cdiv(n, block);
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (unsigned)g;
}

static cf_status plan_check(const cf_spectrum_plan *p) {
  CF_REQUIRE(p && p->N > 0, "null or empty spectrum plan");
  return CF_OK;
}

static cf_status d2z_exec(cf_spectrum_plan *p, double *in, double *out, hipStream_t st) {
  CF_FFT(hipfftSetStream(p->d2z, st));
  CF_FFT(hipfftExecD2Z(p->d2z, (hipfftDoubleReal *)in, (hipfftDoubleComplex *)out));
  return CF_OK;
}

static cf_status z2d_exec(cf_spectrum_plan *p, double *in, double *out, hipStream_t st) {
  CF_FFT(hipfftSetStream(p->z2d, st));
  CF_FFT(hipfftExecZ2D(p->z2d, (hipfftDoubleComplex *)in, (hipfftDoubleReal *)out));
  return CF_OK;
}

// Spectrum factories over s, |k| = sqrt(s)*2*pi/L (random_field.cpp:67,212-254).
static void spectrum_table(int kind, int64_t N, double L, const double *prm, double *P) {
  const double pi = 3.14159265358979323846;
  const int64_t nb = 3 * (N / 2) * (N / 2) + 1;
  double P_k_ast = 0.0, k_ast = 0.0, alpha = 0.0, P_k0 = 0.0;
  if (kind == CF_SPEC_POWER_CUTOFF || kind == CF_SPEC_DERIV_POWER_CUTOFF) {
    // power_law_with_cutoff_given_amplitude_3d, random_field.cpp:212-221
    const double sigma = prm[0];
    k_ast = prm[1];
    alpha = prm[2];
    P_k_ast = alpha == -3.0
                  ? 2 * pi * pi * sigma * sigma / (k_ast * k_ast * k_ast * log(k_ast * L / (2 * pi)))
                  : (3 + alpha) * 2 * pi * pi * sigma * sigma * (L * L * L * pow(k_ast * L, alpha)) /
                        (pow(k_ast * L, alpha + 3) - pow(2 * pi, alpha + 3));
    P_k_ast *= (N / L) * (N / L) * (N / L);
  }
  if (kind == CF_SPEC_SCALE_INV) {
    // scale_invariant_spectrum_3d, random_field.cpp:234-240
    const double k0 = 1.0;
    P_k0 = prm[0] / ((k0 * k0 * k0) / (2 * pi * pi));
    P_k0 *= (N / L) * (N / L) * (N / L);
  }
  for (int64_t s = 0; s < nb; ++s) {
    const double k = sqrt((double)s) * 2 * pi / L;
    double v = 0.0;
    switch (kind) {
      case CF_SPEC_GAUSS: v = (k == 0.0) ? 0.0 : exp(-(k * k) / (2 * prm[0] * prm[0])); break;
      case CF_SPEC_KM3_GAUSS: v = (k == 0.0) ? 0.0 : pow(k, -3) * exp(-(k * k) / 2); break;
      case CF_SPEC_SCALE_INV: v = (k == 0.0) ? 0.0 : (P_k0 * pow(k / 1.0, -3)); break;
      case CF_SPEC_POWER_CUTOFF:
        v = (k == 0.0) ? 0.0 : (P_k_ast * (k <= k_ast ? pow(k / k_ast, alpha) : 0.0));
        break;
      case CF_SPEC_DERIV_POWER_CUTOFF: {
        // to_deriv_spectrum, random_field.cpp:248-250
        const double m = prm[3];
        const double Pf = (k == 0.0) ? 0.0 : (P_k_ast * (k <= k_ast ? pow(k / k_ast, alpha) : 0.0));
        v = (k * k + m * m) * Pf;
        break;
      }
      default: v = 0.0;
    }
    P[s] = v;
  }
}

}  // namespace cf

using namespace cf;

extern "C" {

int64_t cf_spectrum_bins(int64_t N) { return 3 * (N / 2) * (N / 2) + 1; }

cf_status cf_spectrum_plan_destroy(cf_spectrum_plan *p) {
  if (!p) return CF_OK;
  if (p->have_d2z) hipfftDestroy(p->d2z);
  if (p->have_z2d) hipfftDestroy(p->z2d);
  void *bufs[] = {p->work, p->rbuf, p->cbuf0, p->cbuf1, p->roff, p->rows, p->red, p->bins};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  delete p;
  return CF_OK;
}

cf_status cf_spectrum_plan_create(int64_t N, cf_spectrum_plan **out) {
  CF_REQUIRE(out, "cf_spectrum_plan_create: null output");
  *out = nullptr;
  CF_REQUIRE(valid_N(N) && N <= 1024, "cf_spectrum_plan_create: unsupported N=%lld", (long long)N);
  cf_spectrum_plan *p = new cf_spectrum_plan();
  p->N = N;
  p->H = N / 2 + 1;
  p->nb = cf_spectrum_bins(N);
  p->nhalf = N * N * p->H;
  p->nr = 2 * (N / 2) * (N / 2) + 1;
  auto fail = [&](cf_status s) {
    cf_spectrum_plan_destroy(p);
    return s;
  };
  // hipFFT plans sharing one work area (cuda_wrapper.cu:202-222)
  size_t ws1 = 0, ws2 = 0;
  if (hipfftCreate(&p->d2z) != HIPFFT_SUCCESS) { set_error("hipfftCreate failed"); return fail(CF_E_FFT); }
  p->have_d2z = true;
  if (hipfftCreate(&p->z2d) != HIPFFT_SUCCESS) { set_error("hipfftCreate failed"); return fail(CF_E_FFT); }
  p->have_z2d = true;
  hipfftSetAutoAllocation(p->d2z, 0);
  hipfftSetAutoAllocation(p->z2d, 0);
  if (hipfftMakePlan3d(p->d2z, (int)N, (int)N, (int)N, HIPFFT_D2Z, &ws1) != HIPFFT_SUCCESS ||
      hipfftMakePlan3d(p->z2d, (int)N, (int)N, (int)N, HIPFFT_Z2D, &ws2) != HIPFFT_SUCCESS) {
    set_error("hipfftMakePlan3d(N=%lld) failed", (long long)N);
    return fail(CF_E_FFT);
  }
  const size_t ws = ws1 > ws2 ? ws1 : ws2;
  const size_t n3 = (size_t)N * N * N;
  if ((ws && hipMalloc(&p->work, ws) != hipSuccess) ||
      hipMalloc((void **)&p->rbuf, n3 * sizeof(double)) != hipSuccess ||
      hipMalloc((void **)&p->cbuf0, 2 * (size_t)p->nhalf * sizeof(double)) != hipSuccess ||
      hipMalloc((void **)&p->cbuf1, 2 * (size_t)p->nhalf * sizeof(double)) != hipSuccess ||
      hipMalloc((void **)&p->roff, ((size_t)p->nr + 1) * sizeof(int32_t)) != hipSuccess ||
      hipMalloc((void **)&p->rows, (size_t)(N * N) * sizeof(int32_t)) != hipSuccess ||
      hipMalloc((void **)&p->red, (RED_BLOCKS + 1) * sizeof(double)) != hipSuccess ||
      hipMalloc((void **)&p->bins, (size_t)p->nb * sizeof(double)) != hipSuccess) {
    set_error("cf_spectrum_plan_create(N=%lld): device allocation failed", (long long)N);
    return fail(CF_E_NOMEM);
  }
  if (ws) {
    if (hipfftSetWorkArea(p->d2z, p->work) != HIPFFT_SUCCESS ||
        hipfftSetWorkArea(p->z2d, p->work) != HIPFFT_SUCCESS) {
      set_error("hipfftSetWorkArea failed");
      return fail(CF_E_FFT);
    }
  }
  // Row table: stable counting sort of the N^2 rows (a, b) by r = as^2 + bs^2 (the
  // binning's step A; its output Q, H * nr <= N^3 doubles, lives in rbuf).
  {
    const int64_t nr = p->nr;
    std::vector<int32_t> off((size_t)nr + 1, 0);
    auto fold = [&](int64_t v) { return (v <= N / 2) ? v : N - v; };
    for (int64_t a = 0; a < N; ++a)
      for (int64_t b = 0; b < N; ++b) off[fold(a) * fold(a) + fold(b) * fold(b) + 1]++;
    for (int64_t q = 0; q < nr; ++q) off[q + 1] += off[q];
    std::vector<int32_t> pos(off.begin(), off.end() - 1);
    std::vector<int32_t> rws((size_t)(N * N));
    for (int64_t a = 0; a < N; ++a)
      for (int64_t b = 0; b < N; ++b) rws[pos[fold(a) * fold(a) + fold(b) * fold(b)]++] = (int32_t)(a * N + b);
    if (hipMemcpy(p->roff, off.data(), off.size() * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->rows, rws.data(), rws.size() * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess) {
      set_error("cf_spectrum_plan_create: bin table upload failed");
      return fail(CF_E_HIP);
    }
  }
  *out = p;
  return CF_OK;
}

cf_status cf_fft_d2z(cf_spectrum_plan *p, const double *in, double *out, cf_stream stream) {
  CF_TRY(plan_check(p));
  CF_REQUIRE(in && out, "cf_fft_d2z: null buffer");
  const size_t n3 = (size_t)p->N * p->N * p->N;
  CF_HIP(hipMemcpyAsync(p->rbuf, in, n3 * sizeof(double), hipMemcpyDeviceToDevice, S(stream)));
  return d2z_exec(p, p->rbuf, out, S(stream));
}

cf_status cf_fft_z2d(cf_spectrum_plan *p, double *in, double *out, cf_stream stream) {
  CF_TRY(plan_check(p));
  CF_REQUIRE(in && out, "cf_fft_z2d: null buffer");
  return z2d_exec(p, in, out, S(stream));
}

cf_status cf_fft_z2d_inplace(cf_spectrum_plan *p, double *inout, cf_stream stream) {
  CF_TRY(plan_check(p));
  CF_REQUIRE(inout, "cf_fft_z2d_inplace: null buffer");
  // C2R into the plan's real scratch (the half spectrum in inout may be overwritten,
  // as an in-place transform would), then spread the rows into the padded layout.
  CF_TRY(z2d_exec(p, inout, p->rbuf, S(stream)));
  const int64_t N = p->N, total = N * N * N;
  hipLaunchKernelGGL(k_pad_rows, dim3(grid_cap(total, 256)), dim3(256), 0, S(stream), p->rbuf, inout, N, total);
  CF_LAUNCHED();
  return CF_OK;
}

// Step A into the plan's real scratch (free once the transform has run), step B into bins.
static cf_status bin_launch(cf_spectrum_plan *p, const double *fk, const double *dk, double m2, double kir2,
                            double a2, double *bins, hipStream_t st) {
  CF_REQUIRE(p->H <= 64 * BIN_CQ, "binning: N=%lld exceeds the row chunking", (long long)p->N);
  int lrb = 0;  // RB = 2^lrb row classes per block: the largest power of two with RB * H <= BIN_LDS
  while ((2 << lrb) * p->H <= BIN_LDS && lrb < 8) ++lrb;
  const dim3 gA((unsigned)cdiv(p->nr, (int64_t)1 << lrb)), b256(256);
  auto rows_kernel = [&](auto mode, auto cq) {
    hipLaunchKernelGGL((k_bin_rows<decltype(mode)::value, decltype(cq)::value>), gA, b256, 0, st,
                       (const double2 *)fk, (const double2 *)dk, p->roff, p->rows, p->H, p->nr, lrb, m2, kir2, a2,
                       p->rbuf);
  };
  auto by_cq = [&](auto mode) {
    const int64_t cq = cdiv(p->H, 64);  // H <= 64 * BIN_CQ (checked above)
    if (cq <= 1) rows_kernel(mode, std::integral_constant<int, 1>());
    else if (cq <= 2) rows_kernel(mode, std::integral_constant<int, 2>());
    else if (cq <= 3) rows_kernel(mode, std::integral_constant<int, 3>());
    else if (cq <= 4) rows_kernel(mode, std::integral_constant<int, 4>());
    else if (cq <= 5) rows_kernel(mode, std::integral_constant<int, 5>());
    else rows_kernel(mode, std::integral_constant<int, BIN_CQ>());
  };
  if (dk) by_cq(std::true_type());
  else by_cq(std::false_type());
  CF_LAUNCHED();
  hipLaunchKernelGGL(k_bin_cols, dim3((unsigned)cdiv(p->nb, 256)), b256, 0, st, p->rbuf, p->H, p->nr, p->nb, bins);
  CF_LAUNCHED();
  return CF_OK;
}

cf_status cf_bin_power(cf_spectrum_plan *p, const double *fk, double *bins, cf_stream stream) {
  CF_TRY(plan_check(p));
  CF_REQUIRE(fk && bins, "cf_bin_power: null buffer");
  return bin_launch(p, fk, nullptr, 0.0, 0.0, 1.0, bins, S(stream));
}

cf_status cf_bin_mode_power(cf_spectrum_plan *p, const double *phik, const double *dphik, double L,
                            double m, double a_t, double *bins, cf_stream stream) {
  CF_TRY(plan_check(p));
  CF_REQUIRE(phik && dphik && bins, "cf_bin_mode_power: null buffer");
  const double pi = 3.14159265358979323846;
  const double kir2 = (2 * pi / L) * (2 * pi / L);
  return bin_launch(p, phik, dphik, m * m, kir2, a_t * a_t, bins, S(stream));
}

cf_status cf_power_spectrum(cf_spectrum_plan *p, const double *f, double *bins, cf_stream stream) {
  CF_TRY(plan_check(p));
  CF_TRY(cf_fft_d2z(p, f, p->cbuf0, stream));
  return cf_bin_power(p, p->cbuf0, bins, stream);
}

cf_status cf_power_spectrum_consume(cf_spectrum_plan *p, double *f, double *bins, cf_stream stream) {
  CF_TRY(plan_check(p));
  CF_REQUIRE(f && bins, "cf_power_spectrum_consume: null buffer");
  CF_TRY(d2z_exec(p, f, p->cbuf0, S(stream)));
  return cf_bin_power(p, p->cbuf0, bins, stream);
}

cf_status cf_mode_power_spectrum(cf_spectrum_plan *p, const double *state, double L, double m,
                                 double a_t, double *bins, cf_stream stream) {
  CF_TRY(plan_check(p));
  CF_REQUIRE(state && bins, "cf_mode_power_spectrum: null buffer");
  const int64_t N3 = p->N * p->N * p->N;
  CF_TRY(cf_fft_d2z(p, state, p->cbuf0, stream));
  CF_TRY(cf_fft_d2z(p, state + N3, p->cbuf1, stream));
  return cf_bin_mode_power(p, p->cbuf0, p->cbuf1, L, m, a_t, bins, stream);
}

cf_status cf_radial_bins(cf_spectrum_plan *p, const double *sbins, double *radial, cf_stream stream) {
  CF_TRY(plan_check(p));
  CF_REQUIRE(sbins && radial, "cf_radial_bins: null buffer");
  const int64_t nr = p->N / 2;
  hipLaunchKernelGGL(k_radial, dim3((unsigned)cdiv(nr, 64)), dim3(64), 0, S(stream), sbins, p->nb, nr,
                     radial);
  CF_LAUNCHED();
  return CF_OK;
}

cf_status cf_normals_host(uint32_t seed, int64_t n, double *out) {
  CF_REQUIRE(out && n >= 0, "cf_normals_host: bad arguments");
  std::mt19937 mt{seed};
  std::normal_distribution<double> dist{0.0, 1.0};
  for (int64_t i = 0; i < n; ++i) out[i] = dist(mt);
  return CF_OK;
}

cf_status cf_spectrum_table_host(int kind, int64_t N, double L, const double *params, double *P) {
  CF_REQUIRE(P && valid_N(N) && L > 0, "cf_spectrum_table_host: bad arguments");
  CF_REQUIRE(kind >= 0 && kind <= 4, "cf_spectrum_table_host: unknown spectrum kind %d", kind);
  CF_REQUIRE(kind == CF_SPEC_KM3_GAUSS || params, "cf_spectrum_table_host: missing parameters");
  double zero[4] = {0, 0, 0, 0};
  spectrum_table(kind, N, L, params ? params : zero, P);
  return CF_OK;
}

cf_status cf_grf(cf_spectrum_plan *p, double *w, const double *sqrtP, const double *Psi, double *out,
                 cf_stream stream) {
  CF_TRY(plan_check(p));
  CF_REQUIRE(w && sqrtP && out, "cf_grf: null buffer");
  const int64_t N3 = p->N * p->N * p->N;
  hipStream_t st = S(stream);
  if (Psi) {
    hipLaunchKernelGGL(k_mul_exp, dim3(grid_cap(N3, 256)), dim3(256), 0, st, w, Psi, N3);
    CF_LAUNCHED();
  }
  CF_TRY(d2z_exec(p, w, p->cbuf0, st));
  hipLaunchKernelGGL(k_grf_scale, dim3(grid_cap(p->nhalf, 256)), dim3(256), 0, st, (double2 *)p->cbuf0,
                     sqrtP, p->N, p->H);
  CF_LAUNCHED();
  return z2d_exec(p, p->cbuf0, out, st);
}

cf_status cf_scaled_fourier_modes(cf_spectrum_plan *p, const double *f, const double *kernel_by_s, double *out,
                                  cf_stream stream) {
  CF_TRY(plan_check(p));
  CF_REQUIRE(f && kernel_by_s && out, "cf_scaled_fourier_modes: null buffer");
  hipStream_t st = S(stream);
  const size_t n3 = (size_t)p->N * p->N * p->N;
  CF_HIP(hipMemcpyAsync(p->rbuf, f, n3 * sizeof(double), hipMemcpyDeviceToDevice, st));
  CF_TRY(d2z_exec(p, p->rbuf, p->cbuf0, st));
  hipLaunchKernelGGL(k_scale_modes, dim3(grid_cap(p->nhalf, 256)), dim3(256), 0, st, (double2 *)p->cbuf0,
                     kernel_by_s, p->N, p->H);
  CF_LAUNCHED();
  return z2d_exec(p, p->cbuf0, out, st);
}

}  // extern "C"

namespace cf {

// x *= target / rms(x) with a caller-provided reduction scratch (RED_BLOCKS + 1 doubles).
static cf_status rescale_rms_with(double *x, int64_t n, double target, double *red, hipStream_t st) {
  hipLaunchKernelGGL(k_sumsq_partial, dim3(RED_BLOCKS), dim3(256), 0, st, x, n, red);
  hipLaunchKernelGGL(k_rms_factor, dim3(1), dim3(256), 0, st, red, n, target);
  hipLaunchKernelGGL(k_scale_by, dim3(grid_cap(n, 256)), dim3(256), 0, st, x, n, red + RED_BLOCKS);
  CF_LAUNCHED();
  return CF_OK;
}

double *plan_red(cf_spectrum_plan *p) { return p->red; }
double *plan_bins(cf_spectrum_plan *p) { return p->bins; }
int64_t plan_size(const cf_spectrum_plan *p) { return p->N; }

}  // namespace cf

extern "C" {

cf_status cf_rescale_rms(double *x, int64_t n, double target, cf_stream stream) {
  CF_REQUIRE(x && n > 0, "cf_rescale_rms: bad arguments");
  // No plan to borrow from: a stream-ordered scratch of RED_BLOCKS + 1 doubles.
  double *red = nullptr;
  CF_HIP(hipMallocAsync((void **)&red, (RED_BLOCKS + 1) * sizeof(double), S(stream)));
  const cf_status rc = rescale_rms_with(x, n, target, red, S(stream));
  (void)hipFreeAsync(red, S(stream));
  return rc;
}

cf_status cf_init_grf_state(cf_spectrum_plan *p, uint32_t seed, int kind, const double *params, double L,
                            double rms, double *state, cf_stream stream) {
  CF_TRY(plan_check(p));
  CF_REQUIRE(state, "cf_init_grf_state: null state");
  const int64_t N = p->N, N3 = N * N * N;
  hipStream_t st = S(stream);
  std::vector<double> P((size_t)p->nb);
  CF_TRY(cf_spectrum_table_host(kind, N, L, params, P.data()));
  for (auto &v : P) v = sqrt(v);
  double *host = nullptr;
  CF_HIP(hipHostMalloc((void **)&host, (size_t)N3 * sizeof(double), hipHostMallocDefault));
  {
    // random_field.cpp:44-46: w_i = (1/N^3) * normal_i, in index order
    std::mt19937 mt{seed};
    std::normal_distribution<double> dist{0.0, 1.0};
    const double scale = 1.0 / (double)(N * N * N);
    for (int64_t i = 0; i < N3; ++i) host[i] = scale * dist(mt);
  }
  double *sqrtP = nullptr;
  cf_status rc = CF_OK;
  if (hipMalloc((void **)&sqrtP, P.size() * sizeof(double)) != hipSuccess) {
    (void)hipHostFree(host);
    set_error("cf_init_grf_state: allocation failed");
    return CF_E_NOMEM;
  }
  // the white noise goes to the pi half of the state, which is zeroed afterwards
  double *w = state + N3;
  if (hipMemcpyAsync(w, host, (size_t)N3 * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(sqrtP, P.data(), P.size() * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess) {
    rc = CF_E_HIP;
    set_error("cf_init_grf_state: upload failed");
  }
  if (rc == CF_OK) rc = cf_grf(p, w, sqrtP, nullptr, state, stream);
  if (rc == CF_OK && rms > 0.0) rc = rescale_rms_with(state, N3, rms, p->red, st);
  if (rc == CF_OK) rc = cf_fill_zero(w, (size_t)N3 * sizeof(double), stream);
  hipError_t se = hipStreamSynchronize(st);
  (void)hipFree(sqrtP);
  (void)hipHostFree(host);
  if (rc != CF_OK) return rc;
  CF_HIP(se);
  return CF_OK;
}

}  // extern "C"
