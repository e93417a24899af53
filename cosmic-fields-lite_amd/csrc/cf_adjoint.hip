// cf_adjoint.hip -- discrete adjoint of the RK4 lattice evolution and of the
// spectral loss  L = sum_s w_s P_rho(s)  (P_rho: the binned power spectrum of the
// energy density, compute_power_spectrum of fdm3d.cpp:7-34, at the final time).
//
// Loss adjoint. With rho_k the unnormalised DFT of rho and the bins summing the
// full k-cube, L = sum_k w_{s(k)} |rho_k|^2, so dL/drho(x) = 2 sum_k w rho_k e^{ikx}
// = 2 * C2R(w . rho_k): one D2Z, a per-mode multiply and one Z2D. The energy
// density (equations.cpp:106-138 with the lambda term of equations_cuda.cu:354)
//   rho = pi^2/2 + m^2 phi^2/2 + lambda phi^4/4 + (inv/8) sum_d (phi(x+e_d) - phi(x-e_d))^2
// gives
//   dL/dpi(x)  = g(x) pi(x)
//   dL/dphi(x) = g(x) (m^2 phi + lambda phi^3)
//              + (inv/4) sum_d [ g(x-e_d)(phi(x) - phi(x-2e_d)) - g(x+e_d)(phi(x+2e_d) - phi(x)) ]
//   dL/dm     += sum_x g(x) m phi(x)^2.
//
// Step adjoint. For x' = x + h/6 k1 + h/3 k2 + h/3 k3 + h/6 k4 with k_i = f(y_{i-1}),
// y0 = x, y1 = x + h/2 k1, y2 = x + h/2 k2, y3 = x + h k3 (runge_kutta4_classic.hpp:
// 87-138), and lambda' = dL/dx':
//   a4 = h/6 lambda';                 ybar3 = J(y3)^T a4
//   a3 = h/3 lambda' + h ybar3;       ybar2 = J(y2)^T a3
//   a2 = h/3 lambda' + h/2 ybar2;     ybar1 = J(y1)^T a2
//   a1 = h/6 lambda' + h/2 ybar1
//   lambda = lambda' + ybar3 + ybar2 + ybar1 + J(x)^T a1
//   dL/dm += sum_i (df/dm at y_{i-1}) . a_i,  df/dm = [0, -2 m phi]
// where f = [pi, -3H pi - m^2 phi - lambda phi^3 + inv Lap phi] and
//   (J^T a)_phi = (-m^2 - 3 lambda phi^2) a_pi + inv (sum_nbr a_pi - 6 a_pi)
//   (J^T a)_pi  = a_phi - 3H a_pi.
// The stage states y_i are recomputed from x with the forward kernels (same bits).
//
// Checkpointing: the forward pass keeps the state every `ckpt` steps; the
// backward pass recomputes one segment of states at a time. Scalars (L, dL/dm)
// come from fixed-grid block partials summed on the host in a fixed order, so
// every result is bit-reproducible.
#include <cmath>
#include <vector>

#include "cf_march.hpp"

namespace cf {

constexpr int ADJ_BLOCKS = 512;
constexpr int ADJ_THREADS = 256;

__device__ __forceinline__ int64_t wrapi(int64_t i, int64_t N) { return i < 0 ? i + N : (i >= N ? i - N : i); }

// Fixed-order tree sum of one value per thread (tid = linear thread index).
__device__ __forceinline__ void block_partial(double v, double *part_slot, int tid) {
  __shared__ double sm[ADJ_THREADS];
  sm[tid] = v;
  __syncthreads();
  for (int s = ADJ_THREADS / 2; s > 0; s >>= 1) {
    if (tid < s) sm[tid] += sm[tid + s];
    __syncthreads();
  }
  if (tid == 0) *part_slot = sm[0];
}

struct AdjCoef {
  double c_h;   // -3H at the stage time
  double c_m;   // m*m
  double lam;   // lambda
  double inv;   // inv_ah_sqr at the stage time
  double two_m; // 2m
  double ck;    // a = ck*lambda' + dk*ybar_prev
  double dk;
};

// Stage-state recompute: y = x + cy * f(in) (the forward stage kernels' arithmetic,
// so y1..y3 equal the forward step's stage states bit for bit). One thread per (b,c)
// column marching along a, as the forward stage kernels.
template <bool LAM>
__global__ __launch_bounds__(ADJ_THREADS) void k_adj_y(Tiling tl, RhsCoef rk, const double *__restrict__ in,
                                                       const double *__restrict__ x, double *__restrict__ y,
                                                       double cy) {
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
  double prev = in[((a0 == 0) ? N - 1 : a0 - 1) * NN + col];
  double cur = in[a0 * NN + col];
  for (int64_t a = a0; a < a1; ++a) {
    const int64_t ap = (a + 1 == N) ? 0 : a + 1;
    const int64_t row = a * NN, i = row + col;
    const double next = in[ap * NN + col];
    const double piv = in[N3 + i];
    const double kpi = rhs_pi(cur, next, prev, in[row + bp * N + c], in[row + bm * N + c], in[row + b * N + cp],
                              in[row + b * N + cm], piv, rk, LAM);
    y[i] = x[i] + cy * piv;
    y[N3 + i] = x[N3 + i] + cy * kpi;
    prev = cur;
    cur = next;
  }
}

// One backward stage: a = ck*lp + dk*yb_prev (HAS_YB) or ck*lp; out = J(y)^T a.
// MODE 0 and 1: ybar = out (each stage's output kept in its own buffer); MODE 2 (the
// last): acc = (((lp + yb4) + yb3) + yb_prev) + out, the lambda of the step in the
// order of a running sum, without reading and rewriting an accumulator per stage.
// part[blockIdx] = sum over the block's sites of (-2 m phi_y) a_pi, in a fixed order.
// A fixed grid of ADJ_BLOCKS blocks walks the tiles (one thread per (b,c) column
// marching along a, a_pi at a-1, a, a+1 kept in registers).
template <int MODE, bool HAS_YB>
__global__ __launch_bounds__(ADJ_THREADS) void k_adj_back(Tiling tl, const double *__restrict__ lp,
                                                          const double *__restrict__ ybp,
                                                          const double *__restrict__ y, double *__restrict__ ybar,
                                                          double *__restrict__ acc, const double *__restrict__ yb4,
                                                          const double *__restrict__ yb3, AdjCoef k,
                                                          double *__restrict__ part) {
  const int64_t N = tl.N, NN = N * N, N3 = NN * N;
  auto api = [&](int64_t j) {
    const double v = k.ck * lp[N3 + j];
    if constexpr (HAS_YB) return v + k.dk * ybp[N3 + j];
    else return v;
  };
  double dm = 0.0;
  for (int t = blockIdx.x; t < tl.total; t += ADJ_BLOCKS) {
    const int ct = t % tl.n_ct, r = t / tl.n_ct, bt = r % tl.n_bt, ac = r / tl.n_bt;
    const int64_t c = (int64_t)ct * TC + threadIdx.x;
    const int64_t b = (int64_t)bt * TB + threadIdx.y;
    const int64_t a0 = (int64_t)ac * tl.A;
    if (c >= N || b >= N || a0 >= N) continue;
    const int64_t a1 = (a0 + tl.A < N) ? a0 + tl.A : N;
    const int64_t cp = (c + 1 == N) ? 0 : c + 1, cm = (c == 0) ? N - 1 : c - 1;
    const int64_t bp = (b + 1 == N) ? 0 : b + 1, bm = (b == 0) ? N - 1 : b - 1;
    const int64_t col = b * N + c;
    double prev = api(((a0 == 0) ? N - 1 : a0 - 1) * NN + col);
    double cur = api(a0 * NN + col);
    for (int64_t a = a0; a < a1; ++a) {
      const int64_t ap = (a + 1 == N) ? 0 : a + 1;
      const int64_t row = a * NN, i = row + col;
      const double next = api(ap * NN + col);
      double aphi = k.ck * lp[i];
      if constexpr (HAS_YB) aphi = aphi + k.dk * ybp[i];
      const double a0v = cur;
      const double nb = next + prev + api(row + bp * N + c) + api(row + bm * N + c) + api(row + b * N + cp) +
                        api(row + b * N + cm);
      const double phi = y[i];
      const double out_phi = (-k.c_m - 3.0 * k.lam * phi * phi) * a0v + k.inv * (nb - 6.0 * a0v);
      const double out_pi = aphi + k.c_h * a0v;
      if constexpr (MODE == 2) {
        acc[i] = (((lp[i] + yb4[i]) + yb3[i]) + ybp[i]) + out_phi;
        acc[N3 + i] = (((lp[N3 + i] + yb4[N3 + i]) + yb3[N3 + i]) + ybp[N3 + i]) + out_pi;
      } else {
        ybar[i] = out_phi;
        ybar[N3 + i] = out_pi;
      }
      dm += (-k.two_m * phi) * a0v;
      prev = cur;
      cur = next;
    }
  }
  block_partial(dm, part + blockIdx.x, (int)(threadIdx.y * TC + threadIdx.x));
}

// Half-spectrum fk *= 2 w[s].
__global__ void k_weight_modes(double2 *__restrict__ fk, const double *__restrict__ w, int64_t N, int64_t H) {
  const int64_t total = N * N * H;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = e % H, r = e / H, b = r % N, a = r / N;
    const int64_t as = a <= N / 2 ? a : N - a, bs = b <= N / 2 ? b : N - b;
    const double s2 = 2.0 * w[as * as + bs * bs + c * c];
    double2 v = fk[e];
    v.x *= s2;
    v.y *= s2;
    fk[e] = v;
  }
}

// Energy-density adjoint: g (N^3) -> gx (2N^3); part = sum g m phi^2.
__global__ __launch_bounds__(ADJ_THREADS) void k_rho_adjoint(int64_t N, const double *__restrict__ x,
                                                             const double *__restrict__ g, double *__restrict__ gx,
                                                             double m, double lam, double inv,
                                                             double *__restrict__ part) {
  const int64_t NN = N * N, N3 = NN * N;
  double dm = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * ADJ_THREADS + threadIdx.x; i < N3; i += (int64_t)ADJ_BLOCKS * ADJ_THREADS) {
    const int64_t a = i / NN, b = (i / N) % N, c = i % N;
    const int64_t co[3] = {a, b, c};
    const int64_t st[3] = {NN, N, 1};
    const double phi = x[i];
    double gphi = g[i] * (m * m * phi + lam * phi * phi * phi);
    for (int d = 0; d < 3; ++d) {
      const int64_t base = i - co[d] * st[d];
      const int64_t jm = base + wrapi(co[d] - 1, N) * st[d], jp = base + wrapi(co[d] + 1, N) * st[d];
      const int64_t jmm = base + wrapi(co[d] - 2, N) * st[d], jpp = base + wrapi(co[d] + 2, N) * st[d];
      gphi += 0.25 * inv * (g[jm] * (phi - x[jmm]) - g[jp] * (x[jpp] - phi));
    }
    gx[i] = gphi;
    gx[N3 + i] = g[i] * x[N3 + i];
    dm += g[i] * m * phi * phi;
  }
  block_partial(dm, part + blockIdx.x, (int)threadIdx.x);
}

static double host_sum(const std::vector<double> &v, size_t beg, size_t n) {
  double s = 0.0;
  for (size_t i = 0; i < n; ++i) s += v[beg + i];
  return s;
}

}  // namespace cf

using namespace cf;

extern "C" {

size_t cf_adjoint_work_bytes(int64_t N, int64_t n_steps, int64_t ckpt) {
  if (!valid_N(N) || n_steps < 0 || ckpt <= 0) return 0;
  const size_t v = 2 * (size_t)N * N * N * sizeof(double);
  const size_t nck = (size_t)((n_steps + ckpt - 1) / ckpt);
  // checkpoints + one segment of states + y1,y2,y3 + k + lambda, lambda', ybar x2, acc + rk4 work (3)
  return nck * v + (size_t)ckpt * v + 3 * v + v + 5 * v + 3 * v +
         (size_t)(4 * n_steps + 1) * ADJ_BLOCKS * sizeof(double);
}

// L and dL/dx (gx_dev, 2N^3) and the direct dL/dm at a state x; weights_dev has
// cf_spectrum_bins(N) entries. scratch_dev: N^3 doubles. Blocking (L, dL/dm on host).
cf_status cf_loss_spectrum_adjoint(cf_spectrum_plan *plan, const cf_kg_params *p, const double *x, double t,
                                   const double *weights_dev, double *gx_dev, double *scratch_dev, double *loss,
                                   double *dldm, cf_stream stream) {
  CF_REQUIRE(plan && p && x && weights_dev && gx_dev && scratch_dev && loss && dldm,
             "cf_loss_spectrum_adjoint: null argument");
  const int64_t N = p->N, N3 = N * N * N, H = N / 2 + 1, nb = cf_spectrum_bins(N);
  hipStream_t st = S(stream);
  const double a_t = cf_cosmo_a(&p->cosmo, t);
  const double inv = cf_inv_ah_sqr(N, p->L, a_t);
  double *rho = scratch_dev;
  CF_TRY(cf_energy_density_f64(N, x, rho, p->m, p->lambda, inv, stream));
  double *fk = gx_dev;  // the half spectrum (2 N^2 H doubles) fits in gx (2N^3) until rho's gradient is formed
  CF_TRY(cf_fft_d2z(plan, rho, fk, stream));
  double *bins = nullptr, *part = nullptr;
  CF_HIP(hipMallocAsync((void **)&bins, nb * sizeof(double), st));
  CF_HIP(hipMallocAsync((void **)&part, ADJ_BLOCKS * sizeof(double), st));
  cf_status rc = cf_bin_power(plan, fk, bins, stream);
  std::vector<double> hb((size_t)nb), hw((size_t)nb), hp(ADJ_BLOCKS);
  if (rc == CF_OK && (hipMemcpyAsync(hb.data(), bins, nb * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
                      hipMemcpyAsync(hw.data(), weights_dev, nb * sizeof(double), hipMemcpyDeviceToHost, st) !=
                          hipSuccess))
    rc = CF_E_HIP;
  if (rc == CF_OK) {
    hipLaunchKernelGGL(k_weight_modes, dim3(4096), dim3(256), 0, st, (double2 *)fk, weights_dev, N, H);
    if (hipGetLastError() != hipSuccess) rc = CF_E_HIP;
  }
  if (rc == CF_OK) rc = cf_fft_z2d(plan, fk, rho, stream);  // rho <- dL/drho
  if (rc == CF_OK) {
    hipLaunchKernelGGL(k_rho_adjoint, dim3(ADJ_BLOCKS), dim3(ADJ_THREADS), 0, st, N, x, rho, gx_dev, p->m, p->lambda,
                       inv, part);
    if (hipGetLastError() != hipSuccess) rc = CF_E_HIP;
  }
  if (rc == CF_OK && hipMemcpyAsync(hp.data(), part, ADJ_BLOCKS * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess)
    rc = CF_E_HIP;
  hipError_t se = hipStreamSynchronize(st);
  (void)hipFree(bins);
  (void)hipFree(part);
  if (rc != CF_OK) {
    set_error("cf_loss_spectrum_adjoint: device operation failed");
    return rc;
  }
  CF_HIP(se);
  double L = 0.0;
  for (int64_t s = 0; s < nb; ++s) L += hw[s] * hb[s];
  *loss = L;
  *dldm = host_sum(hp, 0, ADJ_BLOCKS);
  (void)N3;
  return CF_OK;
}

// One RK4 step adjoint at state x_n (time t): lam_dev holds dL/dx_{n+1} on entry and
// dL/dx_n on return; the step's dL/dm contributions go to part_dev (4*ADJ_BLOCKS
// doubles, summed by the caller). work_dev: 9 vectors of 2N^3 doubles.
cf_status cf_rk4_step_adjoint_f64(const cf_kg_params *p, const double *x, double t, double dt, double *lam,
                                  double *part_dev, void *work_dev, cf_stream stream) {
  CF_REQUIRE(p && x && lam && part_dev && work_dev, "cf_rk4_step_adjoint_f64: null argument");
  CF_REQUIRE(valid_N(p->N), "cf_rk4_step_adjoint_f64: unsupported N");
  const int64_t N = p->N, n = 2 * N * N * N;
  hipStream_t st = S(stream);
  double *y1 = (double *)work_dev, *y2 = y1 + n, *y3 = y2 + n;
  double *yb4 = y3 + n, *yb3 = yb4 + n, *yb2 = yb3 + n, *acc = yb2 + n;
  const double dh = dt / 2.0, th = t + dh, dt6 = dt / 6.0, dt3 = dt / 3.0;
  const Tiling tl = make_tiling(N, default_chunk(N));
  const dim3 blk(TC, TB);
  static_assert(TC * TB == ADJ_THREADS, "block shape");
  // recompute the stage states with the forward stage arithmetic (same bits)
  auto ystage = [&](const double *in, double tt, double cy, double *y) -> cf_status {
    const double a_t = cf_cosmo_a(&p->cosmo, tt);
    RhsCoef rk;
    rk.c_h = (-3.0) * cf_cosmo_H(&p->cosmo, tt);
    rk.c_m = p->m * p->m;
    rk.lam = p->lambda;
    rk.inv = cf_inv_ah_sqr(N, p->L, a_t);
    if (rk.lam != 0.0)
      hipLaunchKernelGGL(k_adj_y<true>, dim3(tl.total), blk, 0, st, tl, rk, in, x, y, cy);
    else
      hipLaunchKernelGGL(k_adj_y<false>, dim3(tl.total), blk, 0, st, tl, rk, in, x, y, cy);
    CF_LAUNCHED();
    return CF_OK;
  };
  CF_TRY(ystage(x, t, dh, y1));
  CF_TRY(ystage(y1, th, dh, y2));
  CF_TRY(ystage(y2, th, dt, y3));
  auto coef = [&](double tt, double ck, double dk) {
    AdjCoef k;
    const double a_t = cf_cosmo_a(&p->cosmo, tt);
    k.c_h = (-3.0) * cf_cosmo_H(&p->cosmo, tt);
    k.c_m = p->m * p->m;
    k.lam = p->lambda;
    k.inv = cf_inv_ah_sqr(N, p->L, a_t);
    k.two_m = 2.0 * p->m;
    k.ck = ck;
    k.dk = dk;
    return k;
  };
  const dim3 g(ADJ_BLOCKS);
  const double *no = nullptr;
  hipLaunchKernelGGL((k_adj_back<0, false>), g, blk, 0, st, tl, lam, no, y3, yb4, (double *)nullptr, no, no,
                     coef(t + dt, dt6, 0.0), part_dev);
  hipLaunchKernelGGL((k_adj_back<1, true>), g, blk, 0, st, tl, lam, yb4, y2, yb3, (double *)nullptr, no, no,
                     coef(th, dt3, dt), part_dev + ADJ_BLOCKS);
  hipLaunchKernelGGL((k_adj_back<1, true>), g, blk, 0, st, tl, lam, yb3, y1, yb2, (double *)nullptr, no, no,
                     coef(th, dt3, dh), part_dev + 2 * ADJ_BLOCKS);
  hipLaunchKernelGGL((k_adj_back<2, true>), g, blk, 0, st, tl, lam, yb2, x, (double *)nullptr, acc, yb4, yb3,
                     coef(t, dt6, dh), part_dev + 3 * ADJ_BLOCKS);
  CF_LAUNCHED();
  CF_HIP(hipMemcpyAsync(lam, acc, n * sizeof(double), hipMemcpyDeviceToDevice, st));
  return CF_OK;
}

// Forward n_steps from x0 (t_k = t0 + k dt) with checkpoints every ckpt steps, the
// spectral loss at t_end, and the backward pass: grad_x0_dev = dL/dx0 (2N^3),
// *loss = L, *dldm = dL/dm (direct loss term + every step). Blocking.
cf_status cf_adjoint_spectrum_loss(cf_spectrum_plan *plan, const cf_kg_params *p, const double *x0, double t0,
                                   double dt, int64_t n_steps, int64_t ckpt, const double *weights_dev,
                                   double *grad_x0_dev, double *loss, double *dldm, void *work_dev,
                                   cf_stream stream) {
  CF_REQUIRE(plan && p && x0 && weights_dev && grad_x0_dev && loss && dldm && work_dev && n_steps >= 0 && ckpt > 0,
             "cf_adjoint_spectrum_loss: bad arguments");
  CF_REQUIRE(valid_N(p->N), "cf_adjoint_spectrum_loss: unsupported N");
  const int64_t N = p->N, n = 2 * N * N * N;
  const size_t vb = (size_t)n * sizeof(double);
  hipStream_t st = S(stream);
  const int64_t nck = (n_steps + ckpt - 1) / ckpt;
  double *ck = (double *)work_dev;                // nck checkpoints
  double *seg = ck + (size_t)nck * n;              // ckpt states of one segment
  double *sw = seg + (size_t)ckpt * n;             // step-adjoint work: 7 vectors
  double *lamb = sw + 7 * n;                       // lambda
  double *fw = lamb + n;                           // forward work: 3 vectors
  double *parts = fw + 3 * n;                      // (4 n_steps + 1) * ADJ_BLOCKS
  (void)parts;
  // forward with checkpoints
  double *x = seg;  // running state
  CF_HIP(hipMemcpyAsync(x, x0, vb, hipMemcpyDeviceToDevice, st));
  for (int64_t c = 0; c < nck; ++c) {
    CF_HIP(hipMemcpyAsync(ck + (size_t)c * n, x, vb, hipMemcpyDeviceToDevice, st));
    const int64_t k0 = c * ckpt, k1 = (k0 + ckpt < n_steps) ? k0 + ckpt : n_steps;
    CF_TRY(cf_rk4_run_tb_f64(p, x, t0, k0, k1 - k0, dt, 0, fw, stream));  // same bits as the stage kernels
  }
  const double t_end = t0 + static_cast<double>(n_steps) * dt;
  double dm_loss = 0.0;
  CF_TRY(cf_loss_spectrum_adjoint(plan, p, x, t_end, weights_dev, lamb, fw, loss, &dm_loss, stream));
  double *part = (double *)(fw + 3 * n);
  // backward, one segment at a time
  for (int64_t c = nck - 1; c >= 0; --c) {
    const int64_t k0 = c * ckpt, k1 = (k0 + ckpt < n_steps) ? k0 + ckpt : n_steps;
    CF_HIP(hipMemcpyAsync(seg, ck + (size_t)c * n, vb, hipMemcpyDeviceToDevice, st));
    for (int64_t k = k0 + 1; k < k1; ++k) {
      const double *prev = seg + (size_t)(k - 1 - k0) * n;
      double *cur = seg + (size_t)(k - k0) * n;
      CF_TRY(cf_rk4_step_tb_f64(p, prev, cur, t0 + static_cast<double>(k - 1) * dt, dt, 0, stream));
    }
    for (int64_t k = k1 - 1; k >= k0; --k) {
      const double t = t0 + static_cast<double>(k) * dt;
      CF_TRY(cf_rk4_step_adjoint_f64(p, seg + (size_t)(k - k0) * n, t, dt, lamb, part + (size_t)k * 4 * ADJ_BLOCKS,
                                     sw, stream));
    }
  }
  std::vector<double> hp((size_t)n_steps * 4 * ADJ_BLOCKS);
  if (n_steps) CF_HIP(hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  CF_HIP(hipMemcpyAsync(grad_x0_dev, lamb, vb, hipMemcpyDeviceToDevice, st));
  CF_HIP(hipStreamSynchronize(st));
  double dm = dm_loss;
  for (int64_t k = n_steps - 1; k >= 0; --k) dm += host_sum(hp, (size_t)k * 4 * ADJ_BLOCKS, 4 * ADJ_BLOCKS);
  *dldm = dm;
  return CF_OK;
}

}  // extern "C"
