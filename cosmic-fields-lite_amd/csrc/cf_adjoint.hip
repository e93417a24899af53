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
// The backward step runs as two column-marching kernels (k_adj_back_a, k_adj_back_b)
// that recompute the stage states they need from x on the fly with the forward
// stage arithmetic, in the forward kernels' operation order (same bits).
//
// Checkpointing: the forward pass keeps the state every `ckpt` steps; the
// backward pass recomputes one segment of states at a time. Scalars (L, dL/dm)
// come from fixed-grid block partials summed on the host in a fixed order, so
// every result is bit-reproducible.
#include <cmath>
#include <utility>
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

// Coefficients of one backward step from t: the forward right-hand side at t
// (stage 1) and at t + dt/2 (stages 2 and 3), -3H and inv_ah_sqr at t + dt (stage 4).
struct AdjStep {
  RhsCoef r1, r2;
  double ch4, inv4;
  double dh, dt, dt3, dt6;  // dt/2, dt, dt/3, dt/6
  double two_m;             // 2m
};

// Tile of the v-th slot of the fixed grid's walk. Blocks go to the 8 XCDs round
// robin (block mod 8), and a block's slots keep its residue, so when the tile
// count allows, each XCD is given one contiguous range of tiles: the halo rows
// and planes that neighbouring tiles share meet in one L2.
__device__ __forceinline__ int adj_tile(int v, int total) {
  return (total & 7) == 0 ? (v & 7) * (total >> 3) + (v >> 3) : v;
}

// The transposed Jacobian applied to a at a site, phi component:
// (-m^2 - 3 lambda phi_y^2) a_pi + inv (sum of the six neighbours' a_pi - 6 a_pi).
__device__ __forceinline__ double jt_phi(double c_m, double lam, double phi_y, double inv, double a0, double nb) {
  return (-c_m - 3.0 * lam * phi_y * phi_y) * a0 + inv * (nb - 6.0 * a0);
}

// Backward step, first half: stages 4, 3 and the local part of stage 2.
//
// Only phi of each stage state enters the transposed Jacobian, and only at the
// site, and phi of a stage state is local in the previous one:
//   y1.phi = x.phi + h/2 x.pi,  y2.phi = x.phi + h/2 y1.pi,  y3.phi = x.phi + h y2.pi
// with y1.pi = x.pi + h/2 f_pi(x) and y2.pi = x.pi + h/2 f_pi(y1): every y_s.phi at
// a site follows from x within distance 1 (the forward stage arithmetic, so the
// same bits as the stored stage states). Likewise, with ybar_s = J(y_s)^T a_{s+1},
// ybar.pi = a.phi - 3H a.pi is local and only ybar.phi has a stencil, so
//   a4 = h/6 lambda'                         (local)
//   ybar3 (at y3)                            phi: stencil of a4.pi
//   a3 = h/3 lambda' + h ybar3               pi: local
//   ybar2 (at y2)                            phi: stencil of a3.pi
//   a2 = h/3 lambda' + h/2 ybar2             at the site
//   ybar1.pi = a2.phi - 3H a2.pi,  a1.pi = h/6 lambda'.pi + h/2 ybar1.pi
// all within distance 1 of lambda' and x. Writes s = (lambda' + ybar3) + ybar2 (the
// pi half also + ybar1.pi, the order of the running sum) and w = (a2.pi, a1.pi),
// the two fields the second half takes stencils of. part: stage 4, 3, 2 partials
// of dL/dm, ADJ_BLOCKS apart. One thread per (b,c) column marching along a; a fixed
// grid of ADJ_BLOCKS blocks walks the tiles so the partials have a fixed order.
template <bool LAM>
__global__ __launch_bounds__(ADJ_THREADS) void k_adj_back_a(Tiling tl, AdjStep k, const double *__restrict__ x,
                                                            const double *__restrict__ lp, double *__restrict__ s,
                                                            double *__restrict__ w, double *__restrict__ part) {
  const int64_t N = tl.N, NN = N * N, N3 = NN * N;
  const double c_m = k.r1.c_m, lam = k.r1.lam;
  // per-point fields from loaded (x.phi, x.pi) and (lambda'.phi, lambda'.pi)
  auto y1phi = [&](double f, double p) { return f + k.dh * p; };
  auto a4pi = [&](double lpi) { return k.dt6 * lpi; };
  auto a3pi = [&](double lphi, double lpi) {
    const double yb4pi = k.dt6 * lphi + k.ch4 * a4pi(lpi);
    return k.dt3 * lpi + k.dt * yb4pi;
  };
  double dm4 = 0.0, dm3 = 0.0, dm2 = 0.0;
  for (int v = blockIdx.x; v < tl.total; v += ADJ_BLOCKS) {
    const int t = adj_tile(v, tl.total);
    const int ct = t % tl.n_ct, r = t / tl.n_ct, bt = r % tl.n_bt, ac = r / tl.n_bt;
    const int64_t c = (int64_t)ct * TC + threadIdx.x;
    const int64_t b = (int64_t)bt * TB + threadIdx.y;
    const int64_t a0 = (int64_t)ac * tl.A;
    if (c >= N || b >= N || a0 >= N) continue;
    const int64_t a1 = (a0 + tl.A < N) ? a0 + tl.A : N;
    const int64_t cp = (c + 1 == N) ? 0 : c + 1, cm = (c == 0) ? N - 1 : c - 1;
    const int64_t bp = (b + 1 == N) ? 0 : b + 1, bm = (b == 0) ? N - 1 : b - 1;
    const int64_t col = b * N + c;
    const int64_t nbr[4] = {bp * N + c, bm * N + c, b * N + cp, b * N + cm};  // in-plane neighbours
    const int64_t jp = ((a0 == 0) ? N - 1 : a0 - 1) * NN + col, jc = a0 * NN + col;
    double xp = x[jp], xc = x[jc], xpic = x[N3 + jc], lphic = lp[jc], lpic = lp[N3 + jc];
    double yp = y1phi(xp, x[N3 + jp]), yc = y1phi(xc, xpic);
    double fp = a4pi(lp[N3 + jp]), fc = a4pi(lpic);
    double gp = a3pi(lp[jp], lp[N3 + jp]), gc = a3pi(lphic, lpic);
    // the loads of plane a: (x.phi, x.pi, lambda'.phi, lambda'.pi) at a+1 on the
    // column (e[0..3]) and at plane a's four in-plane neighbours (e[4..19]); those of
    // the next plane are issued before plane a's arithmetic, so that they overlap it
    auto load_plane = [&](int64_t a, double *e) {
      const int64_t ap = (a + 1 == N) ? 0 : a + 1;
      const int64_t row = a * NN, jn = ap * NN + col;
      e[0] = x[jn];
      e[1] = x[N3 + jn];
      e[2] = lp[jn];
      e[3] = lp[N3 + jn];
#pragma unroll
      for (int d = 0; d < 4; ++d) {
        e[4 + d] = x[row + nbr[d]];
        e[8 + d] = x[N3 + row + nbr[d]];
        e[12 + d] = lp[row + nbr[d]];
        e[16 + d] = lp[N3 + row + nbr[d]];
      }
    };
    // plane a from e (loaded earlier) while plane a+1 loads into en; unrolled by two
    // so that the two buffers trade places without copies
    auto plane = [&](int64_t a, const double *e, double *en) {
      const int64_t row = a * NN, i = row + col;
      load_plane((a + 1 < a1) ? a + 1 : a, en);
      __builtin_amdgcn_sched_barrier(0);
        const double xn = e[0], xpin = e[1], lphin = e[2], lpin = e[3];
        const double *qf = e + 4, *qp = e + 8, *ql = e + 12, *qm = e + 16;
        const double yn = y1phi(xn, xpin), fn = a4pi(lpin), gn = a3pi(lphin, lpin);
        // stage states at the site
        const double y1pi = xpic + k.dh * rhs_pi(xc, xn, xp, qf[0], qf[1], qf[2], qf[3], xpic, k.r1, LAM);
        const double y2pi = xpic + k.dh * rhs_pi(yc, yn, yp, y1phi(qf[0], qp[0]), y1phi(qf[1], qp[1]),
                                                            y1phi(qf[2], qp[2]), y1phi(qf[3], qp[3]), y1pi, k.r2, LAM);
        const double y2phi = xc + k.dh * y1pi;
        const double y3phi = xc + k.dt * y2pi;
        // stage 4 at y3 (t + dt)
        const double nb4 = fn + fp + a4pi(qm[0]) + a4pi(qm[1]) + a4pi(qm[2]) + a4pi(qm[3]);
        const double yb4phi = jt_phi(c_m, lam, y3phi, k.inv4, fc, nb4);
        const double yb4pi = k.dt6 * lphic + k.ch4 * fc;
        dm4 += (-k.two_m * y3phi) * fc;
        // stage 3 at y2 (t + dt/2)
        const double nb3 =
            gn + gp + a3pi(ql[0], qm[0]) + a3pi(ql[1], qm[1]) + a3pi(ql[2], qm[2]) + a3pi(ql[3], qm[3]);
        const double yb3phi = jt_phi(c_m, lam, y2phi, k.r2.inv, gc, nb3);
        const double yb3pi = (k.dt3 * lphic + k.dt * yb4phi) + k.r2.c_h * gc;
        dm3 += (-k.two_m * y2phi) * gc;
        // stage 2 at y1 (t + dt/2): the local part
        const double a2pi = k.dt3 * lpic + k.dh * yb3pi;
        const double a2phi = k.dt3 * lphic + k.dh * yb3phi;
        const double yb2pi = a2phi + k.r2.c_h * a2pi;
        dm2 += (-k.two_m * yc) * a2pi;
        s[i] = (lphic + yb4phi) + yb3phi;
        s[N3 + i] = ((lpic + yb4pi) + yb3pi) + yb2pi;
        w[i] = a2pi;
        w[N3 + i] = k.dt6 * lpic + k.dh * yb2pi;
        xp = xc; xc = xn; xpic = xpin; lphic = lphin; lpic = lpin;
        yp = yc; yc = yn;
        fp = fc; fc = fn;
        gp = gc; gc = gn;
    };
    double e0[20], e1[20];
    load_plane(a0, e0);
    for (int64_t a = a0; a < a1; a += 2) {
      plane(a, e0, e1);
      if (a + 1 < a1) plane(a + 1, e1, e0);
    }
  }
  const int tid = (int)(threadIdx.y * TC + threadIdx.x);
  block_partial(dm4, part + blockIdx.x, tid);
  block_partial(dm3, part + ADJ_BLOCKS + blockIdx.x, tid);
  block_partial(dm2, part + 2 * ADJ_BLOCKS + blockIdx.x, tid);
}

// Backward step, second half: ybar1.phi (at y1, stencil of a2.pi) and J(x)^T a1
// (stencil of a1.pi), and lambda = ((s + ybar1) + J(x)^T a1); part: the stage 1
// partials of dL/dm. lo must not overlap x, lp, s or w.
__global__ __launch_bounds__(ADJ_THREADS) void k_adj_back_b(Tiling tl, AdjStep k, const double *__restrict__ x,
                                                            const double *__restrict__ lp,
                                                            const double *__restrict__ s,
                                                            const double *__restrict__ w, double *__restrict__ lo,
                                                            double *__restrict__ part) {
  const int64_t N = tl.N, NN = N * N, N3 = NN * N;
  const double c_m = k.r1.c_m, lam = k.r1.lam;
  double dm1 = 0.0;
  for (int v = blockIdx.x; v < tl.total; v += ADJ_BLOCKS) {
    const int t = adj_tile(v, tl.total);
    const int ct = t % tl.n_ct, r = t / tl.n_ct, bt = r % tl.n_bt, ac = r / tl.n_bt;
    const int64_t c = (int64_t)ct * TC + threadIdx.x;
    const int64_t b = (int64_t)bt * TB + threadIdx.y;
    const int64_t a0 = (int64_t)ac * tl.A;
    if (c >= N || b >= N || a0 >= N) continue;
    const int64_t a1 = (a0 + tl.A < N) ? a0 + tl.A : N;
    const int64_t cp = (c + 1 == N) ? 0 : c + 1, cm = (c == 0) ? N - 1 : c - 1;
    const int64_t bp = (b + 1 == N) ? 0 : b + 1, bm = (b == 0) ? N - 1 : b - 1;
    const int64_t col = b * N + c;
    const int64_t jp = ((a0 == 0) ? N - 1 : a0 - 1) * NN + col, jc = a0 * NN + col;
    const int64_t nbr[4] = {bp * N + c, bm * N + c, b * N + cp, b * N + cm};  // in-plane neighbours
    double up = w[jp], uc = w[jc];            // a2.pi
    double vp = w[N3 + jp], vc = w[N3 + jc];  // a1.pi
    for (int64_t a = a0; a < a1; ++a) {
      const int64_t ap = (a + 1 == N) ? 0 : a + 1;
      const int64_t row = a * NN, i = row + col, jn = ap * NN + col;
      // every load of the iteration first, so that they are in flight together
      const double un = w[jn], vn = w[N3 + jn];
      const double xphi = x[i], xpi = x[N3 + i], lphi = lp[i], sphi = s[i], spi = s[N3 + i];
      double qu[4], qv[4];
#pragma unroll
      for (int d = 0; d < 4; ++d) {
        qu[d] = w[row + nbr[d]];
        qv[d] = w[N3 + row + nbr[d]];
      }
      __builtin_amdgcn_sched_barrier(0);
      const double y1 = xphi + k.dh * xpi;
      const double nb2 = un + up + qu[0] + qu[1] + qu[2] + qu[3];
      const double yb2phi = jt_phi(c_m, lam, y1, k.r2.inv, uc, nb2);
      const double nb1 = vn + vp + qv[0] + qv[1] + qv[2] + qv[3];
      const double outphi = jt_phi(c_m, lam, xphi, k.r1.inv, vc, nb1);
      const double outpi = (k.dt6 * lphi + k.dh * yb2phi) + k.r1.c_h * vc;
      lo[i] = (sphi + yb2phi) + outphi;
      lo[N3 + i] = spi + outpi;
      dm1 += (-k.two_m * xphi) * vc;
      up = uc; uc = un;
      vp = vc; vc = vn;
    }
  }
  block_partial(dm1, part + blockIdx.x, (int)(threadIdx.y * TC + threadIdx.x));
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
  // checkpoints + one segment of states + step work (2) + lambda (2) + rk4 work (3)
  return nck * v + (size_t)ckpt * v + 2 * v + 2 * v + 3 * v +
         (size_t)(4 * n_steps + 1) * ADJ_BLOCKS * sizeof(double);
}

// L and dL/dx (gx_dev, 2N^3) and the direct dL/dm at a state x; weights_dev has
// cf_spectrum_bins(N) entries. scratch_dev: N^3 doubles. Blocking (L, dL/dm on host).
cf_status cf_loss_spectrum_adjoint(cf_spectrum_plan *plan, const cf_kg_params *p, const double *x, double t,
                                   const double *weights_dev, double *gx_dev, double *scratch_dev, double *loss,
                                   double *dldm, cf_stream stream) {
  CF_REQUIRE(plan && p && x && weights_dev && gx_dev && scratch_dev && loss && dldm,
             "cf_loss_spectrum_adjoint: null argument");
  CF_REQUIRE(plan_size(plan) == p->N, "cf_loss_spectrum_adjoint: plan is for N=%lld, lattice N=%lld",
             (long long)plan_size(plan), (long long)p->N);
  const int64_t N = p->N, N3 = N * N * N, H = N / 2 + 1, nb = cf_spectrum_bins(N);
  hipStream_t st = S(stream);
  const double a_t = cf_cosmo_a(&p->cosmo, t);
  const double inv = cf_inv_ah_sqr(N, p->L, a_t);
  double *rho = scratch_dev;
  CF_TRY(cf_energy_density_f64(N, x, rho, p->m, p->lambda, inv, stream));
  double *fk = gx_dev;  // the half spectrum (2 N^2 H doubles) fits in gx (2N^3) until rho's gradient is formed
  CF_TRY(cf_fft_d2z(plan, rho, fk, stream));
  static_assert(ADJ_BLOCKS <= 1025, "dL/dm partials live in the plan's reduction buffer");
  double *bins = plan_bins(plan), *part = plan_red(plan);
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

// One backward step from state x at time t: lo = dL/dx_n from li = dL/dx_{n+1};
// part: 4*ADJ_BLOCKS partials of dL/dm (stages 4, 3, 2, 1). sw, ww: 2N^3 doubles
// each. lo overlaps none of x, li, sw, ww.
static cf_status adj_step(const cf_kg_params *p, const double *x, double t, double dt, const double *li, double *lo,
                          double *part, double *sw, double *ww, hipStream_t st) {
  const int64_t N = p->N;
  const double dh = dt / 2.0, th = t + dh;
  auto rhs = [&](double tt) {
    RhsCoef r;
    r.c_h = (-3.0) * cf_cosmo_H(&p->cosmo, tt);
    r.c_m = p->m * p->m;
    r.lam = p->lambda;
    r.inv = cf_inv_ah_sqr(N, p->L, cf_cosmo_a(&p->cosmo, tt));
    return r;
  };
  AdjStep k;
  k.r1 = rhs(t);
  k.r2 = rhs(th);
  const RhsCoef r4 = rhs(t + dt);
  k.ch4 = r4.c_h;
  k.inv4 = r4.inv;
  k.dh = dh;
  k.dt = dt;
  k.dt3 = dt / 3.0;
  k.dt6 = dt / 6.0;
  k.two_m = 2.0 * p->m;
  const Tiling tl = make_tiling(N, default_chunk(N));
  const dim3 g(ADJ_BLOCKS), blk(TC, TB);
  static_assert(TC * TB == ADJ_THREADS, "block shape");
  if (p->lambda != 0.0)
    hipLaunchKernelGGL(k_adj_back_a<true>, g, blk, 0, st, tl, k, x, li, sw, ww, part);
  else
    hipLaunchKernelGGL(k_adj_back_a<false>, g, blk, 0, st, tl, k, x, li, sw, ww, part);
  hipLaunchKernelGGL(k_adj_back_b, g, blk, 0, st, tl, k, x, li, (const double *)sw, (const double *)ww, lo,
                     part + 3 * ADJ_BLOCKS);
  CF_LAUNCHED();
  return CF_OK;
}

// One RK4 step adjoint at state x_n (time t): lam_dev holds dL/dx_{n+1} on entry and
// dL/dx_n on return; the step's dL/dm contributions go to part_dev (4*ADJ_BLOCKS
// doubles, summed by the caller). work_dev: 3 vectors of 2N^3 doubles.
cf_status cf_rk4_step_adjoint_f64(const cf_kg_params *p, const double *x, double t, double dt, double *lam,
                                  double *part_dev, void *work_dev, cf_stream stream) {
  CF_REQUIRE(p && x && lam && part_dev && work_dev, "cf_rk4_step_adjoint_f64: null argument");
  CF_REQUIRE(valid_N(p->N), "cf_rk4_step_adjoint_f64: unsupported N");
  const int64_t N = p->N, n = 2 * N * N * N;
  hipStream_t st = S(stream);
  double *sw = (double *)work_dev, *ww = sw + n, *lo = ww + n;
  CF_TRY(adj_step(p, x, t, dt, lam, lo, part_dev, sw, ww, st));
  CF_HIP(hipMemcpyAsync(lam, lo, n * sizeof(double), hipMemcpyDeviceToDevice, st));
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
  double *sw = seg + (size_t)ckpt * n;             // step-adjoint work: 2 vectors
  double *ww = sw + n;
  double *lamb = ww + n;                           // lambda, two buffers taken in turn
  double *lamb2 = lamb + n;
  double *fw = lamb2 + n;                          // forward work: 3 vectors
  double *part = fw + 3 * n;                       // (4 n_steps + 1) * ADJ_BLOCKS
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
      CF_TRY(adj_step(p, seg + (size_t)(k - k0) * n, t, dt, lamb, lamb2, part + (size_t)k * 4 * ADJ_BLOCKS, sw, ww,
                      st));
      std::swap(lamb, lamb2);
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
