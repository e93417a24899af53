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
// The backward step is one kernel (k_adj_step) that marches tiles of columns along
// a, recomputes the stage states it needs from x on the fly with the forward stage
// arithmetic, in the forward kernels' operation order (same bits), and keeps the
// intermediate adjoint fields in LDS.
//
// Checkpointing: the forward pass keeps the state every `ckpt` steps (the last
// segment's states as they are made); the backward pass recomputes one segment of
// states at a time. Scalars (L, dL/dm) come from fixed-grid block partials summed
// on the host in a fixed order, so every result is bit-reproducible.
#include <cmath>
#include <utility>
#include <vector>

#include "cf_march.hpp"

namespace cf {

constexpr int ADJ_BLOCKS = 512;
constexpr int ADJ_THREADS = 256;

__device__ __forceinline__ int64_t wrapi(int64_t i, int64_t N) { return i < 0 ? i + N : (i >= N ? i - N : i); }

// Fixed-order tree sum of one value per thread over a block of NT threads (tid =
// linear thread index; sm: NT doubles of LDS).
template <int NT>
__device__ __forceinline__ void block_partial(double v, double *part_slot, int tid, double *sm) {
  __syncthreads();  // sm may still be read by an earlier sum
  sm[tid] = v;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (tid < s) sm[tid] += sm[tid + s];
    __syncthreads();
  }
  if (tid == 0) *part_slot = sm[0];
}

__device__ __forceinline__ void block_partial(double v, double *part_slot, int tid) {
  __shared__ double sm[ADJ_THREADS];
  block_partial<ADJ_THREADS>(v, part_slot, tid, sm);
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

// Geometry of the backward step: a block of ADJ_R x ADJ_R threads, one per (b,c)
// column of a region, marches planes [a0, a1) of one tile; the tile is the region
// without a border of two columns (ADJ_T = ADJ_R - 4).
constexpr int ADJ_R = 32;
constexpr int ADJ_T = ADJ_R - 4;
constexpr int ADJ_FT = ADJ_R * ADJ_R;  // threads per block

struct AdjGeom {
  int64_t N;
  int nt;     // tiles per side (b and c)
  int A;      // planes per chunk
  int total;  // nt * nt * chunks
};

// Plane slots of the step in LDS (112 KB: one block per CU).
struct AdjLds {
  double q[2][4][ADJ_FT];  // per column: x.phi, y1.phi, a4.pi, a3.pi of one plane; two planes
  double w[3][2][ADJ_FT];  // per column: a2.pi, a1.pi of one plane; three planes
  double red[ADJ_FT];      // block sums of the dL/dm partials
};

// s_waitcnt immediate (gfx9 encoding) that waits for the wave's LDS operations
// (lgkmcnt 0) and for no vector-memory or export counter (vmcnt 63, expcnt 7).
constexpr int ADJ_WAIT_LDS = 15 | (3 << 14) | (7 << 4);

// Workgroup barrier for LDS plane slots only: every wave's LDS writes complete, then
// s_barrier. Global loads in flight (the next plane's prefetch) stay in flight; a
// __syncthreads() fence would wait for them too. The empty asm statements keep the
// compiler from moving memory operations across the barrier.
__device__ __forceinline__ void adj_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(ADJ_WAIT_LDS);
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// Backward step, fused.
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
// all within distance 1 of lambda' and x (half A). The rest needs the stencils of
// w = (a2.pi, a1.pi) (half B):
//   ybar1.phi (at y1): stencil of a2.pi;  J(x)^T a1: stencil of a1.pi
//   lambda = ((lambda' + ybar4) + ybar3 + ybar1).phi + (J(x)^T a1).phi, and the pi
//   half ((lambda' + ybar4 + ybar3) + ybar2 + ybar1).pi + (J(x)^T a1).pi
// (the order of the running sum s).
//
// Per plane j of the march, every thread loads x and lambda' at its own site of
// plane j+2 (used one iteration later, so the load overlaps a whole iteration) and
//   A(j) on the region minus its outer ring: half A at plane j. The in-plane
//        neighbours' x.phi, y1.phi, a4.pi and a3.pi come from the LDS slot of plane
//        j (written by their owners one iteration earlier), the a-neighbours from the
//        thread's own registers. w(j) goes to a ring of three LDS plane slots, s(j)
//        stays in registers.
//   barrier
//   B(j-1) on the tile: half B at plane j-1, with w's in-plane neighbours from the
//        slot of plane j-1 and its a-neighbours (j-2, j) from registers; writes lambda.
// One barrier per plane suffices: A(j) rewrites the q slot that A(j-1) read and the w
// slot that B(j-3) read, both before the previous barrier. s and w never reach HBM;
// the step reads x and lambda' (2 x 16 B/site, times the region/tile area ratio) and
// writes lambda (16 B/site). part: the dL/dm partials of stages 4, 3, 2, 1 of this
// block, ADJ_BLOCKS apart (block 0 zeroes the slots of blocks beyond the grid). lo
// must not overlap x or lp.
template <bool LAM>
__global__ __launch_bounds__(ADJ_FT) void k_adj_step(AdjGeom g, AdjStep k, const double *__restrict__ x,
                                                     const double *__restrict__ lp, double *__restrict__ lo,
                                                     double *__restrict__ part) {
  __shared__ AdjLds S;
  const int64_t N = g.N, NN = N * N, N3 = NN * N;
  const int me = (int)threadIdx.x, tx = me & (ADJ_R - 1), ty = me / ADJ_R;
  const double c_m = k.r1.c_m, lam = k.r1.lam;
  const bool in_a = tx >= 1 && tx <= ADJ_R - 2 && ty >= 1 && ty <= ADJ_R - 2;
  // in-plane neighbours in a slot, in the rhs_pi argument order: b+1, b-1, c+1, c-1
  const int o0 = me + ADJ_R, o1 = me - ADJ_R, o2 = me + 1, o3 = me - 1;
  auto y1phi = [&](double f, double p) { return f + k.dh * p; };
  auto a4pi = [&](double lpi) { return k.dt6 * lpi; };
  auto a3pi = [&](double lphi, double lpi) {
    const double yb4pi = k.dt6 * lphi + k.ch4 * a4pi(lpi);
    return k.dt3 * lpi + k.dt * yb4pi;
  };
  double dm4 = 0.0, dm3 = 0.0, dm2 = 0.0, dm1 = 0.0;
  for (int v = blockIdx.x; v < g.total; v += gridDim.x) {
    const int t = adj_tile(v, g.total);
    const int ct = t % g.nt, r = t / g.nt, bt = r % g.nt, ac = r / g.nt;
    const int64_t b0 = (int64_t)bt * ADJ_T, c0 = (int64_t)ct * ADJ_T;
    const int64_t a0 = (int64_t)ac * g.A, a1 = (a0 + g.A < N) ? a0 + g.A : N;
    const int64_t nb_ = (N - b0 < ADJ_T) ? N - b0 : ADJ_T, nc_ = (N - c0 < ADJ_T) ? N - c0 : ADJ_T;
    const bool out = ty >= 2 && ty < 2 + nb_ && tx >= 2 && tx < 2 + nc_;
    const int64_t b = ((b0 - 2 + ty) % N + N) % N, c = ((c0 - 2 + tx) % N + N) % N;
    const int64_t col = b * N + c;
    auto site = [&](int64_t a) { return (a < 0 ? a + N : (a >= N ? a - N : a)) * NN + col; };
    // own column: plane j-1 (p), j (c), j+1 (n, loaded), j+2 (loading)
    int64_t i = site(a0 - 2);
    double xp = x[i], yp = y1phi(xp, x[N3 + i]);
    double fp = a4pi(lp[N3 + i]), gp = a3pi(lp[i], lp[N3 + i]);
    i = site(a0 - 1);
    double xc = x[i], xpic = x[N3 + i], lphic = lp[i], lpic = lp[N3 + i];
    double yc = y1phi(xc, xpic), fc = a4pi(lpic), gc = a3pi(lphic, lpic);
    i = site(a0);
    double nx = x[i], npi = x[N3 + i], nlphi = lp[i], nlpi = lp[N3 + i];
    S.q[0][0][me] = xc;
    S.q[0][1][me] = yc;
    S.q[0][2][me] = fc;
    S.q[0][3][me] = gc;
    // B's history: w at j-2, j-1 and s, lambda'.phi at j-1
    double um2 = 0.0, um1 = 0.0, vm2 = 0.0, vm1 = 0.0, sphi_p = 0.0, spi_p = 0.0, lphi_p = 0.0;
    adj_barrier();
    int sq = 0, sw = 0, sw_p = 2;  // slot of plane j in q and w; w slot of plane j-1
    for (int64_t j = a0 - 1; j <= a1; ++j) {
      const double xn = nx, xpin = npi, lphin = nlphi, lpin = nlpi;
      if (j + 2 <= a1 + 1) {  // plane j+2
        i = site(j + 2);
        nx = x[i];
        npi = x[N3 + i];
        nlphi = lp[i];
        nlpi = lp[N3 + i];
      }
      const double yn = y1phi(xn, xpin), fn = a4pi(lpin), gn = a3pi(lphin, lpin);
      S.q[sq ^ 1][0][me] = xn;
      S.q[sq ^ 1][1][me] = yn;
      S.q[sq ^ 1][2][me] = fn;
      S.q[sq ^ 1][3][me] = gn;
      double sphi = 0.0, spi = 0.0, w0 = 0.0, w1 = 0.0;
      if (in_a) {  // A(j)
        const double *Q = S.q[sq][0];
        const double *Y = S.q[sq][1];
        const double *F4 = S.q[sq][2];
        const double *G3 = S.q[sq][3];
        // stage states at the site
        const double y1pi = xpic + k.dh * rhs_pi(xc, xn, xp, Q[o0], Q[o1], Q[o2], Q[o3], xpic, k.r1, LAM);
        const double y2pi = xpic + k.dh * rhs_pi(yc, yn, yp, Y[o0], Y[o1], Y[o2], Y[o3], y1pi, k.r2, LAM);
        const double y2phi = xc + k.dh * y1pi;
        const double y3phi = xc + k.dt * y2pi;
        // stage 4 at y3 (t + dt)
        const double nb4 = fn + fp + F4[o0] + F4[o1] + F4[o2] + F4[o3];
        const double yb4phi = jt_phi(c_m, lam, y3phi, k.inv4, fc, nb4);
        const double yb4pi = k.dt6 * lphic + k.ch4 * fc;
        // stage 3 at y2 (t + dt/2)
        const double nb3 = gn + gp + G3[o0] + G3[o1] + G3[o2] + G3[o3];
        const double yb3phi = jt_phi(c_m, lam, y2phi, k.r2.inv, gc, nb3);
        const double yb3pi = (k.dt3 * lphic + k.dt * yb4phi) + k.r2.c_h * gc;
        // stage 2 at y1 (t + dt/2): the local part
        const double a2pi = k.dt3 * lpic + k.dh * yb3pi;
        const double a2phi = k.dt3 * lphic + k.dh * yb3phi;
        const double yb2pi = a2phi + k.r2.c_h * a2pi;
        if (out && j >= a0 && j < a1) {
          dm4 += (-k.two_m * y3phi) * fc;
          dm3 += (-k.two_m * y2phi) * gc;
          dm2 += (-k.two_m * yc) * a2pi;
        }
        sphi = (lphic + yb4phi) + yb3phi;
        spi = ((lpic + yb4pi) + yb3pi) + yb2pi;
        w0 = a2pi;
        w1 = k.dt6 * lpic + k.dh * yb2pi;
        S.w[sw][0][me] = w0;
        S.w[sw][1][me] = w1;
      }
      adj_barrier();
      if (out && j > a0) {  // B(j-1): x.phi = xp, y1.phi = yp at j-1
        const double *U = S.w[sw_p][0];
        const double *V = S.w[sw_p][1];
        const double nb2 = w0 + um2 + U[o0] + U[o1] + U[o2] + U[o3];
        const double yb2phi = jt_phi(c_m, lam, yp, k.r2.inv, um1, nb2);
        const double nb1 = w1 + vm2 + V[o0] + V[o1] + V[o2] + V[o3];
        const double outphi = jt_phi(c_m, lam, xp, k.r1.inv, vm1, nb1);
        const double outpi = (k.dt6 * lphi_p + k.dh * yb2phi) + k.r1.c_h * vm1;
        const int64_t io = site(j - 1);
        lo[io] = (sphi_p + yb2phi) + outphi;
        lo[N3 + io] = spi_p + outpi;
        dm1 += (-k.two_m * xp) * vm1;
      }
      um2 = um1; um1 = w0;
      vm2 = vm1; vm1 = w1;
      sphi_p = sphi; spi_p = spi; lphi_p = lphic;
      xp = xc; xc = xn; xpic = xpin; lphic = lphin; lpic = lpin;
      yp = yc; yc = yn;
      fp = fc; fc = fn;
      gp = gc; gc = gn;
      sq ^= 1;
      sw_p = sw;
      sw = (sw == 2) ? 0 : sw + 1;
    }
  }
  block_partial<ADJ_FT>(dm4, part + blockIdx.x, me, S.red);
  block_partial<ADJ_FT>(dm3, part + ADJ_BLOCKS + blockIdx.x, me, S.red);
  block_partial<ADJ_FT>(dm2, part + 2 * ADJ_BLOCKS + blockIdx.x, me, S.red);
  block_partial<ADJ_FT>(dm1, part + 3 * ADJ_BLOCKS + blockIdx.x, me, S.red);
  if (blockIdx.x == 0)
    for (int sl = (int)gridDim.x + me; sl < ADJ_BLOCKS; sl += ADJ_FT)
      for (int st = 0; st < 4; ++st) part[st * ADJ_BLOCKS + sl] = 0.0;
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
  // checkpoints + one segment of states + lambda (2) + rk4 work (3)
  return nck * v + (size_t)ckpt * v + 2 * v + 3 * v +
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

// Chunk of the backward step's march. One 1024-thread block runs per CU at a time
// (112 KB of LDS), so k chunks per tile cost about ceil(tiles k / CUs) rounds of
// (N/k + 3) plane iterations (two extra planes of half A and the prologue's loads);
// among the k within 2% of the cheapest the largest is taken (N=128: 25 tiles x 10
// chunks of 13 planes, one round on 256 CUs).
static AdjGeom adj_geom(int64_t N) {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    cus = 256;
  AdjGeom g;
  g.N = N;
  g.nt = (int)cdiv(N, ADJ_T);
  const int64_t tiles = (int64_t)g.nt * g.nt;
  std::vector<int64_t> cost((size_t)N + 1, -1);
  int64_t best = -1;
  for (int64_t k = 1; k <= N; ++k) {
    const int64_t ch = cdiv(N, k);
    if (cdiv(N, ch) != k) continue;  // chunks actually launched for this chunk length
    cost[k] = cdiv(tiles * k, cus) * (ch + 3);
    if (best < 0 || cost[k] < best) best = cost[k];
  }
  int64_t pick = 1;
  for (int64_t k = 1; k <= N; ++k)
    if (cost[k] >= 0 && cost[k] * 100 <= best * 102) pick = k;
  g.A = (int)cdiv(N, pick);
  g.total = (int)(tiles * cdiv(N, g.A));
  return g;
}

// One backward step from state x at time t: lo = dL/dx_n from li = dL/dx_{n+1};
// part: 4*ADJ_BLOCKS partials of dL/dm (stages 4, 3, 2, 1). lo overlaps neither x
// nor li.
static cf_status adj_step(const cf_kg_params *p, const double *x, double t, double dt, const double *li, double *lo,
                          double *part, hipStream_t st) {
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
  const AdjGeom g = adj_geom(N);
  const dim3 grid(g.total < ADJ_BLOCKS ? g.total : ADJ_BLOCKS), blk(ADJ_FT);
  if (p->lambda != 0.0)
    hipLaunchKernelGGL(k_adj_step<true>, grid, blk, 0, st, g, k, x, li, lo, part);
  else
    hipLaunchKernelGGL(k_adj_step<false>, grid, blk, 0, st, g, k, x, li, lo, part);
  CF_LAUNCHED();
  return CF_OK;
}

// One RK4 step adjoint at state x_n (time t): lam_dev holds dL/dx_{n+1} on entry and
// dL/dx_n on return; the step's dL/dm contributions go to part_dev (4*ADJ_BLOCKS
// doubles, summed by the caller). work_dev: 2N^3 doubles.
cf_status cf_rk4_step_adjoint_f64(const cf_kg_params *p, const double *x, double t, double dt, double *lam,
                                  double *part_dev, void *work_dev, cf_stream stream) {
  CF_REQUIRE(p && x && lam && part_dev && work_dev, "cf_rk4_step_adjoint_f64: null argument");
  CF_REQUIRE(valid_N(p->N), "cf_rk4_step_adjoint_f64: unsupported N");
  const int64_t N = p->N, n = 2 * N * N * N;
  hipStream_t st = S(stream);
  double *lo = (double *)work_dev;
  CF_TRY(adj_step(p, x, t, dt, lam, lo, part_dev, st));
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
  double *lamb = seg + (size_t)ckpt * n;           // lambda, two buffers taken in turn
  double *lamb2 = lamb + n;
  double *fw = lamb2 + n;                          // forward work: 3 vectors
  double *part = fw + 3 * n;                       // (4 n_steps + 1) * ADJ_BLOCKS
  // forward with checkpoints; the last segment's states are kept in seg as they are
  // made, so the backward pass does not recompute them
  double *x = seg;  // running state
  CF_HIP(hipMemcpyAsync(x, x0, vb, hipMemcpyDeviceToDevice, st));
  for (int64_t c = 0; c < nck; ++c) {
    const int64_t k0 = c * ckpt, k1 = (k0 + ckpt < n_steps) ? k0 + ckpt : n_steps;
    if (c + 1 < nck) {
      CF_HIP(hipMemcpyAsync(ck + (size_t)c * n, x, vb, hipMemcpyDeviceToDevice, st));
      CF_TRY(cf_rk4_run_tb_f64(p, x, t0, k0, k1 - k0, dt, 0, fw, stream));  // same bits as the stage kernels
    } else {
      // seg[k - k0] -> seg[k + 1 - k0]; the final state goes to lamb2, which the loss
      // adjoint leaves alone and the backward pass overwrites only after the loss
      for (int64_t k = k0; k < k1; ++k) {
        double *nxt = (k + 1 < k1) ? seg + (size_t)(k + 1 - k0) * n : lamb2;
        CF_TRY(cf_rk4_step_tb_f64(p, seg + (size_t)(k - k0) * n, nxt, t0 + static_cast<double>(k) * dt, dt, 0,
                                  stream));
      }
      x = lamb2;
    }
  }
  const double t_end = t0 + static_cast<double>(n_steps) * dt;
  double dm_loss = 0.0;
  CF_TRY(cf_loss_spectrum_adjoint(plan, p, x, t_end, weights_dev, lamb, fw, loss, &dm_loss, stream));
  // backward, one segment at a time
  for (int64_t c = nck - 1; c >= 0; --c) {
    const int64_t k0 = c * ckpt, k1 = (k0 + ckpt < n_steps) ? k0 + ckpt : n_steps;
    if (c + 1 < nck) CF_HIP(hipMemcpyAsync(seg, ck + (size_t)c * n, vb, hipMemcpyDeviceToDevice, st));
    for (int64_t k = k0 + 1; k < k1 && c + 1 < nck; ++k) {
      const double *prev = seg + (size_t)(k - 1 - k0) * n;
      double *cur = seg + (size_t)(k - k0) * n;
      CF_TRY(cf_rk4_step_tb_f64(p, prev, cur, t0 + static_cast<double>(k - 1) * dt, dt, 0, stream));
    }
    for (int64_t k = k1 - 1; k >= k0; --k) {
      const double t = t0 + static_cast<double>(k) * dt;
      CF_TRY(adj_step(p, seg + (size_t)(k - k0) * n, t, dt, lamb, lamb2, part + (size_t)k * 4 * ADJ_BLOCKS, st));
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
