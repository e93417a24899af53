// cf_rk4_tb.hip -- temporally blocked classic RK4 on one lattice: all four
// stages of a step in one kernel, so the state is read once and written once
// per step (the step body, its geometry and its arithmetic are described in
// cf_tb_body.hpp). One block per work unit (output tile x a-chunk).
//
// HBM per site-step: FP64 read x, write x' = 32 B; FP32-comp read hi, lo, write
// hi', lo' = 32 B (plus halo and warm-up re-reads, mostly served by L2).
// The output goes to a second buffer (neighbouring blocks still read the old
// state), so runs ping-pong between two state arrays.
#include <cmath>

#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "cf_tb_body.hpp"
#include "cf_tb_ring.hpp"
#include "cf_tb_run.hpp"

namespace cf {

struct TbGeom {
  int64_t N;
  int Tb, Tc;        // output tile edges along b and c
  int ntb, ntc;      // tiles along b and c
  int A;             // planes per chunk
  int n_ac;          // chunks
  int total;         // ntb * ntc * n_ac
  int64_t NA;        // planes of the state buffers along a (N, or a slab's planes + ghosts)
  int64_t a_lo, a_hi;  // output planes [a_lo, a_hi)
  int pace;            // ticks of the 100 MHz real-time counter per plane iteration, 0: unpaced
  int rb, rc;          // region shape (Tb, Tc plus the halo) the launch was planned for
};

// Work unit of logical block `bid`: tile (tb, tc), chunk ac, c fastest.
__device__ __forceinline__ TbUnit tb_unit_of(const TbGeom &g, int bid) {
  const int tc = bid % g.ntc;
  const int rest = bid / g.ntc;
  const int tb = rest % g.ntb;
  const int ac = rest / g.ntb;
  TbUnit u;
  u.N = g.N;
  u.T = g.Tb;
  u.Tc = g.Tc;
  u.b0 = (int64_t)tb * g.Tb;
  u.c0 = (int64_t)tc * g.Tc;
  u.a0 = g.a_lo + (int64_t)ac * g.A;
  u.a1 = (u.a0 + g.A < g.a_hi) ? u.a0 + g.A : g.a_hi;
  u.NA = g.NA;
  u.pace = g.pace;
  return u;
}

// Region shape (b rows x c columns) of the FP32-comp single-lattice kernel. Measured
// at N=512 (GSite*step/s, same bits): 32 x 32 (24 x 24 tiles) 87.8-88.3; 24 x 40 (16 x 32
// tiles, whose c edges fall on 128-byte lines; 960 threads) 83.0-83.4; 25 x 40 81.4-82.1.
constexpr int TB_F32C_RB = 32, TB_F32C_RC = 32;
// Region shape of the single-precision-stencil kernel: 24 x 40 (16 x 32 tiles, 960 of the
// 1024 threads; the tiles' c edges fall on 128-byte lines, and N=512 makes 512 tiles).
// Measured at N=512 with 256-plane chunks (1024 blocks, four full rounds over 256 CUs;
// GSite*step/s, same bits): unpaced 101.7 against 76.7 for 32 x 32, paced 107.0 (at 115
// ticks per plane) against 97.3 (at 128).
constexpr int TB_F32S_RB = 24, TB_F32S_RC = 40;

template <class T, bool F32C, bool LAM>
__global__ __launch_bounds__(1024) void k_rk4_tb(TbGeom g, TbStep k, const T *__restrict__ xin,
                                                 const float *__restrict__ lin, T *__restrict__ xout,
                                                 float *__restrict__ lout) {
  __shared__ TbLds<T> lds;
  // Blocks go to the 8 XCDs round robin; consecutive logical tiles are mapped
  // to one XCD so that neighbouring tiles' halos meet in its L2.
  const TbUnit u = tb_unit_of(g, tb_xcd_order(blockIdx.x, g.total));
  // Lane shifts for the c neighbours in FP64; LDS in the compensated mode, whose
  // float plane buffers are cheap and whose VALU load is the larger (measured).
  constexpr int RB = F32C ? TB_F32C_RB : TBR, RC = F32C ? TB_F32C_RC : TBR;
  tb_unit<T, F32C, LAM, true, !F32C && RC == TBR, TB_KG, false, RB, RC>(lds, u, k, xin, lin, xout, lout);
}

// The compensated storage with the single-precision stencil (float forces, double
// increment): the same body as the FP32-comp kernel over 24 x 40 regions; PACE: paced
// (cf_tb_body.hpp), at g.pace ticks per plane iteration. Occupancy stays at one
// 1024-thread block per CU, so the compiler may use the 128 VGPRs that allows: with
// its default target it kept 84 and serialised the LDS neighbour reads (read, wait,
// add, read), 77.6 against 80.4 GSite*step/s unpaced at N=512.
template <bool LAM, bool PACE, int RB = TB_F32S_RB, int RC = TB_F32S_RC>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_rk4_tb_sa(
    TbGeom g, TbStep k, TbStepF kf, const float *__restrict__ xin, const float *__restrict__ lin,
    float *__restrict__ xout, float *__restrict__ lout) {
  __shared__ TbLds<float> lds;
  const TbUnit u = tb_unit_of(g, tb_xcd_order(blockIdx.x, g.total));
  tb_unit<float, true, LAM, true, false, TB_KG, false, RB, RC, true, PACE>(lds, u, k, xin, lin, xout, lout, nullptr,
                                                                         TbPsi{}, kf);
}

// FP64: the same step with the state streamed into an LDS ring TB_RING_DEPTH planes
// ahead (cf_tb_ring.hpp; same bits as the register-prefetch body).
constexpr int TB_RING_DEPTH = 4;

template <bool LAM>
__global__ __launch_bounds__(1024) void k_rk4_tb_ring(TbGeom g, TbStep k, const double *__restrict__ xin,
                                                      double *__restrict__ xout) {
  __shared__ __attribute__((aligned(16))) char smem[TbRing<TB_RING_DEPTH>::BYTES];
  const TbUnit u = tb_unit_of(g, tb_xcd_order(blockIdx.x, g.total));
  tb_unit_ring<LAM, TB_RING_DEPTH>(smem, u, k, xin, xout);
}

// The same step for the fixed-curvature (Psi, N^3) and sqrt-potential equation
// variants, FP64.
template <int VAR>
__global__ __launch_bounds__(1024) void k_rk4_tb_var(TbGeom g, TbStep k, const double *__restrict__ xin,
                                                     double *__restrict__ xout, const double *__restrict__ psi,
                                                     TbPsi cps) {
  __shared__ TbLds<double> lds;
  const TbUnit u = tb_unit_of(g, tb_xcd_order(blockIdx.x, g.total));
  tb_unit<double, false, false, true, true, VAR, true>(lds, u, k, xin, nullptr, xout, nullptr, psi, cps);
}

// Tiles of (RB - 8) x (RC - 8) sites (the whole extent when N is smaller) over the
// output planes [a_lo, a_hi) of buffers of NA planes; a periodic lattice: 0, N, N.
static TbGeom tb_geom(int64_t N, int A, int RB = TBR, int RC = TBR, int64_t a_lo = 0, int64_t a_hi = -1,
                      int64_t NA = -1) {
  TbGeom g;
  g.N = N;
  g.a_lo = a_lo;
  g.a_hi = a_hi < 0 ? N : a_hi;
  g.NA = NA < 0 ? N : NA;
  const int64_t planes = g.a_hi - g.a_lo;
  g.Tb = N < RB - 2 * TBH ? (int)N : RB - 2 * TBH;
  g.Tc = N < RC - 2 * TBH ? (int)N : RC - 2 * TBH;
  g.ntb = (int)cdiv(N, g.Tb);
  g.ntc = (int)cdiv(N, g.Tc);
  g.A = (A <= 0 || A > planes) ? (int)planes : A;
  g.n_ac = (int)cdiv(planes, g.A);
  g.total = g.ntb * g.ntc * g.n_ac;
  g.pace = 0;
  g.rb = RB;
  g.rc = RC;
  return g;
}

// Default planes per block along a. One 1024-thread block runs per CU at a time,
// so k chunks per tile cost about ceil(tiles k / CUs) rounds of (N/k + 8) plane
// iterations (8 warm-up planes per chunk). Too few blocks leave CUs idle (N=256 with
// one chunk: 121 blocks for 256 CUs), too many pay warm-ups. Among the k within 2%
// of the cheapest, the largest is taken. Measured on MI355X (FP64 / FP32-comp
// GSite*step/s): N=128 chunk 19 66 / 55 against 16 / 12 with one chunk; N=256 chunk
// 128 103 / 89 against 65 / 51; N=384 one chunk 118 / 88 against 92 / 74 with 256;
// N=512 chunk 256 (the rule's pick) and 171 within 1-3% of each other, one chunk
// 96 / 77.
static int tb_default_chunk(int64_t N, int RB = TBR, int RC = TBR, int64_t planes = -1,
                            int64_t *cost_out = nullptr, std::vector<int> *near = nullptr, int near_pct = 0) {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    cus = 256;
  if (planes < 0) planes = N;  // output planes of the launch
  const TbGeom g1 = tb_geom(N, 0, RB, RC);
  const int64_t tiles = (int64_t)g1.ntb * g1.ntc;
  const int64_t kmax = planes / 8 > 1 ? planes / 8 : 1;
  std::vector<int64_t> cost((size_t)kmax + 1, -1);
  int64_t best = -1;
  for (int64_t k = 1; k <= kmax; ++k) {
    const int64_t ch = cdiv(planes, k);
    const int64_t kk = cdiv(planes, ch);  // chunks actually launched for this chunk length
    if (kk != k) continue;
    cost[k] = cdiv(tiles * k, cus) * (ch + 8);
    if (best < 0 || cost[k] < best) best = cost[k];
  }
  int64_t pick = 1;
  for (int64_t k = 1; k <= kmax; ++k)
    if (cost[k] >= 0 && cost[k] * 100 <= best * 102) pick = k;
  if (cost_out) *cost_out = best;
  if (near)  // every chunk length within near_pct percent of the best cost, longest first
    for (int64_t k = 1; k <= kmax; ++k)
      if (cost[k] >= 0 && cost[k] * 100 <= best * (100 + near_pct)) near->push_back((int)cdiv(planes, k));
  return (int)cdiv(planes, pick);
}

// Where the FP32-comp run puts its second hi and lo arrays in the work buffer: at the
// first address from `from` on with the same residue modulo TB_ALIGN_SLACK as the
// caller's array `like`. Measured at N=512: hi, lo and the second pair at relative
// offsets that are not multiples of 4 KB (one allocation carved 256 B - 2 KB apart)
// ran 72-83 GSite*step/s against 84-90 at multiples of 4 KB. Matching residues keeps
// each array's offset to its partner of the other step a multiple of 4 KB; the work
// buffer carries the slack (cf_rk4_tb_work_bytes).
constexpr size_t TB_ALIGN_SLACK = 4096;
static float *same_residue(void *from, const float *like) {
  const uintptr_t f = reinterpret_cast<uintptr_t>(from), l = reinterpret_cast<uintptr_t>(like);
  return reinterpret_cast<float *>(f + ((l - f) & (TB_ALIGN_SLACK - 1)));
}

// Geometry of the FP32-comp kernel (its region shape), chunk <= 0: the cost model's.
static TbGeom tb_f32c_geom(int64_t N, int chunk) {
  return tb_geom(N, chunk > 0 ? chunk : tb_default_chunk(N, TB_F32C_RB, TB_F32C_RC), TB_F32C_RB, TB_F32C_RC);
}
// Region shape of the single-precision-stencil kernel for `planes` output planes of an
// N^2 cross-section: 24 x 40 unless the cost model's launch over 32 x 32 regions is more
// than TB_F32S_SQUARE_GAIN percent cheaper (at N=192 and 384 its 8 x 8 / 16 x 16 tiles
// fill the CUs in one round where 24 x 40 needs more rounds of shorter chunks; measured
// at N=384: 105.7 GSite*step/s against 97.3). Elsewhere 24 x 40 is faster at equal cost
// (N=256: 109.8 against 95.4; 768: 108.3 against 92.6, a cost ratio of 1.14).
constexpr int TB_F32S_SQUARE_GAIN = 20;
static void tb_f32s_shape(int64_t N, int64_t planes, int *RB, int *RC) {
  int64_t c_strip = 0, c_square = 0;
  tb_default_chunk(N, TB_F32S_RB, TB_F32S_RC, planes, &c_strip);
  tb_default_chunk(N, TB_F32C_RB, TB_F32C_RC, planes, &c_square);
  const bool square = c_square * (100 + TB_F32S_SQUARE_GAIN) < c_strip * 100;
  *RB = square ? TB_F32C_RB : TB_F32S_RB;
  *RC = square ? TB_F32C_RC : TB_F32S_RC;
}
// Default chunk of the single-precision-stencil step per device and lattice size, set by
// cf_rk4_tb_tune_pace (chunk <= 0) or cf_rk4_tb_set_chunk; absent: the cost model's.
static std::mutex g_chunk_mu;
static std::map<std::pair<int, int64_t>, int> g_chunk;

static int tb_f32s_tuned_chunk(int64_t N) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  std::lock_guard<std::mutex> lock(g_chunk_mu);
  const auto it = g_chunk.find(std::make_pair(dev, N));
  return it == g_chunk.end() ? 0 : it->second;
}

// ... and of the single-precision-stencil kernel (sa) or the FP32-comp one.
static TbGeom tb_comp_geom(bool sa, int64_t N, int chunk) {
  if (!sa) return tb_f32c_geom(N, chunk);
  int RB = 0, RC = 0;
  tb_f32s_shape(N, N, &RB, &RC);
  if (chunk <= 0) chunk = tb_f32s_tuned_chunk(N);
  return tb_geom(N, chunk > 0 ? chunk : tb_default_chunk(N, RB, RC), RB, RC);
}

template <class T, bool F32C>
static cf_status tb_launch(const TbGeom &g, const TbStep &k, const T *xin, const float *lin, T *xout,
                           float *lout, hipStream_t st) {
  if constexpr (!F32C) {
    if (k.lam != 0.0)
      hipLaunchKernelGGL(k_rk4_tb_ring<true>, dim3(g.total), dim3(1024), 0, st, g, k, xin, xout);
    else
      hipLaunchKernelGGL(k_rk4_tb_ring<false>, dim3(g.total), dim3(1024), 0, st, g, k, xin, xout);
  } else {
    if (k.lam != 0.0)
      hipLaunchKernelGGL((k_rk4_tb<T, F32C, true>), dim3(g.total), dim3(1024), 0, st, g, k, xin, lin, xout, lout);
    else
      hipLaunchKernelGGL((k_rk4_tb<T, F32C, false>), dim3(g.total), dim3(1024), 0, st, g, k, xin, lin, xout, lout);
  }
  CF_LAUNCHED();
  return CF_OK;
}

// The single-precision stencil's scalars: TbStep's RHS coefficients and stage-input
// steps rounded to float.
static TbStepF tb_step_f32(const TbStep &k) {
  TbStepF f;
  for (int s = 0; s < 4; ++s) {
    f.c_h[s] = (float)k.c_h[s];
    f.inv[s] = (float)k.inv[s];
  }
  f.c_m = (float)k.c_m;
  f.lam = (float)k.lam;
  f.dh = (float)k.dh;
  f.dt = (float)k.dt;
  return f;
}

template <int RB, int RC>
static void tb_launch_sa_shape(const TbGeom &g, const TbStep &k, const TbStepF &kf, const float *xin,
                               const float *lin, float *xout, float *lout, hipStream_t st) {
  const bool lam = k.lam != 0.0, paced = g.pace > 0;
  const dim3 grid(g.total), block(1024);
  if (lam && paced)
    hipLaunchKernelGGL((k_rk4_tb_sa<true, true, RB, RC>), grid, block, 0, st, g, k, kf, xin, lin, xout, lout);
  else if (lam)
    hipLaunchKernelGGL((k_rk4_tb_sa<true, false, RB, RC>), grid, block, 0, st, g, k, kf, xin, lin, xout, lout);
  else if (paced)
    hipLaunchKernelGGL((k_rk4_tb_sa<false, true, RB, RC>), grid, block, 0, st, g, k, kf, xin, lin, xout, lout);
  else
    hipLaunchKernelGGL((k_rk4_tb_sa<false, false, RB, RC>), grid, block, 0, st, g, k, kf, xin, lin, xout, lout);
}

static cf_status tb_launch_sa(const TbGeom &g, const TbStep &k, const float *xin, const float *lin, float *xout,
                              float *lout, hipStream_t st) {
  const TbStepF kf = tb_step_f32(k);
  if (g.rb == TB_F32S_RB && g.rc == TB_F32S_RC)
    tb_launch_sa_shape<TB_F32S_RB, TB_F32S_RC>(g, k, kf, xin, lin, xout, lout, st);
  else if (g.rb == TB_F32C_RB && g.rc == TB_F32C_RC)
    tb_launch_sa_shape<TB_F32C_RB, TB_F32C_RC>(g, k, kf, xin, lin, xout, lout, st);
  else
    CF_REQUIRE(false, "single-precision stencil: no kernel for %d x %d regions", g.rb, g.rc);
  CF_LAUNCHED();
  return CF_OK;
}

// [a, a + na) and [b, b + nb) do not overlap.
static bool disjoint(const void *a, size_t na, const void *b, size_t nb) {
  const char *x = static_cast<const char *>(a), *y = static_cast<const char *>(b);
  return x + na <= y || y + nb <= x;
}
static bool disjoint(const void *a, const void *b, size_t bytes) { return disjoint(a, bytes, b, bytes); }

static cf_status tb_check(const cf_kg_params *p) {
  CF_REQUIRE(p, "null cf_kg_params");
  CF_REQUIRE(valid_N(p->N), "unsupported lattice size N=%lld (even, 4..2048)", (long long)p->N);
  CF_REQUIRE(p->L > 0.0, "box size L must be positive");
  return CF_OK;
}

cf_status tb_step_variant(int var, const cf_kg_params *p, double f_a, const double *psi, const double *xin,
                          double *xout, double t, double dt, int chunk, hipStream_t st, const TbPsi *cps) {
  CF_TRY(tb_check(p));
  const bool comoving = var == TB_COMOVING || var == TB_APPROX_COMOVING;
  CF_REQUIRE(var == TB_SQRT_POTENTIAL || (var == TB_FIXED_CURVATURE && psi) ||
                 (comoving && cps && cps->f[0] && cps->f[1] && cps->f[2] && cps->G > 0),
             "tb_step_variant: bad variant");
  const size_t n = 2 * (size_t)p->N * p->N * p->N;
  CF_REQUIRE(xin && xout && disjoint(xin, xout, n * sizeof(double)),
             "temporally blocked variant step: null or overlapping state buffers");
  const TbGeom g = tb_geom(p->N, chunk > 0 ? chunk : tb_default_chunk(p->N));
  TbStep k = tb_step_host(p, t, dt);
  k.fa2 = f_a * f_a;
  const TbPsi cp = cps ? *cps : TbPsi{};
  if (var == TB_FIXED_CURVATURE)
    hipLaunchKernelGGL(k_rk4_tb_var<TB_FIXED_CURVATURE>, dim3(g.total), dim3(1024), 0, st, g, k, xin, xout, psi, cp);
  else if (var == TB_SQRT_POTENTIAL)
    hipLaunchKernelGGL(k_rk4_tb_var<TB_SQRT_POTENTIAL>, dim3(g.total), dim3(1024), 0, st, g, k, xin, xout, psi, cp);
  else if (var == TB_COMOVING)
    hipLaunchKernelGGL(k_rk4_tb_var<TB_COMOVING>, dim3(g.total), dim3(1024), 0, st, g, k, xin, xout, psi, cp);
  else
    hipLaunchKernelGGL(k_rk4_tb_var<TB_APPROX_COMOVING>, dim3(g.total), dim3(1024), 0, st, g, k, xin, xout, psi,
                       cp);
  CF_LAUNCHED();
  return CF_OK;
}

// Pacing of the single-precision-stencil kernel per device, lattice size and chunk (ticks
// of the 100 MHz real-time counter per plane iteration; absent or 0: unpaced), set by
// cf_rk4_tb_tune_pace or cf_rk4_tb_set_pace.
static std::mutex g_pace_mu;
using TbPaceKey = std::tuple<int, int64_t, int, int>;  // device, N, chunk, region rows
static std::map<TbPaceKey, int> g_pace;

static cf_status tb_pace_key(const TbGeom &g, TbPaceKey *key) {
  int dev = 0;
  CF_HIP(hipGetDevice(&dev));
  *key = std::make_tuple(dev, g.N, g.A, g.rb);
  return CF_OK;
}
static cf_status tb_pace_lookup(TbGeom &g) {
  TbPaceKey key;
  CF_TRY(tb_pace_key(g, &key));
  std::lock_guard<std::mutex> lock(g_pace_mu);
  const auto it = g_pace.find(key);
  g.pace = it == g_pace.end() ? 0 : it->second;
  return CF_OK;
}
static cf_status tb_pace_store(const TbGeom &g, int pace) {
  TbPaceKey key;
  CF_TRY(tb_pace_key(g, &key));
  std::lock_guard<std::mutex> lock(g_pace_mu);
  g_pace[key] = pace;
  return CF_OK;
}

// Runs and single steps of the compensated storage: FP64 arithmetic (sa = false) or
// the single-precision stencil (sa = true).
static cf_status tb_launch_comp(bool sa, const TbGeom &g, const TbStep &k, const float *xin, const float *lin,
                                float *xout, float *lout, hipStream_t st) {
  return sa ? tb_launch_sa(g, k, xin, lin, xout, lout, st) : tb_launch<float, true>(g, k, xin, lin, xout, lout, st);
}

static cf_status tb_run_comp(const char *name, bool sa, const cf_kg_params *p, float *hi, float *lo, double t0,
                             int64_t k0, int64_t n_steps, double dt, int chunk, void *work, hipStream_t st) {
  CF_TRY(tb_check(p));
  const size_t n = 2 * (size_t)p->N * p->N * p->N;
  const size_t wb = 2 * n * sizeof(float) + 2 * TB_ALIGN_SLACK;  // cf_rk4_tb_work_bytes
  CF_REQUIRE(hi && lo && work && disjoint(hi, lo, n * sizeof(float)) &&
                 disjoint(hi, n * sizeof(float), work, wb) && disjoint(lo, n * sizeof(float), work, wb),
             "%s: null or overlapping buffer", name);
  CF_REQUIRE(n_steps >= 0, "%s: negative step count", name);
  TbGeom g = tb_comp_geom(sa, p->N, chunk);
  if (sa) CF_TRY(tb_pace_lookup(g));
  float *his[2] = {hi, same_residue(work, hi)};
  float *los[2] = {lo, same_residue(his[1] + n, lo)};
  for (int64_t s = 0; s < n_steps; ++s) {
    const double t = t0 + static_cast<double>(k0 + s) * dt;
    CF_TRY(tb_launch_comp(sa, g, tb_step_host(p, t, dt), his[s & 1], los[s & 1], his[(s + 1) & 1],
                          los[(s + 1) & 1], st));
  }
  if (n_steps & 1) {
    CF_HIP(hipMemcpyAsync(hi, his[1], n * sizeof(float), hipMemcpyDeviceToDevice, st));
    CF_HIP(hipMemcpyAsync(lo, los[1], n * sizeof(float), hipMemcpyDeviceToDevice, st));
  }
  return CF_OK;
}

static cf_status tb_step_comp(const char *name, bool sa, const cf_kg_params *p, const float *hi_in,
                              const float *lo_in, float *hi_out, float *lo_out, double t, double dt, int chunk,
                              hipStream_t st) {
  CF_TRY(tb_check(p));
  const size_t b = 2 * (size_t)p->N * p->N * p->N * sizeof(float);
  CF_REQUIRE(hi_in && lo_in && hi_out && lo_out, "%s: null buffer", name);
  CF_REQUIRE(disjoint(hi_in, lo_in, b) && disjoint(hi_out, lo_out, b) && disjoint(hi_in, hi_out, b) &&
                 disjoint(hi_in, lo_out, b) && disjoint(lo_in, hi_out, b) && disjoint(lo_in, lo_out, b),
             "%s: overlapping buffers", name);
  TbGeom g = tb_comp_geom(sa, p->N, chunk);
  if (sa) CF_TRY(tb_pace_lookup(g));
  return tb_launch_comp(sa, g, tb_step_host(p, t, dt), hi_in, lo_in, hi_out, lo_out, st);
}

// Ping-pong buffer of the entry point that allocates nothing
// (cf_rk4_run_tb_f64_noalloc): 2 N^3 doubles for N <= CF_NOALLOC_MAX_N, one copy per
// GPU, held until the process ends. In host code `scratch` names a host-side
// stand-in: kernels and copies take the device address from hipGetSymbolAddress.
static __device__ __align__(256) double scratch[2 * CF_NOALLOC_MAX_N * CF_NOALLOC_MAX_N * CF_NOALLOC_MAX_N];

// scratch's device address on each device, looked up once per device with that
// device current.
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[64] = {};

static cf_status scratch_on_current_device(double **out) {
  int dev = 0;
  CF_HIP(hipGetDevice(&dev));
  CF_REQUIRE(dev >= 0 && dev < 64, "cf_rk4_run_tb_f64_noalloc: device ordinal %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    CF_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    CF_REQUIRE(ptr, "cf_rk4_run_tb_f64_noalloc: no device address for the scratch buffer");
    g_scratch_ptr[dev] = ptr;
  }
  *out = static_cast<double *>(g_scratch_ptr[dev]);
  return CF_OK;
}

static cf_status tb_step_comp_slab(const char *name, bool sa, const cf_kg_params *p, int64_t n_planes,
                                   const float *hi_in, const float *lo_in, float *hi_out, float *lo_out, double t,
                                   double dt, int chunk, hipStream_t st) {
  CF_TRY(tb_check(p));
  CF_REQUIRE(n_planes >= 1 && n_planes <= p->N, "%s: n_planes=%lld outside 1..N", name, (long long)n_planes);
  const int64_t NA = n_planes + 2 * TBH;
  const size_t b = 2 * (size_t)NA * p->N * p->N * sizeof(float);
  CF_REQUIRE(hi_in && lo_in && hi_out && lo_out, "%s: null buffer", name);
  CF_REQUIRE(disjoint(hi_in, lo_in, b) && disjoint(hi_out, lo_out, b) && disjoint(hi_in, hi_out, b) &&
                 disjoint(hi_in, lo_out, b) && disjoint(lo_in, hi_out, b) && disjoint(lo_in, lo_out, b),
             "%s: overlapping buffers", name);
  int RB = TB_F32C_RB, RC = TB_F32C_RC;
  if (sa) tb_f32s_shape(p->N, n_planes, &RB, &RC);
  TbGeom g = tb_geom(p->N, chunk > 0 ? chunk : tb_default_chunk(p->N, RB, RC, n_planes), RB, RC, TBH,
                     TBH + n_planes, NA);
  if (sa) CF_TRY(tb_pace_lookup(g));  // the pace tuned for this lattice size and chunk, if any
  return tb_launch_comp(sa, g, tb_step_host(p, t, dt), hi_in, lo_in, hi_out, lo_out, st);
}

}  // namespace cf

using namespace cf;

extern "C" {

size_t cf_rk4_tb_work_bytes(int64_t N, int precision) {
  if (!valid_N(N)) return 0;
  const size_t n = 2 * (size_t)N * N * N;
  return (precision == CF_PREC_F32C || precision == CF_PREC_F32S) ? 2 * n * sizeof(float) + 2 * TB_ALIGN_SLACK : n * sizeof(double);
}

cf_status cf_rk4_run_tb_f64(const cf_kg_params *p, double *x, double t0, int64_t k0, int64_t n_steps, double dt,
                            int chunk, void *work, cf_stream stream) {
  CF_TRY(tb_check(p));
  CF_REQUIRE(x && work && disjoint(x, work, 2 * (size_t)p->N * p->N * p->N * sizeof(double)),
             "cf_rk4_run_tb_f64: null or overlapping buffer");
  CF_REQUIRE(n_steps >= 0, "cf_rk4_run_tb_f64: negative step count");
  const TbGeom g = tb_geom(p->N, chunk > 0 ? chunk : tb_default_chunk(p->N));
  double *bufs[2] = {x, (double *)work};
  for (int64_t s = 0; s < n_steps; ++s) {
    const double t = t0 + static_cast<double>(k0 + s) * dt;
    CF_TRY((tb_launch<double, false>(g, tb_step_host(p, t, dt), bufs[s & 1], nullptr, bufs[(s + 1) & 1], nullptr,
                                      S(stream))));
  }
  if (n_steps & 1) {
    const size_t bytes = 2 * (size_t)p->N * p->N * p->N * sizeof(double);
    CF_HIP(hipMemcpyAsync(x, work, bytes, hipMemcpyDeviceToDevice, S(stream)));
  }
  return CF_OK;
}

cf_status cf_rk4_run_tb_f64_noalloc(const cf_kg_params *p, double *x, double t0, int64_t k0, int64_t n_steps,
                                    double dt, int chunk, cf_stream stream) {
  CF_TRY(tb_check(p));
  CF_REQUIRE(p->N <= CF_NOALLOC_MAX_N,
             "cf_rk4_run_tb_f64_noalloc: N=%lld exceeds the preallocated buffer (N <= %d); use cf_rk4_run_tb_f64",
             (long long)p->N, CF_NOALLOC_MAX_N);
  CF_REQUIRE(x, "cf_rk4_run_tb_f64_noalloc: null state");
  CF_REQUIRE(n_steps >= 0, "cf_rk4_run_tb_f64_noalloc: negative step count");
  double *work = nullptr;
  CF_TRY(scratch_on_current_device(&work));
  const size_t bytes = 2 * (size_t)p->N * p->N * p->N * sizeof(double);
  CF_REQUIRE(disjoint(x, work, bytes), "cf_rk4_run_tb_f64_noalloc: the state overlaps the library's buffer");
  const TbGeom g = tb_geom(p->N, chunk > 0 ? chunk : tb_default_chunk(p->N));
  double *bufs[2] = {x, work};
  for (int64_t s = 0; s < n_steps; ++s) {
    const double t = t0 + static_cast<double>(k0 + s) * dt;
    CF_TRY((tb_launch<double, false>(g, tb_step_host(p, t, dt), bufs[s & 1], nullptr, bufs[(s + 1) & 1], nullptr,
                                      S(stream))));
  }
  if (n_steps & 1) CF_HIP(hipMemcpyAsync(x, work, bytes, hipMemcpyDeviceToDevice, S(stream)));
  return CF_OK;
}

cf_status cf_rk4_run_tb_f32c(const cf_kg_params *p, float *hi, float *lo, double t0, int64_t k0, int64_t n_steps,
                             double dt, int chunk, void *work, cf_stream stream) {
  return tb_run_comp("cf_rk4_run_tb_f32c", false, p, hi, lo, t0, k0, n_steps, dt, chunk, work, S(stream));
}

cf_status cf_rk4_run_tb_f32s(const cf_kg_params *p, float *hi, float *lo, double t0, int64_t k0, int64_t n_steps,
                             double dt, int chunk, void *work, cf_stream stream) {
  return tb_run_comp("cf_rk4_run_tb_f32s", true, p, hi, lo, t0, k0, n_steps, dt, chunk, work, S(stream));
}

cf_status cf_rk4_step_tb_f64(const cf_kg_params *p, const double *x_in, double *x_out, double t, double dt,
                             int chunk, cf_stream stream) {
  CF_TRY(tb_check(p));
  const size_t n = 2 * (size_t)p->N * p->N * p->N;
  CF_REQUIRE(x_in && x_out && disjoint(x_in, x_out, n * sizeof(double)),
             "cf_rk4_step_tb_f64: null or overlapping state buffers");
  const TbGeom g = tb_geom(p->N, chunk > 0 ? chunk : tb_default_chunk(p->N));
  return tb_launch<double, false>(g, tb_step_host(p, t, dt), x_in, nullptr, x_out, nullptr, S(stream));
}

cf_status cf_rk4_step_tb_f64_slab(const cf_kg_params *p, int64_t n_planes, const double *x_in, double *x_out,
                                  double t, double dt, int chunk, cf_stream stream) {
  CF_TRY(tb_check(p));
  CF_REQUIRE(n_planes >= 1 && n_planes <= p->N, "cf_rk4_step_tb_f64_slab: n_planes=%lld outside 1..N",
             (long long)n_planes);
  const int64_t NA = n_planes + 2 * TBH;  // own planes and four ghost planes on each side
  const size_t bytes = 2 * (size_t)NA * p->N * p->N * sizeof(double);
  CF_REQUIRE(x_in && x_out && disjoint(x_in, x_out, bytes), "cf_rk4_step_tb_f64_slab: null or overlapping slabs");
  // A unit reads planes a0 - 4 .. a1 + 3 of its chunk (wrapping at NA only in the ring's
  // dead prefetches past a1 + 3), all inside the slab for chunks within [4, 4 + n_planes).
  const TbGeom g = tb_geom(p->N, chunk > 0 ? chunk : tb_default_chunk(p->N, TBR, TBR, n_planes), TBR, TBR, TBH,
                           TBH + n_planes, NA);
  return tb_launch<double, false>(g, tb_step_host(p, t, dt), x_in, nullptr, x_out, nullptr, S(stream));
}


cf_status cf_rk4_step_tb_f32c(const cf_kg_params *p, const float *hi_in, const float *lo_in, float *hi_out,
                              float *lo_out, double t, double dt, int chunk, cf_stream stream) {
  return tb_step_comp("cf_rk4_step_tb_f32c", false, p, hi_in, lo_in, hi_out, lo_out, t, dt, chunk, S(stream));
}

cf_status cf_rk4_step_tb_f32s(const cf_kg_params *p, const float *hi_in, const float *lo_in, float *hi_out,
                              float *lo_out, double t, double dt, int chunk, cf_stream stream) {
  return tb_step_comp("cf_rk4_step_tb_f32s", true, p, hi_in, lo_in, hi_out, lo_out, t, dt, chunk, S(stream));
}

cf_status cf_rk4_step_tb_f32c_slab(const cf_kg_params *p, int64_t n_planes, const float *hi_in, const float *lo_in,
                                   float *hi_out, float *lo_out, double t, double dt, int chunk, cf_stream stream) {
  return tb_step_comp_slab("cf_rk4_step_tb_f32c_slab", false, p, n_planes, hi_in, lo_in, hi_out, lo_out, t, dt, chunk,
                           S(stream));
}

cf_status cf_rk4_step_tb_f32s_slab(const cf_kg_params *p, int64_t n_planes, const float *hi_in, const float *lo_in,
                                   float *hi_out, float *lo_out, double t, double dt, int chunk, cf_stream stream) {
  return tb_step_comp_slab("cf_rk4_step_tb_f32s_slab", true, p, n_planes, hi_in, lo_in, hi_out, lo_out, t, dt, chunk,
                           S(stream));
}

// Pacing of the single-precision-stencil step: time the step of this lattice at a
// range of paces and keep one with margin. Below the fastest pace the rate falls (over
// 32 x 32 regions it fell off a cliff: blocks that are not held back drift out of step;
// over the 24 x 40 regions the kernel uses now, by 1-3% per tick) and above it declines
// slowly, and where the fall begins moves with the clock: a pace one or two ticks above a
// cliff measured fast while tuning and then ran 10-15% slow. So each pace P is scored by
// the slowest of P - TB_PACE_MARGIN .. P, and the best score wins (at equal scores the
// larger pace); unpaced wins if it is as fast.
constexpr int TB_PACE_MARGIN = 2;

// Chunk lengths within this many percent of the cost model's best that the pace tuner
// tries when the caller leaves the chunk to it. At N=512 (24 x 40 regions) these are 512,
// 256, 171 and 128 planes; measured tuned: 109.0-110.4, 105.9-107.2, 108.4-109.3 and
// 106.9-108.5 GSite*step/s: the drift of the blocks, which the cost model does not
// see, depends on the chunk length and how the rounds of blocks line up.
constexpr int TB_CHUNK_NEAR_PCT = 5;

// Pace of launches over geometry g: the pick and its score (the slowest ms per step over
// the picked pace and the TB_PACE_MARGIN paces below it; the unpaced ms for pick 0).
static cf_status tb_tune_pace_geom(TbGeom g, const TbStep &k, const float *hi, const float *lo, float *ho,
                                   float *lw, hipStream_t st, hipEvent_t e0, hipEvent_t e1, int *pick_out,
                                   float *score_out) {
  std::map<int, float> ms;  // pace -> ms per step
  auto time_pace = [&](int pace) -> cf_status {
    if (ms.count(pace)) return CF_OK;
    g.pace = pace;
    CF_TRY(tb_launch_sa(g, k, hi, lo, ho, lw, st));  // warm-up
    float best = 1e30f;
    for (int w = 0; w < 2; ++w) {  // the faster of two windows of three launches
      CF_HIP(hipEventRecord(e0, st));
      for (int r = 0; r < 3; ++r) CF_TRY(tb_launch_sa(g, k, hi, lo, ho, lw, st));
      CF_HIP(hipEventRecord(e1, st));
      CF_HIP(hipEventSynchronize(e1));
      float el = 0.0f;
      CF_HIP(hipEventElapsedTime(&el, e0, e1));
      best = el / 3.0f < best ? el / 3.0f : best;
    }
    ms[pace] = best;
    return CF_OK;
  };
  // coarse grid 0.64 .. 3.2 us per plane, then every tick from below the coarse best to
  // the margin above it
  for (int pace = 0; pace <= 320; pace += (pace == 0 ? 64 : 8)) CF_TRY(time_pace(pace));
  int coarse = 0;
  for (const auto &kv : ms)
    if (kv.second < ms[coarse]) coarse = kv.first;
  if (coarse > 0)
    for (int pace = coarse - 8; pace <= coarse + 8 + TB_PACE_MARGIN; ++pace) CF_TRY(time_pace(pace));
  int pick = 0;
  float pick_worst = ms[0];
  for (const auto &kv : ms) {
    if (kv.first - TB_PACE_MARGIN < 1) continue;
    float worst = 0.0f;
    bool full = true;
    for (int q = kv.first - TB_PACE_MARGIN; q <= kv.first && full; ++q) {
      const auto it = ms.find(q);
      full = it != ms.end();
      if (full) worst = it->second > worst ? it->second : worst;
    }
    if (full && worst <= pick_worst && (pick > 0 || worst < ms[0])) {
      pick = kv.first;
      pick_worst = worst;
    }
  }
  *pick_out = pick;
  *score_out = pick_worst;
  return CF_OK;
}

cf_status cf_rk4_tb_tune_pace(const cf_kg_params *p, const float *hi, const float *lo, double t, double dt, int chunk,
                              void *work, cf_stream stream, int *pace_out) {
  CF_TRY(tb_check(p));
  const size_t n = 2 * (size_t)p->N * p->N * p->N;
  const size_t wb = 2 * n * sizeof(float) + 2 * TB_ALIGN_SLACK;  // cf_rk4_tb_work_bytes
  CF_REQUIRE(hi && lo && work && disjoint(hi, lo, n * sizeof(float)) &&
                 disjoint(hi, n * sizeof(float), work, wb) && disjoint(lo, n * sizeof(float), work, wb),
             "cf_rk4_tb_tune_pace: null or overlapping buffer");
  hipStream_t st = S(stream);
  int RB = 0, RC = 0;
  tb_f32s_shape(p->N, p->N, &RB, &RC);
  std::vector<int> chunks;
  if (chunk > 0)
    chunks.push_back(chunk);
  else
    tb_default_chunk(p->N, RB, RC, p->N, nullptr, &chunks, TB_CHUNK_NEAR_PCT);
  float *ho = same_residue(work, hi), *lw = same_residue(ho + n, lo);
  const TbStep k = tb_step_host(p, t, dt);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  CF_HIP(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    CF_REQUIRE(false, "cf_rk4_tb_tune_pace: hipEventCreate failed");
  }
  cf_status rc = CF_OK;
  std::vector<TbGeom> picks;
  for (size_t i = 0; i < chunks.size() && rc == CF_OK; ++i) {
    TbGeom g = tb_geom(p->N, chunks[i], RB, RC);
    float score = 0.0f;
    rc = tb_tune_pace_geom(g, k, hi, lo, ho, lw, st, e0, e1, &g.pace, &score);
    if (rc == CF_OK) rc = tb_pace_store(g, g.pace);
    picks.push_back(g);
  }
  // Between chunks: each at its pick, timed again over longer windows (the fastest of
  // three windows of five launches; a pace tuning's short windows and margin rank the
  // paces of one chunk, not chunks whose rate curves differ in steepness).
  auto time_pick = [&](const TbGeom &g, float *ms) -> cf_status {
    *ms = 1e30f;
    for (int w = 0; w < 3; ++w) {
      CF_TRY(tb_launch_sa(g, k, hi, lo, ho, lw, st));  // a gap-free start for the window
      CF_HIP(hipEventRecord(e0, st));
      for (int r = 0; r < 5; ++r) CF_TRY(tb_launch_sa(g, k, hi, lo, ho, lw, st));
      CF_HIP(hipEventRecord(e1, st));
      CF_HIP(hipEventSynchronize(e1));
      float el = 0.0f;
      CF_HIP(hipEventElapsedTime(&el, e0, e1));
      if (el / 5.0f < *ms) *ms = el / 5.0f;
    }
    return CF_OK;
  };
  int best_chunk = 0, best_pace = 0;
  float best_ms = 0.0f;
  for (size_t i = 0; i < picks.size() && rc == CF_OK && chunks.size() > 1; ++i) {
    float ms = 0.0f;
    rc = time_pick(picks[i], &ms);
    if (rc == CF_OK && (i == 0 || ms < best_ms)) {
      best_chunk = picks[i].A;
      best_pace = picks[i].pace;
      best_ms = ms;
    }
  }
  if (chunks.size() == 1 && rc == CF_OK) {
    best_chunk = picks[0].A;
    best_pace = picks[0].pace;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  CF_TRY(rc);
  if (chunk <= 0) {
    int dev = 0;
    CF_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_chunk_mu);
    g_chunk[std::make_pair(dev, p->N)] = best_chunk;
  }
  if (pace_out) *pace_out = best_pace;
  return CF_OK;
}

cf_status cf_rk4_tb_set_pace(int64_t N, int chunk, int pace) {
  CF_REQUIRE(valid_N(N), "cf_rk4_tb_set_pace: unsupported lattice size N=%lld", (long long)N);
  CF_REQUIRE(pace >= 0 && pace <= 100000, "cf_rk4_tb_set_pace: pace %d outside 0..100000", pace);
  return tb_pace_store(tb_comp_geom(true, N, chunk), pace);
}

cf_status cf_rk4_tb_set_chunk(int64_t N, int chunk) {
  CF_REQUIRE(valid_N(N), "cf_rk4_tb_set_chunk: unsupported lattice size N=%lld", (long long)N);
  CF_REQUIRE(chunk >= 0 && chunk <= N, "cf_rk4_tb_set_chunk: chunk %d outside 0..N", chunk);
  int dev = 0;
  CF_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_chunk_mu);
  if (chunk == 0)
    g_chunk.erase(std::make_pair(dev, N));
  else
    g_chunk[std::make_pair(dev, N)] = chunk;
  return CF_OK;
}

int cf_rk4_tb_get_chunk(int64_t N) {
  if (!valid_N(N)) return -1;
  return tb_comp_geom(true, N, 0).A;
}

cf_status cf_rk4_tb_f32s_region(int64_t N, int *rows, int *cols) {
  CF_REQUIRE(valid_N(N), "cf_rk4_tb_f32s_region: unsupported lattice size N=%lld", (long long)N);
  CF_REQUIRE(rows && cols, "cf_rk4_tb_f32s_region: null output");
  tb_f32s_shape(N, N, rows, cols);
  return CF_OK;
}

int cf_rk4_tb_get_pace(int64_t N, int chunk) {
  if (!valid_N(N)) return -1;
  TbGeom g = tb_comp_geom(true, N, chunk);
  if (tb_pace_lookup(g) != CF_OK) return -1;
  return g.pace;
}

}  // extern "C"
