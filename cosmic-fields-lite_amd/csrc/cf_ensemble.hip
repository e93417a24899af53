// cf_ensemble.hip -- ensembles of lattices of uneven N advanced together.
//
// All realisations of an ensemble take their RK4 steps in one kernel per step
// over a table of work units (one output tile of one lattice over a chunk of
// a-planes, the unit of the temporally blocked step, cf_tb_body.hpp). Two launch
// forms: a fixed grid of 1024-thread blocks, each running a list of units that
// the host balanced by longest-processing-time-first over their iteration counts
// (so the mix of N = 64 / 96 / 128 lattices finishes together), or one block per
// unit, longest first. No atomics and no inter-block communication. The block's
// region shape (32 x 32 or 24 x 40 columns, so 24 x 24 or 16 x 32 tiles) and the
// chunk are chosen per lattice size; the autotuner times the candidates.
//
// The step scalars a(t), H(t) of every stage live in a device table, so the
// kernels of a whole run can be captured once into a hipGraph and replayed
// (cf_ensemble_run with use_graph). States ping-pong between two buffers per
// realisation; a step reads buffer p and writes 1-p, which is also what lets the
// autotuner time candidate plans without touching the state.
//
// Each realisation's result equals a single-lattice run of the same steps bit
// for bit (same per-site arithmetic as cf_rk4_run_tb_* and the stage kernels).
#include <hipfft/hipfft.h>

#include <algorithm>
#include <map>
#include <queue>
#include <vector>

#include "cf_tb_body.hpp"

namespace cf {

struct EnsReal {
  int64_t N;
  double L;
  void *x[2];      // FP64: state; F32C: hi
  float *lo[2];    // F32C only
};

struct EnsUnit {
  int32_t real;
  int32_t T, Tc;   // output tile edges along b and c
  int32_t region;  // CF_ENS_REGION_*: the block's region shape
  int64_t b0, c0, a0, a1;
};

// Region shapes of a unit's block (rows along b x columns along c): 32 x 32 (24 x 24
// tiles, 1024 threads) and 24 x 40 (16 x 32 tiles, 960 threads: 1.88 region columns
// per output site against 1.78, but tiles that divide N = 64 and 128 where 24 x 24
// leaves 27% of the covered area outside the lattice).
constexpr int ENS_R1_B = 24, ENS_R1_C = 40;
inline int ens_tile_b(int region, int64_t N) {
  const int t = region == CF_ENS_REGION_24X40 ? ENS_R1_B - 2 * TBH : TBR - 2 * TBH;
  return N < t ? (int)N : t;
}
inline int ens_tile_c(int region, int64_t N) {
  const int t = region == CF_ENS_REGION_24X40 ? ENS_R1_C - 2 * TBH : TBR - 2 * TBH;
  return N < t ? (int)N : t;
}

struct EnsStepCoef {
  double a[4], H[4];  // a(t) and H(t) at the four stage times
};

// A value known to be equal in all lanes (work-unit and lattice descriptors,
// step scalars), read from the first lane so that it lives in scalar registers
// and the index arithmetic on it is scalar.
template <class V>
__device__ __forceinline__ V wave_uniform(const V &v) {
  static_assert(sizeof(V) % 4 == 0, "word-sized fields only");
  struct Words {
    uint32_t w[sizeof(V) / 4];
  };
  Words a = __builtin_bit_cast(Words, v);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(V) / 4); ++i) a.w[i] = __builtin_amdgcn_readfirstlane(a.w[i]);
  return __builtin_bit_cast(V, a);
}

// One work unit of the step: its lattice's step scalars, then the temporally
// blocked body over a region of RB x RC columns.
template <class T, bool F32C, bool LAM, bool DEFER, bool LANE_NBR = true, int RB = TBR, int RC = TBR, bool SA = false,
          bool PACE = false>
__device__ __forceinline__ void ens_unit(TbLds<T> &lds, const EnsReal *__restrict__ reals, const EnsUnit &e,
                                         const EnsStepCoef &sc, int parity, double c_m, double lam, double dt,
                                         int pace = 0) {
  const EnsReal &R = reals[e.real];
  // Select this step's buffers first, then make each pointer uniform: indexing
  // x[parity] after the struct is in registers leaves the compiler treating the
  // pointers as divergent (plane addresses then run on the vector unit and every
  // buffer-resource build re-reads them lane by lane).
  const T *xin = wave_uniform((const T *)(parity ? R.x[1] : R.x[0]));
  T *xout = wave_uniform((T *)(parity ? R.x[0] : R.x[1]));
  const float *lin = wave_uniform(parity ? R.lo[1] : R.lo[0]);
  float *lout = wave_uniform(parity ? R.lo[0] : R.lo[1]);
  const int64_t N = wave_uniform(R.N);
  const double L = wave_uniform(R.L);
  TbStep k;
  // The step scalars are the same for the whole block; moving them to scalar
  // registers keeps the vector registers for the step body.
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    k.c_h[s] = wave_uniform((-3.0) * sc.H[s]);
    k.inv[s] = wave_uniform(1.0 / ((L / N) * (L / N)) / (sc.a[s] * sc.a[s]));  // cf_inv_ah_sqr
  }
  k.c_m = c_m;
  k.lam = lam;
  k.m = 0.0;
  k.fa2 = 0.0;
  k.dh = dt / 2.0;
  k.dt = dt;
  k.dt6 = dt / 6.0;
  k.dt3 = dt / 3.0;
  k = wave_uniform(k);  // the divisions above ran on the vector unit (no scalar FP64)
  TbStepF kf{};  // single-precision stencil: the scalars rounded to float, as on the host
  if constexpr (SA) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      kf.c_h[s] = (float)k.c_h[s];
      kf.inv[s] = (float)k.inv[s];
    }
    kf.c_m = (float)k.c_m;
    kf.lam = (float)k.lam;
    kf.dh = (float)k.dh;
    kf.dt = (float)k.dt;
    kf = wave_uniform(kf);
  }
  TbUnit u;
  u.N = N;
  u.b0 = e.b0;
  u.c0 = e.c0;
  u.a0 = e.a0;
  u.a1 = e.a1;
  u.T = e.T;
  u.Tc = e.Tc;
  u.NA = N;
  u.pace = pace;
  tb_unit<T, F32C, LAM, DEFER, LANE_NBR, TB_KG, false, RB, RC, SA, PACE>(lds, u, k, xin, lin, xout, lout, nullptr,
                                                                       TbPsi{}, kf);
}

// A unit in its region shape (a wave-uniform branch; lane neighbours only in rows of
// 32 columns).
template <class T, bool F32C, bool LAM, bool DEFER, bool LANE_NBR, bool SA = false, bool PACE = false>
__device__ __forceinline__ void ens_unit_any(TbLds<T> &lds, const EnsReal *__restrict__ reals, const EnsUnit &unit,
                                             const EnsStepCoef &sc, int parity, double c_m, double lam, double dt,
                                             int pace = 0) {
  const EnsUnit e = wave_uniform(unit);
  if (e.region == CF_ENS_REGION_24X40)
    ens_unit<T, F32C, LAM, DEFER, false, ENS_R1_B, ENS_R1_C, SA, PACE>(lds, reals, e, sc, parity, c_m, lam, dt,
                                                                        pace);
  else
    ens_unit<T, F32C, LAM, DEFER, LANE_NBR, TBR, TBR, SA, PACE>(lds, reals, e, sc, parity, c_m, lam, dt, pace);
}

// Persistent form: a fixed grid; block b runs units unit_off[b] .. unit_off[b+1]-1
// (an LPT-balanced list). Stores are not deferred: the unit loop leaves less
// register room than a single unit.
template <class T, bool F32C, bool LAM>
__global__ __launch_bounds__(1024) void k_ens_tb(const EnsReal *__restrict__ reals,
                                                 const int32_t *__restrict__ unit_off,
                                                 const EnsUnit *__restrict__ units,
                                                 const EnsStepCoef *__restrict__ steps, int64_t step, int parity,
                                                 double c_m, double lam, double dt) {
  __shared__ TbLds<T> lds;
  const EnsStepCoef sc = steps[step];
  const int beg = unit_off[blockIdx.x], end = unit_off[blockIdx.x + 1];
  for (int i = beg; i < end; ++i) {
    ens_unit_any<T, F32C, LAM, false, true>(lds, reals, units[i], sc, parity, c_m, lam, dt);
    __syncthreads();  // the next unit rewrites the plane buffers
  }
}

// Flat form: one block per unit (units sorted longest first, so the hardware
// dispatcher balances the tail), the body of the single-lattice kernel. Blocks go
// to the 8 XCDs in turn; groups of ENS_XCD_GROUP consecutive units (neighbouring
// tiles of one lattice, whose halo rows overlap) are dealt to one XCD, so those rows
// meet in one L2, while the long-first order stays balanced across XCDs.
constexpr int ENS_XCD_GROUP = 4;

template <class T, bool F32C, bool LAM>
__global__ __launch_bounds__(1024) void k_ens_tb_flat(const EnsReal *__restrict__ reals,
                                                      const EnsUnit *__restrict__ units,
                                                      const EnsStepCoef *__restrict__ steps, int64_t step,
                                                      int parity, double c_m, double lam, double dt) {
  __shared__ TbLds<T> lds;
  int u = blockIdx.x;
  const int total = gridDim.x;
  if (total % (8 * ENS_XCD_GROUP) == 0) {  // a bijection of [0, total); otherwise dispatch order
    const int x = u & 7, i = u >> 3;
    u = ((i / ENS_XCD_GROUP) * 8 + x) * ENS_XCD_GROUP + (i % ENS_XCD_GROUP);
  }
  ens_unit_any<T, F32C, LAM, true, !F32C>(lds, reals, units[u], steps[step], parity, c_m, lam, dt);
}

// The same two forms for the compensated storage with the single-precision stencil
// (CF_PREC_F32S), paced (PACE: every unit's plane iterations follow pace ticks of the
// real-time counter from the unit's start; cf_tb_body.hpp). Occupancy is one 1024-thread
// block per CU regardless, so the compiler may use the 128 VGPRs that allows.
template <bool LAM, bool PACE>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_ens_tb_sa(
    const EnsReal *__restrict__ reals, const int32_t *__restrict__ unit_off, const EnsUnit *__restrict__ units,
    const EnsStepCoef *__restrict__ steps, int64_t step, int parity, double c_m, double lam, double dt, int pace) {
  __shared__ TbLds<float> lds;
  const EnsStepCoef sc = steps[step];
  const int beg = unit_off[blockIdx.x], end = unit_off[blockIdx.x + 1];
  for (int i = beg; i < end; ++i) {
    ens_unit_any<float, true, LAM, false, false, true, PACE>(lds, reals, units[i], sc, parity, c_m, lam, dt, pace);
    __syncthreads();
  }
}

template <bool LAM, bool PACE>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_ens_tb_flat_sa(
    const EnsReal *__restrict__ reals, const EnsUnit *__restrict__ units, const EnsStepCoef *__restrict__ steps,
    int64_t step, int parity, double c_m, double lam, double dt, int pace) {
  __shared__ TbLds<float> lds;
  int u = blockIdx.x;
  const int total = gridDim.x;
  if (total % (8 * ENS_XCD_GROUP) == 0) {
    const int x = u & 7, i = u >> 3;
    u = ((i / ENS_XCD_GROUP) * 8 + x) * ENS_XCD_GROUP + (i % ENS_XCD_GROUP);
  }
  ens_unit_any<float, true, LAM, true, false, true, PACE>(lds, reals, units[u], steps[step], parity, c_m, lam, dt,
                                                          pace);
}

}  // namespace cf

using namespace cf;

struct cf_ensemble {
  int n = 0, prec = CF_PREC_F64;
  int pace = 0;  // CF_PREC_F32S: ticks per plane iteration (0: unpaced)
  double m = 0, lambda = 0;
  cf_cosmology cosmo{1, 0, 0, 1};
  std::vector<EnsReal> reals;
  std::vector<void *> allocs;
  EnsReal *d_reals = nullptr;
  std::map<int64_t, int> chunk;   // planes per unit, per lattice size
  std::map<int64_t, int> region;  // CF_ENS_REGION_*, per lattice size
  int grid = 0;
  int n_units = 0;
  int32_t *d_unit_off = nullptr;
  EnsUnit *d_units = nullptr;
  EnsStepCoef *d_steps = nullptr;
  EnsStepCoef *h_steps = nullptr;  // pinned staging of the step table
  hipEvent_t steps_ev = nullptr;   // recorded after the table's last upload
  int64_t steps_cap = 0;
  int parity = 0;
  hipStream_t st = nullptr;
  hipGraphExec_t exec = nullptr;
  double g_t0 = 0, g_dt = 0;
  int64_t g_k0 = -1, g_n = -1;
  int g_parity = -1;
  std::map<int64_t, cf_spectrum_plan *> plans;
  double *scratch = nullptr;  // 3 N_max^3 doubles: a joined state and a density
};

namespace cf {

static int ens_default_chunk(int64_t N) { return N >= 256 ? 128 : (int)N; }

static void ens_free_plan(cf_ensemble *e) {
  if (e->d_unit_off) (void)hipFree(e->d_unit_off);
  if (e->d_units) (void)hipFree(e->d_units);
  e->d_unit_off = nullptr;
  e->d_units = nullptr;
}

// The graph is dropped whenever the tables it reads (units, step scalars) are
// rewritten; a replay still in flight on the ensemble's stream ends first.
static void ens_drop_graph(cf_ensemble *e) {
  if (e->exec) {
    (void)hipStreamSynchronize(e->st);
    (void)hipGraphExecDestroy(e->exec);
  }
  e->exec = nullptr;
  e->g_k0 = -1;
}

// Build the unit table for the current chunks and grid, LPT-balanced. Launches
// still in flight read the old table, so they end before it is freed.
static cf_status ens_build(cf_ensemble *e) {
  CF_HIP(hipStreamSynchronize(e->st));
  std::vector<EnsUnit> units;
  std::vector<int> cost;
  for (int r = 0; r < e->n; ++r) {
    const int64_t N = e->reals[r].N;
    const int reg = e->region[N];
    const int Tb = ens_tile_b(reg, N), Tc = ens_tile_c(reg, N);
    const int64_t ntb = cdiv(N, Tb), ntc = cdiv(N, Tc);
    const int A = e->chunk[N];
    const int waves = reg == CF_ENS_REGION_24X40 ? (ENS_R1_B * ENS_R1_C + 63) / 64 : TBR * TBR / 64;
    for (int64_t a0 = 0; a0 < N; a0 += A)
      for (int64_t tb = 0; tb < ntb; ++tb)
        for (int64_t tc = 0; tc < ntc; ++tc) {
          EnsUnit u;
          u.real = r;
          u.T = Tb;
          u.Tc = Tc;
          u.region = reg;
          u.b0 = tb * Tb;
          u.c0 = tc * Tc;
          u.a0 = a0;
          u.a1 = std::min<int64_t>(a0 + A, N);
          units.push_back(u);
          cost.push_back(((int)(u.a1 - u.a0) + 8) * waves);
        }
  }
  std::vector<int> order(units.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
  if (e->grid == CF_ENS_GRID_FLAT) {  // one block per unit, longest first
    std::vector<EnsUnit> flat;
    for (int i : order) flat.push_back(units[i]);
    ens_free_plan(e);
    e->n_units = (int)flat.size();
    CF_HIP(hipMalloc((void **)&e->d_units, flat.size() * sizeof(EnsUnit)));
    CF_HIP(hipMemcpy(e->d_units, flat.data(), flat.size() * sizeof(EnsUnit), hipMemcpyHostToDevice));
    ens_drop_graph(e);
    return CF_OK;
  }
  const int G = std::max(1, std::min<int>(e->grid, (int)units.size()));
  typedef std::pair<long, int> Load;  // (load, block)
  std::priority_queue<Load, std::vector<Load>, std::greater<Load>> heap;
  for (int b = 0; b < G; ++b) heap.push(Load(0, b));
  std::vector<std::vector<int>> lists(G);
  for (int i : order) {
    Load l = heap.top();
    heap.pop();
    lists[l.second].push_back(i);
    heap.push(Load(l.first + cost[i], l.second));
  }
  std::vector<int32_t> off(G + 1, 0);
  std::vector<EnsUnit> flat;
  for (int b = 0; b < G; ++b) {
    for (int i : lists[b]) flat.push_back(units[i]);
    off[b + 1] = (int32_t)flat.size();
  }
  ens_free_plan(e);
  e->grid = G;
  e->n_units = (int)flat.size();
  CF_HIP(hipMalloc((void **)&e->d_unit_off, off.size() * sizeof(int32_t)));
  CF_HIP(hipMalloc((void **)&e->d_units, flat.size() * sizeof(EnsUnit)));
  CF_HIP(hipMemcpy(e->d_unit_off, off.data(), off.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  CF_HIP(hipMemcpy(e->d_units, flat.data(), flat.size() * sizeof(EnsUnit), hipMemcpyHostToDevice));
  ens_drop_graph(e);
  return CF_OK;
}

// Step table for steps k0 .. k0+n-1 at t_k = t0 + k*dt. The upload is queued on
// the ensemble's stream, behind every launch of earlier runs that still reads the
// previous table, from a pinned staging buffer that is rewritten only after its
// previous upload has completed. Runs may therefore follow one another without a
// host synchronisation. A captured graph reads the table, so it is dropped.
static cf_status ens_fill_steps(cf_ensemble *e, double t0, int64_t k0, int64_t n, double dt) {
  ens_drop_graph(e);
  if (n > e->steps_cap) {
    CF_HIP(hipStreamSynchronize(e->st));  // launches in flight read the old buffers
    if (e->d_steps) CF_HIP(hipFree(e->d_steps));
    if (e->h_steps) CF_HIP(hipHostFree(e->h_steps));
    e->d_steps = nullptr;
    e->h_steps = nullptr;
    e->steps_cap = 0;
    CF_HIP(hipMalloc((void **)&e->d_steps, (size_t)n * sizeof(EnsStepCoef)));
    CF_HIP(hipHostMalloc((void **)&e->h_steps, (size_t)n * sizeof(EnsStepCoef), hipHostMallocDefault));
    e->steps_cap = n;
  } else {
    CF_HIP(hipEventSynchronize(e->steps_ev));  // the staging buffer's previous upload is done
  }
  EnsStepCoef *h = e->h_steps;
  for (int64_t s = 0; s < n; ++s) {
    const double t = t0 + static_cast<double>(k0 + s) * dt;
    const double dh = dt / 2.0, th = t + dh;
    const double ts[4] = {t, th, th, t + dt};
    for (int q = 0; q < 4; ++q) {
      h[s].a[q] = cf_cosmo_a(&e->cosmo, ts[q]);
      h[s].H[q] = cf_cosmo_H(&e->cosmo, ts[q]);
    }
  }
  CF_HIP(hipMemcpyAsync(e->d_steps, h, (size_t)n * sizeof(EnsStepCoef), hipMemcpyHostToDevice, e->st));
  CF_HIP(hipEventRecord(e->steps_ev, e->st));
  return CF_OK;
}

static cf_status ens_launch_sa(cf_ensemble *e, int64_t step, int parity, double dt, hipStream_t st) {
  const double c_m = e->m * e->m;
  const bool lam = e->lambda != 0.0, paced = e->pace > 0;
  const dim3 blk(1024);
  if (e->grid == CF_ENS_GRID_FLAT) {
    const dim3 grd(e->n_units);
#define CF_ENS_FLAT_SA(L, P)                                                                                        \
  hipLaunchKernelGGL((k_ens_tb_flat_sa<L, P>), grd, blk, 0, st, e->d_reals, e->d_units, e->d_steps, step, parity, \
                     c_m, e->lambda, dt, e->pace)
    if (lam && paced) CF_ENS_FLAT_SA(true, true);
    else if (lam) CF_ENS_FLAT_SA(true, false);
    else if (paced) CF_ENS_FLAT_SA(false, true);
    else CF_ENS_FLAT_SA(false, false);
#undef CF_ENS_FLAT_SA
  } else {
    const dim3 grd(e->grid);
#define CF_ENS_SA(L, P)                                                                                            \
  hipLaunchKernelGGL((k_ens_tb_sa<L, P>), grd, blk, 0, st, e->d_reals, e->d_unit_off, e->d_units, e->d_steps, step, \
                     parity, c_m, e->lambda, dt, e->pace)
    if (lam && paced) CF_ENS_SA(true, true);
    else if (lam) CF_ENS_SA(true, false);
    else if (paced) CF_ENS_SA(false, true);
    else CF_ENS_SA(false, false);
#undef CF_ENS_SA
  }
  CF_LAUNCHED();
  return CF_OK;
}

static cf_status ens_launch(cf_ensemble *e, int64_t step, int parity, double dt, hipStream_t st) {
  if (e->prec == CF_PREC_F32S) return ens_launch_sa(e, step, parity, dt, st);
  dim3 grd(e->grid), blk(1024);
  const double c_m = e->m * e->m;
  const bool lam = e->lambda != 0.0;
  if (e->grid == CF_ENS_GRID_FLAT) {
    grd = dim3(e->n_units);
    if (e->prec == CF_PREC_F32C) {
      if (lam)
        hipLaunchKernelGGL((k_ens_tb_flat<float, true, true>), grd, blk, 0, st, e->d_reals, e->d_units, e->d_steps,
                           step, parity, c_m, e->lambda, dt);
      else
        hipLaunchKernelGGL((k_ens_tb_flat<float, true, false>), grd, blk, 0, st, e->d_reals, e->d_units, e->d_steps,
                           step, parity, c_m, e->lambda, dt);
    } else {
      if (lam)
        hipLaunchKernelGGL((k_ens_tb_flat<double, false, true>), grd, blk, 0, st, e->d_reals, e->d_units,
                           e->d_steps, step, parity, c_m, e->lambda, dt);
      else
        hipLaunchKernelGGL((k_ens_tb_flat<double, false, false>), grd, blk, 0, st, e->d_reals, e->d_units,
                           e->d_steps, step, parity, c_m, e->lambda, dt);
    }
    CF_LAUNCHED();
    return CF_OK;
  }
  if (e->prec == CF_PREC_F32C) {
    if (lam)
      hipLaunchKernelGGL((k_ens_tb<float, true, true>), grd, blk, 0, st, e->d_reals, e->d_unit_off, e->d_units,
                         e->d_steps, step, parity, c_m, e->lambda, dt);
    else
      hipLaunchKernelGGL((k_ens_tb<float, true, false>), grd, blk, 0, st, e->d_reals, e->d_unit_off, e->d_units,
                         e->d_steps, step, parity, c_m, e->lambda, dt);
  } else {
    if (lam)
      hipLaunchKernelGGL((k_ens_tb<double, false, true>), grd, blk, 0, st, e->d_reals, e->d_unit_off, e->d_units,
                         e->d_steps, step, parity, c_m, e->lambda, dt);
    else
      hipLaunchKernelGGL((k_ens_tb<double, false, false>), grd, blk, 0, st, e->d_reals, e->d_unit_off,
                         e->d_units, e->d_steps, step, parity, c_m, e->lambda, dt);
  }
  CF_LAUNCHED();
  return CF_OK;
}

// Join the caller's stream into the ensemble's stream and back.
static cf_status ens_fence(hipStream_t from, hipStream_t to) {
  hipEvent_t ev;
  CF_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  hipError_t a = hipEventRecord(ev, from);
  hipError_t b = hipStreamWaitEvent(to, ev, 0);
  (void)hipEventDestroy(ev);
  CF_HIP(a);
  CF_HIP(b);
  return CF_OK;
}

// Time per launch of step 0 (reads buffer p, writes 1-p): the shorter of two windows
// of `reps` launches each, after one warm launch.
static cf_status ens_time(cf_ensemble *e, int reps, double dt, float *ms) {
  hipEvent_t e0, e1;
  CF_HIP(hipEventCreate(&e0));
  CF_HIP(hipEventCreate(&e1));
  cf_status rc = ens_launch(e, 0, e->parity, dt, e->st);  // warm
  *ms = 1e30f;
  for (int w = 0; w < 2 && rc == CF_OK; ++w) {
    float t = 0.0f;
    if (hipEventRecord(e0, e->st) != hipSuccess) rc = CF_E_HIP;
    for (int i = 0; i < reps && rc == CF_OK; ++i) rc = ens_launch(e, 0, e->parity, dt, e->st);
    if (rc == CF_OK && hipEventRecord(e1, e->st) != hipSuccess) rc = CF_E_HIP;
    if (rc == CF_OK && hipEventSynchronize(e1) != hipSuccess) rc = CF_E_HIP;
    if (rc == CF_OK && hipEventElapsedTime(&t, e0, e1) != hipSuccess) rc = CF_E_HIP;
    if (rc == CF_OK && t / reps < *ms) *ms = t / reps;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != CF_OK) set_error("ensemble timing failed");
  return rc;
}

// CF_PREC_F32S: the pace of the units' plane iterations, as the single-lattice tuner
// (cf_rk4_tb_tune_pace) picks it: a coarse grid of paces, then tick by tick around the
// best, each pace P scored by the slowest of P-2..P (a margin above the fall below the
// fastest pace); unpaced if that is as fast.
static cf_status ens_tune_pace(cf_ensemble *e, int reps, double dt) {
  std::map<int, float> ms;
  auto time_pace = [&](int pace) -> cf_status {
    if (ms.count(pace)) return CF_OK;
    e->pace = pace;
    float t = 0.0f;
    CF_TRY(ens_time(e, reps, dt, &t));
    ms[pace] = t;
    return CF_OK;
  };
  for (int pace = 0; pace <= 320; pace += (pace == 0 ? 64 : 8)) CF_TRY(time_pace(pace));
  int coarse = 0;
  for (const auto &kv : ms)
    if (kv.second < ms[coarse]) coarse = kv.first;
  constexpr int MARGIN = 2;
  if (coarse > 0)
    for (int pace = coarse - 8; pace <= coarse + 8 + MARGIN; ++pace) CF_TRY(time_pace(pace));
  int pick = 0;
  float pick_worst = ms[0];
  for (const auto &kv : ms) {
    if (kv.first - MARGIN < 1) continue;
    float worst = 0.0f;
    bool full = true;
    for (int q = kv.first - MARGIN; q <= kv.first && full; ++q) {
      const auto it = ms.find(q);
      full = it != ms.end();
      if (full) worst = it->second > worst ? it->second : worst;
    }
    if (full && worst <= pick_worst && (pick > 0 || worst < ms[0])) {
      pick = kv.first;
      pick_worst = worst;
    }
  }
  e->pace = pick;
  return CF_OK;
}

}  // namespace cf

extern "C" {

cf_status cf_ensemble_set_pace(cf_ensemble *e, int pace) {
  CF_REQUIRE(e, "cf_ensemble_set_pace: null ensemble");
  CF_REQUIRE(pace >= 0 && pace <= 100000, "cf_ensemble_set_pace: pace %d outside 0..100000", pace);
  CF_REQUIRE(pace == 0 || e->prec == CF_PREC_F32S, "cf_ensemble_set_pace: only CF_PREC_F32S ensembles are paced");
  ens_drop_graph(e);  // a captured graph holds the old pace as a kernel argument
  e->pace = pace;
  return CF_OK;
}

int cf_ensemble_get_pace(const cf_ensemble *e) { return e ? e->pace : -1; }

cf_status cf_ensemble_destroy(cf_ensemble *e) {
  if (!e) return CF_OK;
  ens_drop_graph(e);
  ens_free_plan(e);
  if (e->st) (void)hipStreamSynchronize(e->st);
  if (e->d_steps) (void)hipFree(e->d_steps);
  if (e->h_steps) (void)hipHostFree(e->h_steps);
  if (e->steps_ev) (void)hipEventDestroy(e->steps_ev);
  if (e->d_reals) (void)hipFree(e->d_reals);
  for (void *p : e->allocs) (void)hipFree(p);
  for (auto &kv : e->plans) cf_spectrum_plan_destroy(kv.second);
  if (e->scratch) (void)hipFree(e->scratch);
  if (e->st) (void)hipStreamDestroy(e->st);
  delete e;
  return CF_OK;
}

cf_status cf_ensemble_create(int n, const int64_t *N, const double *L, double m, double lambda,
                             const cf_cosmology *cosmo, int precision, cf_ensemble **out) {
  CF_REQUIRE(out && N && L && cosmo && n > 0, "cf_ensemble_create: bad arguments");
  CF_REQUIRE(precision == CF_PREC_F64 || precision == CF_PREC_F32C || precision == CF_PREC_F32S,
             "cf_ensemble_create: unknown precision");
  *out = nullptr;
  for (int r = 0; r < n; ++r)
    CF_REQUIRE(valid_N(N[r]) && N[r] <= 1024 && L[r] > 0, "cf_ensemble_create: realisation %d: bad N or L", r);
  cf_ensemble *e = new cf_ensemble();
  e->n = n;
  e->prec = precision;
  e->m = m;
  e->lambda = lambda;
  e->cosmo = *cosmo;
  e->grid = 1024;
  auto fail = [&](cf_status s) {
    cf_ensemble_destroy(e);
    return s;
  };
  if (hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&e->steps_ev, hipEventDisableTiming) != hipSuccess) {
    set_error("cf_ensemble_create: stream");
    return fail(CF_E_HIP);
  }
  int64_t nmax = 0;
  for (int r = 0; r < n; ++r) {
    EnsReal R{};
    R.N = N[r];
    R.L = L[r];
    nmax = std::max(nmax, N[r]);
    const size_t cnt = 2 * (size_t)N[r] * N[r] * N[r];
    for (int q = 0; q < 2; ++q) {
      void *p = nullptr;
      const size_t bytes = cnt * (precision != CF_PREC_F64 ? sizeof(float) : sizeof(double));
      if (hipMalloc(&p, bytes) != hipSuccess) {
        set_error("cf_ensemble_create: out of device memory");
        return fail(CF_E_NOMEM);
      }
      e->allocs.push_back(p);
      R.x[q] = p;
      if (precision != CF_PREC_F64) {  // compensated storage: hi in x, lo here
        void *l = nullptr;
        if (hipMalloc(&l, cnt * sizeof(float)) != hipSuccess) {
          set_error("cf_ensemble_create: out of device memory");
          return fail(CF_E_NOMEM);
        }
        e->allocs.push_back(l);
        R.lo[q] = (float *)l;
      }
    }
    e->reals.push_back(R);
    if (!e->chunk.count(N[r])) e->chunk[N[r]] = ens_default_chunk(N[r]);
    if (!e->region.count(N[r])) e->region[N[r]] = CF_ENS_REGION_32X32;
  }
  if (hipMalloc((void **)&e->d_reals, n * sizeof(EnsReal)) != hipSuccess ||
      hipMemcpy(e->d_reals, e->reals.data(), n * sizeof(EnsReal), hipMemcpyHostToDevice) != hipSuccess ||
      hipMalloc((void **)&e->scratch, 3 * (size_t)nmax * nmax * nmax * sizeof(double)) != hipSuccess) {
    set_error("cf_ensemble_create: allocation failed");
    return fail(CF_E_NOMEM);
  }
  cf_status s = ens_build(e);
  if (s != CF_OK) return fail(s);
  *out = e;
  return CF_OK;
}

int64_t cf_ensemble_sites(const cf_ensemble *e) {
  int64_t s = 0;
  if (e)
    for (const auto &R : e->reals) s += R.N * R.N * R.N;
  return s;
}

cf_status cf_ensemble_state(cf_ensemble *e, int r, void **x_dev, float **lo_dev) {
  CF_REQUIRE(e && r >= 0 && r < e->n && x_dev, "cf_ensemble_state: bad arguments");
  *x_dev = e->reals[r].x[e->parity];
  if (lo_dev) *lo_dev = e->prec != CF_PREC_F64 ? e->reals[r].lo[e->parity] : nullptr;
  return CF_OK;
}

cf_status cf_ensemble_set_plan(cf_ensemble *e, int64_t N, int chunk, int grid) {
  CF_REQUIRE(e, "cf_ensemble_set_plan: null ensemble");
  if (N > 0) {
    CF_REQUIRE(e->chunk.count(N), "cf_ensemble_set_plan: no lattice of size %lld", (long long)N);
    e->chunk[N] = (chunk <= 0 || chunk > N) ? (int)N : chunk;
  }
  CF_REQUIRE(grid >= 0 || grid == CF_ENS_GRID_FLAT, "cf_ensemble_set_plan: bad grid %d", grid);
  if (grid != 0) e->grid = grid;
  return ens_build(e);
}

cf_status cf_ensemble_get_plan(const cf_ensemble *e, int64_t N, int *chunk, int *grid) {
  CF_REQUIRE(e && chunk && grid, "cf_ensemble_get_plan: bad arguments");
  auto it = e->chunk.find(N);
  *chunk = it == e->chunk.end() ? 0 : it->second;
  *grid = e->grid;
  return CF_OK;
}

cf_status cf_ensemble_set_region(cf_ensemble *e, int64_t N, int region) {
  CF_REQUIRE(e, "cf_ensemble_set_region: null ensemble");
  CF_REQUIRE(e->region.count(N), "cf_ensemble_set_region: no lattice of size %lld", (long long)N);
  CF_REQUIRE(region == CF_ENS_REGION_32X32 || region == CF_ENS_REGION_24X40,
             "cf_ensemble_set_region: unknown region shape %d", region);
  e->region[N] = region;
  return ens_build(e);
}

cf_status cf_ensemble_get_region(const cf_ensemble *e, int64_t N, int *region) {
  CF_REQUIRE(e && region, "cf_ensemble_get_region: bad arguments");
  auto it = e->region.find(N);
  CF_REQUIRE(it != e->region.end(), "cf_ensemble_get_region: no lattice of size %lld", (long long)N);
  *region = it->second;
  return CF_OK;
}

cf_status cf_ensemble_init_grf(cf_ensemble *e, const uint32_t *seeds, int kind, const double *params, double rms,
                               cf_stream stream) {
  CF_REQUIRE(e && seeds, "cf_ensemble_init_grf: bad arguments");
  for (int r = 0; r < e->n; ++r) {
    EnsReal &R = e->reals[r];
    cf_spectrum_plan *&pl = e->plans[R.N];
    if (!pl) CF_TRY(cf_spectrum_plan_create(R.N, &pl));
    const int64_t cnt = 2 * R.N * R.N * R.N;
    double *dst = e->prec == CF_PREC_F64 ? (double *)R.x[e->parity] : e->scratch;
    CF_TRY(cf_init_grf_state(pl, seeds[r], kind, params, R.L, rms, dst, stream));
    if (e->prec != CF_PREC_F64)
      CF_TRY(cf_split_f64(e->scratch, (float *)R.x[e->parity], R.lo[e->parity], cnt, stream));
  }
  CF_HIP(hipStreamSynchronize(S(stream)));
  return CF_OK;
}

cf_status cf_ensemble_tune(cf_ensemble *e, double dt, int reps, int *grid_out) {
  CF_REQUIRE(e, "cf_ensemble_tune: null ensemble");
  if (reps <= 0) reps = 3;
  CF_TRY(ens_fill_steps(e, 0.0, 0, 1, dt));
  const int grids[] = {256, 512, 768, 1024, 2048, CF_ENS_GRID_FLAT};
  float best = 1e30f, ms = 0;
  int best_g = e->grid;
  for (int g : grids) {
    e->grid = g;
    CF_TRY(ens_build(e));
    CF_TRY(ens_time(e, reps, dt, &ms));
    if (ms < best) {
      best = ms;
      best_g = g;
    }
  }
  e->grid = best_g;
  CF_TRY(ens_build(e));
  // per lattice size, every (region shape, chunk) pair, the other sizes' plans held;
  // a second sweep revisits each size once the others have their tuned plans. Chunks
  // go longest first and a candidate must win by more than 0.5%: within the timing
  // noise the longer chunk (fewer warm-up planes) is kept. A paced ensemble
  // (CF_PREC_F32S) tunes its pace between the two sweeps, so the second one times the
  // plans paced.
  for (int sweep = 0; sweep < 2; ++sweep) {
    if (sweep == 1 && e->prec == CF_PREC_F32S) CF_TRY(ens_tune_pace(e, reps, dt));
    for (auto &kv : e->chunk) {
      const int64_t N = kv.first;
      int best_a = kv.second, best_r = e->region[N];
      CF_TRY(ens_time(e, reps, dt, &best));
      std::vector<int> chunks;
      for (int A : {(int)N, 256, 128, 64, 32, 16, 8})
        if (A <= N && std::find(chunks.begin(), chunks.end(), A) == chunks.end()) chunks.push_back(A);
      for (int reg : {CF_ENS_REGION_32X32, CF_ENS_REGION_24X40})
        for (int A : chunks) {
          if (A == best_a && reg == best_r) continue;
          e->chunk[N] = A;
          e->region[N] = reg;
          CF_TRY(ens_build(e));
          CF_TRY(ens_time(e, reps, dt, &ms));
          if (ms < best * 0.995f) {
            best = ms;
            best_a = A;
            best_r = reg;
          }
        }
      e->chunk[N] = best_a;
      e->region[N] = best_r;
      CF_TRY(ens_build(e));
    }
  }
  if (grid_out) *grid_out = e->grid;
  return CF_OK;
}

cf_status cf_ensemble_run(cf_ensemble *e, double t0, int64_t k0, int64_t n_steps, double dt, int use_graph,
                          cf_stream stream) {
  CF_REQUIRE(e && n_steps >= 0, "cf_ensemble_run: bad arguments");
  if (n_steps == 0) return CF_OK;
  const bool reuse = use_graph && e->exec && e->g_t0 == t0 && e->g_k0 == k0 && e->g_n == n_steps && e->g_dt == dt &&
                     e->g_parity == e->parity;
  if (!reuse) CF_TRY(ens_fill_steps(e, t0, k0, n_steps, dt));
  CF_TRY(ens_fence(S(stream), e->st));
  if (!use_graph) {
    for (int64_t s = 0; s < n_steps; ++s) CF_TRY(ens_launch(e, s, e->parity ^ (int)(s & 1), dt, e->st));
  } else {
    if (!reuse) {
      ens_drop_graph(e);
      hipGraph_t g;
      CF_HIP(hipStreamBeginCapture(e->st, hipStreamCaptureModeThreadLocal));
      cf_status rc = CF_OK;
      for (int64_t s = 0; s < n_steps && rc == CF_OK; ++s) rc = ens_launch(e, s, e->parity ^ (int)(s & 1), dt, e->st);
      hipError_t ce = hipStreamEndCapture(e->st, &g);
      if (ce == hipSuccess && rc != CF_OK) (void)hipGraphDestroy(g);
      if (rc != CF_OK) return rc;
      CF_HIP(ce);
      hipError_t ie = hipGraphInstantiate(&e->exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      CF_HIP(ie);
      e->g_t0 = t0;
      e->g_k0 = k0;
      e->g_n = n_steps;
      e->g_dt = dt;
      e->g_parity = e->parity;
    }
    CF_HIP(hipGraphLaunch(e->exec, e->st));
  }
  CF_TRY(ens_fence(e->st, S(stream)));
  e->parity ^= (int)(n_steps & 1);
  return CF_OK;
}

// Energy-density power spectrum of realisation r at time t (observer.hpp:102-108),
// into bins_dev (cf_spectrum_bins(N) doubles).
cf_status cf_ensemble_density_spectrum(cf_ensemble *e, int r, double t, double *bins_dev, cf_stream stream) {
  CF_REQUIRE(e && r >= 0 && r < e->n && bins_dev, "cf_ensemble_density_spectrum: bad arguments");
  EnsReal &R = e->reals[r];
  cf_spectrum_plan *&pl = e->plans[R.N];
  if (!pl) CF_TRY(cf_spectrum_plan_create(R.N, &pl));
  const int64_t n3 = R.N * R.N * R.N;
  const double *x = (const double *)R.x[e->parity];
  if (e->prec != CF_PREC_F64) {
    CF_TRY(cf_join_f64((const float *)R.x[e->parity], R.lo[e->parity], e->scratch, 2 * n3, stream));
    x = e->scratch;
  }
  double *rho = e->scratch + 2 * n3;  // after the joined state (scratch holds 3 N_max^3)
  const double a_t = cf_cosmo_a(&e->cosmo, t);
  CF_TRY(cf_energy_density_f64(R.N, x, rho, e->m, e->lambda, cf_inv_ah_sqr(R.N, R.L, a_t), stream));
  return cf_power_spectrum_consume(pl, rho, bins_dev, stream);  // rho is this call's scratch
}

}  // extern "C"
