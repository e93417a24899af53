/* cfhip.h -- C ABI of libcfhip.so, the MI355X (gfx950) lattice-evolution library.
 *
 * Drop-in boundary for the Cosmic-Fields-Lite hot path: every entry point below
 * names the reference interface it replaces (paths relative to the reference
 * repository). Plain pointers and sizes only; no C++ or torch types cross.
 *
 * Conventions
 *   - Lattice site (a,b,c) of an N^3 lattice is at N*N*a + N*b + c (c fastest),
 *     as IDX_OF in src/fdm3d.hpp:16. A state is [phi (N^3) | dphi/dt (N^3)].
 *   - All pointers named *_dev are device pointers owned by the caller. The
 *     library allocates device memory in the *_create calls (plans, ensembles,
 *     variants, WKB), in cf_spectrum_plan_create for its own scratch, and, per
 *     call, a stream-ordered scratch of at most 8 KB in cf_norm_inf,
 *     cf_count_nonfinite and cf_rescale_rms (hipMallocAsync / hipFreeAsync on the
 *     caller's stream).
 *   - Launches are asynchronous on the caller's stream (NULL = default stream);
 *     only cf_d2h, cf_sync and the *_host helpers block.
 *   - Every call returns CF_OK or an error code; cf_last_error() gives a
 *     thread-local message. No exception crosses the ABI.
 *   - Handles are not thread-safe (the reference is single-threaded).
 */
#ifndef CFHIP_H
#define CFHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t cf_status;
#define CF_OK 0
#define CF_E_INVALID 1     /* bad argument (size, null pointer, unsupported N) */
#define CF_E_HIP 2         /* HIP runtime error */
#define CF_E_FFT 3         /* hipFFT error */
#define CF_E_NOMEM 4       /* device or host allocation failed */
#define CF_E_UNSUPPORTED 5 /* valid request this library does not implement */

typedef void *cf_stream; /* a hipStream_t; NULL is the default stream */
typedef struct cf_spectrum_plan cf_spectrum_plan; /* FFT plans + bin table of one N (below) */

/* ---------------------------------------------------------------- status */
const char *cf_last_error(void);
const char *cf_version(void);

/* ---------------------------------------------------------------- device, memory
 * Replace thrust::device_vector<double> allocation/fill (src/cuda_wrapper.cuh:39-45),
 * the initializer H2D copy ALGORITHM_NAMESPACE::copy (src/initializer.hpp:24-30,43-44)
 * and copy_vector D2H (src/cuda_wrapper.cu:45-54). */
cf_status cf_device_count(int *count);
cf_status cf_set_device(int device);
cf_status cf_sync(cf_stream stream);
cf_status cf_stream_create(cf_stream *stream);
cf_status cf_stream_destroy(cf_stream stream);
cf_status cf_alloc(size_t bytes, void **ptr_dev);
cf_status cf_free(void *ptr_dev);
cf_status cf_h2d(void *dst_dev, const void *src_host, size_t bytes, cf_stream stream); /* completes before return */
cf_status cf_d2h(void *dst_host, const void *src_dev, size_t bytes, cf_stream stream); /* completes before return */
cf_status cf_d2d(void *dst_dev, const void *src_dev, size_t bytes, cf_stream stream);
cf_status cf_fill_zero(void *ptr_dev, size_t bytes, cf_stream stream); /* a kernel, safe inside graph capture */

/* ---------------------------------------------------------------- parameters */
/* StaticEOSCosmology (src/physics.hpp:15-56): a(t), H(t), eta(t) for a constant
 * equation of state. Minkowski = {1, 0, 0, 1} (physics.hpp:30-31). */
typedef struct {
  double a1, H1, t1, p;
} cf_cosmology;

/* Klein-Gordon in FRW with an optional lambda*phi^4/4 potential:
 *   phi'' + 3H phi' - Lap(phi)/a^2 + m^2 phi + lambda phi^3 = 0
 * KleinGordonEquation (src/equations.cpp:8-36) is cosmology = Minkowski, lambda = 0;
 * KleinGordonEquationInFRW (equations.cpp:71-104, equations_cuda.cu:522-534) has lambda = 0;
 * CudaLambdaEquationInFRW (equations_cuda.cu:575-607) has lambda != 0. */
typedef struct {
  int64_t N;
  double L, m, lambda;
  cf_cosmology cosmo;
} cf_kg_params;

/* Host evaluation of the cosmology (physics.hpp:34-41, 52-55) and of
 * inv_ah_sqr = 1/((L/N)^2)/(a^2) (equations.cpp:80). */
double cf_cosmo_a(const cf_cosmology *c, double t);
double cf_cosmo_H(const cf_cosmology *c, double t);
double cf_cosmo_eta(const cf_cosmology *c, double t);
double cf_inv_ah_sqr(int64_t N, double L, double a_t);

/* ---------------------------------------------------------------- odeint system and ops (FP64)
 * cf_rhs_f64 replaces Equation::operator()(const State &x, State &dxdt, double t)
 * (KG_FRW_equation_kernel / Lambda_FRW_equation_kernel, equations_cuda.cu:15-61, via
 * CudaKleinGordonEquationInFRW::operator() :522-534 and CudaLambdaEquationInFRW :575-589).
 * H and inv_ah2 are the host-evaluated H(t) and cf_inv_ah_sqr(N, L, a(t)). */
cf_status cf_rhs_f64(int64_t N, const double *x_dev, double *dxdt_dev, double m, double lambda,
                     double H, double inv_ah2, cf_stream stream);

/* thrust_operations::scale_sumK (src/odeint_thrust/thrust_operations.cu:4-189, ops :220-353):
 * v0 = sum_{i<K} alpha[i] * v[i], evaluated left to right, 1 <= K <= 14. */
cf_status cf_scale_sum(int K, double *v0_dev, const double *const *v_dev_array_host,
                       const double *alpha_host, int64_t n, cf_stream stream);
cf_status cf_scale_sum2(double *v0_dev, const double *v1_dev, const double *v2_dev, double a1,
                        double a2, int64_t n, cf_stream stream);
cf_status cf_scale_sum5(double *v0_dev, const double *v1_dev, const double *v2_dev,
                        const double *v3_dev, const double *v4_dev, const double *v5_dev, double a1,
                        double a2, double a3, double a4, double a5, int64_t n, cf_stream stream);
/* vector_space_norm_inf (src/odeint_thrust/thrust_algebra.hpp:51-64): max |v|, blocking. */
cf_status cf_norm_inf(const double *v_dev, int64_t n, double *out_host, cf_stream stream);
/* Number of NaN or infinite entries of v, blocking. No reference counterpart (the
 * reference builds with -ffast-math and checks nothing): the observer's optional
 * finite check. */
cf_status cf_count_nonfinite(const double *v_dev, int64_t n, int64_t *count_host, cf_stream stream);

/* ---------------------------------------------------------------- fused RK4 steps
 * One classic RK4 step (runge_kutta4_classic::do_step_impl,
 * external/boost/numeric/odeint/stepper/runge_kutta4_classic.hpp:87-138, with the
 * scale_sum2/scale_sum5 axpys of thrust_operations.cu:236-313) in four fused stage
 * kernels. cf_rk4_run takes n_steps steps at t_k = t0 + (k0 + k) * dt, the time
 * formula of integrate_const (detail/integrate_const.hpp:56-63).
 *
 * FP64: x_dev is the state (2N^3 doubles), updated in place; the results equal the
 * reference arithmetic order (1*x + dt/6 k1 + dt/3 k2 + dt/3 k3 + dt/6 k4).
 * work_dev holds cf_rk4_work_bytes(N, CF_PREC_F64) bytes. */
#define CF_PREC_F64 0
#define CF_PREC_F32C 1 /* float hi+lo state, float stage arrays, double arithmetic */
#define CF_PREC_F32S 2 /* float hi+lo state, float stencil arithmetic, double increment */
size_t cf_rk4_work_bytes(int64_t N, int precision);
cf_status cf_rk4_step_f64(const cf_kg_params *p, double *x_dev, double t, double dt, void *work_dev,
                          cf_stream stream);
cf_status cf_rk4_run_f64(const cf_kg_params *p, double *x_dev, double t0, int64_t k0, int64_t n_steps,
                         double dt, void *work_dev, cf_stream stream);
/* FP32-compensated storage: hi_dev, lo_dev are 2N^3 floats each; the state is
 * hi + lo. Stage inputs and the RK4 increment are stored as float, all arithmetic is
 * double, the update x += increment is compensated into lo. */
cf_status cf_rk4_step_f32c(const cf_kg_params *p, float *hi_dev, float *lo_dev, double t, double dt,
                           void *work_dev, cf_stream stream);
cf_status cf_rk4_run_f32c(const cf_kg_params *p, float *hi_dev, float *lo_dev, double t0, int64_t k0,
                          int64_t n_steps, double dt, void *work_dev, cf_stream stream);
/* Temporally blocked variant: all four stages of a step in one kernel (state read
 * once and written once per step, 32 B/site-step), same arithmetic and bits as
 * cf_rk4_run_f64 / cf_rk4_run_f32c. Steps ping-pong between the state and
 * work_dev (cf_rk4_tb_work_bytes); the result is always left in the state.
 * chunk = planes per block along a (<= 0: default). FP32-comp speed depends on the
 * relative placement of the arrays (up to 20% at N=512): the run places its second hi
 * and lo inside work at the caller's hi and lo addresses modulo 4 KB (work carries 8 KB
 * of slack for it), so hi and lo themselves should be allocated separately or at a
 * relative offset that is a multiple of 4 KB. */
size_t cf_rk4_tb_work_bytes(int64_t N, int precision);
cf_status cf_rk4_run_tb_f64(const cf_kg_params *p, double *x_dev, double t0, int64_t k0, int64_t n_steps,
                            double dt, int chunk, void *work_dev, cf_stream stream);
cf_status cf_rk4_run_tb_f32c(const cf_kg_params *p, float *hi_dev, float *lo_dev, double t0, int64_t k0,
                             int64_t n_steps, double dt, int chunk, void *work_dev, cf_stream stream);
/* Compensated storage with the single-precision stencil (no reference counterpart;
 * the run-A storage of the precision study with float right-hand sides): the state
 * is hi + lo as in cf_rk4_run_tb_f32c, each stage's force is evaluated in float on
 * the float stage inputs (seven-point sum, coefficients rounded to float, fused
 * multiply-adds), the stage inputs are fma(c, k, hi) in float, the increment
 * dt/6 k1 + dt/3 k2 + dt/3 k3 is accumulated in double and the new state
 * (hi + lo) + (increment + dt/6 k4) is formed in double and split into hi and lo. The
 * increment weights stay double, so the scheme has no drift from rounded step
 * lengths. Work buffer and placement as cf_rk4_run_tb_f32c
 * (cf_rk4_tb_work_bytes(N, CF_PREC_F32S)). The launch covers the lattice with 24 x 40
 * or 32 x 32 regions (16 x 32 or 24 x 24 tiles), whichever the chunk cost model favours
 * for N and the device's CU count; both give the same bits. */
cf_status cf_rk4_run_tb_f32s(const cf_kg_params *p, float *hi_dev, float *lo_dev, double t0, int64_t k0,
                             int64_t n_steps, double dt, int chunk, void *work_dev, cf_stream stream);
/* Pacing of the single-precision-stencil step (no reference counterpart). Its blocks
 * run faster when held to a common schedule of plane iterations (neighbouring tiles
 * then read their shared halo rows while these are still in the L2). The pace, in ticks
 * of the GPU's 100 MHz real-time counter per plane iteration (0: unpaced), is kept per
 * device, lattice size, chunk and region shape and used by cf_rk4_run_tb_f32s and
 * cf_rk4_step_tb_f32s (and by cf_rk4_step_tb_f32s_slab where its slab gets the same
 * chunk and shape).
 * cf_rk4_tb_tune_pace times about 350 steps of the lattice of p at paces 0..3.2 us per
 * plane (state hi/lo at time t, unchanged; work as for cf_rk4_run_tb_f32s) and keeps
 * the pace P whose slowest of P-2..P is fastest (the rate falls just below the best
 * pace, so the pick keeps a margin above it; unpaced if that is as fast), and
 * returns it in *pace_out (may be null); it synchronises the stream. With chunk <= 0
 * it does so for each chunk length whose launch the cost model rates within 5% of its
 * best (at N=512 four of them, about 1400 steps in all), keeps every chunk's pace, and
 * makes the chunk with the fastest score the default for N on this device (the chunk
 * that chunk <= 0 means from then on in the f32s runs, steps and pace calls), its pace
 * the one returned. Pacing and chunking change timing only, never results.
 * cf_rk4_tb_get_pace returns -1 for an unsupported N. */
cf_status cf_rk4_tb_tune_pace(const cf_kg_params *p, const float *hi_dev, const float *lo_dev, double t, double dt,
                              int chunk, void *work_dev, cf_stream stream, int *pace_out);
cf_status cf_rk4_tb_set_pace(int64_t N, int chunk, int pace);
int cf_rk4_tb_get_pace(int64_t N, int chunk);
/* Default chunk (planes per block along a) of the f32s step for N on the current device
 * (no reference counterpart): cf_rk4_tb_set_chunk sets it (0: back to the launch cost
 * model's), cf_rk4_tb_get_chunk returns the chunk that chunk <= 0 resolves to (tuned,
 * set or the cost model's), -1 for an unsupported N. */
cf_status cf_rk4_tb_set_chunk(int64_t N, int chunk);
int cf_rk4_tb_get_chunk(int64_t N);
/* Region shape (rows along b x columns along c, halo included) over which
 * cf_rk4_run_tb_f32s and cf_rk4_step_tb_f32s cover an N^3 lattice on the current device:
 * 24 x 40 or 32 x 32 (no reference counterpart; the kernel's name carries it). */
cf_status cf_rk4_tb_f32s_region(int64_t N, int *rows, int *cols);
/* cf_rk4_run_tb_f64 without a work buffer and without any allocation. No counterpart
 * in the reference: it is for callers that cannot allocate device memory (code
 * under hipGraph capture, serving loops, harnesses that forbid hipMalloc). The
 * ping-pong buffer is a static device array of the library (32 MiB, one per GPU,
 * held until the process ends), so N <= CF_NOALLOC_MAX_N; larger N returns
 * CF_E_INVALID. All calls on one GPU share that buffer: calls that overlap on one
 * GPU (other streams, other host threads) race; calls ordered on one stream do not.
 * The buffer needs no clearing (every step overwrites all of it). Same results as
 * cf_rk4_run_tb_f64, bit for bit. */
#define CF_NOALLOC_MAX_N 128
cf_status cf_rk4_run_tb_f64_noalloc(const cf_kg_params *p, double *x_dev, double t0, int64_t k0, int64_t n_steps,
                                    double dt, int chunk, cf_stream stream);
/* One temporally blocked step out of place (x_out = step(x_in); the buffers must not
 * overlap; no work buffer): the form a stepper that swaps its state buffers uses. */
cf_status cf_rk4_step_tb_f64(const cf_kg_params *p, const double *x_in_dev, double *x_out_dev, double t, double dt,
                             int chunk, cf_stream stream);
cf_status cf_rk4_step_tb_f32c(const cf_kg_params *p, const float *hi_in_dev, const float *lo_in_dev,
                              float *hi_out_dev, float *lo_out_dev, double t, double dt, int chunk,
                              cf_stream stream);
cf_status cf_rk4_step_tb_f32s(const cf_kg_params *p, const float *hi_in_dev, const float *lo_in_dev,
                              float *hi_out_dev, float *lo_out_dev, double t, double dt, int chunk,
                              cf_stream stream);
/* One temporally blocked FP64 step of a slab of an N^3 lattice split along a: the
 * share of one GPU in a slab decomposition (no reference counterpart; the reference
 * runs one lattice on one device). x_in_dev and x_out_dev hold n_planes + 8 planes of
 * phi, then as many of pi: [phi (n_planes + 8) N^2 | pi (n_planes + 8) N^2]. Planes
 * 4 .. n_planes + 3 are the slab's own (lattice planes a0 .. a0 + n_planes - 1);
 * planes 0..3 and n_planes + 4 .. n_planes + 7 are ghost copies of the neighbours'
 * planes a0 - 4 .. a0 - 1 and a0 + n_planes .. a0 + n_planes + 3 (periodic), filled
 * before the step. Writes the slab's own planes of x_out_dev and leaves its ghost
 * planes as they were; b and c are periodic as usual. Each own site gets the bits of
 * cf_rk4_step_tb_f64 on the whole lattice. n_planes >= 1; the buffers must not
 * overlap. */
cf_status cf_rk4_step_tb_f64_slab(const cf_kg_params *p, int64_t n_planes, const double *x_in_dev,
                                  double *x_out_dev, double t, double dt, int chunk, cf_stream stream);
/* The same for FP32-comp storage: hi and lo each laid out as the FP64 slab above (n_planes
 * + 8 planes of phi, then of pi). Each own site gets the bits of cf_rk4_step_tb_f32c;
 * no two of the four buffers may overlap. */
cf_status cf_rk4_step_tb_f32c_slab(const cf_kg_params *p, int64_t n_planes, const float *hi_in_dev,
                                   const float *lo_in_dev, float *hi_out_dev, float *lo_out_dev, double t,
                                   double dt, int chunk, cf_stream stream);
/* The same for the float-force form (bits of cf_rk4_step_tb_f32s). It is paced with the
 * pace kept for (N, its chunk): tune it on the whole lattice with the same chunk
 * (cf_rk4_tb_tune_pace) or set it (cf_rk4_tb_set_pace); otherwise it runs unpaced. */
cf_status cf_rk4_step_tb_f32s_slab(const cf_kg_params *p, int64_t n_planes, const float *hi_in_dev,
                                   const float *lo_in_dev, float *hi_out_dev, float *lo_out_dev, double t,
                                   double dt, int chunk, cf_stream stream);

/* Split a double state into hi/lo floats and back (x = hi + lo). */
cf_status cf_split_f64(const double *x_dev, float *hi_dev, float *lo_dev, int64_t n, cf_stream stream);
cf_status cf_join_f64(const float *hi_dev, const float *lo_dev, double *x_dev, int64_t n, cf_stream stream);

/* ---------------------------------------------------------------- ensembles
 * Seeded realisations of uneven N advanced together (no reference counterpart:
 * the reference runs one lattice per process, src/main.cpp:82-196). One persistent
 * kernel per step walks all realisations' work units; a run can be captured once
 * into a hipGraph and replayed. Each realisation's result equals a single-lattice
 * run of the same steps bit for bit (cf_rk4_run_tb_f64 / _f32c / _f32s for precision
 * CF_PREC_F64 / F32C / F32S). All realisations share m, lambda, cosmology. */
typedef struct cf_ensemble cf_ensemble;
cf_status cf_ensemble_create(int n, const int64_t *N, const double *L, double m, double lambda,
                             const cf_cosmology *cosmo, int precision, cf_ensemble **e);
cf_status cf_ensemble_destroy(cf_ensemble *e);
int64_t cf_ensemble_sites(const cf_ensemble *e); /* sum of N^3 */
/* Current state of realisation r: x (FP64) or hi (F32C, F32S) and lo (F32C, F32S, else
 * NULL).
 * Valid until the next run (states ping-pong). */
cf_status cf_ensemble_state(cf_ensemble *e, int r, void **x_dev, float **lo_dev);
/* Initial conditions: realisation r gets the GRF of seeds[r] (cf_init_grf_state). */
cf_status cf_ensemble_init_grf(cf_ensemble *e, const uint32_t *seeds, int kind, const double *params, double rms,
                               cf_stream stream);
/* Work plan: planes per unit for lattices of size N (N <= 0: keep) and the grid:
 * > 0: that many persistent blocks, each running an LPT-balanced list of units;
 * CF_ENS_GRID_FLAT: one block per unit, longest first (the dispatcher balances);
 * 0: keep. */
#define CF_ENS_GRID_FLAT (-1)
cf_status cf_ensemble_set_plan(cf_ensemble *e, int64_t N, int chunk, int grid);
cf_status cf_ensemble_get_plan(const cf_ensemble *e, int64_t N, int *chunk, int *grid);
/* Region shape of the blocks that advance lattices of size N (rows along b x columns
 * along c): CF_ENS_REGION_32X32 (24 x 24 output tiles, the default) or
 * CF_ENS_REGION_24X40 (16 x 32 tiles, which divide N = 64 and 128; 24 x 24 tiles
 * leave 27% of the covered area outside such lattices). Same bits either way. */
#define CF_ENS_REGION_32X32 0
#define CF_ENS_REGION_24X40 1
cf_status cf_ensemble_set_region(cf_ensemble *e, int64_t N, int region);
cf_status cf_ensemble_get_region(const cf_ensemble *e, int64_t N, int *region);
/* Autotuner: times candidate grids (persistent and flat), then per N candidate
 * region shapes and chunks, on the device (a step reads one buffer and writes the
 * other, so the state is untouched), and keeps the fastest plan; a CF_PREC_F32S
 * ensemble also tunes its pace (as cf_rk4_tb_tune_pace) before revisiting the plans.
 * Blocking. */
cf_status cf_ensemble_tune(cf_ensemble *e, double dt, int reps, int *grid_out);
/* Pace of a CF_PREC_F32S ensemble's units (ticks of the 100 MHz real-time counter per
 * plane iteration, 0: unpaced; see cf_rk4_tb_tune_pace). Changes timing only, never
 * results; other precisions accept only 0. cf_ensemble_get_pace returns -1 for null. */
cf_status cf_ensemble_set_pace(cf_ensemble *e, int pace);
int cf_ensemble_get_pace(const cf_ensemble *e);
/* n_steps RK4 steps at t_k = t0 + (k0 + k) dt; use_graph: capture (or reuse) a hipGraph.
 * Asynchronous: ordered after earlier work on `stream`, and later work on `stream`
 * waits for it. Consecutive runs (segments of one evolution) need no host
 * synchronisation between them; the graph is re-captured whenever the step table
 * changes (another t0, k0, n_steps or dt, an eager run, a new plan). */
cf_status cf_ensemble_run(cf_ensemble *e, double t0, int64_t k0, int64_t n_steps, double dt, int use_graph,
                          cf_stream stream);
/* Energy-density power spectrum of realisation r at time t into bins_dev. */
cf_status cf_ensemble_density_spectrum(cf_ensemble *e, int r, double t, double *bins_dev, cf_stream stream);

/* ---------------------------------------------------------------- adjoint
 * Discrete adjoint of the RK4 evolution and of the spectral loss
 * L = sum_s w_s P_rho(s) (P_rho: compute_power_spectrum of the energy density,
 * fdm3d.cpp:7-34, at the final time; w_s given per bin in weights_dev). No
 * reference counterpart (the reference has no gradients); the forward pieces are
 * the ones above, so gradients are of exactly the computed quantities. */
size_t cf_adjoint_work_bytes(int64_t N, int64_t n_steps, int64_t ckpt);
/* L, dL/dx (gx_dev, 2N^3) and the direct dL/dm at state x and time t.
 * scratch_dev: N^3 doubles. Blocking. */
cf_status cf_loss_spectrum_adjoint(cf_spectrum_plan *plan, const cf_kg_params *p, const double *x_dev, double t,
                                   const double *weights_dev, double *gx_dev, double *scratch_dev, double *loss,
                                   double *dldm, cf_stream stream);
/* One step back: lam_dev holds dL/dx_{n+1} on entry, dL/dx_n on return, for the
 * step taken from x_dev at time t; dL/dm contributions go to part_dev (2048
 * doubles, summed by the caller in order). work_dev: 2N^3 doubles. */
cf_status cf_rk4_step_adjoint_f64(const cf_kg_params *p, const double *x_dev, double t, double dt, double *lam_dev,
                                  double *part_dev, void *work_dev, cf_stream stream);
/* Forward n_steps from x0 (checkpoint every ckpt steps), loss at t0 + n_steps dt,
 * backward: grad_x0_dev = dL/dx0 (2N^3), *loss, *dldm (all of m's paths). Blocking.
 * work_dev: cf_adjoint_work_bytes(N, n_steps, ckpt). */
cf_status cf_adjoint_spectrum_loss(cf_spectrum_plan *plan, const cf_kg_params *p, const double *x0_dev, double t0,
                                   double dt, int64_t n_steps, int64_t ckpt, const double *weights_dev,
                                   double *grad_x0_dev, double *loss, double *dldm, void *work_dev,
                                   cf_stream stream);

/* ---------------------------------------------------------------- energy density
 * Equation::compute_energy_density (equations.cpp:106-138; energy_density_kernel,
 * equations_cuda.cu:332-354 via :537-553 and :591-607): rho (N^3) from the state. */
cf_status cf_energy_density_f64(int64_t N, const double *x_dev, double *rho_dev, double m,
                                double lambda, double inv_ah2, cf_stream stream);
/* The observer's density average over a (src/observer.hpp:110-119,
 * rho.reshaped(N*N, N).rowwise().mean()): avg_dev[b*N + c] = (sum over a, in order) / N.
 * The a = 0 slice is simply the first N^2 entries of rho. */
cf_status cf_axis_average(int64_t N, const double *rho_dev, double *avg_dev, cf_stream stream);

/* ---------------------------------------------------------------- equation variants
 * The curvature / potential equations, FP64:
 *   CF_VAR_FIXED_CURVATURE      CudaFixedCurvatureEquationInFRW (equations_cuda.cu:114-137,
 *                               645-675); field = Psi (N^3, real space), copied.
 *   CF_VAR_SQRT_POTENTIAL       CudaSqrtPotentialEquationInFRW (equations_cuda.cu:83-106,
 *                               610-642); V = m^2 f_a^2 (sqrt(1 + phi^2/f_a^2) - 1).
 *   CF_VAR_COMOVING_CURVATURE   ComovingCurvatureEquationInFRW (equations.cpp:141-322) /
 *                               CudaComovingCurvatureEquationInFRW (equations_cuda.cu:678-761);
 *                               field = R_fft, the unnormalised R2C of the comoving
 *                               curvature (N*N*(N/2+1) complex), copied. Psi, dPsi/dt
 *                               are rebuilt at every distinct stage time.
 *   CF_VAR_APPROX_COMOVING_CURVATURE  CudaApproximateComovingCurvatureEquationInFRW
 *                               (equations_cuda.cu:299-329, 764-815): Psi on an M^3 grid
 *                               from the cut-off spectrum (compute_cutoff_fouriers,
 *                               fdm3d_cuda.cu:117-141, 206-219), Taylor-expanded exponentials;
 *                               M must divide N. field = R_fft of the N^3 lattice.
 * The comoving variants need N (and M) <= 1024. A handle keeps a small cache of Psi
 * fields keyed by time: use it from one stream at a time. */
#define CF_VAR_FIXED_CURVATURE 1
#define CF_VAR_SQRT_POTENTIAL 2
#define CF_VAR_COMOVING_CURVATURE 3
#define CF_VAR_APPROX_COMOVING_CURVATURE 4
typedef struct cf_variant_s *cf_variant;
/* p->lambda is ignored. f_a: sqrt potential only. M: approximate variant only. Blocking. */
cf_status cf_variant_create(const cf_kg_params *p, int kind, double f_a, int64_t M, const double *field_dev,
                            cf_variant *out, cf_stream stream);
void cf_variant_destroy(cf_variant v);
/* Equation::operator()(x, dxdt, t): the odeint system call. */
cf_status cf_variant_rhs(cf_variant v, const double *x_dev, double *dxdt_dev, double t, cf_stream stream);
/* One / n fused classic RK4 steps (as cf_rk4_step_f64 / cf_rk4_run_f64);
 * work_dev: cf_rk4_work_bytes(N, CF_PREC_F64). */
cf_status cf_variant_rk4_step(cf_variant v, double *x_dev, double t, double dt, void *work_dev, cf_stream stream);
cf_status cf_variant_rk4_run(cf_variant v, double *x_dev, double t0, int64_t k0, int64_t n_steps, double dt,
                             void *work_dev, cf_stream stream);
/* The same steps temporally blocked (one kernel per step, as cf_rk4_run_tb_f64) for
 * every variant; same bits as cf_variant_rk4_run. The comoving variants build Psi,
 * dPsi/dt at t, t+dt/2 and t+dt first (two new transforms per step, as the stage
 * kernels). work_dev: cf_rk4_tb_work_bytes(N, CF_PREC_F64). */
cf_status cf_variant_rk4_run_tb(cf_variant v, double *x_dev, double t0, int64_t k0, int64_t n_steps, double dt,
                                int chunk, void *work_dev, cf_stream stream);
cf_status cf_variant_rk4_step_tb(cf_variant v, const double *x_in_dev, double *x_out_dev, double t, double dt,
                                 int chunk, cf_stream stream);
/* Equation::compute_energy_density (equations_cuda.cu:422-473, equations.cpp:262-322). */
cf_status cf_variant_energy_density(cf_variant v, const double *x_dev, double *rho_dev, double t,
                                    cf_stream stream);
/* Comoving variants: Psi(t) and dPsi/dt(t) in real space (G^3 each, G = N or M). */
cf_status cf_variant_psi(cf_variant v, double t, double *psi_dev, double *dpsi_dev, cf_stream stream);
/* compute_Comoving_Curvature_Psi_dPsidt_fft (equations.cpp:141-183): half spectra of a G^3
 * grid; factor = (2 pi/L) eta(t)/sqrt(3). */
cf_status cf_comoving_psi_fft(int64_t G, const double *R_fft_dev, double *psi_fft_dev, double *dpsi_fft_dev,
                              double factor, double H, cf_stream stream);
/* compute_cutoff_fouriers (fdm3d.cpp:105-130; fdm3d_cuda.cu:117-141): half spectrum of N^3
 * -> half spectrum of M^3, divided by (N/M)^3. */
cf_status cf_cutoff_fouriers(int64_t N, int64_t M, const double *in_dev, double *out_dev, cf_stream stream);

/* ---------------------------------------------------------------- WKB extrapolation
 * WKBSolutionForKleinGordonEquationInFRW (src/wkb.cpp:5-121): each Fourier mode of the
 * state at t_i advanced to a later t by its closed-form WKB phase, for a
 * radiation-era cosmology (p = 1, t1 = 1/(2 H1), as wkb.cpp:7-8 asserts) and m > 0.
 * The handle keeps the scaled half spectra of the state; evaluate writes the
 * extrapolated state (2N^3) for any t. N <= 1024. */
typedef struct cf_wkb_s *cf_wkb;
cf_status cf_wkb_create(const cf_kg_params *p, const double *state_dev, double t_i, cf_wkb *out, cf_stream stream);
cf_status cf_wkb_evaluate(cf_wkb w, double t, double *state_out_dev, cf_stream stream);
void cf_wkb_destroy(cf_wkb w);
/* The per-s tables the evaluation uses (compute_phase_integral and omega_eff,
 * wkb.cpp:40-96): cos(phase), sin(phase), sqrt(omega_eff(t_i)); phase_host may be NULL.
 * Each table holds cf_spectrum_bins(N) doubles. */
cf_status cf_wkb_tables_host(int64_t N, double L, double m, double a1, double H1, double t_i, double t_f,
                             double *cos_host, double *sin_host, double *sqrt_omega_host, double *phase_host);

/* ---------------------------------------------------------------- FFT and spectra
 * A spectrum plan owns the hipFFT D2Z/Z2D plans of an N^3 lattice with one shared
 * work area (cufftWrapperNoBatching, src/cuda_wrapper.cu:202-255), scratch buffers,
 * and the bin table used by the deterministic binning kernels. */
cf_status cf_spectrum_plan_create(int64_t N, cf_spectrum_plan **plan);
cf_status cf_spectrum_plan_destroy(cf_spectrum_plan *plan);
int64_t cf_spectrum_bins(int64_t N); /* 3*(N/2)^2 + 1 */

/* execute_d2z / execute_z2d (cuda_wrapper.cu:230-235, 245-250; fftw_wrapper.cpp:24-49):
 * out = unnormalised R2C of in (N*N*(N/2+1) interleaved complex); in is preserved.
 * z2d: out = unnormalised C2R of in; in is destroyed (as FFTW / cuFFT). */
cf_status cf_fft_d2z(cf_spectrum_plan *plan, const double *in_dev, double *out_dev, cf_stream stream);
cf_status cf_fft_z2d(cf_spectrum_plan *plan, double *in_dev, double *out_dev, cf_stream stream);
/* execute_inplace_z2d (cuda_wrapper.cu:252-254; fftw_wrapper.hpp:64, fftw_wrapper.cpp:63-68):
 * inout_dev holds the half spectrum (2 N^2 (N/2+1) doubles) and receives its
 * unnormalised C2R in the FFTW padded layout: site (a, b, c) at
 * 2(N/2+1)(N a + b) + c (PADDED_IDX_OF, fdm3d.hpp:24); the two padding doubles of
 * each row are unspecified. */
cf_status cf_fft_z2d_inplace(cf_spectrum_plan *plan, double *inout_dev, cf_stream stream);

/* compute_power_spectrum (src/fdm3d.cpp:7-34; fdm3d_cuda.cu:161-176):
 * bins_dev[s] = sum over the full k-cube with |n|^2 = s of |f_k|^2. Deterministic. */
cf_status cf_power_spectrum(cf_spectrum_plan *plan, const double *f_dev, double *bins_dev,
                            cf_stream stream);
/* The same bins, with f_dev handed to the transform as its input without the copy
 * that preserves it: f_dev's contents are undefined afterwards (rocFFT may use an
 * out-of-place transform's input as scratch). For a caller that owns f_dev and is
 * done with it, e.g. the observer's energy density once its slice and axis average
 * are taken (observer.hpp:102-119): saves an N^3 copy per spectrum. */
cf_status cf_power_spectrum_consume(cf_spectrum_plan *plan, double *f_dev, double *bins_dev, cf_stream stream);
/* compute_mode_power_spectrum (fdm3d.cpp:36-72; fdm3d_cuda.cu:144-159):
 * bins_dev[s] = sum |phi_k|^2 + |dphi_k|^2 / (m^2 + (2 pi/L)^2 s / a^2). */
cf_status cf_mode_power_spectrum(cf_spectrum_plan *plan, const double *state_dev, double L, double m,
                                 double a_t, double *bins_dev, cf_stream stream);
/* The binning kernels alone, on half-spectra (sum_power_kernel / sum_mode_power_kernel,
 * fdm3d_cuda.cu:3-62, without atomics). */
cf_status cf_bin_power(cf_spectrum_plan *plan, const double *fk_dev, double *bins_dev, cf_stream stream);
cf_status cf_bin_mode_power(cf_spectrum_plan *plan, const double *phik_dev, const double *dphik_dev,
                            double L, double m, double a_t, double *bins_dev, cf_stream stream);
/* Radial bins (no reference counterpart; plot_util.py:54-74 rebins s offline):
 * radial_dev[j] = sum of s-bins with floor(sqrt(s)) = j, j < N/2. */
cf_status cf_radial_bins(cf_spectrum_plan *plan, const double *sbins_dev, double *radial_dev,
                         cf_stream stream);

/* ---------------------------------------------------------------- Gaussian random fields
 * RandomNormal + generate_inhomogeneous_gaussian_random_field (src/random_field.cpp:13-89)
 * and the spectrum factories (random_field.cpp:212-254). Normals come from a freshly
 * seeded std::mt19937 + std::normal_distribution<double> on the host. */
#define CF_SPEC_GAUSS 0           /* exp(-k^2/(2 sigma^2)), P(0)=0; params {sigma} */
#define CF_SPEC_KM3_GAUSS 1       /* k^-3 exp(-k^2/2), P(0)=0 */
#define CF_SPEC_SCALE_INV 2       /* scale_invariant_spectrum_3d; params {A_s} */
#define CF_SPEC_POWER_CUTOFF 3    /* power_law_with_cutoff_given_amplitude_3d; {sigma, k_ast, alpha} */
#define CF_SPEC_DERIV_POWER_CUTOFF 4 /* to_deriv_spectrum(m, kind 3); {sigma, k_ast, alpha, m} */
cf_status cf_normals_host(uint32_t seed, int64_t n, double *out_host);
cf_status cf_spectrum_table_host(int kind, int64_t N, double L, const double *params,
                                 double *P_by_s_host);
/* out_dev = GRF from white noise w_dev (N^3, already scaled by 1/N^3 as random_field.cpp:44-46)
 * and sqrt(P) per s (sqrtP_by_s_dev, cf_spectrum_bins(N) entries); Psi_dev may be NULL.
 * w_dev is overwritten. */
cf_status cf_grf(cf_spectrum_plan *plan, double *w_dev, const double *sqrtP_by_s_dev,
                 const double *Psi_dev, double *out_dev, cf_stream stream);
/* compute_field_with_scaled_fourier_modes (src/fdm3d.cpp:76-102): out = C2R(R2C(f) * kernel(s) / N^3),
 * kernel tabulated over s = |n|^2 (cf_spectrum_bins(N) entries, k = sqrt(s) 2 pi / L). f is preserved. */
cf_status cf_scaled_fourier_modes(cf_spectrum_plan *plan, const double *f_dev, const double *kernel_by_s_dev,
                                  double *out_dev, cf_stream stream);
/* x *= target / rms(x), rms from a fixed-order device reduction (bit-reproducible). */
cf_status cf_rescale_rms(double *x_dev, int64_t n, double target, cf_stream stream);
/* Workload initial condition: phi = GRF(seed, spectrum kind/params) rescaled to rms
 * (rms <= 0: no rescale), dphi/dt = 0. state_dev holds 2N^3 doubles. Blocking. */
cf_status cf_init_grf_state(cf_spectrum_plan *plan, uint32_t seed, int kind, const double *params,
                            double L, double rms, double *state_dev, cf_stream stream);

#ifdef __cplusplus
}
#endif
#endif /* CFHIP_H */
