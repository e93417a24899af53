// cf_tb_run.hpp -- library-internal entry of the temporally blocked step for the
// equation variants (defined in cf_rk4_tb.hip, used by cf_variants.hip).
#pragma once

#include "cf_common.hpp"

namespace cf {

// n_steps FP64 steps of the fixed-curvature (var = 1, Psi: N^3) or sqrt-potential
// (var = 2, f_a) equation, one kernel per step; same bits as the variants' four
// stage kernels. work: cf_rk4_tb_work_bytes(N, CF_PREC_F64).
// One such step from xin into xout (2N^3 doubles each, not overlapping).
cf_status tb_step_variant(int var, const cf_kg_params *p, double f_a, const double *psi, const double *xin,
                          double *xout, double t, double dt, int chunk, hipStream_t st);
cf_status tb_run_variant(int var, const cf_kg_params *p, double f_a, const double *psi, double *x, double t0,
                         int64_t k0, int64_t n_steps, double dt, int chunk, void *work, hipStream_t st);

}  // namespace cf
