// cf_tb_run.hpp -- library-internal entry of the temporally blocked step for the
// equation variants (defined in cf_rk4_tb.hip, used by cf_variants.hip).
#pragma once

#include "cf_common.hpp"
#include "cf_tb_body.hpp"

namespace cf {

// One FP64 step of the fixed-curvature (var = 1, psi: N^3), sqrt-potential (var = 2,
// f_a), comoving-curvature (var = 3) or coarse-grid comoving (var = 4) equation from
// xin into xout (2N^3 doubles each, not overlapping), one kernel; the comoving ones
// take cps with Psi | dPsi at t, t + dt/2 and t + dt. Same bits as the variants'
// four stage kernels.
cf_status tb_step_variant(int var, const cf_kg_params *p, double f_a, const double *psi, const double *xin,
                          double *xout, double t, double dt, int chunk, hipStream_t st, const TbPsi *cps);

}  // namespace cf
