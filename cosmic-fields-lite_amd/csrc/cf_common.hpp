// cf_common.hpp -- shared helpers of libcfhip.so: status reporting, checks,
// launch geometry. Internal; not installed.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "cfhip.h"

namespace cf {

// Thread-local message of the last failing call (cf_last_error).
void set_error(const char *fmt, ...);

#define CF_REQUIRE(cond, ...)                \
  do {                                       \
    if (!(cond)) {                           \
      ::cf::set_error(__VA_ARGS__);          \
      return CF_E_INVALID;                   \
    }                                        \
  } while (0)

#define CF_HIP(expr)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) {                                                                 \
      ::cf::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__,      \
                      __LINE__);                                                            \
      return CF_E_HIP;                                                                      \
    }                                                                                       \
  } while (0)

// Check the launch that was just enqueued.
#define CF_LAUNCHED() CF_HIP(hipGetLastError())

#define CF_TRY(expr)                   \
  do {                                 \
    cf_status s_ = (expr);             \
    if (s_ != CF_OK) return s_;        \
  } while (0)

inline hipStream_t S(cf_stream s) { return reinterpret_cast<hipStream_t>(s); }

// Lattice sizes the kernels support: even N (random_field.cpp:37 requires it),
// and N^3 small enough that 2N^3 indexes fit comfortably in int64.
inline bool valid_N(int64_t N) { return N >= 4 && N % 2 == 0 && N <= 2048; }

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Scratch owned by a spectrum plan (cf_spectrum.hip), for library code that runs
// with a plan: a reduction buffer of 1025 doubles and one spectrum of
// cf_spectrum_bins(N) doubles. Calls that share a plan are ordered on one stream.
double *plan_red(cf_spectrum_plan *p);
double *plan_bins(cf_spectrum_plan *p);
int64_t plan_size(const cf_spectrum_plan *p);  // its lattice size N

}  // namespace cf
