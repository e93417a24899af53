// cf_core.hip -- status, device/memory entry points, odeint vector operations
// (scale_sumK, norm_inf) and the host cosmology helpers.
#include <cmath>
#include <cstring>
#include <vector>

#include "cf_common.hpp"

namespace cf {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

// ---------------------------------------------------------------- kernels
__global__ void k_fill_zero_u32(uint32_t *p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = 0u;
}

__global__ void k_fill_zero_u8(uint8_t *p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// v0 = sum alpha_i v_i, left to right (thrust_operations.cu:4-189 evaluation
// order; each product is rounded before the sum: the library is built with
// -ffp-contract=off).
struct SumArgs {
  const double *v[14];
  double a[14];
};

template <int K>
__global__ void k_scale_sum(double *v0, SumArgs args, int64_t n) {  // v0 may alias an input (odeint passes x as both)
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    double acc = args.a[0] * args.v[0][i];
#pragma unroll
    for (int k = 1; k < K; ++k) acc = acc + args.a[k] * args.v[k][i];
    v0[i] = acc;
  }
}

__global__ void k_absmax_partial(const double *__restrict__ v, int64_t n, double *__restrict__ part) {
  __shared__ double sm[256];
  double m = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    m = fmax(m, fabs(v[i]));
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// Number of NaN and infinite values (exponent field all ones, independent of any
// floating-point mode), one partial count per block.
__global__ void k_nonfinite_partial(const double *__restrict__ v, int64_t n, unsigned long long *__restrict__ part) {
  __shared__ unsigned long long sm[256];
  unsigned long long k = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    k += ((__double_as_longlong(v[i]) >> 52) & 0x7ff) == 0x7ff;
  sm[threadIdx.x] = k;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

static int grid_for(int64_t n, int block, int max_blocks = 65536) {
  int64_t g = cdiv(n, block);
  if (g > max_blocks) g = max_blocks;
  if (g < 1) g = 1;
  return (int)g;
}

}  // namespace cf

using namespace cf;

extern "C" {

const char *cf_last_error(void) { return cf::g_err; }

const char *cf_version(void) { return "cfhip 0.1 (gfx950)"; }

cf_status cf_device_count(int *count) {
  CF_REQUIRE(count, "cf_device_count: null pointer");
  CF_HIP(hipGetDeviceCount(count));
  return CF_OK;
}

cf_status cf_set_device(int device) {
  CF_HIP(hipSetDevice(device));
  return CF_OK;
}

cf_status cf_sync(cf_stream stream) {
  CF_HIP(hipStreamSynchronize(S(stream)));
  return CF_OK;
}

cf_status cf_stream_create(cf_stream *stream) {
  CF_REQUIRE(stream, "cf_stream_create: null pointer");
  hipStream_t s;
  CF_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = reinterpret_cast<cf_stream>(s);
  return CF_OK;
}

cf_status cf_stream_destroy(cf_stream stream) {
  if (stream) CF_HIP(hipStreamDestroy(S(stream)));
  return CF_OK;
}

cf_status cf_alloc(size_t bytes, void **ptr) {
  CF_REQUIRE(ptr, "cf_alloc: null pointer");
  *ptr = nullptr;
  if (bytes == 0) return CF_OK;
  hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess) {
    set_error("cf_alloc(%zu bytes): %s", bytes, hipGetErrorString(e));
    *ptr = nullptr;
    return CF_E_NOMEM;
  }
  return CF_OK;
}

cf_status cf_free(void *ptr) {
  if (ptr) CF_HIP(hipFree(ptr));
  return CF_OK;
}

cf_status cf_h2d(void *dst, const void *src, size_t bytes, cf_stream stream) {
  if (bytes == 0) return CF_OK;
  CF_REQUIRE(dst && src, "cf_h2d: null pointer");
  CF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S(stream)));
  CF_HIP(hipStreamSynchronize(S(stream)));
  return CF_OK;
}

cf_status cf_d2h(void *dst, const void *src, size_t bytes, cf_stream stream) {
  if (bytes == 0) return CF_OK;
  CF_REQUIRE(dst && src, "cf_d2h: null pointer");
  CF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S(stream)));
  CF_HIP(hipStreamSynchronize(S(stream)));
  return CF_OK;
}

cf_status cf_d2d(void *dst, const void *src, size_t bytes, cf_stream stream) {
  if (bytes == 0) return CF_OK;
  CF_REQUIRE(dst && src, "cf_d2d: null pointer");
  CF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, S(stream)));
  return CF_OK;
}

cf_status cf_fill_zero(void *ptr, size_t bytes, cf_stream stream) {
  if (bytes == 0) return CF_OK;
  CF_REQUIRE(ptr, "cf_fill_zero: null pointer");
  const uintptr_t addr = reinterpret_cast<uintptr_t>(ptr);
  if (addr % 4 == 0 && bytes % 4 == 0) {
    int64_t n = (int64_t)(bytes / 4);
    hipLaunchKernelGGL(k_fill_zero_u32, dim3(grid_for(n, 256)), dim3(256), 0, S(stream),
                       (uint32_t *)ptr, n);
  } else {
    int64_t n = (int64_t)bytes;
    hipLaunchKernelGGL(k_fill_zero_u8, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, S(stream),
                       (uint8_t *)ptr, n);
  }
  CF_LAUNCHED();
  return CF_OK;
}

// ---------------------------------------------------------------- cosmology (host)
double cf_cosmo_a(const cf_cosmology *c, double t) {
  return c->a1 * pow(1 + (1 + 1 / c->p) * c->H1 * (t - c->t1), c->p / (1 + c->p));
}
double cf_cosmo_H(const cf_cosmology *c, double t) {
  return c->H1 * pow(1 + (1 + 1 / c->p) * c->H1 * (t - c->t1), -1);
}
double cf_cosmo_eta(const cf_cosmology *c, double t) {
  return (c->p / (c->a1 * c->H1)) * pow(1 + (1 + 1 / c->p) * c->H1 * (t - c->t1), 1 / (1 + c->p));
}
double cf_inv_ah_sqr(int64_t N, double L, double a_t) {
  return 1.0 / ((L / N) * (L / N)) / (a_t * a_t);
}

// ---------------------------------------------------------------- odeint ops
cf_status cf_scale_sum(int K, double *v0, const double *const *v, const double *alpha, int64_t n,
                       cf_stream stream) {
  CF_REQUIRE(K >= 1 && K <= 14, "cf_scale_sum: K=%d outside 1..14", K);
  CF_REQUIRE(v0 && v && alpha && n >= 0, "cf_scale_sum: null pointer or negative size");
  if (n == 0) return CF_OK;
  SumArgs args;
  std::memset(&args, 0, sizeof(args));
  for (int k = 0; k < K; ++k) {
    CF_REQUIRE(v[k], "cf_scale_sum: null input vector %d", k);
    args.v[k] = v[k];
    args.a[k] = alpha[k];
  }
  dim3 g(grid_for(n, 256, 1 << 20)), b(256);
  switch (K) {
#define CF_CASE(k) \
  case k: hipLaunchKernelGGL(k_scale_sum<k>, g, b, 0, S(stream), v0, args, n); break;
    CF_CASE(1) CF_CASE(2) CF_CASE(3) CF_CASE(4) CF_CASE(5) CF_CASE(6) CF_CASE(7)
    CF_CASE(8) CF_CASE(9) CF_CASE(10) CF_CASE(11) CF_CASE(12) CF_CASE(13) CF_CASE(14)
#undef CF_CASE
  }
  CF_LAUNCHED();
  return CF_OK;
}

cf_status cf_scale_sum2(double *v0, const double *v1, const double *v2, double a1, double a2,
                        int64_t n, cf_stream stream) {
  const double *v[2] = {v1, v2};
  const double al[2] = {a1, a2};
  return cf_scale_sum(2, v0, v, al, n, stream);
}

cf_status cf_scale_sum5(double *v0, const double *v1, const double *v2, const double *v3,
                        const double *v4, const double *v5, double a1, double a2, double a3,
                        double a4, double a5, int64_t n, cf_stream stream) {
  const double *v[5] = {v1, v2, v3, v4, v5};
  const double al[5] = {a1, a2, a3, a4, a5};
  return cf_scale_sum(5, v0, v, al, n, stream);
}

cf_status cf_norm_inf(const double *v, int64_t n, double *out, cf_stream stream) {
  CF_REQUIRE(v && out && n >= 0, "cf_norm_inf: bad arguments");
  *out = 0.0;
  if (n == 0) return CF_OK;
  const int nb = grid_for(n, 256, 1024);
  double *part = nullptr;
  CF_HIP(hipMallocAsync((void **)&part, sizeof(double) * nb, S(stream)));
  hipLaunchKernelGGL(k_absmax_partial, dim3(nb), dim3(256), 0, S(stream), v, n, part);
  hipError_t le = hipGetLastError();
  std::vector<double> h(nb);
  hipError_t ce = hipMemcpyAsync(h.data(), part, sizeof(double) * nb, hipMemcpyDeviceToHost, S(stream));
  hipError_t se = hipStreamSynchronize(S(stream));
  (void)hipFreeAsync(part, S(stream));
  CF_HIP(le);
  CF_HIP(ce);
  CF_HIP(se);
  double m = 0.0;
  for (double x : h) m = x > m ? x : m;
  *out = m;
  return CF_OK;
}

cf_status cf_count_nonfinite(const double *v, int64_t n, int64_t *count, cf_stream stream) {
  CF_REQUIRE(v && count && n >= 0, "cf_count_nonfinite: bad arguments");
  *count = 0;
  if (n == 0) return CF_OK;
  const int nb = grid_for(n, 256, 1024);
  unsigned long long *part = nullptr;
  CF_HIP(hipMallocAsync((void **)&part, sizeof(unsigned long long) * nb, S(stream)));
  hipLaunchKernelGGL(k_nonfinite_partial, dim3(nb), dim3(256), 0, S(stream), v, n, part);
  hipError_t le = hipGetLastError();
  std::vector<unsigned long long> h(nb);
  hipError_t ce = hipMemcpyAsync(h.data(), part, sizeof(unsigned long long) * nb, hipMemcpyDeviceToHost, S(stream));
  hipError_t se = hipStreamSynchronize(S(stream));
  (void)hipFreeAsync(part, S(stream));
  CF_HIP(le);
  CF_HIP(ce);
  CF_HIP(se);
  unsigned long long k = 0;
  for (unsigned long long x : h) k += x;
  *count = (int64_t)k;
  return CF_OK;
}

}  // extern "C"
