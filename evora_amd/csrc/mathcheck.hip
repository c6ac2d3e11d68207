// mathcheck.hip -- the device arithmetic of the fast forms, evaluated on host
// arrays (evh_math_eval), so that tests can hold it against the host's IEEE
// operations and libm: the general power exp(y ln x), the powers of g1 the
// fluxes and the dual-energy branches take, the shared-reciprocal quotient,
// the correctly rounded reciprocal, the square root, and the FP32 power.
// Compiled like the fused kernels (contraction on), so the functions are
// the ones those kernels inline.
#include <cmath>

#include "evh_internal.h"
#include "physics_fast.h"

namespace {

using namespace evh;

// the KParams fields the powers of g1 read, for a given g1 (as
// make_kparams sets them)
__device__ KParams params_g1(double g1) {
  KParams P = {};
  P.g1 = g1;
  const double delta = 2.0 / 3.0 - g1;
  P.g1_third = fabs(delta) <= 1e-6 ? 1 : 0;
  P.g1_dln2 = delta * 0.69314718055994530942;
  return P;
}

__global__ void __launch_bounds__(256) k_math(int fn, long long n, const double *__restrict__ x,
                                              const double *__restrict__ y, double *__restrict__ out) {
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const double a = x[i], b = y[i];
    double r = 0.0;
    switch (fn) {
      case EVH_MATH_POW: r = fm::pow_pos(a, b); break;
      case EVH_MATH_POW_G1: r = fast::pow_g1(params_g1(b), a); break;
      case EVH_MATH_RPOW_G1: r = fast::rpow_g1(params_g1(b), a); break;
      case EVH_MATH_DIV: r = fm::div_cr(a, b, fm::rcp(b)); break;
      case EVH_MATH_RCP: r = fm::rcp_cr(a); break;
      case EVH_MATH_SQRT: r = fm::sqrt_pos(a); break;
      case EVH_MATH_POWF: r = double(fm::pow_pos(float(a), float(b))); break;
      default: r = __builtin_nan(""); break;
    }
    out[i] = r;
  }
}

}  // namespace

extern "C" int evh_math_eval(int fn, long long n, const double *x, const double *y, double *out) {
  if (fn < EVH_MATH_POW || fn > EVH_MATH_POWF || n < 0 || (n > 0 && (!x || !y || !out))) return EVH_ERR_CONFIG;
  if (n == 0) return EVH_OK;
  double *d = nullptr;
  const size_t bytes = sizeof(double) * (size_t)n;
  if (hipMalloc(&d, 3 * bytes) != hipSuccess) return EVH_ERR_HIP;
  hipError_t e = hipMemcpy(d, x, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d + n, y, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    const long long blocks = std::min((n + 255) / 256, 4096LL);
    hipLaunchKernelGGL(k_math, dim3((int)blocks), dim3(256), 0, 0, fn, n, d, d + n, d + 2 * n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(out, d + 2 * n, bytes, hipMemcpyDeviceToHost);
  hipFree(d);
  return e == hipSuccess ? EVH_OK : EVH_ERR_HIP;
}
