// Shared device/host helpers for the gfx950 Hpsi library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace inqamd {

// complex128 as two interleaved doubles (INQ's complex = thrust::complex<double>,
// src/math/complex.hpp:22-28).
using cplx = double2;

__host__ __device__ inline cplx mk(double re, double im) { return make_double2(re, im); }
__host__ __device__ inline cplx cadd(cplx a, cplx b) { return mk(a.x + b.x, a.y + b.y); }
__host__ __device__ inline cplx csub(cplx a, cplx b) { return mk(a.x - b.x, a.y - b.y); }
__host__ __device__ inline cplx cscale(double s, cplx a) { return mk(s * a.x, s * a.y); }
__host__ __device__ inline cplx cmul(cplx a, cplx b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__host__ __device__ inline cplx cmulconj(cplx a, cplx b) {
  return mk(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// multiply by -i (forward) or +i (backward): sign = -1 gives -i*a
__host__ __device__ inline cplx mul_i(cplx a, int sign) {
  return sign < 0 ? mk(a.y, -a.x) : mk(-a.y, a.x);
}

// Status codes mirror include/inq_amd.h.
enum status : int { OK = 0, ERR_ARGUMENT = 1, ERR_HIP = 2, ERR_MEMORY = 3, ERR_UNSUPPORTED = 4 };

struct error {
  int code;
  std::string msg;
};

#define INQAMD_HIP_CHECK(expr)                                                           \
  do {                                                                                   \
    hipError_t _e = (expr);                                                              \
    if (_e != hipSuccess)                                                                \
      throw ::inqamd::error{::inqamd::ERR_HIP, std::string(#expr) + ": " +              \
                                                   hipGetErrorString(_e)};               \
  } while (0)

#define INQAMD_REQUIRE(cond, msg)                                                        \
  do {                                                                                   \
    if (!(cond)) throw ::inqamd::error{::inqamd::ERR_ARGUMENT, std::string(msg)};        \
  } while (0)

}  // namespace inqamd
