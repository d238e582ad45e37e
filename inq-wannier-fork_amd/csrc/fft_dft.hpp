// Radix-R DFT butterflies on registers and twiddle access shared by the
// pencil-pass FFT kernels.  sign -1 = exp(-2 pi i/R) (forward), +1 backward.
#pragma once

#include "common.hpp"

namespace inqamd {

template <int R>
__device__ __forceinline__ void dft(cplx *v, int sign);

template <>
__device__ __forceinline__ void dft<2>(cplx *v, int) {
  cplx a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <>
__device__ __forceinline__ void dft<3>(cplx *v, int sign) {
  const double c = 0.86602540378443864676372317075294;  // sin(2 pi/3)
  cplx t1 = cadd(v[1], v[2]);
  cplx t2 = csub(v[1], v[2]);
  cplx y0 = cadd(v[0], t1);
  cplx m = mk(v[0].x - 0.5 * t1.x, v[0].y - 0.5 * t1.y);
  cplx s = mul_i(cscale(c, t2), sign);  // sign*i*c*t2
  v[0] = y0;
  v[1] = cadd(m, s);
  v[2] = csub(m, s);
}

template <>
__device__ __forceinline__ void dft<4>(cplx *v, int sign) {
  cplx t0 = cadd(v[0], v[2]);
  cplx t1 = csub(v[0], v[2]);
  cplx t2 = cadd(v[1], v[3]);
  cplx t3 = mul_i(csub(v[1], v[3]), sign);
  v[0] = cadd(t0, t2);
  v[2] = csub(t0, t2);
  v[1] = cadd(t1, t3);
  v[3] = csub(t1, t3);
}

template <>
__device__ __forceinline__ void dft<5>(cplx *v, int sign) {
  const double c1 = 0.30901699437494742410229341718282;   // cos(2 pi/5)
  const double c2 = -0.80901699437494742410229341718282;  // cos(4 pi/5)
  const double s1 = 0.95105651629515357211643933337938;   // sin(2 pi/5)
  const double s2 = 0.58778525229247312916870595463907;   // sin(4 pi/5)
  cplx t1 = cadd(v[1], v[4]);
  cplx t2 = cadd(v[2], v[3]);
  cplx t3 = csub(v[1], v[4]);
  cplx t4 = csub(v[2], v[3]);
  cplx y0 = cadd(v[0], cadd(t1, t2));
  cplx b1 = mk(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
  cplx b2 = mk(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
  cplx d1 = mul_i(mk(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y), sign);
  cplx d2 = mul_i(mk(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y), sign);
  v[0] = y0;
  v[1] = cadd(b1, d1);
  v[4] = csub(b1, d1);
  v[2] = cadd(b2, d2);
  v[3] = csub(b2, d2);
}

template <>
__device__ __forceinline__ void dft<8>(cplx *v, int sign) {
  const double h = 0.70710678118654752440084436210485;  // 1/sqrt(2)
  cplx u[4], w[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    u[k] = cadd(v[k], v[k + 4]);
    w[k] = csub(v[k], v[k + 4]);
  }
  // w[k] *= W8^k ; W8 = (1 + sign i)/sqrt(2)
  w[1] = cscale(h, cadd(w[1], mul_i(w[1], sign)));
  w[2] = mul_i(w[2], sign);
  w[3] = cscale(h, csub(mul_i(w[3], sign), w[3]));
  dft<4>(u, sign);
  dft<4>(w, sign);
#pragma unroll
  for (int k = 0; k < 4; k++) {
    v[2 * k] = u[k];
    v[2 * k + 1] = w[k];
  }
}

__device__ __forceinline__ cplx twiddle(const cplx *__restrict__ tw, int idx, int sign) {
  cplx t = tw[idx];
  if (sign > 0) t.y = -t.y;
  return t;
}


__device__ __forceinline__ int symmetric_index(int i, int n) { return i >= (n + 1) / 2 ? i - n : i; }

}  // namespace inqamd
