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

// cos and sin of 2 pi m / N for the composite butterflies (N = 6, 9, 12),
// the correctly rounded doubles
__host__ __device__ constexpr double unit_cos(int N, int m) {
  constexpr double c30 = 0.86602540378443860;  // cos(pi/6)
  constexpr double c12[12] = {1.0, c30, 0.5, 0.0, -0.5, -c30, -1.0, -c30, -0.5, 0.0, 0.5, c30};
  constexpr double c9[9] = {1.0, 0.76604444311897801, 0.17364817766693036, -0.5,
                            -0.93969262078590843, -0.93969262078590843, -0.5,
                            0.17364817766693036, 0.76604444311897801};
  return N == 9 ? c9[m % 9] : c12[(m * (12 / N)) % 12];
}

__host__ __device__ constexpr double unit_sin(int N, int m) {
  constexpr double c30 = 0.86602540378443860;  // sin(pi/3)
  constexpr double s12[12] = {0.0, 0.5, c30, 1.0, c30, 0.5, 0.0, -0.5, -c30, -1.0, -c30, -0.5};
  constexpr double s9[9] = {0.0, 0.64278760968653936, 0.98480775301220802, c30,
                            0.34202014332566871, -0.34202014332566871, -c30,
                            -0.98480775301220802, -0.64278760968653936};
  return N == 9 ? s9[m % 9] : s12[(m * (12 / N)) % 12];
}

// Radix P*Q butterfly on registers by one Cooley-Tukey split: inputs
// n = Q n1 + n2, radix-P transforms over n1, twiddles W_N^(n2 k1), radix-Q
// transforms over n2, outputs k = k1 + P k2.
template <int P, int Q>
__device__ __forceinline__ void dft_split(cplx *v, int sign) {
  constexpr int N = P * Q;
  cplx a[Q][P];
#pragma unroll
  for (int n2 = 0; n2 < Q; n2++) {
#pragma unroll
    for (int n1 = 0; n1 < P; n1++) a[n2][n1] = v[Q * n1 + n2];
    dft<P>(a[n2], sign);
  }
#pragma unroll
  for (int n2 = 1; n2 < Q; n2++) {
#pragma unroll
    for (int k1 = 1; k1 < P; k1++) {
      const int m = (n2 * k1) % N;
      if (4 * m == N) {
        a[n2][k1] = mul_i(a[n2][k1], sign);
      } else if (2 * m == N) {
        a[n2][k1] = mk(-a[n2][k1].x, -a[n2][k1].y);
      } else if (4 * m == 3 * N) {
        a[n2][k1] = mul_i(a[n2][k1], -sign);
      } else {
        a[n2][k1] = cmul(a[n2][k1], mk(unit_cos(N, m), sign * unit_sin(N, m)));
      }
    }
  }
#pragma unroll
  for (int k1 = 0; k1 < P; k1++) {
    cplx t[Q];
#pragma unroll
    for (int n2 = 0; n2 < Q; n2++) t[n2] = a[n2][k1];
    dft<Q>(t, sign);
#pragma unroll
    for (int k2 = 0; k2 < Q; k2++) v[k1 + P * k2] = t[k2];
  }
}

template <>
__device__ __forceinline__ void dft<6>(cplx *v, int sign) {
  dft_split<3, 2>(v, sign);
}

template <>
__device__ __forceinline__ void dft<9>(cplx *v, int sign) {
  dft_split<3, 3>(v, sign);
}

template <>
__device__ __forceinline__ void dft<12>(cplx *v, int sign) {
  dft_split<4, 3>(v, sign);
}

__device__ __forceinline__ cplx twiddle(const cplx *__restrict__ tw, int idx, int sign) {
  cplx t = tw[idx];
  if (sign > 0) t.y = -t.y;
  return t;
}


__device__ __forceinline__ int symmetric_index(int i, int n) { return i >= (n + 1) / 2 ? i - n : i; }

}  // namespace inqamd
