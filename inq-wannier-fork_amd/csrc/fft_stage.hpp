// Compile-time Stockham stages shared by the plan-specialised FFT kernels.
//
// A stage works on a LEN x B tile (B independent columns).  Thread w owns
// butterfly j = w / B of column w % B; inputs j + r*LEN/R come through the
// `ld(i, c)` functor, outputs go to (j/NS)*NS*R + j%NS + r*NS through
// `st(i, c, v, q, r)` (q, r: the thread's butterfly slot).  Reading LDS and
// writing LDS in place needs `sync_mid` between the two phases.  Threads at
// or above NT (a workgroup sized for a longer plan) only join the barriers.
#pragma once

#include "fft_dft.hpp"

namespace inqamd {

template <int R0, int R1, int R2>
struct plan3 {
  static constexpr int LEN = R0 * R1 * R2;
  static constexpr int NSTAGE = (R0 > 1) + (R1 > 1) + (R2 > 1);
};

constexpr int plan_threads(int len, int B) {
  int t = (len * B / 8 + 63) / 64 * 64;
  return t < 64 ? 64 : (t > 256 ? 256 : t);
}

// One stage with compile-time radix R and span NS over a LEN x B tile.
template <int R, int NS, int LEN, int B, int NT, class LD, class ST>
__device__ __forceinline__ void pstage(const cplx *__restrict__ tw, int sign, LD ld, ST st,
                                       bool sync_mid) {
  constexpr int NB = LEN / R;
  constexpr int NBF = (NB * B + NT - 1) / NT;
  constexpr int TSTRIDE = LEN / (NS * R);
  cplx v[NBF][R];
#pragma unroll
  for (int q = 0; q < NBF; q++) {
    const int w = threadIdx.x + q * NT;
    if ((int)threadIdx.x < NT && (NBF * NT == NB * B || w < NB * B)) {
      const int j = w / B, bb = w % B, k = j % NS;
#pragma unroll
      for (int r = 0; r < R; r++) v[q][r] = ld(j + r * NB, bb);
      if (NS > 1 && k != 0) {
#pragma unroll
        for (int r = 1; r < R; r++) v[q][r] = cmul(v[q][r], twiddle(tw, r * k * TSTRIDE, sign));
      }
      dft<R>(v[q], sign);
    }
  }
  if (sync_mid) __syncthreads();
#pragma unroll
  for (int q = 0; q < NBF; q++) {
    const int w = threadIdx.x + q * NT;
    if ((int)threadIdx.x < NT && (NBF * NT == NB * B || w < NB * B)) {
      const int j = w / B, bb = w % B, k = j % NS;
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; r++) st(base + r * NS, bb, v[q][r], q, r);
    }
  }
}

// Runs the stages of plan P (R0, R1, R2; radix 1 = absent).  The first
// stage reads with `first_ld`, the last writes with `last_st`; stages in
// between go through LDS in place.
template <class P, int R0, int R1, int R2, int B, int NT, class FL, class LL, class LS, class LST>
__device__ __forceinline__ void run_plan(const cplx *tw, int sign, bool first_from_lds, FL first_ld,
                                         LL ll, LS ls, LST last_st, bool last_to_lds) {
  constexpr int LEN = P::LEN;
  constexpr int NSTAGE = P::NSTAGE;
  // stage 0
  if constexpr (NSTAGE == 1) {
    pstage<R0, 1, LEN, B, NT>(tw, sign, first_ld, last_st, first_from_lds && last_to_lds);
  } else {
    pstage<R0, 1, LEN, B, NT>(tw, sign, first_ld, ls, first_from_lds);
    __syncthreads();
    if constexpr (NSTAGE == 2) {
      pstage<R1, R0, LEN, B, NT>(tw, sign, ll, last_st, last_to_lds);
    } else {
      pstage<R1, R0, LEN, B, NT>(tw, sign, ll, ls, true);
      __syncthreads();
      pstage<R2, R0 * R1, LEN, B, NT>(tw, sign, ll, last_st, last_to_lds);
    }
  }
  if (last_to_lds) __syncthreads();
}

}  // namespace inqamd
