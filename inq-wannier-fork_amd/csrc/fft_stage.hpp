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

// Inputs of one stage held in registers: the R inputs of each of the
// thread's NBF butterflies.
template <int R, int LEN, int B, int NT>
struct stage_regs {
  static constexpr int NBF = (LEN / R * B + NT - 1) / NT;
  cplx v[NBF][R];
};

// Loads a stage's inputs (butterfly j = w / B of column w % B takes inputs
// j + r*LEN/R) through `ld(i, c)`.
template <int R, int LEN, int B, int NT, class LD>
__device__ __forceinline__ void pload(stage_regs<R, LEN, B, NT> &in, LD ld) {
  constexpr int NB = LEN / R;
  constexpr int NBF = stage_regs<R, LEN, B, NT>::NBF;
#pragma unroll
  for (int q = 0; q < NBF; q++) {
    const int w = threadIdx.x + q * NT;
    if ((int)threadIdx.x < NT && (NBF * NT == NB * B || w < NB * B)) {
      const int j = w / B, bb = w % B;
#pragma unroll
      for (int r = 0; r < R; r++) in.v[q][r] = ld(j + r * NB, bb);
    }
  }
}

// Twiddles of a stage's butterflies (slot [q][r - 1] for input r >= 1),
// loaded together ahead of their use: one memory round trip per stage
// instead of one per twiddle.
template <int R, int LEN, int B, int NT>
struct stage_tw {
  static constexpr int NBF = stage_regs<R, LEN, B, NT>::NBF;
  cplx t[NBF][R > 1 ? R - 1 : 1];
};

template <int R, int NS, int LEN, int B, int NT>
__device__ __forceinline__ void ptw(stage_tw<R, LEN, B, NT> &t, const cplx *__restrict__ tw,
                                    int sign) {
  if constexpr (NS > 1) {
    constexpr int NBF = stage_regs<R, LEN, B, NT>::NBF;
    constexpr int TSTRIDE = LEN / (NS * R);
#pragma unroll
    for (int q = 0; q < NBF; q++) {
      const int w = threadIdx.x + q * NT;
      const int k = (w / B) % NS;
#pragma unroll
      for (int r = 1; r < R; r++) t.t[q][r - 1] = twiddle(tw, r * k * TSTRIDE, sign);
    }
  }
}

// Twiddles and butterflies of one stage with compile-time radix R and span
// NS over a LEN x B tile, in registers, with the stage's twiddles in t:
// slot (q, r) ends up holding output (j/NS)*NS*R + j%NS + r*NS of column w % B.
template <int R, int NS, int LEN, int B, int NT>
__device__ __forceinline__ void pcompute_t(stage_regs<R, LEN, B, NT> &in,
                                           const stage_tw<R, LEN, B, NT> &t, int sign) {
  constexpr int NB = LEN / R;
  constexpr int NBF = stage_regs<R, LEN, B, NT>::NBF;
#pragma unroll
  for (int q = 0; q < NBF; q++) {
    const int w = threadIdx.x + q * NT;
    if ((int)threadIdx.x < NT && (NBF * NT == NB * B || w < NB * B)) {
      const int j = w / B, k = j % NS;
      if (NS > 1 && k != 0) {
#pragma unroll
        for (int r = 1; r < R; r++) in.v[q][r] = cmul(in.v[q][r], t.t[q][r - 1]);
      }
      dft<R>(in.v[q], sign);
    }
  }
}

#ifndef INQAMD_TW_HOIST
#define INQAMD_TW_HOIST 1
#endif
// Stages whose twiddles are loaded ahead: at most 7 per thread (28
// registers), so that tiles near their register budget do not spill
template <int R, int LEN, int B, int NT, bool TWH>
constexpr bool tw_hoist() {
  return INQAMD_TW_HOIST && TWH && stage_regs<R, LEN, B, NT>::NBF * (R - 1) <= 7;
}

// The same with the twiddles loaded from the table tw: all of the stage's
// ahead of the products (tw_hoist; TWH false: a caller holding many other
// registers through the transform), or each at its product
template <int R, int NS, int LEN, int B, int NT, bool TWH = true>
__device__ __forceinline__ void pcompute(stage_regs<R, LEN, B, NT> &in, const cplx *__restrict__ tw,
                                         int sign) {
  if constexpr (tw_hoist<R, LEN, B, NT, TWH>()) {
    stage_tw<R, LEN, B, NT> t;
    ptw<R, NS, LEN, B, NT>(t, tw, sign);
    pcompute_t<R, NS, LEN, B, NT>(in, t, sign);
  } else {
    constexpr int NB = LEN / R;
    constexpr int NBF = stage_regs<R, LEN, B, NT>::NBF;
    constexpr int TSTRIDE = LEN / (NS * R);
#pragma unroll
    for (int q = 0; q < NBF; q++) {
      const int w = threadIdx.x + q * NT;
      if ((int)threadIdx.x < NT && (NBF * NT == NB * B || w < NB * B)) {
        const int j = w / B, k = j % NS;
        if (NS > 1 && k != 0) {
#pragma unroll
          for (int r = 1; r < R; r++)
            in.v[q][r] = cmul(in.v[q][r], twiddle(tw, r * k * TSTRIDE, sign));
        }
        dft<R>(in.v[q], sign);
      }
    }
  }
}

// Stores of one stage's outputs from registers
template <int R, int NS, int LEN, int B, int NT, class ST>
__device__ __forceinline__ void pstore(stage_regs<R, LEN, B, NT> &in, ST st) {
  constexpr int NB = LEN / R;
  constexpr int NBF = stage_regs<R, LEN, B, NT>::NBF;
#pragma unroll
  for (int q = 0; q < NBF; q++) {
    const int w = threadIdx.x + q * NT;
    if ((int)threadIdx.x < NT && (NBF * NT == NB * B || w < NB * B)) {
      const int j = w / B, bb = w % B, k = j % NS;
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; r++) st(base + r * NS, bb, in.v[q][r], q, r);
    }
  }
}

// Twiddles, butterflies and stores of one stage with compile-time radix R
// and span NS over a LEN x B tile, from inputs already in registers.
template <int R, int NS, int LEN, int B, int NT, bool TWH = true, class ST>
__device__ __forceinline__ void pfinish(stage_regs<R, LEN, B, NT> &in, const cplx *__restrict__ tw,
                                        int sign, ST st, bool sync_mid) {
  pcompute<R, NS, LEN, B, NT, TWH>(in, tw, sign);
  if (sync_mid) __syncthreads();
  pstore<R, NS, LEN, B, NT>(in, st);
}

// One stage: inputs through `ld(i, c)`, outputs to (j/NS)*NS*R + j%NS + r*NS
// through `st(i, c, v, q, r)`.  The twiddle loads go out before the inputs'.
template <int R, int NS, int LEN, int B, int NT, bool TWH = true, class LD, class ST>
__device__ __forceinline__ void pstage(const cplx *__restrict__ tw, int sign, LD ld, ST st,
                                       bool sync_mid) {
  stage_regs<R, LEN, B, NT> in;
  if constexpr (tw_hoist<R, LEN, B, NT, TWH>()) {
    stage_tw<R, LEN, B, NT> t;
    ptw<R, NS, LEN, B, NT>(t, tw, sign);
    pload<R, LEN, B, NT>(in, ld);
    pcompute_t<R, NS, LEN, B, NT>(in, t, sign);
  } else {
    pload<R, LEN, B, NT>(in, ld);
    pcompute<R, NS, LEN, B, NT, false>(in, tw, sign);
  }
  if (sync_mid) __syncthreads();
  pstore<R, NS, LEN, B, NT>(in, st);
}

// Hands a stage's outputs (registers `a`, radix RA, span NSA) to the next
// stage's inputs (registers `b`, radix RB) through an LDS array of doubles
// that holds one component at a time: real parts, then imaginary parts.
// Half the LDS of a complex exchange for twice the barriers, so that twice
// as many workgroups share a CU.  `sa(i, c)` is the slot of element (i, c)
// as stored, `sb(i, c)` as loaded.  Starts with a barrier (earlier readers
// of the array are done); ends after the loads, without one.
template <int RA, int NSA, int RB, int LEN, int B, int NT, class SA, class SB>
__device__ __forceinline__ void xchg_split(stage_regs<RA, LEN, B, NT> &a, stage_regs<RB, LEN, B, NT> &b,
                                           double *lds, SA sa, SB sb) {
  constexpr int NBA = LEN / RA, NFA = stage_regs<RA, LEN, B, NT>::NBF;
  constexpr int NBB = LEN / RB, NFB = stage_regs<RB, LEN, B, NT>::NBF;
#pragma unroll
  for (int part = 0; part < 2; part++) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NFA; q++) {
      const int w = threadIdx.x + q * NT;
      if ((int)threadIdx.x < NT && (NFA * NT == NBA * B || w < NBA * B)) {
        const int j = w / B, bb = w % B;
        const int base = (j / NSA) * NSA * RA + j % NSA;
#pragma unroll
        for (int r = 0; r < RA; r++) lds[sa(base + r * NSA, bb)] = part ? a.v[q][r].y : a.v[q][r].x;
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NFB; q++) {
      const int w = threadIdx.x + q * NT;
      if ((int)threadIdx.x < NT && (NFB * NT == NBB * B || w < NBB * B)) {
        const int j = w / B, bb = w % B;
#pragma unroll
        for (int r = 0; r < RB; r++) {
          const double v = lds[sb(j + r * NBB, bb)];
          if (part)
            b.v[q][r].y = v;
          else
            b.v[q][r].x = v;
        }
      }
    }
  }
}

// Runs the stages of plan P (R0, R1, R2; radix 1 = absent) from the first
// stage's inputs already in registers (pload with the plan's R0); the last
// stage writes with `last_st`, stages in between go through LDS in place.
template <class P, int R0, int R1, int R2, int B, int NT, bool TWH = true, class LL, class LS,
          class LST>
__device__ __forceinline__ void run_plan_from(stage_regs<R0, P::LEN, B, NT> &in, const cplx *tw,
                                              int sign, bool first_from_lds, LL ll, LS ls,
                                              LST last_st, bool last_to_lds) {
  constexpr int LEN = P::LEN;
  constexpr int NSTAGE = P::NSTAGE;
  if constexpr (NSTAGE == 1) {
    pfinish<R0, 1, LEN, B, NT, TWH>(in, tw, sign, last_st, first_from_lds && last_to_lds);
  } else {
    pfinish<R0, 1, LEN, B, NT, TWH>(in, tw, sign, ls, first_from_lds);
    __syncthreads();
    if constexpr (NSTAGE == 2) {
      pstage<R1, R0, LEN, B, NT, TWH>(tw, sign, ll, last_st, last_to_lds);
    } else {
      pstage<R1, R0, LEN, B, NT, TWH>(tw, sign, ll, ls, true);
      __syncthreads();
      pstage<R2, R0 * R1, LEN, B, NT, TWH>(tw, sign, ll, last_st, last_to_lds);
    }
  }
  if (last_to_lds) __syncthreads();
}

// Runs the stages of plan P; the first stage reads with `first_ld`.
template <class P, int R0, int R1, int R2, int B, int NT, bool TWH = true, class FL, class LL,
          class LS, class LST>
__device__ __forceinline__ void run_plan(const cplx *tw, int sign, bool first_from_lds, FL first_ld,
                                         LL ll, LS ls, LST last_st, bool last_to_lds) {
  stage_regs<R0, P::LEN, B, NT> in;
  pload<R0, P::LEN, B, NT>(in, first_ld);
  run_plan_from<P, R0, R1, R2, B, NT, TWH>(in, tw, sign, first_from_lds, ll, ls, last_st,
                                           last_to_lds);
}

}  // namespace inqamd
