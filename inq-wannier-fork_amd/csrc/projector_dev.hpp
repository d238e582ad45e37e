// Device bodies of the projector kernels, shared by the one-system kernels
// (projector.hip, projector_mfma.hip) and the uneven-batch work-queue kernels
// (batch.hip).
#pragma once

#include "internal.hpp"

namespace inqamd {

// e^{+i(k+A).r} for sphere point g and e^{-i(k+A).r} for ownership entry g
// (the polar(1, +-dot(k, pos)) of projector_all.hpp:109, :228).
__device__ __forceinline__ void phase_item(const float *pos, const int *uent, int nsph, int nent,
                                           const double *kA, cplx *phase_sph, cplx *phase_ent,
                                           int g) {
  const double k0 = kA[0], k1 = kA[1], k2 = kA[2];
  if (g < nsph) {
    const float *r = pos + (long)g * 3;
    double sn, cs;
    sincos(k0 * (double)r[0] + k1 * (double)r[1] + k2 * (double)r[2], &sn, &cs);
    phase_sph[g] = mk(cs, sn);
  }
  if (g < nent) {
    const float *r = pos + (long)uent[g] * 3;
    double sn, cs;
    sincos(k0 * (double)r[0] + k1 * (double)r[1] + k2 * (double)r[2], &sn, &cs);
    phase_ent[g] = mk(cs, -sn);
  }
}

// Nonlocal projector on the FP64 matrix cores (v_mfma_f64_16x16x4f64).
//
// One workgroup per (sphere p, tile of 8 bands = 16 real columns):
//   gather  X[i][c] = e^{i(k+A).r_pi} psi(r_pi, band), real-doubled
//   GEMM1   P[lm][c] = dV D[lm] sum_i M[lm][i] X[i][c]   (projector_all.hpp:117-165;
//           MFMA rows = lm, k = 4 sphere points, columns = 16)
//   GEMM2   Q[i][c] = sum_lm M[lm][i] P[lm][c]            (:172-212; MFMA rows =
//           16 sphere points, k = 4 lm, columns = 16 -- a full 16x16x4 tile
//           for nlm = 4)
//   write   Q[i] e^{-i(k+A).r_pi} to its ownership-map entry (:225-231 minus
//           the scatter, which the Hpsi epilogue performs in fixed order).
// The PE_WAVES waves of a workgroup take interleaved point blocks; their
// GEMM1 partial tiles are summed in LDS in wave order, so results are
// reproducible.
typedef double v4d __attribute__((ext_vector_type(4)));

// Operand / result lane layout of v_mfma_f64_16x16x4f64 (checked on the
// device by inqamd_selftest_mfma):
//   A (16x4): lane l holds A[l % 16][l / 16]
//   B (4x16): lane l holds B[l / 16][l % 16]
//   D (16x16): lane l, element v holds D[4 v + l / 16][l % 16]
__device__ __forceinline__ int mfma_row_a(int l) { return l & 15; }
__device__ __forceinline__ int mfma_k(int l) { return l >> 4; }
__device__ __forceinline__ int mfma_col_b(int l) { return l & 15; }
__device__ __forceinline__ int mfma_row_d(int l, int v) { return 4 * v + (l >> 4); }
__device__ __forceinline__ int mfma_col_d(int l) { return l & 15; }

constexpr int PE_WAVES = 8;
constexpr int PE_UNROLL = 4;  // gather steps whose loads are issued together
constexpr int PE_NT = 64 * PE_WAVES;
constexpr int PE_BANDS = 8;

// One workgroup (PE_NT threads) on sphere p and bands [b0, b0 + PE_BANDS).
__device__ __forceinline__ void project_expand_block(const project_expand_args &a, int p, int b0) {
  __shared__ double red[PE_WAVES][16][16];
  __shared__ double Pr[16][16];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int K = a.size[p];
  const int nlm = a.max_nlm;
  const long off = (long)p * a.max_sphere;
  const double *M = a.mat + (long)p * nlm * a.max_sphere;
  const int col = mfma_col_b(lane);
  const int kk = mfma_k(lane);
  const int band = b0 + (col >> 1);
  const bool band_ok = band < a.nst;

  // GEMM1: rows = lm, k = 4 points, columns = real-doubled bands; the point
  // indices, phases, psi values and table entries of PE_UNROLL steps are
  // loaded before their MFMAs so the dependent gathers overlap
  v4d acc = {0.0, 0.0, 0.0, 0.0};
  const int lm_a = mfma_row_a(lane);
  constexpr int STEP = PE_WAVES * 4;
  for (int i0 = wave * 4 + kk; i0 < K; i0 += STEP * PE_UNROLL) {
    int idx[PE_UNROLL];
#pragma unroll
    for (int u = 0; u < PE_UNROLL; u++) {
      const int i = i0 + u * STEP;
      idx[u] = i < K ? a.point[off + i] : -1;
    }
    cplx ps[PE_UNROLL], ph[PE_UNROLL];
    double av[PE_UNROLL];
#pragma unroll
    for (int u = 0; u < PE_UNROLL; u++) {
      const int i = i0 + u * STEP;
      ps[u] = (idx[u] >= 0 && band_ok) ? a.psi[(long)idx[u] * a.ld + band] : mk(0.0, 0.0);
      ph[u] = i < K ? a.phase[off + i] : mk(0.0, 0.0);
      av[u] = (i < K && lm_a < nlm) ? M[(long)lm_a * a.max_sphere + i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < PE_UNROLL; u++) {
      const cplx x = cmul(ph[u], ps[u]);
      const double bv = (col & 1) ? x.y : x.x;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u], bv, acc, 0, 0, 0);
    }
  }
#pragma unroll
  for (int v = 0; v < 4; v++) red[wave][mfma_row_d(lane, v)][mfma_col_d(lane)] = acc[v];
  __syncthreads();
  if (threadIdx.x < 256) {
    const int lm = threadIdx.x >> 4, c = threadIdx.x & 15;  // 16 x 16 tile
    double s = red[0][lm][c];
    for (int w = 1; w < PE_WAVES; w++) s += red[w][lm][c];
    // GEMM alpha = dV (projector_all.hpp:131-139), then x coeff (:160-164)
    s = a.dV * s;
    s = lm < nlm ? s * a.coeff[(long)p * nlm + lm] : 0.0;
    Pr[lm][c] = s;
  }
  __syncthreads();
  if (a.P && threadIdx.x < 256) {
    // P[p][lm][b] for callers that need the coefficients (INQ-shaped project)
    const int lm = threadIdx.x >> 4, c = threadIdx.x & 15;
    const int bb = b0 + (c >> 1);
    if (lm < nlm && bb < a.nst && (c & 1) == 0)
      a.P[((long)p * nlm + lm) * a.nst + bb] = mk(Pr[lm][c], Pr[lm][c + 1]);
  }
  if (!a.Qe) return;

  // GEMM2: rows = 16 points, k = 4 lm per step, columns = real-doubled bands
  const int nkb = (nlm + 3) / 4;
  for (int i0 = wave * 16; i0 < K; i0 += PE_WAVES * 16) {
    v4d q = {0.0, 0.0, 0.0, 0.0};
    const int ia = i0 + mfma_row_a(lane);
    for (int kb = 0; kb < nkb; kb++) {
      const int lm = kb * 4 + kk;
      const double av = (ia < K && lm < nlm) ? M[(long)lm * a.max_sphere + ia] : 0.0;
      const double bv = lm < nlm ? Pr[lm][col] : 0.0;
      q = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, q, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int i = i0 + mfma_row_d(lane, v);
      // partner lane holds the other component of the same complex value
      const double mine = q[v];
      const double other = __shfl_xor(mine, 1);
      if (i < K && band_ok) {
        const cplx ph = a.phase[off + i];  // e^{+i(k+A).r}; apply its conjugate
        const double re = (col & 1) ? other : mine;
        const double im = (col & 1) ? mine : other;
        const double out = (col & 1) ? (im * ph.x - re * ph.y) : (re * ph.x + im * ph.y);
        double *dst = reinterpret_cast<double *>(a.Qe + (long)a.ent[off + i] * a.ldQ + band);
        dst[col & 1] = out;
      }
    }
  }
}

}  // namespace inqamd
