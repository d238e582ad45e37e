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

// GEMM1 over the sphere points [ib, ie) of sphere p for bands
// [b0, b0 + PE_BANDS): rows = lm, k = 4 points, columns = real-doubled
// bands.  The point indices, phases, psi values and table entries of
// PE_UNROLL steps are loaded before their MFMAs so the dependent gathers
// overlap.  The waves' tiles are summed in wave order into out[16][16]
// (LDS or global), unscaled.
__device__ __forceinline__ void gemm1_points(const project_expand_args &a, int p, int b0, int ib,
                                             int ie, double (*red)[16][16], double *out) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nlm = a.max_nlm;
  const long off = (long)p * a.max_sphere;
  const double *M = a.mat + (long)p * nlm * a.max_sphere;
  const int col = mfma_col_b(lane);
  const int kk = mfma_k(lane);
  const int band = b0 + (col >> 1);
  const bool band_ok = band < a.nst;
  v4d acc = {0.0, 0.0, 0.0, 0.0};
  const int lm_a = mfma_row_a(lane);
  constexpr int STEP = PE_WAVES * 4;
  // the loop variable is uniform across the wave (every lane runs every
  // MFMA); the lane's point i = j0 + kk + u STEP is guarded instead
  for (int j0 = ib + wave * 4; j0 < ie; j0 += STEP * PE_UNROLL) {
    int idx[PE_UNROLL];
#pragma unroll
    for (int u = 0; u < PE_UNROLL; u++) {
      const int i = j0 + kk + u * STEP;
      idx[u] = i < ie ? a.point[off + i] : -1;
    }
    cplx ps[PE_UNROLL], ph[PE_UNROLL];
    double av[PE_UNROLL];
#pragma unroll
    for (int u = 0; u < PE_UNROLL; u++) {
      const int i = j0 + kk + u * STEP;
      ps[u] = (idx[u] >= 0 && band_ok) ? a.psi[(long)idx[u] * a.ld + band] : mk(0.0, 0.0);
      ph[u] = i < ie ? a.phase[off + i] : mk(0.0, 0.0);
      av[u] = (i < ie && lm_a < nlm) ? M[(long)lm_a * a.max_sphere + i] : 0.0;
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
    out[lm * 16 + c] = s;
  }
}

// GEMM2 over the sphere points [ib, ie): Q[i][c] = sum_lm M[lm][i] Pr[lm][c],
// written as Q e^{-i(k+A).r} to the point's ownership-map entry.
__device__ __forceinline__ void gemm2_points(const project_expand_args &a, int p, int b0, int ib,
                                             int ie, const double (*Pr)[16]) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nlm = a.max_nlm;
  const long off = (long)p * a.max_sphere;
  const double *M = a.mat + (long)p * nlm * a.max_sphere;
  const int col = mfma_col_b(lane);
  const int kk = mfma_k(lane);
  const int band = b0 + (col >> 1);
  const bool band_ok = band < a.nst;
  const int nkb = (nlm + 3) / 4;
  for (int i0 = ib + wave * 16; i0 < ie; i0 += PE_WAVES * 16) {
    v4d q = {0.0, 0.0, 0.0, 0.0};
    const int ia = i0 + mfma_row_a(lane);
    for (int kb = 0; kb < nkb; kb++) {
      const int lm = kb * 4 + kk;
      const double av = (ia < ie && lm < nlm) ? M[(long)lm * a.max_sphere + ia] : 0.0;
      const double bv = lm < nlm ? Pr[lm][col] : 0.0;
      q = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, q, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int i = i0 + mfma_row_d(lane, v);
      // partner lane holds the other component of the same complex value
      const double mine = q[v];
      const double other = __shfl_xor(mine, 1);
      if (i < ie && band_ok) {
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

// P = dV D (sum of the partial tiles): GEMM alpha = dV (projector_all.hpp:131-139),
// then x coeff (:160-164); and the INQ-shaped coefficients if requested.
__device__ __forceinline__ void scale_coefficients(const project_expand_args &a, int p, int b0,
                                                   double (*Pr)[16]) {
  const int nlm = a.max_nlm;
  if (threadIdx.x < 256) {
    const int lm = threadIdx.x >> 4, c = threadIdx.x & 15;
    double s = a.dV * Pr[lm][c];
    Pr[lm][c] = lm < nlm ? s * a.coeff[(long)p * nlm + lm] : 0.0;
  }
  __syncthreads();
  if (a.P && threadIdx.x < 256) {
    // P[p][lm][b] for callers that need the coefficients (INQ-shaped project)
    const int lm = threadIdx.x >> 4, c = threadIdx.x & 15;
    const int bb = b0 + (c >> 1);
    if (lm < nlm && bb < a.nst && (c & 1) == 0)
      a.P[((long)p * nlm + lm) * a.nst + bb] = mk(Pr[lm][c], Pr[lm][c + 1]);
  }
}

// One workgroup (PE_NT threads) on sphere p and bands [b0, b0 + PE_BANDS).
__device__ __forceinline__ void project_expand_block(const project_expand_args &a, int p, int b0) {
  __shared__ double red[PE_WAVES][16][16];
  __shared__ double Pr[16][16];
  const int K = a.size[p];
  gemm1_points(a, p, b0, 0, K, red, &Pr[0][0]);
  __syncthreads();
  scale_coefficients(a, p, b0, Pr);
  if (!a.Qe) return;
  __syncthreads();
  gemm2_points(a, p, b0, 0, K, Pr);
}

// Point-chunked form for few (sphere, band tile) pairs: chunk c of nc covers
// points [K c / nc, K (c + 1) / nc).  Pass 1 writes each chunk's GEMM1 tile;
// pass 2 sums the chunk tiles of its (sphere, tile) in chunk order, scales
// them and expands its own chunk.
__device__ __forceinline__ int chunk_begin(int K, int c, int nc) {
  return (int)((long)K * c / nc);
}

__device__ __forceinline__ void project_chunk_block(const project_expand_args &a, int p, int b0,
                                                    int c, int nc, double *part) {
  __shared__ double red[PE_WAVES][16][16];
  const int K = a.size[p];
  gemm1_points(a, p, b0, chunk_begin(K, c, nc), chunk_begin(K, c + 1, nc), red, part);
}

__device__ __forceinline__ void expand_chunk_block(const project_expand_args &a, int p, int b0,
                                                   int c, int nc, const double *parts) {
  __shared__ double Pr[16][16];
  const int K = a.size[p];
  if (threadIdx.x < 256) {
    const int e = threadIdx.x;
    double s = parts[e];
    for (int k = 1; k < nc; k++) s += parts[k * 256 + e];
    Pr[e >> 4][e & 15] = s;
  }
  __syncthreads();
  scale_coefficients(a, p, b0, Pr);
  if (!a.Qe) return;
  __syncthreads();
  gemm2_points(a, p, b0, chunk_begin(K, c, nc), chunk_begin(K, c + 1, nc), Pr);
}

// ---------------------------------------------------------------------------
// GEMM1 alone on v_mfma_f64_4x4x4f64: P = dV D M^T (phase psi) for 16-band
// tiles, no padding rows (projector_all.hpp:117-165)
// ---------------------------------------------------------------------------
// v_mfma_f64_4x4x4f64 runs four independent 4x4x4 blocks s = (l >> 2) & 3;
// lane l holds (checked on the device by inqamd_selftest_mfma_4x4)
//   A_s[i][k] with i = l & 3,  k = l >> 4
//   B_s[k][j] with j = l & 3,  k = l >> 4
//   D_s[i][j] with i = l >> 4, j = l & 3
// With A_s = M[lm][4 points] in every block and B_s = X[4 points][bands
// 4s..4s+3], one instruction is 4 points x 16 bands x 4 lm, and lane l
// loads point l >> 4, band l & 15 and M[l & 3][point l >> 4], and receives
// P[lm = l >> 4][band l & 15]: real and imaginary parts are two
// instructions on the same A.
constexpr int P4_WAVES = 8;
constexpr int P4_NT = 64 * P4_WAVES;
constexpr int P4_BANDS = 16;  // bands of one instruction pair
constexpr int P4_UNROLL = 4;

__device__ __forceinline__ double mfma4(double a, double b, double c) {
  return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 0);
}

// One wave's sums over the points j0 + kk + u STEP (j0 = j, j + STEP U, ...
// below ie; kk = lane >> 4 picks one of 4 points of a step) of sphere p for
// bands [b0, b0 + 16 NB) into accr/acci (lane order: lm = l >> 4, band
// b0 + 16 q + (l & 15)).  Each point's M row and phase serve NB band groups.
template <int NB, int STEP>
__device__ __forceinline__ void p4_wave_points(const project_expand_args &a, int p, int b0,
                                               int j, int ie, double *accr, double *acci) {
  // point steps in flight: 8 for one band group (the small per-chunk
  // launches, which have few workgroups), 4 for more (registers); a wave
  // adds its points in increasing order either way
  constexpr int U = NB == 1 ? 8 : P4_UNROLL;
  const int lane = threadIdx.x & 63;
  const int nlm = a.max_nlm;
  const long off = (long)p * a.max_sphere;
  const double *M = a.mat + (long)p * nlm * a.max_sphere;
  const int kk = lane >> 4;
  const int lm_a = lane & 3;
  // the lane's chunk bands b0 + (lane & 15) + 16 q and their offsets in psi
  // (band 0 of the row for a band past the chunk)
  int gband[NB];
  bool bok[NB];
#pragma unroll
  for (int q = 0; q < NB; q++) {
    const int cb = b0 + (lane & 15) + 16 * q;
    bok[q] = cb < a.nbc;
    gband[q] = bok[q] ? cb + (cb >> (3 + a.glog)) * a.gskip : 0;
  }
  const long mrow = (long)(lm_a < nlm ? lm_a : 0) * a.max_sphere;
#pragma unroll
  for (int q = 0; q < NB; q++) accr[q] = acci[q] = 0.0;
  // Every load reads an address in bounds (a lane past the range reads the
  // sphere's point 0, a grid index whenever the range is non-empty; table
  // padding is -1, but for an empty range no psi is gathered) and its value
  // is selected after it arrives, so no load sits in a branch of its own.  The point indices of a step are loaded one
  // step ahead: the gathers of psi wait for no index load, and the MFMAs
  // wait for their own operands only.
  int idx[U];
#pragma unroll
  for (int u = 0; u < U; u++) {
    const int i = j + kk + u * STEP;
    idx[u] = a.point[off + (i < ie ? i : 0)];
  }
  for (int j0 = j; j0 < ie; j0 += STEP * U) {
    cplx ps[U][NB], ph[U];
    double av[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = j0 + kk + u * STEP;
      const int ic = i < ie ? i : 0;
#pragma unroll
      for (int q = 0; q < NB; q++) ps[u][q] = a.psi[(long)idx[u] * a.ld + gband[q]];
      ph[u] = a.phase[off + ic];
      av[u] = M[mrow + ic];
    }
    // the next indices are issued after this step's operands (the scheduler
    // would otherwise mix them, and the MFMAs would wait for every load)
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = j0 + STEP * U + kk + u * STEP;
      idx[u] = a.point[off + (i < ie ? i : 0)];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < U; u++) {
      const bool in = j0 + kk + u * STEP < ie;
      const cplx f = in ? ph[u] : mk(0.0, 0.0);
      const double m = (in && lm_a < nlm) ? av[u] : 0.0;
#pragma unroll
      for (int q = 0; q < NB; q++) {
        const cplx x = cmul(f, (in && bok[q]) ? ps[u][q] : mk(0.0, 0.0));
        accr[q] = mfma4(m, x.x, accr[q]);
        acci[q] = mfma4(m, x.y, acci[q]);
      }
    }
  }
}

// Sums over the sphere points [ib, ie) of sphere p for bands
// [b0, b0 + 16 NB): wave w takes the steps of 4 points at ib + 4 w,
// ib + 4 w + 4 P4_WAVES, ...; the waves' accumulators are added in wave
// order into out[NB][64][2] (lane order of p4_wave_points; real,
// imaginary), unscaled.
template <int NB>
__device__ __forceinline__ void gemm1_p4_points(const project_expand_args &a, int p, int b0,
                                                int ib, int ie, double (*red)[NB][64][2],
                                                double *out) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  double accr[NB], acci[NB];
  p4_wave_points<NB, P4_WAVES * 4>(a, p, b0, ib + wave * 4, ie, accr, acci);
#pragma unroll
  for (int q = 0; q < NB; q++) {
    red[wave][q][lane][0] = accr[q];
    red[wave][q][lane][1] = acci[q];
  }
  __syncthreads();
  for (int t = threadIdx.x; t < NB * 128; t += P4_NT) {
    const int q = t >> 7, l = (t >> 1) & 63, c = t & 1;
    double s = red[0][q][l][c];
    for (int w = 1; w < P4_WAVES; w++) s += red[w][q][l][c];
    out[t] = s;
  }
}

// P[p][lm][b] = dV D[p][lm] x the summed tile (projector_all.hpp:131-139, 160-164)
template <int NB>
__device__ __forceinline__ void p4_store(const project_expand_args &a, int p, int b0,
                                         const double *sum) {
  for (int t = threadIdx.x; t < NB * 64; t += blockDim.x) {
    const int q = t >> 6, l = t & 63;
    const int lm = l >> 4, cb = b0 + 16 * q + (l & 15);
    if (lm < a.max_nlm && cb < a.nbc) {
      const long b = a.pband0 + cb + (cb >> (3 + a.glog)) * a.gskip;
      const double d = a.dV * a.coeff[(long)p * a.max_nlm + lm];
      a.P[((long)p * a.max_nlm + lm) * a.nst + b] = mk(d * sum[2 * t], d * sum[2 * t + 1]);
    }
  }
}

}  // namespace inqamd
