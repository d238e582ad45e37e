// Current observable and the discrete adjoint of the frozen-potential ETRS
// propagation under a Fourier-series vector potential.
//
// Notation (plain sums over grid points and bands, dV = V/N):
//   <a|b> = sum conj(a) b.
//   For A = A0 + s d (d Cartesian), W = dH/ds is
//     (G + k + A).d            in G space (the kinetic 1/2|G+k+A|^2), plus
//     dV_nl/ds:  (dV_nl/ds f)(x_i) = i dV e^{-i(k+A).x_i} sum_lm M_i D (r_lm - w_i p_lm)
//   with w_i = x_i.d, p_lm = sum_j M_j e^{i(k+A).x_j} f_j and r_lm the same
//   sum weighted by w_j (projector_all.hpp:117-165 with an r weight,
//   :319-403).  J.d = -occ dV <psi|W|psi> is observables::current
//   (current.hpp:25-53) along d.
//   The adjoint state lambda satisfies dL = Re <lambda|dpsi>.  For one Taylor
//   half step psi' = T(z, H) phi, T = sum_{m<=M} (zH)^m/m! (exponential.hpp:21-34):
//     lambda_in = T(conj z, H) lambda'
//     dL/dV(r) += Re sum_b sum_j conj(X_j) Y_j,   dL/ds += Re sum_j <X_j|W|Y_j>
//   with Y_j = H^j phi and X_j = sum_{k=0}^{M-1-j} conj(z)^{k+j+1}/(k+j+1)! H^k lambda'.
//   The kinetic part of <X|W|Y> is (1/N) sum_G w(G) conj(X^(G)) Y^(G) with
//   X^ the unnormalised forward transform.

#include "internal.hpp"
#include "inq_amd.h"

#include <cmath>
#include <complex>
#include <vector>

namespace inqamd {

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// Moments of fields on the projector spheres:
//   out[f][p][lm][c][b] = sum_i M[p][lm][i] w_c(i) e^{i(k+A).x_i} f(x_i, b)
// w_0 = 1, w_c = sum_e pos_e L[c-1][e] (c = 1..3), or w_2 = w_1^2 if square2.
// A workgroup takes 64 bands x one point chunk of one sphere and field; with
// several chunks the partial sums go to `partial` and a second kernel adds
// them in chunk order.
constexpr int MOM_MAXF = 16;
struct moments_args {
  const cplx *f[MOM_MAXF];
  int nf;
  long ld;
  int nst;
  int nproj, max_sphere, max_nlm;
  const int *point, *size;
  const double *mat;
  const float *pos;
  const cplx *phase;  // e^{+i(k+A).x} per sphere point
  double L[9];
  int square2;
  int pch;            // point chunks per sphere
  cplx *out;          // field stride out_stride
  long out_stride;
  cplx *partial;      // [nf][pch][nproj][nlm][4][nst] when pch > 1
};

constexpr int MOM_BT = 64;
constexpr int MOM_G = 4;
constexpr int MOM_U = 4;

__global__ __launch_bounds__(MOM_BT *MOM_G) void moments_kernel(const moments_args a) {
  __shared__ cplx red[MOM_G - 1][4][4][MOM_BT];
  const int p = blockIdx.y;
  const int fi = blockIdx.z / a.pch, ci = blockIdx.z % a.pch;
  const int bb = threadIdx.x % MOM_BT, g = threadIdx.x / MOM_BT;
  const int b = blockIdx.x * MOM_BT + bb;
  const int K = a.size[p];
  const int i0 = (int)((long)K * ci / a.pch), i1 = (int)((long)K * (ci + 1) / a.pch);
  const int nlm = a.max_nlm;
  const long off = (long)p * a.max_sphere;
  const double *M = a.mat + (long)p * nlm * a.max_sphere;
  const cplx *f = a.f[fi];
  const long fsize = (long)a.nproj * nlm * 4 * a.nst;
  cplx *dst = a.pch > 1 ? a.partial + ((long)fi * a.pch + ci) * fsize : a.out + fi * a.out_stride;
  for (int lm0 = 0; lm0 < nlm; lm0 += 4) {
    cplx acc[4][4];
#pragma unroll
    for (int l = 0; l < 4; l++)
#pragma unroll
      for (int c = 0; c < 4; c++) acc[l][c] = mk(0.0, 0.0);
    if (b < a.nst) {
      // MOM_U points per pass, all their loads issued before the sums (a
      // point past the chunk end reloads point i and adds zero), so the
      // gathers overlap; the summation order stays i, i + MOM_G, ...
      for (int i = i0 + g; i < i1; i += MOM_G * MOM_U) {
        int pt[MOM_U];
        bool ok[MOM_U];
#pragma unroll
        for (int u = 0; u < MOM_U; u++) {
          ok[u] = i + u * MOM_G < i1;
          pt[u] = ok[u] ? i + u * MOM_G : i;
        }
        cplx x[MOM_U];
        double w[MOM_U][4], m[MOM_U][4];
#pragma unroll
        for (int u = 0; u < MOM_U; u++) {
          const long q = off + pt[u];
          x[u] = f[(long)a.point[q] * a.ld + b];
          const cplx ph = a.phase[q];
          const float *ps = a.pos + q * 3;
          const double p0 = ps[0], p1 = ps[1], p2 = ps[2];
          w[u][0] = 1.0;
#pragma unroll
          for (int c = 1; c < 4; c++)
            w[u][c] = p0 * a.L[(c - 1) * 3 + 0] + p1 * a.L[(c - 1) * 3 + 1] + p2 * a.L[(c - 1) * 3 + 2];
          if (a.square2) w[u][2] = w[u][1] * w[u][1];
#pragma unroll
          for (int l = 0; l < 4; l++)
            m[u][l] = M[(long)(lm0 + l < nlm ? lm0 + l : nlm - 1) * a.max_sphere + pt[u]];
          x[u] = ok[u] ? cmul(ph, x[u]) : mk(0.0, 0.0);
        }
#pragma unroll
        for (int u = 0; u < MOM_U; u++) {
#pragma unroll
          for (int l = 0; l < 4; l++) {
#pragma unroll
            for (int c = 0; c < 4; c++) {
              const double mw = m[u][l] * w[u][c];
              acc[l][c].x += mw * x[u].x;
              acc[l][c].y += mw * x[u].y;
            }
          }
        }
      }
    }
    if (g > 0) {
#pragma unroll
      for (int l = 0; l < 4; l++)
#pragma unroll
        for (int c = 0; c < 4; c++) red[g - 1][l][c][bb] = acc[l][c];
    }
    __syncthreads();
    if (g == 0 && b < a.nst) {
      for (int gg = 0; gg < MOM_G - 1; gg++)
#pragma unroll
        for (int l = 0; l < 4; l++)
#pragma unroll
          for (int c = 0; c < 4; c++) acc[l][c] = cadd(acc[l][c], red[gg][l][c][bb]);
#pragma unroll
      for (int l = 0; l < 4; l++)
        if (lm0 + l < nlm)
#pragma unroll
          for (int c = 0; c < 4; c++)
            dst[(((long)p * nlm + lm0 + l) * 4 + c) * a.nst + b] = acc[l][c];
    }
    __syncthreads();
  }
}

// out[f][e] = sum over chunks of partial[f][chunk][e], in chunk order
__global__ __launch_bounds__(256) void moments_reduce_kernel(const cplx *partial, int nf, int pch,
                                                             long fsize, cplx *out,
                                                             long out_stride) {
  const long total = (long)nf * fsize;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const long fi = e / fsize, r = e - fi * fsize;
    const cplx *src = partial + fi * pch * fsize + r;
    cplx s = src[0];
    for (int c = 1; c < pch; c++) s = cadd(s, src[(long)c * fsize]);
    out[fi * out_stride + r] = s;
  }
}

__device__ __forceinline__ int sym_index(int i, int n) { return i >= (n + 1) / 2 ? i - n : i; }

// Per-block partial sums of conj(X^) Y^ m_c over G and bands, m_0 = 1 and
// m_1..3 the symmetric indices (x, y, z); 8 doubles (4 complex) per block.
__global__ __launch_bounds__(256) void gspace_moments_kernel(const cplx *X, const cplx *Y, long ld,
                                                             int nb, int n0, int n1, int n2,
                                                             double *partial) {
  __shared__ double red[8][256];
  const long total = (long)n0 * n1 * n2 * nb;
  double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const long p = e / nb;
    const int b = (int)(e - p * nb);
    const cplx x = X[p * ld + b], y = Y[p * ld + b];
    const double re = x.x * y.x + x.y * y.y, im = x.x * y.y - x.y * y.x;
    const int iz = (int)(p % n2);
    const long pxy = p / n2;
    const int iy = (int)(pxy % n1), ix = (int)(pxy / n1);
    const double m0 = sym_index(ix, n0), m1 = sym_index(iy, n1), m2 = sym_index(iz, n2);
    s[0] += re;
    s[1] += im;
    s[2] += m0 * re;
    s[3] += m0 * im;
    s[4] += m1 * re;
    s[5] += m1 * im;
    s[6] += m2 * re;
    s[7] += m2 * im;
  }
#pragma unroll
  for (int k = 0; k < 8; k++) red[k][threadIdx.x] = s[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int k = 0; k < 8; k++) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 8) partial[(long)blockIdx.x * 8 + threadIdx.x] = red[threadIdx.x][0];
}

// Per (sphere, lm) sums over bands of the coefficient-weighted moment pairs:
//   A_c = sum_b D conj(mX_0) mY_c,  B_c = sum_b D conj(mX_c) mY_0  (c = 0..3),
//   C = sum_b D conj(mX_1) mY_1;  18 doubles per row.
__global__ __launch_bounds__(64) void nl_pair_kernel(const cplx *mX, const cplx *mY,
                                                     const double *coeff, int nlm, int nst,
                                                     double *partial) {
  __shared__ double red[18][64];
  const int row = blockIdx.x;  // p * nlm + lm
  const double D = coeff[row];
  double s[18];
#pragma unroll
  for (int k = 0; k < 18; k++) s[k] = 0.0;
  for (int b = threadIdx.x; b < nst; b += 64) {
    cplx x[4], y[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
      x[c] = mX[((long)row * 4 + c) * nst + b];
      y[c] = mY[((long)row * 4 + c) * nst + b];
    }
#pragma unroll
    for (int c = 0; c < 4; c++) {
      // conj(x0) y_c and conj(x_c) y0
      s[2 * c] += D * (x[0].x * y[c].x + x[0].y * y[c].y);
      s[2 * c + 1] += D * (x[0].x * y[c].y - x[0].y * y[c].x);
      s[8 + 2 * c] += D * (x[c].x * y[0].x + x[c].y * y[0].y);
      s[8 + 2 * c + 1] += D * (x[c].x * y[0].y - x[c].y * y[0].x);
    }
    s[16] += D * (x[1].x * y[1].x + x[1].y * y[1].y);
    s[17] += D * (x[1].x * y[1].y - x[1].y * y[1].x);
  }
#pragma unroll
  for (int k = 0; k < 18; k++) red[k][threadIdx.x] = s[k];
  __syncthreads();
  for (int w = 32; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int k = 0; k < 18; k++) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 18) partial[(long)row * 18 + threadIdx.x] = red[threadIdx.x][0];
}

// out[k] = sum over rows of partial[row][k] (k < width), one 256-thread
// block: thread t adds rows t, t + 256, ..., then a tree over the threads, a
// fixed order.
constexpr int SUM_ROWS_NT = 256;
__global__ __launch_bounds__(SUM_ROWS_NT) void sum_rows_kernel(const double *partial, int nrows,
                                                               int width, double *out) {
  __shared__ double red[SUM_ROWS_NT];
  for (int k = 0; k < width; k++) {
    double s = 0.0;
    for (int r = threadIdx.x; r < nrows; r += SUM_ROWS_NT) s += partial[(long)r * width + k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = SUM_ROWS_NT / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = red[0];
    __syncthreads();
  }
}

// Record layout (OBS_RECORD doubles): [0..7] G-space S_0..S_3 (complex),
// [8..15] A_0..A_3, [16..23] B_0..B_3, [24..25] C.
struct pair_weights {
  double inv_n;  // 1/N
  double bd[3];  // b_e . d (reciprocal vectors, Cartesian)
  double kad;    // (k + A)_cart . d
  double dV;
};

// Re <X|W|Y> from one record.
__device__ __forceinline__ double pair_value(const double *r, const pair_weights &w) {
  const double kin = w.bd[0] * r[2] + w.bd[1] * r[4] + w.bd[2] * r[6] + w.kad * r[0];
  // Re(i dV (A_1 - B_1)) = -dV Im(A_1 - B_1)
  const double nl = -w.dV * (r[8 + 3] - r[16 + 3]);
  return w.inv_n * kin + nl;
}

struct grad_step_args {
  int nrec;         // records (one per j)
  pair_weights w;
  double t, omega;  // time of the half step
  int ncoef;
  double scale;     // multiplies the sum before it is accumulated
};

// dL/ds of one half step from its records, accumulated into
// gcoef[k] += g sin((k + 1) omega t).
__global__ void grad_step_kernel(const double *rec, const grad_step_args a, double *gcoef) {
  __shared__ double g;
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int j = 0; j < a.nrec; j++) s += pair_value(rec + (long)j * OBS_RECORD, a.w);
    g = a.scale * s;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < a.ncoef; k += blockDim.x)
    gcoef[k] += g * sin((k + 1) * a.omega * a.t);
}

// The explicit dependence of L = -occ dV <psi|W(A_T)|psi> on s at T:
//   dL/ds = -occ dV (|d|^2 <psi|psi> + dV sum D (2|r|^2 - 2 Re conj(p) s2))
// with record moments of psi (c = 0: p, 1: r, 2: the w^2-weighted s2).
struct final_term_args {
  double occ, dV, inv_n, dd;
  double t, omega;
  int ncoef;
  double scale;
};
__global__ void final_term_kernel(const double *rec, const final_term_args a, double *gcoef) {
  __shared__ double g;
  if (threadIdx.x == 0) {
    const double norm = a.inv_n * rec[0];  // Parseval: sum_r |psi|^2 = (1/N) sum_G |psi^|^2
    const double nl = a.dV * (2.0 * rec[24] - 2.0 * rec[8 + 4]);
    g = a.scale * (-a.occ * a.dV * (a.dd * norm + nl));
  }
  __syncthreads();
  for (int k = threadIdx.x; k < a.ncoef; k += blockDim.x)
    gcoef[k] += g * sin((k + 1) * a.omega * a.t);
}

// loss = -occ dV Re <psi|W|psi>
__global__ void loss_kernel(const double *rec, const pair_weights w, double occ, double *loss) {
  if (threadIdx.x == 0) loss[0] = -occ * w.dV * pair_value(rec, w);
}

// G(G, b) *= c (m . bd + kad)   (W's kinetic factor, scaled)
__global__ __launch_bounds__(256) void gspace_weight_kernel(cplx *G, long ld, int nb, int n0,
                                                            int n1, int n2, double bd0, double bd1,
                                                            double bd2, double kad, double c) {
  const long total = (long)n0 * n1 * n2 * nb;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const long p = e / nb;
    const int b = (int)(e - p * nb);
    const int iz = (int)(p % n2);
    const long pxy = p / n2;
    const int iy = (int)(pxy % n1), ix = (int)(pxy / n1);
    const double w =
        sym_index(ix, n0) * bd0 + sym_index(iy, n1) * bd1 + sym_index(iz, n2) * bd2 + kad;
    G[p * ld + b] = cscale(c * w, G[p * ld + b]);
  }
}

// sph[p][i][b] = c * i * sum_lm M[p][lm][i] D[p][lm] (r_lm,b - w_i p_lm,b)
// with w_i = pos_i . L0 (the sphere part of dV_nl/ds psi before the phase).
struct wsphere_args {
  const cplx *mom;  // [nproj][nlm][4][nb] (c = 0: p, 1: r)
  int nb, nproj, max_sphere, max_nlm;
  const double *mat, *coeff;
  const float *pos;
  const int *size;
  double L0[3];
  double c;
  cplx *sph;
};
__global__ __launch_bounds__(256) void wsphere_kernel(const wsphere_args a) {
  const long total = (long)a.nproj * a.max_sphere * a.nb;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const long pi = e / a.nb;
    const int b = (int)(e - pi * a.nb);
    const int p = (int)(pi / a.max_sphere);
    const int i = (int)(pi - (long)p * a.max_sphere);
    cplx v = mk(0.0, 0.0);
    if (i < a.size[p]) {
      const float *ps = a.pos + pi * 3;
      const double w = ps[0] * a.L0[0] + ps[1] * a.L0[1] + ps[2] * a.L0[2];
      for (int lm = 0; lm < a.max_nlm; lm++) {
        const double md = a.mat[((long)p * a.max_nlm + lm) * a.max_sphere + i] *
                          a.coeff[(long)p * a.max_nlm + lm];
        const cplx pp = a.mom[(((long)p * a.max_nlm + lm) * 4 + 0) * a.nb + b];
        const cplx rr = a.mom[(((long)p * a.max_nlm + lm) * 4 + 1) * a.nb + b];
        v.x += md * (rr.x - w * pp.x);
        v.y += md * (rr.y - w * pp.y);
      }
      v = cscale(a.c, mk(-v.y, v.x));  // c * i * v
    }
    a.sph[pi * a.nb + b] = v;
  }
}

// dst_t = c_t src (set) or dst_t += c_t src, for up to 9 targets
struct axpy_args {
  const cplx *src;
  long ld;
  int nb;
  long total;
  int nt;
  int set;
  cplx *dst[9];
  double cre[9], cim[9];
};
__global__ __launch_bounds__(256) void axpy_multi_kernel(const axpy_args a) {
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < a.total; e += (long)gridDim.x * 256) {
    const long p = e / a.nb;
    const int b = (int)(e - p * a.nb);
    const cplx s = a.src[p * a.ld + b];
    for (int t = 0; t < a.nt; t++) {
      const cplx v = mk(a.cre[t] * s.x - a.cim[t] * s.y, a.cre[t] * s.y + a.cim[t] * s.x);
      cplx *d = a.dst[t] + p * a.ld + b;
      *d = a.set ? v : cadd(*d, v);
    }
  }
}

// gV[p] += Re sum_j sum_b conj(X_j(p, b)) Y_j(p, b); one wave per grid point
struct vgrad_args {
  const cplx *X[8];
  const cplx *Y[8];
  int nj;
  long ld;
  int nb;
  long npoint;
  double *gV;
};
__global__ __launch_bounds__(256) void vgrad_kernel(const vgrad_args a) {
  const int lane = threadIdx.x & 63;
  const long p = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (p >= a.npoint) return;
  double s = 0.0;
  for (int j = 0; j < a.nj; j++) {
    const cplx *X = a.X[j] + p * a.ld;
    const cplx *Y = a.Y[j] + p * a.ld;
    for (int b = lane; b < a.nb; b += 64) {
      const cplx x = X[b], y = Y[b];
      s += x.x * y.x + x.y * y.y;
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
  if (lane == 0) a.gV[p] += s;
}

__global__ __launch_bounds__(256) void dzero_kernel(double *x, long n) {
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long)gridDim.x * 256) x[e] = 0.0;
}

// ---------------------------------------------------------------------------
// fused half-step gradient kernels (three-sweep Hpsi)
//
// With A_k = H^k lambda' and Y_j = H^j phi (k, j < M) and the coefficient
// matrix E[k][j] = z^(k+j+1)/(k+j+1)! (zero for k + j >= M), the sums over
// X_j of the half step become
//   dL/dV(r) += Re sum_b sum_kj E[k][j] conj(A_k) Y_j
//   dL/ds    += Re sum_kj E[k][j] <A_k|W|Y_j>
// so no X_j is formed; the forward transforms of A_k and Y_j come from the
// G-space dumps of the Hpsi calls that produce the powers.
// ---------------------------------------------------------------------------
constexpr int TD_MAXM = 8;

struct coef_matrix {
  int M;
  double er[TD_MAXM][TD_MAXM], ei[TD_MAXM][TD_MAXM];
};

// Re(e conj(a) y)
__device__ __forceinline__ double re_e_conj_a_y(double er, double ei, cplx a, cplx y) {
  const double pr = a.x * y.x + a.y * y.y, pi = a.x * y.y - a.y * y.x;
  return er * pr - ei * pi;
}

struct vgrad_multi_args {
  const cplx *A[TD_MAXM];
  const cplx *Y[TD_MAXM];
  long ld;
  int nb;
  long npoint;
  double *gV;
};

// MM = E.M (the Taylor order) as a template argument: the loads of a band's
// M powers of each field are all issued before the sums use them (with a
// runtime M each load sat in a branch of its own and waited for itself)
template <int MM>
__global__ __launch_bounds__(256) void vgrad_multi_kernel(const vgrad_multi_args a,
                                                          const coef_matrix E) {
  const int lane = threadIdx.x & 63;
  const long p = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (p >= a.npoint) return;
  double s = 0.0;
  for (int b = lane; b < a.nb; b += 64) {
    cplx y[MM], x[MM];
#pragma unroll
    for (int j = 0; j < MM; j++) y[j] = a.Y[j][p * a.ld + b];
#pragma unroll
    for (int k = 0; k < MM; k++) x[k] = a.A[k][p * a.ld + b];
#pragma unroll
    for (int k = 0; k < MM; k++) {
#pragma unroll
      for (int j = 0; j < MM; j++)
        if (k + j < MM) s += re_e_conj_a_y(E.er[k][j], E.ei[k][j], x[k], y[j]);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
  if (lane == 0) a.gV[p] += s;
}

struct gpair_args {
  const cplx *A[TD_MAXM];  // band-major [b][kz][kx][ky] forward transforms
  const cplx *Y[TD_MAXM];
  int L;
  long total;  // nb * L^3
  double bd[3], kad;
};

// per-block partial sums of Re sum_kj E[k][j] w(G) conj(A_k^) Y_j^ (MM = E.M,
// as in vgrad_multi_kernel)
template <int MM>
__global__ __launch_bounds__(256) void gpair_kernel(const gpair_args a, const coef_matrix E,
                                                    double *partial) {
  __shared__ double red[256];
  const int L = a.L;
  const long L2 = (long)L * L, L3 = L2 * L;
  double s = 0.0;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < a.total; e += (long)gridDim.x * 256) {
    const long r = e % L3;
    const int kz = (int)(r / L2), kx = (int)((r / L) % L), ky = (int)(r % L);
    const double w = sym_index(kx, L) * a.bd[0] + sym_index(ky, L) * a.bd[1] +
                     sym_index(kz, L) * a.bd[2] + a.kad;
    cplx y[MM], x[MM];
#pragma unroll
    for (int j = 0; j < MM; j++) y[j] = a.Y[j][e];
#pragma unroll
    for (int k = 0; k < MM; k++) x[k] = a.A[k][e];
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < MM; k++) {
#pragma unroll
      for (int j = 0; j < MM; j++)
        if (k + j < MM) t += re_e_conj_a_y(E.er[k][j], E.ei[k][j], x[k], y[j]);
    }
    s += w * t;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// per (sphere, lm): S = sum_b D sum_kj E[k][j] (conj(p^A_k) r^Y_j - conj(r^A_k) p^Y_j)
__global__ __launch_bounds__(64) void nl_multi_kernel(const cplx *mA, const cplx *mY,
                                                      long field_stride, const double *coeff,
                                                      int nb, const coef_matrix E,
                                                      double *partial) {
  __shared__ double red[2][64];
  const int row = blockIdx.x;
  const int M = E.M;
  const double D = coeff[row];
  double sr = 0.0, si = 0.0;
  for (int b = threadIdx.x; b < nb; b += 64) {
    for (int k = 0; k < M; k++) {
      const cplx pa = mA[k * field_stride + ((long)row * 4 + 0) * nb + b];
      const cplx ra = mA[k * field_stride + ((long)row * 4 + 1) * nb + b];
      for (int j = 0; j + k < M; j++) {
        const cplx py = mY[j * field_stride + ((long)row * 4 + 0) * nb + b];
        const cplx ry = mY[j * field_stride + ((long)row * 4 + 1) * nb + b];
        // u = conj(pa) ry - conj(ra) py
        const double ur = (pa.x * ry.x + pa.y * ry.y) - (ra.x * py.x + ra.y * py.y);
        const double ui = (pa.x * ry.y - pa.y * ry.x) - (ra.x * py.y - ra.y * py.x);
        const double er = E.er[k][j], ei = E.ei[k][j];
        sr += D * (er * ur - ei * ui);
        si += D * (er * ui + ei * ur);
      }
    }
  }
  red[0][threadIdx.x] = sr;
  red[1][threadIdx.x] = si;
  __syncthreads();
  for (int w = 32; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      red[0][threadIdx.x] += red[0][threadIdx.x + w];
      red[1][threadIdx.x] += red[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x < 2) partial[(long)row * 2 + threadIdx.x] = red[threadIdx.x][0];
}

struct grad_fused_args {
  double inv_n, dV;
  double t, omega;
  int ncoef;
  double scale;
};

// rec[0] = kinetic sum, rec[1..2] = S of nl_multi; g = inv_n kin + Re(i dV S)
__global__ void grad_fused_kernel(const double *rec, const grad_fused_args a, double *gcoef) {
  __shared__ double g;
  if (threadIdx.x == 0) g = a.scale * (a.inv_n * rec[0] - a.dV * rec[2]);
  __syncthreads();
  for (int k = threadIdx.x; k < a.ncoef; k += blockDim.x)
    gcoef[k] += g * sin((k + 1) * a.omega * a.t);
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
static unsigned blocks_for(long total) {
  long b = (total + 255) / 256;
  if (b > 65536L * 4) b = 65536L * 4;
  return (unsigned)(b < 1 ? 1 : b);
}

long observables_scratch_doubles(const handle &h) {
  const long nl = (long)h.pt.nproj * h.pt.max_nlm * 18;
  const long gs = (long)OBS_GS_BLOCKS * 8;
  return (nl > gs ? nl : gs) + (long)OBS_NRECORDS * OBS_RECORD;
}

static double *obs_partial(handle &h) { return h.obs; }
static double *obs_record(handle &h, int r) {
  const long nl = (long)h.pt.nproj * h.pt.max_nlm * 18;
  const long gs = (long)OBS_GS_BLOCKS * 8;
  return h.obs + (nl > gs ? nl : gs) + (long)r * OBS_RECORD;
}

// Point chunks per sphere so that a moments launch fills the device.
static int moment_chunks(const handle &h, int nb, int nf) {
  const long blocks = (long)((nb + MOM_BT - 1) / MOM_BT) * h.pt.nproj * nf;
  long pch = (2048 + blocks - 1) / (blocks > 0 ? blocks : 1);
  if (pch > MOM_MAX_CHUNKS) pch = MOM_MAX_CHUNKS;
  return (int)(pch < 1 ? 1 : pch);
}

// sphere moments of nf fields (phases of the slot must be current); field
// f's moments go to out + f * out_stride
static void sphere_moments_multi(handle &h, const cplx *const *f, int nf, long ld, int nb,
                                 const double L[9], bool square2, cplx *out, long out_stride) {
  const projector_tables &t = h.pt;
  if (t.nproj == 0) return;
  moments_args a{};
  for (int i = 0; i < nf; i++) a.f[i] = f[i];
  a.nf = nf;
  a.ld = ld;
  a.nst = nb;
  a.nproj = t.nproj;
  a.max_sphere = t.max_sphere;
  a.max_nlm = t.max_nlm;
  a.point = t.point;
  a.size = t.size;
  a.mat = t.mat;
  a.pos = t.pos;
  a.phase = t.phase_sph;
  for (int i = 0; i < 9; i++) a.L[i] = L[i];
  a.square2 = square2 ? 1 : 0;
  a.out = out;
  a.out_stride = out_stride;
  const long fsize = (long)t.nproj * t.max_nlm * 4 * nb;
  a.pch = moment_chunks(h, nb, nf);
  if ((long)a.pch * nf * fsize > h.mom_partial_elems) a.pch = 1;
  a.partial = h.mom_partial;
  dim3 grid((nb + MOM_BT - 1) / MOM_BT, t.nproj, nf * a.pch);
  hipLaunchKernelGGL(moments_kernel, grid, dim3(MOM_BT * MOM_G), 0, h.stream, a);
  INQAMD_HIP_CHECK(hipGetLastError());
  if (a.pch > 1) {
    const long total = nf * fsize;
    hipLaunchKernelGGL(moments_reduce_kernel, dim3(blocks_for(total)), dim3(256), 0, h.stream,
                       (const cplx *)h.mom_partial, nf, a.pch, fsize, out, out_stride);
    INQAMD_HIP_CHECK(hipGetLastError());
  }
}

static void sphere_moments(handle &h, const cplx *f, long ld, int nb, const double L[9],
                           bool square2, cplx *out) {
  sphere_moments_multi(h, &f, 1, ld, nb, L, square2, out, 0);
}

// record[0..7] = G-space moments of (Xh, Yh)
static void gspace_record(handle &h, const cplx *Xh, const cplx *Yh, long ld, int nb, double *rec) {
  double *part = obs_partial(h);
  hipLaunchKernelGGL(gspace_moments_kernel, dim3(OBS_GS_BLOCKS), dim3(256), 0, h.stream, Xh, Yh,
                     ld, nb, h.n[0], h.n[1], h.n[2], part);
  INQAMD_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(sum_rows_kernel, dim3(1), dim3(SUM_ROWS_NT), 0, h.stream,
                     (const double *)part, OBS_GS_BLOCKS, 8, rec);
  INQAMD_HIP_CHECK(hipGetLastError());
}

// record[8..25] = sphere pair sums of (mX, mY), weighted by coeff
// [nproj][max_nlm] (the KB coefficients, or commutator_coeff's)
static void nl_record(handle &h, const cplx *mX, const cplx *mY, int nb, double *rec,
                      const double *coeff) {
  const projector_tables &t = h.pt;
  double *part = obs_partial(h);
  const int nrows = t.nproj * t.max_nlm;
  if (nrows > 0) {
    hipLaunchKernelGGL(nl_pair_kernel, dim3(nrows), dim3(64), 0, h.stream, mX, mY,
                       coeff, t.max_nlm, nb, part);
    INQAMD_HIP_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(sum_rows_kernel, dim3(1), dim3(SUM_ROWS_NT), 0, h.stream,
                     (const double *)part, nrows, 18, rec + 8);
  INQAMD_HIP_CHECK(hipGetLastError());
}

static void zero_doubles(handle &h, double *x, long n) {
  hipLaunchKernelGGL(dzero_kernel, dim3(blocks_for(n)), dim3(256), 0, h.stream, x, n);
  INQAMD_HIP_CHECK(hipGetLastError());
}

static void axpy_multi(handle &h, const cplx *src, long ld, int nb, int nt, cplx *const *dst,
                       const std::complex<double> *c, bool set) {
  if (nt == 0) return;
  axpy_args a{};
  a.src = src;
  a.ld = ld;
  a.nb = nb;
  a.total = h.npoint * nb;
  a.nt = nt;
  a.set = set ? 1 : 0;
  for (int t = 0; t < nt; t++) {
    a.dst[t] = dst[t];
    a.cre[t] = c[t].real();
    a.cim[t] = c[t].imag();
  }
  hipLaunchKernelGGL(axpy_multi_kernel, dim3(blocks_for(a.total)), dim3(256), 0, h.stream, a);
  INQAMD_HIP_CHECK(hipGetLastError());
}

// Cartesian vector from covariant components: v = sum_d cov_d b_d / 2 pi
static void cov_to_cart(const handle &h, const double cov[3], double cart[3]) {
  for (int c = 0; c < 3; c++)
    cart[c] = (cov[0] * h.rlat[0 * 3 + c] + cov[1] * h.rlat[1 * 3 + c] +
               cov[2] * h.rlat[2 * 3 + c]) /
              (2.0 * M_PI);
}

// a_e . d for the lattice vectors a_e
static void lattice_dot(const handle &h, const double d[3], double out[3]) {
  for (int e = 0; e < 3; e++)
    out[e] = h.lattice[3 * e + 0] * d[0] + h.lattice[3 * e + 1] * d[1] + h.lattice[3 * e + 2] * d[2];
}

// Coefficients of the current's [V_nl, r] term in the A-derivative form:
// D itself (INQAMD_COMMUTATOR_CORRECTED) or (1 + D^2)/2, which gives the
// current of position_commutator as the reference writes it
// (INQAMD_COMMUTATOR_INQ, projector_all.hpp:362-366).
static const double *commutator_coeff(const handle &h, int commutator) {
  return commutator == INQAMD_COMMUTATOR_CORRECTED ? h.pt.coeff : h.pt.coeff_inq;
}

static pair_weights make_weights(const handle &h, const double d[3], const double kA_cov[3]) {
  pair_weights w;
  w.inv_n = 1.0 / (double)h.npoint;
  for (int e = 0; e < 3; e++)
    w.bd[e] = h.rlat[3 * e + 0] * d[0] + h.rlat[3 * e + 1] * d[1] + h.rlat[3 * e + 2] * d[2];
  double kc[3];
  cov_to_cart(h, kA_cov, kc);
  w.kad = kc[0] * d[0] + kc[1] * d[1] + kc[2] * d[2];
  w.dV = h.volume / (double)h.npoint;
  return w;
}

}  // namespace inqamd

// ===========================================================================
// propagation object
// ===========================================================================
using namespace inqamd;

struct inqamd_td {
  inqamd_hamiltonian *h = nullptr;
  inqamd_td_desc d{};
  int nseg = 0, nb = 0;
  std::vector<double> coef;   // coefficients of the last forward
  bool have_forward = false;
  long generation = 0;        // the handle's V/k generation at the last forward
  long forwards = 0;          // forwards run so far
  double dir_cov[3] = {0, 0, 0};
  long bytes = 0;
  cplx *ckpt = nullptr;       // [nseg + 1][N][nst]; the last one holds psi(T)
  cplx *tape = nullptr;       // [2K][nstore][N][nb]
  int nstore = 0;             // powers kept per half step (order or 1)
  cplx *cur = nullptr;        // [N][nb]
  cplx *lam = nullptr, *lam_in = nullptr;
  cplx *hk[2] = {nullptr, nullptr};
  cplx *X[8] = {};
  cplx *Ybuf[8] = {};         // recomputed powers when not stored
  cplx *gA = nullptr, *gB = nullptr;  // G-space temporaries [N][nb]
  cplx *sph = nullptr;        // [nproj][max_sphere][nb]
  // fused path (three-sweep Hpsi): the tape holds Y_j and their forward
  // transforms; A_k = H^k lambda' (k >= 1) and their transforms
  bool fused = false;
  cplx *Ak[8] = {};
  cplx *Ah[8] = {};
  cplx *momA = nullptr, *momY = nullptr;  // [order][nproj][nlm][4][nb]
  long mom_stride = 0;
  double *gcoef = nullptr;    // device [ncoef]
  double *gV = nullptr;       // device [N]
  double *loss = nullptr;     // device [1]
};

namespace {

void *td_alloc(inqamd_td *td, size_t bytes) {
  if (bytes == 0) return nullptr;
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess)
    throw error{ERR_MEMORY, "propagation workspace: hipMalloc(" + std::to_string(bytes) +
                                "): " + hipGetErrorString(e)};
  td->bytes += (long)bytes;
  return p;
}

void td_release(inqamd_td *td) {
  std::vector<void *> ps = {td->ckpt, td->tape, td->cur,   td->lam,  td->lam_in, td->hk[0],
                            td->hk[1], td->gA,  td->gB,    td->sph,  td->gcoef,  td->gV,
                            td->loss};
  for (int j = 0; j < 8; j++) {
    ps.push_back(td->X[j]);
    ps.push_back(td->Ybuf[j]);
    ps.push_back(td->Ak[j]);
    ps.push_back(td->Ah[j]);
  }
  ps.push_back(td->momA);
  ps.push_back(td->momY);
  for (void *p : ps)
    if (p) (void)hipFree(p);
}

double series_value(const inqamd_td *td, double t) {
  double f = 0.0;
  for (int j = 0; j < td->d.ncoef; j++) f += td->coef[j] * std::sin((j + 1) * td->d.omega * t);
  return f;
}

// k + A(t), covariant
void ka_at(const inqamd_td *td, double t, double out[3]) {
  const double f = series_value(td, t);
  for (int e = 0; e < 3; e++) out[e] = td->h->k_cov[e] + td->dir_cov[e] * f;
}

// copy bands [b0, b0 + nb) of x (ld) into y (ld = nb)
void slice_copy(handle &h, const cplx *x, long xld, cplx *y, long yld, int nb) {
  copy(h, x, xld, y, yld, nb);
}

// One Taylor half step of `nb` bands (ld = nb), writing H^j phi (j <
// nstore) into the tape slots Y[j] (Y[0] = phi).  The arithmetic is
// exponential_in_place's: each H application adds its term to the new
// state in its last sweep.  With one slot (nstore == 1) the step runs in
// place on cur after copying phi to Y[0]; with the powers stored, phi is
// already in Y[0] and the new state goes to `next` (the first H
// application writes phi + z H phi there, so nothing is copied).
void replay_half(inqamd_td *td, cplx *cur, int nb, cplx *const *Y, cplx *next, int slot) {
  handle &h = *td->h;
  const int order = td->d.order;
  const std::complex<double> z(0.0, td->d.dt / 2.0);
  if (td->nstore == 1) {
    copy(h, cur, nb, Y[0], nb, nb);
    exponential_in_place(h, z, order, cur, nb, nb, slot);
    return;
  }
  if (taylor_chain_ok(h)) {
    // Y[it] = H^it phi for it < order (the last power only feeds `next`);
    // fused: Y[order + it - 1] = the transform of Y[it - 1]
    cplx *outs[TD_MAXM], *dumps[TD_MAXM];
    for (int it = 1; it <= order; it++) {
      outs[it - 1] = it < order ? Y[it] : nullptr;
      dumps[it - 1] = td->fused ? Y[order + it - 1] : nullptr;
    }
    taylor_chain(h, Y[0], nb, nb, order, slot, next, nb, true, z, nullptr, 0, 0.0, outs,
                 td->fused ? dumps : nullptr);
    return;
  }
  std::complex<double> coeff = 1.0;
  for (int it = 1; it <= order; it++) {
    coeff *= z / double(it);
    taylor_acc acc;
    acc.p1 = next;
    acc.ld1 = nb;
    acc.c1re = coeff.real();
    acc.c1im = coeff.imag();
    acc.init1 = it == 1;
    if (td->fused) acc.gdump = Y[order + it - 1];  // transform of Y_{it-1}
    cplx *out = it < order ? Y[it] : td->hk[0];
    hamiltonian_apply(h, Y[it - 1], nb, out, nb, nb, slot, &acc);
  }
}

// The fused half step: Y[0..M) real powers, Y[M..2M) their transforms.
void adjoint_half_fused(inqamd_td *td, int nb, cplx *const *Y, int slot, double t_h,
                        const double kA[3]) {
  handle &h = *td->h;
  const int M = td->d.order;
  const std::complex<double> z(0.0, td->d.dt / 2.0);
  const std::complex<double> zc = std::conj(z);
  // lambda_in = sum_k conj(z)^k/k! H^k lambda' (the k >= 1 terms fused; the
  // first H application writes lambda' + conj(z) H lambda')
  const cplx *A[TD_MAXM];
  A[0] = td->lam;
  if (taylor_chain_ok(h)) {
    // A_k = H^k lambda' (k < M) into Ak[k], their transforms into Ah[k - 1]
    cplx *outs[TD_MAXM], *dumps[TD_MAXM];
    for (int k = 1; k <= M; k++) {
      outs[k - 1] = k < M ? td->Ak[k] : nullptr;
      dumps[k - 1] = td->Ah[k - 1];
      if (k < M) A[k] = td->Ak[k];
    }
    taylor_chain(h, td->lam, nb, nb, M, slot, td->lam_in, nb, true, zc, nullptr, 0, 0.0, outs,
                 dumps);
  } else {
    std::complex<double> ck = 1.0;
    for (int k = 1; k <= M; k++) {
      ck *= zc / double(k);
      taylor_acc acc;
      acc.p1 = td->lam_in;
      acc.ld1 = nb;
      acc.init1 = k == 1;
      acc.c1re = ck.real();
      acc.c1im = ck.imag();
      acc.gdump = td->Ah[k - 1];
      cplx *out = k < M ? td->Ak[k] : td->hk[0];
      hamiltonian_apply(h, A[k - 1], nb, out, nb, nb, slot, &acc);
      if (k < M) A[k] = out;
    }
  }
  coef_matrix E{};
  E.M = M;
  {
    std::vector<std::complex<double>> e(M + 1);
    e[0] = 1.0;
    for (int m = 1; m <= M; m++) e[m] = e[m - 1] * z / double(m);
    for (int k = 0; k < M; k++)
      for (int j = 0; k + j < M; j++) {
        E.er[k][j] = e[k + j + 1].real();
        E.ei[k][j] = e[k + j + 1].imag();
      }
  }
  // dL/dV
  {
    vgrad_multi_args a{};
    for (int k = 0; k < M; k++) {
      a.A[k] = A[k];
      a.Y[k] = Y[k];
    }
    a.ld = nb;
    a.nb = nb;
    a.npoint = h.npoint;
    a.gV = td->gV;
    const dim3 grid((unsigned)((h.npoint + 3) / 4));
    switch (M) {
#define INQAMD_VGRAD(m)                                                                   \
  case m:                                                                                \
    hipLaunchKernelGGL(vgrad_multi_kernel<m>, grid, dim3(256), 0, h.stream, a, E);        \
    break;
      INQAMD_VGRAD(1) INQAMD_VGRAD(2) INQAMD_VGRAD(3) INQAMD_VGRAD(4)
      INQAMD_VGRAD(5) INQAMD_VGRAD(6) INQAMD_VGRAD(7) INQAMD_VGRAD(8)
#undef INQAMD_VGRAD
      default: throw error{ERR_ARGUMENT, "Taylor order outside 1..8"};
    }
    INQAMD_HIP_CHECK(hipGetLastError());
  }
  double *rec = obs_record(h, 0);
  double *part = obs_partial(h);
  const pair_weights w = make_weights(h, td->d.direction, kA);
  // kinetic part of dL/ds from the transforms
  {
    gpair_args a{};
    for (int k = 0; k < M; k++) {
      a.A[k] = td->Ah[k];
      a.Y[k] = Y[M + k];
    }
    a.L = h.n[0];
    a.total = h.npoint * nb;
    for (int e = 0; e < 3; e++) a.bd[e] = w.bd[e];
    a.kad = w.kad;
    switch (M) {
#define INQAMD_GPAIR(m)                                                                   \
  case m:                                                                                \
    hipLaunchKernelGGL(gpair_kernel<m>, dim3(OBS_GS_BLOCKS), dim3(256), 0, h.stream, a, E, \
                       part);                                                            \
    break;
      INQAMD_GPAIR(1) INQAMD_GPAIR(2) INQAMD_GPAIR(3) INQAMD_GPAIR(4)
      INQAMD_GPAIR(5) INQAMD_GPAIR(6) INQAMD_GPAIR(7) INQAMD_GPAIR(8)
#undef INQAMD_GPAIR
      default: throw error{ERR_ARGUMENT, "Taylor order outside 1..8"};
    }
    INQAMD_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(sum_rows_kernel, dim3(1), dim3(SUM_ROWS_NT), 0, h.stream,
                       (const double *)part, OBS_GS_BLOCKS, 1, rec);
    INQAMD_HIP_CHECK(hipGetLastError());
  }
  // projector part of dL/ds from sphere moments of A_k and Y_j
  const projector_tables &t = h.pt;
  const int nrows = t.nproj * t.max_nlm;
  if (nrows > 0) {
    projector_phases(h, slot);
    double L[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    lattice_dot(h, td->d.direction, L);
    sphere_moments_multi(h, A, M, nb, nb, L, false, td->momA, td->mom_stride);
    sphere_moments_multi(h, Y, M, nb, nb, L, false, td->momY, td->mom_stride);
    hipLaunchKernelGGL(nl_multi_kernel, dim3(nrows), dim3(64), 0, h.stream,
                       (const cplx *)td->momA, (const cplx *)td->momY, td->mom_stride,
                       (const double *)t.coeff, nb, E, part);
    INQAMD_HIP_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(sum_rows_kernel, dim3(1), dim3(SUM_ROWS_NT), 0, h.stream,
                     (const double *)part, nrows, 2, rec + 1);
  INQAMD_HIP_CHECK(hipGetLastError());
  grad_fused_args ga;
  ga.inv_n = w.inv_n;
  ga.dV = w.dV;
  ga.t = t_h;
  ga.omega = td->d.omega;
  ga.ncoef = td->d.ncoef;
  ga.scale = 1.0;
  hipLaunchKernelGGL(grad_fused_kernel, dim3(1), dim3(64), 0, h.stream, (const double *)rec, ga,
                     td->gcoef);
  INQAMD_HIP_CHECK(hipGetLastError());
  std::swap(td->lam, td->lam_in);
}

// dL/ds and dL/dV contributions of one half step, and lambda <- T(conj z)^dagger lambda'
void adjoint_half(inqamd_td *td, int nb, cplx *const *Y, int slot, double t_h, const double kA[3],
                  double scale) {
  handle &h = *td->h;
  const int M = td->d.order;
  const std::complex<double> zc(0.0, -td->d.dt / 2.0);  // conj(z)
  std::vector<std::complex<double>> c(M + 2);
  c[0] = 1.0;
  for (int m = 1; m <= M; m++) c[m] = c[m - 1] * zc / double(m);
  // Y_j = H^j phi when the powers were not stored
  cplx *Yp[8];
  Yp[0] = Y[0];
  for (int j = 1; j < M; j++) {
    if (td->nstore > 1) {
      Yp[j] = Y[j];
    } else {
      hamiltonian_apply(h, Yp[j - 1], nb, td->Ybuf[j], nb, nb, slot, nullptr);
      Yp[j] = td->Ybuf[j];
    }
  }
  // lambda_in = lambda, X_j = c_{j+1} lambda
  {
    cplx *dst[9];
    std::complex<double> cf[9];
    dst[0] = td->lam_in;
    cf[0] = 1.0;
    for (int j = 0; j < M; j++) {
      dst[1 + j] = td->X[j];
      cf[1 + j] = c[j + 1];
    }
    axpy_multi(h, td->lam, nb, nb, M + 1, dst, cf, true);
  }
  // H^k lambda (k = 1..M): lambda_in += c_k H^k lambda (fused), X_j += c_{k+j+1} H^k lambda
  const cplx *prev = td->lam;
  for (int k = 1; k <= M; k++) {
    cplx *out = td->hk[k & 1];
    taylor_acc acc;
    acc.p1 = td->lam_in;
    acc.ld1 = nb;
    acc.c1re = c[k].real();
    acc.c1im = c[k].imag();
    hamiltonian_apply(h, prev, nb, out, nb, nb, slot, &acc);
    cplx *dst[8];
    std::complex<double> cf[8];
    int nt = 0;
    for (int j = 0; j + k + 1 <= M; j++) {
      dst[nt] = td->X[j];
      cf[nt] = c[k + j + 1];
      nt++;
    }
    axpy_multi(h, out, nb, nb, nt, dst, cf, false);
    prev = out;
  }
  // dL/dV(r) += Re sum_j sum_b conj(X_j) Y_j
  {
    vgrad_args a{};
    for (int j = 0; j < M; j++) {
      a.X[j] = td->X[j];
      a.Y[j] = Yp[j];
    }
    a.nj = M;
    a.ld = nb;
    a.nb = nb;
    a.npoint = h.npoint;
    a.gV = td->gV;
    hipLaunchKernelGGL(vgrad_kernel, dim3((unsigned)((h.npoint + 3) / 4)), dim3(256), 0, h.stream,
                       a);
    INQAMD_HIP_CHECK(hipGetLastError());
  }
  // dL/ds = Re sum_j <X_j|W|Y_j>
  projector_phases(h, slot);
  double L[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  lattice_dot(h, td->d.direction, L);
  for (int j = 0; j < M; j++) {
    double *rec = obs_record(h, j);
    fft3d(h, td->X[j], nb, td->gA, nb, nb, -1, 1.0);
    fft3d(h, Yp[j], nb, td->gB, nb, nb, -1, 1.0);
    gspace_record(h, td->gA, td->gB, nb, nb, rec);
    sphere_moments(h, td->X[j], nb, nb, L, false, h.mom_a);
    sphere_moments(h, Yp[j], nb, nb, L, false, h.mom_b);
    nl_record(h, h.mom_a, h.mom_b, nb, rec, h.pt.coeff);
  }
  grad_step_args ga;
  ga.nrec = M;
  ga.w = make_weights(h, td->d.direction, kA);
  ga.t = t_h;
  ga.omega = td->d.omega;
  ga.ncoef = td->d.ncoef;
  ga.scale = scale;
  hipLaunchKernelGGL(grad_step_kernel, dim3(1), dim3(64), 0, h.stream, (const double *)obs_record(h, 0),
                     ga, td->gcoef);
  INQAMD_HIP_CHECK(hipGetLastError());
  std::swap(td->lam, td->lam_in);
}

// lambda(T) = -2 occ dV W psi(T) for the bands in cur (ld = nb), and the
// explicit dL/ds(T) term into gcoef.  `slot` holds k + A(T).
void adjoint_start(inqamd_td *td, const cplx *psi, int nb, int slot, const double kA[3],
                   double t, double scale) {
  handle &h = *td->h;
  const double occ = td->d.occ;
  const double dV = h.volume / (double)h.npoint;
  const double *d = td->d.direction;
  pair_weights w = make_weights(h, d, kA);
  projector_phases(h, slot);
  double L[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  lattice_dot(h, d, L);
  // explicit term: moments with (1, w, w^2) and the G-space norm
  double *rec = obs_record(h, 0);
  fft3d(h, psi, nb, td->gA, nb, nb, -1, 1.0);
  gspace_record(h, td->gA, td->gA, nb, nb, rec);
  sphere_moments(h, psi, nb, nb, L, true, h.mom_a);
  const double *lcoeff = commutator_coeff(h, td->d.commutator);
  nl_record(h, h.mom_a, h.mom_a, nb, rec, lcoeff);
  final_term_args fa;
  fa.occ = occ;
  fa.dV = dV;
  fa.inv_n = w.inv_n;
  fa.dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  fa.t = t;
  fa.omega = td->d.omega;
  fa.ncoef = td->d.ncoef;
  fa.scale = scale;
  hipLaunchKernelGGL(final_term_kernel, dim3(1), dim3(64), 0, h.stream, (const double *)rec, fa,
                     td->gcoef);
  INQAMD_HIP_CHECK(hipGetLastError());
  // lambda = c F^-1[w(G) psi^] + c i dV e^{-i(k+A).x} sum_lm M D (r - w p), c = -2 occ dV
  const double c = -2.0 * occ * dV * scale;
  const long total = h.npoint * nb;
  hipLaunchKernelGGL(gspace_weight_kernel, dim3(blocks_for(total)), dim3(256), 0, h.stream, td->gA,
                     (long)nb, nb, h.n[0], h.n[1], h.n[2], w.bd[0], w.bd[1], w.bd[2], w.kad,
                     c / (double)h.npoint);
  INQAMD_HIP_CHECK(hipGetLastError());
  fft3d(h, td->gA, nb, td->lam, nb, nb, +1, 1.0);
  const projector_tables &tb = h.pt;
  if (tb.nproj > 0 && tb.nunique > 0) {
    wsphere_args a;
    a.mom = h.mom_a;
    a.nb = nb;
    a.nproj = tb.nproj;
    a.max_sphere = tb.max_sphere;
    a.max_nlm = tb.max_nlm;
    a.mat = tb.mat;
    a.coeff = lcoeff;
    a.pos = tb.pos;
    a.size = tb.size;
    for (int e = 0; e < 3; e++) a.L0[e] = L[e];
    a.c = c * dV;
    a.sph = td->sph;
    const long n = (long)tb.nproj * tb.max_sphere * nb;
    hipLaunchKernelGGL(wsphere_kernel, dim3(blocks_for(n)), dim3(256), 0, h.stream, a);
    INQAMD_HIP_CHECK(hipGetLastError());
    projector_apply_sphere(h, td->sph, td->lam, nb, nb, slot);
  }
}

}  // namespace

extern "C" {

int inqamd_set_position_commutator(inqamd_hamiltonian_t h, int commutator) {
  if (!h) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    INQAMD_REQUIRE(commutator == INQAMD_COMMUTATOR_INQ || commutator == INQAMD_COMMUTATOR_CORRECTED,
                   "commutator must be INQAMD_COMMUTATOR_INQ or INQAMD_COMMUTATOR_CORRECTED");
    h->commutator = commutator;
  });
}

int inqamd_current(inqamd_hamiltonian_t h, const void *psi, int nst, int ld, double occ,
                   double J_cov[3], double J_cart[3]) {
  if (!h) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    check_orbital(*h, psi, "psi");
    check_bands(*h, nst, ld, true);
    INQAMD_REQUIRE(J_cov && J_cart, "outputs are null");
    const cplx *f = (const cplx *)psi;
    double *rec = obs_record(*h, 0);
    // G-space moments of |psi^|^2 (transform.hpp:337-367 into the workspace)
    fft3d(*h, f, ld, h->work, nst, nst, -1, 1.0);
    gspace_record(*h, h->work, h->work, nst, nst, rec);
    // sphere moments with the Cartesian components of the relative positions
    double L[9];
    for (int c = 0; c < 3; c++)
      for (int e = 0; e < 3; e++) L[c * 3 + e] = h->lattice[e * 3 + c];
    projector_phases(*h, KA_CURRENT);
    sphere_moments(*h, f, ld, nst, L, false, h->mom_a);
    nl_record(*h, h->mom_a, h->mom_a, nst, rec, commutator_coeff(*h, h->commutator));
    double r[OBS_RECORD];
    INQAMD_HIP_CHECK(hipMemcpyAsync(r, rec, sizeof(r), hipMemcpyDeviceToHost, h->stream));
    INQAMD_HIP_CHECK(hipStreamSynchronize(h->stream));
    const double kA[3] = {h->k_cov[0] + h->A_cov[0], h->k_cov[1] + h->A_cov[1],
                          h->k_cov[2] + h->A_cov[2]};
    double kc[3];
    cov_to_cart(*h, kA, kc);
    const double dV = h->volume / (double)h->npoint, inv_n = 1.0 / (double)h->npoint;
    for (int c = 0; c < 3; c++) {
      // <psi|W_c|psi> for d = e_c: kinetic (G + k + A)_c, sphere part through r_c
      const double kin = h->rlat[0 * 3 + c] * r[2] + h->rlat[1 * 3 + c] * r[4] +
                         h->rlat[2 * 3 + c] * r[6] + kc[c] * r[0];
      const double nl = -dV * (r[8 + 2 * (c + 1) + 1] - r[16 + 2 * (c + 1) + 1]);
      J_cart[c] = -occ * dV * (inv_n * kin + nl);
    }
    for (int e = 0; e < 3; e++)
      J_cov[e] = h->lattice[3 * e + 0] * J_cart[0] + h->lattice[3 * e + 1] * J_cart[1] +
                 h->lattice[3 * e + 2] * J_cart[2];
  });
}

int inqamd_td_create(inqamd_hamiltonian_t h, const inqamd_td_desc *desc, inqamd_td_t *out) {
  if (!h || !desc || !out) return INQAMD_ERR_ARGUMENT;
  *out = nullptr;
  inqamd_td *td = new inqamd_td();
  td->h = h;
  int rc = guarded(h, [&] {
    const inqamd_td_desc &d = *desc;
    check_bands(*h, d.nst, d.nst, true);
    INQAMD_REQUIRE(d.nsteps >= 0, "nsteps must be >= 0");
    INQAMD_REQUIRE(d.order >= 1 && d.order <= 8, "order must be in [1, 8]");
    INQAMD_REQUIRE(d.ncoef >= 0 && d.ncoef <= 64, "ncoef must be in [0, 64]");
    INQAMD_REQUIRE(d.checkpoint_every >= 1, "checkpoint_every must be >= 1");
    INQAMD_REQUIRE(d.band_slice >= 0, "band_slice must be >= 0");
    INQAMD_REQUIRE(d.commutator == INQAMD_COMMUTATOR_INQ ||
                       d.commutator == INQAMD_COMMUTATOR_CORRECTED,
                   "commutator must be INQAMD_COMMUTATOR_INQ or INQAMD_COMMUTATOR_CORRECTED");
    td->d = d;
    td->nb = d.band_slice == 0 || d.band_slice > d.nst ? d.nst : d.band_slice;
    td->nseg = (d.nsteps + d.checkpoint_every - 1) / d.checkpoint_every;
    lattice_dot(*h, d.direction, td->dir_cov);
    td->fused = h->three_sweep && d.store_powers;
    td->nstore = td->fused ? 2 * d.order : (d.store_powers ? d.order : 1);
    const long N = h->npoint;
    const size_t full = (size_t)N * d.nst * sizeof(cplx);
    const size_t slice = (size_t)N * td->nb * sizeof(cplx);
    const int K = d.checkpoint_every < (d.nsteps > 0 ? d.nsteps : 1) ? d.checkpoint_every
                                                                      : (d.nsteps > 0 ? d.nsteps : 1);
    td->ckpt = (cplx *)td_alloc(td, full * (td->nseg + 1));
    td->tape = (cplx *)td_alloc(td, slice * 2 * K * td->nstore);
    td->cur = (cplx *)td_alloc(td, slice);
    td->lam = (cplx *)td_alloc(td, slice);
    td->lam_in = (cplx *)td_alloc(td, slice);
    td->hk[0] = (cplx *)td_alloc(td, slice);
    td->gA = (cplx *)td_alloc(td, slice);
    if (td->fused) {
      for (int k = 1; k < d.order; k++) td->Ak[k] = (cplx *)td_alloc(td, slice);
      for (int k = 0; k < d.order; k++) td->Ah[k] = (cplx *)td_alloc(td, slice);
      if (h->pt.nproj > 0) {
        td->mom_stride = (long)h->pt.nproj * h->pt.max_nlm * 4 * td->nb;
        td->momA = (cplx *)td_alloc(td, (size_t)td->mom_stride * d.order * sizeof(cplx));
        td->momY = (cplx *)td_alloc(td, (size_t)td->mom_stride * d.order * sizeof(cplx));
      }
    } else {
      td->hk[1] = (cplx *)td_alloc(td, slice);
      for (int j = 0; j < d.order; j++) td->X[j] = (cplx *)td_alloc(td, slice);
      if (td->nstore == 1)
        for (int j = 1; j < d.order; j++) td->Ybuf[j] = (cplx *)td_alloc(td, slice);
      td->gB = (cplx *)td_alloc(td, slice);
    }
    if (h->pt.nproj > 0)
      td->sph = (cplx *)td_alloc(td, (size_t)h->pt.nproj * h->pt.max_sphere * td->nb * sizeof(cplx));
    td->gcoef = (double *)td_alloc(td, (size_t)(d.ncoef > 0 ? d.ncoef : 1) * sizeof(double));
    td->gV = (double *)td_alloc(td, (size_t)N * sizeof(double));
    td->loss = (double *)td_alloc(td, sizeof(double));
  });
  if (rc != INQAMD_OK) {
    td_release(td);
    delete td;
    return rc;
  }
  *out = td;
  return INQAMD_OK;
}

int inqamd_td_forward(inqamd_td_t td, const double *coeff, const void *psi0, int ld, void *psi_T,
                      double *loss) {
  if (!td || !td->h) return INQAMD_ERR_ARGUMENT;
  inqamd_hamiltonian *h = td->h;
  return guarded(h, [&] {
    const inqamd_td_desc &d = td->d;
    check_orbital(*h, psi0, "psi0");
    check_bands(*h, d.nst, ld, true);
    if (psi_T) check_orbital(*h, psi_T, "psi_T");
    INQAMD_REQUIRE(loss, "loss is null");
    INQAMD_REQUIRE(d.ncoef == 0 || coeff, "coefficients are null");
    td->coef.assign(coeff, coeff + d.ncoef);
    const long N = h->npoint;
    const int nst = d.nst;
    const int K = d.checkpoint_every;
    cplx *state = td->ckpt + (size_t)td->nseg * N * nst;
    copy(*h, (const cplx *)psi0, ld, state, nst, nst);
    const std::complex<double> z(0.0, d.dt / 2.0);  // etrs.hpp:36, :67
    for (int n = 0; n < d.nsteps; n++) {
      if (n % K == 0) copy(*h, state, nst, td->ckpt + (size_t)(n / K) * N * nst, nst, nst);
      double k0[3], k1[3];
      ka_at(td, n * d.dt, k0);
      ka_at(td, n * d.dt + d.dt, k1);
      set_slot(*h, KA_T, k0);
      set_slot(*h, KA_TDT, k1);
      exponential_in_place(*h, z, d.order, state, nst, nst, KA_T);
      exponential_in_place(*h, z, d.order, state, nst, nst, KA_TDT);
    }
    // loss = J(T).d = -occ dV <psi|W(A(T))|psi>
    const double T = d.nsteps * d.dt;
    double kT[3];
    ka_at(td, T, kT);
    set_slot(*h, KA_T, kT);
    projector_phases(*h, KA_T);
    double L[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    lattice_dot(*h, d.direction, L);
    double *rec = obs_record(*h, 0);
    fft3d(*h, state, nst, h->work, nst, nst, -1, 1.0);
    gspace_record(*h, h->work, h->work, nst, nst, rec);
    sphere_moments(*h, state, nst, nst, L, false, h->mom_a);
    nl_record(*h, h->mom_a, h->mom_a, nst, rec, commutator_coeff(*h, d.commutator));
    pair_weights w = make_weights(*h, d.direction, kT);
    hipLaunchKernelGGL(loss_kernel, dim3(1), dim3(64), 0, h->stream, (const double *)rec, w, d.occ,
                       td->loss);
    INQAMD_HIP_CHECK(hipGetLastError());
    if (psi_T) copy(*h, state, nst, (cplx *)psi_T, ld, nst);
    // the handle is left at A(T)
    for (int e = 0; e < 3; e++) h->A_cov[e] = kT[e] - h->k_cov[e];
    update_current_slot(*h);
    INQAMD_HIP_CHECK(hipMemcpyAsync(loss, td->loss, sizeof(double), hipMemcpyDeviceToHost,
                                    h->stream));
    INQAMD_HIP_CHECK(hipStreamSynchronize(h->stream));
    td->have_forward = true;
    td->generation = h->generation;
    td->forwards++;
  });
}

int inqamd_td_backward(inqamd_td_t td, double dloss, double *grad_coeff, double *grad_vloc,
                       void *lambda0) {
  if (!td || !td->h) return INQAMD_ERR_ARGUMENT;
  inqamd_hamiltonian *h = td->h;
  return guarded(h, [&] {
    INQAMD_REQUIRE(td->have_forward, "backward needs a forward first");
    INQAMD_REQUIRE(td->generation == h->generation,
                   "the Hamiltonian's scalar potential or k-point changed since the forward; "
                   "run the forward again before the backward");
    INQAMD_REQUIRE(grad_coeff || td->d.ncoef == 0, "grad_coeff is null");
    if (lambda0) check_orbital(*h, lambda0, "lambda0");
    const inqamd_td_desc &d = td->d;
    const long N = h->npoint;
    const int nst = d.nst, nb = td->nb, K = d.checkpoint_every;
    zero_doubles(*h, td->gcoef, d.ncoef > 0 ? d.ncoef : 1);
    zero_doubles(*h, td->gV, N);
    const size_t slice_elems = (size_t)N * nb;
    const double T = d.nsteps * d.dt;
    for (int b0 = 0; b0 < nst; b0 += nb) {
      const int w = nst - b0 < nb ? nst - b0 : nb;
      // lambda(T) from psi(T)
      double kT[3];
      ka_at(td, T, kT);
      set_slot(*h, KA_T, kT);
      slice_copy(*h, td->ckpt + (size_t)td->nseg * N * nst + b0, nst, td->cur, w, w);
      adjoint_start(td, td->cur, w, KA_T, kT, T, dloss);
      for (int s = td->nseg - 1; s >= 0; s--) {
        const int n0 = s * K, n1 = n0 + K < d.nsteps ? n0 + K : d.nsteps;
        // replay the segment, keeping each half step's input (and powers);
        // with the powers stored the state moves from tape slot to tape slot
        cplx *first = td->nstore == 1 ? td->cur : td->tape;
        slice_copy(*h, td->ckpt + (size_t)s * N * nst + b0, nst, first, w, w);
        for (int n = n0; n < n1; n++)
          for (int half = 0; half < 2; half++) {
            double kv[3];
            ka_at(td, n * d.dt + half * d.dt, kv);
            const int slot = half ? KA_TDT : KA_T;
            set_slot(*h, slot, kv);
            cplx *Y[16];
            const size_t base = ((size_t)(n - n0) * 2 + half) * td->nstore;
            for (int j = 0; j < td->nstore; j++) Y[j] = td->tape + (base + j) * slice_elems;
            const bool last = n == n1 - 1 && half == 1;
            cplx *next = last ? td->cur : td->tape + (base + td->nstore) * slice_elems;
            replay_half(td, td->cur, w, Y, next, slot);
          }
        // back through the segment
        for (int n = n1 - 1; n >= n0; n--)
          for (int half = 1; half >= 0; half--) {
            double kv[3];
            const double th = n * d.dt + half * d.dt;
            ka_at(td, th, kv);
            const int slot = half ? KA_TDT : KA_T;
            set_slot(*h, slot, kv);
            cplx *Y[16];
            const size_t base = ((size_t)(n - n0) * 2 + half) * td->nstore;
            for (int j = 0; j < td->nstore; j++) Y[j] = td->tape + (base + j) * slice_elems;
            if (td->fused)
              adjoint_half_fused(td, w, Y, slot, th, kv);
            else
              adjoint_half(td, w, Y, slot, th, kv, 1.0);
          }
      }
      if (lambda0) copy(*h, td->lam, w, (cplx *)lambda0 + b0, nst, w);
    }
    // the adjoint already carries dloss through lambda(T); gV and gcoef are
    // its sums
    if (grad_vloc)
      INQAMD_HIP_CHECK(hipMemcpyAsync(grad_vloc, td->gV, N * sizeof(double),
                                      hipMemcpyDeviceToDevice, h->stream));
    if (d.ncoef > 0)
      INQAMD_HIP_CHECK(hipMemcpyAsync(grad_coeff, td->gcoef, d.ncoef * sizeof(double),
                                      hipMemcpyDeviceToHost, h->stream));
    INQAMD_HIP_CHECK(hipStreamSynchronize(h->stream));
    update_current_slot(*h);
  });
}

long inqamd_td_memory(inqamd_td_t td) { return td ? td->bytes : 0; }

long inqamd_td_forward_count(inqamd_td_t td) { return td ? td->forwards : -1; }

void inqamd_td_destroy(inqamd_td_t td) {
  if (!td) return;
  if (td->h) (void)hipStreamSynchronize(td->h->stream);
  td_release(td);
  delete td;
}

}  // extern "C"
