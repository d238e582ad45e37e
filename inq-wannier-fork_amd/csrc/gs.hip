// Ground-state pieces: operations::laplacian_add, operations::preconditioner
// (Teter-Payne-Allan), operations::orthogonalize,
// eigensolvers::steepest_descent on G-space orbitals and
// ground_state::subspace_diagonalization, the eigensolver steps of
// ground_state/calculator.hpp:150-180.
//
// Steepest descent keeps hphi, the residual and H residual in a workspace of
// three [N][max_bands] arrays allocated on first use.  Its per-band sums run
// as fused streaming passes (partial sums per point chunk, then a fixed-order
// sum over chunks, so results do not depend on scheduling):
//   pass 1  eig = <hphi|phi>, norm = <phi|phi>
//   pass 2  residual = hphi - (eig/norm) phi, written, with the kinetic sums
//           of the preconditioner sum 1/2|G|^2 |r|^2 and sum |r|^2
//   pass 3  residual *= K(1/2|G|^2 / kinetic)
//   H residual (G-space Hpsi)
//   pass 4  <r|r>, <phi|r>, <r|Hr>, <phi|Hr> and the step lambda per band
//   pass 5  phi += lambda r, hphi += lambda Hr (with the next step's pass-1 sums)
// orthogonalize forms S = dV Phi^H Phi (a split-K strided-batched ZGEMM),
// factors it on the host (S = R R^H, R lower) and applies Phi <- Phi R^-H
// with ZGEMM; the subspace diagonalisation diagonalises dV Phi^H H Phi with
// rocSOLVER ZHEEVD.

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cmath>
#include <complex>
#include <cstdlib>
#include <vector>

#include "internal.hpp"
#include "inq_amd.h"

namespace inqamd {
namespace {

constexpr int GS_BT = 64;       // bands per reduction block
constexpr int GS_LANES = 4;     // point lanes per band
constexpr long GS_CHUNK = 512;  // grid points per partial sum
constexpr int IP_BATCHES = 64;  // point chunks of the split-K inner products

struct gs_state {
  cplx *hphi = nullptr, *res = nullptr, *hres = nullptr;  // [N][max_bands] each
  double *part = nullptr;   // [nchunk][max_bands][8] partial sums
  double *band = nullptr;   // [max_bands][16] per-band results
  rocblas_handle blas = nullptr;
  cplx *S = nullptr;        // [max_bands]^2 device overlap / rotation
  double *eig = nullptr;    // [2 max_bands] eigenvalues + tridiagonal work
  cplx *ip_part = nullptr;  // [IP_BATCHES + 1][max_bands]^2 split-K inner-product partials
  rocblas_int *info = nullptr;
  std::vector<std::complex<double>> hostS;
};

#define GS_BLAS_CHECK(expr)                                                              \
  do {                                                                                   \
    rocblas_status _s = (expr);                                                          \
    if (_s != rocblas_status_success)                                                    \
      throw ::inqamd::error{::inqamd::ERR_HIP, std::string(#expr) + ": " +              \
                                                   rocblas_status_to_string(_s)};        \
  } while (0)

long nchunks(const handle &h) { return (h.npoint + GS_CHUNK - 1) / GS_CHUNK; }

gs_state &state(handle &h) {
  if (!h.gs) h.gs = new gs_state();
  return *static_cast<gs_state *>(h.gs);
}

void *dalloc(size_t bytes) {
  void *p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess)
    throw error{ERR_MEMORY, "device allocation of " + std::to_string(bytes) + " bytes failed"};
  return p;
}

void ensure_reductions(handle &h) {
  gs_state &g = state(h);
  if (!g.part) {
    g.part = (double *)dalloc((size_t)nchunks(h) * h.max_bands * 8 * sizeof(double));
    g.band = (double *)dalloc((size_t)h.max_bands * 16 * sizeof(double));
  }
}

void ensure_workspace(handle &h) {
  gs_state &g = state(h);
  ensure_reductions(h);
  if (!g.hphi) {
    const size_t orb = (size_t)h.npoint * h.max_bands * sizeof(cplx);
    g.hphi = (cplx *)dalloc(orb);
    g.res = (cplx *)dalloc(orb);
    g.hres = (cplx *)dalloc(orb);
  }
}

void ensure_blas(handle &h) {
  gs_state &g = state(h);
  if (!g.blas) GS_BLAS_CHECK(rocblas_create_handle(&g.blas));
  if (!g.S) g.S = (cplx *)dalloc((size_t)h.max_bands * h.max_bands * sizeof(cplx));
  if (!g.eig) g.eig = (double *)dalloc((size_t)2 * h.max_bands * sizeof(double));
  if (!g.info) g.info = (rocblas_int *)dalloc(sizeof(rocblas_int));
  if (!g.ip_part)
    g.ip_part = (cplx *)dalloc((size_t)(IP_BATCHES + 1) * h.max_bands * h.max_bands * sizeof(cplx));
  GS_BLAS_CHECK(rocblas_set_stream(g.blas, h.stream));
}

// Grid geometry for the G-space kernels: G covariant = 2 pi m with m the
// symmetric index (fourier_space.hpp:66-81), |G|^2 through the reciprocal
// lattice rows b_i (which carry the 2 pi).
struct gspace {
  int n0, n1, n2;
  double b[9];
  __device__ void m_of(long p, int &m0, int &m1, int &m2) const {
    const int iz = (int)(p % n2);
    const long t = p / n2;
    const int iy = (int)(t % n1), ix = (int)(t / n1);
    m0 = ix >= (n0 + 1) / 2 ? ix - n0 : ix;
    m1 = iy >= (n1 + 1) / 2 ? iy - n1 : iy;
    m2 = iz >= (n2 + 1) / 2 ? iz - n2 : iz;
  }
  __device__ double g2(int m0, int m1, int m2) const {
    const double gx = m0 * b[0] + m1 * b[3] + m2 * b[6];
    const double gy = m0 * b[1] + m1 * b[4] + m2 * b[7];
    const double gz = m0 * b[2] + m1 * b[5] + m2 * b[8];
    return gx * gx + gy * gy + gz * gz;
  }
};

gspace geometry(const handle &h) {
  gspace g;
  g.n0 = h.n[0];
  g.n1 = h.n[1];
  g.n2 = h.n[2];
  for (int i = 0; i < 9; i++) g.b[i] = h.rlat[i];
  return g;
}

unsigned grid1(long total) {
  long b = (total + 255) / 256;
  return (unsigned)(b > 65536 ? 65536 : b);
}

// ---------------------------------------------------------------------------
// laplacian_add (laplacian.hpp:24-38)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void laplacian_add_kernel(gspace geo, const cplx *__restrict__ ff,
                                                            cplx *__restrict__ out, long ld,
                                                            int nst, long npoint, double factor,
                                                            double c0, double c1, double c2) {
  const long total = npoint * nst;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < total; i += (long)gridDim.x * 256) {
    const long p = i / nst;
    const int b = (int)(i % nst);
    int m0, m1, m2;
    geo.m_of(p, m0, m1, m2);
    const double two_pi = 6.283185307179586476925286766559;
    const double gdot = two_pi * (c0 * m0 + c1 * m1 + c2 * m2);
    const double lapl = factor * (-geo.g2(m0, m1, m2) + gdot);
    const cplx f = ff[p * ld + b];
    cplx o = out[p * ld + b];
    o.x += lapl * f.x;
    o.y += lapl * f.y;
    out[p * ld + b] = o;
  }
}

// ---------------------------------------------------------------------------
// per-band reductions: block (chunk, 64-band group); thread = band x lane
// ---------------------------------------------------------------------------
template <int NV, class F>
__device__ __forceinline__ void band_partials(long npoint, int nst, double *part, F term) {
  __shared__ double red[GS_LANES][GS_BT][NV];
  const int bb = threadIdx.x % GS_BT;
  const int lane = threadIdx.x / GS_BT;
  const int b = blockIdx.y * GS_BT + bb;
  const long p0 = (long)blockIdx.x * GS_CHUNK;
  const long p1 = p0 + GS_CHUNK < npoint ? p0 + GS_CHUNK : npoint;
  double acc[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) acc[v] = 0.0;
  if (b < nst)
    for (long p = p0 + lane; p < p1; p += GS_LANES) term(p, b, acc);
#pragma unroll
  for (int v = 0; v < NV; v++) red[lane][bb][v] = acc[v];
  __syncthreads();
  if (lane == 0 && b < nst) {
#pragma unroll
    for (int v = 0; v < NV; v++) {
      double s = red[0][bb][v];
      for (int l = 1; l < GS_LANES; l++) s += red[l][bb][v];
      part[((long)blockIdx.x * nst + b) * 8 + v] = s;
    }
  }
}

// pass 1: eig = sum conj(hphi) phi, norm = sum |phi|^2
__global__ __launch_bounds__(256) void eig_norm_partial(const cplx *__restrict__ hphi, long hld,
                                                        const cplx *__restrict__ phi, long ld,
                                                        int nst, long npoint, double *part) {
  band_partials<3>(npoint, nst, part, [&](long p, int b, double *a) {
    const cplx h = hphi[p * hld + b], f = phi[p * ld + b];
    a[0] += h.x * f.x + h.y * f.y;
    a[1] += h.x * f.y - h.y * f.x;
    a[2] += f.x * f.x + f.y * f.y;
  });
}

// pass 2: residual = hphi - evn phi (shift.hpp:19-31 with factor -1), and
// the preconditioner's sums 1/2|G|^2 |r|^2, |r|^2 (preconditioner.hpp:42)
__global__ __launch_bounds__(256) void residual_partial(gspace geo, const cplx *__restrict__ hphi,
                                                        long hld, const cplx *__restrict__ phi,
                                                        long ld, cplx *__restrict__ res, long rld,
                                                        const double *__restrict__ band, int nst,
                                                        long npoint, double *part) {
  band_partials<2>(npoint, nst, part, [&](long p, int b, double *a) {
    const double er = band[b * 16 + 3], ei = band[b * 16 + 4];  // evn
    const cplx h = hphi[p * hld + b], f = phi[p * ld + b];
    cplx r;
    r.x = h.x + (-1.0) * (er * f.x - ei * f.y);
    r.y = h.y + (-1.0) * (er * f.y + ei * f.x);
    res[p * rld + b] = r;
    int m0, m1, m2;
    geo.m_of(p, m0, m1, m2);
    const double lap = -0.5 * (-geo.g2(m0, m1, m2));
    const double r2 = r.x * r.x + r.y * r.y;
    a[0] += lap * r2;
    a[1] += r2;
  });
}

// pass 4: m0 = <r|r>, m1 = <phi|r>, m2 = <r|Hr>, m3 = <phi|Hr>
__global__ __launch_bounds__(256) void mm_partial(const cplx *__restrict__ phi, long ld,
                                                  const cplx *__restrict__ res,
                                                  const cplx *__restrict__ hres, long rld, int nst,
                                                  long npoint, double *part) {
  band_partials<8>(npoint, nst, part, [&](long p, int b, double *a) {
    const cplx f = phi[p * ld + b], r = res[p * rld + b], hr = hres[p * rld + b];
    a[0] += r.x * r.x + r.y * r.y;
    a[1] += r.x * r.y - r.y * r.x;
    a[2] += f.x * r.x + f.y * r.y;
    a[3] += f.x * r.y - f.y * r.x;
    a[4] += r.x * hr.x + r.y * hr.y;
    a[5] += r.x * hr.y - r.y * hr.x;
    a[6] += f.x * hr.x + f.y * hr.y;
    a[7] += f.x * hr.y - f.y * hr.x;
  });
}

// Sums of the NV partials of band b over all point chunks, by one 256-thread
// workgroup (strided over chunks, then a tree in LDS: a fixed order).
// Thread 0 gets the totals in `out`.
template <int NV>
__device__ __forceinline__ void band_sums(const double *part, long nchunk, int nst, int b,
                                          double *out) {
  __shared__ double red[256][NV];
  double acc[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) acc[v] = 0.0;
  for (long c = threadIdx.x; c < nchunk; c += 256)
#pragma unroll
    for (int v = 0; v < NV; v++) acc[v] += part[(c * nst + b) * 8 + v];
#pragma unroll
  for (int v = 0; v < NV; v++) red[threadIdx.x][v] = acc[v];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int v = 0; v < NV; v++) red[threadIdx.x][v] += red[threadIdx.x + w][v];
    __syncthreads();
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int v = 0; v < NV; v++) out[v] = red[0][v];
}

// band record: [0..1] eig, [2] norm, [3..4] eig/norm, [5] kinetic, [6] lambda
// (one workgroup of 256 threads per band)
__global__ __launch_bounds__(256) void eig_norm_final(const double *part, long nchunk, int nst,
                                                      double dV, double *band) {
  const int b = blockIdx.x;
  double t[3];
  band_sums<3>(part, nchunk, nst, b, t);
  if (threadIdx.x != 0) return;
  const double er = dV * t[0], ei = dV * t[1], nn = dV * t[2];
  band[b * 16 + 0] = er;
  band[b * 16 + 1] = ei;
  band[b * 16 + 2] = nn;
  band[b * 16 + 3] = er / nn;
  band[b * 16 + 4] = ei / nn;
}

__global__ __launch_bounds__(256) void kinetic_final(const double *part, long nchunk, int nst,
                                                     double *band) {
  const int b = blockIdx.x;
  double t[2];
  band_sums<2>(part, nchunk, nst, b, t);
  // overlap_diagonal_normalized: the volume element cancels
  if (threadIdx.x == 0) band[b * 16 + 5] = t[0] / t[1];
}

// steepest_descent.hpp:56-68
__global__ __launch_bounds__(256) void lambda_final(const double *part, long nchunk, int nst,
                                                    double dV, double *band) {
  const int b = blockIdx.x;
  double t[8];
  band_sums<8>(part, nchunk, nst, b, t);
  if (threadIdx.x != 0) return;
  // m0 = <r|r>, m1 = <phi|r>, m2 = <r|Hr>, m3 = <phi|Hr>, m4 = eig, m5 = norm
  double m[8];
  for (int v = 0; v < 8; v++) m[v] = dV * t[v];
  const double m4r = band[b * 16 + 0], m4i = band[b * 16 + 1], m5r = band[b * 16 + 2];
  // real parts of the complex products
  auto re = [](double ar, double ai, double br, double bi) { return ar * br - ai * bi; };
  const double ca = re(m[0], m[1], m[6], m[7]) - re(m[4], m[5], m[2], m[3]);
  const double cb = re(m5r, 0.0, m[4], m[5]) - re(m4r, m4i, m[0], m[1]);
  const double cc = re(m4r, m4i, m[2], m[3]) - re(m[6], m[7], m5r, 0.0);
  const double den = cb + sqrt(cb * cb - 4.0 * ca * cc);
  band[b * 16 + 6] = fabs(den) < 1e-15 ? 0.0 : 2.0 * cc / den;
}

// K of Teter, Payne and Allan (preconditioner.hpp:30-37)
__device__ double tpa_k(double x) {
  const double num = ((8.0 * x + 12.0) * x + 18.0) * x + 27.0;
  const double den = (((16.0 * x + 8.0) * x + 12.0) * x + 18.0) * x + 27.0;
  return num / den;
}

// pass 3: phi(G) *= K(1/2 |G|^2 / kinetic_b) (preconditioner.hpp:45-52).
// A wave takes one grid point at a time, its lanes the bands: the point's
// G vector is worked out once per point (the index divisions of a flat
// per-element loop cost more than the memory traffic).
constexpr int PRECOND_POINT_WAVES = 65536;  // waves of the grid-stride loop at most

__global__ __launch_bounds__(256) void precondition_kernel(gspace geo, cplx *__restrict__ phi,
                                                           long ld, int nst, long npoint,
                                                           const double *__restrict__ band) {
  const int lane = threadIdx.x & 63;
  const long nwaves = (long)gridDim.x * 4;
  for (long p = ((long)blockIdx.x * 256 + threadIdx.x) >> 6; p < npoint; p += nwaves) {
    int m0, m1, m2;
    geo.m_of(p, m0, m1, m2);
    const double g2 = geo.g2(m0, m1, m2);
    // four bands per lane in flight (a band past nst reads the lane's first
    // band and is not stored)
    for (int b0 = lane; b0 < nst; b0 += 256) {
      cplx f[4];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int b = b0 + 64 * j;
        f[j] = phi[p * ld + (b < nst ? b : b0)];
      }
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int b = b0 + 64 * j;
        if (b < nst) {
          const double k = tpa_k(-0.5 * (-g2) / band[b * 16 + 5]);
          phi[p * ld + b] = mk(f[j].x * k, f[j].y * k);
        }
      }
    }
  }
}

// pass 5 (steepest_descent.hpp:70-77) of the last step, which leaves hphi:
// phi += lambda r over the grid of the partial-sum passes (point chunk x
// 64-band block; the lanes of a point chunk stride its points)
__global__ __launch_bounds__(256) void sd_update_phi(cplx *__restrict__ phi, long ld,
                                                     const cplx *__restrict__ res, long rld,
                                                     int nst, long npoint,
                                                     const double *__restrict__ band) {
  const int bb = threadIdx.x % GS_BT;
  const int lane = threadIdx.x / GS_BT;
  const int b = blockIdx.y * GS_BT + bb;
  if (b >= nst) return;
  const long p0 = (long)blockIdx.x * GS_CHUNK;
  const long p1 = p0 + GS_CHUNK < npoint ? p0 + GS_CHUNK : npoint;
  const double lam = band[b * 16 + 6];
  for (long p = p0 + lane; p < p1; p += GS_LANES) {
    cplx f = phi[p * ld + b];
    const cplx r = res[p * rld + b];
    f.x += lam * r.x;
    f.y += lam * r.y;
    phi[p * ld + b] = f;
  }
}

// pass 5 of a step fused with pass 1 of the next (the partial sums of
// eig_norm_partial over the updated phi and hphi, in its order), for steps
// that update hphi
__global__ __launch_bounds__(256) void sd_update_eig_norm(cplx *__restrict__ phi, long ld,
                                                          const cplx *__restrict__ res,
                                                          cplx *__restrict__ hphi,
                                                          const cplx *__restrict__ hres, long rld,
                                                          int nst, long npoint,
                                                          const double *__restrict__ band,
                                                          double *part) {
  band_partials<3>(npoint, nst, part, [&](long p, int b, double *a) {
    const double lam = band[b * 16 + 6];
    cplx f = phi[p * ld + b];
    const cplx r = res[p * rld + b];
    cplx h = hphi[p * rld + b];
    const cplx hr = hres[p * rld + b];
    f.x += lam * r.x;
    f.y += lam * r.y;
    h.x += lam * hr.x;
    h.y += lam * hr.y;
    phi[p * ld + b] = f;
    hphi[p * rld + b] = h;
    a[0] += h.x * f.x + h.y * f.y;
    a[1] += h.x * f.y - h.y * f.x;
    a[2] += f.x * f.x + f.y * f.y;
  });
}

dim3 red_grid(const handle &h, int nst) {
  return dim3((unsigned)nchunks(h), (unsigned)((nst + GS_BT - 1) / GS_BT));
}

void check_launch() { INQAMD_HIP_CHECK(hipGetLastError()); }

// out[i] = sum over the partial slices of part (fixed order)
__global__ __launch_bounds__(256) void slice_sum_kernel(const cplx *__restrict__ part, int nslice,
                                                        long n, cplx *__restrict__ out) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    cplx s = part[i];
    for (int k = 1; k < nslice; k++) s = cadd(s, part[(long)k * n + i]);
    out[i] = s;
  }
}

// Row-major M[r][c] = alpha sum_p conj(A[p][r]) B[p][c] over all grid points
// (operations::overlap, overlap.hpp:26-60, as a BLAS gemm H(A) B).  The
// point sum is long and the output small, so it is split over IP_BATCHES
// point chunks (a strided-batched ZGEMM; one gemm over all points would keep
// only (nst/64)^2 workgroups busy) and the chunk results are summed in order.
void inner_products(handle &h, const cplx *A, long lda, const cplx *B, long ldb, int nst,
                    double alpha, cplx *out) {
  ensure_blas(h);
  gs_state &g = state(h);
  const long N = h.npoint;
  const int nb = (int)(N >= 256L * IP_BATCHES ? IP_BATCHES : 1);
  const long kc = N / nb, rem = N - kc * nb;
  const rocblas_stride sc = (rocblas_stride)nst * nst;
  const rocblas_double_complex a = {alpha, 0.0}, zero = {0.0, 0.0};
  // column-major C_b(c, r) = sum_{p in chunk b} B[p][c] conj(A[p][r]) is the
  // row-major M_b[r][c]
  GS_BLAS_CHECK(rocblas_zgemm_strided_batched(
      g.blas, rocblas_operation_none, rocblas_operation_conjugate_transpose, nst, nst,
      (rocblas_int)kc, &a, reinterpret_cast<const rocblas_double_complex *>(B), (int)ldb,
      (rocblas_stride)(kc * ldb), reinterpret_cast<const rocblas_double_complex *>(A), (int)lda,
      (rocblas_stride)(kc * lda), &zero, reinterpret_cast<rocblas_double_complex *>(g.ip_part),
      nst, sc, nb));
  int nslice = nb;
  if (rem > 0) {
    GS_BLAS_CHECK(rocblas_zgemm(
        g.blas, rocblas_operation_none, rocblas_operation_conjugate_transpose, nst, nst,
        (rocblas_int)rem, &a, reinterpret_cast<const rocblas_double_complex *>(B + kc * nb * ldb),
        (int)ldb, reinterpret_cast<const rocblas_double_complex *>(A + kc * nb * lda), (int)lda,
        &zero, reinterpret_cast<rocblas_double_complex *>(g.ip_part + (long)nb * sc), nst));
    nslice++;
  }
  hipLaunchKernelGGL(slice_sum_kernel, dim3(grid1(sc)), dim3(256), 0, h.stream,
                     (const cplx *)g.ip_part, nslice, (long)sc, out);
  check_launch();
}

// The K(1/2|G|^2/kinetic) scaling of operations::preconditioner for
// fourier_space orbitals (preconditioner.hpp:39-55), given the residual-type
// kinetic sums already in the band records.
void precondition_apply(handle &h, cplx *phi, long ld, int nst) {
  gs_state &g = state(h);
  hipLaunchKernelGGL(kinetic_final, dim3(nst), dim3(256), 0, h.stream,
                     (const double *)g.part, nchunks(h), nst, g.band);
  check_launch();
  const long waves = h.npoint < PRECOND_POINT_WAVES ? h.npoint : PRECOND_POINT_WAVES;
  hipLaunchKernelGGL(precondition_kernel, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0,
                     h.stream, geometry(h), phi, ld, nst, h.npoint, (const double *)g.band);
  check_launch();
}

// kinetic sums of phi alone (preconditioner.hpp:42) into the partial buffer
__global__ __launch_bounds__(256) void kinetic_partial(gspace geo, const cplx *__restrict__ phi,
                                                       long ld, int nst, long npoint,
                                                       double *part) {
  band_partials<2>(npoint, nst, part, [&](long p, int b, double *a) {
    const cplx f = phi[p * ld + b];
    int m0, m1, m2;
    geo.m_of(p, m0, m1, m2);
    const double lap = -0.5 * (-geo.g2(m0, m1, m2));
    const double r2 = f.x * f.x + f.y * f.y;
    a[0] += lap * r2;
    a[1] += r2;
  });
}

}  // namespace

// ks_hamiltonian::operator() on fourier_space orbitals (ks_hamiltonian.hpp:
// 156-176): to_real, the real-space operator, to_fourier.  Uses hn and hn2.
void hamiltonian_fourier(handle &h, const cplx *phi, long ld, cplx *hphi, long hld, int nst) {
  static const bool sweeps = [] {
    const char *e = getenv("INQAMD_GSPACE_SWEEPS");
    return !(e && atoi(e) == 0);
  }();
  if (h.three_sweep && sweeps) {
    hamiltonian_fourier_sweeps(h, phi, ld, hphi, hld, nst);
    return;
  }
  fft3d(h, phi, ld, h.hn, nst, nst, +1, 1.0 / (double)h.npoint);
  hamiltonian_apply(h, h.hn, nst, h.hn2, nst, nst, KA_CURRENT);
  fft3d(h, h.hn2, nst, hphi, hld, nst, -1, 1.0);
}

void preconditioner(handle &h, cplx *phi, long ld, int nst) {
  ensure_reductions(h);
  gs_state &g = state(h);
  hipLaunchKernelGGL(kinetic_partial, red_grid(h, nst), dim3(256), 0, h.stream, geometry(h),
                     (const cplx *)phi, ld, nst, h.npoint, g.part);
  check_launch();
  precondition_apply(h, phi, ld, nst);
}

// operations::orthogonalize (orthogonalize.hpp:25-32): S = dV Phi^H Phi,
// S = R R^H (R lower: zpotrf on the row-major S, cholesky.hpp:29-80),
// Phi <- Phi R^-H (invert_triangular + rotate, rotate.hpp:147-156).
void orthogonalize(handle &h, cplx *phi, long ld, int nst, double dV) {
  ensure_blas(h);
  gs_state &g = state(h);
  // S[i][j] = dV sum_p conj(phi_pi) phi_pj, row-major
  inner_products(h, phi, ld, phi, ld, nst, dV, g.S);
  g.hostS.resize((size_t)nst * nst);
  INQAMD_HIP_CHECK(hipMemcpyAsync(g.hostS.data(), g.S, (size_t)nst * nst * sizeof(cplx),
                                  hipMemcpyDeviceToHost, h.stream));
  INQAMD_HIP_CHECK(hipStreamSynchronize(h.stream));
  using C = std::complex<double>;
  const std::vector<C> &S = g.hostS;
  std::vector<C> R((size_t)nst * nst, 0.0);
  for (int j = 0; j < nst; j++) {
    double d = std::real(S[(size_t)j * nst + j]);
    for (int k = 0; k < j; k++) d -= std::norm(R[(size_t)j * nst + k]);
    if (!(d > 0.0)) throw error{ERR_ARGUMENT, "orthogonalize: overlap matrix is not positive definite (band " + std::to_string(j) + ")"};
    const double rjj = std::sqrt(d);
    R[(size_t)j * nst + j] = rjj;
    for (int i = j + 1; i < nst; i++) {
      C v = S[(size_t)i * nst + j];
      for (int k = 0; k < j; k++) v -= R[(size_t)i * nst + k] * std::conj(R[(size_t)j * nst + k]);
      R[(size_t)i * nst + j] = v / rjj;
    }
  }
  // R^-1 by forward substitution, then T = (R^-1)^H, upper triangular
  std::vector<C> Ri((size_t)nst * nst, 0.0);
  for (int c = 0; c < nst; c++) {
    Ri[(size_t)c * nst + c] = 1.0 / std::real(R[(size_t)c * nst + c]);
    for (int i = c + 1; i < nst; i++) {
      C v = 0.0;
      for (int k = c; k < i; k++) v -= R[(size_t)i * nst + k] * Ri[(size_t)k * nst + c];
      Ri[(size_t)i * nst + c] = v / std::real(R[(size_t)i * nst + i]);
    }
  }
  std::vector<C> &T = g.hostS;
  for (int i = 0; i < nst; i++)
    for (int j = 0; j < nst; j++) T[(size_t)i * nst + j] = std::conj(Ri[(size_t)j * nst + i]);
  INQAMD_HIP_CHECK(hipMemcpyAsync(g.S, T.data(), (size_t)nst * nst * sizeof(cplx),
                                  hipMemcpyHostToDevice, h.stream));
  // out[p][j] = sum_i phi[p][i] T[i][j]: column-major out^T = T_cm * Phi^T
  // with T_cm (ld nst) the row-major T read column-major
  const rocblas_double_complex one = {1.0, 0.0}, zero = {0.0, 0.0};
  GS_BLAS_CHECK(rocblas_zgemm(g.blas, rocblas_operation_none, rocblas_operation_none, nst,
                              (rocblas_int)h.npoint, nst, &one,
                              reinterpret_cast<const rocblas_double_complex *>(g.S), nst,
                              reinterpret_cast<const rocblas_double_complex *>(phi), (int)ld,
                              &zero, reinterpret_cast<rocblas_double_complex *>(h.hn), nst));
  copy(h, h.hn, nst, phi, ld, nst);
  // the host buffer is reused by the next call only after this copy ran
  INQAMD_HIP_CHECK(hipStreamSynchronize(h.stream));
}

// eigensolvers::steepest_descent (steepest_descent.hpp:23-88) on
// fourier_space orbitals with the G-space Hamiltonian.
void steepest_descent(handle &h, cplx *phi, long ld, int nst, bool precond) {
  ensure_workspace(h);
  gs_state &g = state(h);
  const long rl = nst;
  const long nchunk = nchunks(h);
  // fourier_space::volume_element (fourier_space.hpp:41-44)
  const double dV = h.volume / ((double)h.npoint * (double)h.npoint);
  const dim3 rg = red_grid(h, nst), fg(nst);
  hamiltonian_fourier(h, phi, ld, g.hphi, rl, nst);
  const int num_steps = 5;
  for (int istep = 0; istep < num_steps; istep++) {
    // pass 1 (after the first step, the previous step's update made its sums)
    if (istep == 0) {
      hipLaunchKernelGGL(eig_norm_partial, rg, dim3(256), 0, h.stream, (const cplx *)g.hphi,
                         rl, (const cplx *)phi, ld, nst, h.npoint, g.part);
      check_launch();
    }
    hipLaunchKernelGGL(eig_norm_final, fg, dim3(256), 0, h.stream, (const double *)g.part, nchunk,
                       nst, dV, g.band);
    check_launch();
    hipLaunchKernelGGL(residual_partial, rg, dim3(256), 0, h.stream, geometry(h),
                       (const cplx *)g.hphi, rl, (const cplx *)phi, ld, g.res, rl,
                       (const double *)g.band, nst, h.npoint, g.part);
    check_launch();
    if (precond) precondition_apply(h, g.res, rl, nst);
    hamiltonian_fourier(h, g.res, rl, g.hres, rl, nst);
    hipLaunchKernelGGL(mm_partial, rg, dim3(256), 0, h.stream, (const cplx *)phi, ld,
                       (const cplx *)g.res, (const cplx *)g.hres, rl, nst, h.npoint, g.part);
    check_launch();
    hipLaunchKernelGGL(lambda_final, fg, dim3(256), 0, h.stream, (const double *)g.part, nchunk,
                       nst, dV, g.band);
    check_launch();
    if (istep != num_steps - 1)
      hipLaunchKernelGGL(sd_update_eig_norm, rg, dim3(256), 0, h.stream, phi, ld,
                         (const cplx *)g.res, g.hphi, (const cplx *)g.hres, rl, nst, h.npoint,
                         (const double *)g.band, g.part);
    else
      hipLaunchKernelGGL(sd_update_phi, rg, dim3(256), 0, h.stream, phi, ld,
                         (const cplx *)g.res, rl, nst, h.npoint, (const double *)g.band);
    check_launch();
  }
  orthogonalize(h, phi, ld, nst, dV);
}

// ground_state::subspace_diagonalization (subspace_diagonalization.hpp:16-25)
// on real-space orbitals: H_sub = dV Phi^H (H Phi) (overlap.hpp:26-60), its
// eigenpairs (matrix::diagonalize, here rocSOLVER ZHEEVD), Phi <- Phi V
// (rotate, rotate.hpp:63).  Eigenvalues ascending into eig_host.
void subspace_diagonalization(handle &h, cplx *phi, long ld, int nst, double *eig_host) {
  ensure_workspace(h);
  ensure_blas(h);
  gs_state &g = state(h);
  hamiltonian_apply(h, phi, ld, g.hphi, nst, nst, KA_CURRENT);
  // H_sub[r][c] = dV sum_p conj(phi[p][r]) hphi[p][c], row-major (dV = V/N);
  // read as a column-major matrix that is conj(H_sub)
  inner_products(h, phi, ld, g.hphi, nst, nst, h.volume / (double)h.npoint, g.S);
  const rocblas_double_complex zero = {0.0, 0.0}, one = {1.0, 0.0};
  // eigenvectors w_k of conj(H_sub) are conj(v_k), v_k those of H_sub
  GS_BLAS_CHECK(rocsolver_zheevd(g.blas, rocblas_evect_original, rocblas_fill_upper, nst,
                                 reinterpret_cast<rocblas_double_complex *>(g.S), nst, g.eig,
                                 g.eig + h.max_bands, g.info));
  // out[p][k] = sum_i phi[p][i] v_k[i] = sum_i phi[p][i] conj(W(i, k)):
  // column-major out^T = W^H Phi^T
  GS_BLAS_CHECK(rocblas_zgemm(g.blas, rocblas_operation_conjugate_transpose,
                              rocblas_operation_none, nst, (rocblas_int)h.npoint, nst, &one,
                              reinterpret_cast<const rocblas_double_complex *>(g.S), nst,
                              reinterpret_cast<const rocblas_double_complex *>(phi), (int)ld,
                              &zero, reinterpret_cast<rocblas_double_complex *>(h.hn), nst));
  copy(h, h.hn, nst, phi, ld, nst);
  rocblas_int info = 0;
  INQAMD_HIP_CHECK(hipMemcpyAsync(eig_host, g.eig, (size_t)nst * sizeof(double),
                                  hipMemcpyDeviceToHost, h.stream));
  INQAMD_HIP_CHECK(hipMemcpyAsync(&info, g.info, sizeof(info), hipMemcpyDeviceToHost, h.stream));
  INQAMD_HIP_CHECK(hipStreamSynchronize(h.stream));
  if (info != 0)
    throw error{ERR_ARGUMENT, "subspace_diagonalization: eigensolver did not converge (info " +
                                  std::to_string(info) + ")"};
}

void gs_release(handle &h) {
  if (!h.gs) return;
  gs_state *g = static_cast<gs_state *>(h.gs);
  for (void *p : {(void *)g->hphi, (void *)g->res, (void *)g->hres, (void *)g->part,
                  (void *)g->band, (void *)g->S, (void *)g->eig, (void *)g->info,
                  (void *)g->ip_part})
    if (p) (void)hipFree(p);
  if (g->blas) (void)rocblas_destroy_handle(g->blas);
  delete g;
  h.gs = nullptr;
}

}  // namespace inqamd

using namespace inqamd;

extern "C" {

int inqamd_laplacian_add(inqamd_hamiltonian_t h, const void *ff, void *laplff, int nst, int ld,
                         double factor, const double gradcoeff_contra[3]) {
  if (!h) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    check_orbital(*h, ff, "ff");
    check_orbital(*h, laplff, "laplff");
    check_bands(*h, nst, ld, false);
    INQAMD_REQUIRE(gradcoeff_contra, "gradcoeff is null");
    hipLaunchKernelGGL(laplacian_add_kernel, dim3(grid1(h->npoint * nst)), dim3(256), 0,
                       h->stream, geometry(*h), (const cplx *)ff, (cplx *)laplff, (long)ld, nst,
                       h->npoint, factor, gradcoeff_contra[0], gradcoeff_contra[1],
                       gradcoeff_contra[2]);
    check_launch();
  });
}

int inqamd_preconditioner(inqamd_hamiltonian_t h, void *phi, int nst, int ld) {
  if (!h) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    check_orbital(*h, phi, "phi");
    check_bands(*h, nst, ld, true);
    preconditioner(*h, (cplx *)phi, ld, nst);
  });
}

int inqamd_orthogonalize(inqamd_hamiltonian_t h, void *phi, int nst, int ld, int fourier) {
  if (!h) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    check_orbital(*h, phi, "phi");
    check_bands(*h, nst, ld, true);
    // real_space::volume_element V/N, fourier_space::volume_element V/N^2
    // (real_space.hpp:150-152, fourier_space.hpp:41-44)
    const double dV = fourier ? h->volume / ((double)h->npoint * (double)h->npoint)
                              : h->volume / (double)h->npoint;
    orthogonalize(*h, (cplx *)phi, ld, nst, dV);
  });
}

int inqamd_subspace_diagonalization(inqamd_hamiltonian_t h, void *phi, int nst, int ld,
                                    double *eigenvalues) {
  if (!h) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    check_orbital(*h, phi, "phi");
    check_bands(*h, nst, ld, true);
    INQAMD_REQUIRE(eigenvalues, "eigenvalues is null");
    subspace_diagonalization(*h, (cplx *)phi, ld, nst, eigenvalues);
  });
}

int inqamd_steepest_descent(inqamd_hamiltonian_t h, void *phi, int nst, int ld, int precondition) {
  if (!h) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    check_orbital(*h, phi, "phi");
    check_bands(*h, nst, ld, true);
    steepest_descent(*h, (cplx *)phi, ld, nst, precondition != 0);
  });
}

}  // extern "C"
