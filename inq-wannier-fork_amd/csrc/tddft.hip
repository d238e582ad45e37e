// Interacting real-time TDDFT for static ions and one spin component: the
// density (observables::density), the periodic Hartree potential
// (solvers::poisson), LDA exchange-correlation (LDA_X + LDA_C_PZ, the
// functionals of options::theory::lda), the Kohn-Sham potential update
// (self_consistency::update_hamiltonian) and the ETRS step with its
// self-consistent second half (real_time::etrs, etrs.hpp:22-82).
//
// Fields are [N] doubles on the device.  Reductions (energies, the SCF
// density change) are per-block partial sums then a fixed-order sum, so they
// do not depend on scheduling.

#include <cmath>
#include <complex>
#include <vector>

#include "internal.hpp"
#include "inq_amd.h"

namespace inqamd {

void exponential_2_for_1(handle &h, std::complex<double> z1, std::complex<double> z2, int order,
                         cplx *phi, cplx *expphi, long ld, long eld, int nst, int slot);
void exponential_from(handle &h, std::complex<double> z, int order, const cplx *in, long ild,
                      cplx *out, long old, int nst, int slot);

namespace {

constexpr int TD_RED_BLOCKS = 1024;

struct td_state {
  double *rho = nullptr, *rho_old = nullptr, *vh = nullptr, *vxc = nullptr;  // [N]
  double *occ = nullptr;       // [max_bands]
  double *red = nullptr;       // [TD_RED_BLOCKS][4] partial sums + [4] totals
  cplx *cf = nullptr;          // [N] complex Poisson work
  cplx *save = nullptr;        // [N][max_bands] the ETRS half step
  cplx *g3 = nullptr;          // [N][3] density gradient / GGA flux (spectral)
};

td_state &state(handle &h) {
  if (!h.tddft) h.tddft = new td_state();
  return *static_cast<td_state *>(h.tddft);
}

void *dalloc(size_t bytes) {
  void *p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess)
    throw error{ERR_MEMORY, "device allocation of " + std::to_string(bytes) + " bytes failed"};
  return p;
}

void ensure_fields(handle &h) {
  td_state &t = state(h);
  if (t.rho) return;
  const size_t fb = (size_t)h.npoint * sizeof(double);
  t.rho = (double *)dalloc(fb);
  t.rho_old = (double *)dalloc(fb);
  t.vh = (double *)dalloc(fb);
  t.vxc = (double *)dalloc(fb);
  t.occ = (double *)dalloc((size_t)h.max_bands * sizeof(double));
  t.red = (double *)dalloc((size_t)(TD_RED_BLOCKS + 1) * 4 * sizeof(double));
  t.cf = (cplx *)dalloc((size_t)h.npoint * sizeof(cplx));
}

void ensure_gga(handle &h) {
  td_state &t = state(h);
  if (!t.g3) t.g3 = (cplx *)dalloc((size_t)h.npoint * 3 * sizeof(cplx));
}

void ensure_save(handle &h) {
  td_state &t = state(h);
  if (!t.save) t.save = (cplx *)dalloc((size_t)h.npoint * h.max_bands * sizeof(cplx));
}

unsigned grid1(long total) {
  long b = (total + 255) / 256;
  return (unsigned)(b > 65536 ? 65536 : b);
}

void check_launch() { INQAMD_HIP_CHECK(hipGetLastError()); }

// density.hpp:23-31: rho[p] = sum_b occ[b] |psi[p][b]|^2.  One wave per
// point: lanes take bands b = lane, lane + 64, ... (contiguous 1 KB per wave
// load), then a fixed butterfly over the lanes.
__global__ __launch_bounds__(256) void density_kernel(const cplx *__restrict__ psi, long ld,
                                                      int nst, const double *__restrict__ occ,
                                                      long npoint, double *__restrict__ rho) {
  const int lane = threadIdx.x & 63;
  const long wave = (blockIdx.x * 256L + threadIdx.x) >> 6;
  const long nwave = ((long)gridDim.x * 256) >> 6;
  for (long p = wave; p < npoint; p += nwave) {
    double s = 0.0;
    for (int b = lane; b < nst; b += 64) {
      const cplx f = psi[p * ld + b];
      s += occ[b] * (f.x * f.x + f.y * f.y);
    }
#pragma unroll
    for (int w = 32; w > 0; w >>= 1) s += __shfl_xor(s, w, 64);
    if (lane == 0) rho[p] = s;
  }
}

__global__ __launch_bounds__(256) void real_to_complex(const double *__restrict__ f, long n,
                                                       cplx *__restrict__ c) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256)
    c[i] = mk(f[i], 0.0);
}

__global__ __launch_bounds__(256) void complex_real_part(const cplx *__restrict__ c, long n,
                                                         double *__restrict__ f) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) f[i] = c[i].x;
}

// poisson.hpp:69-84 with the 3-D kernel (:29-35): V(G) *= (-4 pi / N)(-1/|G|^2),
// 0 where |G|^2 < 1e-6
__global__ __launch_bounds__(256) void poisson_kernel(cplx *__restrict__ c, int n0, int n1, int n2,
                                                      double b0, double b1, double b2, double b3,
                                                      double b4, double b5, double b6, double b7,
                                                      double b8) {
  const long N = (long)n0 * n1 * n2;
  const double scal = (-4.0 * 3.14159265358979323846) / (double)N;
  for (long p = blockIdx.x * 256L + threadIdx.x; p < N; p += (long)gridDim.x * 256) {
    const int iz = (int)(p % n2);
    const long t = p / n2;
    const int iy = (int)(t % n1), ix = (int)(t / n1);
    const int m0 = ix >= (n0 + 1) / 2 ? ix - n0 : ix;
    const int m1 = iy >= (n1 + 1) / 2 ? iy - n1 : iy;
    const int m2 = iz >= (n2 + 1) / 2 ? iz - n2 : iz;
    const double gx = m0 * b0 + m1 * b3 + m2 * b6;
    const double gy = m0 * b1 + m1 * b4 + m2 * b7;
    const double gz = m0 * b2 + m1 * b5 + m2 * b8;
    const double g2 = gx * gx + gy * gy + gz * gz;
    const double ker = g2 < 1e-6 ? 0.0 : -1.0 / g2;
    c[p] = cscale(scal * ker, c[p]);
  }
}

// Unpolarised Dirac exchange + Perdew-Zunger 1981 correlation (libxc's
// LDA_X and LDA_C_PZ): energy per particle e and v = d(n e)/dn; densities
// below 1e-15 give 0.
__device__ void lda_point(double n, double &e, double &v) {
  e = v = 0.0;
  if (n < 1e-15) return;
  const double pi = 3.14159265358979323846;
  const double cx = cbrt(3.0 / pi);
  const double n13 = cbrt(n);
  const double ex = -0.75 * cx * n13;
  const double vx = -cx * n13;
  const double rs = cbrt(3.0 / (4.0 * pi * n));
  double ec, vc;
  if (rs >= 1.0) {
    const double gam = -0.1423, b1 = 1.0529, b2 = 0.3334;
    const double sq = sqrt(rs);
    const double den = 1.0 + b1 * sq + b2 * rs;
    ec = gam / den;
    vc = ec * (1.0 + (7.0 / 6.0) * b1 * sq + (4.0 / 3.0) * b2 * rs) / den;
  } else {
    const double A = 0.0311, B = -0.048, C = 0.0020, D = -0.0116;
    const double lr = log(rs);
    ec = A * lr + B + C * rs * lr + D * rs;
    vc = A * lr + (B - A / 3.0) + (2.0 / 3.0) * C * rs * lr + (2.0 * D - C) / 3.0 * rs;
  }
  e = ex + ec;
  v = vx + vc;
}

// PBE (Perdew, Burke, Ernzerhof 1996), one spin component: energy per
// particle e, vn = d(n e)/dn and vs = d(n e)/dsigma (sigma = |grad n|^2);
// the uniform-gas correlation is Perdew-Wang 1992 with libxc's LDA_C_PW_MOD
// parameters, as GGA_C_PBE uses it.
__device__ void pbe_exchange(double n, double sigma, double &e, double &vn, double &vs) {
  const double kappa = 0.8040, mu = 0.2195149727645171;
  const double ax = -0.75 * cbrt(3.0 / M_PI);
  const double n13 = cbrt(n);
  const double n83 = n13 * n13 * n * n;                                // n^(8/3)
  const double c = 4.0 * pow(3.0 * M_PI * M_PI, 2.0 / 3.0);
  const double x = sigma / (c * n83);                                  // s^2
  const double den = 1.0 + mu * x / kappa;
  const double F = 1.0 + kappa - kappa / den, dF = mu / (den * den);
  e = ax * n13 * F;
  vn = ax * n13 * ((4.0 / 3.0) * F - (8.0 / 3.0) * x * dF);
  vs = ax * n13 * n * dF / (c * n83);
}

__device__ void pbe_correlation(double n, double sigma, double &e, double &vn, double &vs) {
  const double beta = 0.06672455060314922, gamma = (1.0 - M_LN2) / (M_PI * M_PI);
  const double A0 = 0.0310907, a1 = 0.21370, b1 = 7.5957, b2 = 3.5876, b3 = 1.6382,
               b4 = 0.49294;
  const double rs = cbrt(3.0 / (4.0 * M_PI * n)), sq = sqrt(rs);
  const double q0 = -2.0 * A0 * (1.0 + a1 * rs);
  const double q1 = 2.0 * A0 * (b1 * sq + b2 * rs + b3 * rs * sq + b4 * rs * rs);
  const double q1p = A0 * (b1 / sq + 2.0 * b2 + 3.0 * b3 * sq + 4.0 * b4 * rs);
  const double lg = log(1.0 + 1.0 / q1);
  const double ec = q0 * lg;
  const double dec = (-2.0 * A0 * a1 * lg - q0 * q1p / (q1 * q1 + q1)) * (-rs / (3.0 * n));
  const double kf = cbrt(3.0 * M_PI * M_PI * n);
  const double ty = M_PI / (16.0 * kf * n * n);  // t^2 / sigma
  const double y = sigma * ty;
  const double ex = exp(-ec / gamma);
  const double A = (beta / gamma) / (ex - 1.0);
  const double dA = (beta / (gamma * gamma)) * ex / ((ex - 1.0) * (ex - 1.0));
  const double P = 1.0 + A * y, D = 1.0 + A * y + A * A * y * y;
  const double N = (beta / gamma) * y * P / D;
  const double H = gamma * log(1.0 + N);
  const double dNy = (beta / gamma) * ((P + A * y) * D - y * P * (A + 2.0 * A * A * y)) / (D * D);
  const double dNA = (beta / gamma) * y * (y * D - P * (y + 2.0 * A * y * y)) / (D * D);
  const double g = gamma / (1.0 + N);
  e = ec + H;
  vn = e + n * (dec + g * dNA * dA * dec - g * dNy * (7.0 / 3.0) * y / n);
  vs = n * g * dNy * ty;
}

// i G_d f(G) / N into g3[p][d] (operations::gradient in Fourier space)
__global__ __launch_bounds__(256) void grad_spec_kernel(const cplx *__restrict__ f, int n0, int n1,
                                                        int n2, double b0, double b1, double b2,
                                                        double b3, double b4, double b5,
                                                        double b6, double b7, double b8,
                                                        double inv_n, cplx *__restrict__ g3) {
  const long N = (long)n0 * n1 * n2;
  for (long p = blockIdx.x * 256L + threadIdx.x; p < N; p += (long)gridDim.x * 256) {
    const int iz = (int)(p % n2);
    const long t = p / n2;
    const int iy = (int)(t % n1), ix = (int)(t / n1);
    const int m0 = ix >= (n0 + 1) / 2 ? ix - n0 : ix, m1 = iy >= (n1 + 1) / 2 ? iy - n1 : iy,
              m2 = iz >= (n2 + 1) / 2 ? iz - n2 : iz;
    const double g[3] = {m0 * b0 + m1 * b3 + m2 * b6, m0 * b1 + m1 * b4 + m2 * b7,
                         m0 * b2 + m1 * b5 + m2 * b8};
    const cplx v = f[p];
    for (int d = 0; d < 3; d++) g3[p * 3 + d] = mk(-g[d] * v.y * inv_n, g[d] * v.x * inv_n);
  }
}

// sum_d i G_d g3[p][d] / N into f (the divergence in Fourier space)
__global__ __launch_bounds__(256) void div_spec_kernel(const cplx *__restrict__ g3, int n0, int n1,
                                                       int n2, double b0, double b1, double b2,
                                                       double b3, double b4, double b5, double b6,
                                                       double b7, double b8, double inv_n,
                                                       cplx *__restrict__ f) {
  const long N = (long)n0 * n1 * n2;
  for (long p = blockIdx.x * 256L + threadIdx.x; p < N; p += (long)gridDim.x * 256) {
    const int iz = (int)(p % n2);
    const long t = p / n2;
    const int iy = (int)(t % n1), ix = (int)(t / n1);
    const int m0 = ix >= (n0 + 1) / 2 ? ix - n0 : ix, m1 = iy >= (n1 + 1) / 2 ? iy - n1 : iy,
              m2 = iz >= (n2 + 1) / 2 ? iz - n2 : iz;
    const double g[3] = {m0 * b0 + m1 * b3 + m2 * b6, m0 * b1 + m1 * b4 + m2 * b7,
                         m0 * b2 + m1 * b5 + m2 * b8};
    double re = 0.0, im = 0.0;
    for (int d = 0; d < 3; d++) {
      const cplx v = g3[p * 3 + d];
      re -= g[d] * v.y;
      im += g[d] * v.x;
    }
    f[p] = mk(re * inv_n, im * inv_n);
  }
}

// Block partial sums of up to 4 values per point into red[block][4].
template <class F>
__device__ __forceinline__ void block_sums4(long n, double *red, F term) {
  __shared__ double sh[256][4];
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) term(i, a);
  for (int v = 0; v < 4; v++) sh[threadIdx.x][v] = a[v];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int v = 0; v < 4; v++) sh[threadIdx.x][v] += sh[threadIdx.x + w][v];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int v = 0; v < 4; v++) red[blockIdx.x * 4 + v] = sh[0][v];
}

// Totals of the block partials: one wave per value, lanes strided over the
// blocks, then a fixed butterfly.
__global__ __launch_bounds__(256) void final_sums4(double *red, int nblocks) {
  const int v = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double s = 0.0;
  for (int b = lane; b < nblocks; b += 64) s += red[b * 4 + v];
#pragma unroll
  for (int w = 32; w > 0; w >>= 1) s += __shfl_xor(s, w, 64);
  if (lane == 0) red[(long)TD_RED_BLOCKS * 4 + v] = s;
}

// xc_term.hpp:120-142 for one spin and no core density: n = max(rho, 0),
// vxc[p] = v(n); sums of e(n) n and rho vxc
__global__ __launch_bounds__(256) void lda_kernel(const double *__restrict__ rho, long npoint,
                                                  double *__restrict__ vxc, double *red) {
  block_sums4(npoint, red, [&](long p, double *a) {
    const double r = rho[p];
    const double n = r < 0.0 ? 0.0 : r;
    double e, v;
    lda_point(n, e, v);
    vxc[p] = v;
    a[0] += e * n;
    a[1] += r * v;
  });
}

// xc_term.hpp:267-292 for a GGA: n = max(rho, 0), sigma from the real part
// of the gradient; vxc = vn, the flux -2 vs grad n back into g3; sums of
// e n.  which: 1 = exchange, 2 = correlation, 3 = both.
__global__ __launch_bounds__(256) void gga_point_kernel(const double *__restrict__ rho,
                                                        long npoint, int which,
                                                        double *__restrict__ vxc,
                                                        cplx *__restrict__ g3, double *red) {
  block_sums4(npoint, red, [&](long p, double *a) {
    const double r = rho[p];
    const double n = r < 0.0 ? 0.0 : r;
    const double gx = g3[p * 3].x, gy = g3[p * 3 + 1].x, gz = g3[p * 3 + 2].x;
    const double sigma = gx * gx + gy * gy + gz * gz;
    double e = 0.0, vn = 0.0, vs = 0.0;
    if (n >= 1e-15) {
      double e1, v1, s1;
      if (which & 1) {
        pbe_exchange(n, sigma, e1, v1, s1);
        e += e1, vn += v1, vs += s1;
      }
      if (which & 2) {
        pbe_correlation(n, sigma, e1, v1, s1);
        e += e1, vn += v1, vs += s1;
      }
    }
    vxc[p] = vn;
    g3[p * 3] = mk(-2.0 * vs * gx, 0.0);
    g3[p * 3 + 1] = mk(-2.0 * vs * gy, 0.0);
    g3[p * 3 + 2] = mk(-2.0 * vs * gz, 0.0);
    a[0] += e * n;
  });
}

// vxc += Re f (the divergence term); sums of rho vxc
__global__ __launch_bounds__(256) void gga_finish_kernel(const cplx *__restrict__ f,
                                                         const double *__restrict__ rho,
                                                         long npoint, double *__restrict__ vxc,
                                                         double *red) {
  block_sums4(npoint, red, [&](long p, double *a) {
    const double v = vxc[p] + f[p].x;
    vxc[p] = v;
    a[0] += rho[p] * v;
  });
}

// LDA exchange alone (LDA_X): sums of e n and rho v
__global__ __launch_bounds__(256) void lda_x_kernel(const double *__restrict__ rho, long npoint,
                                                    double *__restrict__ vxc, double *red) {
  block_sums4(npoint, red, [&](long p, double *a) {
    const double r = rho[p];
    const double n = r < 0.0 ? 0.0 : r;
    const double n13 = cbrt(n), cx = cbrt(3.0 / M_PI);
    const double e = n < 1e-15 ? 0.0 : -0.75 * cx * n13, v = n < 1e-15 ? 0.0 : -cx * n13;
    vxc[p] = v;
    a[0] += e * n;
    a[1] += r * v;
  });
}

// self_consistency.hpp:96-145: V_KS = vion (+ V_H) (+ V_xc) and the sums of
// rho vion and rho V_H
__global__ __launch_bounds__(256) void ks_kernel(const double *__restrict__ vion,
                                                 const double *__restrict__ rho,
                                                 const double *__restrict__ vh,
                                                 const double *__restrict__ vxc, long npoint,
                                                 double *__restrict__ vks, double *red) {
  block_sums4(npoint, red, [&](long p, double *a) {
    double v = vion[p];
    a[0] += rho[p] * vion[p];
    if (vh) {
      v += vh[p];
      a[1] += rho[p] * vh[p];
    }
    if (vxc) v += vxc[p];
    vks[p] = v;
  });
}

// operations::integral_sum_absdiff (integral.hpp:113-127) without the dV
__global__ __launch_bounds__(256) void absdiff_kernel(const double *__restrict__ a,
                                                      const double *__restrict__ b, long n,
                                                      double *red) {
  block_sums4(n, red, [&](long i, double *s) { s[0] += fabs(a[i] - b[i]); });
}

// observables::dipole of a density (dipole.hpp:24-40): sums of r_cart rho,
// r from the symmetric-range contravariant coordinates (real_space.hpp:76-95)
__global__ __launch_bounds__(256) void dipole_kernel(const double *__restrict__ rho, int n0, int n1,
                                                     int n2, double a0, double a1, double a2,
                                                     double a3, double a4, double a5, double a6,
                                                     double a7, double a8, double *red) {
  const long N = (long)n0 * n1 * n2;
  block_sums4(N, red, [&](long p, double *s) {
    const int iz = (int)(p % n2);
    const long t = p / n2;
    const int iy = (int)(t % n1), ix = (int)(t / n1);
    const double c0 = (double)(ix >= (n0 + 1) / 2 ? ix - n0 : ix) / n0;
    const double c1 = (double)(iy >= (n1 + 1) / 2 ? iy - n1 : iy) / n1;
    const double c2 = (double)(iz >= (n2 + 1) / 2 ? iz - n2 : iz) / n2;
    const double v = rho[p];
    s[0] += ((c0 * a0 + c1 * a3) + c2 * a6) * v;
    s[1] += ((c0 * a1 + c1 * a4) + c2 * a7) * v;
    s[2] += ((c0 * a2 + c1 * a5) + c2 * a8) * v;
  });
}

unsigned red_blocks(long n) {
  long b = (n + 255) / 256;
  return (unsigned)(b > TD_RED_BLOCKS ? TD_RED_BLOCKS : b);
}

// run a block_sums4 kernel, then its fixed-order total; copy the 4 totals to host
template <typename K, typename... A>
void reduce4(handle &h, long n, double out[4], K kernel, A... args) {
  td_state &t = state(h);
  const unsigned nb = red_blocks(n);
  hipLaunchKernelGGL(kernel, dim3(nb), dim3(256), 0, h.stream, args..., t.red);
  check_launch();
  hipLaunchKernelGGL(final_sums4, dim3(1), dim3(256), 0, h.stream, t.red, (int)nb);
  check_launch();
  INQAMD_HIP_CHECK(hipMemcpyAsync(out, t.red + (long)TD_RED_BLOCKS * 4, 4 * sizeof(double),
                                  hipMemcpyDeviceToHost, h.stream));
  INQAMD_HIP_CHECK(hipStreamSynchronize(h.stream));
}

}  // namespace

void density(handle &h, const cplx *psi, long ld, int nst, const double *occ_dev, double *rho) {
  hipLaunchKernelGGL(density_kernel, dim3(grid1(h.npoint * 64)), dim3(256), 0, h.stream, psi, ld,
                     nst, occ_dev, h.npoint, rho);
  check_launch();
}

void poisson(handle &h, const double *rho, double *vh) {
  ensure_fields(h);
  td_state &t = state(h);
  hipLaunchKernelGGL(real_to_complex, dim3(grid1(h.npoint)), dim3(256), 0, h.stream, rho,
                     h.npoint, t.cf);
  check_launch();
  fft3d(h, t.cf, 1, t.cf, 1, 1, -1, 1.0);
  const double *b = h.rlat;
  hipLaunchKernelGGL(poisson_kernel, dim3(grid1(h.npoint)), dim3(256), 0, h.stream, t.cf, h.n[0],
                     h.n[1], h.n[2], b[0], b[1], b[2], b[3], b[4], b[5], b[6], b[7], b[8]);
  check_launch();
  fft3d(h, t.cf, 1, t.cf, 1, 1, +1, 1.0);  // to_real without normalisation
  hipLaunchKernelGGL(complex_real_part, dim3(grid1(h.npoint)), dim3(256), 0, h.stream,
                     (const cplx *)t.cf, h.npoint, vh);
  check_launch();
}

void dipole(handle &h, const double *rho, double out[3]) {
  ensure_fields(h);
  const double dV = h.volume / (double)h.npoint;
  const double *a = h.lattice;
  double s[4];
  reduce4(h, h.npoint, s, dipole_kernel, rho, h.n[0], h.n[1], h.n[2], a[0], a[1], a[2], a[3], a[4],
          a[5], a[6], a[7], a[8]);
  for (int d = 0; d < 3; d++) out[d] = dV * s[d];
}

void lda_xc(handle &h, const double *rho, double *vxc, double *exc, double *nvxc) {
  ensure_fields(h);
  const double dV = h.volume / (double)h.npoint;
  double s[4];
  reduce4(h, h.npoint, s, lda_kernel, rho, h.npoint, vxc);
  *exc = dV * s[0];
  *nvxc = dV * s[1];
}

// hamiltonian::xc_term::evaluate_functional for a GGA (xc_term.hpp:267-300):
// grad rho = Re to_real(i G to_fourier rho), the point forms, then
// vxc = vn + div(-2 vs grad rho) by the same spectral derivative
void gga_xc(handle &h, const double *rho, int which, double *vxc, double *exc, double *nvxc) {
  ensure_fields(h);
  ensure_gga(h);
  td_state &t = state(h);
  const double dV = h.volume / (double)h.npoint, inv_n = 1.0 / (double)h.npoint;
  const double *b = h.rlat;
  hipLaunchKernelGGL(real_to_complex, dim3(grid1(h.npoint)), dim3(256), 0, h.stream, rho,
                     h.npoint, t.cf);
  check_launch();
  fft3d(h, t.cf, 1, t.cf, 1, 1, -1, 1.0);
  hipLaunchKernelGGL(grad_spec_kernel, dim3(grid1(h.npoint)), dim3(256), 0, h.stream,
                     (const cplx *)t.cf, h.n[0], h.n[1], h.n[2], b[0], b[1], b[2], b[3], b[4],
                     b[5], b[6], b[7], b[8], inv_n, t.g3);
  check_launch();
  fft3d(h, t.g3, 3, t.g3, 3, 3, +1, 1.0);
  double s[4];
  reduce4(h, h.npoint, s, gga_point_kernel, rho, h.npoint, which, vxc, t.g3);
  *exc = dV * s[0];
  fft3d(h, t.g3, 3, t.g3, 3, 3, -1, 1.0);
  hipLaunchKernelGGL(div_spec_kernel, dim3(grid1(h.npoint)), dim3(256), 0, h.stream,
                     (const cplx *)t.g3, h.n[0], h.n[1], h.n[2], b[0], b[1], b[2], b[3], b[4],
                     b[5], b[6], b[7], b[8], inv_n, t.cf);
  check_launch();
  fft3d(h, t.cf, 1, t.cf, 1, 1, +1, 1.0);
  reduce4(h, h.npoint, s, gga_finish_kernel, (const cplx *)t.cf, rho, h.npoint, vxc);
  *nvxc = dV * s[0];
}

// One functional: 1 = LDA_X, 2 = LDA_X + LDA_C_PZ, 3 = GGA_X_PBE,
// 4 = GGA_C_PBE, 5 = both PBE parts
void xc_functional(handle &h, const double *rho, int kind, double *vxc, double *exc,
                   double *nvxc) {
  if (kind == 2) {
    lda_xc(h, rho, vxc, exc, nvxc);
  } else if (kind == 1) {
    ensure_fields(h);
    double s[4];
    reduce4(h, h.npoint, s, lda_x_kernel, rho, h.npoint, vxc);
    const double dV = h.volume / (double)h.npoint;
    *exc = dV * s[0];
    *nvxc = dV * s[1];
  } else {
    gga_xc(h, rho, kind - 2, vxc, exc, nvxc);
  }
}

// update_hamiltonian: the handle's V <- vion + V_H[rho] + V_xc[rho];
// energies {external, hartree, exc, nvxc}; xc: 0 none, 1 LDA
// (options::theory::lda), 2 PBE (the theory's default)
void ks_potential(handle &h, const double *vion, const double *rho, bool hartree, int xc,
                  double energies[4]) {
  ensure_fields(h);
  td_state &t = state(h);
  const double dV = h.volume / (double)h.npoint;
  energies[2] = energies[3] = 0.0;
  if (hartree) poisson(h, rho, t.vh);
  if (xc) xc_functional(h, rho, xc == 2 ? 5 : 2, t.vxc, &energies[2], &energies[3]);
  double s[4];
  reduce4(h, h.npoint, s, ks_kernel, vion, rho, hartree ? (const double *)t.vh : nullptr,
          xc ? (const double *)t.vxc : nullptr, h.npoint, h.V);
  h.generation++;
  energies[0] = dV * s[0];
  energies[1] = hartree ? 0.5 * dV * s[1] : 0.0;
}

void upload_occ(handle &h, const double *occ_host, int nst) {
  ensure_fields(h);
  INQAMD_HIP_CHECK(hipMemcpyAsync(state(h).occ, occ_host, (size_t)nst * sizeof(double),
                                  hipMemcpyHostToDevice, h.stream));
}

// real_time::etrs with the self-consistent second half (etrs.hpp:22-82),
// static ions, no induced vector potential.  The handle's V is V_KS of
// H(t) on entry and of H(t+dt) on return.
int etrs_scf(handle &h, double dt, const double A_t[3], const double A_tdt[3], int order,
             cplx *psi, long ld, int nst, const double *vion, bool hartree, int xc) {
  ensure_fields(h);
  ensure_save(h);
  td_state &t = state(h);
  const int nscf = 5;
  const double scf_threshold = 5e-5;
  const double dV = h.volume / (double)h.npoint;
  double kt[3], ktdt[3];
  for (int d = 0; d < 3; d++) {
    kt[d] = h.k_cov[d] + A_t[d];
    ktdt[d] = h.k_cov[d] + A_tdt[d];
  }
  set_slot(h, KA_T, kt);
  set_slot(h, KA_TDT, ktdt);
  // :33-38 save = T(i dt/2, H(t)) psi, psi <- T(i dt, H(t)) psi
  // (the saved half step is compact: leading dimension nst)
  exponential_2_for_1(h, std::complex<double>(0.0, dt / 2.0), std::complex<double>(0.0, dt), order,
                      psi, t.save, ld, nst, nst, KA_T);
  // :44 the density of the full step, :53 H(t + dt)
  density(h, psi, ld, nst, t.occ, t.rho);
  double en[4];
  ks_potential(h, vion, t.rho, hartree, xc, en);
  for (int d = 0; d < 3; d++) h.A_cov[d] = A_tdt[d];
  update_current_slot(h);
  // :56-58 restore the half step; :61-80 the second half, self-consistently
  // (each pass psi = T(i dt/2, H(t+dt)) save, written from save directly)
  int iscf = 0;
  for (; iscf < nscf; iscf++) {
    exponential_from(h, std::complex<double>(0.0, dt / 2.0), order, t.save, nst, psi, ld, nst,
                     KA_TDT);
    std::swap(t.rho, t.rho_old);
    density(h, psi, ld, nst, t.occ, t.rho);
    double s[4];
    reduce4(h, h.npoint, s, absdiff_kernel, (const double *)t.rho_old, (const double *)t.rho,
            h.npoint);
    const bool done = dV * s[0] < scf_threshold || iscf == nscf - 1;
    ks_potential(h, vion, t.rho, hartree, xc, en);
    if (done) break;
  }
  return iscf + 1;
}

void td_release(handle &h) {
  if (!h.tddft) return;
  td_state *t = static_cast<td_state *>(h.tddft);
  for (void *p : {(void *)t->rho, (void *)t->rho_old, (void *)t->vh, (void *)t->vxc,
                  (void *)t->occ, (void *)t->red, (void *)t->cf, (void *)t->save,
                  (void *)t->g3})
    if (p) (void)hipFree(p);
  delete t;
  h.tddft = nullptr;
}

}  // namespace inqamd

using namespace inqamd;

extern "C" {

int inqamd_density(inqamd_hamiltonian_t h, const void *psi, int nst, int ld, const double *occ,
                   double *rho) {
  if (!h) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    check_orbital(*h, psi, "psi");
    check_bands(*h, nst, ld, true);
    INQAMD_REQUIRE(occ && rho, "occ and rho are required");
    upload_occ(*h, occ, nst);
    density(*h, (const cplx *)psi, ld, nst, state(*h).occ, rho);
  });
}

int inqamd_poisson(inqamd_hamiltonian_t h, const double *rho, double *vh) {
  if (!h) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    INQAMD_REQUIRE(rho && vh, "rho and vh are required");
    poisson(*h, rho, vh);
  });
}

int inqamd_dipole(inqamd_hamiltonian_t h, const double *rho, double dipole_cart[3]) {
  if (!h) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    INQAMD_REQUIRE(rho && dipole_cart, "rho and dipole are required");
    dipole(*h, rho, dipole_cart);
  });
}

int inqamd_lda_xc(inqamd_hamiltonian_t h, const double *rho, double *vxc, double *exc,
                  double *nvxc) {
  if (!h) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    INQAMD_REQUIRE(rho && vxc && exc && nvxc, "rho, vxc, exc and nvxc are required");
    lda_xc(*h, rho, vxc, exc, nvxc);
  });
}

int inqamd_xc(inqamd_hamiltonian_t h, const double *rho, int kind, double *vxc, double *exc,
              double *nvxc) {
  if (!h) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    INQAMD_REQUIRE(rho && vxc && exc && nvxc, "rho, vxc, exc and nvxc are required");
    INQAMD_REQUIRE(kind >= 1 && kind <= 5, "kind must be 1 (LDA_X), 2 (LDA), 3 (PBE_X), 4 "
                                           "(PBE_C) or 5 (PBE)");
    xc_functional(*h, rho, kind, vxc, exc, nvxc);
  });
}

int inqamd_ks_potential(inqamd_hamiltonian_t h, const double *vion, const double *rho,
                        int hartree, int xc, double energies[4]) {
  if (!h) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    INQAMD_REQUIRE(vion && rho && energies, "vion, rho and energies are required");
    INQAMD_REQUIRE(xc >= 0 && xc <= 2, "xc must be 0 (none), 1 (LDA) or 2 (PBE)");
    ks_potential(*h, vion, rho, hartree != 0, xc, energies);
  });
}

int inqamd_etrs_scf_step(inqamd_hamiltonian_t h, double dt, const double A_t_cov[3],
                         const double A_tdt_cov[3], int order, void *psi, int nst, int ld,
                         const double *vion, const double *occ, int hartree, int xc, int *nscf) {
  if (!h) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    check_orbital(*h, psi, "psi");
    check_bands(*h, nst, ld, true);
    INQAMD_REQUIRE(A_t_cov && A_tdt_cov, "vector potentials are required");
    INQAMD_REQUIRE(vion && occ, "vion and occ are required");
    INQAMD_REQUIRE(order >= 1, "order must be >= 1");
    upload_occ(*h, occ, nst);
    INQAMD_REQUIRE(xc >= 0 && xc <= 2, "xc must be 0 (none), 1 (LDA) or 2 (PBE)");
    const int n = etrs_scf(*h, dt, A_t_cov, A_tdt_cov, order, (cplx *)psi, ld, nst, vion,
                           hartree != 0, xc);
    if (nscf) *nscf = n;
  });
}

}  // extern "C"
