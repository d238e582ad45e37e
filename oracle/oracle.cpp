/*
 * oracle.cpp -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * A plain, serial-per-band FP64 restatement of INQ's CPU branches for the
 * Hpsi / Taylor / ETRS path.  Every function cites the reference lines it
 * follows.  Its FFT is a recursive decimation-in-time mixed-radix transform
 * with an O(p^2) combination step for any prime factor p, and twiddles
 * evaluated in long double -- deliberately a different algorithm from the
 * HIP library's Stockham passes so that the two do not share mistakes.
 */
#include "oracle.h"

#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>
#include <cstdint>
#include <algorithm>
#include <memory>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using cplx = std::complex<double>;

/* ---------------------------------------------------------------------------
 * Cell metric (src/systems/cell.hpp:36-53, 146-224)
 * ------------------------------------------------------------------------ */
struct vec3 {
  double v[3];
  double &operator[](int i) { return v[i]; }
  double operator[](int i) const { return v[i]; }
};

static vec3 operator+(vec3 a, vec3 b) { return {{a[0] + b[0], a[1] + b[1], a[2] + b[2]}}; }
static vec3 operator-(vec3 a, vec3 b) { return {{a[0] - b[0], a[1] - b[1], a[2] - b[2]}}; }
static vec3 operator*(double s, vec3 a) { return {{s * a[0], s * a[1], s * a[2]}}; }
static double dot(vec3 a, vec3 b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static vec3 cross(vec3 a, vec3 b) {
  return {{a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]}};
}

struct cell {
  vec3 lat[3];
  vec3 rlat[3];
  double volume;

  explicit cell(const double *lattice) {
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) lat[i][j] = lattice[3 * i + j];
    volume = dot(lat[0], cross(lat[1], lat[2]));
    rlat[0] = (2.0 * M_PI / volume) * cross(lat[1], lat[2]);
    rlat[1] = (2.0 * M_PI / volume) * cross(lat[2], lat[0]);
    rlat[2] = (2.0 * M_PI / volume) * cross(lat[0], lat[1]);
  }
  /* cell.hpp:215-217 */
  vec3 contra_to_cart(vec3 v) const { return (v[0] * lat[0] + v[1] * lat[1]) + v[2] * lat[2]; }
  /* cell.hpp:220-222 */
  vec3 cov_to_cart(vec3 v) const {
    vec3 s = (v[0] * rlat[0] + v[1] * rlat[1]) + v[2] * rlat[2];
    return {{s[0] / (2.0 * M_PI), s[1] / (2.0 * M_PI), s[2] / (2.0 * M_PI)}};
  }
  /* cell.hpp:205-207 */
  vec3 cart_to_contra(vec3 w) const {
    return {{dot(rlat[0], w) / (2.0 * M_PI), dot(rlat[1], w) / (2.0 * M_PI),
             dot(rlat[2], w) / (2.0 * M_PI)}};
  }
  /* cell.hpp:185-187 */
  vec3 cart_to_cov(vec3 w) const { return {{dot(lat[0], w), dot(lat[1], w), dot(lat[2], w)}}; }
  double norm_cov(vec3 v) const { vec3 c = cov_to_cart(v); return dot(c, c); }
  vec3 cov_to_contra(vec3 v) const { return cart_to_contra(cov_to_cart(v)); }
};

/* grid.hpp:111-117 */
static int symmetric(int i, int n) { return i >= (n + 1) / 2 ? i - n : i; }

/* ---------------------------------------------------------------------------
 * 1-D FFT: recursive DIT, smallest-prime-first, O(p^2) combination.
 * ------------------------------------------------------------------------ */
struct twiddles {
  int n;
  std::vector<cplx> w; /* w[k] = exp(-2 pi i k / n) */
};

static const twiddles &get_twiddles(int n) {
  static std::mutex mtx;
  static std::map<int, std::unique_ptr<twiddles>> cache;
  std::lock_guard<std::mutex> lock(mtx);
  auto it = cache.find(n);
  if (it != cache.end()) return *it->second;
  auto tw = std::make_unique<twiddles>();
  tw->n = n;
  tw->w.resize(n);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int k = 0; k < n; k++) {
    long double ang = -2.0L * pi * (long double)k / (long double)n;
    tw->w[k] = cplx((double)cosl(ang), (double)sinl(ang));
  }
  auto &ref = *tw;
  cache[n] = std::move(tw);
  return ref;
}

static int smallest_factor(int n) {
  for (int p : {2, 3, 5}) if (n % p == 0) return p;
  for (int p = 7; p * p <= n; p += 2) if (n % p == 0) return p;
  return n;
}

/* DFT of length nsub of in[0], in[istride], ... into out[0..nsub).
 * tw_step = N / nsub so that exp(-2 pi i j / nsub) = w[j * tw_step]. */
static void fft_rec(const cplx *in, long istride, cplx *out, int nsub, int tw_step,
                    const std::vector<cplx> &w, int N, bool inverse, cplx *scratch) {
  if (nsub == 1) { out[0] = in[0]; return; }
  int p = smallest_factor(nsub);
  int m = nsub / p;
  for (int q = 0; q < p; q++)
    fft_rec(in + q * istride, istride * p, out + q * m, m, tw_step * p, w, N, inverse, scratch);
  auto W = [&](long idx) { cplx c = w[idx % N]; return inverse ? std::conj(c) : c; };
  for (int k = 0; k < m; k++) {
    for (int q = 0; q < p; q++) scratch[q] = out[q * m + k] * W((long)q * k * tw_step);
    for (int s = 0; s < p; s++) {
      cplx acc = 0.0;
      for (int q = 0; q < p; q++) acc += scratch[q] * W((long)((q * s) % p) * m * tw_step);
      scratch[p + s] = acc;
    }
    for (int s = 0; s < p; s++) out[k + s * m] = scratch[p + s];
  }
}

static void fft1d(int n, const cplx *in, cplx *out, int sign, cplx *scratch) {
  const twiddles &tw = get_twiddles(n);
  fft_rec(in, 1, out, n, 1, tw.w, n, sign > 0, scratch);
}

static int max_prime(int n) {
  int best = 1;
  while (n > 1) { int p = smallest_factor(n); best = std::max(best, p); n /= p; }
  return best;
}

/* Batched 3-D DFT over [x][y][z][b] (src/operations/transform.hpp:132, 265). */
static void fft3d(const int n[3], int nst, int ld, const cplx *in, cplx *out, int sign,
                  bool normalize) {
  const long N = (long)n[0] * n[1] * n[2];
  /* dense working copy [point][band] */
  std::vector<cplx> work((size_t)N * nst);
  for (long p = 0; p < N; p++)
    for (int b = 0; b < nst; b++) work[p * nst + b] = in[p * ld + b];
  const long stride[3] = {(long)n[1] * n[2], (long)n[2], 1};
  for (int axis = 0; axis < 3; axis++) {
    const int a1 = (axis + 1) % 3, a2 = (axis + 2) % 3;
    const int len = n[axis];
    const long nlines = (long)n[a1] * n[a2] * nst;
    int pmax = max_prime(len);
#pragma omp parallel
    {
      std::vector<cplx> line(len), res(len), scratch(2 * pmax + 2);
#pragma omp for schedule(static)
      for (long l = 0; l < nlines; l++) {
        int b = (int)(l % nst);
        long rest = l / nst;
        int i1 = (int)(rest % n[a1]);
        int i2 = (int)(rest / n[a1]);
        long base = i1 * stride[a1] + i2 * stride[a2];
        for (int i = 0; i < len; i++) line[i] = work[(base + i * stride[axis]) * nst + b];
        fft1d(len, line.data(), res.data(), sign, scratch.data());
        for (int i = 0; i < len; i++) work[(base + i * stride[axis]) * nst + b] = res[i];
      }
    }
  }
  const double factor = 1.0 / (double)N;
  for (long p = 0; p < N; p++)
    for (int b = 0; b < nst; b++) {
      cplx v = work[p * nst + b];
      /* transform.hpp:327-330: ar[ip] = factor*ar[ip] */
      out[p * ld + b] = normalize ? factor * v : v;
    }
}

/* ---------------------------------------------------------------------------
 * Hamiltonian pieces
 * ------------------------------------------------------------------------ */
static long npoints(const orc_system *s) { return (long)s->n[0] * s->n[1] * s->n[2]; }

/* projector_all::project (src/hamiltonian/projector_all.hpp:98-216), CPU
 * branches (:117-125, :172-180). */
static void project(const orc_system *s, const double kA_cov[3], int nst, int ld,
                    const cplx *psi, cplx *sphere) {
  cell c(s->lattice);
  const long maxK = s->max_sphere;
  const double dV = c.volume / (double)npoints(s); /* real_space.hpp:150-152 */
#pragma omp parallel for schedule(dynamic)
  for (int ip = 0; ip < s->nproj; ip++) {
    cplx *sph = sphere + (long)ip * maxK * nst;
    /* gather with phase polar(1, dot(k, pos)) (:104-115) */
    for (long i = 0; i < maxK; i++) {
      const int *poi = s->points + ((long)ip * maxK + i) * 3;
      if (poi[0] >= 0) {
        const float *pos = s->positions + ((long)ip * maxK + i) * 3;
        double arg = kA_cov[0] * (double)pos[0] + kA_cov[1] * (double)pos[1] +
                     kA_cov[2] * (double)pos[2];
        cplx phase = std::polar(1.0, arg);
        long gp = ((long)poi[0] * s->n[1] + poi[1]) * s->n[2] + poi[2];
        for (int b = 0; b < nst; b++) sph[i * nst + b] = phase * psi[gp * ld + b];
      } else {
        for (int b = 0; b < nst; b++) sph[i * nst + b] = 0.0;
      }
    }
    if (s->sphere_size[ip] == 0 || s->nlm[ip] == 0) continue; /* locally_empty_ */
    const double *M = s->matrices + (long)ip * s->max_nlm * maxK;
    /* GEMM1: P[lm][b] = dV sum_i M[lm][i] sph[i][b] (:117-125) */
    std::vector<cplx> P((size_t)s->max_nlm * nst, 0.0);
    for (int lm = 0; lm < s->max_nlm; lm++)
      for (long i = 0; i < maxK; i++) {
        double m = M[lm * maxK + i];
        for (int b = 0; b < nst; b++) P[lm * nst + b] += m * sph[i * nst + b];
      }
    for (int lm = 0; lm < s->max_nlm; lm++)
      for (int b = 0; b < nst; b++) P[lm * nst + b] *= dV;
    /* scale by the KB coefficients (:158-165) */
    for (int lm = 0; lm < s->max_nlm; lm++)
      for (int b = 0; b < nst; b++) P[lm * nst + b] *= s->coeff[(long)ip * s->max_nlm + lm];
    /* GEMM2: sph[i][b] = sum_lm M[lm][i] P[lm][b] (:172-180) */
    for (long i = 0; i < maxK; i++)
      for (int b = 0; b < nst; b++) {
        cplx acc = 0.0;
        for (int lm = 0; lm < s->max_nlm; lm++) acc += M[lm * maxK + i] * P[lm * nst + b];
        sph[i * nst + b] = acc;
      }
  }
}

/* projector_all::apply (src/hamiltonian/projector_all.hpp:220-232), with the
 * scatter in fixed (projector, point) order instead of atomics. */
static void apply(const orc_system *s, const double kA_cov[3], int nst, int ld,
                  const cplx *sphere, cplx *hpsi) {
  const long maxK = s->max_sphere;
  for (int ip = 0; ip < s->nproj; ip++) {
    if (s->sphere_size[ip] == 0 || s->nlm[ip] == 0) continue;
    const cplx *sph = sphere + (long)ip * maxK * nst;
    for (long i = 0; i < maxK; i++) {
      const int *poi = s->points + ((long)ip * maxK + i) * 3;
      if (poi[0] < 0) continue;
      const float *pos = s->positions + ((long)ip * maxK + i) * 3;
      double arg = kA_cov[0] * (double)pos[0] + kA_cov[1] * (double)pos[1] +
                   kA_cov[2] * (double)pos[2];
      cplx phase = std::polar(1.0, -arg);
      long gp = ((long)poi[0] * s->n[1] + poi[1]) * s->n[2] + poi[2];
      for (int b = 0; b < nst; b++) hpsi[gp * ld + b] += phase * sph[i * nst + b];
    }
  }
}

/* laplacian(ff, -0.5, -2 (k+A)^contra) G-space branch
 * (src/operations/laplacian.hpp:70-79), in place on a dense [N][nst] array. */
static void laplacian_kA(const orc_system *s, const double kA_cov[3], int nst, cplx *ff) {
  cell c(s->lattice);
  vec3 kA{{kA_cov[0], kA_cov[1], kA_cov[2]}};
  vec3 gc = c.cov_to_contra(kA);
  gc = -2.0 * gc;
  const double factor = -0.5;
#pragma omp parallel for schedule(static)
  for (int ix = 0; ix < s->n[0]; ix++)
    for (int iy = 0; iy < s->n[1]; iy++)
      for (int iz = 0; iz < s->n[2]; iz++) {
        /* fourier_space.hpp:66-73: G covariant = 2 pi * symmetric index */
        vec3 g{{2.0 * M_PI * symmetric(ix, s->n[0]), 2.0 * M_PI * symmetric(iy, s->n[1]),
                2.0 * M_PI * symmetric(iz, s->n[2])}};
        double g2 = c.norm_cov(g);
        double lapl = factor * (-g2 + (gc[0] * g[0] + gc[1] * g[1] + gc[2] * g[2]));
        long p = ((long)ix * s->n[1] + iy) * s->n[2] + iz;
        for (int b = 0; b < nst; b++) ff[p * nst + b] = lapl * ff[p * nst + b];
      }
}

static void hamiltonian(const orc_system *s, const double kA_cov[3], int nst, int ld,
                        const cplx *psi, cplx *hpsi) {
  cell c(s->lattice);
  const long N = npoints(s);
  const long maxK = s->max_sphere;
  /* ks_hamiltonian.hpp:136 */
  std::vector<cplx> sphere((size_t)std::max<long>(1, s->nproj * maxK * nst));
  if (s->nproj > 0) project(s, kA_cov, nst, ld, psi, sphere.data());
  /* :138-144 */
  std::vector<cplx> fs((size_t)N * nst);
  fft3d(s->n, nst, ld, psi, fs.data(), -1, false);
  laplacian_kA(s, kA_cov, nst, fs.data());
  std::vector<cplx> rs((size_t)N * nst);
  fft3d(s->n, nst, nst, fs.data(), rs.data(), +1, true);
  /* :146 scalar_potential_add with shift 0.5 |k+A|^2 (scalar_potential.hpp:29-34) */
  vec3 kA{{kA_cov[0], kA_cov[1], kA_cov[2]}};
  const double shift = 0.5 * c.norm_cov(kA);
#pragma omp parallel for schedule(static)
  for (long p = 0; p < N; p++)
    for (int b = 0; b < nst; b++)
      hpsi[p * ld + b] = rs[p * nst + b] + (s->vloc[p] + shift) * psi[p * ld + b];
  /* :149 */
  if (s->nproj > 0) apply(s, kA_cov, nst, ld, sphere.data(), hpsi);
}

/* ks_hamiltonian::operator()(orbital_set<fourier_space>) (ks_hamiltonian.hpp:156-176):
 * phi_rs = to_real(phi); h_rs = (V + 1/2|k+A|^2) phi_rs + V_nl phi_rs;
 * h = to_fourier(h_rs); h += laplacian(phi, -1/2, -2 (k+A)^contra) (laplacian_add). */
static void hamiltonian_fourier(const orc_system *s, const double kA_cov[3], int nst, int ld,
                                const cplx *phi, cplx *hphi) {
  cell c(s->lattice);
  const long N = npoints(s);
  const long maxK = s->max_sphere;
  std::vector<cplx> rs((size_t)N * nst), hrs((size_t)N * nst);
  fft3d(s->n, nst, ld, phi, rs.data(), +1, true);
  std::vector<cplx> sphere((size_t)std::max<long>(1, s->nproj * maxK * nst));
  if (s->nproj > 0) project(s, kA_cov, nst, nst, rs.data(), sphere.data());
  vec3 kA{{kA_cov[0], kA_cov[1], kA_cov[2]}};
  const double shift = 0.5 * c.norm_cov(kA);
#pragma omp parallel for schedule(static)
  for (long p = 0; p < N; p++)
    for (int b = 0; b < nst; b++) hrs[p * nst + b] = (s->vloc[p] + shift) * rs[p * nst + b];
  if (s->nproj > 0) apply(s, kA_cov, nst, nst, sphere.data(), hrs.data());
  std::vector<cplx> fs((size_t)N * nst);
  fft3d(s->n, nst, nst, hrs.data(), fs.data(), -1, false);
  std::vector<cplx> lap((size_t)N * nst);
  for (long p = 0; p < N; p++)
    for (int b = 0; b < nst; b++) lap[p * nst + b] = phi[p * ld + b];
  laplacian_kA(s, kA_cov, nst, lap.data());
  for (long p = 0; p < N; p++)
    for (int b = 0; b < nst; b++) hphi[p * ld + b] = fs[p * nst + b] + lap[p * nst + b];
}

/* operations::shift (src/operations/shift.hpp:33-44): phi += factor*shift */
static void shift_add(long N, int nst, int ld, cplx factor, const cplx *x, cplx *y) {
#pragma omp parallel for schedule(static)
  for (long p = 0; p < N; p++)
    for (int b = 0; b < nst; b++) y[p * ld + b] += factor * x[p * ld + b];
}

/* exponential.hpp:21-34 */
static void exponential_in_place(const orc_system *s, const double kA[3], cplx z, int order,
                                 int nst, int ld, cplx *phi) {
  const long N = npoints(s);
  std::vector<cplx> hn((size_t)N * ld), tmp((size_t)N * ld);
  hamiltonian(s, kA, nst, ld, phi, hn.data());
  cplx coeff = 1.0;
  for (int iter = 1; iter <= order; iter++) {
    if (iter > 1) { hamiltonian(s, kA, nst, ld, hn.data(), tmp.data()); std::swap(hn, tmp); }
    coeff *= z / (double)iter;
    shift_add(N, nst, ld, coeff, hn.data(), phi);
  }
}

/* ---------------------------------------------------------------------------
 * pcg32 (O'Neill's PCG XSH-RR 64/32 as pcg-cpp's setseq engine with the
 * default stream): state' = state * 6364136223846793005 + 1442695040888963407,
 * output from the previous state, seeded with bump(seed + increment).
 * ------------------------------------------------------------------------ */
struct pcg32 {
  static constexpr uint64_t mult = 6364136223846793005ULL;
  static constexpr uint64_t inc = 1442695040888963407ULL;
  uint64_t state;
  explicit pcg32(uint64_t seed) : state((seed + inc) * mult + inc) {}
  uint32_t operator()() {
    uint64_t old = state;
    state = old * mult + inc;
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
  }
  // jump ahead by delta LCG steps (Brown, "Random number generation with
  // arbitrary strides", 1994): accumulate the affine map by squaring
  void discard(uint64_t delta) {
    uint64_t cur_mult = mult, cur_plus = inc, acc_mult = 1u, acc_plus = 0u;
    while (delta > 0) {
      if (delta & 1u) {
        acc_mult *= cur_mult;
        acc_plus = acc_plus * cur_mult + cur_plus;
      }
      cur_plus = (cur_mult + 1u) * cur_plus;
      cur_mult *= cur_mult;
      delta >>= 1;
    }
    state = acc_mult * state + acc_plus;
  }
};

/* gradient(phi, -1, k+A) then current density (current.hpp:34-42) plus the
 * [V_nl, r] term; returns J_cov.  commutator 0 is position_commutator as the
 * reference writes it (projector_all.hpp:319-403: the plain projection scaled
 * by the KB coefficient twice, :362-366, the r-weighted one not at all);
 * commutator 1 scales both projections once, which makes the term exactly
 * -dV_nl/dA. */
static void current(const orc_system *s, const double kA_cov[3], double occ, int nst, int ld,
                    const cplx *psi, int commutator, double J[3]) {
  cell c(s->lattice);
  const long N = npoints(s);
  std::vector<cplx> fs((size_t)N * nst);
  fft3d(s->n, nst, ld, psi, fs.data(), -1, false);
  // g phi = -i (G + k + A) phi, three covariant components
  std::vector<cplx> g[3];
  for (int d = 0; d < 3; d++) {
    std::vector<cplx> gd((size_t)N * nst);
    for (int ix = 0; ix < s->n[0]; ix++)
      for (int iy = 0; iy < s->n[1]; iy++)
        for (int iz = 0; iz < s->n[2]; iz++) {
          const int m[3] = {symmetric(ix, s->n[0]), symmetric(iy, s->n[1]), symmetric(iz, s->n[2])};
          const double gcov = 2.0 * M_PI * m[d] + kA_cov[d];
          const long p = ((long)ix * s->n[1] + iy) * s->n[2] + iz;
          for (int b = 0; b < nst; b++) gd[p * nst + b] = cplx(0.0, -1.0) * gcov * fs[p * nst + b];
        }
    g[d].resize((size_t)N * nst);
    fft3d(s->n, nst, nst, gd.data(), g[d].data(), +1, true);
  }
  // position_commutator: [V_nl, r] phi, r relative to the atom, covariant
  if (s->nproj > 0) {
    const long maxK = s->max_sphere;
    const double dV = c.volume / (double)N;
    for (int ip = 0; ip < s->nproj; ip++) {
      if (s->sphere_size[ip] == 0 || s->nlm[ip] == 0) continue;
      const double *M = s->matrices + (long)ip * s->max_nlm * maxK;
      std::vector<cplx> sph((size_t)maxK * nst, 0.0), rsph((size_t)maxK * 3 * nst, 0.0);
      for (long i = 0; i < maxK; i++) {
        const int *poi = s->points + ((long)ip * maxK + i) * 3;
        if (poi[0] < 0) continue;
        const float *pos = s->positions + ((long)ip * maxK + i) * 3;
        vec3 rr{{(double)pos[0], (double)pos[1], (double)pos[2]}};
        double arg = kA_cov[0] * rr[0] + kA_cov[1] * rr[1] + kA_cov[2] * rr[2];
        cplx ph = std::polar(1.0, arg);
        long gp = ((long)poi[0] * s->n[1] + poi[1]) * s->n[2] + poi[2];
        for (int b = 0; b < nst; b++) {
          sph[i * nst + b] = ph * psi[gp * ld + b];
          for (int d = 0; d < 3; d++) rsph[(i * 3 + d) * nst + b] = rr[d] * sph[i * nst + b];
        }
      }
      std::vector<cplx> P((size_t)s->max_nlm * nst, 0.0), R((size_t)s->max_nlm * 3 * nst, 0.0);
      for (int lm = 0; lm < s->max_nlm; lm++)
        for (long i = 0; i < maxK; i++) {
          double m = M[lm * maxK + i];
          for (int b = 0; b < nst; b++) {
            P[lm * nst + b] += dV * m * sph[i * nst + b];
            for (int d = 0; d < 3; d++) R[(lm * 3 + d) * nst + b] += dV * m * rsph[(i * 3 + d) * nst + b];
          }
        }
      for (int lm = 0; lm < s->max_nlm; lm++) {
        double co = s->coeff[(long)ip * s->max_nlm + lm];
        for (int b = 0; b < nst; b++) {
          if (commutator == 0) {
            // projector_all.hpp:362-366: proj *= coe; proj *= coe; rproj untouched
            P[lm * nst + b] *= co;
            P[lm * nst + b] *= co;
          } else {
            P[lm * nst + b] *= co;
            for (int d = 0; d < 3; d++) R[(lm * 3 + d) * nst + b] *= co;
          }
        }
      }
      for (long i = 0; i < maxK; i++) {
        const int *poi = s->points + ((long)ip * maxK + i) * 3;
        if (poi[0] < 0) continue;
        const float *pos = s->positions + ((long)ip * maxK + i) * 3;
        vec3 rr{{(double)pos[0], (double)pos[1], (double)pos[2]}};
        double arg = kA_cov[0] * rr[0] + kA_cov[1] * rr[1] + kA_cov[2] * rr[2];
        cplx ph = std::polar(1.0, -arg);
        long gp = ((long)poi[0] * s->n[1] + poi[1]) * s->n[2] + poi[2];
        for (int b = 0; b < nst; b++) {
          cplx sphi = 0.0, srphi[3] = {0.0, 0.0, 0.0};
          for (int lm = 0; lm < s->max_nlm; lm++) {
            double m = M[lm * maxK + i];
            sphi += m * P[lm * nst + b];
            for (int d = 0; d < 3; d++) srphi[d] += m * R[(lm * 3 + d) * nst + b];
          }
          // to_covariant(srphi - rr * sphi) (projector_all.hpp:398)
          cplx contra[3];
          for (int d = 0; d < 3; d++) contra[d] = srphi[d] - rr[d] * sphi;
          // covariant = lattice . cartesian, cartesian = sum_e lat_e contra_e
          cplx cartv[3];
          for (int d = 0; d < 3; d++)
            cartv[d] = c.lat[0][d] * contra[0] + c.lat[1][d] * contra[1] + c.lat[2][d] * contra[2];
          for (int d = 0; d < 3; d++) {
            cplx cov = c.lat[d][0] * cartv[0] + c.lat[d][1] * cartv[1] + c.lat[d][2] * cartv[2];
            g[d][gp * nst + b] += ph * cov;
          }
        }
      }
    }
  }
  const double dV = c.volume / (double)N;
  for (int d = 0; d < 3; d++) {
    double sum = 0.0;
    for (long p = 0; p < N; p++)
      for (int b = 0; b < nst; b++) {
        cplx ph = psi[p * ld + b], gp = g[d][p * nst + b];
        sum += 0.5 * occ * std::imag(std::conj(ph) * gp - std::conj(gp) * ph);
      }
    J[d] = dV * sum;  /* operations::integral (integral.hpp:24-29) */
  }
}

} // namespace

/* ===========================================================================
 * C ABI
 * ======================================================================== */
/* ---------------------------------------------------------------------------
 * Ground state: laplacian_add, the TPA preconditioner, orthogonalize and the
 * steepest-descent eigensolver (the GS path of ground_state/calculator.hpp:
 * 170-180).
 * ------------------------------------------------------------------------- */

/* operations::laplacian_add, G-space branch (src/operations/laplacian.hpp:24-38):
 * laplff += factor (-|G|^2 + gradcoeff . G) ff with G covariant = 2 pi m
 * (fourier_space.hpp:66-81) and gradcoeff contravariant. */
static void laplacian_add(const int n[3], const cell &c, double factor, const double gradcoeff[3],
                          int nst, int ld, const cplx *ff, cplx *laplff) {
#pragma omp parallel for schedule(static)
  for (int ix = 0; ix < n[0]; ix++)
    for (int iy = 0; iy < n[1]; iy++)
      for (int iz = 0; iz < n[2]; iz++) {
        vec3 g{{2.0 * M_PI * symmetric(ix, n[0]), 2.0 * M_PI * symmetric(iy, n[1]),
                2.0 * M_PI * symmetric(iz, n[2])}};
        const double lapl =
            factor * (-c.norm_cov(g) + (gradcoeff[0] * g[0] + gradcoeff[1] * g[1] + gradcoeff[2] * g[2]));
        const long p = ((long)ix * n[1] + iy) * n[2] + iz;
        for (int b = 0; b < nst; b++) laplff[p * ld + b] += lapl * ff[p * ld + b];
      }
}

/* operations::overlap_diagonal (src/operations/overlap_diagonal.hpp:25-60):
 * out[b] = dV sum_p conj(a[p][b]) b[p][b]. */
static void overlap_diag(long N, int nst, int ld, double dV, const cplx *a, const cplx *bb,
                         cplx *out) {
  for (int b = 0; b < nst; b++) {
    cplx v = 0.0;
    for (long p = 0; p < N; p++) v += std::conj(a[p * ld + b]) * bb[p * ld + b];
    out[b] = dV * v;
  }
}

/* The Teter-Payne-Allan function of operations::preconditioner
 * (src/operations/preconditioner.hpp:30-37). */
static double tpa_k(double x) {
  const double num = ((8.0 * x + 12.0) * x + 18.0) * x + 27.0;
  const double den = (((16.0 * x + 8.0) * x + 12.0) * x + 18.0) * x + 27.0;
  return num / den;
}

/* operations::preconditioner on fourier_space orbitals (preconditioner.hpp:39-55):
 * kinetic_b = Re<laplacian(phi, -1/2)|phi> / <phi|phi>, then
 * phi(G) *= K(1/2 |G|^2 / kinetic_b). */
static void precondition(const int n[3], const cell &c, int nst, int ld, cplx *phi) {
  const long N = (long)n[0] * n[1] * n[2];
  std::vector<double> g2h((size_t)N);
  for (int ix = 0; ix < n[0]; ix++)
    for (int iy = 0; iy < n[1]; iy++)
      for (int iz = 0; iz < n[2]; iz++) {
        vec3 g{{2.0 * M_PI * symmetric(ix, n[0]), 2.0 * M_PI * symmetric(iy, n[1]),
                2.0 * M_PI * symmetric(iz, n[2])}};
        g2h[((long)ix * n[1] + iy) * n[2] + iz] = -0.5 * (-c.norm_cov(g));
      }
  std::vector<double> kin((size_t)nst);
  for (int b = 0; b < nst; b++) {
    cplx v = 0.0, nn = 0.0;
    for (long p = 0; p < N; p++) {
      v += std::conj(g2h[p] * phi[p * ld + b]) * phi[p * ld + b];
      nn += std::conj(phi[p * ld + b]) * phi[p * ld + b];
    }
    kin[b] = std::real(v / nn);
  }
#pragma omp parallel for schedule(static)
  for (long p = 0; p < N; p++)
    for (int b = 0; b < nst; b++) phi[p * ld + b] *= tpa_k(g2h[p] / kin[b]);
}

/* operations::orthogonalize (src/operations/orthogonalize.hpp:25-32):
 * S = dV Phi^H Phi (overlap.hpp:26-60); zpotrf on the row-major S gives
 * S = R R^H with R lower triangular (matrix/cholesky.hpp:29-80);
 * invert_triangular and rotate_trs (rotate.hpp:147-156) then set
 * Phi <- Phi R^-H, i.e. Gram-Schmidt in band order. */
static void orthogonalize(long N, int nst, int ld, double dV, cplx *phi) {
  std::vector<cplx> S((size_t)nst * nst);
  for (int i = 0; i < nst; i++)
    for (int j = 0; j <= i; j++) {
      cplx v = 0.0;
      for (long p = 0; p < N; p++) v += std::conj(phi[p * ld + i]) * phi[p * ld + j];
      S[(size_t)i * nst + j] = dV * v;
      S[(size_t)j * nst + i] = std::conj(dV * v);
    }
  /* S[i][j] = sum_k R[i][k] conj(R[j][k]) */
  std::vector<cplx> R((size_t)nst * nst, 0.0);
  for (int j = 0; j < nst; j++) {
    double d = std::real(S[(size_t)j * nst + j]);
    for (int k = 0; k < j; k++) d -= std::norm(R[(size_t)j * nst + k]);
    const double rjj = std::sqrt(d);
    R[(size_t)j * nst + j] = rjj;
    for (int i = j + 1; i < nst; i++) {
      cplx v = S[(size_t)i * nst + j];
      for (int k = 0; k < j; k++) v -= R[(size_t)i * nst + k] * std::conj(R[(size_t)j * nst + k]);
      R[(size_t)i * nst + j] = v / rjj;
    }
  }
  /* each point's row x solves x R^H = phi_row by forward substitution */
#pragma omp parallel for schedule(static)
  for (long p = 0; p < N; p++) {
    cplx *row = phi + p * ld;
    for (int j = 0; j < nst; j++) {
      cplx v = row[j];
      for (int i = 0; i < j; i++) v -= row[i] * std::conj(R[(size_t)j * nst + i]);
      row[j] = v / std::real(R[(size_t)j * nst + j]);
    }
  }
}

/* eigensolvers::steepest_descent (src/eigensolvers/steepest_descent.hpp:23-88):
 * five line minimisations along the (preconditioned) residual, each with the
 * exact step of the 2x2 Rayleigh-Ritz problem, then orthogonalize. */
template <class Op, class Prec>
static void steepest_descent(long N, int nst, int ld, double dV, Op ham, Prec prec, cplx *phi) {
  const int num_steps = 5;
  const size_t sz = (size_t)N * ld;
  std::vector<cplx> hphi(sz), residual(sz), hres(sz);
  ham(phi, hphi.data());
  std::vector<cplx> eig(nst), norm(nst), mm0(nst), mm1(nst), mm2(nst), mm3(nst);
  std::vector<double> lam(nst);
  for (int istep = 0; istep < num_steps; istep++) {
    residual = hphi;
    overlap_diag(N, nst, ld, dV, residual.data(), phi, eig.data());
    overlap_diag(N, nst, ld, dV, phi, phi, norm.data());
    std::vector<cplx> evn(nst);
    for (int b = 0; b < nst; b++) evn[b] = eig[b] / std::real(norm[b]);
    /* shift(-1, evnorm, phi, residual): residual -= evnorm phi (shift.hpp:19-31) */
    for (long p = 0; p < N; p++)
      for (int b = 0; b < nst; b++) residual[p * ld + b] += -1.0 * evn[b] * phi[p * ld + b];
    prec(residual.data());
    ham(residual.data(), hres.data());
    overlap_diag(N, nst, ld, dV, residual.data(), residual.data(), mm0.data());
    overlap_diag(N, nst, ld, dV, phi, residual.data(), mm1.data());
    overlap_diag(N, nst, ld, dV, residual.data(), hres.data(), mm2.data());
    overlap_diag(N, nst, ld, dV, phi, hres.data(), mm3.data());
    for (int b = 0; b < nst; b++) {
      const double ca = std::real(mm0[b] * mm3[b] - mm2[b] * mm1[b]);
      const double cb = std::real(norm[b] * mm2[b] - eig[b] * mm0[b]);
      const double cc = std::real(eig[b] * mm1[b] - mm3[b] * norm[b]);
      const double den = cb + std::sqrt(cb * cb - 4.0 * ca * cc);
      lam[b] = std::fabs(den) < 1e-15 ? 0.0 : 2.0 * cc / den;
    }
    for (long p = 0; p < N; p++)
      for (int b = 0; b < nst; b++) {
        phi[p * ld + b] += lam[b] * residual[p * ld + b];
        if (istep != num_steps - 1) hphi[p * ld + b] += lam[b] * hres[p * ld + b];
      }
  }
  orthogonalize(N, nst, ld, dV, phi);
}

/* ---------------------------------------------------------------------------
 * Interacting TDDFT pieces: density, Hartree (periodic Poisson), LDA XC,
 * the Kohn-Sham potential update and the self-consistent ETRS step.
 * ------------------------------------------------------------------------- */

/* observables::density::calculate_add (src/observables/density.hpp:23-31),
 * one spin component: rho[p] += sum_b occ[b] |psi[p][b]|^2 */
static void density_add(long N, int nst, int ld, const double *occ, const cplx *psi, double *rho) {
#pragma omp parallel for schedule(static)
  for (long p = 0; p < N; p++)
    for (int b = 0; b < nst; b++) rho[p] += occ[b] * std::norm(psi[p * ld + b]);
}

/* solvers::poisson, 3-D periodic (src/solvers/poisson.hpp:29-35, 69-97,
 * 224-227): to_fourier, V(G) *= (-4 pi / N) * (-1/|G|^2) (0 for |G|^2 < 1e-6),
 * to_real without normalisation; the real part. */
static void poisson3d(const int n[3], const cell &c, const double *rho, double *vh) {
  const long N = (long)n[0] * n[1] * n[2];
  std::vector<cplx> f((size_t)N), g((size_t)N);
  for (long p = 0; p < N; p++) f[p] = rho[p];
  fft3d(n, 1, 1, f.data(), g.data(), -1, false);
  const double scal = (-4.0 * M_PI) / (double)N;
#pragma omp parallel for schedule(static)
  for (int ix = 0; ix < n[0]; ix++)
    for (int iy = 0; iy < n[1]; iy++)
      for (int iz = 0; iz < n[2]; iz++) {
        vec3 gv{{2.0 * M_PI * symmetric(ix, n[0]), 2.0 * M_PI * symmetric(iy, n[1]),
                 2.0 * M_PI * symmetric(iz, n[2])}};
        const double g2 = c.norm_cov(gv);
        const double ker = g2 < 1e-6 ? 0.0 : -1.0 / g2;
        g[((long)ix * n[1] + iy) * n[2] + iz] *= scal * ker;
      }
  fft3d(n, 1, 1, g.data(), f.data(), +1, false);
  for (long p = 0; p < N; p++) vh[p] = std::real(f[p]);
}

/* Unpolarised LDA exchange (Dirac/Slater) and Perdew-Zunger 1981
 * correlation, the functionals libxc calls LDA_X and LDA_C_PZ that
 * options::theory::lda() selects (src/options/theory.hpp:48-53).  Energy
 * per particle e and potential v = d(n e)/dn; densities below 1e-15 give 0. */
static void lda_x_point(double n, double &e, double &v) {
  e = v = 0.0;
  if (n < 1e-15) return;
  const double cx = std::cbrt(3.0 / M_PI);
  const double n13 = std::cbrt(n);
  e = -0.75 * cx * n13;
  v = -cx * n13;
}

static void lda_point(double n, double &e, double &v) {
  e = v = 0.0;
  if (n < 1e-15) return;
  double ex, vx;
  lda_x_point(n, ex, vx);
  const double rs = std::cbrt(3.0 / (4.0 * M_PI * n));
  double ec, vc;
  if (rs >= 1.0) {
    const double gam = -0.1423, b1 = 1.0529, b2 = 0.3334;
    const double sq = std::sqrt(rs);
    const double den = 1.0 + b1 * sq + b2 * rs;
    ec = gam / den;
    vc = ec * (1.0 + (7.0 / 6.0) * b1 * sq + (4.0 / 3.0) * b2 * rs) / den;
  } else {
    const double A = 0.0311, B = -0.048, C = 0.0020, D = -0.0116;
    const double lr = std::log(rs);
    ec = A * lr + B + C * rs * lr + D * rs;
    vc = A * lr + (B - A / 3.0) + (2.0 / 3.0) * C * rs * lr + (2.0 * D - C) / 3.0 * rs;
  }
  e = ex + ec;
  v = vx + vc;
}

/* Unpolarised PBE (Perdew, Burke, Ernzerhof, PRL 77, 3865 (1996)), the
 * functionals libxc calls GGA_X_PBE and GGA_C_PBE that options::theory's
 * default selects (src/options/theory.hpp:40-45, 73-77).  For a point with
 * density n and sigma = |grad n|^2: energy per particle e and the partial
 * derivatives vn = d(n e)/dn, vs = d(n e)/dsigma.  The uniform-gas
 * correlation is Perdew-Wang 1992 (PRB 45, 13244) with the parameters of
 * libxc's LDA_C_PW_MOD (A = 0.0310907), as GGA_C_PBE uses it. */
static void pw92_point(double rs, double &ec, double &dec) {
  const double A = 0.0310907, a1 = 0.21370, b1 = 7.5957, b2 = 3.5876, b3 = 1.6382,
               b4 = 0.49294;
  const double sq = std::sqrt(rs);
  const double q0 = -2.0 * A * (1.0 + a1 * rs);
  const double q1 = 2.0 * A * (b1 * sq + b2 * rs + b3 * rs * sq + b4 * rs * rs);
  const double q1p = A * (b1 / sq + 2.0 * b2 + 3.0 * b3 * sq + 4.0 * b4 * rs);
  const double lg = std::log(1.0 + 1.0 / q1);
  ec = q0 * lg;
  dec = -2.0 * A * a1 * lg - q0 * q1p / (q1 * q1 + q1);
}

static void pbe_x_point(double n, double sigma, double &e, double &vn, double &vs) {
  e = vn = vs = 0.0;
  if (n < 1e-15) return;
  const double kappa = 0.8040, mu = 0.2195149727645171;
  const double ax = -0.75 * std::cbrt(3.0 / M_PI);
  const double n13 = std::cbrt(n);
  const double c = 4.0 * std::pow(3.0 * M_PI * M_PI, 2.0 / 3.0);
  const double x = sigma / (c * n13 * n13 * n13 * n13 * n13 * n13 * n13 * n13);  // s^2
  const double den = 1.0 + mu * x / kappa;
  const double F = 1.0 + kappa - kappa / den;
  const double dF = mu / (den * den);
  e = ax * n13 * F;
  vn = (4.0 / 3.0) * ax * n13 * F + ax * n13 * n * dF * (-8.0 / 3.0) * x / n;
  vs = ax * n13 * n * dF / (c * n13 * n13 * n13 * n13 * n13 * n13 * n13 * n13);
}

static void pbe_c_point(double n, double sigma, double &e, double &vn, double &vs) {
  e = vn = vs = 0.0;
  if (n < 1e-15) return;
  const double beta = 0.06672455060314922, gamma = (1.0 - std::log(2.0)) / (M_PI * M_PI);
  const double rs = std::cbrt(3.0 / (4.0 * M_PI * n));
  double ec, decdrs;
  pw92_point(rs, ec, decdrs);
  const double dec = decdrs * (-rs / (3.0 * n));  // d ec / dn
  const double kf = std::cbrt(3.0 * M_PI * M_PI * n);
  const double y = sigma * M_PI / (16.0 * kf * n * n);  // t^2
  const double ex = std::exp(-ec / gamma);
  const double A = (beta / gamma) / (ex - 1.0);
  const double dAde = (beta / (gamma * gamma)) * ex / ((ex - 1.0) * (ex - 1.0));
  const double P = 1.0 + A * y, D = 1.0 + A * y + A * A * y * y;
  const double N = (beta / gamma) * y * P / D;
  const double H = gamma * std::log(1.0 + N);
  const double dNdy = (beta / gamma) * ((P + A * y) * D - y * P * (A + 2.0 * A * A * y)) / (D * D);
  const double dNdA = (beta / gamma) * y * (y * D - P * (y + 2.0 * A * y * y)) / (D * D);
  const double dHdy = gamma / (1.0 + N) * dNdy, dHdA = gamma / (1.0 + N) * dNdA;
  e = ec + H;
  vn = e + n * (dec + dHdA * dAde * dec + dHdy * (-7.0 / 3.0) * y / n);
  vs = n * dHdy * M_PI / (16.0 * kf * n * n);
}

/* hamiltonian::xc_term::evaluate_functional for a GGA, one spin component
 * (src/hamiltonian/xc_term.hpp:267-300): grad n = Re to_real(i G to_fourier
 * n) (operations::gradient, src/operations/gradient.hpp:22-45), sigma =
 * |grad n|^2, the point forms, then v = vn - 2 div(vs grad n) with the same
 * spectral divergence; e_xc = dV sum e n.  which: 1 = exchange, 2 =
 * correlation, 3 = both. */
static void gga_xc(const int n[3], const cell &c, const double *rho, int which, double *vxc,
                   double *exc) {
  const long N = (long)n[0] * n[1] * n[2];
  std::vector<cplx> f((size_t)N), g((size_t)N), h((size_t)N);
  for (long p = 0; p < N; p++) f[p] = rho[p];
  fft3d(n, 1, 1, f.data(), g.data(), -1, false);
  auto gcart = [&](int ix, int iy, int iz) {
    vec3 gv{{2.0 * M_PI * symmetric(ix, n[0]), 2.0 * M_PI * symmetric(iy, n[1]),
             2.0 * M_PI * symmetric(iz, n[2])}};
    return c.cov_to_cart(gv);
  };
  std::vector<double> grad[3];
  for (int d = 0; d < 3; d++) {
    grad[d].resize((size_t)N);
    for (int ix = 0; ix < n[0]; ix++)
      for (int iy = 0; iy < n[1]; iy++)
        for (int iz = 0; iz < n[2]; iz++) {
          const long p = ((long)ix * n[1] + iy) * n[2] + iz;
          h[p] = cplx(0.0, gcart(ix, iy, iz)[d]) * g[p] / (double)N;
        }
    fft3d(n, 1, 1, h.data(), f.data(), +1, false);
    for (long p = 0; p < N; p++) grad[d][p] = std::real(f[p]);
  }
  std::vector<double> vs((size_t)N);
  double es = 0.0;
  for (long p = 0; p < N; p++) {
    const double nn = rho[p] < 0.0 ? 0.0 : rho[p];
    const double sigma = grad[0][p] * grad[0][p] + grad[1][p] * grad[1][p] + grad[2][p] * grad[2][p];
    double e = 0.0, vn = 0.0, vsig = 0.0;
    if (which & 1) {
      double e1, v1, s1;
      pbe_x_point(nn, sigma, e1, v1, s1);
      e += e1, vn += v1, vsig += s1;
    }
    if (which & 2) {
      double e1, v1, s1;
      pbe_c_point(nn, sigma, e1, v1, s1);
      e += e1, vn += v1, vsig += s1;
    }
    vxc[p] = vn;
    vs[p] = vsig;
    es += e * nn;
  }
  // div(-2 vs grad n)
  std::vector<cplx> acc((size_t)N, cplx(0.0, 0.0));
  for (int d = 0; d < 3; d++) {
    for (long p = 0; p < N; p++) f[p] = -2.0 * vs[p] * grad[d][p];
    fft3d(n, 1, 1, f.data(), g.data(), -1, false);
    for (int ix = 0; ix < n[0]; ix++)
      for (int iy = 0; iy < n[1]; iy++)
        for (int iz = 0; iz < n[2]; iz++) {
          const long p = ((long)ix * n[1] + iy) * n[2] + iz;
          acc[p] += cplx(0.0, gcart(ix, iy, iz)[d]) * g[p] / (double)N;
        }
  }
  fft3d(n, 1, 1, acc.data(), f.data(), +1, false);
  for (long p = 0; p < N; p++) vxc[p] += std::real(f[p]);
  *exc = c.volume / (double)N * es;
}

/* hamiltonian::xc_term::operator() (src/hamiltonian/xc_term.hpp:120-142) for
 * one spin component and no core density: n = max(rho, 0); vxc[p] = v(n);
 * exc = dV sum e(n) n; nvxc = dV sum rho vxc. */
static void lda_xc(long N, double dV, const double *rho, double *vxc, double *exc, double *nvxc) {
  double es = 0.0, ns = 0.0;
  for (long p = 0; p < N; p++) {
    const double n = rho[p] < 0.0 ? 0.0 : rho[p];
    double e, v;
    lda_point(n, e, v);
    vxc[p] = v;
    es += e * n;
    ns += rho[p] * v;
  }
  *exc = dV * es;
  *nvxc = dV * ns;
}

/* observables::dipole of a density (src/observables/dipole.hpp:24-40):
 * dV sum_p r_cart(p) rho[p], r from the symmetric-range contravariant point
 * coordinates (real_space.hpp:76-95). */
static void dipole(const int n[3], const cell &c, const double *rho, double out[3]) {
  const long N = (long)n[0] * n[1] * n[2];
  const double dV = c.volume / (double)N;
  double s[3] = {0.0, 0.0, 0.0};
  for (int ix = 0; ix < n[0]; ix++)
    for (int iy = 0; iy < n[1]; iy++)
      for (int iz = 0; iz < n[2]; iz++) {
        vec3 rc{{(double)symmetric(ix, n[0]) / n[0], (double)symmetric(iy, n[1]) / n[1],
                 (double)symmetric(iz, n[2]) / n[2]}};
        vec3 r = c.contra_to_cart(rc);
        const double v = rho[((long)ix * n[1] + iy) * n[2] + iz];
        for (int d = 0; d < 3; d++) s[d] += r[d] * v;
      }
  for (int d = 0; d < 3; d++) out[d] = dV * s[d];
}

/* self_consistency::update_hamiltonian (src/hamiltonian/self_consistency.hpp:
 * 88-162) for static ions, one spin, no electric field: V_KS = vion + V_H[rho]
 * (hartree) + V_xc[rho] (xc).  energies: external, hartree, exc, nvxc. */
static void ks_potential(const int n[3], const cell &c, const double *vion, const double *rho,
                         int hartree, int xc, double *vks, double *energies) {
  const long N = (long)n[0] * n[1] * n[2];
  const double dV = c.volume / (double)N;
  double ext = 0.0;
  for (long p = 0; p < N; p++) ext += rho[p] * vion[p];
  energies[0] = dV * ext;
  for (long p = 0; p < N; p++) vks[p] = vion[p];
  energies[1] = 0.0;
  if (hartree) {
    std::vector<double> vh((size_t)N);
    poisson3d(n, c, rho, vh.data());
    double eh = 0.0;
    for (long p = 0; p < N; p++) eh += rho[p] * vh[p];
    energies[1] = 0.5 * dV * eh;
    for (long p = 0; p < N; p++) vks[p] += vh[p];
  }
  energies[2] = energies[3] = 0.0;
  if (xc) {
    // xc 1: LDA (options::theory::lda); 2: PBE (the default theory)
    std::vector<double> vx((size_t)N);
    if (xc == 2) {
      gga_xc(n, c, rho, 3, vx.data(), &energies[2]);
      double s = 0.0;
      for (long p = 0; p < N; p++) s += rho[p] * vx[p];
      energies[3] = dV * s;
    } else {
      lda_xc(N, dV, rho, vx.data(), &energies[2], &energies[3]);
    }
    for (long p = 0; p < N; p++) vks[p] += vx[p];
  }
}

extern "C" {

void orc_set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

int orc_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

int orc_fft3d(const int n[3], int nst, int ld, const double *in, double *out, int sign,
              int normalize) {
  if (!n || n[0] <= 0 || n[1] <= 0 || n[2] <= 0 || nst <= 0 || ld < nst) return 1;
  fft3d(n, nst, ld, reinterpret_cast<const cplx *>(in), reinterpret_cast<cplx *>(out), sign,
        normalize != 0);
  return 0;
}

int orc_fft1d(int n, const double *in, double *out, int sign) {
  if (n <= 0) return 1;
  std::vector<cplx> scratch(2 * max_prime(n) + 2), src(n);
  for (int i = 0; i < n; i++) src[i] = cplx(in[2 * i], in[2 * i + 1]);
  fft1d(n, src.data(), reinterpret_cast<cplx *>(out), sign, scratch.data());
  return 0;
}

int orc_containing_cube(const double lattice[9], const int n[3], const double pos_[3],
                        double radius, int lo[3], int hi[3]) {
  cell c(lattice);
  vec3 pos{{pos_[0], pos_[1], pos_[2]}};
  for (int idir = 0; idir < 3; idir++) {
    vec3 rec = c.rlat[idir];
    double len = std::sqrt(dot(rec, rec));
    vec3 lointer = pos - (radius / len) * rec;
    vec3 hiinter = pos + (radius / len) * rec;
    double dlo = c.cart_to_contra(lointer)[idir];
    double dhi = c.cart_to_contra(hiinter)[idir];
    double spacing = 1.0 / n[idir]; /* grid.hpp:37 */
    lo[idir] = (int)lround(floor(dlo / spacing));
    hi[idir] = (int)lround(ceil(dhi / spacing)) + 1;
    int rb = -n[idir] / 2, re = n[idir] / 2 + n[idir] % 2; /* grid.hpp:142-148 */
    lo[idir] = std::min(std::max(lo[idir], rb), re);
    hi[idir] = std::min(std::max(hi[idir], rb), re);
  }
  return 0;
}

int orc_sphere_build(const double lattice[9], const int n[3], const double center[3],
                     double radius, int capacity, int *coords, float *distance, float *relpos,
                     int *count) {
  cell c(lattice);
  /* periodic_replicas (src/ionic/periodic_replicas.hpp:27-55) with
   * position_in_cell (src/systems/cell.hpp:238-250) */
  vec3 pos{{center[0], center[1], center[2]}};
  vec3 crys = c.cart_to_contra(pos);
  for (int d = 0; d < 3; d++) {
    crys[d] -= floor(crys[d]);
    if (crys[d] >= 0.5) crys[d] -= 1.0;
  }
  pos = c.contra_to_cart(crys);
  double rl[3];
  for (int d = 0; d < 3; d++) rl[d] = std::sqrt(dot(c.lat[d], c.lat[d]));
  double range = std::sqrt(rl[0] * rl[0] + rl[1] * rl[1] + rl[2] * rl[2]); /* grid.hpp:62-64 */
  std::vector<vec3> reps;
  reps.push_back(pos);
  int nm[3];
  for (int d = 0; d < 3; d++) nm[d] = 1 + (int)lround(range / std::sqrt(dot(c.lat[d], c.lat[d])));
  for (int ix = -nm[0]; ix <= nm[0]; ix++)
    for (int iy = -nm[1]; iy <= nm[1]; iy++)
      for (int iz = -nm[2]; iz <= nm[2]; iz++) {
        if (ix == 0 && iy == 0 && iz == 0) continue;
        vec3 r = ((pos + (double)ix * c.lat[0]) + (double)iy * c.lat[1]) + (double)iz * c.lat[2];
        reps.push_back(r);
      }
  /* spherical_grid::initialize (src/basis/spherical_grid.hpp:58-158) */
  int cnt = 0;
  const double r2max = radius * radius;
  for (const vec3 &re : reps) {
    int lo[3], hi[3];
    double rev[3] = {re[0], re[1], re[2]};
    orc_containing_cube(lattice, n, rev, radius, lo, hi);
    int d0 = hi[0] - lo[0], d1 = hi[1] - lo[1], d2 = hi[2] - lo[2];
    if (d0 * d1 * d2 == 0) continue;
    for (int ix = 0; ix < d0; ix++)
      for (int iy = 0; iy < d1; iy++)
        for (int iz = 0; iz < d2; iz++) {
          int ii[3] = {lo[0] + ix, lo[1] + iy, lo[2] + iz};
          for (int d = 0; d < 3; d++) if (ii[d] < 0) ii[d] += n[d]; /* grid.hpp:123-128 */
          /* real_space.hpp:76-79: symmetric index times 1/n, then to cartesian */
          vec3 rc{{symmetric(ii[0], n[0]) * (1.0 / n[0]), symmetric(ii[1], n[1]) * (1.0 / n[1]),
                   symmetric(ii[2], n[2]) * (1.0 / n[2])}};
          vec3 rpoint = c.contra_to_cart(rc);
          vec3 dr = rpoint - re;
          double n2 = dot(dr, dr);
          if (n2 > r2max) continue;
          if (cnt < capacity) {
            coords[3 * cnt + 0] = ii[0];
            coords[3 * cnt + 1] = ii[1];
            coords[3 * cnt + 2] = ii[2];
            distance[cnt] = (float)std::sqrt(n2);
            vec3 rp = c.cart_to_contra(dr);
            relpos[3 * cnt + 0] = (float)rp[0];
            relpos[3 * cnt + 1] = (float)rp[1];
            relpos[3 * cnt + 2] = (float)rp[2];
          }
          cnt++;
        }
  }
  *count = cnt;
  return cnt > capacity ? 2 : 0;
}

int orc_gaussian_projector_matrix(const double lattice[9], int npoint, const float *distance,
                                  const float *relpos, double sigma, int nlm, double *matrix,
                                  int ld) {
  if (nlm != 1 && nlm != 4) return 1;
  cell c(lattice);
  const double y00 = 1.0 / std::sqrt(4.0 * M_PI);
  const double y1 = std::sqrt(3.0 / (4.0 * M_PI));
  for (int i = 0; i < npoint; i++) {
    double d = (double)distance[i];
    double f = std::exp(-d * d / (sigma * sigma));
    vec3 rc{{(double)relpos[3 * i], (double)relpos[3 * i + 1], (double)relpos[3 * i + 2]}};
    vec3 x = c.contra_to_cart(rc);
    double r = std::sqrt(dot(x, x));
    matrix[0 * (long)ld + i] = f * y00;
    if (nlm == 4) {
      double inv = r > 0.0 ? 1.0 / r : 0.0;
      matrix[1 * (long)ld + i] = f * y1 * x[1] * inv; /* m = -1 : y */
      matrix[2 * (long)ld + i] = f * y1 * x[2] * inv; /* m =  0 : z */
      matrix[3 * (long)ld + i] = f * y1 * x[0] * inv; /* m = +1 : x */
    }
  }
  return 0;
}

int orc_hamiltonian_apply(const orc_system *sys, const double kA_cov[3], int nst, int ld,
                          const double *psi, double *hpsi) {
  if (!sys || nst <= 0 || ld < nst) return 1;
  hamiltonian(sys, kA_cov, nst, ld, reinterpret_cast<const cplx *>(psi),
              reinterpret_cast<cplx *>(hpsi));
  return 0;
}

/* energy::calculate (src/hamiltonian/energy.hpp:59-107) for one orbital set:
 * residual = H phi; eps = overlap_diagonal_normalized(residual, phi) (real part);
 * residual -= eps phi; normres = overlap_diagonal(residual) = dV sum |residual|^2;
 * nl = overlap_diagonal_normalized(non_local(phi), phi) (real part). */
int orc_energy(const orc_system *sys, const double kA_cov[3], int nst, int ld, const double *psi_,
               double *eps, double *normres, double *nl) {
  if (!sys || nst <= 0 || ld < nst) return 1;
  const cplx *psi = reinterpret_cast<const cplx *>(psi_);
  cell c(sys->lattice);
  const long N = npoints(sys);
  const double dV = c.volume / (double)N;
  std::vector<cplx> res((size_t)N * ld), vnl((size_t)N * ld, 0.0);
  hamiltonian(sys, kA_cov, nst, ld, psi, res.data());
  if (sys->nproj > 0) {
    std::vector<cplx> sphere((size_t)sys->nproj * sys->max_sphere * nst);
    project(sys, kA_cov, nst, ld, psi, sphere.data());
    apply(sys, kA_cov, nst, ld, sphere.data(), vnl.data());
  }
  for (int b = 0; b < nst; b++) {
    cplx v = 0.0, nn = 0.0, w = 0.0;
    for (long p = 0; p < N; p++) {
      v += std::conj(res[p * ld + b]) * psi[p * ld + b];
      w += std::conj(vnl[p * ld + b]) * psi[p * ld + b];
      nn += std::conj(psi[p * ld + b]) * psi[p * ld + b];
    }
    eps[b] = std::real(v / nn);
    nl[b] = std::real(w / nn);
    double r = 0.0;
    for (long p = 0; p < N; p++) r += std::norm(res[p * ld + b] - eps[b] * psi[p * ld + b]);
    normres[b] = dV * r;
  }
  return 0;
}

int orc_hamiltonian_apply_fourier(const orc_system *sys, const double kA_cov[3], int nst, int ld,
                                  const double *phi, double *hphi) {
  if (!sys || nst <= 0 || ld < nst) return 1;
  hamiltonian_fourier(sys, kA_cov, nst, ld, reinterpret_cast<const cplx *>(phi),
                      reinterpret_cast<cplx *>(hphi));
  return 0;
}

int orc_projector_project(const orc_system *sys, const double kA_cov[3], int nst, int ld,
                          const double *psi, double *sphere) {
  if (!sys || nst <= 0 || ld < nst) return 1;
  project(sys, kA_cov, nst, ld, reinterpret_cast<const cplx *>(psi),
          reinterpret_cast<cplx *>(sphere));
  return 0;
}

int orc_projector_apply(const orc_system *sys, const double kA_cov[3], int nst, int ld,
                        const double *sphere, double *hpsi) {
  if (!sys || nst <= 0 || ld < nst) return 1;
  apply(sys, kA_cov, nst, ld, reinterpret_cast<const cplx *>(sphere),
        reinterpret_cast<cplx *>(hpsi));
  return 0;
}

int orc_band_energies(int npoint, int nst, int ld, const double *hpsi_, const double *psi_,
                      double *eps) {
  const cplx *hpsi = reinterpret_cast<const cplx *>(hpsi_);
  const cplx *psi = reinterpret_cast<const cplx *>(psi_);
  for (int b = 0; b < nst; b++) {
    cplx value = 0.0;
    cplx norm = 0.0;
    for (long p = 0; p < npoint; p++) {
      /* overlap_diagonal.hpp:124-126: {conj(mat1)*mat2, conj(mat2)*mat2} */
      value += std::conj(hpsi[p * ld + b]) * psi[p * ld + b];
      norm += std::conj(psi[p * ld + b]) * psi[p * ld + b];
    }
    eps[b] = std::real(value / norm); /* :167 with real_part */
  }
  return 0;
}

int orc_exponential_in_place(const orc_system *sys, const double kA_cov[3], double zre,
                             double zim, int order, int nst, int ld, double *psi) {
  if (!sys || nst <= 0 || ld < nst || order < 0) return 1;
  exponential_in_place(sys, kA_cov, cplx(zre, zim), order, nst, ld,
                       reinterpret_cast<cplx *>(psi));
  return 0;
}

int orc_exponential_2_for_1(const orc_system *sys, const double kA_cov[3], double z1re,
                            double z1im, double z2re, double z2im, int order, int nst, int ld,
                            double *phi_, double *expphi_) {
  if (!sys || nst <= 0 || ld < nst || order < 0) return 1;
  const long N = npoints(sys);
  cplx *phi = reinterpret_cast<cplx *>(phi_);
  cplx *expphi = reinterpret_cast<cplx *>(expphi_);
  for (long p = 0; p < N; p++)
    for (int b = 0; b < nst; b++) expphi[p * ld + b] = phi[p * ld + b];
  std::vector<cplx> hn((size_t)N * ld), tmp((size_t)N * ld);
  hamiltonian(sys, kA_cov, nst, ld, phi, hn.data());
  cplx c1 = 1.0, c2 = 1.0;
  for (int iter = 1; iter <= order; iter++) {
    if (iter > 1) { hamiltonian(sys, kA_cov, nst, ld, hn.data(), tmp.data()); std::swap(hn, tmp); }
    c1 *= cplx(z1re, z1im) / (double)iter;
    c2 *= cplx(z2re, z2im) / (double)iter;
    shift_add(N, nst, ld, c1, hn.data(), expphi);
    shift_add(N, nst, ld, c2, hn.data(), phi);
  }
  return 0;
}

int orc_randomize(const int n[3], int nst, int index, double *psi_) {
  cplx *psi = reinterpret_cast<cplx *>(psi_);
  const uint64_t N = (uint64_t)n[0] * n[1] * n[2];
  const uint64_t seed = N * (uint64_t)nst + (uint64_t)index;
  const double maxv = 4294967296.0;  /* numeric_limits<uint32_t>::max() + 1 */
  for (uint64_t p = 0; p < N; p++)
    for (int b = 0; b < nst; b++) {
      uint64_t step = (uint64_t)b + (uint64_t)nst * p;
      pcg32 rng(seed);
      rng.discard(step * 2u);
      double re = 2.0 * rng() / maxv - 1.0;
      double im = 2.0 * rng() / maxv - 1.0;
      psi[p * nst + b] = cplx(re, im);
    }
  return 0;
}

int orc_current(const orc_system *sys, const double kA_cov[3], double occ, int nst, int ld,
                const double *psi, int commutator, double J_cov[3], double J_cart[3]) {
  if (!sys || nst <= 0 || ld < nst || commutator < 0 || commutator > 1) return 1;
  current(sys, kA_cov, occ, nst, ld, reinterpret_cast<const cplx *>(psi), commutator, J_cov);
  cell c(sys->lattice);
  vec3 jc{{J_cov[0], J_cov[1], J_cov[2]}};
  vec3 cart = c.cov_to_cart(jc);
  for (int d = 0; d < 3; d++) J_cart[d] = cart[d];
  return 0;
}

int orc_td_current(const orc_system *sys, int ncoef, const double *coef, double omega,
                   const double direction[3], double dt, int nsteps, int order, double occ,
                   int nst, int ld, double *psi_, int commutator, double *loss) {
  if (!sys || nst <= 0 || ld < nst || commutator < 0 || commutator > 1) return 1;
  cell c(sys->lattice);
  cplx *psi = reinterpret_cast<cplx *>(psi_);
  auto A_cov = [&](double t, double out[3]) {
    double f = 0.0;
    for (int j = 0; j < ncoef; j++) f += coef[j] * std::sin((j + 1) * omega * t);
    vec3 cart{{direction[0] * f, direction[1] * f, direction[2] * f}};
    vec3 cov = c.cart_to_cov(cart);
    for (int d = 0; d < 3; d++) out[d] = cov[d];
  };
  double At[3], Atdt[3];
  for (int n = 0; n < nsteps; n++) {
    const double t = n * dt;
    A_cov(t, At);
    A_cov(t + dt, Atdt);
    exponential_in_place(sys, At, cplx(0.0, dt / 2.0), order, nst, ld, psi);
    exponential_in_place(sys, Atdt, cplx(0.0, dt / 2.0), order, nst, ld, psi);
  }
  double AN[3], Jc[3], Jcart[3];
  A_cov(nsteps > 0 ? (nsteps - 1) * dt + dt : 0.0, AN);
  orc_current(sys, AN, occ, nst, ld, psi_, commutator, Jc, Jcart);
  *loss = Jcart[0] * direction[0] + Jcart[1] * direction[1] + Jcart[2] * direction[2];
  return 0;
}

int orc_etrs_step(const orc_system *sys, const double kA_t[3], const double kA_tdt[3], double dt,
                  int order, int nst, int ld, double *psi_) {
  if (!sys || nst <= 0 || ld < nst) return 1;
  cplx *psi = reinterpret_cast<cplx *>(psi_);
  /* etrs.hpp:36: halfstep = T(i dt/2, H(t)) psi, (full step only feeds the
   * density, which a frozen potential ignores); etrs.hpp:57-67: restore the
   * half step and apply T(i dt/2, H(t+dt)). */
  exponential_in_place(sys, kA_t, cplx(0.0, dt / 2.0), order, nst, ld, psi);
  exponential_in_place(sys, kA_tdt, cplx(0.0, dt / 2.0), order, nst, ld, psi);
  return 0;
}

int orc_laplacian_add(const orc_system *sys, double factor, const double gradcoeff_contra[3],
                      int nst, int ld, const double *ff, double *laplff) {
  if (!sys || nst <= 0 || ld < nst) return 1;
  cell c(sys->lattice);
  laplacian_add(sys->n, c, factor, gradcoeff_contra, nst, ld, reinterpret_cast<const cplx *>(ff),
                reinterpret_cast<cplx *>(laplff));
  return 0;
}

int orc_preconditioner(const orc_system *sys, int nst, int ld, double *phi) {
  if (!sys || nst <= 0 || ld < nst) return 1;
  cell c(sys->lattice);
  precondition(sys->n, c, nst, ld, reinterpret_cast<cplx *>(phi));
  return 0;
}

int orc_orthogonalize(long npoint, int nst, int ld, double dV, double *phi) {
  if (npoint <= 0 || nst <= 0 || ld < nst) return 1;
  orthogonalize(npoint, nst, ld, dV, reinterpret_cast<cplx *>(phi));
  return 0;
}

int orc_steepest_descent(const orc_system *sys, const double kA_cov[3], int precond, int nst,
                         int ld, double *phi_) {
  if (!sys || nst <= 0 || ld < nst) return 1;
  cell c(sys->lattice);
  const long N = npoints(sys);
  /* fourier_space::volume_element (fourier_space.hpp:41-44) */
  const double dV = c.volume / ((double)N * (double)N);
  auto ham = [&](const cplx *in, cplx *out) { hamiltonian_fourier(sys, kA_cov, nst, ld, in, out); };
  auto prec = [&](cplx *r) {
    if (precond) precondition(sys->n, c, nst, ld, r);
  };
  steepest_descent(N, nst, ld, dV, ham, prec, reinterpret_cast<cplx *>(phi_));
  return 0;
}

int orc_steepest_descent_dense(int npoint, int nvec, const double *matrix_, double dV, int niter,
                               double *phi_, double *eigenvalues_) {
  if (npoint <= 0 || nvec <= 0 || niter < 0) return 1;
  const cplx *M = reinterpret_cast<const cplx *>(matrix_);
  cplx *phi = reinterpret_cast<cplx *>(phi_);
  cplx *ev = reinterpret_cast<cplx *>(eigenvalues_);
  /* operations::matrix_operator (matrix_operator.hpp:32-43): M phi */
  auto ham = [&](const cplx *in, cplx *out) {
    for (int i = 0; i < npoint; i++)
      for (int b = 0; b < nvec; b++) {
        cplx v = 0.0;
        for (int j = 0; j < npoint; j++) v += M[(size_t)i * npoint + j] * in[(size_t)j * nvec + b];
        out[(size_t)i * nvec + b] = v;
      }
  };
  auto prec = [](cplx *) {};
  for (int it = 0; it < niter; it++) steepest_descent(npoint, nvec, nvec, dV, ham, prec, phi);
  std::vector<cplx> mphi((size_t)npoint * nvec);
  ham(phi, mphi.data());
  overlap_diag(npoint, nvec, nvec, dV, phi, mphi.data(), ev);
  return 0;
}

int orc_density(long npoint, int nst, int ld, const double *occ, const double *psi, double *rho) {
  if (npoint <= 0 || nst <= 0 || ld < nst) return 1;
  for (long p = 0; p < npoint; p++) rho[p] = 0.0;
  density_add(npoint, nst, ld, occ, reinterpret_cast<const cplx *>(psi), rho);
  return 0;
}

int orc_poisson(const orc_system *sys, const double *rho, double *vh) {
  if (!sys) return 1;
  cell c(sys->lattice);
  poisson3d(sys->n, c, rho, vh);
  return 0;
}

int orc_lda_xc(long npoint, double dV, const double *rho, double *vxc, double *exc, double *nvxc) {
  if (npoint <= 0) return 1;
  lda_xc(npoint, dV, rho, vxc, exc, nvxc);
  return 0;
}

int orc_xc(const orc_system *sys, const double *rho, int kind, double *vxc, double *exc) {
  if (!sys || kind < 1 || kind > 5) return 1;
  cell c(sys->lattice);
  const long N = (long)sys->n[0] * sys->n[1] * sys->n[2];
  if (kind >= 3) {
    gga_xc(sys->n, c, rho, kind - 2, vxc, exc);
    return 0;
  }
  double es = 0.0;
  for (long p = 0; p < N; p++) {
    const double nn = rho[p] < 0.0 ? 0.0 : rho[p];
    double e, v;
    if (kind == 1)
      lda_x_point(nn, e, v);
    else
      lda_point(nn, e, v);
    vxc[p] = v;
    es += e * nn;
  }
  *exc = c.volume / (double)N * es;
  return 0;
}

int orc_dipole(const orc_system *sys, const double *rho, double *out) {
  if (!sys) return 1;
  cell c(sys->lattice);
  dipole(sys->n, c, rho, out);
  return 0;
}

int orc_ks_potential(const orc_system *sys, const double *vion, const double *rho, int hartree,
                     int xc, double *vks, double *energies) {
  if (!sys) return 1;
  cell c(sys->lattice);
  ks_potential(sys->n, c, vion, rho, hartree, xc, vks, energies);
  return 0;
}

/* real_time::etrs with the self-consistent second half (src/real_time/etrs.hpp:
 * 22-82), static ions, no induced vector potential.  sys->vloc is V_KS of H(t);
 * on return vks holds V_KS of H(t+dt) and *nscf the SCF passes run. */
int orc_etrs_scf(const orc_system *sys, const double *vion, const double *occ, const double kA_t[3],
                 const double kA_tdt[3], double dt, int order, int hartree, int xc, int nst, int ld,
                 double *psi_, double *vks, int *nscf) {
  if (!sys || nst <= 0 || ld < nst) return 1;
  cplx *psi = reinterpret_cast<cplx *>(psi_);
  cell c(sys->lattice);
  const long N = npoints(sys);
  const double dV = c.volume / (double)N;
  const int nscf_max = 5;
  const double scf_threshold = 5e-5;
  orc_system s = *sys;
  std::vector<double> V(sys->vloc, sys->vloc + N);
  s.vloc = V.data();
  /* :33-38 halfstep_phi = exponential_2_for_1(ham, i dt/2, i dt, phi): save
   * holds the half step, psi the full step */
  std::vector<cplx> save((size_t)N * ld);
  orc_exponential_2_for_1(&s, kA_t, 0.0, dt / 2.0, 0.0, dt, order, nst, ld, psi_,
                          reinterpret_cast<double *>(save.data()));
  /* :44 density of the full step */
  std::vector<double> rho((size_t)N, 0.0), old((size_t)N);
  density_add(N, nst, ld, occ, psi, rho.data());
  /* :53 update_hamiltonian(ham, ..., time + dt) */
  double en[4];
  ks_potential(sys->n, c, vion, rho.data(), hartree, xc, V.data(), en);
  /* :56-58 restore the half step */
  for (long p = 0; p < N; p++)
    for (int b = 0; b < nst; b++) psi[p * ld + b] = save[p * ld + b];
  /* :61-80 */
  int iscf = 0;
  for (; iscf < nscf_max; iscf++) {
    if (iscf != 0)
      for (long p = 0; p < N; p++)
        for (int b = 0; b < nst; b++) psi[p * ld + b] = save[p * ld + b];
    exponential_in_place(&s, kA_tdt, cplx(0.0, dt / 2.0), order, nst, ld, psi);
    old = rho;
    std::fill(rho.begin(), rho.end(), 0.0);
    density_add(N, nst, ld, occ, psi, rho.data());
    double delta = 0.0;
    for (long p = 0; p < N; p++) delta += std::fabs(old[p] - rho[p]);
    delta *= dV;
    const bool done = delta < scf_threshold || iscf == nscf_max - 1;
    ks_potential(sys->n, c, vion, rho.data(), hartree, xc, V.data(), en);
    if (done) break;
  }
  for (long p = 0; p < N; p++) vks[p] = V[p];
  *nscf = iscf + 1;
  return 0;
}

} // extern "C"
