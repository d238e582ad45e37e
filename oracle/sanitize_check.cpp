/*
 * sanitize_check.cpp -- TEST INFRASTRUCTURE ONLY.
 *
 * Host-side driver built with AddressSanitizer and UndefinedBehaviorSanitizer
 * (oracle/Makefile, target `sanitize`): it runs the library's host geometry
 * builders (inq-wannier-fork_amd/csrc/sphere.cpp) and the oracle's functions
 * on small systems -- spheres crossing the cell boundary, a non-orthogonal
 * cell, prime-length FFT axes, projector tables with padded rows -- so that
 * any out-of-bounds access, leak or undefined operation in them stops the
 * run with a report.  Prints SANITIZE-OK at the end.
 */
#include "oracle.h"
#include "inq_amd.h"

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

namespace {

int failures = 0;

void check(bool ok, const char *what) {
  if (!ok) {
    std::printf("FAILED: %s\n", what);
    failures++;
  }
}

struct tables {
  int nproj = 0, max_sphere = 0, max_nlm = 4;
  std::vector<int> points, nlm, size;
  std::vector<float> pos;
  std::vector<double> mat, coeff;
};

// Projector tables of Gaussian projectors on the given atoms, built by the
// library's host builders (sphere.cpp) exactly as the Python side does.
tables build_tables(const double lat[9], const int n[3], const std::vector<double> &atoms_frac,
                    double radius, double sigma) {
  tables t;
  t.nproj = (int)atoms_frac.size() / 3;
  std::vector<std::vector<int>> co(t.nproj);
  std::vector<std::vector<float>> di(t.nproj), rp(t.nproj);
  for (int a = 0; a < t.nproj; a++) {
    double c[3];
    for (int d = 0; d < 3; d++)
      c[d] = atoms_frac[3 * a] * lat[0 * 3 + d] + atoms_frac[3 * a + 1] * lat[1 * 3 + d] +
             atoms_frac[3 * a + 2] * lat[2 * 3 + d];
    int count = 0;
    // capacity 0: only counts (INQAMD_ERR_MEMORY when the sphere holds points)
    const int rc = inqamd_sphere_build(lat, n, c, radius, 0, nullptr, nullptr, nullptr, &count);
    check(rc == INQAMD_OK || rc == INQAMD_ERR_MEMORY, "sphere count");
    co[a].resize(3 * (size_t)count);
    di[a].resize(count);
    rp[a].resize(3 * (size_t)count);
    int again = 0;
    check(inqamd_sphere_build(lat, n, c, radius, count, co[a].data(), di[a].data(), rp[a].data(),
                              &again) == 0 && again == count,
          "sphere build");
    if (count > t.max_sphere) t.max_sphere = count;
  }
  const int K = t.max_sphere;
  t.points.assign((size_t)t.nproj * K * 3, -1);
  t.pos.assign((size_t)t.nproj * K * 3, 0.0f);
  t.mat.assign((size_t)t.nproj * t.max_nlm * K, 0.0);
  t.coeff.assign((size_t)t.nproj * t.max_nlm, 0.0);
  t.nlm.assign(t.nproj, t.max_nlm);
  t.size.assign(t.nproj, 0);
  std::mt19937 rng(7);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  for (int a = 0; a < t.nproj; a++) {
    const int cnt = (int)di[a].size();
    t.size[a] = cnt;
    for (int i = 0; i < 3 * cnt; i++) {
      t.points[(size_t)a * K * 3 + i] = co[a][i];
      t.pos[(size_t)a * K * 3 + i] = rp[a][i];
    }
    check(inqamd_gaussian_projector_matrix(lat, cnt, di[a].data(), rp[a].data(), sigma, t.max_nlm,
                                           t.mat.data() + (size_t)a * t.max_nlm * K, K) == 0,
          "projector matrix");
    for (int lm = 0; lm < t.max_nlm; lm++) t.coeff[(size_t)a * t.max_nlm + lm] = u(rng);
  }
  return t;
}

orc_system make_system(const int n[3], const double lat[9], const std::vector<double> &v,
                       const tables &t) {
  orc_system s{};
  for (int d = 0; d < 3; d++) s.n[d] = n[d];
  for (int i = 0; i < 9; i++) s.lattice[i] = lat[i];
  s.vloc = v.data();
  s.nproj = t.nproj;
  s.max_sphere = t.max_sphere;
  s.max_nlm = t.max_nlm;
  s.points = t.points.data();
  s.positions = t.pos.data();
  s.matrices = t.mat.data();
  s.coeff = t.coeff.data();
  s.nlm = t.nlm.data();
  s.sphere_size = t.size.data();
  return s;
}

void run_case(const int n[3], const double lat[9], const std::vector<double> &atoms, int nst) {
  const long N = (long)n[0] * n[1] * n[2];
  const tables t = build_tables(lat, n, atoms, 1.8, 0.9);
  std::vector<double> v(N);
  std::mt19937 rng(11);
  std::normal_distribution<double> g(0.0, 1.0);
  for (auto &x : v) x = -0.5 + 0.1 * g(rng);
  const orc_system s = make_system(n, lat, v, t);
  std::vector<double> psi(2 * N * nst), out(2 * N * nst), tmp(2 * N * nst);
  for (auto &x : psi) x = g(rng);
  const double kA[3] = {0.05, -0.03, 0.02};
  check(orc_fft3d(n, nst, nst, psi.data(), tmp.data(), -1, 0) == 0, "fft forward");
  check(orc_fft3d(n, nst, nst, tmp.data(), out.data(), +1, 1) == 0, "fft backward");
  double err = 0.0, ref = 0.0;
  for (long i = 0; i < 2 * N * nst; i++) {
    err += (out[i] - psi[i]) * (out[i] - psi[i]);
    ref += psi[i] * psi[i];
  }
  check(std::sqrt(err / ref) < 1e-13, "fft round trip");
  check(orc_hamiltonian_apply(&s, kA, nst, nst, psi.data(), out.data()) == 0, "hamiltonian");
  check(orc_hamiltonian_apply_fourier(&s, kA, nst, nst, tmp.data(), out.data()) == 0,
        "hamiltonian fourier");
  std::vector<double> sph(2 * (size_t)t.nproj * t.max_sphere * nst);
  check(orc_projector_project(&s, kA, nst, nst, psi.data(), sph.data()) == 0, "project");
  check(orc_projector_apply(&s, kA, nst, nst, sph.data(), out.data()) == 0, "apply");
  std::vector<double> eps(nst), nr(nst), nl(nst);
  check(orc_energy(&s, kA, nst, nst, psi.data(), eps.data(), nr.data(), nl.data()) == 0, "energy");
  double jc[3], jx[3];
  for (int mode = 0; mode < 2; mode++)
    check(orc_current(&s, kA, 2.0, nst, nst, psi.data(), mode, jc, jx) == 0, "current");
  tmp = psi;
  check(orc_etrs_step(&s, kA, kA, 0.01, 4, nst, nst, tmp.data()) == 0, "etrs step");
  const double coef[3] = {0.01, -0.02, 0.015};
  const double dir[3] = {1.0, 0.0, 0.0};
  double loss = 0.0;
  tmp = psi;
  check(orc_td_current(&s, 3, coef, 0.1, dir, 0.01, 2, 4, 2.0, nst, nst, tmp.data(), 0, &loss) == 0,
        "td current");
  std::vector<double> rho(N), vh(N), vxc(N), occ(nst, 2.0);
  check(orc_density(N, nst, nst, occ.data(), psi.data(), rho.data()) == 0, "density");
  check(orc_poisson(&s, rho.data(), vh.data()) == 0, "poisson");
  double exc = 0.0;
  for (int kind = 1; kind <= 5; kind++) check(orc_xc(&s, rho.data(), kind, vxc.data(), &exc) == 0, "xc");
  double en[4];
  std::vector<double> vks(N);
  check(orc_ks_potential(&s, v.data(), rho.data(), 1, 2, vks.data(), en) == 0, "ks potential");
  int lo[3], hi[3];
  const double c0[3] = {0.1, 0.2, 0.3};
  check(inqamd_containing_cube(lat, n, c0, 1.8, lo, hi) == 0, "containing cube");
  check(orc_containing_cube(lat, n, c0, 1.8, lo, hi) == 0, "oracle containing cube");
}

}  // namespace

int main() {
  {  // orthorhombic cell, atoms on and across the cell boundary, prime axis
    const int n[3] = {12, 13, 10};
    const double lat[9] = {6.0, 0, 0, 0, 6.5, 0, 0, 0, 5.0};
    run_case(n, lat, {0.0, 0.0, 0.0, 0.99, 0.5, 0.02, 0.3, 0.97, 0.5}, 3);
  }
  {  // non-orthogonal cell
    const int n[3] = {10, 10, 12};
    const double lat[9] = {7.0, 0.0, 0.0, 0.3, 6.5, 0.0, 0.0, -0.4, 7.2};
    run_case(n, lat, {0.03, 0.01, -0.04, 0.41, -0.32, 0.22}, 2);
  }
  if (failures) {
    std::printf("%d checks failed\n", failures);
    return 1;
  }
  std::printf("SANITIZE-OK\n");
  return 0;
}
