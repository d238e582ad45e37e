// Host-side builders of the projector geometry: containing_cube,
// periodic_replicas and spherical_grid (src/basis/containing_cube.hpp,
// src/ionic/periodic_replicas.hpp, src/basis/spherical_grid.hpp) and the
// Gaussian projector table of projector::build (src/hamiltonian/projector.hpp).
//
// The sphere geometry is FP32 in the reference (spherical_grid.hpp:40-43);
// this file evaluates it in the reference's operation order and is compiled
// without floating-point contraction so the FP32 values are reproducible.

#include "inq_amd.h"

#include <cmath>
#include <vector>

namespace {

struct metric {
  double a[3][3];  // lattice rows
  double b[3][3];  // reciprocal rows (with 2 pi)

  explicit metric(const double *lat) {
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) a[i][j] = lat[3 * i + j];
    double c[3][3];
    for (int i = 0; i < 3; i++) {
      const double *u = a[(i + 1) % 3], *v = a[(i + 2) % 3];
      c[i][0] = u[1] * v[2] - u[2] * v[1];
      c[i][1] = u[2] * v[0] - u[0] * v[2];
      c[i][2] = u[0] * v[1] - u[1] * v[0];
    }
    const double vol = a[0][0] * c[0][0] + a[0][1] * c[0][1] + a[0][2] * c[0][2];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) b[i][j] = 2.0 * M_PI / vol * c[i][j];
  }
  // cell_metric::to_contravariant(cartesian) (cell.hpp:204-207)
  void to_contra(const double w[3], double out[3]) const {
    for (int i = 0; i < 3; i++)
      out[i] = (b[i][0] * w[0] + b[i][1] * w[1] + b[i][2] * w[2]) / (2.0 * M_PI);
  }
  // cell_metric::to_cartesian(contravariant) (cell.hpp:214-217)
  void to_cart(const double v[3], double out[3]) const {
    for (int j = 0; j < 3; j++) out[j] = (a[0][j] * v[0] + a[1][j] * v[1]) + a[2][j] * v[2];
  }
};

inline int sym(int i, int n) { return i >= (n + 1) / 2 ? i - n : i; }

void cube(const metric &m, const int n[3], const double pos[3], double radius, int lo[3],
          int hi[3]) {
  for (int d = 0; d < 3; d++) {
    const double *rec = m.b[d];
    const double len = std::sqrt(rec[0] * rec[0] + rec[1] * rec[1] + rec[2] * rec[2]);
    const double f = radius / len;
    double lw[3], hw[3], lc[3], hc[3];
    for (int j = 0; j < 3; j++) {
      lw[j] = pos[j] - f * rec[j];
      hw[j] = pos[j] + f * rec[j];
    }
    m.to_contra(lw, lc);
    m.to_contra(hw, hc);
    const double spacing = 1.0 / n[d];
    int l = (int)std::lround(std::floor(lc[d] / spacing));
    int h = (int)std::lround(std::ceil(hc[d] / spacing)) + 1;
    const int rb = -n[d] / 2, re = n[d] / 2 + n[d] % 2;
    lo[d] = l < rb ? rb : (l > re ? re : l);
    hi[d] = h < rb ? rb : (h > re ? re : h);
  }
}

}  // namespace

extern "C" {

int inqamd_containing_cube(const double lattice[9], const int n[3], const double pos[3],
                           double radius, int lo[3], int hi[3]) {
  if (!lattice || !n || !pos || !lo || !hi || radius < 0.0) return INQAMD_ERR_ARGUMENT;
  metric m(lattice);
  cube(m, n, pos, radius, lo, hi);
  return INQAMD_OK;
}

int inqamd_sphere_build(const double lattice[9], const int n[3], const double center[3],
                        double radius, int capacity, int *coords, float *distance, float *relpos,
                        int *count) {
  if (!lattice || !n || !center || !count || radius < 0.0) return INQAMD_ERR_ARGUMENT;
  if (capacity > 0 && (!coords || !distance || !relpos)) return INQAMD_ERR_ARGUMENT;
  metric m(lattice);
  // periodic_replicas: the centre folded into [-1/2, 1/2) (cell.hpp:238-250)
  double crys[3], home[3];
  m.to_contra(center, crys);
  for (int d = 0; d < 3; d++) {
    crys[d] -= std::floor(crys[d]);
    if (crys[d] >= 0.5) crys[d] -= 1.0;
  }
  m.to_cart(crys, home);
  double alen[3];
  for (int d = 0; d < 3; d++)
    alen[d] = std::sqrt(m.a[d][0] * m.a[d][0] + m.a[d][1] * m.a[d][1] + m.a[d][2] * m.a[d][2]);
  const double range = std::sqrt(alen[0] * alen[0] + alen[1] * alen[1] + alen[2] * alen[2]);
  int nm[3];
  for (int d = 0; d < 3; d++) nm[d] = 1 + (int)std::lround(range / alen[d]);
  std::vector<double> reps;  // xyz triplets, home first (periodic_replicas.hpp:35-53)
  reps.insert(reps.end(), home, home + 3);
  for (int i = -nm[0]; i <= nm[0]; i++)
    for (int j = -nm[1]; j <= nm[1]; j++)
      for (int k = -nm[2]; k <= nm[2]; k++) {
        if (i == 0 && j == 0 && k == 0) continue;
        for (int c = 0; c < 3; c++)
          reps.push_back(((home[c] + (double)i * m.a[0][c]) + (double)j * m.a[1][c]) +
                         (double)k * m.a[2][c]);
      }
  const double r2 = radius * radius;
  int cnt = 0;
  for (size_t rp = 0; rp < reps.size(); rp += 3) {
    const double *re = &reps[rp];
    int lo[3], hi[3];
    cube(m, n, re, radius, lo, hi);
    if ((hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]) == 0) continue;
    for (int x = lo[0]; x < hi[0]; x++)
      for (int y = lo[1]; y < hi[1]; y++)
        for (int z = lo[2]; z < hi[2]; z++) {
          const int g[3] = {x < 0 ? x + n[0] : x, y < 0 ? y + n[1] : y, z < 0 ? z + n[2] : z};
          const double rc[3] = {sym(g[0], n[0]) * (1.0 / n[0]), sym(g[1], n[1]) * (1.0 / n[1]),
                                sym(g[2], n[2]) * (1.0 / n[2])};
          double rpoint[3], dr[3];
          m.to_cart(rc, rpoint);
          for (int c = 0; c < 3; c++) dr[c] = rpoint[c] - re[c];
          const double d2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
          if (d2 > r2) continue;
          if (cnt < capacity) {
            double rel[3];
            m.to_contra(dr, rel);
            for (int c = 0; c < 3; c++) {
              coords[3 * cnt + c] = g[c];
              relpos[3 * cnt + c] = (float)rel[c];
            }
            distance[cnt] = (float)std::sqrt(d2);
          }
          cnt++;
        }
  }
  *count = cnt;
  return cnt > capacity ? INQAMD_ERR_MEMORY : INQAMD_OK;
}

int inqamd_gaussian_projector_matrix(const double lattice[9], int npoint, const float *distance,
                                     const float *relpos, double sigma, int nlm, double *matrix,
                                     int ld) {
  if (!lattice || (npoint > 0 && (!distance || !relpos || !matrix)) || sigma <= 0.0 || ld < npoint)
    return INQAMD_ERR_ARGUMENT;
  if (nlm != 1 && nlm != 4) return INQAMD_ERR_UNSUPPORTED;
  metric m(lattice);
  const double y00 = 1.0 / std::sqrt(4.0 * M_PI);
  const double y1 = std::sqrt(3.0 / (4.0 * M_PI));
  for (int i = 0; i < npoint; i++) {
    const double d = (double)distance[i];
    const double f = std::exp(-d * d / (sigma * sigma));
    const double rc[3] = {(double)relpos[3 * i], (double)relpos[3 * i + 1],
                          (double)relpos[3 * i + 2]};
    double x[3];
    m.to_cart(rc, x);
    const double r = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    matrix[i] = f * y00;
    if (nlm == 4) {
      const double inv = r > 0.0 ? 1.0 / r : 0.0;
      matrix[1L * ld + i] = f * y1 * x[1] * inv;
      matrix[2L * ld + i] = f * y1 * x[2] * inv;
      matrix[3L * ld + i] = f * y1 * x[0] * inv;
    }
  }
  return INQAMD_OK;
}

}  // extern "C"
