// Internal state of a Hamiltonian handle and the launchers shared between
// the translation units of the library.
#pragma once

#include "common.hpp"
#include "fft.hpp"

#include <string>
#include <vector>

namespace inqamd {

// Device-resident k + A slots (covariant).  Kernels read k+A from device
// memory so a captured graph can change it between replays.
enum ka_slot : int { KA_CURRENT = 0, KA_T = 1, KA_TDT = 2, KA_NSLOTS = 4 };

struct projector_tables {
  int nproj = 0;
  int max_sphere = 0;
  int max_nlm = 0;
  int *point = nullptr;    // [nproj][max_sphere] linear grid index, -1 padded
  float *pos = nullptr;    // [nproj][max_sphere][3] contravariant, FP32
  double *mat = nullptr;   // [nproj][max_nlm][max_sphere], zero padded
  double *coeff = nullptr; // [nproj][max_nlm], zero where lm >= nlm
  int *size = nullptr;     // [nproj] points (0 for locally empty spheres)
  // ownership map of the scatter: unique grid points and their contributors
  int nunique = 0;
  int *upoint = nullptr;   // [nunique] linear grid index
  int *uoff = nullptr;     // [nunique + 1]
  int *uent = nullptr;     // [sum K] p * max_sphere + i, sorted by (point, p, i)
};

}  // namespace inqamd

struct inqamd_hamiltonian {
  int n[3] = {0, 0, 0};
  long npoint = 0;
  double lattice[9] = {};
  double rlat[9] = {};
  double volume = 0.0;
  hipStream_t stream = nullptr;
  int max_bands = 0;

  inqamd::fft_axis_plan axis[3];
  size_t max_lds = 65536;       // opt-in LDS per workgroup on this device
  bool three_sweep = false;     // cubic grid with fft3 kernels available

  double *V = nullptr;          // [N]
  double *kA = nullptr;         // [KA_NSLOTS][3]
  double k_cov[3] = {0, 0, 0};
  double A_cov[3] = {0, 0, 0};

  inqamd::cplx *work = nullptr; // [N][max_bands] FFT intermediate
  inqamd::cplx *hn = nullptr;   // [N][max_bands] Taylor H^n psi
  inqamd::cplx *hn2 = nullptr;  // [N][max_bands] Taylor ping-pong
  inqamd::cplx *proj = nullptr; // [nproj][max_nlm][max_bands]
  double *partial = nullptr;    // reduction partials
  long npartial = 0;

  inqamd::projector_tables pt;

  std::string last_error;
};

namespace inqamd {

using handle = inqamd_hamiltonian;

// projector.hip
void projector_upload(handle &h, int nproj, int max_sphere, int max_nlm, const int *points,
                      const float *positions, const double *matrices, const double *coeff,
                      const int *nlm, const int *sphere_size);
void projector_free(handle &h);
// P[p][lm][b] = dV * D[p][lm] * sum_i M[p][lm][i] e^{i(k+A).r_pi} psi(r_pi, b)
void projector_project_coefficients(handle &h, const cplx *psi, long ld, int nst, int slot);
// hpsi(r, b) += sum_{(p,i) at r} e^{-i(k+A).r_pi} sum_lm M[p][lm][i] P[p][lm][b]
void projector_apply_coefficients(handle &h, cplx *hpsi, long ld, int nst, int slot);
// sphere[p][i][b] = sum_lm M[p][lm][i] P[p][lm][b]   (INQ-shaped project output)
void projector_expand(handle &h, cplx *sphere, int nst);
// hpsi(r, b) += sum_{(p,i) at r} e^{-i(k+A).r_pi} sphere[p][i][b]
void projector_apply_sphere(handle &h, const cplx *sphere, cplx *hpsi, long ld, int nst, int slot);

}  // namespace inqamd
