// Batched complex128 3-D FFT over INQ's [x][y][z][band] layout, built from
// one-axis "pencil" passes.  Each pass stages a tile of `len` points x B
// contiguous bands in LDS, runs a mixed-radix Stockham FFT there and writes
// the tile back, optionally fused with the kinetic multiply (fwd, x T(G), bwd
// in one pass) or with the Hpsi epilogue (x 1/N + V psi).
#pragma once

#include "common.hpp"

#include <vector>

namespace inqamd {

constexpr int FFT_MAX_STAGES = 12;
constexpr int FFT_MAX_GENERIC_RADIX = 31;  // largest radix with a register-resident butterfly
constexpr int FFT_BLOCK = 256;

enum fft_mode : int { FFT_PLAIN = 0, FFT_KINETIC = 1, FFT_EPILOGUE = 2 };

// Ownership-map word of a grid point: first entry << OWN_COUNT_BITS | entry count.
constexpr int OWN_COUNT_BITS = 8;
constexpr int OWN_MAX_COUNT = (1 << OWN_COUNT_BITS) - 1;

struct fft_axis_plan {
  int len = 0;
  int nstages = 0;
  int radix[FFT_MAX_STAGES] = {};
  cplx *tw = nullptr;  // device: W_len^k = exp(-2 pi i k / len), k < len
};

// Arguments of one pencil pass (passed by value to the kernel).
struct fft_pass_args {
  const cplx *in;
  long in_ld;
  cplx *out;
  long out_ld;
  int nst;
  int len;
  long stride;  // point stride along the transformed axis
  int c1;       // other axis 1 (fastest varying over workgroups)
  long s1;
  int c2;
  long s2;
  int nstages;
  int radix[FFT_MAX_STAGES];
  const cplx *tw;
  double scale;
  int sign;
  // FFT_KINETIC: element (i, o1, o2) sits at grid coords c[ax]=i, c[ax1]=o1, c[ax2]=o2
  int n[3];
  int ax, ax1, ax2;
  double rlat[9];      // reciprocal lattice rows (include 2 pi)
  const double *kA;    // device, covariant k + A
  // FFT_EPILOGUE: out = scale * fft + V[p] * psi[p][b]
  const double *V;
  const cplx *psi;
  long psi_ld;
};

// Factor len into radices (8, 4, 2, 3, 5, then the remaining primes; those
// above FFT_MAX_GENERIC_RADIX run a direct O(R^2) stage from LDS).
// Returns false if that takes more than FFT_MAX_STAGES stages.
bool fft_factor(int len, std::vector<int> &radices);

// Host twiddle table exp(-2 pi i k/len), evaluated in long double.
std::vector<cplx> fft_twiddles(int len);

// Band tile chosen for a given nst and axis length.
int fft_band_tile(int nst, int len);

// Launch one pencil pass of the given mode on `stream`.
void fft_launch_pass(const fft_pass_args &a, int mode, hipStream_t stream);

// Register-resident pass (fft_reg.hip): plans with radices in {2,3,4,5,8},
// len <= 128.  `variant` selects the (band tile, threads) shape.
bool fft_reg_supported(const fft_pass_args &a);
void fft_reg_launch(const fft_pass_args &a, int mode, int variant, hipStream_t stream);

// Plan-specialised pass (fft_plan.hip) for the tabulated lengths and band
// tiles B in {8, 16}; returns false when (len, B) has no instantiation.
bool fft_plan_launch(const fft_pass_args &a, int mode, int B, hipStream_t stream);

// Arguments of the three-sweep local Hamiltonian (fft3.hip) on a band chunk
// of a cubic L^3 grid.
struct fft3_args {
  const cplx *psi;   // chunk's first band, INQ layout, leading dimension ld
  long ld;
  cplx *hpsi;        // output, leading dimension hld
  long hld;
  cplx *W;           // band-major intermediate [nb][L][L][L]
  int nb;
  // psi, hpsi, Qe and the accumulators hold chunk band cb at offset
  // cb + (cb >> (3 + glog)) * gskip: groups of 8 << glog bands
  // (gskip 0: the chunk's bands are contiguous)
  int gskip, glog;
  const double *V;   // [N]
  const double *kA;  // device covariant k + A
  double rlat[9];
  const cplx *tw;    // W_L^k table
  double scale;      // 1/N
  // projector apply fused into the last sweep (own == nullptr: not fused):
  // hpsi(p, b) += sum_{j < cnt} Qe[e0 + j][b] with own[p] = (e0 << OWN_COUNT_BITS) | cnt
  // (own[p] = -1 where no sphere reaches)
  const int *own;
  const cplx *Qe;        // [nent][ldQ], offset to the chunk's first band
  long ldQ;
  // or (Pc != nullptr, Qe unused) V_nl psi expanded at the owned points
  // from the projector coefficients: hpsi(p, b) += sum_{j < cnt}
  // phase_ent[e] sum_{lm < nlmP} ent_mat[e][lm] Pc[(ent_sph[e] nlmP + lm) ldPc + b],
  // e = e0 + j (Pc offset to the chunk's first band; ent_sph[e0] = own_sph[p])
  const cplx *Pc;
  long ldPc;
  int nlmP;
  const int *ent_sph;     // [nent] sphere of each ownership entry
  const int *own_sph;     // [N] sphere of a point's first entry (0 where unowned)
  const double *ent_mat;  // [nent][4] projector values at the entry's point
  const cplx *phase_ent;  // [nent] e^{-i(k+A).r} of the call
  // Taylor accumulations acc += c * hpsi (p == nullptr: unused), chunk-offset
  cplx *acc1;
  long acc1_ld;
  double c1re, c1im;
  int acc1_init;  // acc1 = psi + c1 * hpsi instead of acc1 += c1 * hpsi
  cplx *acc2;
  long acc2_ld;
  double c2re, c2im;
  // optional copy of psi's forward transform (before the kinetic factor),
  // band-major [nb][kz][kx][ky] like W, chunk-offset (nullptr: none)
  cplx *gdump;
  // hpsi not written at all (the last term of a Taylor chain: only the
  // accumulators need it)
  int skip_store;
  // launch shapes (0: the process default): z tiles of z_ty y values, x-y
  // planes with plane_threads(L) (xy_wide 1) or twice as many (2)
  int z_ty, xy_wide;
  // G-space Hpsi sweeps: G-space orbitals whose 1/2 |G+k+A|^2 multiple the
  // last sweep adds (fft3_fourier_hamiltonian)
  const cplx *kin_in;
  long kin_ld;
};

// G-space Hpsi on the chunk's bands from psi = IFFT(phi) (INQ layout, `psi`)
// and phi itself (`kin_in`): hpsi = FFT(V psi + V_nl psi) + 1/2 |G+k+A|^2 phi,
// V_nl psi from the ownership entries Qe, in three sweeps (z forward of the
// potential terms, x-y forward, a transposing write that adds the kinetic
// term).
void fft3_fourier_hamiltonian(const fft3_args &a, int L, hipStream_t stream);

// True when the three-sweep kernels exist for length L and fit in max_lds.
bool fft3_supported(int L, size_t max_lds);
// Sets the kernels' LDS attributes for length L (once, before any launch).
void fft3_prepare(int L);
// hpsi = IFFT(1/2|G+k+A|^2 FFT(psi)) / N + V psi for the chunk's bands: all
// three sweeps (sweep < 0) or only sweep 0 (z forward), 1 (x-y) or 2 (z back).
void fft3_local_hamiltonian(const fft3_args &a, int L, hipStream_t stream, int sweep = -1,
                            bool fwd_next = false);

}  // namespace inqamd
