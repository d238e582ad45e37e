// Internal state of a Hamiltonian handle and the launchers shared between
// the translation units of the library.
#pragma once

#include "common.hpp"
#include "fft.hpp"

#include <complex>
#include <cstdlib>
#include <exception>
#include <new>
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
  // (1 + D^2)/2 per coefficient: with it the A-derivative form of [V_nl, r]
  // gives the current of position_commutator as the reference writes it
  // (projector_all.hpp:362-366)
  double *coeff_inq = nullptr;
  int *size = nullptr;     // [nproj] points (0 for locally empty spheres)
  // ownership map of the scatter: unique grid points and their contributors
  int nunique = 0;
  int *upoint = nullptr;   // [nunique] linear grid index
  int *uoff = nullptr;     // [nunique + 1]
  int *uent = nullptr;     // [sum K] p * max_sphere + i, sorted by (point, p, i)
  int nent = 0;            // sum K (entries of uent)
  int *owner = nullptr;    // [N] row of the ownership map holding a grid point, -1 if none
  // e^{+i(k+A).r} per sphere point and e^{-i(k+A).r} per ownership entry,
  // recomputed for the k+A of each Hpsi call
  cplx *phase_sph = nullptr;  // [nproj][max_sphere]
  cplx *phase_ent = nullptr;  // [nent]
  int *ent = nullptr;         // [nproj][max_sphere] ownership entry of each sphere point
  int *own = nullptr;         // [N] (first entry << OWN_COUNT_BITS) | count, -1 if not owned
  // per ownership entry, for expanding V_nl psi from the coefficients in the
  // Hpsi sweep (max_nlm <= 4 only; nullptr otherwise)
  int *ent_sph = nullptr;     // [nent] sphere
  int *own_sph = nullptr;     // [N] sphere of the point's first entry, 0 where unowned
  double *ent_mat = nullptr;  // [nent][4] M[sphere][lm][point], zero for lm >= max_nlm
};

struct project_expand_args {
  const cplx *psi;
  long ld;
  int nst;
  int max_sphere, max_nlm;
  const int *point, *size;
  const double *mat, *coeff;
  const cplx *phase;  // e^{+i(k+A).r} per sphere point
  const int *ent;
  double dV;
  cplx *P;            // [nproj][max_nlm][nst] or nullptr
  cplx *Qe;           // [nent][ldQ] V_nl psi per ownership entry, or nullptr
  long ldQ;
  // the 4x4x4 coefficient kernels only: nbc bands of a chunk whose band cb
  // sits at psi[cb + (cb >> (3 + glog)) gskip] (psi offset to the chunk's
  // first band) and is written to P at band pband0 + the same offset
  int nbc, gskip, glog;
  long pband0;
};

// psi_acc += c * (H psi) accumulations fused into the Hpsi epilogue (Taylor
// sums); p == nullptr disables a slot.
struct taylor_acc {
  cplx *p1 = nullptr;
  long ld1 = 0;
  double c1re = 0.0, c1im = 0.0;
  bool init1 = false;  // p1 = psi + c1 hpsi (p1's old values unused)
  cplx *p2 = nullptr;
  long ld2 = 0;
  double c2re = 0.0, c2im = 0.0;
  // three-sweep Hpsi only: the forward transform of psi, band-major
  // [nst][kz][kx][ky] (nullptr: not written)
  cplx *gdump = nullptr;
};

}  // namespace inqamd

struct inqamd_hamiltonian {
  int n[3] = {0, 0, 0};
  long npoint = 0;
  double lattice[9] = {};
  double rlat[9] = {};
  double volume = 0.0;
  hipStream_t stream = nullptr;
  int device = 0;                   // the device current at creation
  hipStream_t aux = nullptr;        // second stream: odd band chunks of Hpsi
  hipStream_t aux2 = nullptr;       // third stream (INQAMD_STREAMS=3)
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_join2 = nullptr, ev_proj = nullptr;
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
  inqamd::cplx *vnl = nullptr;  // [nent][max_bands] V_nl psi per ownership entry
  double *partial = nullptr;    // reduction partials
  long npartial = 0;
  inqamd::cplx *mom_a = nullptr; // [nproj][max_nlm][4][max_bands] sphere moments
  inqamd::cplx *mom_b = nullptr; // second operand of pair reductions
  double *obs = nullptr;        // observable partials and result records
  double *proj_part = nullptr;  // point-chunk GEMM1 tiles of the projector
  inqamd::cplx *mom_partial = nullptr;  // point-chunk partial sums of sphere moments
  long mom_partial_elems = 0;

  inqamd::projector_tables pt;
  // per-handle tiling of the three-sweep Hpsi (inqamd_set_tuning; 0: the
  // process default)
  int tune_z_ty = 0;            // z tiles of 8 or 16 y
  int tune_xy_wide = 0;         // x-y planes: 1 = plane_threads(L), 2 = twice that
  int tune_chunk = 0;           // bands per chunk of the sweeps
  int tune_p4 = 0;              // bands per projector coefficient workgroup (16/32/64)
  int commutator = 0;           // INQAMD_COMMUTATOR_* of inqamd_current
  long generation = 0;          // bumped whenever V or k changes (stale-state checks)

  void *gs = nullptr;           // steepest-descent workspace and BLAS handle (gs.hip)
  void *tddft = nullptr;        // density / potential fields and the ETRS half step (tddft.hip)

  std::string last_error;
};

namespace inqamd {

using handle = inqamd_hamiltonian;

// projector.hip
void projector_upload(handle &h, int nproj, int max_sphere, int max_nlm, const int *points,
                      const float *positions, const double *matrices, const double *coeff,
                      const int *nlm, const int *sphere_size);
void projector_free(handle &h);
// phase_sph / phase_ent for the k+A in `slot`
void projector_phases(handle &h, int slot);
// P[p][lm][b] = dV * D[p][lm] * sum_i M[p][lm][i] e^{i(k+A).r_pi} psi(r_pi, b)
void projector_project_coefficients(handle &h, const cplx *psi, long ld, int nst, int slot);
// hpsi(r, b) += sum_{(p,i) at r} e^{-i(k+A).r_pi} sum_lm M[p][lm][i] P[p][lm][b]
void projector_apply_coefficients(handle &h, cplx *hpsi, long ld, int nst, int slot);
// Q[u][b] = the same sum for every owned grid point u (compact, ld = nst)
void projector_owned_sum(handle &h, cplx *Q, int nst, int slot);
// projector_mfma.hip: gather + MFMA GEMM1 (-> P if write_P) + MFMA GEMM2 ->
// per-entry V_nl psi in Qe (if Qe != nullptr)
void projector_project_expand(handle &h, const cplx *psi, long ld, int nst, cplx *Qe, long ldQ,
                              bool write_P);
void mfma_selftest(double *out_device, hipStream_t stream);
// P coefficients of all spheres and bands on the 4x4x4 FP64 MFMA (max_nlm <= 4)
void projector_project_p4(handle &h, const cplx *psi, long ld, int nst);
// The same for the nb bands of one band chunk (psi offset to its first band
// b0; band groups as in fft3_args) on `stream`: P at bands b0 + offset
// (part_slot: which of PROJ_PART_SLOTS regions of the partial-tile buffer
// its point-chunked form uses; concurrent calls take different slots)
void projector_project_p4_chunk(handle &h, const cplx *psi, long ld, int nst, int b0, int nb,
                                int gskip, int glog, hipStream_t stream, int part_slot);
void mfma4_selftest(double *out_device, hipStream_t stream);

// hamiltonian.hip
void set_slot(handle &h, int slot, const double v[3]);
void update_current_slot(handle &h);
// G-space Hpsi on a three-sweep grid (phi and hphi distinct from the
// handle's work arrays)
void hamiltonian_fourier_sweeps(handle &h, const cplx *phi, long ld, cplx *hphi, long hld,
                                int nst);
void fft3d(handle &h, const cplx *in, long in_ld, cplx *out, long out_ld, int nst, int sign,
           double scale);
void hamiltonian_apply(handle &h, const cplx *psi, long ld, cplx *hpsi, long hld, int nst,
                       int slot, const taylor_acc *acc = nullptr);
void shift(handle &h, std::complex<double> c, const cplx *x, long xld, cplx *y, long yld,
           int nst);
void copy(handle &h, const cplx *x, long xld, cplx *y, long yld, int nst);
// Taylor chains on three-sweep grids (hamiltonian.hip): H^n in for n = 1..order
// with acc1 += z1^n/n! H^n in (acc1 = in + z1 H in when init1) and acc2 +=
// z2^n/n! H^n in, chunk by chunk, each order's last sweep transforming its
// output forward for the next; outs/dumps optional per order (see there)
bool taylor_chain_ok(const handle &h);
void taylor_chain(handle &h, const cplx *in, long ild, int nst, int order, int slot, cplx *p1,
                  long ld1, bool init1, std::complex<double> z1, cplx *p2, long ld2,
                  std::complex<double> z2, cplx *const *outs, cplx *const *dumps);
void exponential_in_place(handle &h, std::complex<double> z, int order, cplx *psi, long ld,
                          int nst, int slot);
// gs.hip
void hamiltonian_fourier(handle &h, const cplx *phi, long ld, cplx *hphi, long hld, int nst);
void gs_release(handle &h);
// tddft.hip
void td_release(handle &h);
// td.hip: observables scratch of a handle (sizes from max_bands and tables)
constexpr int OBS_GS_BLOCKS = 1024;  // blocks (and partial rows) of the G-space reductions
constexpr int OBS_RECORD = 32;       // doubles per result record
constexpr int OBS_NRECORDS = 16;
constexpr int MOM_MAX_CHUNKS = 16;     // point chunks per sphere in the moments kernel
constexpr int PROJ_MAX_CHUNKS = 16;    // point chunks per sphere in the projector
constexpr long PROJ_PART_TILES = 4096; // 16 x 16 GEMM1 tiles the chunk buffer holds
constexpr int PROJ_PART_SLOTS = 4;     // regions of it for concurrent coefficient launches
constexpr long MOM_PARTIAL_BLOCKS = 4096;  // workgroups whose partials the buffer holds
long observables_scratch_doubles(const handle &h);
// sphere[p][i][b] = sum_lm M[p][lm][i] P[p][lm][b]   (INQ-shaped project output)
void projector_expand(handle &h, cplx *sphere, int nst);
// hpsi(r, b) += sum_{(p,i) at r} e^{-i(k+A).r_pi} sphere[p][i][b]
void projector_apply_sphere(handle &h, const cplx *sphere, cplx *hpsi, long ld, int nst, int slot);


// Whether the three-sweep Hpsi of this handle expands V_nl psi in its last
// sweep from the projector coefficients (needs max_nlm <= 4; INQAMD_VNL=entries
// keeps V_nl psi per ownership entry instead)
inline bool vnl_from_coefficients(const handle &h) {
  static const bool entries = [] {
    const char *e = getenv("INQAMD_VNL");
    return e && std::string(e) == "entries";
  }();
  return !entries && h.three_sweep && h.pt.nproj > 0 && h.pt.nunique > 0 && h.pt.ent_mat;
}

// C ABI plumbing shared by the translation units that export functions
template <typename F>
int guarded(inqamd_hamiltonian *h, F &&f) {
  try {
    f();
    if (h) h->last_error.clear();
    return OK;
  } catch (const error &e) {
    if (h) h->last_error = e.msg;
    return e.code;
  } catch (const std::bad_alloc &) {
    if (h) h->last_error = "host allocation failed";
    return ERR_MEMORY;
  } catch (const std::exception &e) {
    if (h) h->last_error = e.what();
    return ERR_ARGUMENT;
  }
}

inline void check_orbital(const handle &h, const void *p, const char *name) {
  if (p == nullptr) throw error{ERR_ARGUMENT, std::string(name) + " is null"};
  if (reinterpret_cast<uintptr_t>(p) % 16 != 0)
    throw error{ERR_ARGUMENT, std::string(name) + " is not 16-byte aligned"};
  (void)h;
}

inline void check_bands(const handle &h, int nst, int ld, bool needs_workspace) {
  if (nst <= 0) throw error{ERR_ARGUMENT, "nst must be positive"};
  if (ld < nst) throw error{ERR_ARGUMENT, "ld must be >= nst"};
  if (needs_workspace && nst > h.max_bands)
    throw error{ERR_ARGUMENT, "nst (" + std::to_string(nst) + ") exceeds max_bands (" +
                                  std::to_string(h.max_bands) + ")"};
}

}  // namespace inqamd
