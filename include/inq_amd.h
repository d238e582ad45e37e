/*
 * inq_amd.h -- C ABI of the MI355X (gfx950) plane-wave Hpsi / real-time
 * propagator library.
 *
 * The library replaces, for one GPU, the hot path of INQ's
 *   hamiltonian::ks_hamiltonian   (src/hamiltonian/ks_hamiltonian.hpp:35-228)
 *   operations::transform         (src/operations/transform.hpp:337-400)
 *                                 (also reachable as operations::space)
 *   hamiltonian::projector_all    (src/hamiltonian/projector_all.hpp:23-421)
 *   operations::exponential*      (src/operations/exponential.hpp:21-69)
 *   real_time::etrs               (src/real_time/etrs.hpp:22-82)
 *   observables::current          (src/observables/current.hpp:25-53)
 * INQ's host code keeps its own containers and calls these functions with
 * plain device pointers.
 *
 * Conventions
 *   - Orbitals are complex128 (two interleaved doubles) in INQ's layout
 *     [x][y][z][band], band index contiguous, `ld` elements between grid
 *     points (ld >= nst), exactly field_set::matrix() (field_set.hpp:51).
 *   - Vectors suffixed _cov are covariant components (cell.hpp:179-192).
 *   - All orbital pointers are DEVICE pointers owned by the caller.
 *     Tables passed at creation are HOST pointers, copied by the library.
 *   - Every call is asynchronous on the handle's stream; nothing
 *     synchronises the device unless the function says so.
 *   - Return value: 0 on success, non-zero status otherwise; the message is
 *     available from inqamd_last_error(handle).  The library never aborts.
 *   - One handle per stream; a handle must not be used from two host threads
 *     at once without external locking.
 */
#ifndef INQ_AMD_H
#define INQ_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
  INQAMD_OK = 0,
  INQAMD_ERR_ARGUMENT = 1,
  INQAMD_ERR_HIP = 2,
  INQAMD_ERR_MEMORY = 3,
  INQAMD_ERR_UNSUPPORTED = 4
};

typedef struct inqamd_hamiltonian *inqamd_hamiltonian_t;

/* Everything ks_hamiltonian needs for a frozen (non-interacting) potential:
 * the grid (basis::real_space, real_space.hpp:30-33), V_KS
 * (ks_hamiltonian.hpp:45) and the packed projector tables of projector_all
 * (projector_all.hpp:408-419).  Projector tables may be absent (nproj = 0). */
typedef struct {
  int n[3];                /* FFT grid */
  double lattice[9];       /* a0, a1, a2 in bohr, row-major */
  const double *vloc;      /* host [N] scalar potential (Ha) */
  int nproj;               /* number of projector spheres (atoms) */
  int max_sphere;          /* max_sphere_size_ */
  int max_nlm;             /* max_nlm_ (<= 16) */
  const int *points;       /* host [nproj][max_sphere][3] grid coords, -1 padded */
  const float *positions;  /* host [nproj][max_sphere][3] contravariant r - R_atom (FP32) */
  const double *matrices;  /* host [nproj][max_nlm][max_sphere] beta_lm(r_i) */
  const double *coeff;     /* host [nproj][max_nlm] KB coefficients */
  const int *nlm;          /* host [nproj] projectors per sphere */
  const int *sphere_size;  /* host [nproj] points per sphere */
  int max_bands;           /* largest nst any later call will pass */
} inqamd_system_desc;

/* ---- handle --------------------------------------------------------------
 * Replaces ks_hamiltonian's constructor (ks_hamiltonian.hpp:79-89) and
 * projector_all(projectors) (projector_all.hpp:86-94).  `stream` is a
 * hipStream_t (NULL = default stream).  Allocates every workspace the later
 * calls need, so they allocate nothing -- except the ground-state calls
 * (orthogonalize, preconditioner, steepest_descent, subspace_diagonalization)
 * and the interacting TDDFT calls (density ... etrs_scf_step), which allocate their own
 * workspace on first use. */
int inqamd_hamiltonian_create(const inqamd_system_desc *desc, void *stream,
                              inqamd_hamiltonian_t *out);
void inqamd_hamiltonian_destroy(inqamd_hamiltonian_t h);
int inqamd_set_stream(inqamd_hamiltonian_t h, void *stream);
const char *inqamd_last_error(inqamd_hamiltonian_t h);
const char *inqamd_version(void);
/* Self-test of the FP64 matrix-core lane layout the projector kernels use:
 * writes D = A B (16 x 16, row-major) for A[i][k] = 1 + i + 16k,
 * B[k][j] = 2 + 3j - 5k into the DEVICE array out[256].  No reference
 * counterpart; a diagnostic for the projector contraction. */
int inqamd_selftest_mfma(double *out_device, void *stream);
/* Diagnostic: the lane layout of the 4x4x4 FP64 matrix instruction the
 * projector coefficients use; block s of D = A B with A_s[i][k] = 1 + i + 4k +
 * 16s and B_s[k][j] = 2 + 3j - 5k + 7s into out_device[s][i][j] (64 doubles). */
int inqamd_selftest_mfma_4x4(double *out_device, void *stream);
/* Streaming-bandwidth probe (diagnostic, no reference counterpart): one
 * launch of blocks_per_cu x CUs workgroups moving 16-byte complex128 elements
 * with plain vector loads/stores -- kind 0 read, 1 write, 2 copy (src -> dst,
 * same positions) -- over rows x nb elements of rows of ld elements: the
 * first nb of each row, or with grouped != 0 the 8-element groups 0,
 * nchunks, 2 nchunks, ... (chunk 0 of nchunks interleaved band groups).
 * ld == nb is a flat contiguous range.  DEVICE src/dst of rows x ld
 * elements (the read probe also needs dst).  Writes the launch's duration
 * (HIP events on `stream`) to *ms; synchronises. */
int inqamd_stream_probe(int kind, const void *src, void *dst, long rows, int ld, int nb,
                        int grouped, int nchunks, int blocks_per_cu, void *stream, float *ms);

/* orbital_set::kpoint() (orbital_set.hpp) -- covariant k. */
/* Per-handle tiling of the three-sweep Hpsi: the autotuner's knobs, no
 * reference counterpart.  0 keeps the library's choice for that knob.
 *   z_tile_y         z tiles of 8 y x 8 bands or 16 y x 4 bands (the plain
 *                    Hpsi's last sweep at 96 runs half-width tiles: 4 y x 8
 *                    bands or 8 y x 4 bands)
 *   xy_threads       x-y plane workgroups of 1 or 2 x plane_threads(L)
 *   band_chunk       bands per chunk of the sweeps (multiple of 8, <= max_bands)
 *   projector_bands  bands per workgroup of the projector coefficients (16/32/64)
 * z_tile_y, band_chunk and projector_bands change no result bit; xy_threads
 * selects another instantiation of the plane kernel, whose floating-point
 * contraction may differ in the last bits. */
typedef struct {
  int z_tile_y;
  int xy_threads;
  int band_chunk;
  int projector_bands;
} inqamd_tuning;
int inqamd_set_tuning(inqamd_hamiltonian_t h, const inqamd_tuning *t);
int inqamd_get_tuning(inqamd_hamiltonian_t h, inqamd_tuning *t);
int inqamd_set_kpoint(inqamd_hamiltonian_t h, const double k_cov[3]);
/* ks_hamiltonian::uniform_vector_potential() = A (ks_hamiltonian.hpp:216-221),
 * as self_consistency::update_hamiltonian sets it (self_consistency.hpp:152-156). */
int inqamd_set_uniform_vector_potential(inqamd_hamiltonian_t h, const double A_cov[3]);
/* ks_hamiltonian::scalar_potential() (ks_hamiltonian.hpp:200-202): replace V
 * from a DEVICE array [N] (stream ordered). */
int inqamd_set_scalar_potential(inqamd_hamiltonian_t h, const double *vloc_device);

/* ---- Hpsi ------------------------------------------------------------------
 * ks_hamiltonian::operator()(orbital_set<real_space, complex>)
 * (ks_hamiltonian.hpp:132-152): hpsi = T(k+A) psi + V psi + V_nl(k+A) psi.
 * psi and hpsi must not alias. */
int inqamd_hamiltonian_apply(inqamd_hamiltonian_t h, const void *psi, void *hpsi, int nst,
                             int ld);
/* ks_hamiltonian::operator()(orbital_set<fourier_space, complex>)
 * (ks_hamiltonian.hpp:156-176): phi and hphi in G space (the unnormalised
 * forward transform of the real-space orbitals), same layout.  Must not alias. */
int inqamd_hamiltonian_apply_fourier(inqamd_hamiltonian_t h, const void *phi, void *hphi,
                                     int nst, int ld);
/* Diagnostic: one Hpsi (as inqamd_hamiltonian_apply, three-sweep path only)
 * with HIP events around each kernel on the handle's stream.  ms[0]: the
 * projector (phases + gather/MFMA/expand), ms[1..3]: the z-forward, x-y and
 * z-backward sweeps summed over band chunks; launches[] the kernel launches
 * each figure covers.  Synchronises.  No reference counterpart. */
int inqamd_hamiltonian_profile(inqamd_hamiltonian_t h, const void *psi, void *hpsi, int nst,
                               int ld, double ms[4], int launches[4]);
/* ks_hamiltonian::non_local(orbital_set<real_space, complex>)
 * (ks_hamiltonian.hpp:104-128): vnlpsi = V_nl(k+A) psi. */
int inqamd_non_local(inqamd_hamiltonian_t h, const void *psi, void *vnlpsi, int nst, int ld);

/* ---- operations::transform (alias operations::space) ----------------------
 * to_fourier (transform.hpp:337-367): unnormalised forward 3-D DFT.
 * to_real (transform.hpp:371-400): backward 3-D DFT, x 1/N if normalize.
 * in and out may be the same pointer. */
int inqamd_transform_to_fourier(inqamd_hamiltonian_t h, const void *in, void *out, int nst,
                                int ld);
int inqamd_transform_to_real(inqamd_hamiltonian_t h, const void *in, void *out, int nst, int ld,
                             int normalize);

/* ---- projector_all ----------------------------------------------------------
 * project (projector_all.hpp:98-216): sphere[p][i][b] = V_nl psi restricted
 * to sphere p ([nproj][max_sphere][nst] complex, device).
 * apply (projector_all.hpp:220-232): hpsi(r_pi, b) += e^{-i(k+A).r} sphere.
 * Scatter order is fixed (no atomics): results are bitwise reproducible. */
int inqamd_projector_project(inqamd_hamiltonian_t h, const void *psi, void *sphere, int nst,
                             int ld);
int inqamd_projector_apply(inqamd_hamiltonian_t h, const void *sphere, void *hpsi, int nst,
                           int ld);

/* ---- reductions --------------------------------------------------------------
 * energy::calculate eigenvalues (energy.hpp:74-75 via
 * overlap_diagonal_normalized, overlap_diagonal.hpp:144-171):
 * eps[b] = Re<hpsi_b|psi_b> / <psi_b|psi_b>.  eps is a DEVICE array [nst]. */
int inqamd_band_energies(inqamd_hamiltonian_t h, const void *hpsi, const void *psi,
                         double *eps, int nst, int ld);
/* The same band energies without any device allocation, for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  Its
 * partial sums live in a fixed 8 MiB device array that each GPU holds until
 * the process ends, instead of the handle's workspace; nst is limited to
 * 2^19 / ceil(N / 256) bands (512 at 64^3, 151 at 96^3; larger nst returns
 * INQAMD_ERR_ARGUMENT with the limit in the message) but not by max_bands.
 * Results equal inqamd_band_energies' bit for bit.  Every call on one GPU
 * shares that array: calls that overlap on one GPU (two streams, two
 * threads) race and must be serialised by the caller. */
int inqamd_band_energies_noalloc(inqamd_hamiltonian_t h, const void *hpsi, const void *psi,
                                 double *eps, int nst, int ld);
/* The per-band terms of energy::calculate (energy.hpp:59-87) for one set of
 * orbitals, with the handle's current k+A:
 *   eps[b]     = Re<H psi_b|psi_b> / <psi_b|psi_b>             (:74-75)
 *   normres[b] = dV |H psi_b - eps_b psi_b|^2                  (:76-77)
 *   nl[b]      = Re<V_nl psi_b|psi_b> / <psi_b|psi_b>          (:81-85)
 * All three are DEVICE arrays [nst]; the occupation-weighted sums
 * (occ_sum) are the caller's.  Uses the handle's work buffers, so psi must
 * not be one of them. */
int inqamd_energy(inqamd_hamiltonian_t h, const void *psi, int nst, int ld, double *eps,
                  double *normres, double *nl);

/* ---- ground state (G-space orbitals) --------------------------------------------
 * operations::laplacian_add, fourier_space branch (laplacian.hpp:24-38):
 * laplff += factor (-|G|^2 + gradcoeff . G) ff, G covariant = 2 pi m,
 * gradcoeff contravariant. */
int inqamd_laplacian_add(inqamd_hamiltonian_t h, const void *ff, void *laplff, int nst, int ld,
                         double factor, const double gradcoeff_contra[3]);
/* operations::preconditioner on fourier_space orbitals (preconditioner.hpp:
 * 25-55, Teter-Payne-Allan): phi(G) *= K(1/2|G|^2 / kinetic_b), in place. */
int inqamd_preconditioner(inqamd_hamiltonian_t h, void *phi, int nst, int ld);
/* operations::orthogonalize (orthogonalize.hpp:25-32): Cholesky of
 * S = dV Phi^H Phi and Phi <- Phi R^-H (Gram-Schmidt in band order), in
 * place; dV of real space (fourier = 0) or G space (fourier != 0).
 * Synchronises the handle's stream (the factorisation runs on the host). */
int inqamd_orthogonalize(inqamd_hamiltonian_t h, void *phi, int nst, int ld, int fourier);
/* eigensolvers::steepest_descent (steepest_descent.hpp:23-88) on fourier_space
 * orbitals with ks_hamiltonian's G-space operator (ks_hamiltonian.hpp:156-176)
 * and, if precondition != 0, operations::preconditioner: five line
 * minimisations then orthogonalize, in place.  Allocates three [N][max_bands]
 * work arrays on first use (kept until destroy). */
int inqamd_steepest_descent(inqamd_hamiltonian_t h, void *phi, int nst, int ld,
                            int precondition);
/* ground_state::subspace_diagonalization (subspace_diagonalization.hpp:16-25)
 * on real-space orbitals: H_sub = dV Phi^H (H Phi), its eigenpairs
 * (matrix::diagonalize), Phi <- Phi V in place; eigenvalues (ascending) into
 * the HOST array eigenvalues[nst].  Eigenvector phases (and the basis of a
 * degenerate eigenspace) are the eigensolver's.  Synchronises. */
int inqamd_subspace_diagonalization(inqamd_hamiltonian_t h, void *phi, int nst, int ld,
                                    double *eigenvalues);

/* ---- interacting TDDFT (static ions, one spin, LDA) -------------------------------
 * Fields are DEVICE arrays [N] of doubles on the handle's grid.
 * observables::density::calculate_add (density.hpp:23-31):
 * rho[p] = sum_b occ[b] |psi[p][b]|^2; occ is a HOST array [nst]. */
int inqamd_density(inqamd_hamiltonian_t h, const void *psi, int nst, int ld, const double *occ,
                   double *rho);
/* observables::dipole of a density (dipole.hpp:24-40): dV sum_p r_cart rho,
 * the positions in the symmetric range; host result.  Synchronises. */
int inqamd_dipole(inqamd_hamiltonian_t h, const double *rho, double dipole_cart[3]);
/* solvers::poisson, 3-D periodic (poisson.hpp:29-97, 224-227):
 * V_H(G) = 4 pi rho(G) / |G|^2, zero mean. */
int inqamd_poisson(inqamd_hamiltonian_t h, const double *rho, double *vh);
/* The LDA of options::theory::lda (theory.hpp:48-53): Dirac exchange +
 * Perdew-Zunger 1981 correlation (libxc's LDA_X, LDA_C_PZ, restated from the
 * published forms) through xc_term::operator() (xc_term.hpp:120-142):
 * vxc[N], exc = int e(n) n, nvxc = int rho vxc (host scalars).  Synchronises. */
int inqamd_lda_xc(inqamd_hamiltonian_t h, const double *rho, double *vxc, double *exc,
                  double *nvxc);
/* hamiltonian::xc_term::evaluate_functional (xc_term.hpp:254-300) for one
 * functional of one spin component: kind 1 = LDA_X, 2 = LDA_X + LDA_C_PZ,
 * 3 = GGA_X_PBE, 4 = GGA_C_PBE, 5 = both PBE parts (the default theory,
 * theory.hpp:40-45).  GGAs take the density gradient and the divergence of
 * -2 vsigma grad rho spectrally (gradient.hpp:22-45).  vxc (device, N);
 * exc = int e rho, nvxc = int rho vxc (host).  Synchronises. */
int inqamd_xc(inqamd_hamiltonian_t h, const double *rho, int kind, double *vxc, double *exc,
              double *nvxc);
/* self_consistency::update_hamiltonian (self_consistency.hpp:88-162) for
 * static ions: the handle's scalar potential <- vion + V_H[rho] (hartree) +
 * V_xc[rho] (xc: 0 none, 1 LDA, 2 PBE); energies[4] (host) = {external,
 * hartree, exc, nvxc}.
 * Synchronises. */
int inqamd_ks_potential(inqamd_hamiltonian_t h, const double *vion, const double *rho,
                        int hartree, int xc, double energies[4]);
/* real_time::etrs with its self-consistent second half (etrs.hpp:22-82):
 * 2-for-1 Taylor step with H(t), density, H(t+dt) from it, then up to 5
 * half steps from the saved half step until int |drho| < 5e-5.  The
 * handle's scalar potential must be V_KS of H(t) on entry; it holds V_KS of
 * H(t+dt) on return.  occ is a HOST array [nst]; xc as for
 * inqamd_ks_potential; *nscf receives the SCF passes.  Allocates an
 * [N][max_bands] half-step array on first use.
 * Synchronises. */
int inqamd_etrs_scf_step(inqamd_hamiltonian_t h, double dt, const double A_t_cov[3],
                         const double A_tdt_cov[3], int order, void *psi, int nst, int ld,
                         const double *vion, const double *occ, int hartree, int xc, int *nscf);

/* ---- checkpoint files ------------------------------------------------------------
 * operations::io::save / load of an orbital_set (io.hpp:163-197): directory
 * `dirname` with one file per band, "%010d_%010d" (first_band + b, spinor),
 * each the band's N complex doubles in grid order, raw native-endian binary
 * (one process's basis part: the whole grid).  save creates the directory;
 * load fails (non-zero status) if a file is missing or short, as INQ's
 * load returns false.  Both synchronise. */
int inqamd_save_states(inqamd_hamiltonian_t h, const void *psi, int nst, int ld, long first_band,
                       int spinor, const char *dirname);
int inqamd_load_states(inqamd_hamiltonian_t h, void *psi, int nst, int ld, long first_band,
                       int spinor, const char *dirname);

/* ---- propagation --------------------------------------------------------------
 * operations::shift (shift.hpp:33-44): y += (re + i im) x. */
int inqamd_shift(inqamd_hamiltonian_t h, double re, double im, const void *x, void *y, int nst,
                 int ld);
/* operations::exponential_in_place (exponential.hpp:21-34):
 * psi <- sum_{n<=order} (z H)^n psi / n!  with the handle's current k+A. */
int inqamd_exponential_in_place(inqamd_hamiltonian_t h, double zre, double zim, int order,
                                void *psi, int nst, int ld);
/* operations::exponential_2_for_1 (exponential.hpp:50-69):
 * expphi <- T(z1) phi, phi <- T(z2) phi, sharing the H^n phi terms. */
int inqamd_exponential_2_for_1(inqamd_hamiltonian_t h, double z1re, double z1im, double z2re,
                               double z2im, int order, void *phi, void *expphi, int nst, int ld);
/* real_time::etrs for a frozen potential (etrs.hpp:22-82):
 * psi <- T(i dt/2, H(A_tdt)) T(i dt/2, H(A_t)) psi; leaves A = A_tdt. */
int inqamd_etrs_step(inqamd_hamiltonian_t h, double dt, const double A_t_cov[3],
                     const double A_tdt_cov[3], int order, void *psi, int nst, int ld);

/* ---- graph-captured ETRS --------------------------------------------------------
 * One frozen-potential ETRS step on a fixed orbital buffer, captured once as a
 * hipGraph (no allocation, no host synchronisation at replay).  The vector
 * potential is read from device memory at replay:
 *  - inqamd_etrs_graph_launch(g, A_t, A_tdt) writes k + A(t) and k + A(t+dt)
 *    into the device slots on the handle's stream, then launches the graph;
 *  - with a Fourier-series A(t) = direction * sum_{j=1..n} c_j sin(j w t)
 *    (Cartesian) the graph evaluates A at t = step*dt and t + dt itself and
 *    advances a device step counter, so A_t/A_tdt may be NULL at launch.
 * Replaces the per-step host loop of real_time::propagate (propagate.hpp:93-98)
 * around etrs (etrs.hpp:22-82). */
typedef struct inqamd_graph *inqamd_graph_t;
typedef struct {
  int n;                    /* number of coefficients (<= 64) */
  const double *coeff;      /* host [n] */
  double omega;             /* base angular frequency */
  double direction[3];      /* Cartesian polarisation */
} inqamd_vector_potential_series;

int inqamd_etrs_graph_create(inqamd_hamiltonian_t h, void *psi, int nst, int ld, double dt,
                             int order, const inqamd_vector_potential_series *series,
                             inqamd_graph_t *out);
int inqamd_etrs_graph_launch(inqamd_graph_t g, const double A_t_cov[3], const double A_tdt_cov[3]);
/* Step counter of a series graph (time = step * dt); stream ordered. */
int inqamd_etrs_graph_set_step(inqamd_graph_t g, long step);
void inqamd_graph_destroy(inqamd_graph_t g);

/* ---- current and the adjoint of a propagation ----------------------------------
 * observables::current (current.hpp:25-53) for bands of equal occupation occ:
 *   J = occ sum_b Im <phi_b| g |phi_b>,  g = -i(grad + i(k+A)) + [V_nl, r]
 * (gradient.hpp:43-48 with factor -1; projector_all::position_commutator,
 * projector_all.hpp:319-403).  The [V_nl, r] term has two forms:
 *   INQAMD_COMMUTATOR_INQ (the default): as the reference writes it, the
 *     plain projection multiplied by the KB coefficient D twice and the
 *     r-weighted one never (:362-366) -- identical results to INQ;
 *   INQAMD_COMMUTATOR_CORRECTED: each projection scaled by D once, which
 *     makes J = -dE/dA exactly (the INQ form equals it with every D replaced
 *     by (1 + D^2)/2).
 * inqamd_set_position_commutator picks the form for the handle.  J_cov is
 * the covariant vector current() returns, J_cart its Cartesian form
 * (viewables.hpp:100-102).  nst <= max_bands.  HOST outputs: synchronises
 * the handle's stream. */
enum { INQAMD_COMMUTATOR_INQ = 0, INQAMD_COMMUTATOR_CORRECTED = 1 };
int inqamd_set_position_commutator(inqamd_hamiltonian_t h, int commutator);
int inqamd_current(inqamd_hamiltonian_t h, const void *psi, int nst, int ld, double occ,
                   double J_cov[3], double J_cart[3]);

/* Frozen-potential propagation and its adjoint.  The forward runs nsteps
 * ETRS steps (etrs.hpp:22-82, Taylor order `order` per half step, exponent
 * +i dt/2; the second half uses A(t + dt) as etrs.hpp:54 does) under
 *   A(t) = direction * sum_{j=1..ncoef} c_j sin(j omega t)   (Cartesian)
 * from psi0 at t = 0, the loop of real_time::propagate (propagate.hpp:93-131)
 * without observables, and returns the loss L = J_cart(T) . direction
 * (the current of current.hpp:25-53 at A(T), T = nsteps dt).  It keeps psi
 * every checkpoint_every steps.  The backward returns dL/dc_j and dL/dV(r)
 * by the discrete adjoint of exactly the forward's arithmetic: each segment
 * is replayed from its checkpoint, one band slice at a time (bands are
 * independent under a frozen potential), and the adjoint state is carried
 * back through every Taylor half step.  The reference has no adjoint; these
 * functions replace its finite-difference use of propagate.  k is the
 * handle's k-point; V is the handle's scalar potential at forward time and
 * must not change before the backward. */
typedef struct inqamd_td *inqamd_td_t;
typedef struct {
  int nst;                  /* bands (<= max_bands) */
  int nsteps;               /* ETRS steps */
  double dt;                /* time step (atomic units) */
  int order;                /* Taylor order of each half step, 1..8 */
  double occ;               /* occupation of every band */
  int ncoef;                /* number of series coefficients (<= 64) */
  double omega;             /* base angular frequency */
  double direction[3];      /* Cartesian polarisation of A and of the loss */
  int checkpoint_every;     /* steps per checkpoint segment (>= 1) */
  int band_slice;           /* bands per backward slice (0: all) */
  int store_powers;         /* 1: keep H^j phi of a segment's half steps; 0: recompute */
  int commutator;           /* the loss's [V_nl, r] form: INQAMD_COMMUTATOR_INQ (0) or
                               INQAMD_COMMUTATOR_CORRECTED (see inqamd_current) */
} inqamd_td_desc;

/* Allocates checkpoints and adjoint workspaces (device memory). */
int inqamd_td_create(inqamd_hamiltonian_t h, const inqamd_td_desc *desc, inqamd_td_t *out);
/* Forwards run on this object so far (-1 for NULL).  The backward always
 * differentiates the LAST forward; a caller that keeps several losses alive
 * (e.g. autograd) records this number at its forward and compares it before
 * its backward.  No reference counterpart. */
long inqamd_td_forward_count(inqamd_td_t td);
/* psi0: DEVICE [N][ld]; coeff: HOST [ncoef]; psi_T: DEVICE [N][ld] or NULL;
 * loss: HOST.  Synchronises the handle's stream. */
int inqamd_td_forward(inqamd_td_t td, const double *coeff, const void *psi0, int ld,
                      void *psi_T, double *loss);
/* Gradients of dloss * L at the last forward's inputs: grad_coeff HOST
 * [ncoef], grad_vloc DEVICE [N] (overwritten; may be NULL), lambda0 DEVICE
 * [N][nst] or NULL: the adjoint state at t = 0 (dL = Re sum conj(lambda0)
 * dpsi0 over grid points and bands).  Fails with INQAMD_ERR_ARGUMENT if the
 * handle's scalar potential or k-point was set after the last forward.
 * Synchronises the handle's stream. */
int inqamd_td_backward(inqamd_td_t td, double dloss, double *grad_coeff, double *grad_vloc,
                       void *lambda0);
/* Device bytes held by the propagation object. */
long inqamd_td_memory(inqamd_td_t td);
void inqamd_td_destroy(inqamd_td_t td);

/* ---- uneven batches ----------------------------------------------------------
 * Hpsi of many independent systems (different grids, band counts and
 * projector sets, each its own handle) in one call: every phase of Hpsi is a
 * single persistent launch over the whole batch that drains a queue of
 * per-system work items (projector blocks, then the three sweeps' tiles of
 * band pieces packed into cache-sized chunks).  Every band goes through the
 * same tile bodies as in inqamd_hamiltonian_apply (agreement to rounding,
 * 1e-14 relative in the tests).  Replaces
 * the per-system loop a caller would write around ks_hamiltonian::operator()
 * (ks_hamiltonian.hpp:132-152); the reference has no batched form.
 * Systems whose grid has no three-sweep kernels run their one-system Hpsi
 * inside the call.  Pointer arrays are HOST arrays of DEVICE pointers; work
 * tables are rebuilt when they change.  Uses each handle's workspaces: a
 * handle must not run other work concurrently with a batch call. */
typedef struct inqamd_batch *inqamd_batch_t;
int inqamd_batch_create(inqamd_hamiltonian_t const *handles, int nsys, void *stream,
                        inqamd_batch_t *out);
int inqamd_batch_apply(inqamd_batch_t b, const void *const *psi, void *const *hpsi,
                       const int *nst, const int *ld);
/* Tiling choices (the autotuner's knobs): W bytes per batch chunk, the y
 * extent of full z tiles (8, 16 or 32 with 64/ty bands), whether band
 * remainders use narrower tiles, whether each sweep of a chunk is one
 * launch over all its grids (0) or one launch per grid length (1), and over
 * how many streams (1..4, default priority) the chunks are dealt. */
int inqamd_batch_set_tuning(inqamd_batch_t b, long chunk_bytes, int ty_main,
                            int narrow_remainders, int group_by_grid, int nstreams);
int inqamd_batch_set_stream(inqamd_batch_t b, void *stream);
/* Cache chunks, kernel launches per apply, band pieces and one-system
 * fallbacks of the last work tables. */
int inqamd_batch_info(inqamd_batch_t b, int *chunks, int *launches, int *pieces, int *fallback);
const char *inqamd_batch_last_error(inqamd_batch_t b);
void inqamd_batch_destroy(inqamd_batch_t b);

/* ---- host-side system builders (no GPU needed) --------------------------------
 * containing_cube (containing_cube.hpp:20-45). */
int inqamd_containing_cube(const double lattice[9], const int n[3], const double pos[3],
                           double radius, int lo[3], int hi[3]);
/* spherical_grid(basis, center, radius) (spherical_grid.hpp:58-158): grid
 * coords, FP32 distance and FP32 contravariant relative positions, in the
 * reference's point order.  Writes at most `capacity` points and stores the
 * full count in *count; returns INQAMD_ERR_MEMORY if capacity was short. */
int inqamd_sphere_build(const double lattice[9], const int n[3], const double center[3],
                        double radius, int capacity, int *coords, float *distance,
                        float *relpos, int *count);
/* projector::build for a Gaussian radial function exp(-r^2/sigma^2) and real
 * harmonics l=0 (nlm=1) or l=0,1 ordered y,z,x (nlm=4)
 * (projector.hpp:34-50): matrix[lm*ld + i]. */
int inqamd_gaussian_projector_matrix(const double lattice[9], int npoint, const float *distance,
                                     const float *relpos, double sigma, int nlm, double *matrix,
                                     int ld);

#ifdef __cplusplus
}
#endif
#endif
