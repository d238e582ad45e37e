/*
 * oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * FP64 CPU restatement of INQ's CPU path for the plane-wave Hpsi and the
 * real-time Taylor/ETRS step.  It is the checker the HIP library is compared
 * against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it; the product library never links or calls it.
 *
 * Conventions (all follow INQ):
 *   - orbitals are complex128 (interleaved re/im doubles) in the layout
 *     [x][y][z][band] with the band index contiguous, leading dimension `ld`
 *     (src/basis/field_set.hpp:51, :188-194);
 *   - to_fourier is the unnormalised forward DFT (sign -1), to_real is the
 *     backward DFT (sign +1) followed by x 1/N (src/operations/transform.hpp:
 *     132, 265, 325-331);
 *   - G and r indices use to_symmetric_range (src/basis/grid.hpp:111-117);
 *   - vectors tagged _cov are covariant (src/systems/cell.hpp:146-224).
 *
 * Parity pins (each function's checks live in tests/):
 *   orc_fft3d           numpy.fft golden vectors, brute-force DFT, round trip,
 *                       Gaussian closed form (transform.hpp:490-545)
 *   orc_containing_cube reference lo/hi known answers (containing_cube.hpp:77-117)
 *   orc_sphere_build    reference sphere counts 257/13758 and volume ratios
 *                       (spherical_grid.hpp:272-396)
 *   orc_hamiltonian_apply
 *                       reference harmonic-oscillator answer 0.0051420503
 *                       (ks_hamiltonian.hpp:356-397), plane-wave eigenvalues
 *                       with k+A, laplacian plane waves (laplacian.hpp:104-297),
 *                       dense brute-force V_nl, Hermiticity
 *   orc_hamiltonian_apply_fourier
 *                       = to_fourier(H to_real(phi)) by the real-space pins;
 *                       G-space plane-wave eigenvalues with k+A
 *   orc_band_energies   plane-wave eigenvalues, Rayleigh quotient closed form
 *   orc_energy          eigenstates (zero residual), |H phi|^2 - eps^2 |phi|^2
 *                       identity, non-local part = H(V=0) - H(V=0, no projectors)
 *   orc_exponential_*   diagonal closed forms in the plane-wave basis
 *                       (exponential.hpp:94-233 pattern), scipy expm
 *   orc_etrs_step       composition of two Taylor-4 half steps; norm drift
 *   orc_laplacian_add   plane-wave closed form with factor and gradient term
 *   orc_preconditioner  TPA closed form on a plane wave (K(1) = 65/81)
 *   orc_orthogonalize   numpy QR of sqrt(dV) Phi (Q / sqrt(dV), R > 0)
 *   orc_steepest_descent_dense
 *                       reference known answers of the 100 x 100 diagonal
 *                       problem (steepest_descent.hpp:120-166)
 *   orc_steepest_descent the dense pins above (same template) with the pinned
 *                       G-space Hamiltonian and preconditioner
 *   orc_density         plane waves (rho = sum occ / V), normalisation
 *   orc_poisson         reference point-charge known answers (poisson.hpp:
 *                       251-314), plane-wave closed form 4 pi/G^2
 *   orc_lda_xc          Dirac exchange closed form, Perdew-Zunger published
 *                       values and continuity at rs = 1, v = d(n e)/dn by
 *                       finite differences (libxc's own rounding: unpinned)
 *   orc_dipole          reference known answers of dipole.hpp:60-115
 *   orc_xc              reference known answers of xc_term.hpp:340-430 (LDA_X and
 *                       GGA_C_PBE energies and potentials on the reference's density:
 *                       PW92 gas, FFT gradient and divergence); GGA_X_PBE by its
 *                       closed-form limits; every kind's potential as the functional
 *                       derivative of its energy (central differences)
 *   orc_ks_potential    sum of the pinned parts; E_H = 1/2 int rho V_H
 *   orc_etrs_scf        frozen limit = orc_etrs_step; time reversal; energy
 *                       conservation of the interacting step
 */
#ifndef INQ_AMD_ORACLE_H
#define INQ_AMD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* A frozen-potential Kohn-Sham system as INQ's ks_hamiltonian holds it. */
typedef struct {
  int n[3];               /* grid sizes */
  double lattice[9];      /* a0, a1, a2 (cartesian, bohr), row-major */
  const double *vloc;     /* [N] local potential, Ha */
  int nproj;              /* projector_all::nprojs_ */
  int max_sphere;         /* projector_all::max_sphere_size_ */
  int max_nlm;            /* projector_all::max_nlm_ */
  const int *points;      /* [nproj][max_sphere][3], -1 padded */
  const float *positions; /* [nproj][max_sphere][3] contravariant, FP32 */
  const double *matrices; /* [nproj][max_nlm][max_sphere] */
  const double *coeff;    /* [nproj][max_nlm] */
  const int *nlm;         /* [nproj] */
  const int *sphere_size; /* [nproj] */
} orc_system;

/* Batched 3-D DFT over the [x][y][z][band] layout (in and out may alias).
 * sign = -1 forward, +1 backward; normalize multiplies by 1/N afterwards. */
int orc_fft3d(const int n[3], int nst, int ld, const double *in, double *out,
              int sign, int normalize);

/* Plain 1-D DFT of one contiguous complex vector (used by tests). */
int orc_fft1d(int n, const double *in, double *out, int sign);

/* containing_cube (src/basis/containing_cube.hpp:20-45). */
int orc_containing_cube(const double lattice[9], const int n[3],
                        const double pos[3], double radius, int lo[3], int hi[3]);

/* spherical_grid::initialize (src/basis/spherical_grid.hpp:58-158).  Writes at
 * most `capacity` points; returns the number of points in *count (which may
 * exceed capacity, then call again with a larger buffer). */
int orc_sphere_build(const double lattice[9], const int n[3], const double center[3],
                     double radius, int capacity, int *coords, float *distance,
                     float *relpos, int *count);

/* Gaussian-projector table M[lm][i] = exp(-d^2/sigma^2) Y_lm(r_i) for l=0,1
 * (real harmonics ordered y, z, x for l=1), evaluated from the FP32 sphere
 * geometry as projector::build does (src/hamiltonian/projector.hpp:44-50). */
int orc_gaussian_projector_matrix(const double lattice[9], int npoint,
                                  const float *distance, const float *relpos,
                                  double sigma, int nlm, double *matrix, int ld);

/* ks_hamiltonian::operator()(orbital_set<real_space, complex>)
 * (src/hamiltonian/ks_hamiltonian.hpp:132-152). kA_cov = k + A (covariant). */
int orc_hamiltonian_apply(const orc_system *sys, const double kA_cov[3], int nst,
                          int ld, const double *psi, double *hpsi);

/* ks_hamiltonian::operator()(orbital_set<fourier_space, complex>)
 * (src/hamiltonian/ks_hamiltonian.hpp:156-176) on orbitals in G space. */
int orc_hamiltonian_apply_fourier(const orc_system *sys, const double kA_cov[3], int nst,
                                  int ld, const double *phi, double *hphi);

/* projector_all::project / apply (src/hamiltonian/projector_all.hpp:98-232).
 * sphere is [nproj][max_sphere][nst] complex. */
int orc_projector_project(const orc_system *sys, const double kA_cov[3], int nst,
                          int ld, const double *psi, double *sphere);
int orc_projector_apply(const orc_system *sys, const double kA_cov[3], int nst,
                        int ld, const double *sphere, double *hpsi);

/* energy::calculate eigenvalue part: eps_b = Re<Hpsi_b|psi_b>/<psi_b|psi_b>
 * (src/hamiltonian/energy.hpp:74-75, overlap_diagonal.hpp:144-171). */
int orc_band_energies(int npoint, int nst, int ld, const double *hpsi,
                      const double *psi, double *eps);

/* energy::calculate per band (src/hamiltonian/energy.hpp:59-107): eigenvalue,
 * residual norm dV |H phi - eps phi|^2, non-local expectation value. */
int orc_energy(const orc_system *sys, const double kA_cov[3], int nst, int ld,
               const double *psi, double *eps, double *normres, double *nl);

/* operations::exponential_in_place (src/operations/exponential.hpp:21-34). */
int orc_exponential_in_place(const orc_system *sys, const double kA_cov[3],
                             double zre, double zim, int order, int nst, int ld,
                             double *psi);

/* operations::exponential_2_for_1 (src/operations/exponential.hpp:50-69):
 * phi <- T(z2) phi, expphi <- T(z1) phi (old phi). */
int orc_exponential_2_for_1(const orc_system *sys, const double kA_cov[3],
                            double z1re, double z1im, double z2re, double z2im,
                            int order, int nst, int ld, double *phi, double *expphi);

/* real_time::etrs with a frozen potential (src/real_time/etrs.hpp:22-82):
 * psi(t+dt) = T4(i dt/2, H(t+dt)) T4(i dt/2, H(t)) psi(t). */
int orc_etrs_step(const orc_system *sys, const double kA_t[3], const double kA_tdt[3],
                  double dt, int order, int nst, int ld, double *psi);

/* operations::randomize (src/operations/randomize.hpp:48-70): pcg32 seeded with
 * N * nst + index, advanced by 2 (b + nst * point) per complex element;
 * uniform values 2 g / 2^32 - 1. */
int orc_randomize(const int n[3], int nst, int index, double *psi);

/* observables::current (src/observables/current.hpp:25-53):
 * j = sum_b occ Im(conj(phi) g phi), g phi = -i (G + k + A) phi + [V_nl, r] phi
 * (gradient.hpp:43-48 with factor -1; projector_all.hpp:319-403).
 * commutator 0: position_commutator as written (the plain projection scaled by
 * the KB coefficient twice, the r-weighted one unscaled, :362-366);
 * commutator 1: each projection scaled once (the A-derivative of V_nl).
 * J_cov = dV sum_r j; J_cart = to_cartesian(J_cov) (viewables.hpp:100-102). */
int orc_current(const orc_system *sys, const double kA_cov[3], double occ, int nst, int ld,
                const double *psi, int commutator, double J_cov[3], double J_cart[3]);

/* Frozen-potential propagation under A(t) = direction * sum_{j=1..n} c_j
 * sin(j w t) (Cartesian): nsteps ETRS steps from psi (in place, A(t) evaluated
 * at t = step*dt and t + dt as etrs.hpp:54), then returns the final current
 * projected on `direction` (Cartesian), with occupation occ per band and the
 * commutator form of orc_current. */
int orc_td_current(const orc_system *sys, int ncoef, const double *coef, double omega,
                   const double direction[3], double dt, int nsteps, int order, double occ,
                   int nst, int ld, double *psi, int commutator, double *loss);

/* Set the number of OpenMP threads the oracle uses (<=0: library default). */
/* operations::laplacian_add, G-space (laplacian.hpp:24-38) */
int orc_laplacian_add(const orc_system *sys, double factor, const double gradcoeff_contra[3],
                      int nst, int ld, const double *ff, double *laplff);
/* operations::preconditioner on G-space orbitals (preconditioner.hpp:25-55), in place */
int orc_preconditioner(const orc_system *sys, int nst, int ld, double *phi);
/* operations::orthogonalize (orthogonalize.hpp:25-32) with volume element dV */
int orc_orthogonalize(long npoint, int nst, int ld, double dV, double *phi);
/* eigensolvers::steepest_descent (steepest_descent.hpp:23-88) on G-space
 * orbitals with the G-space ks_hamiltonian and (precond != 0) the TPA
 * preconditioner; in place */
int orc_steepest_descent(const orc_system *sys, const double kA_cov[3], int precond, int nst,
                         int ld, double *phi);
/* niter steepest_descent calls with a dense operator M [npoint][npoint] and no
 * preconditioner, then eigenvalues[b] = dV <phi_b|M phi_b> (complex) */
int orc_steepest_descent_dense(int npoint, int nvec, const double *matrix, double dV, int niter,
                               double *phi, double *eigenvalues);

/* observables::density::calculate_add (density.hpp:23-31), one spin:
 * rho[p] = sum_b occ[b] |psi[p][b]|^2 */
int orc_density(long npoint, int nst, int ld, const double *occ, const double *psi, double *rho);
/* solvers::poisson, 3-D periodic (poisson.hpp:29-97, 224-227) of a real density */
int orc_poisson(const orc_system *sys, const double *rho, double *vh);
/* LDA_X + LDA_C_PZ (options::theory::lda) through xc_term::operator()
 * (xc_term.hpp:120-142): vxc[N], exc = int e n, nvxc = int rho vxc */
int orc_lda_xc(long npoint, double dV, const double *rho, double *vxc, double *exc, double *nvxc);
/* observables::dipole of a density (dipole.hpp:24-40): dV sum r_cart rho */
int orc_dipole(const orc_system *sys, const double *rho, double *out);
/* Exchange-correlation of one spin component: kind 1 = LDA_X, 2 = LDA_X +
 * LDA_C_PZ, 3 = GGA_X_PBE, 4 = GGA_C_PBE, 5 = both PBE parts; vxc[N] and
 * exc = dV sum e n (xc_term.hpp:254-300) */
int orc_xc(const orc_system *sys, const double *rho, int kind, double *vxc, double *exc);
/* self_consistency::update_hamiltonian (self_consistency.hpp:88-162):
 * vks = vion + V_H (hartree) + V_xc (xc); energies = {external, hartree, exc, nvxc} */
int orc_ks_potential(const orc_system *sys, const double *vion, const double *rho, int hartree,
                     int xc, double *vks, double *energies);
/* real_time::etrs with its SCF re-check (etrs.hpp:22-82); sys->vloc is V_KS
 * of H(t), vks receives V_KS of H(t+dt), nscf the SCF passes */
int orc_etrs_scf(const orc_system *sys, const double *vion, const double *occ, const double kA_t[3],
                 const double kA_tdt[3], double dt, int order, int hartree, int xc, int nst, int ld,
                 double *psi, double *vks, int *nscf);

void orc_set_threads(int nthreads);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
