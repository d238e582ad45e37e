// Hamiltonian handle, Hpsi, transforms, reductions and the Taylor/ETRS
// propagator, exported through the C ABI of include/inq_amd.h.

#include "internal.hpp"
#include "inq_amd.h"

#include <cmath>
#include <cstdlib>
#include <complex>
#include <mutex>
#include <new>
#include <vector>

namespace inqamd {

// ---------------------------------------------------------------------------
// small kernels
// ---------------------------------------------------------------------------
__global__ void set_vec3_kernel(double *dst, double x, double y, double z) {
  if (threadIdx.x == 0) {
    dst[0] = x;
    dst[1] = y;
    dst[2] = z;
  }
}

// y[p][b] += c * x[p][b]  (operations::shift, shift.hpp:39-43)
__global__ __launch_bounds__(256) void shift_kernel(cplx *y, long yld, const cplx *x, long xld,
                                                    int nst, long total, double cre, double cim) {
  for (long g = (long)blockIdx.x * blockDim.x + threadIdx.x; g < total;
       g += (long)gridDim.x * blockDim.x) {
    const long p = g / nst;
    const int b = (int)(g - p * nst);
    const cplx xv = x[p * xld + b];
    cplx yv = y[p * yld + b];
    yv.x += cre * xv.x - cim * xv.y;
    yv.y += cre * xv.y + cim * xv.x;
    y[p * yld + b] = yv;
  }
}

__global__ __launch_bounds__(256) void copy_kernel(cplx *y, long yld, const cplx *x, long xld,
                                                   int nst, long total) {
  for (long g = (long)blockIdx.x * blockDim.x + threadIdx.x; g < total;
       g += (long)gridDim.x * blockDim.x) {
    const long p = g / nst;
    const int b = (int)(g - p * nst);
    y[p * yld + b] = x[p * xld + b];
  }
}

// A(t) = direction * sum_j c_j sin(j w t) at t = step*dt and t + dt
// (etrs.hpp:54: the Hamiltonian is updated to time + dt), covariant, plus k;
// then step += 1.
struct series_args {
  int n;
  double c[64];
  double omega;
  double dt;
  double dir_cov[3];
  double k_cov[3];
};

__global__ void advance_series_kernel(const series_args s, double *step, double *kA) {
  if (threadIdx.x != 0) return;
  const double t0 = step[0] * s.dt;
  const double t1 = t0 + s.dt;
  double f0 = 0.0, f1 = 0.0;
  for (int j = 0; j < s.n; j++) {
    f0 += s.c[j] * sin((j + 1) * s.omega * t0);
    f1 += s.c[j] * sin((j + 1) * s.omega * t1);
  }
  for (int d = 0; d < 3; d++) {
    kA[3 * 1 + d] = s.k_cov[d] + s.dir_cov[d] * f0;  // KA_T
    kA[3 * 2 + d] = s.k_cov[d] + s.dir_cov[d] * f1;  // KA_TDT
    kA[3 * 0 + d] = s.k_cov[d] + s.dir_cov[d] * f1;  // KA_CURRENT: H(t+dt) after the step
  }
  step[0] += 1.0;
}

__global__ void set_scalar_kernel(double *dst, double v) {
  if (threadIdx.x == 0) dst[0] = v;
}

constexpr int EN_BT = 64;
constexpr int EN_LANES = 4;
constexpr long EN_CHUNK = 256;

// Per-chunk partial sums of Re(conj(hpsi) psi) and |psi|^2 for every band
// (overlap_diagonal_normalized_mult, overlap_diagonal.hpp:118-128).
__global__ __launch_bounds__(256) void energies_partial_kernel(const cplx *hpsi, long hld,
                                                               const cplx *psi, long ld, int nst,
                                                               long npoint, double *partial) {
  __shared__ double red[2][EN_LANES][EN_BT];
  const int bb = threadIdx.x % EN_BT;
  const int lane = threadIdx.x / EN_BT;
  const int b = blockIdx.y * EN_BT + bb;
  const long p0 = (long)blockIdx.x * EN_CHUNK;
  const long p1 = p0 + EN_CHUNK < npoint ? p0 + EN_CHUNK : npoint;
  double v = 0.0, nn = 0.0;
  if (b < nst) {
    for (long p = p0 + lane; p < p1; p += EN_LANES) {
      const cplx h = hpsi[p * hld + b];
      const cplx s = psi[p * ld + b];
      v += h.x * s.x + h.y * s.y;
      nn += s.x * s.x + s.y * s.y;
    }
  }
  red[0][lane][bb] = v;
  red[1][lane][bb] = nn;
  __syncthreads();
  if (lane == 0 && b < nst) {
    for (int l = 1; l < EN_LANES; l++) {
      v += red[0][l][bb];
      nn += red[1][l][bb];
    }
    partial[((long)blockIdx.x * nst + b) * 2 + 0] = v;
    partial[((long)blockIdx.x * nst + b) * 2 + 1] = nn;
  }
}

// Per-chunk partial sums of |hpsi - eps_b psi|^2 (the residual of
// energy.hpp:74-77) in column 0 of the partial layout above.
__global__ __launch_bounds__(256) void residual_partial_kernel(const cplx *hpsi, long hld,
                                                               const cplx *psi, long ld,
                                                               const double *eps, int nst,
                                                               long npoint, double *partial) {
  __shared__ double red[EN_LANES][EN_BT];
  const int bb = threadIdx.x % EN_BT;
  const int lane = threadIdx.x / EN_BT;
  const int b = blockIdx.y * EN_BT + bb;
  const long p0 = (long)blockIdx.x * EN_CHUNK;
  const long p1 = p0 + EN_CHUNK < npoint ? p0 + EN_CHUNK : npoint;
  double r = 0.0;
  if (b < nst) {
    const double e = eps[b];
    for (long p = p0 + lane; p < p1; p += EN_LANES) {
      const cplx h = hpsi[p * hld + b];
      const cplx s = psi[p * ld + b];
      const double dx = h.x - e * s.x, dy = h.y - e * s.y;
      r += dx * dx + dy * dy;
    }
  }
  red[lane][bb] = r;
  __syncthreads();
  if (lane == 0 && b < nst) {
    for (int l = 1; l < EN_LANES; l++) r += red[l][bb];
    partial[((long)blockIdx.x * nst + b) * 2 + 0] = r;
  }
}

__global__ void column_sum_kernel(const double *partial, int nchunk, int nst, double scale,
                                  double *out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nst) return;
  double v = 0.0;
  for (int c = 0; c < nchunk; c++) v += partial[((long)c * nst + b) * 2 + 0];
  out[b] = scale * v;
}

__global__ void energies_final_kernel(const double *partial, int nchunk, int nst, double *eps) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nst) return;
  double v = 0.0, nn = 0.0;
  for (int c = 0; c < nchunk; c++) {
    v += partial[((long)c * nst + b) * 2 + 0];
    nn += partial[((long)c * nst + b) * 2 + 1];
  }
  eps[b] = v / nn;
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
// Streams the three-sweep band chunks are dealt over: 2, or 3 with
// INQAMD_STREAMS=3 (read once per process; the third stream is created only
// then).
static int pipeline_streams() {
  static const int n = [] {
    const char *e = getenv("INQAMD_STREAMS");
    return e && atoi(e) == 3 ? 3 : 2;
  }();
  return n;
}

static unsigned grid_for(long total) {
  long blocks = (total + 255) / 256;
  if (blocks > 65536L * 8) blocks = 65536L * 8;
  if (blocks < 1) blocks = 1;
  return (unsigned)blocks;
}

void set_slot(handle &h, int slot, const double v[3]) {
  hipLaunchKernelGGL(set_vec3_kernel, dim3(1), dim3(64), 0, h.stream, h.kA + 3 * slot, v[0], v[1],
                     v[2]);
  INQAMD_HIP_CHECK(hipGetLastError());
}

void update_current_slot(handle &h) {
  double kA[3] = {h.k_cov[0] + h.A_cov[0], h.k_cov[1] + h.A_cov[1], h.k_cov[2] + h.A_cov[2]};
  set_slot(h, KA_CURRENT, kA);
}

static const long *point_strides(const handle &h, long ps[3]) {
  ps[0] = (long)h.n[1] * h.n[2];
  ps[1] = h.n[2];
  ps[2] = 1;
  return ps;
}

// Pencil pass along `axis`; the fastest-varying other axis enumerates the
// workgroups right after the band tiles (neighbouring memory).
static fft_pass_args pass_args(const handle &h, int axis, const cplx *in, long in_ld, cplx *out,
                               long out_ld, int nst, int sign, double scale) {
  static const int others[3][2] = {{2, 1}, {2, 0}, {1, 0}};
  long ps[3];
  point_strides(h, ps);
  fft_pass_args a{};
  a.in = in;
  a.in_ld = in_ld;
  a.out = out;
  a.out_ld = out_ld;
  a.nst = nst;
  a.len = h.n[axis];
  a.stride = ps[axis];
  a.ax = axis;
  a.ax1 = others[axis][0];
  a.ax2 = others[axis][1];
  a.c1 = h.n[a.ax1];
  a.s1 = ps[a.ax1];
  a.c2 = h.n[a.ax2];
  a.s2 = ps[a.ax2];
  a.nstages = h.axis[axis].nstages;
  for (int s = 0; s < FFT_MAX_STAGES; s++) a.radix[s] = h.axis[axis].radix[s];
  a.tw = h.axis[axis].tw;
  a.scale = scale;
  a.sign = sign;
  for (int d = 0; d < 3; d++) a.n[d] = h.n[d];
  for (int d = 0; d < 9; d++) a.rlat[d] = h.rlat[d];
  a.kA = h.kA;
  return a;
}

// 3-D transform as three pencil passes (x, y, z); the last one applies scale.
void fft3d(handle &h, const cplx *in, long in_ld, cplx *out, long out_ld, int nst,
                  int sign, double scale) {
  fft_launch_pass(pass_args(h, 0, in, in_ld, out, out_ld, nst, sign, 1.0), FFT_PLAIN, h.stream);
  fft_launch_pass(pass_args(h, 1, out, out_ld, out, out_ld, nst, sign, 1.0), FFT_PLAIN, h.stream);
  fft_launch_pass(pass_args(h, 2, out, out_ld, out, out_ld, nst, sign, scale), FFT_PLAIN, h.stream);
}

// The projector step of the three-sweep Hpsi (after projector_phases).  Where
// vnl_from_coefficients holds, the last sweep expands V_nl psi from the
// projector coefficients P (the projector writes only P, 1 MB at 64^3 x 256,
// on the 4x4x4 FP64 MFMA), rather than from per-entry V_nl psi values that
// the projector writes and the sweep reads back (16 B per sphere point and
// band each way): 64^3 x 256 Hpsi 1.70 -> 1.65 ms, the 96^3 ETRS step -1.2 %.
static void hpsi_projector(handle &h, const cplx *psi, long ld, int nst) {
  if (vnl_from_coefficients(h))
    projector_project_p4(h, psi, ld, nst);
  else
    projector_project_expand(h, psi, ld, nst, h.vnl, nst, false);
}

// The local part T + V on the bands [b0, b0 + nb) of psi (pointers already
// offset to b0).  The three-sweep path also adds the projector apply and the
// Taylor accumulations `acc` (offset to b0) in its last sweep; the pencil
// path leaves both to the caller.
static void local_passes(handle &h, const cplx *psi, long ld, cplx *hpsi, long hld, int nb,
                         int slot, int b0, long ldP, const taylor_acc *acc, int sweep = -1,
                         hipStream_t stream = nullptr, cplx *W = nullptr, int gskip = 0,
                         int glog = 0, bool fwd_next = false, bool skip_store = false) {
  if (h.three_sweep) {
    fft3_args a{};
    a.gskip = gskip;
    a.glog = glog;
    a.psi = psi;
    a.ld = ld;
    a.hpsi = hpsi;
    a.hld = hld;
    a.W = W ? W : h.work;
    a.nb = nb;
    a.V = h.V;
    a.kA = h.kA + 3 * slot;
    for (int d = 0; d < 9; d++) a.rlat[d] = h.rlat[d];
    a.tw = h.axis[0].tw;
    a.scale = 1.0 / (double)h.npoint;
    const projector_tables &t = h.pt;
    if (vnl_from_coefficients(h)) {
      a.own = t.own;
      a.Pc = h.proj + b0;
      a.ldPc = ldP;
      a.nlmP = t.max_nlm;
      a.ent_sph = t.ent_sph;
      a.own_sph = t.own_sph;
      a.ent_mat = t.ent_mat;
      a.phase_ent = t.phase_ent;
    } else if (t.nproj > 0 && t.nunique > 0) {
      a.own = t.own;
      a.Qe = h.vnl + b0;
      a.ldQ = ldP;
    }
    if (acc && acc->p1) {
      a.acc1 = acc->p1 + b0;
      a.acc1_ld = acc->ld1;
      a.c1re = acc->c1re;
      a.c1im = acc->c1im;
      a.acc1_init = acc->init1 ? 1 : 0;
    }
    if (acc && acc->p2) {
      a.acc2 = acc->p2 + b0;
      a.acc2_ld = acc->ld2;
      a.c2re = acc->c2re;
      a.c2im = acc->c2im;
    }
    if (acc && acc->gdump) {
      if (gskip) throw error{ERR_UNSUPPORTED, "band-group chunks with a transform dump"};
      a.gdump = acc->gdump + (size_t)b0 * h.npoint;
    }
    a.skip_store = skip_store ? 1 : 0;
    a.z_ty = h.tune_z_ty;
    a.xy_wide = h.tune_xy_wide;
    fft3_local_hamiltonian(a, h.n[0], stream ? stream : h.stream, sweep, fwd_next);
    return;
  }
  cplx *w = h.work;
  const long wl = nb;
  fft_launch_pass(pass_args(h, 0, psi, ld, w, wl, nb, -1, 1.0), FFT_PLAIN, h.stream);
  fft_launch_pass(pass_args(h, 1, w, wl, w, wl, nb, -1, 1.0), FFT_PLAIN, h.stream);
  fft_pass_args kz = pass_args(h, 2, w, wl, w, wl, nb, -1, 1.0);
  kz.kA = h.kA + 3 * slot;
  fft_launch_pass(kz, FFT_KINETIC, h.stream);
  fft_launch_pass(pass_args(h, 1, w, wl, w, wl, nb, +1, 1.0), FFT_PLAIN, h.stream);
  fft_pass_args ep = pass_args(h, 0, w, wl, hpsi, hld, nb, +1, 1.0 / (double)h.npoint);
  ep.V = h.V;
  ep.psi = psi;
  ep.psi_ld = ld;
  fft_launch_pass(ep, FFT_EPILOGUE, h.stream);
}

// Bands per chunk of the local passes (0: all bands at once).  A chunk's
// intermediate array can stay cache-resident between passes.
static int band_chunk(const handle &h, int nst) {
  static int forced = [] {
    const char *e = getenv("INQAMD_BAND_CHUNK");
    return e ? atoi(e) : -1;
  }();
  if (h.tune_chunk > 0) return h.tune_chunk < nst ? h.tune_chunk : nst;
  if (forced >= 0) return forced == 0 ? nst : forced;
  // at least 192 MiB of intermediate and 24 bands per chunk: measured
  // fastest at 64^3 x 256 (48 bands; 32 and 64 were 6 % slower) and at
  // 96^3 x 512 (24 bands; 8 and 16 were 4-9 % slower)
  long c = (192L << 20) / (h.npoint * (long)sizeof(cplx));
  c = c < 24 ? 24 : c / 8 * 8;
  return c < nst ? (int)c : nst;
}

// Band chunks of the three-sweep Hpsi.  Where rows are a multiple of 4 KB
// (ld % 256 == 0) and the bands come in whole groups of 8, chunk c of n
// takes the groups c, c + n, c + 2n, ... rather than a contiguous range: a
// contiguous chunk's reads and writes all fall in one 768-byte window of
// every row and crowd a few HBM channels (64^3 x 256: z_fwd 4.4 TB/s,
// against 5.0 at ld = 240 or 288), while groups spread over the row do not.
// Other strides keep contiguous chunks (groups were 3 % slower at ld = 128).
// INQAMD_BAND_GROUPS=0 keeps them always.
struct band_chunks {
  int nst = 0, chunk = 0, n = 1, gskip = 0, glog = 0;
  int gsize() const { return 8 << glog; }
  int b0(int c) const { return gskip ? gsize() * c : c * chunk; }
  int nb(int c) const {
    if (gskip) return gsize() * ((nst / gsize() - c + n - 1) / n);
    return nst - c * chunk < chunk ? nst - c * chunk : chunk;
  }
};

static band_chunks chunk_plan(const handle &h, int nst, long ld, bool groups) {
  static const bool enabled = [] {
    const char *e = getenv("INQAMD_BAND_GROUPS");
    return !(e && atoi(e) == 0);
  }();
  static const int glog = [] {
    const char *e = getenv("INQAMD_BAND_GROUP_SIZE");
    const int g = e ? atoi(e) : 8;
    return g == 32 ? 2 : (g == 16 ? 1 : 0);
  }();
  band_chunks p;
  p.nst = nst;
  p.chunk = band_chunk(h, nst);
  p.n = (nst + p.chunk - 1) / p.chunk;
  const int gs = 8 << glog;
  if (enabled && groups && h.three_sweep && ld % 256 == 0 && p.n > 1 && nst % gs == 0 &&
      p.chunk % gs == 0 && p.n <= nst / gs) {
    p.glog = glog;
    p.gskip = gs * (p.n - 1);
  }
  return p;
}

void shift(handle &h, std::complex<double> c, const cplx *x, long xld, cplx *y, long yld,
           int nst);

// ks_hamiltonian::operator() (ks_hamiltonian.hpp:132-152):
//   phases of e^{+-i(k+A).r} for this call's k+A
//   project -> P coefficients                         :136
//   per band chunk: FFT, x 1/2|G+k+A|^2, inverse FFT   :138-144
//     x 1/N + V psi (the 1/2|k+A|^2 shift of :146 is in the kinetic factor)
//     apply (ordered, atomic-free scatter)             :149
// and, when `acc` is given, acc_j += c_j * hpsi (the Taylor shift of
// exponential.hpp:32, :64-65) in the same write.
void hamiltonian_apply(handle &h, const cplx *psi, long ld, cplx *hpsi, long hld, int nst,
                       int slot, const taylor_acc *acc) {
  const band_chunks plan = chunk_plan(h, nst, ld, !(acc && acc->gdump));
  const int chunk = plan.chunk;
  // three-sweep chunks alternate between two streams (and two halves of the
  // intermediate) so one chunk's launches fill the tails of the other's.
  // Called eagerly, the first chunk (on the second stream) also runs its two
  // FFT sweeps beside the projector and waits for it only before the sweep
  // that adds V_nl psi.  Under graph capture the projector leads: replays of
  // that forked shape ran 15 % slower than the plain alternation.
  static const bool pipeline = [] {
    const char *e = getenv("INQAMD_PIPELINE");
    return !(e && atoi(e) == 0);
  }();
  const bool two = pipeline && h.three_sweep && h.aux && chunk < nst && 2 * chunk <= h.max_bands;
  bool overlap = false;
  if (two) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    INQAMD_HIP_CHECK(hipStreamIsCapturing(h.stream, &cs));
    overlap = cs == hipStreamCaptureStatusNone;
  }
  // INQAMD_STREAMS=3 deals the chunks over three streams (and three slots
  // of the intermediate) where three chunks fit
  const int ns =
      two && pipeline_streams() == 3 && h.aux2 && plan.n >= 3 && 3 * chunk <= h.max_bands ? 3 : 2;
  auto fork = [&] {
    INQAMD_HIP_CHECK(hipEventRecord(h.ev_fork, h.stream));
    INQAMD_HIP_CHECK(hipStreamWaitEvent(h.aux, h.ev_fork, 0));
    if (ns == 3) INQAMD_HIP_CHECK(hipStreamWaitEvent(h.aux2, h.ev_fork, 0));
  };
  if (overlap) fork();
  projector_phases(h, slot);
  if (h.three_sweep)
    hpsi_projector(h, psi, ld, nst);
  else
    projector_project_expand(h, psi, ld, nst, nullptr, 0, true);
  if (overlap) INQAMD_HIP_CHECK(hipEventRecord(h.ev_proj, h.stream));
  if (two && !overlap) fork();
  const int gs = plan.gskip;  // 0 unless three-sweep
  bool waited[3] = {true, false, false};  // the main stream runs after the projector
  for (int c = 0; c < plan.n; c++) {
    const int b0 = plan.b0(c), nb = plan.nb(c);
    if (two) {
      const int si = (c + (overlap ? 1 : 0)) % ns;
      hipStream_t s = si == 0 ? h.stream : (si == 1 ? h.aux : h.aux2);
      cplx *W = h.work + (size_t)(c % ns) * chunk * h.npoint;
      if (overlap && !waited[si]) {
        waited[si] = true;
        local_passes(h, psi + b0, ld, hpsi + b0, hld, nb, slot, b0, nst, acc, 0, s, W, gs, plan.glog);
        local_passes(h, psi + b0, ld, hpsi + b0, hld, nb, slot, b0, nst, acc, 1, s, W, gs, plan.glog);
        INQAMD_HIP_CHECK(hipStreamWaitEvent(s, h.ev_proj, 0));
        local_passes(h, psi + b0, ld, hpsi + b0, hld, nb, slot, b0, nst, acc, 2, s, W, gs, plan.glog);
      } else {
        local_passes(h, psi + b0, ld, hpsi + b0, hld, nb, slot, b0, nst, acc, -1, s, W, gs, plan.glog);
      }
    } else {
      local_passes(h, psi + b0, ld, hpsi + b0, hld, nb, slot, b0, nst, acc, -1, nullptr, nullptr,
                   gs, plan.glog);
    }
  }
  if (two) {
    INQAMD_HIP_CHECK(hipEventRecord(h.ev_join, h.aux));
    INQAMD_HIP_CHECK(hipStreamWaitEvent(h.stream, h.ev_join, 0));
    if (ns == 3) {
      INQAMD_HIP_CHECK(hipEventRecord(h.ev_join2, h.aux2));
      INQAMD_HIP_CHECK(hipStreamWaitEvent(h.stream, h.ev_join2, 0));
    }
  }
  if (!h.three_sweep) {
    projector_apply_coefficients(h, hpsi, hld, nst, slot);
    if (acc && acc->p1 && acc->init1) copy(h, psi, ld, acc->p1, acc->ld1, nst);
    if (acc && acc->p1) shift(h, {acc->c1re, acc->c1im}, hpsi, hld, acc->p1, acc->ld1, nst);
    if (acc && acc->p2) shift(h, {acc->c2re, acc->c2im}, hpsi, hld, acc->p2, acc->ld2, nst);
  }
}

// ks_hamiltonian::operator() on G-space orbitals (ks_hamiltonian.hpp:156-176)
// for three-sweep grids: psi = to_real(phi), the projector on psi, then
// H phi = to_fourier(V psi + V_nl psi) + 1/2 |G+k+A|^2 phi in three sweeps
// per band chunk (z forward, x-y forward, transposing write with the kinetic
// term).  The kinetic term acts on phi directly: the real-space
// Hpsi followed by to_fourier would transform psi forward and back again.
void hamiltonian_fourier_sweeps(handle &h, const cplx *phi, long ld, cplx *hphi, long hld,
                                int nst) {
  fft3d(h, phi, ld, h.hn, nst, nst, +1, 1.0 / (double)h.npoint);
  projector_phases(h, KA_CURRENT);
  projector_project_expand(h, h.hn, nst, nst, h.vnl, nst, false);
  const band_chunks plan = chunk_plan(h, nst, ld, true);
  for (int c = 0; c < plan.n; c++) {
    const int b0 = plan.b0(c), nb = plan.nb(c);
    fft3_args a{};
    a.gskip = plan.gskip;
    a.glog = plan.glog;
    a.psi = h.hn + b0;
    a.ld = nst;
    a.hpsi = hphi + b0;
    a.hld = hld;
    a.W = h.work;
    a.nb = nb;
    a.V = h.V;
    a.kA = h.kA + 3 * KA_CURRENT;
    for (int d = 0; d < 9; d++) a.rlat[d] = h.rlat[d];
    a.tw = h.axis[0].tw;
    a.scale = 1.0;
    const projector_tables &t = h.pt;
    if (t.nproj > 0 && t.nunique > 0) {
      a.own = t.own;
      a.Qe = h.vnl + b0;
      a.ldQ = nst;
    }
    a.kin_in = phi + b0;
    a.kin_ld = ld;
    fft3_fourier_hamiltonian(a, h.n[0], h.stream);
  }
}

void shift(handle &h, std::complex<double> c, const cplx *x, long xld, cplx *y, long yld,
                  int nst) {
  const long total = h.npoint * nst;
  hipLaunchKernelGGL(shift_kernel, dim3(grid_for(total)), dim3(256), 0, h.stream, y, yld, x, xld,
                     nst, total, c.real(), c.imag());
  INQAMD_HIP_CHECK(hipGetLastError());
}

void copy(handle &h, const cplx *x, long xld, cplx *y, long yld, int nst) {
  const long total = h.npoint * nst;
  hipLaunchKernelGGL(copy_kernel, dim3(grid_for(total)), dim3(256), 0, h.stream, y, yld, x, xld,
                     nst, total);
  INQAMD_HIP_CHECK(hipGetLastError());
}

// Taylor chains on three-sweep grids (exponential.hpp:21-69): for
// n = 1..order, H^n psi = H (H^{n-1} psi) with acc1 += c1_n H^n psi (acc1 =
// in + c1_1 H in when init1) and acc2 += c2_n H^n psi, c_n = z^n / n!.  Run
// band chunk by band chunk with every order of a chunk on one stream: the
// projector takes the chunk's bands of H^{n-1} psi, and each order's last
// sweep also transforms its output forward in z into W for the next order,
// so no z_fwd sweep reads H^n psi back (16 B per point and band per order);
// the last order writes only the accumulators.  Chunks alternate between two
// streams.  The same arithmetic as one Hpsi per order: results are equal bit
// for bit.  Needs V_nl psi from coefficients (or no projectors);
// INQAMD_TAYLOR_CHAIN=0 runs one Hpsi per order.
bool taylor_chain_ok(const handle &h) {
  static const bool enabled = [] {
    const char *e = getenv("INQAMD_TAYLOR_CHAIN");
    return !(e && atoi(e) == 0);
  }();
  return enabled && h.three_sweep &&
         (h.pt.nproj == 0 || h.pt.nunique == 0 || vnl_from_coefficients(h));
}

// outs[n - 1] (ld nst) receives H^n in when given (nullptr: the handle's
// ping-pong buffers, and for the last order nothing); dumps[n - 1] receives
// the forward transform of order n's input (H^{n-1} in), band-major, as the
// x-y sweep computes it (taylor_acc::gdump; contiguous band chunks then).
void taylor_chain(handle &h, const cplx *in, long ild, int nst, int order, int slot, cplx *p1,
                  long ld1, bool init1, std::complex<double> z1, cplx *p2, long ld2,
                  std::complex<double> z2, cplx *const *outs, cplx *const *dumps) {
  std::vector<std::complex<double>> c1(order), c2(order);
  {
    std::complex<double> a = 1.0, b = 1.0;
    for (int n = 1; n <= order; n++) {
      a *= z1 / double(n);
      b *= z2 / double(n);
      c1[n - 1] = a;
      c2[n - 1] = b;
    }
  }
  const band_chunks plan = chunk_plan(h, nst, ild, dumps == nullptr);
  static const bool pipeline = [] {
    const char *e = getenv("INQAMD_PIPELINE");
    return !(e && atoi(e) == 0);
  }();
  const bool two = pipeline && h.aux && plan.chunk < nst && 2 * plan.chunk <= h.max_bands;
  const bool proj = h.pt.nproj > 0 && h.pt.nunique > 0;
  projector_phases(h, slot);
  if (two) {
    INQAMD_HIP_CHECK(hipEventRecord(h.ev_fork, h.stream));
    INQAMD_HIP_CHECK(hipStreamWaitEvent(h.aux, h.ev_fork, 0));
  }
  cplx *buf[2] = {h.hn, h.hn2};
  for (int c = 0; c < plan.n; c++) {
    const int b0 = plan.b0(c), nb = plan.nb(c);
    hipStream_t s = two && (c & 1) ? h.aux : h.stream;
    cplx *W = two ? h.work + (size_t)(c & 1) * plan.chunk * h.npoint : h.work;
    const cplx *src = in + b0;
    long sld = ild;
    for (int n = 1; n <= order; n++) {
      cplx *dst = (outs && outs[n - 1] ? outs[n - 1] : buf[n & 1]) + b0;
      const bool keep = n < order || (outs && outs[n - 1]);
      if (proj)
        projector_project_p4_chunk(h, src, sld, nst, b0, nb, plan.gskip, plan.glog, s,
                                   two ? (c & 1) : 0);
      taylor_acc acc;
      acc.p1 = p1;
      acc.ld1 = ld1;
      acc.c1re = c1[n - 1].real();
      acc.c1im = c1[n - 1].imag();
      acc.init1 = init1 && n == 1;
      acc.p2 = p2;
      acc.ld2 = ld2;
      acc.c2re = c2[n - 1].real();
      acc.c2im = c2[n - 1].imag();
      acc.gdump = dumps ? dumps[n - 1] : nullptr;
      if (n == 1)
        local_passes(h, src, sld, dst, nst, nb, slot, b0, nst, &acc, 0, s, W, plan.gskip,
                     plan.glog);
      local_passes(h, src, sld, dst, nst, nb, slot, b0, nst, &acc, 1, s, W, plan.gskip, plan.glog);
      local_passes(h, src, sld, dst, nst, nb, slot, b0, nst, &acc, 2, s, W, plan.gskip, plan.glog,
                   n < order, !keep);
      src = dst;
      sld = nst;
    }
  }
  if (two) {
    INQAMD_HIP_CHECK(hipEventRecord(h.ev_join, h.aux));
    INQAMD_HIP_CHECK(hipStreamWaitEvent(h.stream, h.ev_join, 0));
  }
}

// exponential_in_place (exponential.hpp:21-34): the shift of each order is
// carried by the Hpsi that produces the term.
void exponential_in_place(handle &h, std::complex<double> z, int order, cplx *psi, long ld,
                                 int nst, int slot) {
  if (taylor_chain_ok(h)) {
    taylor_chain(h, psi, ld, nst, order, slot, psi, ld, false, z, nullptr, 0, 0.0, nullptr,
                 nullptr);
    return;
  }
  cplx *hn = h.hn, *hn2 = h.hn2;
  const long tl = nst;
  std::complex<double> coeff = 1.0;
  for (int iter = 1; iter <= order; iter++) {
    coeff *= z / double(iter);
    taylor_acc acc;
    acc.p1 = psi;
    acc.ld1 = ld;
    acc.c1re = coeff.real();
    acc.c1im = coeff.imag();
    if (iter == 1) {
      hamiltonian_apply(h, psi, ld, hn, tl, nst, slot, &acc);
    } else {
      hamiltonian_apply(h, hn, tl, hn2, tl, nst, slot, &acc);
      std::swap(hn, hn2);
    }
  }
}

// out = T(z, H) in (exponential_in_place on a copy of in, without the
// copy: the first H application writes in + z H in to out); in and out have
// their own leading dimensions
void exponential_from(handle &h, std::complex<double> z, int order, const cplx *in, long ild,
                      cplx *out, long old, int nst, int slot) {
  if (taylor_chain_ok(h)) {
    taylor_chain(h, in, ild, nst, order, slot, out, old, true, z, nullptr, 0, 0.0, nullptr,
                 nullptr);
    return;
  }
  cplx *hn = h.hn, *hn2 = h.hn2;
  const long tl = nst;
  std::complex<double> coeff = 1.0;
  for (int iter = 1; iter <= order; iter++) {
    coeff *= z / double(iter);
    taylor_acc acc;
    acc.p1 = out;
    acc.ld1 = old;
    acc.c1re = coeff.real();
    acc.c1im = coeff.imag();
    acc.init1 = iter == 1;
    if (iter == 1) {
      hamiltonian_apply(h, in, ild, hn, tl, nst, slot, &acc);
    } else {
      hamiltonian_apply(h, hn, tl, hn2, tl, nst, slot, &acc);
      std::swap(hn, hn2);
    }
  }
}

// exponential_2_for_1 (exponential.hpp:50-69); expphi (leading dimension
// eld) starts as phi + z1 H phi from the first H application (no copy of phi)
void exponential_2_for_1(handle &h, std::complex<double> z1, std::complex<double> z2,
                                int order, cplx *phi, cplx *expphi, long ld, long eld, int nst,
                                int slot) {
  if (taylor_chain_ok(h)) {
    taylor_chain(h, phi, ld, nst, order, slot, expphi, eld, true, z1, phi, ld, z2, nullptr,
                 nullptr);
    return;
  }
  cplx *hn = h.hn, *hn2 = h.hn2;
  const long tl = nst;
  std::complex<double> c1 = 1.0, c2 = 1.0;
  for (int iter = 1; iter <= order; iter++) {
    c1 *= z1 / double(iter);
    c2 *= z2 / double(iter);
    taylor_acc acc;
    acc.p1 = expphi;
    acc.ld1 = eld;
    acc.c1re = c1.real();
    acc.c1im = c1.imag();
    acc.init1 = iter == 1;
    acc.p2 = phi;
    acc.ld2 = ld;
    acc.c2re = c2.real();
    acc.c2im = c2.imag();
    if (iter == 1) {
      hamiltonian_apply(h, phi, ld, hn, tl, nst, slot, &acc);
    } else {
      hamiltonian_apply(h, hn, tl, hn2, tl, nst, slot, &acc);
      std::swap(hn, hn2);
    }
  }
}

// The two launches of the band energies with their [nchunk][nst][2] partial
// sums in `partial` (every entry the final kernel reads is written first)
static void band_energies_with(handle &h, const cplx *hpsi, long hld, const cplx *psi, long ld,
                               int nst, double *eps, double *partial) {
  const int nchunk = (int)((h.npoint + EN_CHUNK - 1) / EN_CHUNK);
  dim3 grid(nchunk, (nst + EN_BT - 1) / EN_BT);
  hipLaunchKernelGGL(energies_partial_kernel, grid, dim3(256), 0, h.stream, hpsi, hld, psi, ld,
                     nst, h.npoint, partial);
  INQAMD_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(energies_final_kernel, dim3((nst + 63) / 64), dim3(64), 0, h.stream,
                     (const double *)partial, nchunk, nst, eps);
  INQAMD_HIP_CHECK(hipGetLastError());
}

void band_energies(handle &h, const cplx *hpsi, long hld, const cplx *psi, long ld, int nst,
                   double *eps) {
  band_energies_with(h, hpsi, hld, psi, ld, nst, eps, h.partial);
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free band energies
// ---------------------------------------------------------------------------
// The partial sums of energies_partial_kernel live in this file-scope device
// array instead of the handle's workspace: 8 MiB that each GPU the library
// runs on holds until the process ends.  On the host the array's name is a
// stand-in, not a device address: host code only ever uses the address that
// hipGetSymbolAddress returns for a device (scratch_address), resolved once
// per device when a handle is created there.
constexpr long NOALLOC_SCRATCH_DOUBLES = 1L << 20;
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

constexpr int SCRATCH_MAX_DEVICES = 64;
static std::mutex scratch_mutex;
static double *scratch_dev[SCRATCH_MAX_DEVICES] = {};

// The device address of `scratch` on device `dev`, which must be current
// the first time it is asked for (handle creation makes sure of that).
static double *scratch_address(int dev) {
  INQAMD_REQUIRE(dev >= 0 && dev < SCRATCH_MAX_DEVICES, "device index out of range");
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if (!scratch_dev[dev]) {
    int cur = -1;
    INQAMD_HIP_CHECK(hipGetDevice(&cur));
    INQAMD_REQUIRE(cur == dev, "the scratch of a device is resolved with that device current");
    void *p = nullptr;
    INQAMD_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    INQAMD_REQUIRE(p != nullptr, "hipGetSymbolAddress returned no address for the scratch");
    scratch_dev[dev] = static_cast<double *>(p);
  }
  return scratch_dev[dev];
}

// Bands whose partial sums fit the scratch on a grid of npoint points
static long noalloc_max_bands(long npoint) {
  const long nchunk = (npoint + EN_CHUNK - 1) / EN_CHUNK;
  return NOALLOC_SCRATCH_DOUBLES / (2 * nchunk);
}

// ---------------------------------------------------------------------------
// create / destroy
// ---------------------------------------------------------------------------
static void *dalloc(size_t bytes) {
  void *p = nullptr;
  if (bytes == 0) return nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess)
    throw error{ERR_MEMORY, std::string("hipMalloc(") + std::to_string(bytes) + "): " +
                                hipGetErrorString(e)};
  return p;
}

static void release(handle &h) {
  gs_release(h);
  td_release(h);
  for (int a = 0; a < 3; a++)
    if (h.axis[a].tw) (void)hipFree(h.axis[a].tw), h.axis[a].tw = nullptr;
  for (void *p : {(void *)h.V, (void *)h.kA, (void *)h.work, (void *)h.hn, (void *)h.hn2,
                  (void *)h.proj, (void *)h.vnl, (void *)h.partial, (void *)h.mom_a,
                  (void *)h.mom_b, (void *)h.obs, (void *)h.mom_partial, (void *)h.proj_part})
    if (p) (void)hipFree(p);
  h.V = nullptr;
  h.kA = nullptr;
  h.work = h.hn = h.hn2 = h.proj = h.vnl = nullptr;
  h.partial = nullptr;
  h.mom_a = h.mom_b = nullptr;
  h.obs = nullptr;
  h.mom_partial = nullptr;
  h.proj_part = nullptr;
  if (h.ev_fork) (void)hipEventDestroy(h.ev_fork);
  if (h.ev_join) (void)hipEventDestroy(h.ev_join);
  if (h.ev_join2) (void)hipEventDestroy(h.ev_join2);
  if (h.aux2) (void)hipStreamDestroy(h.aux2);
  if (h.ev_proj) (void)hipEventDestroy(h.ev_proj);
  if (h.aux) (void)hipStreamDestroy(h.aux);
  h.ev_fork = h.ev_join = h.ev_join2 = h.ev_proj = nullptr;
  h.aux = h.aux2 = nullptr;
  projector_free(h);
}

static void create(handle &h, const inqamd_system_desc &d, hipStream_t stream) {
  for (int a = 0; a < 3; a++) INQAMD_REQUIRE(d.n[a] > 0, "grid sizes must be positive");
  INQAMD_REQUIRE(d.max_bands > 0, "max_bands must be positive");
  INQAMD_REQUIRE(d.vloc != nullptr, "vloc is required");
  INQAMD_REQUIRE(d.max_nlm >= 0 && d.max_nlm <= 16, "max_nlm must be in [0, 16]");
  INQAMD_REQUIRE(d.nproj >= 0 && d.max_sphere >= 0, "negative projector sizes");
  if (d.nproj > 0)
    INQAMD_REQUIRE(d.points && d.positions && d.matrices && d.coeff && d.nlm && d.sphere_size,
                   "projector tables are incomplete");
  h.stream = stream;
  for (int a = 0; a < 3; a++) h.n[a] = d.n[a];
  h.npoint = (long)d.n[0] * d.n[1] * d.n[2];
  h.max_bands = d.max_bands;
  for (int i = 0; i < 9; i++) h.lattice[i] = d.lattice[i];
  // cell.hpp:46-52
  const double *a0 = d.lattice, *a1 = d.lattice + 3, *a2 = d.lattice + 6;
  auto cross = [](const double *u, const double *v, double *w) {
    w[0] = u[1] * v[2] - u[2] * v[1];
    w[1] = u[2] * v[0] - u[0] * v[2];
    w[2] = u[0] * v[1] - u[1] * v[0];
  };
  double c12[3], c20[3], c01[3];
  cross(a1, a2, c12);
  cross(a2, a0, c20);
  cross(a0, a1, c01);
  h.volume = a0[0] * c12[0] + a0[1] * c12[1] + a0[2] * c12[2];
  INQAMD_REQUIRE(h.volume > 1e-10, "the lattice volume is not positive");
  for (int j = 0; j < 3; j++) {
    h.rlat[0 * 3 + j] = 2.0 * M_PI / h.volume * c12[j];
    h.rlat[1 * 3 + j] = 2.0 * M_PI / h.volume * c20[j];
    h.rlat[2 * 3 + j] = 2.0 * M_PI / h.volume * c01[j];
  }
  for (int a = 0; a < 3; a++) {
    std::vector<int> rad;
    if (!fft_factor(h.n[a], rad))
      throw error{ERR_UNSUPPORTED, "grid size " + std::to_string(h.n[a]) +
                                       " has too many prime factors"};
    if (2L * h.n[a] * fft_band_tile(1, h.n[a]) * (long)sizeof(cplx) > 65536)
      throw error{ERR_UNSUPPORTED, "grid size " + std::to_string(h.n[a]) + " is too long"};
    h.axis[a].len = h.n[a];
    h.axis[a].nstages = (int)rad.size();
    for (size_t s = 0; s < rad.size(); s++) h.axis[a].radix[s] = rad[s];
    std::vector<cplx> tw = fft_twiddles(h.n[a]);
    h.axis[a].tw = (cplx *)dalloc(tw.size() * sizeof(cplx));
    INQAMD_HIP_CHECK(hipMemcpy(h.axis[a].tw, tw.data(), tw.size() * sizeof(cplx),
                               hipMemcpyHostToDevice));
  }
  {
    int dev = 0, lds = 0;
    INQAMD_HIP_CHECK(hipGetDevice(&dev));
    h.device = dev;
    // the allocation-free entry points' scratch on this device (its
    // address, resolved with the device current)
    (void)scratch_address(dev);
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeSharedMemPerBlockOptin, dev) != hipSuccess ||
        lds <= 0)
      INQAMD_HIP_CHECK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    h.max_lds = (size_t)lds;
    const char *path = getenv("INQAMD_HPSI_PATH");
    const bool pencil_forced = path && std::string(path) == "pencil";
    h.three_sweep = !pencil_forced && h.n[0] == h.n[1] && h.n[1] == h.n[2] &&
                    fft3_supported(h.n[0], h.max_lds);
    // kernel attributes are set now so that no later call (or capture) has to
    if (h.three_sweep) {
      fft3_prepare(h.n[0]);
      INQAMD_HIP_CHECK(hipStreamCreateWithFlags(&h.aux, hipStreamNonBlocking));
      INQAMD_HIP_CHECK(hipEventCreateWithFlags(&h.ev_fork, hipEventDisableTiming));
      INQAMD_HIP_CHECK(hipEventCreateWithFlags(&h.ev_join, hipEventDisableTiming));
      if (pipeline_streams() == 3) {
        INQAMD_HIP_CHECK(hipStreamCreateWithFlags(&h.aux2, hipStreamNonBlocking));
        INQAMD_HIP_CHECK(hipEventCreateWithFlags(&h.ev_join2, hipEventDisableTiming));
      }
      INQAMD_HIP_CHECK(hipEventCreateWithFlags(&h.ev_proj, hipEventDisableTiming));
    }
  }
  h.V = (double *)dalloc(h.npoint * sizeof(double));
  INQAMD_HIP_CHECK(hipMemcpy(h.V, d.vloc, h.npoint * sizeof(double), hipMemcpyHostToDevice));
  h.kA = (double *)dalloc(KA_NSLOTS * 3 * sizeof(double));
  std::vector<double> zeros(KA_NSLOTS * 3, 0.0);
  INQAMD_HIP_CHECK(hipMemcpy(h.kA, zeros.data(), zeros.size() * sizeof(double),
                             hipMemcpyHostToDevice));
  const size_t orb = (size_t)h.npoint * h.max_bands * sizeof(cplx);
  h.work = (cplx *)dalloc(orb);
  h.hn = (cplx *)dalloc(orb);
  h.hn2 = (cplx *)dalloc(orb);
  const long nchunk = (h.npoint + EN_CHUNK - 1) / EN_CHUNK;
  h.npartial = nchunk * h.max_bands * 2;
  h.partial = (double *)dalloc(h.npartial * sizeof(double));
  projector_upload(h, d.nproj, d.max_sphere, d.max_nlm, d.points, d.positions, d.matrices,
                   d.coeff, d.nlm, d.sphere_size);
  if (h.pt.nproj > 0)
    h.proj = (cplx *)dalloc((size_t)h.pt.nproj * h.pt.max_nlm * h.max_bands * sizeof(cplx));
  if (h.pt.nproj > 0 && h.pt.nunique > 0 && h.three_sweep)
    h.vnl = (cplx *)dalloc((size_t)h.pt.nent * h.max_bands * sizeof(cplx));
  if (h.pt.nproj > 0) {
    const size_t mom = (size_t)h.pt.nproj * h.pt.max_nlm * 4 * h.max_bands * sizeof(cplx);
    h.mom_a = (cplx *)dalloc(mom);
    h.mom_b = (cplx *)dalloc(mom);
    // 64 bands x nlm x 4 weights per workgroup of the moments kernel
    h.mom_partial_elems = MOM_PARTIAL_BLOCKS * 64 * h.pt.max_nlm * 4;
    h.mom_partial = (cplx *)dalloc(h.mom_partial_elems * sizeof(cplx));
    h.proj_part = (double *)dalloc(PROJ_PART_TILES * 256 * sizeof(double));
  }
  h.obs = (double *)dalloc(observables_scratch_doubles(h) * sizeof(double));
}

}  // namespace inqamd

// ===========================================================================
// C ABI
// ===========================================================================
using namespace inqamd;

namespace {

thread_local std::string g_create_error;

}  // namespace

extern "C" {

const char *inqamd_version(void) { return "inq_amd 0.1 (gfx950)"; }

int inqamd_selftest_mfma(double *out_device, void *stream) {
  if (!out_device) return INQAMD_ERR_ARGUMENT;
  try {
    mfma_selftest(out_device, (hipStream_t)stream);
  } catch (const inqamd::error &e) {
    g_create_error = e.msg;
    return e.code;
  }
  return INQAMD_OK;
}

int inqamd_selftest_mfma_4x4(double *out_device, void *stream) {
  if (!out_device) return INQAMD_ERR_ARGUMENT;
  try {
    mfma4_selftest(out_device, (hipStream_t)stream);
  } catch (const inqamd::error &e) {
    g_create_error = e.msg;
    return e.code;
  }
  return INQAMD_OK;
}

const char *inqamd_last_error(inqamd_hamiltonian_t h) {
  if (!h) return g_create_error.c_str();
  return h->last_error.c_str();
}

int inqamd_hamiltonian_create(const inqamd_system_desc *desc, void *stream,
                              inqamd_hamiltonian_t *out) {
  if (!desc || !out) {
    g_create_error = "null argument";
    return INQAMD_ERR_ARGUMENT;
  }
  *out = nullptr;
  inqamd_hamiltonian *h = nullptr;
  try {
    h = new inqamd_hamiltonian();
    create(*h, *desc, (hipStream_t)stream);
  } catch (const inqamd::error &e) {
    g_create_error = e.msg;
    if (h) { release(*h); delete h; }
    return e.code;
  } catch (const std::bad_alloc &) {
    g_create_error = "host allocation failed";
    if (h) { release(*h); delete h; }
    return INQAMD_ERR_MEMORY;
  }
  g_create_error.clear();
  *out = h;
  return INQAMD_OK;
}

void inqamd_hamiltonian_destroy(inqamd_hamiltonian_t h) {
  if (!h) return;
  (void)hipStreamSynchronize(h->stream);
  release(*h);
  delete h;
}

int inqamd_set_stream(inqamd_hamiltonian_t h, void *stream) {
  if (!h) return INQAMD_ERR_ARGUMENT;
  h->stream = (hipStream_t)stream;
  return INQAMD_OK;
}

int inqamd_set_tuning(inqamd_hamiltonian_t h, const inqamd_tuning *t) {
  if (!h || !t) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    INQAMD_REQUIRE(t->z_tile_y == 0 || t->z_tile_y == 8 || t->z_tile_y == 16,
                   "z_tile_y must be 0, 8 or 16");
    INQAMD_REQUIRE(t->xy_threads >= 0 && t->xy_threads <= 2, "xy_threads must be 0, 1 or 2");
    INQAMD_REQUIRE(t->band_chunk >= 0 && t->band_chunk % 8 == 0,
                   "band_chunk must be 0 or a positive multiple of 8");
    INQAMD_REQUIRE(t->projector_bands == 0 || t->projector_bands == 16 ||
                       t->projector_bands == 32 || t->projector_bands == 64,
                   "projector_bands must be 0, 16, 32 or 64");
    // two band chunks of the pipeline share the handle's max_bands-wide
    // intermediate, so a chunk may hold at most max_bands bands
    INQAMD_REQUIRE(t->band_chunk <= h->max_bands, "band_chunk exceeds max_bands");
    h->tune_z_ty = t->z_tile_y;
    h->tune_xy_wide = t->xy_threads;
    h->tune_chunk = t->band_chunk;
    h->tune_p4 = t->projector_bands;
  });
}

int inqamd_get_tuning(inqamd_hamiltonian_t h, inqamd_tuning *t) {
  if (!h || !t) return INQAMD_ERR_ARGUMENT;
  t->z_tile_y = h->tune_z_ty;
  t->xy_threads = h->tune_xy_wide;
  t->band_chunk = h->tune_chunk;
  t->projector_bands = h->tune_p4;
  return INQAMD_OK;
}

int inqamd_set_kpoint(inqamd_hamiltonian_t h, const double k_cov[3]) {
  if (!h || !k_cov) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    for (int d = 0; d < 3; d++) h->k_cov[d] = k_cov[d];
    h->generation++;
    update_current_slot(*h);
  });
}

int inqamd_set_uniform_vector_potential(inqamd_hamiltonian_t h, const double A_cov[3]) {
  if (!h || !A_cov) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    for (int d = 0; d < 3; d++) h->A_cov[d] = A_cov[d];
    update_current_slot(*h);
  });
}

int inqamd_set_scalar_potential(inqamd_hamiltonian_t h, const double *vloc_device) {
  if (!h) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    INQAMD_REQUIRE(vloc_device != nullptr, "vloc is null");
    INQAMD_HIP_CHECK(hipMemcpyAsync(h->V, vloc_device, h->npoint * sizeof(double),
                                    hipMemcpyDeviceToDevice, h->stream));
    h->generation++;
  });
}

int inqamd_hamiltonian_apply(inqamd_hamiltonian_t h, const void *psi, void *hpsi, int nst,
                             int ld) {
  if (!h) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    check_orbital(*h, psi, "psi");
    check_orbital(*h, hpsi, "hpsi");
    check_bands(*h, nst, ld, true);
    INQAMD_REQUIRE(psi != hpsi, "psi and hpsi must not alias");
    hamiltonian_apply(*h, (const cplx *)psi, ld, (cplx *)hpsi, ld, nst, KA_CURRENT);
  });
}

int inqamd_hamiltonian_apply_fourier(inqamd_hamiltonian_t h, const void *phi, void *hphi, int nst,
                                     int ld) {
  if (!h) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    check_orbital(*h, phi, "phi");
    check_orbital(*h, hphi, "hphi");
    check_bands(*h, nst, ld, true);
    INQAMD_REQUIRE(phi != hphi, "phi and hphi must not alias");
    // ks_hamiltonian.hpp:156-176 is F H F^-1 phi: to_real, the real-space
    // operator, to_fourier (the kinetic part then passes through the
    // transforms instead of being added in G space; equal to rounding)
    hamiltonian_fourier(*h, (const cplx *)phi, ld, (cplx *)hphi, ld, nst);
  });
}

int inqamd_non_local(inqamd_hamiltonian_t h, const void *psi, void *vnlpsi, int nst, int ld) {
  if (!h) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    check_orbital(*h, psi, "psi");
    check_orbital(*h, vnlpsi, "vnlpsi");
    check_bands(*h, nst, ld, true);
    INQAMD_REQUIRE(psi != vnlpsi, "psi and vnlpsi must not alias");
    // vnlphi.fill(0) then project + apply (ks_hamiltonian.hpp:119-126)
    INQAMD_HIP_CHECK(hipMemsetAsync(vnlpsi, 0, (size_t)h->npoint * ld * sizeof(cplx), h->stream));
    projector_phases(*h, KA_CURRENT);
    projector_project_expand(*h, (const cplx *)psi, ld, nst, nullptr, 0, true);
    projector_apply_coefficients(*h, (cplx *)vnlpsi, ld, nst, KA_CURRENT);
  });
}

int inqamd_transform_to_fourier(inqamd_hamiltonian_t h, const void *in, void *out, int nst,
                                int ld) {
  if (!h) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    check_orbital(*h, in, "in");
    check_orbital(*h, out, "out");
    check_bands(*h, nst, ld, false);
    fft3d(*h, (const cplx *)in, ld, (cplx *)out, ld, nst, -1, 1.0);
  });
}

int inqamd_transform_to_real(inqamd_hamiltonian_t h, const void *in, void *out, int nst, int ld,
                             int normalize) {
  if (!h) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    check_orbital(*h, in, "in");
    check_orbital(*h, out, "out");
    check_bands(*h, nst, ld, false);
    fft3d(*h, (const cplx *)in, ld, (cplx *)out, ld, nst, +1,
          normalize ? 1.0 / (double)h->npoint : 1.0);
  });
}

int inqamd_projector_project(inqamd_hamiltonian_t h, const void *psi, void *sphere, int nst,
                             int ld) {
  if (!h) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    check_orbital(*h, psi, "psi");
    check_orbital(*h, sphere, "sphere");
    check_bands(*h, nst, ld, true);
    projector_phases(*h, KA_CURRENT);
    projector_project_expand(*h, (const cplx *)psi, ld, nst, nullptr, 0, true);
    projector_expand(*h, (cplx *)sphere, nst);
  });
}

int inqamd_projector_apply(inqamd_hamiltonian_t h, const void *sphere, void *hpsi, int nst,
                           int ld) {
  if (!h) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    check_orbital(*h, sphere, "sphere");
    check_orbital(*h, hpsi, "hpsi");
    check_bands(*h, nst, ld, false);
    projector_phases(*h, KA_CURRENT);
    projector_apply_sphere(*h, (const cplx *)sphere, (cplx *)hpsi, ld, nst, KA_CURRENT);
  });
}

int inqamd_band_energies(inqamd_hamiltonian_t h, const void *hpsi, const void *psi, double *eps,
                         int nst, int ld) {
  if (!h) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    check_orbital(*h, hpsi, "hpsi");
    check_orbital(*h, psi, "psi");
    INQAMD_REQUIRE(eps != nullptr, "eps is null");
    check_bands(*h, nst, ld, true);
    band_energies(*h, (const cplx *)hpsi, ld, (const cplx *)psi, ld, nst, eps);
  });
}

int inqamd_band_energies_noalloc(inqamd_hamiltonian_t h, const void *hpsi, const void *psi,
                                 double *eps, int nst, int ld) {
  if (!h) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    check_orbital(*h, hpsi, "hpsi");
    check_orbital(*h, psi, "psi");
    INQAMD_REQUIRE(eps != nullptr, "eps is null");
    check_bands(*h, nst, ld, false);
    const long most = noalloc_max_bands(h->npoint);
    if (nst > most)
      throw error{ERR_ARGUMENT, "band_energies_noalloc: " + std::to_string(nst) +
                                    " bands exceed its 8 MiB scratch, which holds at most " +
                                    std::to_string(most) + " bands on this grid"};
    band_energies_with(*h, (const cplx *)hpsi, ld, (const cplx *)psi, ld, nst, eps,
                       scratch_address(h->device));
  });
}

int inqamd_energy(inqamd_hamiltonian_t h, const void *psi, int nst, int ld, double *eps,
                  double *normres, double *nl) {
  if (!h) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    check_orbital(*h, psi, "psi");
    check_bands(*h, nst, ld, true);
    INQAMD_REQUIRE(eps && normres && nl, "outputs are null");
    const cplx *f = (const cplx *)psi;
    // residual = H phi; eps = Re<residual|phi>/<phi|phi> (energy.hpp:74-75)
    hamiltonian_apply(*h, f, ld, h->hn, nst, nst, KA_CURRENT);
    band_energies(*h, h->hn, nst, f, ld, nst, eps);
    // normres = dV |residual - eps phi|^2 (:76-77)
    const int nchunk = (int)((h->npoint + EN_CHUNK - 1) / EN_CHUNK);
    hipLaunchKernelGGL(residual_partial_kernel, dim3(nchunk, (nst + EN_BT - 1) / EN_BT),
                       dim3(256), 0, h->stream, (const cplx *)h->hn, (long)nst, f, (long)ld,
                       (const double *)eps, nst, h->npoint, h->partial);
    INQAMD_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(column_sum_kernel, dim3((nst + 63) / 64), dim3(64), 0, h->stream,
                       (const double *)h->partial, nchunk, nst, h->volume / (double)h->npoint,
                       normres);
    INQAMD_HIP_CHECK(hipGetLastError());
    // nl = Re<V_nl phi|phi>/<phi|phi> (:81-85)
    INQAMD_HIP_CHECK(hipMemsetAsync(h->hn2, 0, (size_t)h->npoint * nst * sizeof(cplx), h->stream));
    projector_phases(*h, KA_CURRENT);
    projector_project_expand(*h, f, ld, nst, nullptr, 0, true);
    projector_apply_coefficients(*h, h->hn2, nst, nst, KA_CURRENT);
    band_energies(*h, h->hn2, nst, f, ld, nst, nl);
  });
}

int inqamd_shift(inqamd_hamiltonian_t h, double re, double im, const void *x, void *y, int nst,
                 int ld) {
  if (!h) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    check_orbital(*h, x, "x");
    check_orbital(*h, y, "y");
    check_bands(*h, nst, ld, false);
    shift(*h, std::complex<double>(re, im), (const cplx *)x, ld, (cplx *)y, ld, nst);
  });
}

int inqamd_exponential_in_place(inqamd_hamiltonian_t h, double zre, double zim, int order,
                                void *psi, int nst, int ld) {
  if (!h) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    check_orbital(*h, psi, "psi");
    check_bands(*h, nst, ld, true);
    INQAMD_REQUIRE(order >= 1, "order must be >= 1");
    exponential_in_place(*h, std::complex<double>(zre, zim), order, (cplx *)psi, ld, nst,
                         KA_CURRENT);
  });
}

int inqamd_exponential_2_for_1(inqamd_hamiltonian_t h, double z1re, double z1im, double z2re,
                               double z2im, int order, void *phi, void *expphi, int nst, int ld) {
  if (!h) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    check_orbital(*h, phi, "phi");
    check_orbital(*h, expphi, "expphi");
    check_bands(*h, nst, ld, true);
    INQAMD_REQUIRE(order >= 1, "order must be >= 1");
    INQAMD_REQUIRE(phi != expphi, "phi and expphi must not alias");
    exponential_2_for_1(*h, std::complex<double>(z1re, z1im), std::complex<double>(z2re, z2im),
                        order, (cplx *)phi, (cplx *)expphi, ld, ld, nst, KA_CURRENT);
  });
}

int inqamd_hamiltonian_profile(inqamd_hamiltonian_t h, const void *psi, void *hpsi, int nst,
                               int ld, double ms[4], int launches[4]) {
  if (!h) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    check_orbital(*h, psi, "psi");
    check_orbital(*h, hpsi, "hpsi");
    check_bands(*h, nst, ld, true);
    INQAMD_REQUIRE(ms && launches, "outputs are null");
    if (!h->three_sweep)
      throw error{ERR_UNSUPPORTED, "per-sweep timing needs the three-sweep Hpsi (cubic grid)"};
    const band_chunks plan = chunk_plan(*h, nst, ld, true);
    const int nchunk = plan.n;
    std::vector<hipEvent_t> ev(2 + 3 * nchunk, nullptr);
    auto cleanup = [&] {
      for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    };
    try {
      for (auto &e : ev) INQAMD_HIP_CHECK(hipEventCreate(&e));
      INQAMD_HIP_CHECK(hipEventRecord(ev[0], h->stream));
      projector_phases(*h, KA_CURRENT);
      hpsi_projector(*h, (const cplx *)psi, ld, nst);
      INQAMD_HIP_CHECK(hipEventRecord(ev[1], h->stream));
      int k = 1;
      for (int c = 0; c < nchunk; c++) {
        const int b0 = plan.b0(c), nb = plan.nb(c);
        for (int sweep = 0; sweep < 3; sweep++) {
          local_passes(*h, (const cplx *)psi + b0, ld, (cplx *)hpsi + b0, ld, nb, KA_CURRENT, b0,
                       nst, nullptr, sweep, nullptr, nullptr, plan.gskip, plan.glog);
          INQAMD_HIP_CHECK(hipEventRecord(ev[++k], h->stream));
        }
      }
      INQAMD_HIP_CHECK(hipEventSynchronize(ev[k]));
      float t = 0.0f;
      INQAMD_HIP_CHECK(hipEventElapsedTime(&t, ev[0], ev[1]));
      ms[0] = t;
      launches[0] = h->pt.nproj > 0 ? 2 : 0;
      for (int sweep = 0; sweep < 3; sweep++) {
        ms[1 + sweep] = 0.0;
        launches[1 + sweep] = nchunk;
      }
      for (int c = 0; c < nchunk; c++)
        for (int sweep = 0; sweep < 3; sweep++) {
          INQAMD_HIP_CHECK(hipEventElapsedTime(&t, ev[1 + 3 * c + sweep], ev[2 + 3 * c + sweep]));
          ms[1 + sweep] += t;
        }
    } catch (...) {
      cleanup();
      throw;
    }
    cleanup();
  });
}

int inqamd_etrs_step(inqamd_hamiltonian_t h, double dt, const double A_t_cov[3],
                     const double A_tdt_cov[3], int order, void *psi, int nst, int ld) {
  if (!h) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    check_orbital(*h, psi, "psi");
    check_bands(*h, nst, ld, true);
    INQAMD_REQUIRE(A_t_cov && A_tdt_cov, "vector potentials are required");
    INQAMD_REQUIRE(order >= 1, "order must be >= 1");
    double kt[3], ktdt[3];
    for (int d = 0; d < 3; d++) {
      kt[d] = h->k_cov[d] + A_t_cov[d];
      ktdt[d] = h->k_cov[d] + A_tdt_cov[d];
    }
    set_slot(*h, KA_T, kt);
    set_slot(*h, KA_TDT, ktdt);
    const std::complex<double> z(0.0, dt / 2.0);  // etrs.hpp:36, :67
    exponential_in_place(*h, z, order, (cplx *)psi, ld, nst, KA_T);
    exponential_in_place(*h, z, order, (cplx *)psi, ld, nst, KA_TDT);
    for (int d = 0; d < 3; d++) h->A_cov[d] = A_tdt_cov[d];
    update_current_slot(*h);
  });
}

}  // extern "C"

struct inqamd_graph {
  inqamd_hamiltonian *h = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t capture = nullptr;
  double *step = nullptr;
  bool series = false;
};

namespace {

void graph_release(inqamd_graph *g) {
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->capture) (void)hipStreamDestroy(g->capture);
  if (g->step) (void)hipFree(g->step);
}

}  // namespace

extern "C" {

int inqamd_etrs_graph_create(inqamd_hamiltonian_t h, void *psi, int nst, int ld, double dt,
                             int order, const inqamd_vector_potential_series *series,
                             inqamd_graph_t *out) {
  if (!h || !out) return INQAMD_ERR_ARGUMENT;
  *out = nullptr;
  inqamd_graph *g = new inqamd_graph();
  g->h = h;
  int rc = guarded(h, [&] {
    check_orbital(*h, psi, "psi");
    check_bands(*h, nst, ld, true);
    INQAMD_REQUIRE(order >= 1, "order must be >= 1");
    series_args sa{};
    if (series) {
      INQAMD_REQUIRE(series->n >= 0 && series->n <= 64, "at most 64 series coefficients");
      INQAMD_REQUIRE(series->n == 0 || series->coeff, "series coefficients are null");
      sa.n = series->n;
      for (int j = 0; j < sa.n; j++) sa.c[j] = series->coeff[j];
      sa.omega = series->omega;
      sa.dt = dt;
      // to_covariant(cartesian) = lattice . v (cell.hpp:184-187)
      for (int i = 0; i < 3; i++) {
        sa.dir_cov[i] = h->lattice[3 * i + 0] * series->direction[0] +
                        h->lattice[3 * i + 1] * series->direction[1] +
                        h->lattice[3 * i + 2] * series->direction[2];
        sa.k_cov[i] = h->k_cov[i];
      }
      g->series = true;
      INQAMD_HIP_CHECK(hipMalloc(&g->step, sizeof(double)));
      const double zero = 0.0;
      INQAMD_HIP_CHECK(hipMemcpy(g->step, &zero, sizeof(double), hipMemcpyHostToDevice));
    }
    INQAMD_HIP_CHECK(hipStreamCreateWithFlags(&g->capture, hipStreamNonBlocking));
    // make the pending work of the handle's stream (tables, V) visible first
    INQAMD_HIP_CHECK(hipStreamSynchronize(h->stream));
    hipStream_t saved = h->stream;
    h->stream = g->capture;
    INQAMD_HIP_CHECK(hipStreamBeginCapture(g->capture, hipStreamCaptureModeThreadLocal));
    try {
      if (g->series) {
        hipLaunchKernelGGL(advance_series_kernel, dim3(1), dim3(64), 0, g->capture, sa, g->step,
                           h->kA);
        INQAMD_HIP_CHECK(hipGetLastError());
      }
      const std::complex<double> z(0.0, dt / 2.0);  // etrs.hpp:36, :67
      exponential_in_place(*h, z, order, (cplx *)psi, ld, nst, KA_T);
      exponential_in_place(*h, z, order, (cplx *)psi, ld, nst, KA_TDT);
    } catch (...) {
      hipGraph_t broken = nullptr;
      (void)hipStreamEndCapture(g->capture, &broken);
      if (broken) (void)hipGraphDestroy(broken);
      h->stream = saved;
      throw;
    }
    h->stream = saved;
    INQAMD_HIP_CHECK(hipStreamEndCapture(g->capture, &g->graph));
    INQAMD_HIP_CHECK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
  });
  if (rc != INQAMD_OK) {
    graph_release(g);
    delete g;
    return rc;
  }
  *out = g;
  return INQAMD_OK;
}

int inqamd_etrs_graph_launch(inqamd_graph_t g, const double A_t_cov[3], const double A_tdt_cov[3]) {
  if (!g || !g->h) return INQAMD_ERR_ARGUMENT;
  inqamd_hamiltonian *h = g->h;
  return guarded(h, [&] {
    if (!g->series || A_t_cov || A_tdt_cov) {
      INQAMD_REQUIRE(A_t_cov && A_tdt_cov, "vector potentials are required for this graph");
      INQAMD_REQUIRE(!g->series, "a series graph computes A(t) itself");
      double kt[3], ktdt[3];
      for (int d = 0; d < 3; d++) {
        kt[d] = h->k_cov[d] + A_t_cov[d];
        ktdt[d] = h->k_cov[d] + A_tdt_cov[d];
      }
      set_slot(*h, KA_T, kt);
      set_slot(*h, KA_TDT, ktdt);
      for (int d = 0; d < 3; d++) h->A_cov[d] = A_tdt_cov[d];
    }
    INQAMD_HIP_CHECK(hipGraphLaunch(g->exec, h->stream));
    if (!g->series) update_current_slot(*h);
  });
}

int inqamd_etrs_graph_set_step(inqamd_graph_t g, long step) {
  if (!g || !g->h) return INQAMD_ERR_ARGUMENT;
  return guarded(g->h, [&] {
    INQAMD_REQUIRE(g->series, "only series graphs have a step counter");
    hipLaunchKernelGGL(set_scalar_kernel, dim3(1), dim3(64), 0, g->h->stream, g->step,
                       (double)step);
    INQAMD_HIP_CHECK(hipGetLastError());
  });
}

void inqamd_graph_destroy(inqamd_graph_t g) {
  if (!g) return;
  if (g->h) (void)hipStreamSynchronize(g->h->stream);
  graph_release(g);
  delete g;
}

}  // extern "C"
