// Register-resident pencil-pass FFT for plans whose radices are all in
// {2, 3, 4, 5, 8}.
//
// A pass transforms a tile of `len` points x B contiguous bands.  Each thread
// owns whole radix-R butterflies: the first Stockham stage reads its inputs
// straight from HBM into registers and the last stage writes its outputs
// straight from registers to HBM, so LDS is only used to exchange data
// between stages -- in place, with a barrier between the reads and the
// writes of a stage.  A 64-point axis (8 x 8) therefore makes one LDS round
// trip per pass, and the LDS footprint is one len x B tile.
//
// Modes: FFT_PLAIN (one transform), FFT_KINETIC (forward, x 1/2|G+k+A|^2,
// backward, all in the same pass), FFT_EPILOGUE (scale x FFT + V psi).

#include "fft.hpp"
#include "fft_dft.hpp"

namespace inqamd {

enum : int { DST_LDS = 0, DST_GLOBAL = 1, DST_KINETIC = 2 };

template <int R, int B, int NT, int MAXLEN, class LD, class ST>
__device__ __forceinline__ void reg_stage(int len, int Ns, const cplx *__restrict__ tw, int sign,
                                          LD ld, ST st, bool sync_mid) {
  constexpr int NBF = (MAXLEN * B / R + NT - 1) / NT;
  const int nb = len / R;
  const int nbf = nb * B;
  const int tstride = len / (Ns * R);
  cplx v[NBF][R];
#pragma unroll
  for (int q = 0; q < NBF; q++) {
    const int w = threadIdx.x + q * NT;
    if (w < nbf) {
      const int j = w / B, bb = w % B, k = j % Ns;
#pragma unroll
      for (int r = 0; r < R; r++) v[q][r] = ld(j + r * nb, bb);
      if (k != 0) {
#pragma unroll
        for (int r = 1; r < R; r++) v[q][r] = cmul(v[q][r], twiddle(tw, r * k * tstride, sign));
      }
      dft<R>(v[q], sign);
    }
  }
  if (sync_mid) __syncthreads();
#pragma unroll
  for (int q = 0; q < NBF; q++) {
    const int w = threadIdx.x + q * NT;
    if (w < nbf) {
      const int j = w / B, bb = w % B, k = j % Ns;
      const int base = (j / Ns) * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; r++) st(base + r * Ns, bb, v[q][r]);
    }
  }
}

template <int R, int B, int NT, int MAXLEN, class GL, class LL, class GS, class LS, class KS>
__device__ __forceinline__ void reg_dispatch(bool src_global, int dst, int len, int Ns,
                                             const cplx *tw, int sign, GL gl, LL ll, GS gs, LS ls,
                                             KS ks) {
  if (src_global) {
    if (dst == DST_GLOBAL) reg_stage<R, B, NT, MAXLEN>(len, Ns, tw, sign, gl, gs, false);
    else if (dst == DST_KINETIC) reg_stage<R, B, NT, MAXLEN>(len, Ns, tw, sign, gl, ks, false);
    else reg_stage<R, B, NT, MAXLEN>(len, Ns, tw, sign, gl, ls, false);
  } else {
    if (dst == DST_GLOBAL) reg_stage<R, B, NT, MAXLEN>(len, Ns, tw, sign, ll, gs, false);
    else if (dst == DST_KINETIC) reg_stage<R, B, NT, MAXLEN>(len, Ns, tw, sign, ll, ks, true);
    else reg_stage<R, B, NT, MAXLEN>(len, Ns, tw, sign, ll, ls, true);
  }
}

template <int B, int NT, int MAXLEN, class GL, class LL, class GS, class LS, class KS>
__device__ __forceinline__ void reg_sequence(const fft_pass_args &a, int sign, bool from_global,
                                             int to, GL gl, LL ll, GS gs, LS ls, KS ks) {
  int Ns = 1;
  for (int s = 0; s < a.nstages; s++) {
    const int R = a.radix[s];
    const bool src_global = from_global && s == 0;
    const int dst = (s == a.nstages - 1) ? to : DST_LDS;
    switch (R) {
      case 8: reg_dispatch<8, B, NT, MAXLEN>(src_global, dst, a.len, Ns, a.tw, sign, gl, ll, gs, ls, ks); break;
      case 4: reg_dispatch<4, B, NT, MAXLEN>(src_global, dst, a.len, Ns, a.tw, sign, gl, ll, gs, ls, ks); break;
      case 2: reg_dispatch<2, B, NT, MAXLEN>(src_global, dst, a.len, Ns, a.tw, sign, gl, ll, gs, ls, ks); break;
      case 3: reg_dispatch<3, B, NT, MAXLEN>(src_global, dst, a.len, Ns, a.tw, sign, gl, ll, gs, ls, ks); break;
      case 5: reg_dispatch<5, B, NT, MAXLEN>(src_global, dst, a.len, Ns, a.tw, sign, gl, ll, gs, ls, ks); break;
      default: break;  // excluded by fft_reg_supported()
    }
    if (dst != DST_GLOBAL) __syncthreads();
    Ns *= R;
  }
}

template <int B, int NT, int MAXLEN, int MODE>
__global__ __launch_bounds__(NT) void fft_reg_kernel(const fft_pass_args a) {
  __shared__ cplx lds[MAXLEN * B];

  const int nbt = (a.nst + B - 1) / B;
  long t = blockIdx.x;
  const int bt = (int)(t % nbt);
  t /= nbt;
  const int o1 = (int)(t % a.c1);
  const int o2 = (int)(t / a.c1);
  const long base = o1 * a.s1 + o2 * a.s2;
  const int b0 = bt * B;

  auto gl = [&](int i, int bb) -> cplx {
    const int b = b0 + bb;
    return b < a.nst ? a.in[(base + (long)i * a.stride) * a.in_ld + b] : mk(0.0, 0.0);
  };
  auto ll = [&](int i, int bb) -> cplx { return lds[i * B + bb]; };
  auto ls = [&](int i, int bb, cplx v) { lds[i * B + bb] = v; };
  auto gs = [&](int i, int bb, cplx v) {
    const int b = b0 + bb;
    if (b >= a.nst) return;
    const long p = base + (long)i * a.stride;
    v = cscale(a.scale, v);
    if constexpr (MODE == FFT_EPILOGUE) {
      const double vp = a.V[p];
      const cplx ps = a.psi[p * a.psi_ld + b];
      v.x += vp * ps.x;
      v.y += vp * ps.y;
    }
    a.out[p * a.out_ld + b] = v;
  };

  if constexpr (MODE == FFT_KINETIC) {
    // 1/2 |G + k + A|^2 with G = sum_d m_d b_d (cartesian); the two fixed
    // coordinates of the pencil and k + A are folded into gfix.
    const double inv2pi = 0.15915494309189533576888376337251;
    const int m1 = symmetric_index(o1, a.n[a.ax1]);
    const int m2 = symmetric_index(o2, a.n[a.ax2]);
    double gfix[3], bax[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const double kc = (a.rlat[0 * 3 + d] * a.kA[0] + a.rlat[1 * 3 + d] * a.kA[1] +
                         a.rlat[2 * 3 + d] * a.kA[2]) * inv2pi;
      gfix[d] = m1 * a.rlat[a.ax1 * 3 + d] + m2 * a.rlat[a.ax2 * 3 + d] + kc;
      bax[d] = a.rlat[a.ax * 3 + d];
    }
    const int nax = a.n[a.ax];
    auto ks = [&](int i, int bb, cplx v) {
      const int m = symmetric_index(i, nax);
      const double q0 = m * bax[0] + gfix[0], q1 = m * bax[1] + gfix[1], q2 = m * bax[2] + gfix[2];
      lds[i * B + bb] = cscale(0.5 * (q0 * q0 + q1 * q1 + q2 * q2), v);
    };
    reg_sequence<B, NT, MAXLEN>(a, -1, true, DST_KINETIC, gl, ll, gs, ls, ks);
    reg_sequence<B, NT, MAXLEN>(a, +1, false, DST_GLOBAL, gl, ll, gs, ls, ks);
  } else {
    reg_sequence<B, NT, MAXLEN>(a, a.sign, true, DST_GLOBAL, gl, ll, gs, ls, ls);
  }
}

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------
bool fft_reg_supported(const fft_pass_args &a) {
  for (int s = 0; s < a.nstages; s++) {
    const int r = a.radix[s];
    if (r != 2 && r != 3 && r != 4 && r != 5 && r != 8) return false;
  }
  return a.len <= 128;
}

template <int B, int NT, int MAXLEN, int MODE>
static void reg_launch(const fft_pass_args &a, hipStream_t stream) {
  const long nbt = (a.nst + B - 1) / B;
  const long nblocks = nbt * a.c1 * a.c2;
  hipLaunchKernelGGL((fft_reg_kernel<B, NT, MAXLEN, MODE>), dim3((unsigned)nblocks), dim3(NT), 0,
                     stream, a);
  INQAMD_HIP_CHECK(hipGetLastError());
}

template <int MODE>
static void reg_launch_mode(const fft_pass_args &a, int variant, hipStream_t stream) {
  if (a.len <= 64) {
    switch (variant) {
      case 1: reg_launch<32, 256, 64, MODE>(a, stream); return;
      case 2: reg_launch<16, 256, 64, MODE>(a, stream); return;
      case 3: reg_launch<8, 64, 64, MODE>(a, stream); return;
      default: reg_launch<16, 128, 64, MODE>(a, stream); return;
    }
  }
  switch (variant) {
    case 3: reg_launch<8, 128, 128, MODE>(a, stream); return;
    default: reg_launch<16, 256, 128, MODE>(a, stream); return;
  }
}

void fft_reg_launch(const fft_pass_args &a, int mode, int variant, hipStream_t stream) {
  switch (mode) {
    case FFT_PLAIN: reg_launch_mode<FFT_PLAIN>(a, variant, stream); break;
    case FFT_KINETIC: reg_launch_mode<FFT_KINETIC>(a, variant, stream); break;
    case FFT_EPILOGUE: reg_launch_mode<FFT_EPILOGUE>(a, variant, stream); break;
    default: throw error{ERR_ARGUMENT, "unknown FFT pass mode"};
  }
}

}  // namespace inqamd
