// Pencil-pass FFT kernels for gfx950 (see fft.hpp).
//
// Stockham autosort formulation: stage s with radix R and span Ns reads
// butterfly inputs j + r*len/R from one LDS buffer, applies twiddles
// W_len^{r (j mod Ns) len/(Ns R)}, a radix-R DFT, and writes outputs to
// (j/Ns)*Ns*R + (j mod Ns) + r*Ns of the other buffer, so the result ends in
// natural order without a bit-reversal pass.  All arithmetic is FP64;
// twiddles come from a host table evaluated in long double.

#include "fft.hpp"
#include "fft_dft.hpp"

#include <cmath>
#include <cstdlib>
#include <stdexcept>

namespace inqamd {

// One Stockham stage with a compile-time radix.
template <int R, int B>
__device__ __forceinline__ void stage(const cplx *__restrict__ src, cplx *__restrict__ dst,
                                      int len, int Ns, const cplx *__restrict__ tw, int sign) {
  const int nb = len / R;
  const int tstride = len / (Ns * R);
  for (int w = threadIdx.x; w < nb * B; w += blockDim.x) {
    const int j = w / B;
    const int bb = w % B;
    const int k = j % Ns;
    cplx v[R];
#pragma unroll
    for (int r = 0; r < R; r++) v[r] = src[(j + r * nb) * B + bb];
    if (k != 0) {
#pragma unroll
      for (int r = 1; r < R; r++) v[r] = cmul(v[r], twiddle(tw, r * k * tstride, sign));
    }
    dft<R>(v, sign);
    const int base = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; r++) dst[(base + r * Ns) * B + bb] = v[r];
  }
}

// Any radix, O(R^2) per butterfly: up to FFT_MAX_GENERIC_RADIX through a
// register array of the twiddled inputs; larger primes read each input from
// LDS once per output (the same products in the same order).
template <int B>
__device__ __noinline__ void stage_generic(const cplx *__restrict__ src, cplx *__restrict__ dst,
                                           int len, int R, int Ns, const cplx *__restrict__ tw,
                                           int sign) {
  const int nb = len / R;
  const int tstride = len / (Ns * R);
  if (R > FFT_MAX_GENERIC_RADIX) {
    for (int w = threadIdx.x; w < nb * B; w += blockDim.x) {
      const int j = w / B;
      const int bb = w % B;
      const int k = j % Ns;
      const int base = (j / Ns) * Ns * R + k;
      for (int s = 0; s < R; s++) {
        cplx acc = src[j * B + bb];
        for (int q = 1; q < R; q++) {
          cplx x = src[(j + q * nb) * B + bb];
          if (k != 0) x = cmul(x, twiddle(tw, q * k * tstride, sign));
          acc = cadd(acc, cmul(x, twiddle(tw, ((q * s) % R) * nb, sign)));
        }
        dst[(base + s * Ns) * B + bb] = acc;
      }
    }
    return;
  }
  for (int w = threadIdx.x; w < nb * B; w += blockDim.x) {
    const int j = w / B;
    const int bb = w % B;
    const int k = j % Ns;
    cplx v[FFT_MAX_GENERIC_RADIX];
    for (int r = 0; r < R; r++) {
      cplx x = src[(j + r * nb) * B + bb];
      v[r] = (r == 0 || k == 0) ? x : cmul(x, twiddle(tw, r * k * tstride, sign));
    }
    const int base = (j / Ns) * Ns * R + k;
    for (int s = 0; s < R; s++) {
      cplx acc = v[0];
      for (int q = 1; q < R; q++) acc = cadd(acc, cmul(v[q], twiddle(tw, ((q * s) % R) * nb, sign)));
      dst[(base + s * Ns) * B + bb] = acc;
    }
  }
}

template <int B, bool GENERIC>
__device__ __forceinline__ int run_stages(cplx *buf0, cplx *buf1, int cur, const fft_pass_args &a,
                                          int sign) {
  int Ns = 1;
  for (int s = 0; s < a.nstages; s++) {
    const int R = a.radix[s];
    const cplx *src = cur ? buf1 : buf0;
    cplx *dst = cur ? buf0 : buf1;
    switch (R) {
      case 2: stage<2, B>(src, dst, a.len, Ns, a.tw, sign); break;
      case 3: stage<3, B>(src, dst, a.len, Ns, a.tw, sign); break;
      case 4: stage<4, B>(src, dst, a.len, Ns, a.tw, sign); break;
      case 5: stage<5, B>(src, dst, a.len, Ns, a.tw, sign); break;
      case 8: stage<8, B>(src, dst, a.len, Ns, a.tw, sign); break;
      default:
        if constexpr (GENERIC) stage_generic<B>(src, dst, a.len, R, Ns, a.tw, sign);
        break;
    }
    __syncthreads();
    cur ^= 1;
    Ns *= R;
  }
  return cur;
}


template <int B, int MODE, bool GENERIC>
__global__ __launch_bounds__(FFT_BLOCK) void fft_pass_kernel(const fft_pass_args a) {
  extern __shared__ cplx lds[];
  cplx *buf0 = lds;
  cplx *buf1 = lds + a.len * B;

  const int nbt = (a.nst + B - 1) / B;
  long t = blockIdx.x;
  const int bt = (int)(t % nbt);
  t /= nbt;
  const int o1 = (int)(t % a.c1);
  const int o2 = (int)(t / a.c1);
  const long base = o1 * a.s1 + o2 * a.s2;
  const int b0 = bt * B;
  const int nel = a.len * B;

  for (int e = threadIdx.x; e < nel; e += blockDim.x) {
    const int i = e / B;
    const int b = b0 + e % B;
    cplx v = mk(0.0, 0.0);
    if (b < a.nst) v = a.in[(base + i * a.stride) * a.in_ld + b];
    buf0[e] = v;
  }
  __syncthreads();

  int cur = run_stages<B, GENERIC>(buf0, buf1, 0, a, MODE == FFT_KINETIC ? -1 : a.sign);

  if constexpr (MODE == FFT_KINETIC) {
    // x 1/2 |G + k + A|^2 : laplacian(-0.5, -2(k+A)) (laplacian.hpp:70-79) plus the
    // real-space 1/2|k+A|^2 shift (ks_hamiltonian.hpp:146), folded into one factor.
    const double inv2pi = 0.15915494309189533576888376337251;
    double kc[3];
#pragma unroll
    for (int d = 0; d < 3; d++)
      kc[d] = (a.rlat[0 * 3 + d] * a.kA[0] + a.rlat[1 * 3 + d] * a.kA[1] + a.rlat[2 * 3 + d] * a.kA[2]) *
              inv2pi;
    int c[3];
    c[a.ax1] = o1;
    c[a.ax2] = o2;
    cplx *buf = cur ? buf1 : buf0;
    for (int e = threadIdx.x; e < nel; e += blockDim.x) {
      c[a.ax] = e / B;
      const int m0 = symmetric_index(c[0], a.n[0]);
      const int m1 = symmetric_index(c[1], a.n[1]);
      const int m2 = symmetric_index(c[2], a.n[2]);
      double q2 = 0.0;
#pragma unroll
      for (int d = 0; d < 3; d++) {
        double q = m0 * a.rlat[0 * 3 + d] + m1 * a.rlat[1 * 3 + d] + m2 * a.rlat[2 * 3 + d] + kc[d];
        q2 += q * q;
      }
      buf[e] = cscale(0.5 * q2, buf[e]);
    }
    __syncthreads();
    cur = run_stages<B, GENERIC>(buf0, buf1, cur, a, +1);
  }

  const cplx *res = cur ? buf1 : buf0;
  for (int e = threadIdx.x; e < nel; e += blockDim.x) {
    const int i = e / B;
    const int b = b0 + e % B;
    if (b >= a.nst) continue;
    const long p = base + i * a.stride;
    cplx v = cscale(a.scale, res[e]);
    if constexpr (MODE == FFT_EPILOGUE) {
      const double vp = a.V[p];
      const cplx ps = a.psi[p * a.psi_ld + b];
      v.x += vp * ps.x;
      v.y += vp * ps.y;
    }
    a.out[p * a.out_ld + b] = v;
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
bool fft_factor(int len, std::vector<int> &radices) {
  radices.clear();
  int n = len;
  while (n % 8 == 0) { radices.push_back(8); n /= 8; }
  if (n % 4 == 0) { radices.push_back(4); n /= 4; }
  if (n % 2 == 0) { radices.push_back(2); n /= 2; }
  while (n % 3 == 0) { radices.push_back(3); n /= 3; }
  while (n % 5 == 0) { radices.push_back(5); n /= 5; }
  for (int p = 7; n > 1 && (p <= FFT_MAX_GENERIC_RADIX || p * p <= n); p += 2) {
    while (n % p == 0) { radices.push_back(p); n /= p; }
  }
  if (n > 1) radices.push_back(n);  // a prime above FFT_MAX_GENERIC_RADIX
  return (int)radices.size() <= FFT_MAX_STAGES;
}

std::vector<cplx> fft_twiddles(int len) {
  std::vector<cplx> tw(len);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int k = 0; k < len; k++) {
    long double ang = -2.0L * pi * (long double)k / (long double)len;
    tw[k] = mk((double)cosl(ang), (double)sinl(ang));
  }
  return tw;
}

int fft_band_tile(int nst, int len) {
  int B = nst >= 16 ? 16 : (nst >= 8 ? 8 : 4);
  // two LDS buffers of len*B complex must stay within 64 KiB
  while (B > 4 && 2L * len * B * (long)sizeof(cplx) > 65536) B /= 2;
  return B;
}

template <int B, int MODE, bool GENERIC>
static void launch_t(const fft_pass_args &a, hipStream_t stream) {
  const long nbt = (a.nst + B - 1) / B;
  const long nblocks = nbt * a.c1 * a.c2;
  const size_t lds = 2 * (size_t)a.len * B * sizeof(cplx);
  hipLaunchKernelGGL((fft_pass_kernel<B, MODE, GENERIC>), dim3((unsigned)nblocks), dim3(FFT_BLOCK),
                     lds, stream, a);
  INQAMD_HIP_CHECK(hipGetLastError());
}

template <int MODE, bool GENERIC>
static void launch_gen(const fft_pass_args &a, int B, hipStream_t stream) {
  switch (B) {
    case 16: launch_t<16, MODE, GENERIC>(a, stream); break;
    case 8: launch_t<8, MODE, GENERIC>(a, stream); break;
    case 4: launch_t<4, MODE, GENERIC>(a, stream); break;
    default: throw error{ERR_UNSUPPORTED, "unsupported band tile"};
  }
}

// Plans whose radices are all in {2, 3, 4, 5, 8} run a kernel without the
// generic-radix path (which needs private arrays and so scratch memory).
template <int MODE>
static void launch_mode(const fft_pass_args &a, int B, hipStream_t stream) {
  bool generic = false;
  for (int s = 0; s < a.nstages; s++) {
    const int r = a.radix[s];
    if (r != 2 && r != 3 && r != 4 && r != 5 && r != 8) generic = true;
  }
  if (generic)
    launch_gen<MODE, true>(a, B, stream);
  else
    launch_gen<MODE, false>(a, B, stream);
}

// Tile shape of the register-resident pass: band tile 16 for wide band
// sets, 8 below 16 bands.  INQAMD_FFT_VARIANT overrides the choice (tuning).
static int reg_variant(const fft_pass_args &a) {
  static int forced = [] {
    const char *e = getenv("INQAMD_FFT_VARIANT");
    return e ? atoi(e) : -1;
  }();
  if (forced >= 0) return forced;
  return a.nst >= 16 ? 0 : 3;
}

void fft_launch_pass(const fft_pass_args &a, int mode, hipStream_t stream) {
  const int v = reg_variant(a);
  if (v < 10 && fft_plan_launch(a, mode, a.nst >= 16 ? 16 : 8, stream)) return;
  if (fft_reg_supported(a)) {
    fft_reg_launch(a, mode, v % 10, stream);
    return;
  }
  const int B = fft_band_tile(a.nst, a.len);
  if (2L * a.len * B * (long)sizeof(cplx) > 65536)
    throw error{ERR_UNSUPPORTED, "FFT axis too long for the LDS pencil pass"};
  switch (mode) {
    case FFT_PLAIN: launch_mode<FFT_PLAIN>(a, B, stream); break;
    case FFT_KINETIC: launch_mode<FFT_KINETIC>(a, B, stream); break;
    case FFT_EPILOGUE: launch_mode<FFT_EPILOGUE>(a, B, stream); break;
    default: throw error{ERR_ARGUMENT, "unknown FFT pass mode"};
  }
}

}  // namespace inqamd
