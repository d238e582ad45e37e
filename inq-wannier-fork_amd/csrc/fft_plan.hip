// Plan-specialised register-resident pencil passes.
//
// Same algorithm as fft_reg.hip (first Stockham stage reads HBM into
// registers, last stage writes registers to HBM, LDS only between stages),
// but the radix sequence is a template parameter, so every stage's span,
// butterfly count and twiddle stride are compile-time constants and a kernel
// holds only the registers its own radices need.  The Hpsi epilogue loads
// the psi and V values of the elements a thread will write before the FFT
// stages start, so their latency overlaps the transform.
//
// Instantiated for the grid lengths of the workloads (24 ... 96) and a few
// small test sizes; other lengths use the runtime-plan kernels.

#include "fft.hpp"
#include "fft_dft.hpp"
#include "fft_stage.hpp"

namespace inqamd {

template <int B, int MODE, int R0, int R1, int R2>
__global__ __launch_bounds__(plan_threads(R0 * R1 * R2, B)) void fft_plan_kernel(
    const fft_pass_args a) {
  using P = plan3<R0, R1, R2>;
  constexpr int LEN = P::LEN;
  constexpr int NT = plan_threads(LEN, B);
  constexpr int RL = P::NSTAGE == 1 ? R0 : (P::NSTAGE == 2 ? R1 : R2);  // last radix
  constexpr int NBL = LEN / RL;
  constexpr int NBFL = (NBL * B + NT - 1) / NT;
  __shared__ cplx lds[LEN * B];

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
  auto ls = [&](int i, int bb, cplx v, int, int) { lds[i * B + bb] = v; };

  if constexpr (MODE == FFT_EPILOGUE) {
    // prefetch V psi for the outputs of the last stage: thread element
    // (q, r) is pencil point j + r*NBL with j = w / B (span NBL = LEN/RL)
    cplx pre[NBFL][RL];
#pragma unroll
    for (int q = 0; q < NBFL; q++) {
      const int w = threadIdx.x + q * NT;
      const int j = w / B, bb = w % B, b = b0 + bb;
#pragma unroll
      for (int r = 0; r < RL; r++) {
        pre[q][r] = mk(0.0, 0.0);
        if ((NBFL * NT == NBL * B || w < NBL * B) && b < a.nst) {
          const long p = base + (long)(j + r * NBL) * a.stride;
          const cplx ps = a.psi[p * a.psi_ld + b];
          const double vp = a.V[p];
          pre[q][r] = mk(vp * ps.x, vp * ps.y);
        }
      }
    }
    // the last stage stores butterfly slot (q, r) at pencil point j + r*NBL,
    // the point whose V psi was prefetched into pre[q][r]
    auto gs = [&](int i, int bb, cplx v, int q, int r) {
      const int b = b0 + bb;
      const cplx add = pre[q][r];
      if (b >= a.nst) return;
      const long p = base + (long)i * a.stride;
      v = cscale(a.scale, v);
      v.x += add.x;
      v.y += add.y;
      a.out[p * a.out_ld + b] = v;
    };
    run_plan<P, R0, R1, R2, B, NT>(a.tw, a.sign, false, gl, ll, ls, gs, false);
  } else {
    auto gs = [&](int i, int bb, cplx v, int, int) {
      const int b = b0 + bb;
      if (b >= a.nst) return;
      const long p = base + (long)i * a.stride;
      a.out[p * a.out_ld + b] = cscale(a.scale, v);
    };
    if constexpr (MODE == FFT_KINETIC) {
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
      auto ks = [&](int i, int bb, cplx v, int, int) {
        const int m = symmetric_index(i, LEN);
        const double q0 = m * bax[0] + gfix[0], q1 = m * bax[1] + gfix[1], q2 = m * bax[2] + gfix[2];
        lds[i * B + bb] = cscale(0.5 * (q0 * q0 + q1 * q1 + q2 * q2), v);
      };
      run_plan<P, R0, R1, R2, B, NT>(a.tw, -1, false, gl, ll, ls, ks, true);
      run_plan<P, R0, R1, R2, B, NT>(a.tw, +1, true, ll, ll, ls, gs, false);
    } else {
      run_plan<P, R0, R1, R2, B, NT>(a.tw, a.sign, false, gl, ll, ls, gs, false);
    }
  }
}

// ---------------------------------------------------------------------------
// host: (length -> radix sequence) table
// ---------------------------------------------------------------------------
template <int B, int MODE, int R0, int R1, int R2>
static void plan_launch(const fft_pass_args &a, hipStream_t stream) {
  constexpr int NT = plan_threads(R0 * R1 * R2, B);
  const long nbt = (a.nst + B - 1) / B;
  const long nblocks = nbt * a.c1 * a.c2;
  hipLaunchKernelGGL((fft_plan_kernel<B, MODE, R0, R1, R2>), dim3((unsigned)nblocks), dim3(NT), 0,
                     stream, a);
  INQAMD_HIP_CHECK(hipGetLastError());
}

#define INQAMD_PLAN_TABLE(X) \
  X(8, 8, 1, 1)              \
  X(12, 4, 3, 1)             \
  X(16, 4, 4, 1)             \
  X(20, 4, 5, 1)             \
  X(24, 8, 3, 1)             \
  X(30, 2, 3, 5)             \
  X(32, 8, 4, 1)             \
  X(36, 4, 3, 3)             \
  X(40, 8, 5, 1)             \
  X(48, 8, 2, 3)             \
  X(60, 4, 3, 5)             \
  X(64, 8, 8, 1)             \
  X(72, 8, 3, 3)             \
  X(80, 8, 2, 5)             \
  X(96, 8, 4, 3)

template <int B, int MODE>
static bool plan_dispatch_b(const fft_pass_args &a, hipStream_t stream) {
  switch (a.len) {
#define INQAMD_PLAN_CASE(L, R0, R1, R2) \
  case L: plan_launch<B, MODE, R0, R1, R2>(a, stream); return true;
    INQAMD_PLAN_TABLE(INQAMD_PLAN_CASE)
#undef INQAMD_PLAN_CASE
    default: return false;
  }
}

template <int MODE>
static bool plan_dispatch(const fft_pass_args &a, int B, hipStream_t stream) {
  switch (B) {
    case 16: return plan_dispatch_b<16, MODE>(a, stream);
    case 8: return plan_dispatch_b<8, MODE>(a, stream);
    default: return false;
  }
}

bool fft_plan_launch(const fft_pass_args &a, int mode, int B, hipStream_t stream) {
  switch (mode) {
    case FFT_PLAIN: return plan_dispatch<FFT_PLAIN>(a, B, stream);
    case FFT_KINETIC: return plan_dispatch<FFT_KINETIC>(a, B, stream);
    case FFT_EPILOGUE: return plan_dispatch<FFT_EPILOGUE>(a, B, stream);
    default: return false;
  }
}

}  // namespace inqamd
