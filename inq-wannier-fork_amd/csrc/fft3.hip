// Three-sweep local part of Hpsi (T + V) for one system on a cubic grid
// L x L x L: one launch per sweep over a band chunk (tile bodies in
// fft3_tiles.hpp).  The projector's V_nl psi (precomputed per owned point)
// and the Taylor accumulations acc += c * hpsi are added in the last sweep's
// write.
//
// Traffic per band: 2 x 16N (z_fwd) + 2 x 16N (xy_kin) + 3 x 16N (z_bwd_epi)
// = 112 N bytes, against 176 N for five pencil passes; run per band chunk
// so W stays resident in the Infinity Cache between sweeps.

#include "fft3_tiles.hpp"

#include <cstdlib>

namespace inqamd {

// Shapes of the one-system kernels: z tiles of TY y x 64/TY bands, x-y
// planes with plane_threads(L) or twice as many threads.  The defaults (8,
// 2x: measured fastest at 64^3 x 256) can be overridden through INQAMD_Z_TY
// (8 or 16) and INQAMD_XY_WIDE (0 or 1).

template <int R0, int R1, int R2, int TY>
__global__ __launch_bounds__(z_threads(R0 * R1 * R2)) void z_fwd_kernel(const fft3_args a) {
  extern __shared__ cplx lds[];
  z_fwd_tile<R0, R1, R2, TY, z_threads(R0 * R1 * R2)>(a, blockIdx.x, lds);
}

constexpr int wide_threads(int L) {
  return 2 * plane_threads(L) > 1024 ? 1024 : 2 * plane_threads(L);
}

template <int R0, int R1, int R2, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(xy_kin_waves<R0, R1, R2, NT>())))
void xy_kin_kernel(const fft3_args a) {
  extern __shared__ cplx lds[];
  xy_kin_plane<R0, R1, R2, NT>(a, blockIdx.x, lds);
}

// TY: y values of a 64-column tile (8 or 16); a 32-column tile takes TY / 2
// (the same bands per tile)
constexpr int z_bwd_ty(int L, bool acc, int ty) { return z_bwd_cols(L, acc) == Z_COLS ? ty : ty / 2; }

template <int R0, int R1, int R2, int TY, bool ACC, bool FWD = false>
__global__ __launch_bounds__(z_bwd_threads(R0 * R1 * R2, ACC))
__attribute__((amdgpu_waves_per_eu(z_bwd_waves(R0 * R1 * R2))))
void z_bwd_epi_kernel(const fft3_args a) {
  extern __shared__ cplx lds[];
  constexpr int L = R0 * R1 * R2;
  // the plain Hpsi's tiles dealt per XCD (64^3 x 256 Hpsi 1.495 -> 1.480
  // ms); the Taylor forms keep the natural order (their steps ran 1 % slower
  // dealt per XCD)
  constexpr int TB = z_bwd_cols(L, ACC) / z_bwd_ty(L, ACC, TY);
  const long t = ACC ? (long)blockIdx.x : xcd_tile(blockIdx.x, gridDim.x, (a.nb + TB - 1) / TB);
  z_bwd_epi_tile<R0, R1, R2, z_bwd_ty(L, ACC, TY), z_bwd_threads(L, ACC), ACC, false, FWD,
                 z_bwd_cols(L, ACC)>(a, t, lds);
}

// Launch shape of a z_bwd_epi variant (acc: the Taylor forms)
template <int R0, int R1, int R2, int TY, bool ACC>
struct z_bwd_shape {
  static constexpr int L = R0 * R1 * R2, ZC = z_bwd_cols(L, ACC);
  static long tiles(int nb) { return z_tiles<z_bwd_ty(L, ACC, TY), ZC>(L, nb); }
  static constexpr size_t lds = z_bwd_lds(L, ZC);
  static constexpr int threads = z_bwd_threads(L, ACC);
};

// G-space Hpsi sweeps
template <int R0, int R1, int R2, int TY>
__global__ __launch_bounds__(z_threads(R0 * R1 * R2)) void z_pot_kernel(const fft3_args a) {
  extern __shared__ cplx lds[];
  // tiles dealt per XCD: one (x, y rows) pair's band tiles share V and the
  // ownership words through one L2
  constexpr int TB = Z_COLS / TY;
  z_fwd_tile<R0, R1, R2, TY, z_threads(R0 * R1 * R2), true>(
      a, xcd_tile(blockIdx.x, gridDim.x, (a.nb + TB - 1) / TB), lds);
}

template <int R0, int R1, int R2, int NT>
__global__ __launch_bounds__(NT) void xy_fwd_kernel(const fft3_args a) {
  extern __shared__ cplx lds[];
  xy_fwd_plane<R0, R1, R2, NT>(a, blockIdx.x, lds);
}

template <int R0, int R1, int R2, int TY>
__global__ __launch_bounds__(z_threads(R0 * R1 * R2))
__attribute__((amdgpu_waves_per_eu(z_bwd_waves(R0 * R1 * R2))))
void z_kin_kernel(const fft3_args a) {
  extern __shared__ cplx lds[];
  z_bwd_epi_tile<R0, R1, R2, TY, z_threads(R0 * R1 * R2), false, true>(a, blockIdx.x, lds);
}

static int z_ty_choice() {
  static const int ty = [] {
    const char *e = getenv("INQAMD_Z_TY");
    return e && atoi(e) == 16 ? 16 : 8;
  }();
  return ty;
}

static bool xy_wide_choice() {
  static const bool w = [] {
    const char *e = getenv("INQAMD_XY_WIDE");
    return !(e && atoi(e) == 0);
  }();
  return w;
}

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------
template <typename K>
static void ensure_lds(K kernel, size_t bytes) {
  if (bytes > 65536)
    INQAMD_HIP_CHECK(hipFuncSetAttribute((const void *)kernel,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

template <int R0, int R1, int R2>
static void prepare3() {
  constexpr int L = R0 * R1 * R2;
  const size_t zl = (size_t)L * Z_COLS * sizeof(cplx), pl = (size_t)L * (L + 1) * sizeof(cplx);
  ensure_lds(z_fwd_kernel<R0, R1, R2, 8>, zl);
  ensure_lds(z_fwd_kernel<R0, R1, R2, 16>, zl);
  ensure_lds(xy_kin_kernel<R0, R1, R2, plane_threads(L)>, xy_kin_lds<R0, R1, R2>());
  ensure_lds(xy_kin_kernel<R0, R1, R2, wide_threads(L)>, xy_kin_lds<R0, R1, R2>());
  ensure_lds(z_pot_kernel<R0, R1, R2, 8>, zl);
  ensure_lds(z_kin_kernel<R0, R1, R2, 8>, zl);
  ensure_lds(xy_fwd_kernel<R0, R1, R2, wide_threads(L)>, pl);
  const size_t zbp = z_bwd_lds(L, z_bwd_cols(L, false)), zba = z_bwd_lds(L, z_bwd_cols(L, true));
  ensure_lds(z_bwd_epi_kernel<R0, R1, R2, 8, false>, zbp);
  ensure_lds(z_bwd_epi_kernel<R0, R1, R2, 16, false>, zbp);
  ensure_lds(z_bwd_epi_kernel<R0, R1, R2, 8, true>, zba);
  ensure_lds(z_bwd_epi_kernel<R0, R1, R2, 16, true>, zba);
  ensure_lds(z_bwd_epi_kernel<R0, R1, R2, 8, true, true>, zba);
  ensure_lds(z_bwd_epi_kernel<R0, R1, R2, 16, true, true>, zba);
}

// sweep: -1 all three, 0/1/2 one of them; fwd_next: the last sweep also
// transforms its output forward in z into W (Taylor chains)
template <int R0, int R1, int R2, int TY, int NT>
static void launch3v(const fft3_args &a, hipStream_t stream, int sweep, bool fwd_next) {
  constexpr int L = R0 * R1 * R2;
  const long ztiles = z_tiles<TY>(L, a.nb);
  const size_t zlds = (size_t)L * Z_COLS * sizeof(cplx);
  const size_t plds = xy_kin_lds<R0, R1, R2>();
  if (sweep < 0 || sweep == 0) {
    hipLaunchKernelGGL((z_fwd_kernel<R0, R1, R2, TY>), dim3((unsigned)ztiles), dim3(z_threads(L)), zlds,
                       stream, a);
    INQAMD_HIP_CHECK(hipGetLastError());
  }
  if (sweep < 0 || sweep == 1) {
    hipLaunchKernelGGL((xy_kin_kernel<R0, R1, R2, NT>), dim3((unsigned)((long)a.nb * L)), dim3(NT),
                       plds, stream, a);
    INQAMD_HIP_CHECK(hipGetLastError());
  }
  if (sweep < 0 || sweep == 2) {
    using SA = z_bwd_shape<R0, R1, R2, TY, true>;
    using SP = z_bwd_shape<R0, R1, R2, TY, false>;
    if (fwd_next)
      hipLaunchKernelGGL((z_bwd_epi_kernel<R0, R1, R2, TY, true, true>), dim3((unsigned)SA::tiles(a.nb)),
                         dim3(SA::threads), SA::lds, stream, a);
    else if (a.acc1 || a.acc2)
      hipLaunchKernelGGL((z_bwd_epi_kernel<R0, R1, R2, TY, true>), dim3((unsigned)SA::tiles(a.nb)),
                         dim3(SA::threads), SA::lds, stream, a);
    else
      hipLaunchKernelGGL((z_bwd_epi_kernel<R0, R1, R2, TY, false>), dim3((unsigned)SP::tiles(a.nb)),
                         dim3(SP::threads), SP::lds, stream, a);
    INQAMD_HIP_CHECK(hipGetLastError());
  }
}

template <int R0, int R1, int R2>
static void launch3(const fft3_args &a, hipStream_t stream, int sweep, bool fwd_next) {
  constexpr int L = R0 * R1 * R2;
  const bool wide = a.xy_wide ? a.xy_wide == 2 : xy_wide_choice();
  const int ty = a.z_ty ? a.z_ty : z_ty_choice();
  if (ty == 16) {
    if (wide)
      launch3v<R0, R1, R2, 16, wide_threads(L)>(a, stream, sweep, fwd_next);
    else
      launch3v<R0, R1, R2, 16, plane_threads(L)>(a, stream, sweep, fwd_next);
  } else {
    if (wide)
      launch3v<R0, R1, R2, 8, wide_threads(L)>(a, stream, sweep, fwd_next);
    else
      launch3v<R0, R1, R2, 8, plane_threads(L)>(a, stream, sweep, fwd_next);
  }
}

template <int R0, int R1, int R2>
static void launch_fourier(const fft3_args &a, hipStream_t stream) {
  constexpr int L = R0 * R1 * R2;
  const long ztiles = z_tiles<8>(L, a.nb);
  const size_t zlds = (size_t)L * Z_COLS * sizeof(cplx);
  const size_t plds = (size_t)L * (L + 1) * sizeof(cplx);
  hipLaunchKernelGGL((z_pot_kernel<R0, R1, R2, 8>), dim3((unsigned)ztiles), dim3(z_threads(L)), zlds,
                     stream, a);
  INQAMD_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL((xy_fwd_kernel<R0, R1, R2, wide_threads(L)>), dim3((unsigned)((long)a.nb * L)),
                     dim3(wide_threads(L)), plds, stream, a);
  INQAMD_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL((z_kin_kernel<R0, R1, R2, 8>), dim3((unsigned)ztiles), dim3(z_threads(L)), zlds,
                     stream, a);
  INQAMD_HIP_CHECK(hipGetLastError());
}

void fft3_fourier_hamiltonian(const fft3_args &a, int L, hipStream_t stream) {
  switch (L) {
#define INQAMD_FFT3_CASE(LL, R0, R1, R2) \
  case LL: launch_fourier<R0, R1, R2>(a, stream); return;
    INQAMD_FFT3_TABLE(INQAMD_FFT3_CASE)
#undef INQAMD_FFT3_CASE
    default: throw error{ERR_UNSUPPORTED, "G-space three-sweep Hpsi: unsupported grid length"};
  }
}

bool fft3_supported(int L, size_t max_lds) {
  switch (L) {
#define INQAMD_FFT3_CASE(LL, R0, R1, R2) case LL:
    INQAMD_FFT3_TABLE(INQAMD_FFT3_CASE)
#undef INQAMD_FFT3_CASE
      return (size_t)L * (L + 1) * sizeof(cplx) <= max_lds && z_bwd_lds(L) <= max_lds;
    default: return false;
  }
}

void fft3_prepare(int L) {
  switch (L) {
#define INQAMD_FFT3_CASE(LL, R0, R1, R2) \
  case LL: prepare3<R0, R1, R2>(); return;
    INQAMD_FFT3_TABLE(INQAMD_FFT3_CASE)
#undef INQAMD_FFT3_CASE
    default: return;
  }
}

void fft3_local_hamiltonian(const fft3_args &a, int L, hipStream_t stream, int sweep,
                            bool fwd_next) {
  switch (L) {
#define INQAMD_FFT3_CASE(LL, R0, R1, R2) \
  case LL: launch3<R0, R1, R2>(a, stream, sweep, fwd_next); return;
    INQAMD_FFT3_TABLE(INQAMD_FFT3_CASE)
#undef INQAMD_FFT3_CASE
    default: throw error{ERR_UNSUPPORTED, "three-sweep Hpsi: unsupported grid length"};
  }
}

}  // namespace inqamd
