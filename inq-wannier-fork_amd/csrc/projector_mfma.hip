// Nonlocal projector on the FP64 matrix cores: one launch per Hpsi over
// (band tile, sphere); the block body is project_expand_block.

#include "projector_dev.hpp"

namespace inqamd {

__global__ __launch_bounds__(PE_NT) void project_expand_kernel(const project_expand_args a) {
  project_expand_block(a, blockIdx.y, blockIdx.x * PE_BANDS);
}

// grid (band tiles, spheres, point chunks); tile (p, t, c) at part[((p T + t) nc + c) 256]
__global__ __launch_bounds__(PE_NT) void project_chunk_kernel(const project_expand_args a, int nc,
                                                              double *part) {
  const long tile = (long)blockIdx.y * gridDim.x + blockIdx.x;
  project_chunk_block(a, blockIdx.y, blockIdx.x * PE_BANDS, blockIdx.z, nc,
                      part + (tile * nc + blockIdx.z) * 256);
}

__global__ __launch_bounds__(PE_NT) void expand_chunk_kernel(const project_expand_args a, int nc,
                                                             const double *part) {
  const long tile = (long)blockIdx.y * gridDim.x + blockIdx.x;
  project_expand_args b = a;
  if (blockIdx.z != 0) b.P = nullptr;  // one writer of the coefficients
  expand_chunk_block(b, blockIdx.y, blockIdx.x * PE_BANDS, blockIdx.z, nc, part + tile * nc * 256);
}

// P coefficients only (the Hpsi sweep expands V_nl psi itself): grid
// (tiles of 16 NB bands, spheres)
template <int NB>
__global__ __launch_bounds__(P4_NT) void project_p4_kernel(const project_expand_args a) {
  __shared__ double red[P4_WAVES][NB][64][2];
  __shared__ double sum[NB * 128];
  const int p = blockIdx.y, b0 = blockIdx.x * P4_BANDS * NB;
  gemm1_p4_points<NB>(a, p, b0, 0, a.size[p], red, sum);
  __syncthreads();
  p4_store<NB>(a, p, b0, sum);
}

// Point-chunked form for few (sphere, tile) pairs: grid (16-band tiles,
// spheres, chunks); chunk c of tile t of sphere p at part[((p T + t) nc + c) 128]
__global__ __launch_bounds__(P4_NT) void project_p4_chunk_kernel(const project_expand_args a,
                                                                 int nc, double *part) {
  __shared__ double red[P4_WAVES][1][64][2];
  const int p = blockIdx.y, c = blockIdx.z;
  const long tile = (long)p * gridDim.x + blockIdx.x;
  const int K = a.size[p];
  gemm1_p4_points<1>(a, p, blockIdx.x * P4_BANDS, chunk_begin(K, c, nc),
                     chunk_begin(K, c + 1, nc), red, part + (tile * nc + c) * 128);
}

// Sums a tile's chunks in chunk order and stores P: grid (tiles, spheres)
__global__ __launch_bounds__(128) void project_p4_finish_kernel(const project_expand_args a,
                                                                int nc, const double *part) {
  __shared__ double sum[128];
  const int p = blockIdx.y;
  const long tile = (long)p * gridDim.x + blockIdx.x;
  const double *q = part + tile * nc * 128;
  double s = q[threadIdx.x];
  for (int c = 1; c < nc; c++) s += q[c * 128 + threadIdx.x];
  sum[threadIdx.x] = s;
  __syncthreads();
  p4_store<1>(a, p, blockIdx.x * P4_BANDS, sum);
}

// Self-test of the lane layout of v_mfma_f64_4x4x4f64: block s of D = A B with
// A_s[i][k] = 1 + i + 4 k + 16 s and B_s[k][j] = 2 + 3 j - 5 k + 7 s, written as
// out[s][i][j] (64 doubles).
__global__ void mfma4_selftest_kernel(double *out) {
  const int l = threadIdx.x;
  const int s = (l >> 2) & 3;
  const double av = 1.0 + (l & 3) + 4.0 * (l >> 4) + 16.0 * s;
  const double bv = 2.0 + 3.0 * (l & 3) - 5.0 * (l >> 4) + 7.0 * s;
  const double d = mfma4(av, bv, 0.0);
  out[s * 16 + (l >> 4) * 4 + (l & 3)] = d;
}

// Self-test of the lane layout of projector_dev.hpp: D = A B with A[i][k] = 1 + i + 16 k and
// B[k][j] = 2 + 3 j - 5 k, written row-major into out[16 * 16].
__global__ void mfma_selftest_kernel(double *out) {
  const int l = threadIdx.x;
  const double av = 1.0 + mfma_row_a(l) + 16.0 * mfma_k(l);
  const double bv = 2.0 + 3.0 * mfma_col_b(l) - 5.0 * mfma_k(l);
  v4d d = {0.0, 0.0, 0.0, 0.0};
  d = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, d, 0, 0, 0);
#pragma unroll
  for (int v = 0; v < 4; v++) out[mfma_row_d(l, v) * 16 + mfma_col_d(l)] = d[v];
}

void projector_project_expand(handle &h, const cplx *psi, long ld, int nst, cplx *Qe, long ldQ,
                              bool write_P) {
  const projector_tables &t = h.pt;
  if (t.nproj == 0) return;
  project_expand_args a;
  a.psi = psi;
  a.ld = ld;
  a.nst = nst;
  a.max_sphere = t.max_sphere;
  a.max_nlm = t.max_nlm;
  a.point = t.point;
  a.size = t.size;
  a.mat = t.mat;
  a.coeff = t.coeff;
  a.phase = t.phase_sph;
  a.ent = t.ent;
  a.dV = h.volume / (double)h.npoint;
  a.P = write_P ? h.proj : nullptr;
  a.Qe = Qe;
  a.ldQ = ldQ;
  const int nbt = (nst + PE_BANDS - 1) / PE_BANDS;
  const long tiles = (long)nbt * t.nproj;
  // few (sphere, band tile) pairs: split the sphere points over workgroups
  int nc = 1;
  if (tiles < 512 && h.proj_part) {
    nc = (int)((1024 + tiles - 1) / tiles);
    const int by_points = t.max_sphere / 256 > 1 ? t.max_sphere / 256 : 1;
    if (nc > by_points) nc = by_points;
    if (nc > PROJ_MAX_CHUNKS) nc = PROJ_MAX_CHUNKS;
    if (tiles * nc > PROJ_PART_TILES) nc = (int)(PROJ_PART_TILES / tiles);
    if (nc < 1) nc = 1;
  }
  if (nc == 1) {
    dim3 grid(nbt, t.nproj);
    hipLaunchKernelGGL(project_expand_kernel, grid, dim3(PE_NT), 0, h.stream, a);
    INQAMD_HIP_CHECK(hipGetLastError());
    return;
  }
  dim3 grid(nbt, t.nproj, nc);
  hipLaunchKernelGGL(project_chunk_kernel, grid, dim3(PE_NT), 0, h.stream, a, nc, h.proj_part);
  INQAMD_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(expand_chunk_kernel, grid, dim3(PE_NT), 0, h.stream, a, nc,
                     (const double *)h.proj_part);
  INQAMD_HIP_CHECK(hipGetLastError());
}

// P = dV D M^T (phase psi) for all spheres and bands on the 4x4x4 FP64 MFMA
// (the three-sweep Hpsi with V_nl psi expanded in its last sweep)
static void project_p4_launch(handle &h, const project_expand_args &a, int nst, int nst_all,
                              hipStream_t st, double *part, long part_doubles);

void projector_project_p4(handle &h, const cplx *psi, long ld, int nst) {
  projector_project_p4_chunk(h, psi, ld, nst, 0, nst, 0, 0, h.stream, 0);
}

void projector_project_p4_chunk(handle &h, const cplx *psi, long ld, int nst, int b0, int nb,
                                int gskip, int glog, hipStream_t stream, int part_slot) {
  const projector_tables &t = h.pt;
  if (t.nproj == 0) return;
  project_expand_args a{};
  a.psi = psi;
  a.ld = ld;
  a.nst = nst;
  a.max_sphere = t.max_sphere;
  a.max_nlm = t.max_nlm;
  a.point = t.point;
  a.size = t.size;
  a.mat = t.mat;
  a.coeff = t.coeff;
  a.phase = t.phase_sph;
  a.ent = t.ent;
  a.dV = h.volume / (double)h.npoint;
  a.P = h.proj;
  a.nbc = nb;
  a.gskip = gskip;
  a.glog = glog;
  a.pband0 = b0;
  INQAMD_REQUIRE(t.max_nlm <= 4, "the 4x4x4 projector needs max_nlm <= 4");
  INQAMD_REQUIRE(part_slot >= 0 && part_slot < PROJ_PART_SLOTS, "projector partial slot");
  const long slot_doubles = PROJ_PART_TILES * 256 / PROJ_PART_SLOTS;
  project_p4_launch(h, a, nb, nst, stream,
                    h.proj_part ? h.proj_part + part_slot * slot_doubles : nullptr, slot_doubles);
}

// Launches of the 4x4x4 coefficient kernels over the nst (chunk) bands of `a`;
// the point-chunked form keeps its partial tiles in part[part_doubles].  The
// point chunking follows from all nst_all bands, so a band's coefficients are
// the same bits whichever chunk computes them.
static void project_p4_launch(handle &h, const project_expand_args &a, int nst, int nst_all,
                              hipStream_t st, double *part, long part_doubles) {
  const projector_tables &t = h.pt;
  // bands per workgroup: 32 (INQAMD_P4_BANDS = 16, 32 or 64; at 64^3 x 256
  // the projector took 62, 61 and 60 us) where that leaves 256 workgroups
  static const int nb16_env = [] {
    const char *e = getenv("INQAMD_P4_BANDS");
    const int b = e ? atoi(e) : 32;
    return b >= 64 ? 4 : (b >= 32 ? 2 : 1);
  }();
  const int nb16 = h.tune_p4 ? h.tune_p4 / 16 : nb16_env;
  const int nbt = (nst + P4_BANDS - 1) / P4_BANDS;
  const long tiles = (long)((nst_all + P4_BANDS - 1) / P4_BANDS) * t.nproj;
  int nc = 1;
  if (tiles < 512 && part) {
    nc = (int)((1024 + tiles - 1) / tiles);
    const int by_points = t.max_sphere / 256 > 1 ? t.max_sphere / 256 : 1;
    if (nc > by_points) nc = by_points;
    if (nc > PROJ_MAX_CHUNKS) nc = PROJ_MAX_CHUNKS;
    if (tiles * nc * 128 > part_doubles) nc = (int)(part_doubles / 128 / tiles);
    if (nc < 1) nc = 1;
  }
  if (nc == 1) {
    const long t32 = (long)((nst + 31) / 32) * t.nproj, t64 = (long)((nst + 63) / 64) * t.nproj;
    if (nb16 == 4 && t64 >= 256)
      hipLaunchKernelGGL(project_p4_kernel<4>, dim3((nst + 63) / 64, t.nproj), dim3(P4_NT), 0,
                         st, a);
    else if (nb16 >= 2 && t32 >= 256)
      hipLaunchKernelGGL(project_p4_kernel<2>, dim3((nst + 31) / 32, t.nproj), dim3(P4_NT), 0,
                         st, a);
    else
      hipLaunchKernelGGL(project_p4_kernel<1>, dim3(nbt, t.nproj), dim3(P4_NT), 0, st, a);
    INQAMD_HIP_CHECK(hipGetLastError());
    return;
  }
  hipLaunchKernelGGL(project_p4_chunk_kernel, dim3(nbt, t.nproj, nc), dim3(P4_NT), 0, st,
                     a, nc, part);
  INQAMD_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(project_p4_finish_kernel, dim3(nbt, t.nproj), dim3(128), 0, st, a, nc,
                     (const double *)part);
  INQAMD_HIP_CHECK(hipGetLastError());
}

void mfma4_selftest(double *out_device, hipStream_t stream) {
  hipLaunchKernelGGL(mfma4_selftest_kernel, dim3(1), dim3(64), 0, stream, out_device);
  INQAMD_HIP_CHECK(hipGetLastError());
}

void mfma_selftest(double *out_device, hipStream_t stream) {
  hipLaunchKernelGGL(mfma_selftest_kernel, dim3(1), dim3(64), 0, stream, out_device);
  INQAMD_HIP_CHECK(hipGetLastError());
}

}  // namespace inqamd
