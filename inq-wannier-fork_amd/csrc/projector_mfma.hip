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

void mfma_selftest(double *out_device, hipStream_t stream) {
  hipLaunchKernelGGL(mfma_selftest_kernel, dim3(1), dim3(64), 0, stream, out_device);
  INQAMD_HIP_CHECK(hipGetLastError());
}

}  // namespace inqamd
