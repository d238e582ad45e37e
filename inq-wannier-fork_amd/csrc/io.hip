// Orbital files in INQ's checkpoint format (operations::io::save/load for an
// orbital_set, src/operations/io.hpp:163-197): a directory with one file per
// (band, spinor component), named "%010d_%010d", each holding that band's N
// complex doubles in grid order ([x][y][z], the basis part of one process)
// as raw native-endian binary.  The device layout [x][y][z][band] is
// transposed band-major on the GPU (through the handle's work array), then
// moved one band at a time.

#include <sys/stat.h>

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "internal.hpp"
#include "inq_amd.h"

namespace inqamd {
namespace {

constexpr int TT = 32;  // transpose tile: 32 points x 32 bands

// out[b * N + p] = in[p * ld + b] for b < nst (to_band_major) or the inverse
template <bool TO_BAND_MAJOR>
__global__ __launch_bounds__(256) void transpose_kernel(const cplx *__restrict__ in,
                                                        cplx *__restrict__ out, long npoint,
                                                        long ld, int nst) {
  __shared__ cplx tile[TT][TT + 1];
  const long p0 = (long)blockIdx.x * TT;
  const int b0 = blockIdx.y * TT;
  const int tx = threadIdx.x % TT, ty = threadIdx.x / TT;  // 32 x 8
  for (int r = ty; r < TT; r += 8) {
    if (TO_BAND_MAJOR) {
      // read point p0 + r, band b0 + tx (band contiguous)
      const long p = p0 + r;
      const int b = b0 + tx;
      if (p < npoint && b < nst) tile[r][tx] = in[p * ld + b];
    } else {
      // read band b0 + r, point p0 + tx (point contiguous)
      const long p = p0 + tx;
      const int b = b0 + r;
      if (p < npoint && b < nst) tile[tx][r] = in[(long)b * npoint + p];
    }
  }
  __syncthreads();
  for (int r = ty; r < TT; r += 8) {
    if (TO_BAND_MAJOR) {
      const long p = p0 + tx;
      const int b = b0 + r;
      if (p < npoint && b < nst) out[(long)b * npoint + p] = tile[tx][r];
    } else {
      const long p = p0 + r;
      const int b = b0 + tx;
      if (p < npoint && b < nst) out[p * ld + b] = tile[r][tx];
    }
  }
}

dim3 tgrid(long npoint, int nst) {
  return dim3((unsigned)((npoint + TT - 1) / TT), (unsigned)((nst + TT - 1) / TT));
}

// utils::num_to_str (src/utils/num_str.hpp:20-24)
std::string num10(long v) {
  char buf[32];
  std::snprintf(buf, sizeof(buf), "%010ld", v);
  return buf;
}

void make_dirs(const std::string &dir) {
  std::string cur;
  for (size_t i = 0; i <= dir.size(); i++) {
    if (i == dir.size() || dir[i] == '/') {
      if (!cur.empty() && ::mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST)
        throw error{ERR_ARGUMENT, "cannot create directory '" + cur + "': " + std::strerror(errno)};
    }
    if (i < dir.size()) cur.push_back(dir[i]);
  }
}

std::string band_file(const std::string &dir, long band, int spinor) {
  return dir + "/" + num10(band) + "_" + num10(spinor);
}

}  // namespace

void save_states(handle &h, const cplx *psi, long ld, int nst, long first_band, int spinor,
                 const std::string &dir) {
  make_dirs(dir);
  hipLaunchKernelGGL(transpose_kernel<true>, tgrid(h.npoint, nst), dim3(256), 0, h.stream, psi,
                     h.hn, h.npoint, ld, nst);
  INQAMD_HIP_CHECK(hipGetLastError());
  std::vector<cplx> buf((size_t)h.npoint);
  for (int b = 0; b < nst; b++) {
    INQAMD_HIP_CHECK(hipMemcpyAsync(buf.data(), h.hn + (size_t)b * h.npoint,
                                    (size_t)h.npoint * sizeof(cplx), hipMemcpyDeviceToHost,
                                    h.stream));
    INQAMD_HIP_CHECK(hipStreamSynchronize(h.stream));
    const std::string name = band_file(dir, first_band + b, spinor);
    FILE *f = std::fopen(name.c_str(), "wb");
    if (!f) throw error{ERR_ARGUMENT, "cannot create restart file '" + name + "'"};
    const size_t n = std::fwrite(buf.data(), sizeof(cplx), (size_t)h.npoint, f);
    const bool ok = n == (size_t)h.npoint && std::fclose(f) == 0;
    if (!ok) throw error{ERR_ARGUMENT, "cannot write restart file '" + name + "'"};
  }
}

void load_states(handle &h, cplx *psi, long ld, int nst, long first_band, int spinor,
                 const std::string &dir) {
  std::vector<cplx> buf((size_t)h.npoint);
  for (int b = 0; b < nst; b++) {
    const std::string name = band_file(dir, first_band + b, spinor);
    FILE *f = std::fopen(name.c_str(), "rb");
    if (!f) throw error{ERR_ARGUMENT, "cannot open restart file '" + name + "'"};
    const size_t n = std::fread(buf.data(), sizeof(cplx), (size_t)h.npoint, f);
    std::fclose(f);
    if (n != (size_t)h.npoint)
      throw error{ERR_ARGUMENT, "restart file '" + name + "' holds fewer than " +
                                    std::to_string(h.npoint) + " points"};
    INQAMD_HIP_CHECK(hipMemcpyAsync(h.hn + (size_t)b * h.npoint, buf.data(),
                                    (size_t)h.npoint * sizeof(cplx), hipMemcpyHostToDevice,
                                    h.stream));
    INQAMD_HIP_CHECK(hipStreamSynchronize(h.stream));
  }
  hipLaunchKernelGGL(transpose_kernel<false>, tgrid(h.npoint, nst), dim3(256), 0, h.stream,
                     (const cplx *)h.hn, psi, h.npoint, ld, nst);
  INQAMD_HIP_CHECK(hipGetLastError());
  INQAMD_HIP_CHECK(hipStreamSynchronize(h.stream));
}

}  // namespace inqamd

using namespace inqamd;

extern "C" {

int inqamd_save_states(inqamd_hamiltonian_t h, const void *psi, int nst, int ld, long first_band,
                       int spinor, const char *dirname) {
  if (!h) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    check_orbital(*h, psi, "psi");
    check_bands(*h, nst, ld, true);
    INQAMD_REQUIRE(dirname && *dirname, "dirname is empty");
    INQAMD_REQUIRE(first_band >= 0 && spinor >= 0, "band and spinor indices must be >= 0");
    save_states(*h, (const cplx *)psi, ld, nst, first_band, spinor, dirname);
  });
}

int inqamd_load_states(inqamd_hamiltonian_t h, void *psi, int nst, int ld, long first_band,
                       int spinor, const char *dirname) {
  if (!h) return INQAMD_ERR_ARGUMENT;
  return guarded(h, [&] {
    check_orbital(*h, psi, "psi");
    check_bands(*h, nst, ld, true);
    INQAMD_REQUIRE(dirname && *dirname, "dirname is empty");
    INQAMD_REQUIRE(first_band >= 0 && spinor >= 0, "band and spinor indices must be >= 0");
    load_states(*h, (cplx *)psi, ld, nst, first_band, spinor, dirname);
  });
}

}  // extern "C"
