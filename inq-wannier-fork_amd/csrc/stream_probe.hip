// Streaming-bandwidth probe: what plain 16-byte vector loads and stores reach
// on this GPU, contiguously and in the access patterns of the Hpsi sweeps
// (a band chunk of INQ's [x][y][z][band] rows), as the ceiling the sweep
// kernels are measured against.  Diagnostic only; no reference counterpart.

#include "internal.hpp"
#include "inq_amd.h"

namespace inqamd {

// Element e of the probed set (e < rows * nb) lives at
//   row * ld + col(e % nb),  col(c) = off + c                 (contiguous window)
//                            col(c) = (c / 8) * stride8 + c % 8 (8-band groups,
//                                     stride8 = 8 * nchunks: chunk 0's groups)
struct probe_args {
  const cplx *src;
  cplx *dst;
  long rows;
  int ld, nb, grouped, stride8;
  long total;
};

__device__ __forceinline__ long probe_index(const probe_args &a, long e) {
  if (a.ld == a.nb) return e;
  const long row = e / a.nb;
  const int c = (int)(e - row * a.nb);
  const int col = a.grouped ? (c >> 3) * a.stride8 + (c & 7) : c;
  return row * a.ld + col;
}

constexpr int PROBE_NT = 256;
constexpr int PROBE_U = 4;

// kind 0 read, 1 write, 2 copy.  Each thread handles PROBE_U elements a
// pass, PROBE_NT apart, with all loads issued before any use.
template <int KIND>
__global__ __launch_bounds__(PROBE_NT) void probe_kernel(const probe_args a) {
  const long step = (long)gridDim.x * PROBE_NT * PROBE_U;
  double2 acc = make_double2(0.0, 0.0);
  for (long base = (long)blockIdx.x * PROBE_NT * PROBE_U + threadIdx.x; base < a.total;
       base += step) {
    long idx[PROBE_U];
    bool ok[PROBE_U];
#pragma unroll
    for (int u = 0; u < PROBE_U; u++) {
      const long e = base + (long)u * PROBE_NT;
      ok[u] = e < a.total;
      idx[u] = ok[u] ? probe_index(a, e) : 0;
    }
    if (KIND == 1) {
#pragma unroll
      for (int u = 0; u < PROBE_U; u++)
        if (ok[u]) a.dst[idx[u]] = make_double2((double)idx[u], 1.0);
    } else {
      double2 v[PROBE_U];
#pragma unroll
      for (int u = 0; u < PROBE_U; u++) v[u] = ok[u] ? a.src[idx[u]] : make_double2(0.0, 0.0);
      if (KIND == 2) {
#pragma unroll
        for (int u = 0; u < PROBE_U; u++)
          if (ok[u]) a.dst[idx[u]] = v[u];
      } else {
#pragma unroll
        for (int u = 0; u < PROBE_U; u++) {
          acc.x += v[u].x;
          acc.y += v[u].y;
        }
      }
    }
  }
  // keeps the reads alive; the sum of the probe's inputs is never this value
  if (KIND == 0 && acc.x == -1.2345e300) a.dst[blockIdx.x * PROBE_NT + threadIdx.x] = acc;
}

}  // namespace inqamd

using namespace inqamd;

extern "C" int inqamd_stream_probe(int kind, const void *src, void *dst, long rows, int ld, int nb,
                                   int grouped, int nchunks, int blocks_per_cu, void *stream,
                                   float *ms) {
  try {
    INQAMD_REQUIRE(kind >= 0 && kind <= 2, "kind must be 0 (read), 1 (write) or 2 (copy)");
    INQAMD_REQUIRE(rows > 0 && nb > 0 && ld >= nb, "need rows > 0 and ld >= nb > 0");
    INQAMD_REQUIRE(blocks_per_cu >= 1 && blocks_per_cu <= 64, "blocks_per_cu must be in [1, 64]");
    INQAMD_REQUIRE(!grouped || (nb % 8 == 0 && nchunks >= 1 && (nb / 8 - 1) * 8 * nchunks + 8 <= ld),
                   "grouped windows must fit the row");
    INQAMD_REQUIRE(kind == 1 || src, "src is null");
    INQAMD_REQUIRE(kind == 0 || dst, "dst is null");
    INQAMD_REQUIRE(ms, "ms is null");
    int dev = 0, cus = 0;
    INQAMD_HIP_CHECK(hipGetDevice(&dev));
    INQAMD_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    probe_args a;
    a.src = (const cplx *)src;
    a.dst = (cplx *)dst;
    a.rows = rows;
    a.ld = ld;
    a.nb = nb;
    a.grouped = grouped;
    a.stride8 = 8 * (nchunks > 0 ? nchunks : 1);
    a.total = rows * nb;
    // the read kernel's keep-alive store may land in dst[0 .. grid * PROBE_NT)
    const long grid = (long)blocks_per_cu * cus;
    INQAMD_REQUIRE(kind != 0 || (dst && grid * PROBE_NT <= rows * (long)ld),
                   "the read probe needs dst of at least the probed extent");
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0, e1;
    INQAMD_HIP_CHECK(hipEventCreate(&e0));
    INQAMD_HIP_CHECK(hipEventCreate(&e1));
    INQAMD_HIP_CHECK(hipEventRecord(e0, s));
    if (kind == 0)
      hipLaunchKernelGGL(probe_kernel<0>, dim3((unsigned)grid), dim3(PROBE_NT), 0, s, a);
    else if (kind == 1)
      hipLaunchKernelGGL(probe_kernel<1>, dim3((unsigned)grid), dim3(PROBE_NT), 0, s, a);
    else
      hipLaunchKernelGGL(probe_kernel<2>, dim3((unsigned)grid), dim3(PROBE_NT), 0, s, a);
    INQAMD_HIP_CHECK(hipGetLastError());
    INQAMD_HIP_CHECK(hipEventRecord(e1, s));
    INQAMD_HIP_CHECK(hipEventSynchronize(e1));
    INQAMD_HIP_CHECK(hipEventElapsedTime(ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  } catch (const error &e) {
    return e.code;
  }
  return INQAMD_OK;
}
