// Uneven batches: many independent systems (different grids, band counts and
// projector sets) through one Hpsi call.
//
// Every phase of Hpsi runs as ONE persistent launch for the whole batch: a
// grid sized to the resident capacity of the device drains a work queue
// (a device counter) whose items are the per-system work units laid end to
// end -- projector (sphere, band tile) blocks of every system, then for each
// batch chunk the z tiles, x-y planes and z tiles of the three sweeps of
// every system piece in it.  A work item finds its system piece by binary
// search over the per-piece item offsets and dispatches the tile body of
// that piece's grid length and z-tile shape.
//
// Systems are cut into pieces of bands and the pieces are packed into batch
// chunks whose intermediate W stays in the Infinity Cache between the
// sweeps; small systems share a chunk, large ones span several.  The sweeps
// of a chunk are ordered by the launch boundaries, so no work item waits on
// another inside a launch.  Band remainders (nst mod 8) run with narrower z
// tiles (4 or 2 bands x 16 or 32 y) instead of padding to 8 bands.
//
// The arithmetic of every band is that of the one-system Hpsi (same tile
// bodies, same projector blocks); batched and per-system results agree to
// rounding (the compiler may contract multiply-adds differently in the
// batch kernels).

#include "fft3_tiles.hpp"
#include "projector_dev.hpp"
#include "inq_amd.h"

#include <algorithm>
#include <climits>
#include <string>
#include <vector>

namespace inqamd {

struct batch_phase_sys {
  const float *pos;
  const int *uent;
  int nsph, nent;
  const double *kA;
  cplx *phase_sph, *phase_ent;
  int item0;
};

struct batch_proj_sys {
  project_expand_args a;
  int nbt;  // band tiles of PE_BANDS (per-entry V_nl psi) or P4_BANDS (coefficients)
  int item0;
};

struct batch_piece {
  fft3_args a;
  int L;
  int ty;      // z-tile y extent: 8, 16 or 32
  int z0;      // first z-sweep item of the piece in its chunk
  int xy0;     // first x-y item of the piece in its chunk
};

// Grid lengths the batch kernels cover (larger grids run the one-system
// Hpsi), with the plans of INQAMD_FFT3_TABLE
#define INQAMD_BATCH_TABLE(X) \
  X(24, 8, 3, 1)              \
  X(30, 6, 5, 1)              \
  X(32, 8, 4, 1)              \
  X(36, 6, 6, 1)              \
  X(40, 8, 5, 1)              \
  X(48, 8, 6, 1)              \
  X(60, 12, 5, 1)             \
  X(64, 8, 8, 1)              \
  X(72, 8, 9, 1)

constexpr int XY_MAX_NT = 384;  // plane_threads(72)

// Dynamic LDS of an x-y plane of grid length L in these kernels
static size_t batch_xy_lds(int L) {
  switch (L) {
#define INQAMD_BATCH_XY_LDS(LL, R0, R1, R2) \
  case LL: return xy_kin_lds<R0, R1, R2>();
    INQAMD_BATCH_TABLE(INQAMD_BATCH_XY_LDS)
#undef INQAMD_BATCH_XY_LDS
    default: return (size_t)L * (L + 1) * sizeof(cplx);
  }
}

// Out-of-line tile bodies for the mixed-grid kernels: each (grid length,
// tile shape) variant is register-allocated on its own instead of as one
// merged switch.
template <int R0, int R1, int R2, int TY>
__device__ __attribute__((noinline)) void z_fwd_call(const fft3_args *a, long t, cplx *lds) {
  const fft3_args c = *a;
  z_fwd_tile<R0, R1, R2, TY>(c, t, lds);
}
template <int R0, int R1, int R2>
__device__ __attribute__((noinline)) void xy_call(const fft3_args *a, long t, cplx *lds) {
  const fft3_args c = *a;
  xy_kin_plane<R0, R1, R2>(c, t, lds);
}
template <int R0, int R1, int R2, int TY>
__device__ __attribute__((noinline)) void z_bwd_call(const fft3_args *a, long t, cplx *lds) {
  const fft3_args c = *a;
  z_bwd_epi_tile<R0, R1, R2, TY>(c, t, lds);
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Last index i in [0, n) with off(i) <= item (off(0) = 0).
template <class OFF>
__device__ __forceinline__ int find_segment(int n, int item, OFF off) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off(mid) <= item)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Drains the queue: every workgroup takes items until the counter passes
// `total`; the next item is claimed while the current one runs, and all
// waves of a workgroup leave together.
template <class F>
__device__ __forceinline__ void drain_queue(unsigned long long *counter, int total, F &&work) {
  __shared__ int s_next;
  if (threadIdx.x == 0) s_next = (int)atomicAdd(counter, 1ULL);
  __syncthreads();
  int cur = uniform(s_next);
  while (cur < total) {
    int next = 0;
    if (threadIdx.x == 0) next = (int)atomicAdd(counter, 1ULL);
    work(cur);
    __syncthreads();  // the item's LDS use and every read of s_next are done
    if (threadIdx.x == 0) s_next = next;
    __syncthreads();
    cur = uniform(s_next);
  }
}

// Items of equal cost: workgroup g takes items g, g + grid, g + 2 grid, ...
template <class F>
__device__ __forceinline__ void drain_strided(int total, F &&work) {
  for (int item = blockIdx.x; item < total; item += gridDim.x) {
    work(item);
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void batch_phase_kernel(const batch_phase_sys *__restrict__ sys, int nsys,
                                                          int total) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const int s = find_segment(nsys, g, [&](int i) { return sys[i].item0; });
  const batch_phase_sys &q = sys[s];
  phase_item(q.pos, q.uent, q.nsph, q.nent, q.kA, q.phase_sph, q.phase_ent, g - q.item0);
}

// Projector items (sphere, band tile): one workgroup per item.
__global__ __launch_bounds__(PE_NT) void batch_project_kernel(
    const batch_proj_sys *__restrict__ sys, int nsys, unsigned long long *, int) {
  const int item = blockIdx.x;
  const int s = find_segment(nsys, item, [&](int i) { return sys[i].item0; });
  const int local = item - sys[s].item0;
  const int nbt = sys[s].nbt;
  project_expand_block(sys[s].a, local / nbt, (local % nbt) * PE_BANDS);
}

// Projector coefficients only (the sweeps expand V_nl psi): items (sphere,
// 16-band tile) on the 4x4x4 FP64 MFMA, one workgroup per item.
__global__ __launch_bounds__(P4_NT) void batch_project_p4_kernel(
    const batch_proj_sys *__restrict__ sys, int nsys, unsigned long long *, int) {
  __shared__ double red[P4_WAVES][1][64][2];
  __shared__ double sum[128];
  const int item = blockIdx.x;
  const int s = find_segment(nsys, item, [&](int i) { return sys[i].item0; });
  const int local = item - sys[s].item0;
  const int nbt = sys[s].nbt;
  const project_expand_args &a = sys[s].a;
  const int p = local / nbt, b0 = (local % nbt) * P4_BANDS;
  gemm1_p4_points<1>(a, p, b0, 0, a.size[p], red, sum);
  __syncthreads();
  p4_store<1>(a, p, b0, sum);
}

__global__ __launch_bounds__(Z_NT) void batch_z_fwd_kernel(const batch_piece *__restrict__ pc, int np,
                                                           unsigned long long *counter,
                                                           int total) {
  extern __shared__ cplx lds[];
  drain_queue(counter, total, [&](int item) {
    const int i = uniform(find_segment(np, item, [&](int k) { return pc[k].z0; }));
    const batch_piece &p = pc[i];
    const long local = item - p.z0;
    const int key = p.L * 4 + (p.ty == 8 ? 0 : p.ty == 16 ? 1 : 2);
    switch (key) {
#define INQAMD_BZF(LL, R0, R1, R2)                                       \
  case LL * 4 + 0: z_fwd_call<R0, R1, R2, 8>(&p.a, local, lds); break;  \
  case LL * 4 + 1: z_fwd_call<R0, R1, R2, 16>(&p.a, local, lds); break; \
  case LL * 4 + 2: z_fwd_call<R0, R1, R2, 32>(&p.a, local, lds); break;
      INQAMD_BATCH_TABLE(INQAMD_BZF)
#undef INQAMD_BZF
      default: break;
    }
  });
}

__global__ __launch_bounds__(XY_MAX_NT) void batch_xy_kernel(const batch_piece *__restrict__ pc, int np,
                                                             unsigned long long *counter,
                                                             int total) {
  extern __shared__ cplx lds[];
  drain_queue(counter, total, [&](int item) {
    const int i = uniform(find_segment(np, item, [&](int k) { return pc[k].xy0; }));
    const batch_piece &p = pc[i];
    const long local = item - p.xy0;
    switch (p.L) {
#define INQAMD_BXY(LL, R0, R1, R2) \
  case LL: xy_call<R0, R1, R2>(&p.a, local, lds); break;
      INQAMD_BATCH_TABLE(INQAMD_BXY)
#undef INQAMD_BXY
      default: break;
    }
  });
}

__global__ __launch_bounds__(Z_NT) void batch_z_bwd_kernel(const batch_piece *__restrict__ pc, int np,
                                                           unsigned long long *counter,
                                                           int total) {
  extern __shared__ cplx lds[];
  drain_queue(counter, total, [&](int item) {
    const int i = uniform(find_segment(np, item, [&](int k) { return pc[k].z0; }));
    const batch_piece &p = pc[i];
    const long local = item - p.z0;
    const int key = p.L * 4 + (p.ty == 8 ? 0 : p.ty == 16 ? 1 : 2);
    switch (key) {
#define INQAMD_BZB(LL, R0, R1, R2)                                           \
  case LL * 4 + 0: z_bwd_call<R0, R1, R2, 8>(&p.a, local, lds); break;  \
  case LL * 4 + 1: z_bwd_call<R0, R1, R2, 16>(&p.a, local, lds); break; \
  case LL * 4 + 2: z_bwd_call<R0, R1, R2, 32>(&p.a, local, lds); break;
      INQAMD_BATCH_TABLE(INQAMD_BZB)
#undef INQAMD_BZB
      default: break;
    }
  });
}

// One-variant kernels: the pieces of a launch share grid length and tile
// shape, so their items cost the same; one workgroup per item (the item is
// blockIdx, uniform by construction, so the piece table is read through
// scalar loads).
// (z tiles of z_threads(L) threads, as in the one-system kernels)
template <int R0, int R1, int R2, int TY>
__global__ __launch_bounds__(z_threads(R0 * R1 * R2)) void group_z_fwd_kernel(
    const batch_piece *__restrict__ pc, int np, unsigned long long *, int) {
  extern __shared__ cplx lds[];
  const int item = blockIdx.x;
  const int i = find_segment(np, item, [&](int k) { return pc[k].z0; });
  z_fwd_tile<R0, R1, R2, TY, z_threads(R0 * R1 * R2)>(pc[i].a, item - pc[i].z0, lds);
}

constexpr int group_xy_threads(int L) {
  return 2 * plane_threads(L) > 1024 ? 1024 : 2 * plane_threads(L);
}

template <int R0, int R1, int R2>
__global__ __launch_bounds__(group_xy_threads(R0 * R1 * R2))
__attribute__((amdgpu_waves_per_eu(xy_kin_waves<R0, R1, R2, group_xy_threads(R0 * R1 * R2)>())))
void group_xy_kernel(
    const batch_piece *__restrict__ pc, int np, unsigned long long *, int) {
  extern __shared__ cplx lds[];
  const int item = blockIdx.x;
  const int i = find_segment(np, item, [&](int k) { return pc[k].xy0; });
  xy_kin_plane<R0, R1, R2, group_xy_threads(R0 * R1 * R2)>(pc[i].a, item - pc[i].xy0, lds);
}

template <int R0, int R1, int R2, int TY>
__global__ __launch_bounds__(z_threads(R0 * R1 * R2))
__attribute__((amdgpu_waves_per_eu(z_bwd_waves(R0 * R1 * R2))))
void group_z_bwd_kernel(const batch_piece *__restrict__ pc, int np, unsigned long long *,
                        int total) {
  extern __shared__ cplx lds[];
  const int item = blockIdx.x;
  const int i = find_segment(np, item, [&](int k) { return pc[k].z0; });
  // the piece's tiles dealt per XCD as in the one-system kernel (workgroups
  // b, b + 8, ... share an XCD whatever the piece's first item)
  constexpr int TB = Z_COLS / TY;
  const long z0 = pc[i].z0, zend = i + 1 < np ? pc[i + 1].z0 : total;
  const long t = xcd_tile(item - z0, zend - z0, (pc[i].a.nb + TB - 1) / TB);
  z_bwd_epi_tile<R0, R1, R2, TY, z_threads(R0 * R1 * R2)>(pc[i].a, t, lds);
}

// kernels of one launch group
struct group_kernels {
  const void *zf = nullptr, *xy = nullptr, *zb = nullptr;
};

static group_kernels group_kernels_for(int L, int ty) {
  group_kernels g;
  switch (L) {
#define INQAMD_GK(LL, R0, R1, R2)                                                    \
  case LL:                                                                         \
    g.xy = (const void *)group_xy_kernel<R0, R1, R2>;                              \
    if (ty == 8) {                                                                 \
      g.zf = (const void *)group_z_fwd_kernel<R0, R1, R2, 8>;                      \
      g.zb = (const void *)group_z_bwd_kernel<R0, R1, R2, 8>;                      \
    } else if (ty == 16) {                                                         \
      g.zf = (const void *)group_z_fwd_kernel<R0, R1, R2, 16>;                     \
      g.zb = (const void *)group_z_bwd_kernel<R0, R1, R2, 16>;                     \
    } else {                                                                       \
      g.zf = (const void *)group_z_fwd_kernel<R0, R1, R2, 32>;                     \
      g.zb = (const void *)group_z_bwd_kernel<R0, R1, R2, 32>;                     \
    }                                                                              \
    break;
    INQAMD_BATCH_TABLE(INQAMD_GK)
#undef INQAMD_GK
    default: break;
  }
  return g;
}

__global__ void zero_counters_kernel(unsigned long long *c, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) c[i] = 0ULL;
}

static long z_items(int L, int ty, int nb) {
  switch (ty) {
    case 8: return z_tiles<8>(L, nb);
    case 16: return z_tiles<16>(L, nb);
    default: return z_tiles<32>(L, nb);
  }
}

static bool batch_supported_length(int L) {
  switch (L) {
#define INQAMD_BL(LL, R0, R1, R2) case LL:
    INQAMD_BATCH_TABLE(INQAMD_BL)
#undef INQAMD_BL
    return true;
    default: return false;
  }
}

}  // namespace inqamd

using namespace inqamd;

struct inqamd_batch {
  std::vector<inqamd_hamiltonian *> hs;
  hipStream_t stream = nullptr;
  std::string last_error;
  int cus = 256;
  // tuning
  long chunk_bytes = 128L << 20;  // W bytes per batch chunk
  int ty_main = 8;                 // z-tile y extent of full band tiles
  bool narrow_remainders = true;   // nst mod 8 bands in 4- or 2-band tiles
  // tables of the last pointer set
  bool built = false;
  std::vector<const void *> k_psi, k_hpsi;
  std::vector<int> k_nst, k_ld;
  batch_phase_sys *d_phase = nullptr;
  int nphase = 0, phase_items = 0;
  batch_proj_sys *d_proj = nullptr;
  int nproj = 0, proj_items = 0;
  batch_piece *d_pieces = nullptr;
  bool group_by_grid = false;      // one launch per grid length instead of per chunk
  int nstreams = 2;                // streams the chunks are dealt over (after the projector)
  std::vector<hipStream_t> aux;    // streams 1.. (stream 0 is `stream`)
  std::vector<hipEvent_t> aux_done;
  hipEvent_t proj_done = nullptr;
  hipEvent_t start = nullptr;      // counters cleared: the aux streams may begin
  struct group {  // pieces sharing one launch of each sweep
    int p0, np;
    int L, ty;
    int zitems, xyitems;
    size_t zlds, zlds_b, xylds;  // z forward, z backward (tile + twiddles), x-y
    int znt, xynt;
    int zgrid_f, zgrid_b, xygrid;
    group_kernels k;
  };
  struct chunk {  // launch groups whose W is cache resident together
    int g0, ng;
  };
  std::vector<group> groups;
  std::vector<chunk> chunks;
  int proj_grid = 0;
  // V_nl psi expanded in the sweeps from the projector coefficients (every
  // member with projectors allows it), else per ownership entry
  bool coefficients = false;
  unsigned long long *counters = nullptr;
  int ncounters = 0;
  std::vector<int> fallback;  // systems run through the one-system Hpsi
  std::vector<hipEvent_t> events;
  hipEvent_t done = nullptr;
};

namespace {

std::string &g_batch_create_error() {
  thread_local std::string msg;
  return msg;
}

void batch_free_tables(inqamd_batch *b) {
  for (void *p : {(void *)b->d_phase, (void *)b->d_proj, (void *)b->d_pieces, (void *)b->counters})
    if (p) (void)hipFree(p);
  b->d_phase = nullptr;
  b->d_proj = nullptr;
  b->d_pieces = nullptr;
  b->counters = nullptr;
  b->built = false;
}

template <class T>
T *upload_table(const std::vector<T> &v) {
  if (v.empty()) return nullptr;
  T *d = nullptr;
  INQAMD_HIP_CHECK(hipMalloc(&d, v.size() * sizeof(T)));
  INQAMD_HIP_CHECK(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return d;
}

int resident_grid(const void *kernel, int threads, size_t lds, int cus, long items) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, lds) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  long g = (long)per_cu * cus;
  if (g > items) g = items;
  return (int)(g < 1 ? 1 : g);
}

// Work tables for one pointer set.
void batch_build(inqamd_batch *b, const void *const *psi, void *const *hpsi, const int *nst,
                 const int *ld) {
  batch_free_tables(b);
  const int nsys = (int)b->hs.size();
  std::vector<batch_phase_sys> ph;
  std::vector<batch_proj_sys> pj;
  b->fallback.clear();
  int phase_items = 0, proj_items = 0;
  struct pend {
    int s, b0, nb, ty;
  };
  std::vector<pend> pieces;
  b->coefficients = true;
  for (int s = 0; s < nsys; s++) {
    const handle &h = *b->hs[s];
    if (h.three_sweep && batch_supported_length(h.n[0]) && h.pt.nproj > 0 &&
        h.pt.nunique > 0 && !vnl_from_coefficients(h))
      b->coefficients = false;
  }
  for (int s = 0; s < nsys; s++) {
    handle &h = *b->hs[s];
    if (!h.three_sweep || !batch_supported_length(h.n[0])) {
      b->fallback.push_back(s);
      continue;
    }
    const projector_tables &t = h.pt;
    if (t.nproj > 0) {
      batch_phase_sys q;
      q.pos = t.pos;
      q.uent = t.uent;
      q.nsph = t.nproj * t.max_sphere;
      q.nent = t.nent;
      q.kA = h.kA + 3 * KA_CURRENT;
      q.phase_sph = t.phase_sph;
      q.phase_ent = t.phase_ent;
      q.item0 = phase_items;
      phase_items += std::max(q.nsph, q.nent);
      ph.push_back(q);
    }
    if (t.nproj > 0 && t.nunique > 0) {
      batch_proj_sys q{};
      project_expand_args &a = q.a;
      a.psi = (const cplx *)psi[s];
      a.ld = ld[s];
      a.nst = nst[s];
      a.max_sphere = t.max_sphere;
      a.max_nlm = t.max_nlm;
      a.point = t.point;
      a.size = t.size;
      a.mat = t.mat;
      a.coeff = t.coeff;
      a.phase = t.phase_sph;
      a.ent = t.ent;
      a.dV = h.volume / (double)h.npoint;
      if (b->coefficients) {
        a.P = h.proj;
        a.Qe = nullptr;
        a.nbc = nst[s];
        q.nbt = (nst[s] + P4_BANDS - 1) / P4_BANDS;
      } else {
        a.P = nullptr;
        a.Qe = h.vnl;
        a.ldQ = nst[s];
        q.nbt = (nst[s] + PE_BANDS - 1) / PE_BANDS;
      }
      q.item0 = proj_items;
      proj_items += q.nbt * t.nproj;
      pj.push_back(q);
    }
    // band pieces: chunk-sized multiples of 8, then the remainder
    const long per_band = h.npoint * (long)sizeof(cplx);
    int maxb = (int)(b->chunk_bytes / per_band) / 8 * 8;
    if (maxb < 8) maxb = 8;
    const int rem = b->narrow_remainders ? nst[s] % 8 : 0;
    const int full = nst[s] - rem;
    for (int b0 = 0; b0 < full; b0 += maxb)
      pieces.push_back({s, b0, std::min(maxb, full - b0), b->ty_main});
    if (rem > 0) pieces.push_back({s, full, rem, rem <= 2 ? 32 : (rem <= 4 ? 16 : 8)});
  }
  // pack pieces into chunks in order (W of a chunk stays cache resident)
  std::vector<std::vector<pend>> chunk_pieces;
  long used = 0;
  for (const pend &p : pieces) {
    const long bytes = b->hs[p.s]->npoint * (long)p.nb * (long)sizeof(cplx);
    if (chunk_pieces.empty() || used + bytes > b->chunk_bytes) {
      chunk_pieces.emplace_back();
      used = 0;
    }
    chunk_pieces.back().push_back(p);
    used += bytes;
  }
  // launch groups: a whole chunk (mixed grids) or its pieces of one
  // (grid length, tile shape)
  std::vector<batch_piece> pc;
  b->chunks.clear();
  b->groups.clear();
  for (auto &cp : chunk_pieces) {
    if (b->group_by_grid)
      std::stable_sort(cp.begin(), cp.end(), [&](const pend &x, const pend &y) {
        const int lx = b->hs[x.s]->n[0], ly = b->hs[y.s]->n[0];
        return lx != ly ? lx < ly : x.ty < y.ty;
      });
    inqamd_batch::chunk ch;
    ch.g0 = (int)b->groups.size();
    for (size_t i = 0; i < cp.size(); i++) {
      const pend &p = cp[i];
      handle &h = *b->hs[p.s];
      const bool split = b->groups.size() == (size_t)ch.g0 ||
                         (b->group_by_grid && (b->groups.back().L != h.n[0] ||
                                               b->groups.back().ty != p.ty));
      if (split) {
        inqamd_batch::group g{};
        g.p0 = (int)pc.size();
        g.L = h.n[0];
        g.ty = p.ty;
        b->groups.push_back(g);
      }
      inqamd_batch::group &g = b->groups.back();
      batch_piece q{};
      fft3_args &a = q.a;
      a.psi = (const cplx *)psi[p.s] + p.b0;
      a.ld = ld[p.s];
      a.hpsi = (cplx *)hpsi[p.s] + p.b0;
      a.hld = ld[p.s];
      a.W = h.work + (size_t)p.b0 * h.npoint;
      a.nb = p.nb;
      a.V = h.V;
      a.kA = h.kA + 3 * KA_CURRENT;
      for (int d = 0; d < 9; d++) a.rlat[d] = h.rlat[d];
      a.tw = h.axis[0].tw;
      a.scale = 1.0 / (double)h.npoint;
      if (h.pt.nproj > 0 && h.pt.nunique > 0 && b->coefficients) {
        a.own = h.pt.own;
        a.Pc = h.proj + p.b0;
        a.ldPc = nst[p.s];
        a.nlmP = h.pt.max_nlm;
        a.ent_sph = h.pt.ent_sph;
        a.own_sph = h.pt.own_sph;
        a.ent_mat = h.pt.ent_mat;
        a.phase_ent = h.pt.phase_ent;
      } else if (h.pt.nproj > 0 && h.pt.nunique > 0) {
        a.own = h.pt.own;
        a.Qe = h.vnl + p.b0;
        a.ldQ = nst[p.s];
      }
      q.L = h.n[0];
      q.ty = p.ty;
      q.z0 = g.zitems;
      q.xy0 = g.xyitems;
      const long zi = z_items(q.L, q.ty, p.nb);
      const long xi = (long)p.nb * q.L;
      if (g.zitems + zi > INT32_MAX || g.xyitems + xi > INT32_MAX)
        throw error{ERR_UNSUPPORTED, "batch launch group has too many work items"};
      g.zitems += (int)zi;
      g.xyitems += (int)xi;
      g.zlds = std::max(g.zlds, (size_t)q.L * Z_COLS * sizeof(cplx));
      g.zlds_b = std::max(g.zlds_b, z_bwd_lds(q.L));
      g.xylds = std::max(g.xylds, batch_xy_lds(q.L));
      g.xynt = std::max(g.xynt, plane_threads(q.L));
      g.np++;
      pc.push_back(q);
    }
    ch.ng = (int)b->groups.size() - ch.g0;
    b->chunks.push_back(ch);
  }
  for (auto &g : b->groups) {
    if (b->group_by_grid) {
      g.k = group_kernels_for(g.L, g.ty);
    } else {
      g.k.zf = (const void *)batch_z_fwd_kernel;
      g.k.xy = (const void *)batch_xy_kernel;
      g.k.zb = (const void *)batch_z_bwd_kernel;
    }
    if (g.zlds > 65536)
      INQAMD_HIP_CHECK(hipFuncSetAttribute(g.k.zf, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)g.zlds));
    if (g.zlds_b > 65536)
      INQAMD_HIP_CHECK(hipFuncSetAttribute(g.k.zb, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)g.zlds_b));
    if (g.xylds > 65536)
      INQAMD_HIP_CHECK(hipFuncSetAttribute(g.k.xy, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)g.xylds));
    if (b->group_by_grid) {
      // one workgroup per item
      g.zgrid_f = g.zgrid_b = g.zitems;
      g.xygrid = g.xyitems;
      g.xynt = group_xy_threads(g.L);
      g.znt = z_threads(g.L);
    } else {
      g.znt = Z_NT;
      g.zgrid_f = resident_grid(g.k.zf, Z_NT, g.zlds, b->cus, g.zitems);
      g.zgrid_b = resident_grid(g.k.zb, Z_NT, g.zlds_b, b->cus, g.zitems);
      g.xygrid = resident_grid(g.k.xy, g.xynt, g.xylds, b->cus, g.xyitems);
    }
  }
  b->proj_grid = std::max(1, proj_items);
  b->d_phase = upload_table(ph);
  b->nphase = (int)ph.size();
  b->phase_items = phase_items;
  b->d_proj = upload_table(pj);
  b->nproj = (int)pj.size();
  b->proj_items = proj_items;
  b->d_pieces = upload_table(pc);
  b->ncounters = 1 + 3 * (int)b->groups.size();
  INQAMD_HIP_CHECK(hipMalloc(&b->counters, b->ncounters * sizeof(unsigned long long)));
  b->k_psi.assign(psi, psi + nsys);
  b->k_hpsi.assign(hpsi, hpsi + nsys);
  b->k_nst.assign(nst, nst + nsys);
  b->k_ld.assign(ld, ld + nsys);
  b->built = true;
}

bool batch_same(const inqamd_batch *b, const void *const *psi, void *const *hpsi, const int *nst,
                const int *ld) {
  if (!b->built) return false;
  for (size_t s = 0; s < b->hs.size(); s++)
    if (b->k_psi[s] != psi[s] || b->k_hpsi[s] != hpsi[s] || b->k_nst[s] != nst[s] ||
        b->k_ld[s] != ld[s])
      return false;
  return true;
}

void batch_run(inqamd_batch *b) {
  hipStream_t st = b->stream;
  hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(256), 0, st, b->counters, b->ncounters);
  INQAMD_HIP_CHECK(hipGetLastError());
  // Chunks are dealt over the streams, the first to an aux stream: the z
  // forward and x-y sweeps need no V_nl psi, so those of the aux streams'
  // chunks run beside the projector, and each aux stream waits for the
  // projector only before its first z backward sweep.
  const int ns = std::min<int>(b->nstreams, 1 + (int)b->aux.size());
  if (ns > 1) {
    INQAMD_HIP_CHECK(hipEventRecord(b->start, st));
    for (int k = 1; k < ns; k++) INQAMD_HIP_CHECK(hipStreamWaitEvent(b->aux[k - 1], b->start, 0));
  }
  if (b->phase_items > 0) {
    hipLaunchKernelGGL(batch_phase_kernel, dim3((b->phase_items + 255) / 256), dim3(256), 0, st,
                       (const batch_phase_sys *)b->d_phase, b->nphase, b->phase_items);
    INQAMD_HIP_CHECK(hipGetLastError());
  }
  if (b->proj_items > 0 && b->coefficients) {
    hipLaunchKernelGGL(batch_project_p4_kernel, dim3(b->proj_grid), dim3(P4_NT), 0, st,
                       (const batch_proj_sys *)b->d_proj, b->nproj, b->counters, b->proj_items);
    INQAMD_HIP_CHECK(hipGetLastError());
  } else if (b->proj_items > 0) {
    hipLaunchKernelGGL(batch_project_kernel, dim3(b->proj_grid), dim3(PE_NT), 0, st,
                       (const batch_proj_sys *)b->d_proj, b->nproj, b->counters, b->proj_items);
    INQAMD_HIP_CHECK(hipGetLastError());
  }
  if (ns > 1) INQAMD_HIP_CHECK(hipEventRecord(b->proj_done, st));
  bool waited[4] = {true, false, false, false};  // stream 0 runs after the projector
  int ci = 0;
  for (const inqamd_batch::chunk &ch : b->chunks) {
    const int si = (ci + 1) % ns;
    const hipStream_t cs = si == 0 ? st : b->aux[si - 1];
    ci++;
    // z forward of every group of the chunk, then x-y, then z backward
    for (int sweep = 0; sweep < 3; sweep++) {
      if (sweep == 2 && !waited[si]) {
        INQAMD_HIP_CHECK(hipStreamWaitEvent(cs, b->proj_done, 0));
        waited[si] = true;
      }
      for (int gi = ch.g0; gi < ch.g0 + ch.ng; gi++) {
        const inqamd_batch::group &g = b->groups[gi];
        const batch_piece *pc = b->d_pieces + g.p0;
        unsigned long long *cnt = b->counters + 1 + 3 * gi + sweep;
        const void *k = sweep == 0 ? g.k.zf : (sweep == 1 ? g.k.xy : g.k.zb);
        const int grid = sweep == 0 ? g.zgrid_f : (sweep == 1 ? g.xygrid : g.zgrid_b);
        const int nt = sweep == 1 ? g.xynt : g.znt;
        const size_t lds = sweep == 0 ? g.zlds : (sweep == 1 ? g.xylds : g.zlds_b);
        int np = g.np, total = sweep == 1 ? g.xyitems : g.zitems;
        void *args[] = {(void *)&pc, (void *)&np, (void *)&cnt, (void *)&total};
        INQAMD_HIP_CHECK(hipLaunchKernel(k, dim3(grid), dim3(nt), args, lds, cs));
      }
    }
  }
  for (int k = 1; k < ns; k++) {
    // an aux stream that got no z backward sweep still joins after the projector
    if (!waited[k]) INQAMD_HIP_CHECK(hipStreamWaitEvent(b->aux[k - 1], b->proj_done, 0));
    INQAMD_HIP_CHECK(hipEventRecord(b->aux_done[k - 1], b->aux[k - 1]));
    INQAMD_HIP_CHECK(hipStreamWaitEvent(st, b->aux_done[k - 1], 0));
  }
}

template <typename F>
int batch_guarded(inqamd_batch *b, F &&f) {
  try {
    f();
    b->last_error.clear();
    return INQAMD_OK;
  } catch (const inqamd::error &e) {
    b->last_error = e.msg;
    return e.code;
  } catch (const std::bad_alloc &) {
    b->last_error = "host allocation failed";
    return INQAMD_ERR_MEMORY;
  } catch (const std::exception &e) {
    b->last_error = e.what();
    return INQAMD_ERR_ARGUMENT;
  }
}

void batch_release(inqamd_batch *b) {
  batch_free_tables(b);
  for (hipEvent_t e : b->events)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : b->aux_done)
    if (e) (void)hipEventDestroy(e);
  for (hipStream_t q : b->aux)
    if (q) (void)hipStreamDestroy(q);
  if (b->done) (void)hipEventDestroy(b->done);
  if (b->proj_done) (void)hipEventDestroy(b->proj_done);
  if (b->start) (void)hipEventDestroy(b->start);
}

}  // namespace

extern "C" {

int inqamd_batch_create(inqamd_hamiltonian_t const *handles, int nsys, void *stream,
                        inqamd_batch_t *out) {
  if (!handles || !out || nsys <= 0) return INQAMD_ERR_ARGUMENT;
  *out = nullptr;
  inqamd_batch *b = new inqamd_batch();
  const int rc = batch_guarded(b, [&] {
    for (int s = 0; s < nsys; s++) {
      INQAMD_REQUIRE(handles[s] != nullptr, "null handle in the batch");
      b->hs.push_back(handles[s]);
    }
    b->stream = (hipStream_t)stream;
    int dev = 0;
    INQAMD_HIP_CHECK(hipGetDevice(&dev));
    INQAMD_HIP_CHECK(hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, dev));
    size_t zmax = 0, zbmax = 0, xymax = 0;
    for (auto *h : b->hs)
      if (h->three_sweep) {
        zmax = std::max(zmax, (size_t)h->n[0] * Z_COLS * sizeof(cplx));
        zbmax = std::max(zbmax, z_bwd_lds(h->n[0]));
        xymax = std::max(xymax, (size_t)h->n[0] * (h->n[0] + 1) * sizeof(cplx));
      }
    if (zmax > 65536)
      INQAMD_HIP_CHECK(hipFuncSetAttribute((const void *)batch_z_fwd_kernel,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)zmax));
    if (zbmax > 65536)
      INQAMD_HIP_CHECK(hipFuncSetAttribute((const void *)batch_z_bwd_kernel,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)zbmax));
    if (xymax > 65536)
      INQAMD_HIP_CHECK(hipFuncSetAttribute((const void *)batch_xy_kernel,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)xymax));
    b->events.assign(nsys, nullptr);
    for (int s = 0; s < nsys; s++)
      INQAMD_HIP_CHECK(hipEventCreateWithFlags(&b->events[s], hipEventDisableTiming));
    INQAMD_HIP_CHECK(hipEventCreateWithFlags(&b->done, hipEventDisableTiming));
    INQAMD_HIP_CHECK(hipEventCreateWithFlags(&b->proj_done, hipEventDisableTiming));
    INQAMD_HIP_CHECK(hipEventCreateWithFlags(&b->start, hipEventDisableTiming));
    b->aux.assign(3, nullptr);
    b->aux_done.assign(3, nullptr);
    for (int k = 0; k < 3; k++) {
      INQAMD_HIP_CHECK(hipStreamCreateWithFlags(&b->aux[k], hipStreamNonBlocking));
      INQAMD_HIP_CHECK(hipEventCreateWithFlags(&b->aux_done[k], hipEventDisableTiming));
    }
  });
  if (rc != INQAMD_OK) {
    g_batch_create_error() = b->last_error;
    batch_release(b);
    delete b;
    return rc;
  }
  *out = b;
  return INQAMD_OK;
}

int inqamd_batch_set_tuning(inqamd_batch_t b, long chunk_bytes, int ty_main,
                            int narrow_remainders, int group_by_grid, int nstreams) {
  if (!b) return INQAMD_ERR_ARGUMENT;
  return batch_guarded(b, [&] {
    INQAMD_REQUIRE(chunk_bytes >= (1L << 20), "chunk_bytes must be at least 1 MiB");
    INQAMD_REQUIRE(ty_main == 8 || ty_main == 16 || ty_main == 32, "ty_main must be 8, 16 or 32");
    INQAMD_REQUIRE(nstreams >= 1 && nstreams <= 4, "nstreams must be in [1, 4]");
    b->chunk_bytes = chunk_bytes;
    b->ty_main = ty_main;
    b->narrow_remainders = narrow_remainders != 0;
    b->group_by_grid = group_by_grid != 0;
    b->nstreams = nstreams;
    b->built = false;
  });
}

int inqamd_batch_apply(inqamd_batch_t b, const void *const *psi, void *const *hpsi, const int *nst,
                       const int *ld) {
  if (!b) return INQAMD_ERR_ARGUMENT;
  return batch_guarded(b, [&] {
    INQAMD_REQUIRE(psi && hpsi && nst && ld, "null argument arrays");
    const int nsys = (int)b->hs.size();
    for (int s = 0; s < nsys; s++) {
      handle &h = *b->hs[s];
      check_orbital(h, psi[s], "psi");
      check_orbital(h, hpsi[s], "hpsi");
      check_bands(h, nst[s], ld[s], true);
      INQAMD_REQUIRE(psi[s] != hpsi[s], "psi and hpsi must not alias");
    }
    if (!batch_same(b, psi, hpsi, nst, ld)) batch_build(b, psi, hpsi, nst, ld);
    // order the batch after the handles' own streams, and them after it
    for (int s = 0; s < nsys; s++)
      if (b->hs[s]->stream != b->stream) {
        INQAMD_HIP_CHECK(hipEventRecord(b->events[s], b->hs[s]->stream));
        INQAMD_HIP_CHECK(hipStreamWaitEvent(b->stream, b->events[s], 0));
      }
    batch_run(b);
    for (int s : b->fallback) {
      handle &h = *b->hs[s];
      hipStream_t saved = h.stream;
      h.stream = b->stream;
      try {
        hamiltonian_apply(h, (const cplx *)psi[s], ld[s], (cplx *)hpsi[s], ld[s], nst[s],
                          KA_CURRENT, nullptr);
      } catch (...) {
        h.stream = saved;
        throw;
      }
      h.stream = saved;
    }
    bool other = false;
    for (auto *h : b->hs) other = other || h->stream != b->stream;
    if (other) {
      INQAMD_HIP_CHECK(hipEventRecord(b->done, b->stream));
      for (auto *h : b->hs)
        if (h->stream != b->stream) INQAMD_HIP_CHECK(hipStreamWaitEvent(h->stream, b->done, 0));
    }
  });
}

int inqamd_batch_set_stream(inqamd_batch_t b, void *stream) {
  if (!b) return INQAMD_ERR_ARGUMENT;
  b->stream = (hipStream_t)stream;
  return INQAMD_OK;
}

int inqamd_batch_info(inqamd_batch_t b, int *chunks, int *launches, int *pieces,
                      int *fallback) {
  if (!b) return INQAMD_ERR_ARGUMENT;
  if (chunks) *chunks = (int)b->chunks.size();
  if (launches)
    *launches = 1 + (b->phase_items > 0) + (b->proj_items > 0) + 3 * (int)b->groups.size();
  int np = 0;
  for (auto &g : b->groups) np += g.np;
  if (pieces) *pieces = np;
  if (fallback) *fallback = (int)b->fallback.size();
  return INQAMD_OK;
}

const char *inqamd_batch_last_error(inqamd_batch_t b) {
  return b ? b->last_error.c_str() : g_batch_create_error().c_str();
}

void inqamd_batch_destroy(inqamd_batch_t b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  else (void)hipDeviceSynchronize();
  batch_release(b);
  delete b;
}

}  // extern "C"
