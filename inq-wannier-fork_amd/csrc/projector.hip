// Nonlocal projector kernels (projector_all, src/hamiltonian/projector_all.hpp).
//
// project: one workgroup per (sphere, band tile) gathers psi at the sphere
// points with the e^{i(k+A).r} phase and reduces against the projector table
// straight into the nlm x nst coefficient block P = dV D M^T (phase psi)
// (:104-165); the K x nst sphere array of the reference never goes to HBM.
// apply: every grid point touched by any sphere is owned by one thread per
// band, which sums its contributors in a fixed (sphere, point) order -- no
// atomics, bitwise reproducible (replaces the atomic scatter of :225-231).

#include "projector_dev.hpp"

#include <algorithm>
#include <numeric>
#include <tuple>

namespace inqamd {

constexpr int PROJ_BT = 16;     // bands per workgroup
constexpr int PROJ_LANES = 16;  // point lanes per workgroup (PROJ_BT * PROJ_LANES = 256)

struct project_args {
  const cplx *psi;
  long ld;
  int nst;
  int max_sphere;
  int max_nlm;
  const int *point;
  const float *pos;
  const double *mat;
  const double *coeff;
  const int *size;
  const cplx *phase;  // e^{+i(k+A).r} per sphere point
  double dV;
  cplx *P;  // [nproj][max_nlm][nst]
};

template <int NLM>
__global__ __launch_bounds__(256) void project_kernel(const project_args a) {
  __shared__ cplx red[PROJ_LANES][PROJ_BT];
  const int p = blockIdx.y;
  const int bb = threadIdx.x % PROJ_BT;
  const int lane = threadIdx.x / PROJ_BT;
  const int b = blockIdx.x * PROJ_BT + bb;
  const int K = a.size[p];
  const int nlm = a.max_nlm;

  cplx acc[NLM];
#pragma unroll
  for (int lm = 0; lm < NLM; lm++) acc[lm] = mk(0.0, 0.0);

  if (b < a.nst) {
    const long off = (long)p * a.max_sphere;
    const double *M = a.mat + (long)p * a.max_nlm * a.max_sphere;
    for (int i = lane; i < K; i += PROJ_LANES) {
      const long gp = a.point[off + i];
      const cplx v = cmul(a.phase[off + i], a.psi[gp * a.ld + b]);
#pragma unroll
      for (int lm = 0; lm < NLM; lm++) {
        if (lm < nlm) {
          const double m = M[(long)lm * a.max_sphere + i];
          acc[lm].x += m * v.x;
          acc[lm].y += m * v.y;
        }
      }
    }
  }

#pragma unroll
  for (int lm = 0; lm < NLM; lm++) {
    if (lm >= nlm) continue;  // uniform over the workgroup
    red[lane][bb] = acc[lm];
    __syncthreads();
    if (lane == 0 && b < a.nst) {
      cplx s = red[0][bb];
      for (int l = 1; l < PROJ_LANES; l++) s = cadd(s, red[l][bb]);
      // GEMM alpha = dV (projector_all.hpp:131-139), then x coeff (:160-164)
      s = cscale(a.dV, s);
      s = cscale(a.coeff[(long)p * a.max_nlm + lm], s);
      a.P[((long)p * a.max_nlm + lm) * a.nst + b] = s;
    }
    __syncthreads();
  }
}

struct apply_args {
  cplx *hpsi;
  long ld;
  int nst;
  int max_sphere;
  int max_nlm;
  const double *mat;
  const cplx *phase;   // e^{-i(k+A).r} per ownership entry
  const cplx *P;       // coefficients [nproj][max_nlm][nst]  (mode 0)
  const cplx *sphere;  // sphere array [nproj][max_sphere][nst] (mode 1)
  int nunique;
  const int *upoint;
  const int *uoff;
  const int *uent;
};

// FROM_SPHERE: contributions come from an INQ-shaped sphere array instead of
// the P coefficients.  TO_Q: store the per-point sum into the compact
// [nunique][nst] buffer a.hpsi (ld = nst) instead of adding it to hpsi.
template <int FROM_SPHERE, int TO_Q = 0>
__global__ __launch_bounds__(256) void apply_kernel(const apply_args a) {
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long u = gid / a.nst;
  const int b = (int)(gid % a.nst);
  if (u >= a.nunique) return;
  cplx sum = mk(0.0, 0.0);
  const int e1 = a.uoff[u + 1];
  for (int e = a.uoff[u]; e < e1; e++) {
    const int g = a.uent[e];
    const int p = g / a.max_sphere;
    const int i = g - p * a.max_sphere;
    cplx s;
    if constexpr (FROM_SPHERE) {
      s = a.sphere[(long)g * a.nst + b];
    } else {
      s = mk(0.0, 0.0);
      const double *M = a.mat + (long)p * a.max_nlm * a.max_sphere + i;
      const cplx *P = a.P + (long)p * a.max_nlm * a.nst + b;
      for (int lm = 0; lm < a.max_nlm; lm++) {
        const double m = M[(long)lm * a.max_sphere];
        const cplx c = P[(long)lm * a.nst];
        s.x += m * c.x;
        s.y += m * c.y;
      }
    }
    // polar(1, -dot(k, r)) (projector_all.hpp:228)
    sum = cadd(sum, cmul(a.phase[e], s));
  }
  if constexpr (TO_Q) {
    a.hpsi[u * a.nst + b] = sum;
  } else {
    cplx *dst = a.hpsi + (long)a.upoint[u] * a.ld + b;
    *dst = cadd(*dst, sum);
  }
}

// e^{+i(k+A).r} for every sphere point and e^{-i(k+A).r} for every ownership
// entry.
__global__ __launch_bounds__(256) void phase_kernel(const float *pos, const int *uent, int nsph,
                                                    int nent, const double *kA, cplx *phase_sph,
                                                    cplx *phase_ent) {
  phase_item(pos, uent, nsph, nent, kA, phase_sph, phase_ent,
             blockIdx.x * blockDim.x + threadIdx.x);
}

struct expand_args {
  cplx *sphere;
  int nst;
  int max_sphere;
  int max_nlm;
  const double *mat;
  const int *size;
  const cplx *P;
  int nproj;
};

__global__ __launch_bounds__(256) void expand_kernel(const expand_args a) {
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long total = (long)a.nproj * a.max_sphere * a.nst;
  if (gid >= total) return;
  const int b = (int)(gid % a.nst);
  const long g = gid / a.nst;
  const int p = (int)(g / a.max_sphere);
  const int i = (int)(g - (long)p * a.max_sphere);
  cplx s = mk(0.0, 0.0);
  if (i < a.size[p]) {
    for (int lm = 0; lm < a.max_nlm; lm++) {
      const double m = a.mat[((long)p * a.max_nlm + lm) * a.max_sphere + i];
      const cplx c = a.P[((long)p * a.max_nlm + lm) * a.nst + b];
      s.x += m * c.x;
      s.y += m * c.y;
    }
  }
  a.sphere[gid] = s;
}

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------
template <typename T>
static T *upload(const std::vector<T> &v) {
  if (v.empty()) return nullptr;
  T *d = nullptr;
  INQAMD_HIP_CHECK(hipMalloc(&d, v.size() * sizeof(T)));
  INQAMD_HIP_CHECK(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return d;
}

void projector_upload(handle &h, int nproj, int max_sphere, int max_nlm, const int *points,
                      const float *positions, const double *matrices, const double *coeff,
                      const int *nlm, const int *sphere_size) {
  projector_tables &t = h.pt;
  t.nproj = nproj;
  t.max_sphere = max_sphere;
  t.max_nlm = max_nlm;
  if (nproj == 0 || max_sphere == 0 || max_nlm == 0) {
    t.nproj = 0;
    return;
  }
  const long nk = (long)nproj * max_sphere;
  std::vector<int> point(nk, -1), size(nproj, 0);
  std::vector<float> pos(nk * 3, 0.0f);
  std::vector<double> mat((size_t)nk * max_nlm, 0.0), co((size_t)nproj * max_nlm, 0.0);
  std::vector<std::tuple<int, int, int>> owners;  // (grid point, p, i)
  for (int p = 0; p < nproj; p++) {
    const int K = sphere_size[p];
    const int L = nlm[p];
    INQAMD_REQUIRE(K >= 0 && K <= max_sphere, "sphere_size exceeds max_sphere");
    INQAMD_REQUIRE(L >= 0 && L <= max_nlm, "nlm exceeds max_nlm");
    // projector::locally_empty (projector.hpp:89-91): nothing to do
    const bool empty = (K == 0 || L == 0);
    size[p] = empty ? 0 : K;
    for (int i = 0; i < K && !empty; i++) {
      const long g = (long)p * max_sphere + i;
      const int *c = points + g * 3;
      INQAMD_REQUIRE(c[0] >= 0 && c[0] < h.n[0] && c[1] >= 0 && c[1] < h.n[1] && c[2] >= 0 &&
                         c[2] < h.n[2],
                     "sphere point outside the grid");
      point[g] = (c[0] * h.n[1] + c[1]) * h.n[2] + c[2];
      for (int d = 0; d < 3; d++) pos[g * 3 + d] = positions[g * 3 + d];
      for (int lm = 0; lm < L; lm++)
        mat[((long)p * max_nlm + lm) * max_sphere + i] =
            matrices[((long)p * max_nlm + lm) * max_sphere + i];
      owners.emplace_back(point[g], p, i);
    }
    for (int lm = 0; lm < L && !empty; lm++) co[(long)p * max_nlm + lm] = coeff[(long)p * max_nlm + lm];
  }
  // Im(conj p r) - D^2 Im(conj r p) = 2 ((1 + D^2)/2) Im(conj p r): every live
  // projector has one, D = 0 included (the unscaled r-weighted projection)
  std::vector<double> co_inq(co.size(), 0.0);
  for (int p = 0; p < nproj; p++)
    for (int lm = 0; lm < nlm[p] && size[p] > 0; lm++) {
      const double d = co[(size_t)p * max_nlm + lm];
      co_inq[(size_t)p * max_nlm + lm] = 0.5 * (1.0 + d * d);
    }
  std::sort(owners.begin(), owners.end());
  std::vector<int> upoint, uoff, uent;
  for (size_t e = 0; e < owners.size(); e++) {
    const int gp = std::get<0>(owners[e]);
    if (upoint.empty() || upoint.back() != gp) {
      upoint.push_back(gp);
      uoff.push_back((int)e);
    }
    uent.push_back(std::get<1>(owners[e]) * max_sphere + std::get<2>(owners[e]));
  }
  uoff.push_back((int)owners.size());

  std::vector<int> owner(h.npoint, -1), own(h.npoint, -1), ent(nk, -1);
  for (size_t u = 0; u < upoint.size(); u++) {
    owner[upoint[u]] = (int)u;
    const int cnt = uoff[u + 1] - uoff[u];
    if (cnt > OWN_MAX_COUNT)
      throw error{ERR_UNSUPPORTED, "more than " + std::to_string(OWN_MAX_COUNT) +
                                       " projector sphere points share a grid point"};
    if (uoff[u] >= (1 << (31 - OWN_COUNT_BITS)))
      throw error{ERR_UNSUPPORTED, "too many projector sphere points for the ownership map"};
    own[upoint[u]] = (uoff[u] << OWN_COUNT_BITS) | cnt;
  }
  for (size_t e = 0; e < uent.size(); e++) ent[uent[e]] = (int)e;
  if (max_nlm <= 4 && !uent.empty()) {
    std::vector<int> esph(uent.size());
    std::vector<double> emat(uent.size() * 4, 0.0);
    for (size_t e = 0; e < uent.size(); e++) {
      const int p = uent[e] / max_sphere, i = uent[e] - p * max_sphere;
      esph[e] = p;
      for (int lm = 0; lm < max_nlm; lm++)
        emat[e * 4 + lm] = mat[((size_t)p * max_nlm + lm) * max_sphere + i];
    }
    std::vector<int> osph(h.npoint, 0);
    for (size_t u = 0; u < upoint.size(); u++) osph[upoint[u]] = esph[uoff[u]];
    t.ent_sph = upload(esph);
    t.own_sph = upload(osph);
    t.ent_mat = upload(emat);
  }
  t.own = upload(own);
  t.ent = upload(ent);
  t.nent = (int)uent.size();
  t.owner = upload(owner);
  INQAMD_HIP_CHECK(hipMalloc(&t.phase_sph, (size_t)nk * sizeof(cplx)));
  INQAMD_HIP_CHECK(hipMalloc(&t.phase_ent, (size_t)(t.nent > 0 ? t.nent : 1) * sizeof(cplx)));
  t.point = upload(point);
  t.pos = upload(pos);
  t.mat = upload(mat);
  t.coeff = upload(co);
  t.coeff_inq = upload(co_inq);
  t.size = upload(size);
  t.nunique = (int)upoint.size();
  t.upoint = upload(upoint);
  t.uoff = upload(uoff);
  t.uent = upload(uent);
}

void projector_free(handle &h) {
  projector_tables &t = h.pt;
  for (void *p : {(void *)t.point, (void *)t.pos, (void *)t.mat, (void *)t.coeff,
                  (void *)t.coeff_inq, (void *)t.size,
                  (void *)t.upoint, (void *)t.uoff, (void *)t.uent, (void *)t.owner,
                  (void *)t.phase_sph, (void *)t.phase_ent, (void *)t.own, (void *)t.ent,
                  (void *)t.ent_sph, (void *)t.own_sph, (void *)t.ent_mat})
    if (p) (void)hipFree(p);
  t = projector_tables{};
}

void projector_phases(handle &h, int slot) {
  const projector_tables &t = h.pt;
  if (t.nproj == 0) return;
  const int nsph = t.nproj * t.max_sphere;
  const int n = nsph > t.nent ? nsph : t.nent;
  hipLaunchKernelGGL(phase_kernel, dim3((n + 255) / 256), dim3(256), 0, h.stream, t.pos, t.uent,
                     nsph, t.nent, h.kA + 3 * slot, t.phase_sph, t.phase_ent);
  INQAMD_HIP_CHECK(hipGetLastError());
}

void projector_project_coefficients(handle &h, const cplx *psi, long ld, int nst, int slot) {
  const projector_tables &t = h.pt;
  if (t.nproj == 0) return;
  project_args a;
  a.psi = psi;
  a.ld = ld;
  a.nst = nst;
  a.max_sphere = t.max_sphere;
  a.max_nlm = t.max_nlm;
  a.point = t.point;
  a.pos = t.pos;
  a.mat = t.mat;
  a.coeff = t.coeff;
  a.size = t.size;
  a.phase = t.phase_sph;
  (void)slot;
  a.dV = h.volume / (double)h.npoint;
  a.P = h.proj;
  dim3 grid((nst + PROJ_BT - 1) / PROJ_BT, t.nproj);
  if (t.max_nlm <= 4)
    hipLaunchKernelGGL(project_kernel<4>, grid, dim3(256), 0, h.stream, a);
  else
    hipLaunchKernelGGL(project_kernel<16>, grid, dim3(256), 0, h.stream, a);
  INQAMD_HIP_CHECK(hipGetLastError());
}

static apply_args make_apply(handle &h, cplx *hpsi, long ld, int nst, int slot) {
  const projector_tables &t = h.pt;
  apply_args a;
  a.hpsi = hpsi;
  a.ld = ld;
  a.nst = nst;
  a.max_sphere = t.max_sphere;
  a.max_nlm = t.max_nlm;
  a.mat = t.mat;
  a.phase = t.phase_ent;
  (void)slot;
  a.P = h.proj;
  a.sphere = nullptr;
  a.nunique = t.nunique;
  a.upoint = t.upoint;
  a.uoff = t.uoff;
  a.uent = t.uent;
  return a;
}

void projector_apply_coefficients(handle &h, cplx *hpsi, long ld, int nst, int slot) {
  if (h.pt.nproj == 0 || h.pt.nunique == 0) return;
  apply_args a = make_apply(h, hpsi, ld, nst, slot);
  const long total = (long)a.nunique * nst;
  hipLaunchKernelGGL(apply_kernel<0>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     h.stream, a);
  INQAMD_HIP_CHECK(hipGetLastError());
}

void projector_owned_sum(handle &h, cplx *Q, int nst, int slot) {
  if (h.pt.nproj == 0 || h.pt.nunique == 0) return;
  apply_args a = make_apply(h, Q, nst, nst, slot);
  const long total = (long)a.nunique * nst;
  hipLaunchKernelGGL((apply_kernel<0, 1>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     h.stream, a);
  INQAMD_HIP_CHECK(hipGetLastError());
}

void projector_expand(handle &h, cplx *sphere, int nst) {
  const projector_tables &t = h.pt;
  if (t.nproj == 0) return;
  expand_args a;
  a.sphere = sphere;
  a.nst = nst;
  a.max_sphere = t.max_sphere;
  a.max_nlm = t.max_nlm;
  a.mat = t.mat;
  a.size = t.size;
  a.P = h.proj;
  a.nproj = t.nproj;
  const long total = (long)t.nproj * t.max_sphere * nst;
  hipLaunchKernelGGL(expand_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, h.stream,
                     a);
  INQAMD_HIP_CHECK(hipGetLastError());
}

void projector_apply_sphere(handle &h, const cplx *sphere, cplx *hpsi, long ld, int nst,
                            int slot) {
  if (h.pt.nproj == 0 || h.pt.nunique == 0) return;
  apply_args a = make_apply(h, hpsi, ld, nst, slot);
  a.sphere = sphere;
  const long total = (long)a.nunique * nst;
  hipLaunchKernelGGL(apply_kernel<1>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     h.stream, a);
  INQAMD_HIP_CHECK(hipGetLastError());
}

}  // namespace inqamd
