// Tile bodies of the three-sweep local Hpsi (T + V) for cubic grids L^3,
// shared by the one-system kernels (fft3.hip) and the uneven-batch
// work-queue kernels (batch.hip).
//
//   z_fwd_tile     psi [x][y][z][band] (INQ layout) -> W[band][z][x][y]
//                  tile: one x, TY y, TB = 64/TY bands, all z.  The first
//                  stage reads psi band-fastest straight into registers, a
//                  z-FFT over the 64 (y, band) columns, and a last stage that
//                  writes TY consecutive y per (band, z).
//   xy_kin_plane   one (band, z) plane W[b][z][.][.] (contiguous L*L):
//                  x-FFT, y-FFT, x 1/2|G+k+A|^2, inverse y, inverse x, in a
//                  padded LDS plane; first/last stage touch HBM directly.
//   z_bwd_epi_tile W rows and twiddles copied into LDS by DMA while V psi and
//                  V_nl psi of the owned points are formed in registers, the
//                  inverse z-FFT, then a transposing write hpsi = FFT/N +
//                  V psi + V_nl psi (+ Taylor accumulations) in INQ layout.
#pragma once

#include "fft.hpp"
#include "fft_dft.hpp"
#include "fft_stage.hpp"

namespace inqamd {

constexpr int Z_COLS = 64;  // FFT columns per z tile: TY y values x TB bands
constexpr int Z_NT = 512;

// LDS slot of column c in a 64-column (1 KB) row of a z tile.  The XOR
// spreads the transposed (band-fastest) accesses of the psi load and of the
// epilogue over all 16 bank groups of 16-byte elements, and keeps the
// FFT stages' contiguous column accesses conflict-free.
__device__ __forceinline__ int zswz(int c) { return c ^ (((c >> 4) & 3) << 1); }

// The same for rows of ZC = 32 or 64 columns (32: the bit above a 16-column
// block flips bit 1, so the transposed accesses of a quarter wave, 8 bands x
// 2 y, fall in distinct bank groups; the FFT stages' accesses of 16
// consecutive columns stay in their aligned 16-slot block)
template <int ZC>
__device__ __forceinline__ int zswz_n(int c) {
  static_assert(ZC == 32 || ZC == 64, "z tile width");
  return c ^ (((c >> 4) & (ZC / 16 - 1)) << 1);
}

// s_waitcnt immediate that waits for every vector memory operation of the
// wave (vmcnt 0) and leaves the export and LDS/scalar counters alone
constexpr unsigned WAITCNT_VM0 = 0x0F70;

constexpr int plane_threads(int L) {
  int t = (L * L / 16 + 63) / 64 * 64;
  return t < 64 ? 64 : (t > 1024 ? 1024 : t);
}

// ---------------------------------------------------------------------------
// sweep 1: z forward, INQ layout -> band-major W
// ---------------------------------------------------------------------------
// POT: the tile transforms V psi + V_nl psi (ownership entries Qe) instead
// of psi, formed as the first stage loads psi.
template <int R0, int R1, int R2, int Z_TY, int NT = Z_NT, bool POT = false>
__device__ __forceinline__ void z_fwd_tile(const fft3_args &a, long t, cplx *lds) {
  using P = plan3<R0, R1, R2>;
  constexpr int L = P::LEN;
  constexpr int Z_TB = Z_COLS / Z_TY;
  // lds: [L][Z_COLS], column = b * Z_TY + y

  const int nyt = (L + Z_TY - 1) / Z_TY;
  const int nbt = (a.nb + Z_TB - 1) / Z_TB;
  const int bt = (int)(t % nbt);
  t /= nbt;
  const int yt = (int)(t % nyt);
  const int x = (int)(t / nyt);
  const int y0 = yt * Z_TY, cb0 = bt * Z_TB;
  const int gb0 = cb0 + (cb0 >> (3 + a.glog)) * a.gskip;  // the tile's first band in psi

  // The first stage reads psi straight into registers: a wave takes one
  // butterfly input of all 64 columns (TY y rows x TB bands), i.e. TY runs
  // of TB * 16 contiguous bytes, each used whole.
  auto gl = [&](int i, int c) -> cplx {
    const int b = c / Z_TY, y = c % Z_TY;
    if (!(y0 + y < L && cb0 + b < a.nb)) return mk(0.0, 0.0);
    const long p = ((long)x * L + y0 + y) * L + i;
    cplx v = a.psi[p * a.ld + gb0 + b];
    if constexpr (POT) {
      const double vp = a.V[p];
      v = mk(vp * v.x, vp * v.y);
      // own[p] = -1 outside every sphere
      const int w = a.own ? a.own[p] : -1;
      if (w >= 0) {
        const int e0 = w >> OWN_COUNT_BITS, cnt = w & OWN_MAX_COUNT;
        for (int j = 0; j < cnt; j++) v = cadd(v, a.Qe[(long)(e0 + j) * a.ldQ + gb0 + b]);
      }
    }
    return v;
  };
  auto ll = [&](int i, int c) -> cplx { return lds[i * Z_COLS + zswz(c)]; };
  auto ls = [&](int i, int c, cplx v, int, int) { lds[i * Z_COLS + zswz(c)] = v; };
  auto gs = [&](int i, int c, cplx v, int, int) {
    const int b = c / Z_TY, y = c % Z_TY;
    if (y0 + y < L && cb0 + b < a.nb)
      a.W[(((long)(cb0 + b) * L + i) * L + x) * L + y0 + y] = v;
  };
  run_plan<P, R0, R1, R2, Z_COLS, NT>(a.tw, -1, false, gl, ll, ls, gs, false);
}

// ---------------------------------------------------------------------------
// sweep 2: 2-D x-y transform of one plane with the kinetic factor
// ---------------------------------------------------------------------------
// Radices of the x-y plane transforms: the grid's plan (R0, R1, R2), except
// at 96 where the plane runs two stages (8 x 12: fewer LDS passes and
// barriers per plane) and the z tiles keep three (their radix-12 stage costs
// more in the transposing epilogue than it saves).
constexpr int xy_radix(int i, int R0, int R1, int R2) {
  const int L = R0 * R1 * R2;
  const int r[3] = {R0, R1, R2};
  const int r96[3] = {8, 12, 1};
  return L == 96 ? r96[i] : r[i];
}

template <int R0, int R1, int R2>
struct xy_plan {
  static constexpr int X0 = xy_radix(0, R0, R1, R2), X1 = xy_radix(1, R0, R1, R2),
                       X2 = xy_radix(2, R0, R1, R2);
  using P = plan3<X0, X1, X2>;
  static_assert(P::LEN == R0 * R1 * R2, "x-y radices must multiply to the grid length");
};

// Two-stage x-y planes whose complex LDS array (L x (L+1) x 16 B) leaves room
// for one workgroup per CU exchange real and imaginary parts in turn through
// half of it (xchg_split), so that two share a CU: at 72 (80 registers at
// 768 threads).  At 96 two workgroups' registers would spill (64 or 96 per
// thread against 110-128), so it keeps the complex exchange.
template <int R0, int R1, int R2>
constexpr bool xy_split() {
  constexpr int L = R0 * R1 * R2;
  return xy_plan<R0, R1, R2>::P::NSTAGE == 2 && 2 * L * (L + 1) * 16 > 160 * 1024 && L < 96;
}

// LDS of an x-y kinetic plane's data
template <int R0, int R1, int R2>
constexpr size_t xy_plane_lds() {
  constexpr size_t L = R0 * R1 * R2;
  return L * (L + 1) * (xy_split<R0, R1, R2>() ? sizeof(double) : sizeof(cplx));
}

// The long planes (72, 80, 96: one to three per CU) take their twiddle table
// into LDS behind the plane.  (Shorter planes ran their Hpsi slower with it,
// though the plane kernel alone ran faster.)
template <int R0, int R1, int R2>
constexpr bool xy_tw_lds() {
  return R0 * R1 * R2 >= 72;
}

// Dynamic LDS of an x-y kinetic plane
template <int R0, int R1, int R2>
constexpr size_t xy_kin_lds() {
  constexpr size_t L = R0 * R1 * R2;
  return xy_plane_lds<R0, R1, R2>() + (xy_tw_lds<R0, R1, R2>() ? L * sizeof(cplx) : 0);
}

// Occupancy floor (amdgpu_waves_per_eu) of an NT-thread x-y kinetic kernel:
// planes with split exchanges ask for the registers of two workgroups per
// CU, the others for one (the floor the launch bound implies by itself)
template <int R0, int R1, int R2, int NT>
constexpr int xy_kin_waves() {
  return ((xy_split<R0, R1, R2>() ? 2 : 1) * (NT / 64) + 3) / 4;
}

// Registers of the plane's first stage (the x-forward inputs)
template <int R0, int R1, int R2, int NT>
using xy_regs = stage_regs<xy_plan<R0, R1, R2>::X0, R0 * R1 * R2, R0 * R1 * R2, NT>;

template <int R0, int R1, int R2, int NT>
__device__ __forceinline__ void xy_load(const fft3_args &a, long plane, xy_regs<R0, R1, R2, NT> &in) {
  constexpr int L = R0 * R1 * R2;
  const cplx *W = a.W + plane * L * L;
  pload<xy_plan<R0, R1, R2>::X0, L, L, NT>(in, [&](int i, int c) -> cplx { return W[i * L + c]; });
}

// Cartesian k + A (the covariant vector through the reciprocal lattice / 2 pi)
__device__ __forceinline__ void xy_kc(const fft3_args &a, double kc[3]) {
  const double inv2pi = 0.15915494309189533576888376337251;
  const double k0 = a.kA[0], k1 = a.kA[1], k2 = a.kA[2];
#pragma unroll
  for (int d = 0; d < 3; d++)
    kc[d] = (a.rlat[0 * 3 + d] * k0 + a.rlat[1 * 3 + d] * k1 + a.rlat[2 * 3 + d] * k2) * inv2pi;
}

// Kinetic factor 1/2 |G+k+A|^2 on the registers of the y-forward last
// stage (radix X1, span X0), where butterfly j < X0 of column c holds
// outputs (y = j + r X0, x = c), r < X1; they are exactly the inputs of the
// first stage of the inverse run with the radices in reverse order (X1, X0),
// which therefore starts from the same registers (one LDS round trip and a
// barrier fewer).  G (optional) receives the transform before the factor.
template <int X0, int X1, int L, int NT>
__device__ __forceinline__ void xy_kin_regs(stage_regs<X1, L, L, NT> &r, cplx *G, const double b0[3],
                                            const double b1[3], const double gz[3]) {
  constexpr int NB = L / X1;  // = X0
  constexpr int NBF = stage_regs<X1, L, L, NT>::NBF;
#pragma unroll
  for (int q = 0; q < NBF; q++) {
    const int w = threadIdx.x + q * NT;
    if ((int)threadIdx.x < NT && (NBF * NT == NB * L || w < NB * L)) {
      const int j = w / L, c = w % L;
#pragma unroll
      for (int rr = 0; rr < X1; rr++) {
        const int i = j + rr * X0;
        if (G) G[c * L + i] = r.v[q][rr];
        const int my = symmetric_index(i, L), mx = symmetric_index(c, L);
        const double q0 = mx * b0[0] + my * b1[0] + gz[0];
        const double q1 = mx * b0[1] + my * b1[1] + gz[1];
        const double q2 = mx * b0[2] + my * b1[2] + gz[2];
        r.v[q][rr] = cscale(0.5 * (q0 * q0 + q1 * q1 + q2 * q2), r.v[q][rr]);
      }
    }
  }
}

// The plane transform from its first-stage inputs already in registers
template <int R0, int R1, int R2, int NT>
__device__ __forceinline__ void xy_kin_plane_from(const fft3_args &a, long plane, cplx *lds,
                                                  xy_regs<R0, R1, R2, NT> &in,
                                                  const cplx *__restrict__ tw,
                                                  const double kc[3]) {
  using XP = xy_plan<R0, R1, R2>;
  using P = typename XP::P;
  constexpr int X0 = XP::X0, X1 = XP::X1, X2 = XP::X2;
  constexpr int L = P::LEN;
  constexpr int PAD = L + 1;  // row stride of the LDS plane (breaks bank aliasing)
  // lds: [x][PAD]; plane = band * L + z
  const int z = (int)(plane % L);
  cplx *W = a.W + plane * L * L;

  // x-axis accessors: element (i = x, c = y)
  auto lx = [&](int i, int c) -> cplx { return lds[i * PAD + c]; };
  auto sx = [&](int i, int c, cplx v, int, int) { lds[i * PAD + c] = v; };
  auto wx = [&](int i, int c, cplx v, int, int) { W[i * L + c] = v; };
  // y-axis accessors: element (i = y, c = x)
  auto ly = [&](int i, int c) -> cplx { return lds[c * PAD + i]; };
  auto sy = [&](int i, int c, cplx v, int, int) { lds[c * PAD + i] = v; };

  // 1/2 |G + k + A|^2 with G = m_x b0 + m_y b1 + m_z b2 (cartesian)
  const int mz = symmetric_index(z, L);
  double gz[3], b0[3], b1[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    gz[d] = mz * a.rlat[2 * 3 + d] + kc[d];
    b0[d] = a.rlat[0 * 3 + d];
    b1[d] = a.rlat[1 * 3 + d];
  }
  cplx *G = a.gdump ? a.gdump + plane * L * L : nullptr;
  auto ky = [&](int i, int c, cplx v, int, int) {
    if (G) G[c * L + i] = v;
    const int my = symmetric_index(i, L), mx = symmetric_index(c, L);
    const double q0 = mx * b0[0] + my * b1[0] + gz[0];
    const double q1 = mx * b0[1] + my * b1[1] + gz[1];
    const double q2 = mx * b0[2] + my * b1[2] + gz[2];
    lds[c * PAD + i] = cscale(0.5 * (q0 * q0 + q1 * q1 + q2 * q2), v);
  };

  if constexpr (xy_split<R0, R1, R2>()) {
    // every stage from registers; the exchanges between them go through
    // the double array one component at a time.  Slots of element (x, y):
    // x * PAD + y, as (i = x, c = y) on the x axis, (i = y, c = x) on y.
    double *ld = reinterpret_cast<double *>(lds);
    auto xs = [&](int i, int c) { return i * PAD + c; };
    auto ys = [&](int i, int c) { return c * PAD + i; };
    pcompute<X0, 1, L, L, NT>(in, tw, -1);  // x forward
    stage_regs<X1, L, L, NT> r1;
    xchg_split<X0, 1, X1, L, L, NT>(in, r1, ld, xs, xs);
    pcompute<X1, X0, L, L, NT>(r1, tw, -1);
    stage_regs<X0, L, L, NT> r2;  // y forward
    xchg_split<X1, X0, X0, L, L, NT>(r1, r2, ld, xs, ys);
    pcompute<X0, 1, L, L, NT>(r2, tw, -1);
    stage_regs<X1, L, L, NT> r3;
    xchg_split<X0, 1, X1, L, L, NT>(r2, r3, ld, ys, ys);
    pcompute<X1, X0, L, L, NT>(r3, tw, -1);
    xy_kin_regs<X0, X1, L, NT>(r3, G, b0, b1, gz);
    pcompute<X1, 1, L, L, NT>(r3, tw, +1);  // y backward, plan (X1, X0)
    stage_regs<X0, L, L, NT> r4;
    xchg_split<X1, 1, X0, L, L, NT>(r3, r4, ld, ys, ys);
    pcompute<X0, X1, L, L, NT>(r4, tw, +1);
    stage_regs<X0, L, L, NT> r5;  // x backward, plan (X0, X1)
    xchg_split<X0, X1, X0, L, L, NT>(r4, r5, ld, ys, xs);
    pcompute<X0, 1, L, L, NT>(r5, tw, +1);
    stage_regs<X1, L, L, NT> r6;
    xchg_split<X0, 1, X1, L, L, NT>(r5, r6, ld, xs, xs);
    pcompute<X1, X0, L, L, NT>(r6, tw, +1);
    // last stage (radix X1, span X0): butterfly j < X0 of column y holds
    // x = j + r X0
    constexpr int NB = L / X1;
    constexpr int NBF = stage_regs<X1, L, L, NT>::NBF;
#pragma unroll
    for (int q = 0; q < NBF; q++) {
      const int w = threadIdx.x + q * NT;
      if ((int)threadIdx.x < NT && (NBF * NT == NB * L || w < NB * L)) {
        const int j = w / L, c = w % L;
#pragma unroll
        for (int rr = 0; rr < X1; rr++) W[(j + rr * X0) * L + c] = r6.v[q][rr];
      }
    }
    return;
  }

  run_plan_from<P, X0, X1, X2, L, NT>(in, tw, -1, false, lx, sx, sx, true);  // x forward
  if constexpr (P::NSTAGE == 2) {
    pstage<X0, 1, L, L, NT>(tw, -1, ly, sy, true);  // y forward
    __syncthreads();
    stage_regs<X1, L, L, NT> r;
    pload<X1, L, L, NT>(r, ly);
    pcompute<X1, X0, L, L, NT>(r, tw, -1);
    // the factor of xy_kin_regs, written out: through the call the compiler
    // schedules these planes' code differently (fewer FMAs, 2 % more
    // instructions at 64)
    constexpr int NB = L / X1;  // = X0
    constexpr int NBF = stage_regs<X1, L, L, NT>::NBF;
#pragma unroll
    for (int q = 0; q < NBF; q++) {
      const int w = threadIdx.x + q * NT;
      if ((int)threadIdx.x < NT && (NBF * NT == NB * L || w < NB * L)) {
        const int j = w / L, c = w % L;
#pragma unroll
        for (int rr = 0; rr < X1; rr++) {
          const int i = j + rr * X0;
          if (G) G[c * L + i] = r.v[q][rr];
          const int my = symmetric_index(i, L), mx = symmetric_index(c, L);
          const double q0 = mx * b0[0] + my * b1[0] + gz[0];
          const double q1 = mx * b0[1] + my * b1[1] + gz[1];
          const double q2 = mx * b0[2] + my * b1[2] + gz[2];
          r.v[q][rr] = cscale(0.5 * (q0 * q0 + q1 * q1 + q2 * q2), r.v[q][rr]);
        }
      }
    }
    pfinish<X1, 1, L, L, NT>(r, tw, +1, sy, true);  // y backward, plan (X1, X0)
    __syncthreads();
    pstage<X0, X1, L, L, NT>(tw, +1, ly, sy, true);
    __syncthreads();
  } else {
    run_plan<P, X0, X1, X2, L, NT>(tw, -1, true, ly, ly, sy, ky, true);  // y forward, x T
    run_plan<P, X0, X1, X2, L, NT>(tw, +1, true, ly, ly, sy, sy, true);  // y backward
  }
  run_plan<P, X0, X1, X2, L, NT>(tw, +1, true, lx, lx, sx, wx, false);  // x backward
}

template <int R0, int R1, int R2, int NT = plane_threads(R0 * R1 * R2)>
__device__ __forceinline__ void xy_kin_plane(const fft3_args &a, long plane, cplx *lds) {
  constexpr int L = R0 * R1 * R2;
  xy_regs<R0, R1, R2, NT> in;
  xy_load<R0, R1, R2, NT>(a, plane, in);
  const cplx *tw = a.tw;
  if constexpr (xy_tw_lds<R0, R1, R2>()) {
    // the twiddles go straight into LDS (16 B per lane) with the plane's
    // loads in flight; the first stage needs none, and the barrier after it
    // makes every wave's copy visible
    static_assert(NT >= 64 * ((L + 63) / 64), "waves for the twiddle copy");
    cplx *ltw = reinterpret_cast<cplx *>(reinterpret_cast<char *>(lds) + xy_plane_lds<R0, R1, R2>());
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (64 * wave < L && 64 * wave + lane < L)
      __builtin_amdgcn_global_load_lds(a.tw + 64 * wave + lane, ltw + 64 * wave, 16, 0, 0);
    __builtin_amdgcn_s_waitcnt(WAITCNT_VM0);
    tw = ltw;
  }
  double kc[3];
  xy_kc(a, kc);
  xy_kin_plane_from<R0, R1, R2, NT>(a, plane, lds, in, tw, kc);
}

// The forward x-y transform of one plane (G-space Hpsi), in place in W
template <int R0, int R1, int R2, int NT = plane_threads(R0 * R1 * R2)>
__device__ __forceinline__ void xy_fwd_plane(const fft3_args &a, long plane, cplx *lds) {
  using XP = xy_plan<R0, R1, R2>;
  using P = typename XP::P;
  constexpr int X0 = XP::X0, X1 = XP::X1, X2 = XP::X2;
  constexpr int L = P::LEN;
  constexpr int PAD = L + 1;
  cplx *W = a.W + plane * L * L;
  auto gx = [&](int i, int c) -> cplx { return W[i * L + c]; };          // (i = x, c = y)
  auto lx = [&](int i, int c) -> cplx { return lds[i * PAD + c]; };
  auto sx = [&](int i, int c, cplx v, int, int) { lds[i * PAD + c] = v; };
  auto ly = [&](int i, int c) -> cplx { return lds[c * PAD + i]; };    // (i = y, c = x)
  auto sy = [&](int i, int c, cplx v, int, int) { lds[c * PAD + i] = v; };
  auto wy = [&](int i, int c, cplx v, int, int) { W[c * L + i] = v; };
  run_plan<P, X0, X1, X2, L, NT>(a.tw, -1, false, gx, lx, sx, sx, true);  // x forward
  run_plan<P, X0, X1, X2, L, NT>(a.tw, -1, true, ly, ly, sy, wy, false);  // y forward
}

// ---------------------------------------------------------------------------
// sweep 3: z backward, x 1/N + V psi, band-major W -> INQ layout
// ---------------------------------------------------------------------------
// ACC: the caller may pass Taylor accumulators (acc1/acc2); without it they
// must be null and the epilogue is a plain store.
// GS (G-space Hpsi): no z transform (the first sweep made it), only the
// transposing write, whose epilogue adds 1/2 |G+k+A|^2 kin_in instead of
// V psi (no ownership entries, no scale).
// FWD (Taylor chains): the tile then transforms its own hpsi forward in z
// into W, in place of the next application's z_fwd sweep over the same tile
// (the chunk's W region the tile read is the region it writes).
// Position of the n-th (from 0) set bit of m, which has more than n set bits
__device__ __forceinline__ int nth_set_bit(unsigned long long m, int n) {
  int pos = 0;
#pragma unroll
  for (int s = 32; s >= 1; s >>= 1) {
    const int c = __popcll(m & ((1ull << s) - 1));
    if (n >= c) {
      n -= c;
      m >>= s;
      pos += s;
    }
  }
  return pos;
}

// projector_all::apply (projector_all.hpp:172-212, 225-231) expanded from
// the coefficients P[p][lm][b] at a z tile's owned outputs (output k of this
// thread: own[k] >= 0, its first entry's sphere in osp[k]), each output's
// ownership entries in fixed (sphere, point) order, added to pre[k].  About
// one output in five is owned, so each wave compacts its owned outputs with
// ballots (queue position: outputs of slice k = 0, 1, ... in lane order) and
// expands them 64 at a time, the entry-table and coefficient loads of a round
// all in flight together.  Outputs and results move between lanes by
// cross-lane permutes, which leave the LDS tile to the W rows loading into it.
template <int NE, int Z_TB>
__device__ __forceinline__ void vnl_expand_wave(const fft3_args &a, int gb0, const int (&own)[NE],
                                                const int (&osp)[NE], cplx (&pre)[NE]) {
  const int lane = threadIdx.x & 63;
  const int wave0 = threadIdx.x & ~63;  // first thread of the wave
  const unsigned long long lt = __lanemask_lt();
  unsigned long long m[NE];
  int P[NE + 1];  // queue position of slice k's first owned output (uniform)
  P[0] = 0;
#pragma unroll
  for (int k = 0; k < NE; k++) {
    m[k] = __ballot(own[k] >= 0);
    P[k + 1] = P[k] + __popcll(m[k]);
  }
  const int T = P[NE];
  for (int r0 = 0; r0 < T; r0 += 64) {
    // this lane expands queue position r0 + lane: owned output (src, ks)
    const int pos = r0 + lane;
    int ks = 0;
#pragma unroll
    for (int k = 1; k < NE; k++)
      if (P[k] <= pos) ks = k;
    unsigned long long mks = m[0];
    int pks = 0;
#pragma unroll
    for (int k = 1; k < NE; k++)
      if (ks == k) {
        mks = m[k];
        pks = P[k];
      }
    const int src = pos < T ? nth_set_bit(mks, pos - pks) : lane;
    int w = 0, sp0 = 0;
#pragma unroll
    for (int k = 0; k < NE; k++) {
      if (P[k] < r0 + 64 && P[k + 1] > r0) {  // slice k has positions in this round
        const int ow = __shfl(own[k], src);
        const int os = __shfl(osp[k], src);
        if (ks == k) {
          w = ow;
          sp0 = os;
        }
      }
    }
    cplx v = mk(0.0, 0.0);
    if (pos < T) {
      const int cb = gb0 + (wave0 + src) % Z_TB;
      const int e0 = w >> OWN_COUNT_BITS, cnt = w & OWN_MAX_COUNT;
      for (int j = 0; j < cnt; j++) {
        const int en = e0 + j;
        const int sp = j == 0 ? sp0 : a.ent_sph[en];
        const double4 mm = reinterpret_cast<const double4 *>(a.ent_mat)[en];
        const double mv[4] = {mm.x, mm.y, mm.z, mm.w};
        const cplx f = a.phase_ent[en];
        const cplx *Pp = a.Pc + (long)sp * a.nlmP * a.ldPc + cb;
        // the four coefficient loads together (lm past nlmP rereads lm 0), then
        // the sum over the sphere's lm in order
        cplx c[4];
#pragma unroll
        for (int lm = 0; lm < 4; lm++) c[lm] = Pp[(long)(lm < a.nlmP ? lm : 0) * a.ldPc];
        cplx u = mk(0.0, 0.0);
#pragma unroll
        for (int lm = 0; lm < 4; lm++) {
          if (lm < a.nlmP) {
            u.x += mv[lm] * c[lm].x;
            u.y += mv[lm] * c[lm].y;
          }
        }
        v = cadd(v, cmul(f, u));
      }
    }
    // each owned output of the round takes its value from the lane that expanded it
#pragma unroll
    for (int k = 0; k < NE; k++) {
      if (P[k] < r0 + 64 && P[k + 1] > r0) {
        const int j = P[k] + __popcll(m[k] & lt) - r0;
        const bool here = own[k] >= 0 && j >= 0 && j < 64;
        const int from = here ? j : lane;
        const double vx = __shfl(v.x, from), vy = __shfl(v.y, from);
        if (here) pre[k] = cadd(pre[k], mk(vx, vy));
      }
    }
  }
}

// ZC: columns of the tile (TY y values x ZC / TY bands), 64 or 32
template <int R0, int R1, int R2, int Z_TY, int NT = Z_NT, bool ACC = false, bool GS = false,
          bool FWD = false, int ZC = Z_COLS>
__device__ __forceinline__ void z_bwd_epi_tile(const fft3_args &a, long t, cplx *lds) {
  using P = plan3<R0, R1, R2>;
  constexpr int L = P::LEN;
  constexpr int Z_COLS = ZC;  // (shadows the namespace constant in this tile)
  constexpr int Z_TB = Z_COLS / Z_TY;
  constexpr int RPI = 64 / Z_COLS;  // tile rows one wave's 16-byte DMA fills
  static_assert(L % RPI == 0, "whole rows per DMA instruction");
  auto zswz = [](int c) { return zswz_n<ZC>(c); };
  static_assert(NT % Z_TB == 0, "thread -> band mapping of the epilogue");
  constexpr int NE = (L * Z_COLS + NT - 1) / NT;  // output elements per thread

  const int nyt = (L + Z_TY - 1) / Z_TY;
  const int nbt = (a.nb + Z_TB - 1) / Z_TB;
  const int bt = (int)(t % nbt);
  t /= nbt;
  const int yt = (int)(t % nyt);
  const int x = (int)(t / nyt);
  const int y0 = yt * Z_TY, cb0 = bt * Z_TB;
  const int gb0 = cb0 + (cb0 >> (3 + a.glog)) * a.gskip;  // the tile's first band in psi, hpsi

  // The tile's W rows go straight from HBM into the LDS tile (16 B per lane:
  // lane l of a wave's load fills slot l of 64 / ZC consecutive rows, so it
  // loads row l / ZC, column zswz(l % ZC), zswz being its own inverse),
  // issued first so that they are in flight with the prefetch below and the
  // V_nl expansion, which use no LDS.  Columns outside the grid or the chunk
  // load W[0] instead; their transforms are never stored.
  // The twiddle table (L entries) goes the same way into the L slots after
  // the tile, so that the transforms' twiddles are LDS reads.
  cplx *ltw = lds + L * Z_COLS;
  if constexpr (!GS) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane / Z_COLS, c = zswz(lane % Z_COLS);
    const int b = c / Z_TY, y = c % Z_TY;
    const bool ok = y0 + y < L && cb0 + b < a.nb;
    const long col = ok ? (((long)(cb0 + b) * L) * L + x) * L + y0 + y : 0;
    const long row = ok ? (long)L * L : 0;
    for (int i = wave * RPI; i < L; i += (NT / 64) * RPI)
      __builtin_amdgcn_global_load_lds(a.W + col + (i + lr) * row, lds + i * Z_COLS, 16, 0, 0);
    if (64 * wave < L && 64 * wave + lane < L)
      __builtin_amdgcn_global_load_lds(a.tw + 64 * wave + lane, ltw + 64 * wave, 16, 0, 0);
  }

  // prefetch V psi (and the ownership entries) of this thread's outputs,
  // or 1/2 |G+k+A|^2 kin_in.  Every load of the prefetch is issued before
  // any product waits for one: an output outside the grid or the chunk loads
  // from point 0, band gb0 (in bounds) and is zeroed afterwards, so the loads
  // need no branches, each of which would wait for its own loads.
  cplx pre[NE];
  int own[NE];
  int osp[NE];  // sphere of the first ownership entry (coefficient expansion)
  int pk[NE];   // point of output k (0 outside the tile's valid outputs)
  int bk[NE];   // band offset of output k in psi (gb0 + b, or gb0)
  bool vk[NE];
#pragma unroll
  for (int k = 0; k < NE; k++) {
    const int e = threadIdx.x + k * NT;
    const int b = e % Z_TB, y = (e / Z_TB) % Z_TY, z = e / Z_COLS;
    vk[k] = e < L * Z_COLS && y0 + y < L && cb0 + b < a.nb;
    pk[k] = vk[k] ? (x * L + y0 + y) * L + z : 0;
    bk[k] = gb0 + (vk[k] ? b : 0);
  }
  if constexpr (GS) {
    double kc[3], gx[3];
    xy_kc(a, kc);
    const int mx = symmetric_index(x, L);
#pragma unroll
    for (int d = 0; d < 3; d++) gx[d] = mx * a.rlat[0 * 3 + d] + kc[d];
    cplx kin[NE];
#pragma unroll
    for (int k = 0; k < NE; k++) kin[k] = a.kin_in[(long)pk[k] * a.kin_ld + bk[k]];
#pragma unroll
    for (int k = 0; k < NE; k++) {
      const int e = threadIdx.x + k * NT;
      const int y = (e / Z_TB) % Z_TY, z = e / Z_COLS;
      const int my = symmetric_index(y0 + y, L), mz = symmetric_index(z, L);
      const double q0 = gx[0] + my * a.rlat[3] + mz * a.rlat[6];
      const double q1 = gx[1] + my * a.rlat[4] + mz * a.rlat[7];
      const double q2 = gx[2] + my * a.rlat[5] + mz * a.rlat[8];
      pre[k] = vk[k] ? cscale(0.5 * (q0 * q0 + q1 * q1 + q2 * q2), kin[k]) : mk(0.0, 0.0);
      own[k] = -1;
      osp[k] = 0;
    }
  } else {
    cplx ps[NE];
    double vp[NE];
#pragma unroll
    for (int k = 0; k < NE; k++) {
      ps[k] = a.psi[(long)pk[k] * a.ld + bk[k]];
      vp[k] = a.V[pk[k]];
    }
#pragma unroll
    for (int k = 0; k < NE; k++) own[k] = -1;
    if (a.own && a.Pc) {
#pragma unroll
      for (int k = 0; k < NE; k++) own[k] = a.own[pk[k]];
    }
#pragma unroll
    for (int k = 0; k < NE; k++) osp[k] = 0;
    if (a.Pc) {
#pragma unroll
      for (int k = 0; k < NE; k++) osp[k] = a.own_sph[pk[k]];
    }
#pragma unroll
    for (int k = 0; k < NE; k++) {
      pre[k] = vk[k] ? mk(vp[k] * ps[k].x, vp[k] * ps[k].y) : mk(0.0, 0.0);
      if (!vk[k]) own[k] = -1;
    }
  }
  if constexpr (!GS) {
    if (a.Pc) vnl_expand_wave<NE, Z_TB>(a, gb0, own, osp, pre);
  }

  auto gl = [&](int i, int c) -> cplx {
    const int b = c / Z_TY, y = c % Z_TY;
    return (y0 + y < L && cb0 + b < a.nb) ? a.W[(((long)(cb0 + b) * L + i) * L + x) * L + y0 + y]
                                          : mk(0.0, 0.0);
  };
  auto ll = [&](int i, int c) -> cplx { return lds[i * Z_COLS + zswz(c)]; };
  auto ls = [&](int i, int c, cplx v, int, int) { lds[i * Z_COLS + zswz(c)] = v; };
  if constexpr (GS) {
    // the first sweep transformed z already: this sweep only transposes
    for (int e = threadIdx.x; e < L * Z_COLS; e += NT) {
      const int i = e / Z_COLS, c = e % Z_COLS;
      lds[i * Z_COLS + zswz(c)] = gl(i, c);
    }
    __syncthreads();
  } else {
    // the W rows and twiddles loaded into LDS at the start: this wave's
    // loads are complete (vmcnt 0), the barrier makes every wave's visible
    __builtin_amdgcn_s_waitcnt(WAITCNT_VM0);
    __syncthreads();
    // (twiddles read at their products: pre[] is held through this transform)
    run_plan<P, R0, R1, R2, Z_COLS, NT, false>(ltw, +1, true, ll, ll, ls, ls, true);
  }

  // projector_all::apply (projector_all.hpp:225-231): V_nl psi of the owned
  // points, its ownership entries summed in fixed (sphere, point) order.
  // The first entries of all this thread's outputs are loaded back to back
  // (one memory round trip rather than one per owned output); further
  // entries (points in overlapping spheres) follow.
  if constexpr (!GS) {
    if (!a.Pc && a.own) {
      // the ownership words again (not held through the transform)
#pragma unroll
      for (int k = 0; k < NE; k++) {
        const int e = threadIdx.x + k * NT;
        const int b = e % Z_TB, y = (e / Z_TB) % Z_TY, z = e / Z_COLS;
        own[k] = e < L * Z_COLS && y0 + y < L && cb0 + b < a.nb
                     ? a.own[((long)x * L + y0 + y) * L + z]
                     : -1;
      }
      cplx q[NE];
#pragma unroll
      for (int k = 0; k < NE; k++) {
        q[k] = mk(0.0, 0.0);
        if (own[k] >= 0 && (own[k] & OWN_MAX_COUNT) > 0) {
          const int cb = gb0 + (int)(threadIdx.x + k * NT) % Z_TB;
          q[k] = a.Qe[(long)(own[k] >> OWN_COUNT_BITS) * a.ldQ + cb];
        }
      }
#pragma unroll
      for (int k = 0; k < NE; k++) {
        if (own[k] >= 0) {
          const int cb = gb0 + (int)(threadIdx.x + k * NT) % Z_TB;
          const int e0 = own[k] >> OWN_COUNT_BITS, cnt = own[k] & OWN_MAX_COUNT;
          if (cnt > 0) pre[k] = cadd(pre[k], q[k]);
          for (int j = 1; j < cnt; j++) pre[k] = cadd(pre[k], a.Qe[(long)(e0 + j) * a.ldQ + cb]);
        }
      }
    }
  }

  // Output k of this thread: element e = thread + k NT of the tile, band b,
  // row y, plane z; valid inside the grid and the chunk.
  auto result = [&](int k, int b, int y, int z) {
    cplx v = cscale(a.scale, lds[z * Z_COLS + zswz(b * Z_TY + y)]);
    v.x += pre[k].x;
    v.y += pre[k].y;
    return v;
  };
  if constexpr (!ACC) {
#pragma unroll
    for (int k = 0; k < NE; k++) {
      const int e = threadIdx.x + k * NT;
      const int b = e % Z_TB, y = (e / Z_TB) % Z_TY, z = e / Z_COLS;
      if (e < L * Z_COLS && y0 + y < L && cb0 + b < a.nb) {
        const cplx v = result(k, b, y, z);
        if (!a.skip_store) a.hpsi[(((long)x * L + y0 + y) * L + z) * a.hld + gb0 + b] = v;
        if constexpr (FWD) lds[z * Z_COLS + zswz(b * Z_TY + y)] = v;
      }
    }
  } else {
    // The Taylor accumulators of G outputs are all loaded (an invalid output
    // loads point 0, band gb0 and stores nothing), then the outputs and the
    // updated accumulators are formed, then stored, each kind of store in a
    // loop of its own: no store waits for a load behind it.
    constexpr int G = NE <= 6 ? NE : (R0 >= 12 ? 2 : (NE + 1) / 2);
    const cplx *s1 = a.acc1_init ? a.psi : a.acc1;
    const long l1 = a.acc1_init ? a.ld : a.acc1_ld;
#pragma unroll
    for (int k0 = 0; k0 < NE; k0 += G) {
      constexpr int GG = G;
      cplx o1[GG], o2[GG], v[GG];
      long pp[GG];
      int cbk[GG];
      bool ok[GG];
#pragma unroll
      for (int g = 0; g < GG; g++) {
        const int e = threadIdx.x + (k0 + g) * NT;
        const int b = e % Z_TB, y = (e / Z_TB) % Z_TY, z = e / Z_COLS;
        ok[g] = k0 + g < NE && e < L * Z_COLS && y0 + y < L && cb0 + b < a.nb;
        pp[g] = ok[g] ? ((long)x * L + y0 + y) * L + z : 0;
        cbk[g] = gb0 + (ok[g] ? b : 0);
      }
      if (a.acc1) {
#pragma unroll
        for (int g = 0; g < GG; g++) o1[g] = s1[pp[g] * l1 + cbk[g]];
      }
      if (a.acc2) {
#pragma unroll
        for (int g = 0; g < GG; g++) o2[g] = a.acc2[pp[g] * a.acc2_ld + cbk[g]];
      }
#pragma unroll
      for (int g = 0; g < GG; g++) {
        const int e = threadIdx.x + (k0 + g) * NT;
        const int b = e % Z_TB, y = (e / Z_TB) % Z_TY, z = e / Z_COLS;
        v[g] = ok[g] ? result(k0 + g, b, y, z) : mk(0.0, 0.0);
      }
      if (a.acc1) {
#pragma unroll
        for (int g = 0; g < GG; g++) {
          o1[g].x += a.c1re * v[g].x - a.c1im * v[g].y;
          o1[g].y += a.c1re * v[g].y + a.c1im * v[g].x;
        }
      }
      if (a.acc2) {
#pragma unroll
        for (int g = 0; g < GG; g++) {
          o2[g].x += a.c2re * v[g].x - a.c2im * v[g].y;
          o2[g].y += a.c2re * v[g].y + a.c2im * v[g].x;
        }
      }
      if (!a.skip_store) {
#pragma unroll
        for (int g = 0; g < GG; g++)
          if (ok[g]) a.hpsi[pp[g] * a.hld + cbk[g]] = v[g];
      }
      if constexpr (FWD) {
#pragma unroll
        for (int g = 0; g < GG; g++) {
          const int e = threadIdx.x + (k0 + g) * NT;
          const int b = e % Z_TB, y = (e / Z_TB) % Z_TY, z = e / Z_COLS;
          if (ok[g]) lds[z * Z_COLS + zswz(b * Z_TY + y)] = v[g];
        }
      }
      if (a.acc1) {
#pragma unroll
        for (int g = 0; g < GG; g++)
          if (ok[g]) a.acc1[pp[g] * a.acc1_ld + cbk[g]] = o1[g];
      }
      if (a.acc2) {
#pragma unroll
        for (int g = 0; g < GG; g++)
          if (ok[g]) a.acc2[pp[g] * a.acc2_ld + cbk[g]] = o2[g];
      }
    }
  }
  if constexpr (FWD) {
    // each thread rewrote only the slots it read; the forward plan mixes them
    __syncthreads();
    auto gs = [&](int i, int c, cplx v, int, int) {
      const int b = c / Z_TY, y = c % Z_TY;
      if (y0 + y < L && cb0 + b < a.nb)
        a.W[(((long)(cb0 + b) * L + i) * L + x) * L + y0 + y] = v;
    };
    run_plan<P, R0, R1, R2, Z_COLS, NT>(ltw, -1, true, ll, ll, ls, gs, false);
  }
}

// Tile of workgroup b of T when the nbt band tiles of one (x, y rows) pair
// (consecutive in the natural order) should run on one XCD, whose L2 then
// serves their shared V, ownership and sphere words once: workgroups go to
// the 8 XCDs in turn, so each group of 8 pairs x nbt band tiles is dealt
// with pair p's band tiles on workgroups b, b + 8, b + 16, ...  The tail
// beyond whole groups keeps the natural order.
__device__ __forceinline__ long xcd_tile(long b, long T, int nbt) {
  constexpr int G = 8;
  const long full = T / ((long)G * nbt) * G * nbt;
  if (b >= full) return b;
  const long q = b / (G * nbt), r = b % (G * nbt);
  return (q * G + r % G) * nbt + r / G;
}

// Dynamic LDS of a z_bwd_epi tile: the tile and the twiddle table
constexpr size_t z_bwd_lds(int L, int zc = Z_COLS) { return ((size_t)L * zc + L) * sizeof(cplx); }


// Occupancy floor of the z_bwd_epi kernels (amdgpu_waves_per_eu): at least
// 4 waves per SIMD (two 512-thread z tiles per CU, or one of 1024) where the
// tile's registers fit in 128 without spilling; the 72-long tile needs more
// and keeps 3.
constexpr int z_bwd_min_waves(int L) { return L == 72 ? 1 : 4; }

// Threads of a z tile: twice Z_NT on the long grids, whose 96 KB+ of LDS
// leaves room for one workgroup per CU; 768 at 72, where the 512-thread
// z_bwd_epi tile took 146 registers (two radix-8 butterflies and nine
// outputs per thread) and so held a CU alone with 8 waves, while at 768 it
// takes 124 and 12 waves share the CU (z_bwd_epi 0.58 vs 0.66 ms at
// 72^3 x 128).
constexpr int z_threads(int L) { return L >= 80 ? 1024 : (L == 72 ? 768 : Z_NT); }
constexpr int z_bwd_waves(int L) { return L == 72 ? 4 : z_bwd_min_waves(L); }

// Columns and threads of the one-system z_bwd_epi tiles: 32 columns of 512
// threads for the plain Hpsi at 96 (48 KB of LDS: two tiles share a CU,
// where a 64-column tile of 96 KB held a CU alone: z_bwd_epi 6.30 -> 5.30 ms
// per 96^3 x 512 Hpsi), else 64 columns of z_threads(L).  The Taylor forms
// (acc) keep 64 columns: with 32 the 96^3 ETRS step ran 103 -> 113 ms (the
// form that transforms its output forward writes W in 64-byte runs of 4 y)
// and 103.4 -> 104.3 ms with only the last order's form at 32.
constexpr int z_bwd_cols(int L, bool acc) { return L == 96 && !acc ? 32 : Z_COLS; }
constexpr int z_bwd_threads(int L, bool acc) { return L == 96 && !acc ? Z_NT : z_threads(L); }

// Grid lengths with three-sweep kernels: (L, radices of the plan)
#define INQAMD_FFT3_TABLE(X) \
  X(24, 8, 3, 1)             \
  X(30, 6, 5, 1)             \
  X(32, 8, 4, 1)             \
  X(36, 6, 6, 1)             \
  X(40, 8, 5, 1)             \
  X(48, 8, 6, 1)             \
  X(60, 12, 5, 1)            \
  X(64, 8, 8, 1)             \
  X(72, 8, 9, 1)             \
  X(80, 8, 2, 5)             \
  X(96, 8, 4, 3)

// Work items of a band chunk of nb bands (tiles of ZC columns)
template <int Z_TY, int ZC = Z_COLS>
inline long z_tiles(int L, int nb) {
  constexpr int Z_TB = ZC / Z_TY;
  return (long)L * ((L + Z_TY - 1) / Z_TY) * ((nb + Z_TB - 1) / Z_TB);
}

}  // namespace inqamd
