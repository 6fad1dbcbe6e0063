// pt_solver.hip -- stream-function multigrid on gfx950.
//
// Kernels (one HIP kernel per reference loop nest or fused group):
//   k_stage<R>      fused stream_derivatives + stream_space + RKC1 stage update
//                   + enforceBCs_stream (ghost write-through), LDS tile with a
//                   3-wide halo       (PS/iterative_solver.f90:973-1018, 890-911,
//                                      1561-1597, 1616-1664)
//   k_residual<R>   stream_derivatives + compute_residual with per-block |r| sums
//                                      (PS/iterative_solver.f90:415-448)
//   k_restrict      full weighting     (PS/iterative_solver.f90:1107-1145)
//   k_prolong       bilinear prolongation + BCs (PS/iterative_solver.f90:343-396)
//   k_rhs1          compute_RHS level 1 (PS/iterative_solver.f90:1075-1101)
//   k_visf / k_vis_* viscosity hierarchy (PS/rheology.f90:2-23, PS/smoother.f90:454-580,
//                   PS/iterative_solver.f90:472-690, 1147-1211)
//   k_coarse_*      banded coarse operator, DGBEQU row scaling, DGBTF2 LU on one
//                   wavefront, and the solve (DGBTRS) as a dense interior operator
//                   built once per factorization (PS/iterative_solver.f90:1701-1887),
//                   applied by k_coarse_apply (FP64) or k_coarse_apply_wg (mixed)
//   k_smooth_small, k_smooth_cb, k_strip_cb
//                   whole RKC1 sweeps of coarse levels in LDS (one workgroup, or
//                   column strips with several fused stages per launch)
//   k_velocities    compute_velocities + velocityBCs (PS/momentum.f90:655-710,
//                   PS/composition.f90:1295-1325)
// FP64 kernels are compiled without FMA contraction so each point is
// evaluated in the reference's operation order.
#include <cmath>
#include <cstring>
#include <algorithm>
#include <vector>
#include "pt_internal.hpp"

namespace pt {

#define IDX(i, k, pitch) ((long)((k) + 3) * (pitch) + ((i) + 3))

// ---------------------------------------------------------------------------
// stencil evaluation of stream_space on an LDS tile
// ---------------------------------------------------------------------------
struct Coef {
  const void *vis, *vxm, *vxz, *vx, *vz, *rhs, *dts;
};

template <typename R>
struct StageArgs {
  const R* Sx;      // stencil input
  const R* S0;      // pointwise
  const R* S2;      // pointwise
  R* Y;             // output (+ mirrored ghosts)
  const R *vis, *vxm, *vxz, *vx, *vz, *rhs, *dts;
  R c0, c1, c2, cf; // Y = ((c0*S0 + c1*S1) + c2*S2) + ((cf*dt)*F)   (stage >= 2)
  int stage1;       // Y = S0 + (cf*dt)*F
  int s0zero;       // S0 == 0 (mixed-precision correction from zero)
  int ivis0;
  int n, m, pitch;
  R sz;             // ghost sign in z (free-slip -1, rigid +1)
  R dr2, dr3, dr4, ds2, ds3, ds4, drds, dr2ds, drds2, dr2ds2;
  R rdr2, rdr3, rdr4, rds2, rds3, rds4, rdrds, rdr2ds, rdrds2, rdr2ds2;
  R D1[5], D2[5], D3[7], D4[7];
  R Drs[5][5], Drss[5][5], Drrs[5][5], Drrss[5][5];
};

// Accumulation in the reference's order: separate multiply and add for the
// FP64 path (bit-faithful); fused multiply-add for the FP32 smoother.
__device__ __forceinline__ double acc_(double s, double c, double v) { return s + c * v; }
__device__ __forceinline__ float acc_(float s, float c, float v) { return __fmaf_rn(c, v, s); }
// derivative scaling: FP64 divides by the grid power as the reference does;
// FP32 multiplies by the precomputed reciprocal
__device__ __forceinline__ double scale_(double s, double d, double rd) { (void)rd; return s / d; }
__device__ __forceinline__ float scale_(float s, float d, float rd) { (void)d; return s * rd; }

constexpr int TX = 64, TZ = 8, PZ = 1;
constexpr size_t SMALL_LDS_BYTES = 160 * 1024;
constexpr int TW = TX + 6, TH = TZ * PZ + 6;
constexpr int TILE_PER = (TH * TW + TX * TZ - 1) / (TX * TZ);   // tile elements per thread

// pointwise operands of stream_space at one point (read once per launch)
template <typename R>
struct PointCoef {
  R rhs, vis, vxm, vxz, vx, vz;
};

template <typename R, bool IV0>
__device__ __forceinline__ PointCoef<R> load_pc(const StageArgs<R>& a, long g) {
  PointCoef<R> c;
  c.rhs = a.rhs[g];
  c.vis = a.vis[g];
  if constexpr (!IV0) {
    c.vxm = a.vxm[g]; c.vxz = a.vxz[g]; c.vx = a.vx[g]; c.vz = a.vz[g];
  } else {
    c.vxm = c.vxz = c.vx = c.vz = (R)0;
  }
  return c;
}

// stream_derivatives (uniform grid) + stream_space at one point, reference
// summation order per derivative, zero stencil weights skipped.  `at(b, q)`
// reads psi(i+q, k+b).  With IV0 (isoviscous, eta_x = eta_z = eta_xz =
// eta_xx-eta_zz = 0) only the three derivatives stream_space uses are formed;
// each is accumulated exactly as in the general case.
template <typename R, bool IV0, typename At>
__device__ __forceinline__ R stream_space_pt(At at, const StageArgs<R>& a, const PointCoef<R>& pc) {
  R sxx = 0, szz = 0, sxxx = 0, szzz = 0, sxxxx = 0, szzzz = 0;
  if constexpr (!IV0) {
#pragma unroll
    for (int q = -2; q <= 2; q++) sxx = acc_(sxx, a.D2[q + 2], at(0, q));
#pragma unroll
    for (int q = -3; q <= 3; q++) if (q != 0) sxxx = acc_(sxxx, a.D3[q + 3], at(0, q));
  }
#pragma unroll
  for (int q = -3; q <= 3; q++) sxxxx = acc_(sxxxx, a.D4[q + 3], at(0, q));
  if constexpr (!IV0) {
#pragma unroll
    for (int q = -2; q <= 2; q++) szz = acc_(szz, a.D2[q + 2], at(q, 0));
#pragma unroll
    for (int q = -3; q <= 3; q++) if (q != 0) szzz = acc_(szzz, a.D3[q + 3], at(q, 0));
  }
#pragma unroll
  for (int q = -3; q <= 3; q++) szzzz = acc_(szzzz, a.D4[q + 3], at(q, 0));
  R sxz = 0, sxxz = 0, sxzz = 0, sxxzz = 0;
#pragma unroll
  for (int b = -2; b <= 2; b++)
#pragma unroll
    for (int q = -2; q <= 2; q++) {
      R v = at(b, q);
      if constexpr (!IV0) {
        if (b != 0 && q != 0) sxz = acc_(sxz, a.Drs[b + 2][q + 2], v);
        if (b != 0) sxxz = acc_(sxxz, a.Drrs[b + 2][q + 2], v);
        if (q != 0) sxzz = acc_(sxzz, a.Drss[b + 2][q + 2], v);
      }
      sxxzz = acc_(sxxzz, a.Drrss[b + 2][q + 2], v);
    }
  R SFxxxx = scale_(sxxxx, a.dr4, a.rdr4), SFzzzz = scale_(szzzz, a.ds4, a.rds4);
  R SFxxzz = scale_(sxxzz, a.dr2ds2, a.rdr2ds2);
  R sv;
  if constexpr (IV0) {
    R t2 = pc.vis * (SFxxxx + SFzzzz + (R)2 * SFxxzz);
    sv = t2 - pc.rhs;
  } else {
    R SFxx = scale_(sxx, a.dr2, a.rdr2), SFxxx = scale_(sxxx, a.dr3, a.rdr3);
    R SFzz = scale_(szz, a.ds2, a.rds2), SFzzz = scale_(szzz, a.ds3, a.rds3);
    R SFxz = scale_(sxz, a.drds, a.rdrds), SFxxz = scale_(sxxz, a.dr2ds, a.rdr2ds);
    R SFxzz = scale_(sxzz, a.drds2, a.rdrds2);
    R t1 = pc.vxm * (SFxx - SFzz);
    R t2 = pc.vis * (SFxxxx + SFzzzz + (R)2 * SFxxzz);
    R t4 = ((R)4 * pc.vxz) * SFxz;
    R t5 = ((R)2 * pc.vx) * (SFxxx + SFxzz);
    R t6 = ((R)2 * pc.vz) * (SFzzz + SFxxz);
    sv = t1 + t2 + t4 + t5 + t6 - pc.rhs;
  }
  return -sv;
}

template <typename R>
struct TileRegs {
  R v[TILE_PER];
  bool in[TILE_PER];
};
// the thread's tile elements, loaded in two phases so that every global
// load of the launch (tile and pointwise operands) is in flight at once:
// tile_issue loads from addresses clamped to the stored range (so the loads
// are unconditional), tile_commit writes LDS, zero outside the level
template <typename R>
__device__ __forceinline__ TileRegs<R> tile_issue(const R* S, int i0, int k0, int n, int m, int pitch) {
  TileRegs<R> r;
  const int tid = threadIdx.y * TX + threadIdx.x;
#pragma unroll
  for (int u = 0; u < TILE_PER; u++) {
    const int idx = min(tid + u * TX * TZ, TH * TW - 1);
    const int rr = idx / TW, cc = idx - rr * TW;
    const int gi = i0 - 3 + cc, gk = k0 - 3 + rr;
    r.in[u] = gi <= n + 3 && gk <= m + 3;
    r.v[u] = S[IDX(min(gi, n + 3), min(gk, m + 3), pitch)];
  }
  return r;
}
template <typename R>
__device__ __forceinline__ void tile_commit(R (*t)[TW], const TileRegs<R>& r) {
  const int tid = threadIdx.y * TX + threadIdx.x;
#pragma unroll
  for (int u = 0; u < TILE_PER; u++) {
    const int idx = tid + u * TX * TZ;
    if (idx < TH * TW) {
      const int rr = idx / TW, cc = idx - rr * TW;
      t[rr][cc] = r.in[u] ? r.v[u] : (R)0;
    }
  }
}

// store Y(i,k) and its ghost images (enforceBCs_stream: antisymmetric side
// walls, sz-symmetric top/bottom, psi = 0 on the walls which are never written)
template <typename R>
__device__ __forceinline__ void store_with_ghosts(R* Y, int i, int k, int n, int m, int pitch, R y, R sz) {
  Y[IDX(i, k, pitch)] = y;
  bool lo_i = i <= 3, hi_i = i >= n - 3, lo_k = k <= 3, hi_k = k >= m - 3;
  if (lo_i) Y[IDX(-i, k, pitch)] = -y;
  if (hi_i) Y[IDX(2 * n - i, k, pitch)] = -y;
  if (lo_k) {
    Y[IDX(i, -k, pitch)] = sz * y;
    if (lo_i) Y[IDX(-i, -k, pitch)] = -(sz * y);
    if (hi_i) Y[IDX(2 * n - i, -k, pitch)] = -(sz * y);
  }
  if (hi_k) {
    Y[IDX(i, 2 * m - k, pitch)] = sz * y;
    if (lo_i) Y[IDX(-i, 2 * m - k, pitch)] = -(sz * y);
    if (hi_i) Y[IDX(2 * n - i, 2 * m - k, pitch)] = -(sz * y);
  }
}

// one RKC1 stage on a TX x (TZ*PZ) tile; the pointwise operands are loaded
// before the tile barrier so their latency overlaps the tile fill
template <typename R, bool IV0>
__global__ void __launch_bounds__(TX* TZ) k_stage(StageArgs<R> a) {
  __shared__ R tile[TH][TW];
  const int i0 = 1 + blockIdx.x * TX, k0 = 1 + blockIdx.y * TZ * PZ;
  const TileRegs<R> tr = tile_issue<R>(a.Sx, i0, k0, a.n, a.m, a.pitch);
  R dt[PZ], s0[PZ], s2[PZ];
  PointCoef<R> pc[PZ];
  const int i = i0 + threadIdx.x;
#pragma unroll
  for (int pz = 0; pz < PZ; pz++) {
    int k = k0 + threadIdx.y + pz * TZ;
    dt[pz] = s0[pz] = s2[pz] = (R)0;
    if (i > a.n - 1 || k > a.m - 1) continue;
    long g = IDX(i, k, a.pitch);
    dt[pz] = a.dts[g];
    if (!a.s0zero) s0[pz] = a.S0[g];
    if (!a.stage1) s2[pz] = a.S2[g];
    pc[pz] = load_pc<R, IV0>(a, g);
  }
  tile_commit<R>(tile, tr);
  __syncthreads();
#pragma unroll
  for (int pz = 0; pz < PZ; pz++) {
    int k = k0 + threadIdx.y + pz * TZ;
    if (i > a.n - 1 || k > a.m - 1) continue;
    int ti = threadIdx.x + 3, tk = threadIdx.y + pz * TZ + 3;
    R F = stream_space_pt<R, IV0>([&](int b, int q) { return tile[tk + b][ti + q]; }, a, pc[pz]);
    R y;
    if (a.stage1) {
      y = s0[pz] + (a.cf * dt[pz]) * F;
    } else {
      R s1 = tile[tk][ti];
      y = ((a.c0 * s0[pz] + a.c1 * s1) + a.c2 * s2[pz]) + ((a.cf * dt[pz]) * F);
    }
    store_with_ghosts<R>(a.Y, i, k, a.n, a.m, a.pitch, y, a.sz);
  }
}

// residual(i,k) = stream_space(i,k,S); block partial sums of |residual|
template <typename R, bool IV0>
__global__ void __launch_bounds__(TX* TZ) k_residual(StageArgs<R> a, R* res, double* partial) {
  __shared__ R tile[TH][TW];
  __shared__ double red[TX * TZ];
  const int i0 = 1 + blockIdx.x * TX, k0 = 1 + blockIdx.y * TZ * PZ;
  const TileRegs<R> tr = tile_issue<R>(a.Sx, i0, k0, a.n, a.m, a.pitch);
  PointCoef<R> pc[PZ];
  const int i = i0 + threadIdx.x;
#pragma unroll
  for (int pz = 0; pz < PZ; pz++) {
    int k = k0 + threadIdx.y + pz * TZ;
    if (i > a.n - 1 || k > a.m - 1) continue;
    pc[pz] = load_pc<R, IV0>(a, IDX(i, k, a.pitch));
  }
  tile_commit<R>(tile, tr);
  __syncthreads();
  double acc = 0.0;
#pragma unroll
  for (int pz = 0; pz < PZ; pz++) {
    int k = k0 + threadIdx.y + pz * TZ;
    if (i > a.n - 1 || k > a.m - 1) continue;
    int ti = threadIdx.x + 3, tk = threadIdx.y + pz * TZ + 3;
    R F = stream_space_pt<R, IV0>([&](int b, int q) { return tile[tk + b][ti + q]; }, a, pc[pz]);
    res[IDX(i, k, a.pitch)] = F;
    acc = acc + fabs((double)F);
  }
  int tid = threadIdx.y * TX + threadIdx.x;
  red[tid] = acc;
  __syncthreads();
  for (int s = TX * TZ / 2; s > 0; s >>= 1) {
    if (tid < s) red[tid] = red[tid] + red[tid + s];
    __syncthreads();
  }
  if (tid == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

// one RK1 smoothing iteration (RKC1=0, PS/iterative_solver.f90:1468-1485,
// :1503-1513) from F = stream_space(S0): psi = S0 + (c dt) F with the local
// step c = 2/(-lb-la+(lb-la) temp), la = -1, lb = -2 r^2/(1+r^2) from the
// point's eigenvalue ratio r (-2 where r is not below 1e99), c = 1 on the
// way up the cycle; then the ghost images (enforceBCs_stream).  Each thread
// reads and writes only its own interior point, so the update is in place.
__global__ void k_rk1_update(double* psi, const double* F, const double* dts, const double* ratio, int n, int m,
                             int pitch, double temp, int dir, double sz) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x, k = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i > n - 1 || k > m - 1) return;
  const long g = IDX(i, k, pitch);
  double cc = 1.0;
  if (dir == 0) {
    const double la = -1.0, lb_user = -2.0, r = ratio[g];
    double lb = lb_user;
    if (r < 1.0e99) lb = lb_user * (r * r) / (1.0 + r * r);
    cc = 2.0 / (-lb - la + (lb - la) * temp);
  }
  const double y = psi[g] + (cc * dts[g]) * F[g];
  store_with_ghosts<double>(psi, i, k, n, m, pitch, y, sz);
}

// ---------------------------------------------------------------------------
// small levels: the whole RKC1 sweep (all stages) and compute_residual in one
// workgroup with the three RKC buffers resident in LDS.  Each thread owns up
// to PPT interior points; their indices and pointwise operands are fixed for
// the sweep and kept in registers.
// ---------------------------------------------------------------------------
struct RkcCoef {
  double mewt[64], mew[64], vv[64], cc[64];
};
// the RKC1 stage count and coefficients of a level's sweeps: the reference's
// nstageS everywhere, or s_coarse stages on the mixed path's levels >= 2
struct LevelRkc {
  int s;
  const double *mewt, *mew, *vv, *cc;
  double stability;
};
static LevelRkc level_rkc(const Context& c, int l, bool fp32) {
  const Tables& t = c.T;
  if (fp32 && l >= 2 && t.s_coarse >= 2) return {t.s_coarse, t.mewt_c, t.mew_c, t.vv_c, t.cc_c, t.RKC1_stability_c};
  return {c.P.nstageS, t.mewt, t.mew, t.vv, t.cc, t.RKC1_stability};
}
static RkcCoef rkc_coef(const LevelRkc& rk) {
  RkcCoef co;
  std::memcpy(co.mewt, rk.mewt, sizeof co.mewt); std::memcpy(co.mew, rk.mew, sizeof co.mew);
  std::memcpy(co.vv, rk.vv, sizeof co.vv); std::memcpy(co.cc, rk.cc, sizeof co.cc);
  return co;
}
// the FP32 sweep's stage coefficients, rounded once on the host: c0 = cc,
// c1 = mew, c2 = vv, cf = mewt * cpar (the product in FP64, as the device
// would form it); read at a uniform stage index they are scalar loads of
// the kernel arguments
struct RkcCoefF {
  float c0[64], c1[64], c2[64], cf[64];
};
static RkcCoefF rkc_f32(const RkcCoef& co, double cpar) {
  RkcCoefF f;
  for (int j = 0; j < 64; j++) {
    f.c0[j] = (float)co.cc[j];
    f.c1[j] = (float)co.mew[j];
    f.c2[j] = (float)co.vv[j];
    f.cf[j] = (float)(co.mewt[j] * cpar);
  }
  return f;
}
constexpr int SMALL_THREADS = 1024;

// RKC1 coefficients of stages 0..63 held one stage per lane of every
// wavefront, read at a (uniform) stage index with v_readlane: no memory
// access in the stage loop
struct LaneRkc {
  double cc, mew, vv, mewt;
};
__device__ __forceinline__ LaneRkc lane_rkc(const RkcCoef& co) {
  const int j = threadIdx.x & 63;
  return LaneRkc{co.cc[j], co.mew[j], co.vv[j], co.mewt[j]};
}
__device__ __forceinline__ double lane_at(double v, int j) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), j);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), j);
  return __hiloint2double(hi, lo);
}

// a level array into LDS (and two more LDS arrays zeroed) by a whole
// workgroup of NT threads, B elements per thread per batch: the batch's
// loads (indices clamped into the array) are all issued before its first
// LDS write
template <typename R, int NT, int B>
__device__ __forceinline__ void copy_in3(const R* src, R* d0, R* d1, R* d2, int ne) {
  for (int e0 = threadIdx.x; e0 < ne; e0 += B * NT) {
    R v[B];
#pragma unroll
    for (int u = 0; u < B; u++) v[u] = src[min(e0 + u * NT, ne - 1)];
#pragma unroll
    for (int u = 0; u < B; u++) {
      const int e = e0 + u * NT;
      if (e < ne) {
        d0[e] = v[u];
        d1[e] = (R)0;
        d2[e] = (R)0;
      }
    }
  }
}

template <typename R, bool IV0, int PPT>
__device__ __forceinline__ void sweep_small(const StageArgs<R>& a, R* err, R* res, int nstage, double cpar,
                                            const RkcCoef& co, R* buf0) {
  const int n = a.n, m = a.m, p = a.pitch;
  const int ne = p * (m + 7);
  R* buf1 = buf0 + ne;
  R* buf2 = buf1 + ne;
  const int tid = threadIdx.x;
  copy_in3<R, SMALL_THREADS, 8>(err, buf0, buf1, buf2, ne);
  // point e of the thread's PPT: row e / tcol, column e % tcol, with rows
  // of tcol points padded to whole wavefronts when the workgroup has room,
  // so that no wave spans two rows (lanes a row lower hit other lanes' LDS
  // banks)
  const int nxi = n - 1;
  const int padded = (nxi + 63) & ~63;
  const int tcol = padded * (m - 1) <= PPT * SMALL_THREADS ? padded : nxi;
  const int npts = tcol * (m - 1);
  int pi[PPT], pk[PPT], pg[PPT];
  R pdt[PPT];
  PointCoef<R> pc[PPT];
#pragma unroll
  for (int q = 0; q < PPT; q++) {
    int e = tid + q * SMALL_THREADS;
    pi[q] = 0;
    pk[q] = 0;
    pg[q] = 0;
    pdt[q] = (R)0;
    const int k = e / tcol, ic = e - k * tcol;
    if (e < npts && ic < nxi) {
      pi[q] = 1 + ic;
      pk[q] = 1 + k;
      pg[q] = (int)IDX(pi[q], pk[q], p);
      pdt[q] = a.dts[pg[q]];
      pc[q] = load_pc<R, IV0>(a, pg[q]);
    }
  }
  const LaneRkc lk = lane_rkc(co);
  __syncthreads();
  // stage 1: S1 = S0 + mewt1*c*dt*F(S0) -> buf1
  {
    const R cf = (R)(lane_at(lk.mewt, 1) * cpar);
#pragma unroll
    for (int q = 0; q < PPT; q++) {
      if (pi[q] == 0) continue;
      const int g = pg[q];
      R F = stream_space_pt<R, IV0>([&](int b, int c) { return buf0[g + b * p + c]; }, a, pc[q]);
      R y = buf0[g] + (cf * pdt[q]) * F;
      store_with_ghosts<R>(buf1, pi[q], pk[q], n, m, p, y, a.sz);
    }
  }
  __syncthreads();
  R* cur = buf1;
  R* prev = buf0;   // S2 = S0 at stage 2
  R* spare = buf2;
  for (int j = 2; j <= nstage; j++) {
    R* out = (prev == buf0) ? spare : prev;
    const R c0 = (R)lane_at(lk.cc, j), c1 = (R)lane_at(lk.mew, j), c2 = (R)lane_at(lk.vv, j);
    const R cf = (R)(lane_at(lk.mewt, j) * cpar);
#pragma unroll
    for (int q = 0; q < PPT; q++) {
      if (pi[q] == 0) continue;
      const int g = pg[q];
      R F = stream_space_pt<R, IV0>([&](int b, int c) { return cur[g + b * p + c]; }, a, pc[q]);
      R y = ((c0 * buf0[g] + c1 * cur[g]) + c2 * prev[g]) + ((cf * pdt[q]) * F);
      store_with_ghosts<R>(out, pi[q], pk[q], n, m, p, y, a.sz);
    }
    __syncthreads();
    R* np = cur;
    cur = out;
    prev = np;
    if (out == spare) spare = nullptr;
  }
  // cur holds the sweep result (psi correction): write back, then the residual
  for (int e = tid; e < ne; e += SMALL_THREADS) err[e] = cur[e];
#pragma unroll
  for (int q = 0; q < PPT; q++) {
    if (pi[q] == 0) continue;
    const int g = pg[q];
    res[g] = stream_space_pt<R, IV0>([&](int b, int c) { return cur[g + b * p + c]; }, a, pc[q]);
  }
}

template <typename R, bool IV0, int PPT>
__global__ void __launch_bounds__(SMALL_THREADS) k_smooth_small(StageArgs<R> a, R* err, R* res, int nstage,
                                                               double cpar, RkcCoef co) {
  extern __shared__ unsigned char smem_raw[];
  sweep_small<R, IV0, PPT>(a, err, res, nstage, cpar, co, reinterpret_cast<R*>(smem_raw));
}

// F = RHS - L psi (isoviscous, FP32) at the RB vertically adjacent points
// (i, k0..k0+RB-1) of an LDS array with row pitch W whose column 0 is global
// column `lo` and whose row 0 is global row -3.  The x-D2 sums of the RB+4
// rows the 5x5 neighbourhood spans are formed once and reused by every
// point's z-D2 sum (psi_xxzz = D2_z(D2_x psi), the product weights D2(b)D2(q)
// factorised); each LDS value is read once.  Rows beyond rmax (points past
// the last interior row of a partial block) are clamped; their F is unused.
template <int RB>
__device__ __forceinline__ void eval_cb(const float* S, int W, int lo, int i, int k0, int rmax,
                                        const StageArgs<float>& a, const float* rhsv, const float* visv, float* F) {
  auto at = [&](int ii, int kk) { return S[(kk + 3) * W + (ii - lo)]; };
  float v[5][RB + 4];
#pragma unroll
  for (int w = 0; w < RB + 4; w++) {
    const int row = min(k0 - 2 + w, rmax);
#pragma unroll
    for (int c = 0; c < 5; c++) v[c][w] = at(i - 2 + c, row);
  }
  float colv[RB + 6];
  colv[0] = at(i, k0 - 3);
#pragma unroll
  for (int w = 0; w < RB + 4; w++) colv[w + 1] = v[2][w];
  colv[RB + 5] = at(i, min(k0 + RB + 2, rmax));
  float t2[RB + 4];
#pragma unroll
  for (int w = 0; w < RB + 4; w++) {
    float acc = 0.0f;
#pragma unroll
    for (int c = 0; c < 5; c++) acc = __fmaf_rn(a.D2[c], v[c][w], acc);
    t2[w] = acc;
  }
#pragma unroll
  for (int r = 0; r < RB; r++) {
    const int w = r + 2;
    const int row = min(k0 + r, rmax);
    float sxxxx = a.D4[0] * at(i - 3, row);
#pragma unroll
    for (int c = 0; c < 5; c++) sxxxx = __fmaf_rn(a.D4[c + 1], v[c][w], sxxxx);
    sxxxx = __fmaf_rn(a.D4[6], at(i + 3, row), sxxxx);
    float szzzz = 0.0f;
#pragma unroll
    for (int b = 0; b < 7; b++) szzzz = __fmaf_rn(a.D4[b], colv[r + b], szzzz);
    float sxxzz = 0.0f;
#pragma unroll
    for (int b = 0; b < 5; b++) sxxzz = __fmaf_rn(a.D2[b], t2[w - 2 + b], sxxzz);
    const float lap2 = sxxxx * a.rdr4 + szzzz * a.rds4 + 2.0f * (sxxzz * a.rdr2ds2);
    F[r] = rhsv[r] - visv[r] * lap2;
  }
}

// FP32 isoviscous single-workgroup sweep for the mixed-precision smoother,
// column-blocked: each thread owns RB vertically adjacent points of one
// column (eval_cb); S0 and the pointwise operands stay in registers.
// Factorised sums round differently from the reference's 5x5 product
// weights, so only the FP32 path uses this.
template <int RB>
__device__ __forceinline__ void sweep_cb(const StageArgs<float>& a, float* err, float* res, int nstage,
                                         const RkcCoefF& co, float* buf0) {
  const int n = a.n, m = a.m, p = a.pitch;
  const int ne = p * (m + 7);
  float* buf1 = buf0 + ne;
  float* buf2 = buf1 + ne;
  const int tid = threadIdx.x;
  // threads of a row block: the n-1 interior columns, padded to whole
  // wavefronts when the workgroup has room, so that no wave spans two row
  // blocks (lanes one block lower hit other lanes' LDS banks)
  const int nxi = n - 1, nkb = (m - 1 + RB - 1) / RB;
  const int padded = (nxi + 63) & ~63;
  const int tcol = padded * nkb <= SMALL_THREADS ? padded : nxi;
  const int kb = tid / tcol, ic = tid - kb * tcol;
  const bool active = kb < nkb && ic < nxi;
  int i = 1, k0 = 1;
  float s0v[RB], dtv[RB], rhsv[RB], visv[RB];
  bool on[RB];
  if (active) {
    i = 1 + ic;
    k0 = 1 + kb * RB;
  }
  // the pointwise operands from clamped (valid) points, zero where the
  // thread has no point; issued before the copy-in so that both are in
  // flight together
#pragma unroll
  for (int r = 0; r < RB; r++) {
    int k = k0 + r;
    on[r] = active && k <= m - 1;
    const long g = IDX(i, min(k, m - 1), p);
    const float e0 = err[g], d0 = a.dts[g], r0 = a.rhs[g], v0 = a.vis[g];
    s0v[r] = on[r] ? e0 : 0.0f;
    dtv[r] = on[r] ? d0 : 0.0f;
    rhsv[r] = on[r] ? r0 : 0.0f;
    visv[r] = on[r] ? v0 : 0.0f;
  }
  copy_in3<float, SMALL_THREADS, 8>(err, buf0, buf1, buf2, ne);
  const int rmax = m + 3;   // last stored row
  __syncthreads();
  {
    const float cf = co.cf[1];
    float F[RB];
    eval_cb<RB>(buf0, p, -3, i, k0, rmax, a, rhsv, visv, F);
#pragma unroll
    for (int r = 0; r < RB; r++)
      if (on[r]) store_with_ghosts<float>(buf1, i, k0 + r, n, m, p, s0v[r] + (cf * dtv[r]) * F[r], a.sz);
  }
  __syncthreads();
  float* cur = buf1;
  float* prev = buf0;   // S2 = S0 at stage 2
  float* spare = buf2;
  for (int j = 2; j <= nstage; j++) {
    float* out = (prev == buf0) ? spare : prev;
    const float c0 = co.c0[j], c1 = co.c1[j], c2 = co.c2[j], cf = co.cf[j];
    float F[RB];
    eval_cb<RB>(cur, p, -3, i, k0, rmax, a, rhsv, visv, F);
#pragma unroll
    for (int r = 0; r < RB; r++) {
      if (!on[r]) continue;
      long g = IDX(i, k0 + r, p);
      float y = ((c0 * s0v[r] + c1 * cur[g]) + c2 * prev[g]) + ((cf * dtv[r]) * F[r]);
      store_with_ghosts<float>(out, i, k0 + r, n, m, p, y, a.sz);
    }
    __syncthreads();
    float* np = cur;
    cur = out;
    prev = np;
    if (out == spare) spare = nullptr;
  }
  for (int e = tid; e < ne; e += SMALL_THREADS) err[e] = cur[e];
  float F[RB];
  eval_cb<RB>(cur, p, -3, i, k0, rmax, a, rhsv, visv, F);
#pragma unroll
  for (int r = 0; r < RB; r++)
    if (on[r]) res[IDX(i, k0 + r, p)] = F[r];
}

template <int RB>
__global__ void __launch_bounds__(SMALL_THREADS) k_smooth_cb(StageArgs<float> a, float* err, float* res, int nstage,
                                                            RkcCoefF co) {
  extern __shared__ unsigned char smem_raw[];
  sweep_cb<RB>(a, err, res, nstage, co, reinterpret_cast<float*>(smem_raw));
}

// ---------------------------------------------------------------------------
// FP32 isoviscous RKC1 sweep segments on column strips (temporal blocking),
// for levels too large for one workgroup.  Workgroup b owns the core columns
// [c0, c1) and keeps whole columns (rows -3..m+3) of the core widened by 3
// columns per fused stage in LDS.  Stage t of a segment is computed on the
// core widened by 3 columns per remaining stage, so after the segment the
// core holds its last stage with no exchange between workgroups (overlap
// columns are computed by both neighbours).  Per point the arithmetic is
// eval_cb's, so the result does not depend on the strip decomposition.
// Threads: x = strip column (64 lanes), y = 16 blocks of 4 rows (m <= 65).
// ---------------------------------------------------------------------------
constexpr int STRIP_TX = 64, STRIP_TY = 16, STRIP_RB = 4;

struct StripArgs {
  const float* S0;    // sweep start (pointwise; stencil source of stage 1)
  const float* Gcur;  // stage j0-1, read when j0 >= 3
  const float* Gprev; // stage j0-2, used when j0 >= 3 (an allocated level buffer: always loaded)
  float* Ocur;        // stage j1 on the core columns (+ ghost images)
  float* Oprev;       // stage j1-1 on the core columns; null on the last segment
  int j0, j1;         // stages of this segment (1-based, j1 > j0)
  int cw;             // core columns per workgroup
  int W;              // LDS strip width in columns
  RkcCoefF co;        // stage coefficients, cf with the sweep's Courant factor
};

__global__ void __launch_bounds__(STRIP_TX* STRIP_TY) k_strip_cb(StageArgs<float> a, StripArgs s) {
  extern __shared__ unsigned char smem_raw[];
  constexpr int RB = STRIP_RB;
  const int n = a.n, m = a.m, p = a.pitch;
  const int halo = 3 * (s.j1 - s.j0 + 1);
  const int c0 = 1 + blockIdx.x * s.cw, c1 = min(c0 + s.cw, n);   // core [c0, c1)
  const int lo = max(-3, c0 - halo), hi = min(n + 3, c1 - 1 + halo);
  const int W = s.W, rows = m + 7;
  float* const buf = reinterpret_cast<float*>(smem_raw);   // three strips of W*rows
  const int bsz = W * rows;
  const int tid = threadIdx.y * STRIP_TX + threadIdx.x, nthr = STRIP_TX * STRIP_TY;
  const float* cur_src = (s.j0 == 1) ? s.S0 : s.Gcur;
  const bool has_prev = s.j0 >= 3;
  // this thread's column and row block; pointwise operands in registers,
  // loaded from clamped (valid) points and zero where the thread has none
  const int ic = c0 - halo + 3 + threadIdx.x;
  const int k0 = 1 + threadIdx.y * RB;
  float s0v[RB], dtv[RB], rhsv[RB], visv[RB];
  bool on[RB];
#pragma unroll
  for (int r = 0; r < RB; r++) {
    const int k = k0 + r;
    on[r] = ic >= 1 && ic <= n - 1 && k <= m - 1;
    const long g = IDX(min(max(ic, 1), n - 1), min(k, m - 1), p);
    const float e0 = s.S0[g], d0 = a.dts[g], r0 = a.rhs[g], v0 = a.vis[g];
    s0v[r] = on[r] ? e0 : 0.0f;
    dtv[r] = on[r] ? d0 : 0.0f;
    rhsv[r] = on[r] ? r0 : 0.0f;
    visv[r] = on[r] ? v0 : 0.0f;
  }
  // the strips: each batch's loads (columns clamped to the strip, both
  // source buffers always allocated) issued before its LDS writes
  constexpr int SB = 8;
  const int nfill = W * rows;
  for (int e0 = tid; e0 < nfill; e0 += SB * nthr) {
    float vc[SB], vp[SB];
#pragma unroll
    for (int u = 0; u < SB; u++) {
      const int e = min(e0 + u * nthr, nfill - 1);
      const int rr = e / W, lc = e - rr * W;
      const long g = IDX(min(lo + lc, hi), rr - 3, p);
      vc[u] = cur_src[g];
      vp[u] = s.Gprev[g];
    }
#pragma unroll
    for (int u = 0; u < SB; u++) {
      const int e = e0 + u * nthr;
      if (e < nfill) {
        const int rr = e / W, lc = e - rr * W;
        const bool in = lo + lc <= hi;
        buf[e] = in ? vc[u] : 0.0f;
        buf[bsz + e] = (in && has_prev) ? vp[u] : 0.0f;
        buf[2 * bsz + e] = 0.0f;
      }
    }
  }
  const int rmax = m + 3;
  auto put = [&](float* B, int ii, int kk, float v) {
    if (ii >= lo && ii <= hi) B[(kk + 3) * W + (ii - lo)] = v;
  };
  __syncthreads();
  int ci = 0, pi = 1;
  for (int j = s.j0; j <= s.j1; j++) {
    const int t = j - s.j0;
    const int rlo = max(1, c0 - halo + 3 * (t + 1)), rhi = min(n - 1, c1 - 1 + halo - 3 * (t + 1));
    const int oi = 3 - ci - pi;
    const float* cur = buf + ci * bsz;
    const float* prev = buf + pi * bsz;
    float* out = buf + oi * bsz;
    const float cc0 = s.co.c0[j], cc1 = s.co.c1[j], cc2 = s.co.c2[j], cf = s.co.cf[j];
    if (ic >= rlo && ic <= rhi && k0 <= m - 1) {
      float F[RB];
      eval_cb<RB>(cur, W, lo, ic, k0, rmax, a, rhsv, visv, F);
#pragma unroll
      for (int r = 0; r < RB; r++) {
        if (!on[r]) continue;
        const int k = k0 + r;
        const int lg = (k + 3) * W + (ic - lo);
        float y;
        if (j == 1) y = s0v[r] + (cf * dtv[r]) * F[r];
        else y = ((cc0 * s0v[r] + cc1 * cur[lg]) + cc2 * prev[lg]) + ((cf * dtv[r]) * F[r]);
        // store_with_ghosts into the strip
        const float sz = a.sz;
        const bool lo_i = ic <= 3, hi_i = ic >= n - 3, lo_k = k <= 3, hi_k = k >= m - 3;
        put(out, ic, k, y);
        if (lo_i) put(out, -ic, k, -y);
        if (hi_i) put(out, 2 * n - ic, k, -y);
        if (lo_k) {
          put(out, ic, -k, sz * y);
          if (lo_i) put(out, -ic, -k, -(sz * y));
          if (hi_i) put(out, 2 * n - ic, -k, -(sz * y));
        }
        if (hi_k) {
          put(out, ic, 2 * m - k, sz * y);
          if (lo_i) put(out, -ic, 2 * m - k, -(sz * y));
          if (hi_i) put(out, 2 * n - ic, 2 * m - k, -(sz * y));
        }
      }
    }
    __syncthreads();
    pi = ci;
    ci = oi;
  }
  // core columns of the last two stages to global memory, with ghost images
  const int cwc = c1 - c0, npts = cwc * (m - 1);
  for (int e = tid; e < npts; e += nthr) {
    const int k = 1 + e / cwc, i = c0 + (e - (k - 1) * cwc);
    const int lg = (k + 3) * W + (i - lo);
    store_with_ghosts<float>(s.Ocur, i, k, n, m, p, buf[ci * bsz + lg], a.sz);
    if (s.Oprev) store_with_ghosts<float>(s.Oprev, i, k, n, m, p, buf[pi * bsz + lg], a.sz);
  }
}

// ---------------------------------------------------------------------------
// coarsest grid as a dense operator: G(:,q) = A^-1 (RowB(q) e_q) from the band
// LU factors (DGBTRS's forward and backward sweeps on a unit vector),
// interior unknowns only.  One wavefront per column q, the work vector in
// LDS: each sweep step's band updates are independent across rows and
// spread over the lanes; every element sees the same operations in the same
// order as the sequential sweep.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64) k_coarse_inverse(const double* ab, const int32_t* ipiv, const double* RowB,
                                                       double* G, int n, int kl, int ku, int nxc, int nzc) {
  extern __shared__ double Y[];   // n entries
  const int nci = (nxc - 1) * (nzc - 1);
  const int t = blockIdx.x, lane = threadIdx.x;
  const int ld = 2 * kl + ku + 1, kd = ku + kl + 1, N = nzc + 1;
  const int iq = 1 + t % (nxc - 1), kq = 1 + t / (nxc - 1);
  const int q = N * iq + kq + 1;   // 1-based unknown index
#define AB(r, c) ab[(long)((c) - 1) * ld + ((r) - 1)]
  for (int pp = lane; pp < n; pp += 64) Y[pp] = 0.0;
  __syncthreads();
  if (lane == 0) Y[q - 1] = RowB[q - 1];
  __syncthreads();
  for (int j = 1; j <= n - 1; j++) {
    const int lm = min(kl, n - j);
    const int l = ipiv[j - 1];
    if (l != j && lane == 0) {
      double tmp = Y[l - 1];
      Y[l - 1] = Y[j - 1];
      Y[j - 1] = tmp;
    }
    __syncthreads();
    const double temp = -1.0 * Y[j - 1];
    for (int u = 1 + lane; u <= lm; u += 64) Y[j + u - 1] = Y[j + u - 1] + AB(kd + u, j) * temp;
    __syncthreads();
  }
  const int kk = kl + ku, kp1 = kk + 1;
  for (int j = n; j >= 1; j--) {
    const double bj = Y[j - 1];
    if (bj != 0.0) {
      const int l = kp1 - j;
      const double x = bj / AB(kp1, j);
      const int lo = max(1, j - kk);
      __syncthreads();   // every lane has read Y(j) before it is replaced
      if (lane == 0) Y[j - 1] = x;
      for (int i = j - 1 - lane; i >= lo; i -= 64) Y[i - 1] = Y[i - 1] - x * AB(l + i, j);
    }
    __syncthreads();
  }
  for (int pi = lane; pi < nci; pi += 64) {
    const int ip = 1 + pi % (nxc - 1), kp = 1 + pi / (nxc - 1);
    G[(long)t * nci + pi] = Y[N * ip + kp];
  }
#undef AB
}

// coarse correction: err(p) = sum_q G(p,q) rhs(q), fixed order in q
template <typename RI, typename RO>
__global__ void k_coarse_apply(const double* G, const RI* rhs, RO* err, int nxc, int nzc, int pitch, RO sz) {
  extern __shared__ double rv[];
  const int nci = (nxc - 1) * (nzc - 1);
  for (int q = threadIdx.x; q < nci; q += blockDim.x) {
    int iq = 1 + q % (nxc - 1), kq = 1 + q / (nxc - 1);
    rv[q] = (double)rhs[IDX(iq, kq, pitch)];
  }
  __syncthreads();
  int pi = blockIdx.x * blockDim.x + threadIdx.x;
  if (pi >= nci) return;
  double acc = 0.0;
  for (int q = 0; q < nci; q++) acc = acc + G[(long)q * nci + pi] * rv[q];
  int ip = 1 + pi % (nxc - 1), kp = 1 + pi / (nxc - 1);
  store_with_ghosts<RO>(err, ip, kp, nxc, nzc, pitch, (RO)acc, sz);
}

// ---------------------------------------------------------------------------
// reductions: fixed-shape block partials then one block in fixed order
// ---------------------------------------------------------------------------
__global__ void k_reduce2d(const double* d, int n, int m, int pitch, int off, int op, double* partial) {
  __shared__ double red[256];
  int tid = threadIdx.x;
  const bool sum = op == RED_SUM_ABS || op == RED_SUM_SQ;
  double acc = sum ? 0.0 : (op == RED_MAX ? -INFINITY : INFINITY);
  long total = (long)n * m;
  for (long idx = (long)blockIdx.x * 256 + tid; idx < total; idx += (long)gridDim.x * 256) {
    int i = (int)(idx % n), k = (int)(idx / n);
    double v = d[(long)(k + off) * pitch + (i + off)];
    if (op == RED_SUM_ABS) acc = acc + fabs(v);
    else if (op == RED_SUM_SQ) acc = acc + v * v;
    else if (op == RED_MAX) acc = fmax(acc, v);
    else acc = fmin(acc, v);
  }
  red[tid] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s) {
      double x = red[tid], y = red[tid + s];
      red[tid] = sum ? x + y : (op == RED_MAX ? fmax(x, y) : fmin(x, y));
    }
    __syncthreads();
  }
  if (tid == 0) partial[blockIdx.x] = red[0];
}
__global__ void k_reduce_final(const double* partial, int nb, int op, double* out) {
  __shared__ double red[256];
  int tid = threadIdx.x;
  const bool sum = op == RED_SUM_ABS || op == RED_SUM_SQ;
  double acc = sum ? 0.0 : (op == RED_MAX ? -INFINITY : INFINITY);
  for (int j = tid; j < nb; j += 256) {
    double v = partial[j];
    acc = sum ? acc + v : (op == RED_MAX ? fmax(acc, v) : fmin(acc, v));
  }
  red[tid] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s) {
      double x = red[tid], y = red[tid + s];
      red[tid] = sum ? x + y : (op == RED_MAX ? fmax(x, y) : fmin(x, y));
    }
    __syncthreads();
  }
  if (tid == 0) *out = red[0];
}
int reduce_to_host(Context& c, const double* d, int n, int m, int pitch, int off, RedOp op, double* out) {
  long total = (long)n * m;
  int nb = (int)std::min<long>(1024, (total + 255) / 256);
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(k_reduce2d, dim3(nb), dim3(256), 0, c.stream, d, n, m, pitch, off, (int)op, c.red);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, c.stream, c.red, nb, (int)op, c.red_out);
  PT_HIP(hipMemcpyAsync(c.h_red, c.red_out, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  PT_HIP(hipStreamSynchronize(c.stream));
  *out = c.h_red[0];
  return PT_OK;
}

// ---------------------------------------------------------------------------
// grid transfer
// ---------------------------------------------------------------------------
// restrict / restrict_coefficients: out(i,k) (coarse interior) from in (fine)
template <typename RI>
__device__ __forceinline__ double restrict_val(const RI* in, int fpitch, int i, int k) {
  int ii = 2 * i, kk = 2 * k;
  double corners = ((double)in[IDX(ii - 1, kk - 1, fpitch)] + (double)in[IDX(ii - 1, kk + 1, fpitch)] +
                    (double)in[IDX(ii + 1, kk + 1, fpitch)] + (double)in[IDX(ii + 1, kk - 1, fpitch)]) / 16.0;
  double sides = ((double)in[IDX(ii - 1, kk, fpitch)] + (double)in[IDX(ii, kk - 1, fpitch)] +
                  (double)in[IDX(ii + 1, kk, fpitch)] + (double)in[IDX(ii, kk + 1, fpitch)]) / 8.0;
  double centre = (double)in[IDX(ii, kk, fpitch)] / 4.0;
  return corners + sides + centre;
}
template <typename RI, typename RO>
__device__ __forceinline__ void restrict_pt(const RI* in, int fpitch, RO* out, int cpitch, int i, int k) {
  out[IDX(i, k, cpitch)] = (RO)restrict_val<RI>(in, fpitch, i, k);
}
template <typename RI, typename RO>
__global__ void k_restrict(const RI* in, int fpitch, RO* out, int cn, int cm, int cpitch) {
  int i = 1 + blockIdx.x * blockDim.x + threadIdx.x, k = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i > cn - 1 || k > cm - 1) return;
  restrict_pt<RI, RO>(in, fpitch, out, cpitch, i, k);
}

// prolong: fine += bilinear(coarse), then the fine ghosts (four parity classes
// of PS/iterative_solver.f90:364-394 evaluated per fine point)
template <typename RC, typename RF>
__device__ __forceinline__ RF prolong_val(const RC* ec, int cpitch, const RF* ef, int fpitch, int i, int k) {
  RF add;
  bool ie = (i % 2) == 0, ke = (k % 2) == 0;
  if (ie && ke) add = (RF)ec[IDX(i / 2, k / 2, cpitch)];
  else if (ie && !ke) add = ((RF)ec[IDX(i / 2, (k - 1) / 2, cpitch)] + (RF)ec[IDX(i / 2, (k + 1) / 2, cpitch)]) / (RF)2;
  else if (!ie && ke) add = ((RF)ec[IDX((i - 1) / 2, k / 2, cpitch)] + (RF)ec[IDX((i + 1) / 2, k / 2, cpitch)]) / (RF)2;
  else
    add = ((RF)ec[IDX((i - 1) / 2, (k - 1) / 2, cpitch)] + (RF)ec[IDX((i + 1) / 2, (k - 1) / 2, cpitch)] +
           (RF)ec[IDX((i + 1) / 2, (k + 1) / 2, cpitch)] + (RF)ec[IDX((i - 1) / 2, (k + 1) / 2, cpitch)]) / (RF)4;
  return ef[IDX(i, k, fpitch)] + add;
}
template <typename RC, typename RF>
__device__ __forceinline__ void prolong_pt(const RC* ec, int cpitch, RF* ef, int fn, int fm, int fpitch, RF sz, int i,
                                           int k) {
  store_with_ghosts<RF>(ef, i, k, fn, fm, fpitch, prolong_val<RC, RF>(ec, cpitch, ef, fpitch, i, k), sz);
}
template <typename RC, typename RF>
__global__ void k_prolong(const RC* ec, int cpitch, RF* ef, int fn, int fm, int fpitch, RF sz) {
  int i = 1 + blockIdx.x * blockDim.x + threadIdx.x, k = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i > fn - 1 || k > fm - 1) return;
  prolong_pt<RC, RF>(ec, cpitch, ef, fn, fm, fpitch, sz, i, k);
}

// coarse correction err = G rhs by one workgroup (mixed path): the column
// range is split in CA_SEG fixed segments, each summed in order by its own
// thread group (loads issued 8 at a time), and the segment sums are added in
// segment order.  LDS: (1 + CA_SEG) * nci doubles.
constexpr int CA_SEG = 4, CA_ROWS = SMALL_THREADS / CA_SEG;
__device__ __forceinline__ void coarse_apply_wg(const double* G, const float* rhs, float* err, int nxc, int nzc,
                                                int pitch, float sz, double* sm) {
  const int nci = (nxc - 1) * (nzc - 1);
  double* rv = sm;
  double* part = sm + nci;
  for (int q = threadIdx.x; q < nci; q += SMALL_THREADS) {
    int iq = 1 + q % (nxc - 1), kq = 1 + q / (nxc - 1);
    rv[q] = (double)rhs[IDX(iq, kq, pitch)];
  }
  __syncthreads();
  const int seg = threadIdx.x / CA_ROWS, qs = (nci + CA_SEG - 1) / CA_SEG;
  const int q0 = seg * qs, q1 = min(nci, q0 + qs);
  for (int pi = threadIdx.x % CA_ROWS; pi < nci; pi += CA_ROWS) {
    double acc = 0.0;
    int q = q0;
    for (; q + 8 <= q1; q += 8) {
      double g[8];
#pragma unroll
      for (int u = 0; u < 8; u++) g[u] = G[(long)(q + u) * nci + pi];
#pragma unroll
      for (int u = 0; u < 8; u++) acc = acc + g[u] * rv[q + u];
    }
    for (; q < q1; q++) acc = acc + G[(long)q * nci + pi] * rv[q];
    part[seg * nci + pi] = acc;
  }
  __syncthreads();
  for (int pi = threadIdx.x; pi < nci; pi += SMALL_THREADS) {
    double acc = part[pi];
    for (int sgi = 1; sgi < CA_SEG; sgi++) acc = acc + part[sgi * nci + pi];
    int ip = 1 + pi % (nxc - 1), kp = 1 + pi / (nxc - 1);
    store_with_ghosts<float>(err, ip, kp, nxc, nzc, pitch, (float)acc, sz);
  }
}
__global__ void __launch_bounds__(SMALL_THREADS) k_coarse_apply_wg(const double* G, const float* rhs, float* err,
                                                                   int nxc, int nzc, int pitch, float sz) {
  extern __shared__ double ca_sm[];
  coarse_apply_wg(G, rhs, err, nxc, nzc, pitch, sz, ca_sm);
}

// enforceBCs_stream (PS/iterative_solver.f90:1616-1664) on a whole level-1
// field: walls (with their ghost extensions) zero, ghosts from the interior
__global__ void k_stream_bcs(double* S, int n, int m, int p, double sz) {
  const int i = -3 + (int)(blockIdx.x * blockDim.x + threadIdx.x), k = -3 + (int)(blockIdx.y * blockDim.y + threadIdx.y);
  if (i > n + 3 || k > m + 3) return;
  if (i == 0 || i == n || k == 0 || k == m) {
    S[IDX(i, k, p)] = 0.0;
    return;
  }
  if (i >= 1 && i <= n - 1 && k >= 1 && k <= m - 1) store_with_ghosts<double>(S, i, k, n, m, p, S[IDX(i, k, p)], sz);
}

// compute_RHS(1): RaT*d(Tbuoy)/dx - RaC*d(Cbuoy)/dx (5-point D1)
__global__ void k_rhs1(const double* Tb, const double* Cb, double* rhs, int nx, int nz, int p3, int p5, double dr,
                       double RaT, double RaC, double d0, double d1, double d2, double d3, double d4) {
  int i = 1 + blockIdx.x * blockDim.x + threadIdx.x, k = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i > nx - 1 || k > nz - 1) return;
  const double D[5] = {d0, d1, d2, d3, d4};
  double r = 0.0;
  if (RaT != 0.0) {
    double s = 0.0;
    for (int q = -2; q <= 2; q++) s = s + D[q + 2] * Tb[IDX(i + q, k, p3)];
    double Tr = s / dr;
    double Tx = Tr / 1.0;
    r = RaT * Tx;
  }
  double s = 0.0;
  for (int q = -2; q <= 2; q++) s = s + D[q + 2] * Cb[(long)(k + 5) * p5 + (i + q + 5)];
  double Cr = s / dr;
  double Cx = Cr / 1.0;
  r = r - RaC * Cx;
  rhs[IDX(i, k, p3)] = r;
}

// compute_velocities: w = -D1_x psi / dr, u = D1_z psi / ds on (0:nx,0:nz)
__global__ void k_velocities(const double* SF, double* u, double* w, int nx, int nz, int p, double dr, double ds,
                             double d0, double d1, double d2, double d3, double d4) {
  int i = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y * blockDim.y + threadIdx.y;
  if (i > nx || k > nz) return;
  const double D[5] = {d0, d1, d2, d3, d4};
  double sx = 0.0, sz = 0.0;
  for (int q = -2; q <= 2; q++) sx = sx + D[q + 2] * SF[IDX(i + q, k, p)];
  for (int q = -2; q <= 2; q++) sz = sz + D[q + 2] * SF[IDX(i, k + q, p)];
  w[IDX(i, k, p)] = -(sx / dr);
  u[IDX(i, k, p)] = sz / ds;
}

// velocityBCs (PS/composition.f90:1295-1325) as a gather over the ghost frame:
// every ghost takes its value from the interior image with the reference's sign
__global__ void k_velocity_bcs(double* u, double* w, int nx, int nz, int p, int Vbc) {
  int i = -3 + (int)(blockIdx.x * blockDim.x + threadIdx.x), k = -3 + (int)(blockIdx.y * blockDim.y + threadIdx.y);
  if (i > nx + 3 || k > nz + 3) return;
  bool gi = i < 0 || i > nx, gk = k < 0 || k > nz;
  int mi = i < 0 ? -i : (i > nx ? 2 * nx - i : i);
  int mk = k < 0 ? -k : (k > nz ? 2 * nz - k : k);
  // w: symmetric in x, antisymmetric in z; w=0 on k=0,nz (all columns)
  double wv, uv;
  if (!gi && !gk) {
    wv = w[IDX(i, k, p)]; uv = u[IDX(i, k, p)];
  } else {
    double wm = w[IDX(mi, mk, p)], um = u[IDX(mi, mk, p)];
    wv = gk ? -wm : wm;
    double su = (Vbc == 0) ? 1.0 : -1.0;
    uv = gi ? -um : um;
    if (gk) uv = su * uv;
  }
  if (k == 0 || k == nz) wv = 0.0;
  if (i == 0 || i == nx) uv = 0.0;
  if (Vbc == 1 && (k == 0 || k == nz)) uv = 0.0;
  if (gi || gk || k == 0 || k == nz || i == 0 || i == nx) {
    w[IDX(i, k, p)] = wv;
    u[IDX(i, k, p)] = uv;
  }
}

// ---------------------------------------------------------------------------
// viscosity hierarchy
// ---------------------------------------------------------------------------
__global__ void k_visf(const double* Tvis, const double* Cvis, double* vs, int nx, int nz, int p3, int p5, int ivis,
                       double lvT, double lvC, double lvP, double ds, int RaTnz) {
  int i = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y * blockDim.y + threadIdx.y;
  if (i > nx || k > nz) return;
  double zg = ds * (double)k;
  if (k == nz) zg = 1.0;
  double v = 1.0;
  if (ivis == 1) v = exp(-(Tvis[IDX(i, k, p3)] * lvT) + (1.0 - zg) * lvP);
  else if (ivis == 2) {
    double dp = 0.0 + Cvis[(long)(k + 5) * p5 + (i + 5)] * lvC;
    if (RaTnz) v = exp(-(Tvis[IDX(i, k, p3)] * lvT) - dp + (1.0 - zg) * lvP);
    else v = exp(-dp + (1.0 - zg) * lvP);
  }
  vs[IDX(i, k, p3)] = v;
}

// viscosityBCs (PS/iterative_solver.f90:666-690): side ghosts on rows
// -3..m+3 then top/bottom ghosts on columns 0..n.  The side pass reads the
// ghost rows before the top/bottom pass rewrites them, as the reference.
__global__ void k_vis_bcs_side(double* v, int n, int m, int p) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;   // over rows -3..m+3, 3 ghosts each side
  int rows = m + 7;
  if (t >= rows * 3) return;
  int k = -3 + t / 3, i = 1 + t % 3;
  v[IDX(-i, k, p)] = v[IDX(i, k, p)];
  v[IDX(n + i, k, p)] = v[IDX(n - i, k, p)];
}
__global__ void k_vis_bcs_tb(double* v, int n, int m, int p) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (n + 1) * 3) return;
  int i = t / 3, k = 1 + t % 3;
  v[IDX(i, -k, p)] = v[IDX(i, k, p)];
  v[IDX(i, m + k, p)] = v[IDX(i, m - k, p)];
}

// viscosityBCs_initialize (PS/iterative_solver.f90:636-664)
__global__ void k_vis_bcs_init_boundary(double* v, const double* f, int n, int m, int p, int fn, int fm, int fp) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t <= m - 2) {   // k = 1..m-1
    int k = t + 1;
    v[IDX(0, k, p)] = f[IDX(0, 2 * k, fp)];
    v[IDX(n, k, p)] = f[IDX(fn, 2 * k, fp)];
  }
  if (t <= n) {
    int i = t;
    v[IDX(i, 0, p)] = f[IDX(2 * i, 0, fp)];
    v[IDX(i, m, p)] = f[IDX(2 * i, fm, fp)];
  }
}
__global__ void k_vis_bcs_init_side(double* v, int n, int m, int p) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (m + 1) * 3) return;
  int k = t / 3, i = 1 + t % 3;
  v[IDX(-i, k, p)] = v[IDX(i, k, p)];
  v[IDX(n + i, k, p)] = v[IDX(n - i, k, p)];
}
__global__ void k_vis_bcs_init_tb(double* v, int n, int m, int p) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (n + 7) * 3) return;
  int i = -3 + t / 3, k = 1 + t % 3;
  v[IDX(i, -k, p)] = v[IDX(i, k, p)];
  v[IDX(i, m + k, p)] = v[IDX(i, m - k, p)];
}

// viscosity_gradients, uniform grid (PS/iterative_solver.f90:488-520)
__global__ void k_vis_grad(VisSet s, int n, int m, int p, int full, int ivis0, double dr, double ds, double dr2,
                           double ds2, double drds, double dtv) {
  int i = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y * blockDim.y + threadIdx.y;
  if (i > n || k > m) return;
  long g = IDX(i, k, p);
  if (ivis0) {
    s.vx[g] = 0; s.vz[g] = 0; s.vxx[g] = 0; s.vzz[g] = 0; s.vxz[g] = 0; s.vxxpzz[g] = 0; s.vxxmzz[g] = 0;
    return;
  }
  const double* v = s.vis;
  if (full) {
    // Drs_so(:,-1)=(.25,0,-.25), (:,0)=0, (:,1)=(-.25,0,.25); sum in column-major order
    double sxz = 0.0;
    sxz = sxz + 0.25 * v[IDX(i - 1, k - 1, p)];
    sxz = sxz + 0.0 * v[IDX(i, k - 1, p)];
    sxz = sxz + (-0.25) * v[IDX(i + 1, k - 1, p)];
    sxz = sxz + 0.0 * v[IDX(i - 1, k, p)];
    sxz = sxz + 0.0 * v[IDX(i, k, p)];
    sxz = sxz + 0.0 * v[IDX(i + 1, k, p)];
    sxz = sxz + (-0.25) * v[IDX(i - 1, k + 1, p)];
    sxz = sxz + 0.0 * v[IDX(i, k + 1, p)];
    sxz = sxz + 0.25 * v[IDX(i + 1, k + 1, p)];
    s.vxz[g] = sxz / drds;
    double sx = 0.0;
    sx = sx + (-0.5) * v[IDX(i - 1, k, p)];
    sx = sx + 0.0 * v[IDX(i, k, p)];
    sx = sx + 0.5 * v[IDX(i + 1, k, p)];
    s.vx[g] = sx / dr;
  }
  double sxx = 0.0;
  sxx = sxx + 1.0 * v[IDX(i - 1, k, p)];
  sxx = sxx + (-2.0) * v[IDX(i, k, p)];
  sxx = sxx + 1.0 * v[IDX(i + 1, k, p)];
  double vxx = sxx / dr2;
  s.vxx[g] = vxx;
  if (full) {
    double sz = 0.0;
    sz = sz + (-0.5) * v[IDX(i, k - 1, p)];
    sz = sz + 0.0 * v[IDX(i, k, p)];
    sz = sz + 0.5 * v[IDX(i, k + 1, p)];
    s.vz[g] = sz / ds;
  }
  double szz = 0.0;
  szz = szz + 1.0 * v[IDX(i, k - 1, p)];
  szz = szz + (-2.0) * v[IDX(i, k, p)];
  szz = szz + 1.0 * v[IDX(i, k + 1, p)];
  double vzz = szz / ds2;
  s.vzz[g] = vzz;
  double xp = vxx + vzz;
  s.vxxpzz[g] = xp;
  if (full) s.vxxmzz[g] = vxx - vzz;
  s.vsmooth[g] = dtv * xp;
}

// test_eigenvalues, RKC1 branch (PS/iterative_solver.f90:1187-1208) + block min of ratio
__global__ void k_test_eig(VisSet s, double* ratio, double* dts, int n, int m, int p, double SR1, double SR2,
                           double SR4, double SR5, double SR6, double SI1, double SI2, double SI4, double SI5,
                           double SI6, double ae, double be, double ctemp, double* partial) {
  __shared__ double red[256];
  int tid = threadIdx.y * blockDim.x + threadIdx.x;
  int i = 1 + blockIdx.x * blockDim.x + threadIdx.x, k = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  double rmin = INFINITY;
  if (i <= n - 1 && k <= m - 1) {
    long g = IDX(i, k, p);
    double vis = s.vis[g];
    double eI = fabs(s.vxxpzz[g] * SI1) + vis * SI2 + fabs(s.vxz[g] * SI4) + fabs(s.vx[g] * SI5) + fabs(s.vz[g] * SI6);
    double eR = s.vxxmzz[g] * SR1 + vis * SR2 + s.vxz[g] * SR4 + s.vx[g] * SR5 + s.vz[g] * SR6;
    double r = fabs(eR / eI);
    ratio[g] = r;
    double qa = eR / ae, qb = eI / be;
    dts[g] = ctemp * eR / (qa * qa + qb * qb);
    rmin = r;
  }
  red[tid] = rmin;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (tid < st) red[tid] = fmin(red[tid], red[tid + st]);
    __syncthreads();
  }
  if (tid == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

__global__ void k_vis_update(double* v, const double* sm, int n, int m, int p, double c) {
  int i = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y * blockDim.y + threadIdx.y;
  if (i > n || k > m) return;
  long g = IDX(i, k, p);
  v[g] = v[g] + c * sm[g];
}

// ---------------------------------------------------------------------------
// coarsest grid: banded operator, row equilibration, LU, solve
// ---------------------------------------------------------------------------
struct CoarseArgs {
  const double *vis, *vxm, *vxz, *vx, *vz;
  int nxc, nzc, p, kl, ku, ngc, ivis0, Vbc;
  double SM1[7][7], SM2[7][7], SM4[7][7], SM5[7][7], SM6[7][7];
};

__device__ double stream_space_matrix(const CoarseArgs& a, int ip, int kp, int iq, int kq, int sval) {
  int di = iq - ip, dk = kq - kp;
  long g = IDX(ip, kp, a.p);
  double v;
  if (a.ivis0) v = a.vis[g] * a.SM2[dk + 3][di + 3];
  else
    v = a.vxm[g] * a.SM1[dk + 3][di + 3] + a.vis[g] * a.SM2[dk + 3][di + 3] + a.vxz[g] * a.SM4[dk + 3][di + 3] +
        a.vx[g] * a.SM5[dk + 3][di + 3] + a.vz[g] * a.SM6[dk + 3][di + 3];
  if (sval == -1) v = -v;
  return v;
}

// one thread per matrix row p (Compute_Matrix_coarse, analysis=0)
__global__ void k_coarse_assemble(CoarseArgs a, double* Mc, double* RowB) {
  int p = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (p > a.ngc) return;
  const int limit[7] = {0, 2, 2, 3, 2, 2, 0};
  int ld = 2 * a.kl + a.ku + 1, kk1 = a.kl + a.ku + 1, N = a.nzc + 1;
  int kp = p - N * ((p - 1) / N) - 1, ip = (p - 1) / N;
  auto M = [&](int r, int q) -> double& { return Mc[(long)(q - 1) * ld + (r - 1)]; };
  if (ip > 0 && ip < a.nxc && kp > 0 && kp < a.nzc) {
    for (int j = -3; j <= 3; j++) {
      int iq = j + ip;
      if (iq >= 0 && iq <= a.nxc)
        for (int kq = kp - limit[j + 3]; kq <= kp + limit[j + 3]; kq++)
          if (kq >= 0 && kq <= a.nzc) {
            int q = N * iq + kq + 1;
            M(kk1 + p - q, q) = stream_space_matrix(a, ip, kp, iq, kq, 1);
          }
    }
    for (int j = -3; j <= 3; j++) {
      int iq = j + ip;
      for (int kq = kp - limit[j + 3]; kq <= kp + limit[j + 3]; kq++)
        if (iq < 0 || iq > a.nxc || kq < 0 || kq > a.nzc) {
          int iiq, kkq, xv, zv;
          if (iq < 0) { iiq = -iq; xv = -1; }
          else if (iq > a.nxc) { iiq = 2 * a.nxc - iq; xv = -1; }
          else { iiq = iq; xv = 1; }
          if (kq < 0) { kkq = -kq; zv = a.Vbc == 0 ? -1 : 1; }
          else if (kq > a.nzc) { kkq = 2 * a.nzc - kq; zv = a.Vbc == 0 ? -1 : 1; }
          else { kkq = kq; zv = 1; }
          int q = N * iiq + kkq + 1;
          M(kk1 + p - q, q) = M(kk1 + p - q, q) + stream_space_matrix(a, ip, kp, iq, kq, xv * zv);
        }
    }
  } else {
    M(kk1, p) = 1.0;   // impermeable boundary rows
  }
}

// DGBEQU row scale factors and row scaling; one thread per row
__global__ void k_coarse_equilibrate(double* Mc, double* RowB, int n, int kl, int ku) {
  int p = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (p > n) return;
  int ld = 2 * kl + ku + 1, kk1 = kl + ku + 1;
  double r = 0.0;
  int qlo = std::max(1, p - kl), qhi = std::min(n, p + ku);
  for (int q = qlo; q <= qhi; q++) {
    double a = fabs(Mc[(long)(q - 1) * ld + (kk1 + p - q - 1)]);
    if (a > r) r = a;
  }
  const double smlnum = 2.2250738585072014e-308, bignum = 1.0 / smlnum;
  r = fmin(fmax(r, smlnum), bignum);
  double rb = 1.0 / r;
  RowB[p - 1] = rb;
  for (int q = qlo; q <= qhi; q++) {
    double& m = Mc[(long)(q - 1) * ld + (kk1 + p - q - 1)];
    m = rb * m;
  }
}

// DGBTF2 (LAPACK unblocked band LU with partial pivoting) on one wavefront
__global__ void __launch_bounds__(64) k_band_lu(double* ab, int32_t* ipiv, int n, int kl, int ku, int32_t* info) {
  const int ld = 2 * kl + ku + 1, kv = ku + kl;
  const int lane = threadIdx.x;
#define AB(r, q) ab[(long)((q) - 1) * ld + ((r) - 1)]
  for (int j = ku + 2 + lane; j <= std::min(kv, n); j += 64)
    for (int i = kv - j + 2; i <= kl; i++) AB(i, j) = 0.0;
  __syncthreads();
  int ju = 1, inf = 0;
  for (int j = 1; j <= n; j++) {
    if (j + kv <= n)
      for (int i = 1 + lane; i <= kl; i += 64) AB(i, j + kv) = 0.0;
    int km = std::min(kl, n - j);
    // IDAMAX over AB(kv+1 .. kv+1+km, j): first index of the maximum |.|
    double best = -1.0; int bi = 0x7fffffff;
    for (int t = lane; t <= km; t += 64) {
      double a = fabs(AB(kv + 1 + t, j));
      if (a > best) { best = a; bi = t; }
    }
    for (int off = 32; off > 0; off >>= 1) {
      double ob = __shfl_down(best, off);
      int oi = __shfl_down(bi, off);
      if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
    }
    int jp = __shfl(bi, 0) + 1;
    __syncthreads();
    if (lane == 0) ipiv[j - 1] = jp + j - 1;
    double piv = AB(kv + jp, j);
    if (piv != 0.0) {
      ju = std::max(ju, std::min(j + ku + jp - 1, n));
      if (jp != 1)
        for (int t = lane; t <= ju - j; t += 64) {
          double* pa = &ab[(long)(j - 1 + t) * ld + (kv + jp - 1 - t)];
          double* pb = &ab[(long)(j - 1 + t) * ld + (kv - t)];
          double tmp = *pa; *pa = *pb; *pb = tmp;
        }
      __syncthreads();
      if (km > 0) {
        double rcp = 1.0 / AB(kv + 1, j);
        for (int t = 1 + lane; t <= km; t += 64) AB(kv + 1 + t, j) = rcp * AB(kv + 1 + t, j);
        __syncthreads();
        // DGER: columns are independent (the pivot-row element y(jj) is not
        // updated), so the (jj, t) pairs are spread over the wavefront
        const int total = (ju - j) * km;
        for (int idx = lane; idx < total; idx += 64) {
          int jj = 1 + idx / km, t = 1 + idx % km;
          double yv = ab[(long)(j - 1 + jj) * ld + (kv - jj)];
          double temp = -1.0 * yv;
          double& x = ab[(long)(j - 1 + jj) * ld + (kv - jj + t)];
          x = x + AB(kv + 1 + t, j) * temp;
        }
      }
    } else if (inf == 0) {
      inf = j;
    }
    __syncthreads();
  }
  if (lane == 0) *info = inf;
#undef AB
}

// ---------------------------------------------------------------------------
// host orchestration
// ---------------------------------------------------------------------------
static dim3 grid2(int ni, int nk, dim3 b) { return dim3((ni + b.x - 1) / b.x, (nk + b.y - 1) / b.y); }

template <typename R>
static StageArgs<R> stage_base(Context& c, int l, bool fp32) {
  StageArgs<R> a;
  std::memset(&a, 0, sizeof a);
  Level& L = c.L[l];
  const LevelGeom& g = L.g;
  a.n = g.n; a.m = g.m; a.pitch = g.pitch;
  a.sz = (R)(c.P.Vbc == 0 ? -1.0 : 1.0);
  a.dr2 = (R)g.dr2; a.dr3 = (R)g.dr3; a.dr4 = (R)g.dr4; a.ds2 = (R)g.ds2; a.ds3 = (R)g.ds3; a.ds4 = (R)g.ds4;
  a.drds = (R)g.drds; a.dr2ds = (R)g.dr2ds; a.drds2 = (R)g.drds2; a.dr2ds2 = (R)g.dr2ds2;
  a.rdr2 = (R)(1.0 / g.dr2); a.rdr3 = (R)(1.0 / g.dr3); a.rdr4 = (R)(1.0 / g.dr4);
  a.rds2 = (R)(1.0 / g.ds2); a.rds3 = (R)(1.0 / g.ds3); a.rds4 = (R)(1.0 / g.ds4);
  a.rdrds = (R)(1.0 / g.drds); a.rdr2ds = (R)(1.0 / g.dr2ds); a.rdrds2 = (R)(1.0 / g.drds2);
  a.rdr2ds2 = (R)(1.0 / g.dr2ds2);
  for (int q = 0; q < 5; q++) { a.D1[q] = (R)c.st.D1[q]; a.D2[q] = (R)c.st.D2[q]; }
  for (int q = 0; q < 7; q++) { a.D3[q] = (R)c.st.D3[q]; a.D4[q] = (R)c.st.D4[q]; }
  for (int b = 0; b < 5; b++)
    for (int q = 0; q < 5; q++) {
      a.Drs[b][q] = (R)c.st.Drs[b][q]; a.Drss[b][q] = (R)c.st.Drss[b][q];
      a.Drrs[b][q] = (R)c.st.Drrs[b][q]; a.Drrss[b][q] = (R)c.st.Drrss[b][q];
    }
  a.ivis0 = c.P.ivis == 0;
  if (!fp32) {
    a.vis = (const R*)L.v.vis; a.vxm = (const R*)L.v.vxxmzz; a.vxz = (const R*)L.v.vxz;
    a.vx = (const R*)L.v.vx; a.vz = (const R*)L.v.vz; a.rhs = (const R*)L.rhs; a.dts = (const R*)L.dts;
  } else {
    a.vis = (const R*)L.fvis; a.vxm = (const R*)L.fvxm; a.vxz = (const R*)L.fvxz;
    a.vx = (const R*)L.fvx; a.vz = (const R*)L.fvz; a.rhs = (const R*)L.frhs; a.dts = (const R*)L.fdts;
  }
  return a;
}

template <typename R, bool IV0>
static void launch_stage(Context& c, dim3 grd, dim3 blk, const StageArgs<R>& a) {
  hipLaunchKernelGGL((k_stage<R, IV0>), grd, blk, 0, c.stream, a);
}
template <typename R>
static void launch_stage(Context& c, dim3 grd, dim3 blk, const StageArgs<R>& a) {
  if (c.P.ivis == 0) launch_stage<R, true>(c, grd, blk, a);
  else launch_stage<R, false>(c, grd, blk, a);
}
template <typename R>
static void launch_residual(Context& c, dim3 grd, dim3 blk, const StageArgs<R>& a, R* res, double* partial) {
  if (c.P.ivis == 0) hipLaunchKernelGGL((k_residual<R, true>), grd, blk, 0, c.stream, a, res, partial);
  else hipLaunchKernelGGL((k_residual<R, false>), grd, blk, 0, c.stream, a, res, partial);
}

// points per thread of the single-workgroup smoother, 0 when the level is
// too large for it: beyond 4 points (isoviscous) the per-thread state no
// longer fits in 128 VGPRs; with the full variable-viscosity operator one
// workgroup is slower than a launch per stage beyond 2 points (FP32 level
// 127x31: 248 us per 35-stage sweep against 35 x 4.9 us)
static int small_ppt(int n, int m, bool iv0) {
  const int npts = (n - 1) * (m - 1);
  const int pmax = iv0 ? 4 : 2;
  for (int ppt = 1; ppt <= pmax; ppt *= 2)
    if (npts <= ppt * SMALL_THREADS) return ppt;
  return 0;
}

template <typename R, bool IV0, int PPT>
static int launch_small_t(Context& c, size_t lds, const StageArgs<R>& a, R* err, R* res, int s, double cc,
                          const RkcCoef& co) {
  // set once per process (a thread-safe static initialisation)
  static const hipError_t attr_rc = hipFuncSetAttribute((const void*)k_smooth_small<R, IV0, PPT>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)SMALL_LDS_BYTES);
  PT_HIP(attr_rc);
  hipLaunchKernelGGL((k_smooth_small<R, IV0, PPT>), dim3(1), dim3(SMALL_THREADS), lds, c.stream, a, err, res, s, cc,
                     co);
  return PT_OK;
}
template <typename R, bool IV0>
static int launch_small_v(Context& c, int ppt, size_t lds, const StageArgs<R>& a, R* err, R* res, int s, double cc,
                          const RkcCoef& co) {
  switch (ppt) {
    case 1: return launch_small_t<R, IV0, 1>(c, lds, a, err, res, s, cc, co);
    case 2: return launch_small_t<R, IV0, 2>(c, lds, a, err, res, s, cc, co);
    default: return launch_small_t<R, IV0, 4>(c, lds, a, err, res, s, cc, co);
  }
}

// rows per thread of the column-blocked FP32 sweep: 4 when that keeps at
// least half the workgroup busy with one column block per thread, 1 when
// every interior point has a thread of its own, else 0
static int cb_rows(int n, int m) {
  const int b4 = (n - 1) * ((m - 1 + 3) / 4);
  if (b4 <= SMALL_THREADS && 2 * b4 >= SMALL_THREADS) return 4;
  return (n - 1) * (m - 1) <= SMALL_THREADS ? 1 : 0;
}
template <int RB>
static int launch_cb_t(Context& c, size_t lds, const StageArgs<float>& a, float* err, float* res, int s, double cc,
                       const RkcCoef& co) {
  // set once per process (a thread-safe static initialisation)
  static const hipError_t attr_rc = hipFuncSetAttribute((const void*)k_smooth_cb<RB>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)SMALL_LDS_BYTES);
  PT_HIP(attr_rc);
  hipLaunchKernelGGL(k_smooth_cb<RB>, dim3(1), dim3(SMALL_THREADS), lds, c.stream, a, err, res, s, rkc_f32(co, cc));
  return PT_OK;
}
static int launch_cb(Context& c, int rb, size_t lds, const StageArgs<float>& a, float* err, float* res, int s,
                     double cc, const RkcCoef& co) {
  if (rb == 1) return launch_cb_t<1>(c, lds, a, err, res, s, cc, co);
  return launch_cb_t<4>(c, lds, a, err, res, s, cc, co);
}

// strip sweep of an FP32 isoviscous level: core columns per workgroup and
// stages per segment (cw == 0: not eligible).  The widest stage region (core
// + 6 columns per further stage) spans at most the 64 column lanes; the
// three strip buffers must fit LDS; rows are covered by 16 blocks of 4.
// The most fused stages come first (9 stages on a 16-column core: 4
// launches per 35-stage sweep): a workgroup's stage takes about the same
// time for any core width, since its 16 waves wait on LDS and the barrier
// rather than on issue, and each launch pays a strip load and store.
struct StripCfg {
  int cw = 0, ns = 0, W = 0;
};
static StripCfg strip_cfg(const LevelGeom& g) {
  StripCfg sc;
  if (g.m - 1 > STRIP_TY * STRIP_RB) return sc;
  const size_t rows = (size_t)g.m + 7;
  for (int ns = (STRIP_TX + 6 - 16) / 6; ns >= 2; ns--) {
    for (int cw = STRIP_TX + 6 - 6 * ns; cw >= 16; cw -= 4) {
      const int W = std::min(cw + 6 * ns, g.n + 7);
      if (3 * (size_t)W * rows * sizeof(float) <= SMALL_LDS_BYTES) {
        sc.cw = cw; sc.ns = ns; sc.W = W;
        return sc;
      }
    }
  }
  return sc;
}

// one FP32 RKC1 sweep of level l as segments of fused stages.  The error
// buffer holds S0 throughout; segments read one buffer pair and write the
// other; the sweep result replaces the error buffer by a pointer swap.
static int strip_sweep(Context& c, int l, const StripCfg& sc, double cpar, const RkcCoef& co) {
  // set once per process (a thread-safe static initialisation)
  static const hipError_t attr_rc = hipFuncSetAttribute((const void*)k_strip_cb, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)SMALL_LDS_BYTES);
  PT_HIP(attr_rc);
  Level& L = c.L[l];
  float** f[5] = {&L.ferr, &L.fB, &L.fC, &L.fD, &L.fE};
  const LevelGeom& g = L.g;
  const int s = level_rkc(c, l, true).s;
  const int nseg = (s + sc.ns - 1) / sc.ns, base = s / nseg, rem = s % nseg;
  if (base < 2) {
    set_error("strip sweep needs at least 2 stages per segment");
    return PT_ERR_UNSUPPORTED;
  }
  const int nblk = (g.n - 1 + sc.cw - 1) / sc.cw;
  const size_t lds = 3 * (size_t)sc.W * (g.m + 7) * sizeof(float);
  StageArgs<float> a = stage_base<float>(c, l, true);
  StripArgs sa;
  sa.S0 = *f[0]; sa.cw = sc.cw; sa.W = sc.W; sa.co = rkc_f32(co, cpar);
  int rd = 1, wr = 3;   // read pair f[rd], f[rd+1]; write pair f[wr], f[wr+1]
  int j0 = 1;
  for (int seg = 0; seg < nseg; seg++) {
    const int len = base + (seg < rem ? 1 : 0), j1 = j0 + len - 1;
    sa.j0 = j0; sa.j1 = j1;
    sa.Gcur = *f[rd]; sa.Gprev = *f[rd + 1];
    sa.Ocur = *f[wr]; sa.Oprev = (j1 == s) ? nullptr : *f[wr + 1];
    hipLaunchKernelGGL(k_strip_cb, dim3(nblk), dim3(STRIP_TX, STRIP_TY), lds, c.stream, a, sa);
    std::swap(rd, wr);
    j0 = j1 + 1;
  }
  std::swap(*f[0], *f[rd]);   // the last segment's output becomes the error buffer
  return PT_OK;
}

// first level of the fused sub-cycle, 0 when it does not apply: levels
// ls..grids-1 are single-workgroup sweeps whose argument table fits the
// kernel arguments, the coarse operator fits one workgroup, and in the mixed
// path level ls-1 is an FP32 level (ls >= 3)
static int launch_coarse_apply_wg(Context& c, const float* rhs, float* err, int n, int m, int pitch, float sz) {
  // set once per process (a thread-safe static initialisation)
  static const hipError_t attr_rc = hipFuncSetAttribute((const void*)k_coarse_apply_wg, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)SMALL_LDS_BYTES);
  PT_HIP(attr_rc);
  const size_t lds = sizeof(double) * (1 + CA_SEG) * c.nci;
  if (lds > SMALL_LDS_BYTES) {
    set_error("coarsest grid too large for the single-workgroup coarse operator");
    return PT_ERR_UNSUPPORTED;
  }
  hipLaunchKernelGGL(k_coarse_apply_wg, dim3(1), dim3(SMALL_THREADS), lds, c.stream, c.Gc, rhs, err, n, m, pitch, sz);
  return PT_OK;
}

// iterate_stream (RKC1, Nmulti sweeps) followed by compute_residual
// (PS/iterative_solver.f90:1418-1614, 415-448).  `s0zero` runs the smoother
// on a correction that starts from zero (mixed precision).
// residual of the smoothed field
template <typename R>
static int smoothed_residual(Context& c, int l, dim3 grd, dim3 blk, bool fp32, bool l1_sum) {
  Level& L = c.L[l];
  StageArgs<R> a = stage_base<R>(c, l, fp32);
  a.Sx = fp32 ? (R*)L.ferr : (R*)L.err;
  R* res = fp32 ? (R*)L.fB : (R*)L.res;   // FP32 residual reuses the B buffer
  int nb = grd.x * grd.y;
  launch_residual<R>(c, grd, blk, a, res, c.red);
  if (l1_sum) {
    // the L1 norm of the residual into device memory; level1_residual copies
    // it to the host after the cycle's launches (or graph replay)
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, c.stream, c.red, nb, (int)RED_SUM_ABS, c.red_out);
  }
  return PT_OK;
}

template <typename R>
static int iterate_stream_t(Context& c, int l, int p, bool fp32, bool l1_sum) {
  Level& L = c.L[l];
  const LevelGeom& g = L.g;
  const LevelRkc rk = level_rkc(c, l, fp32);
  const int s = rk.s;
  dim3 blk(TX, TZ), grd((g.n - 1 + TX - 1) / TX, (g.m - 1 + TZ * PZ - 1) / (TZ * PZ));
  R* err = fp32 ? (R*)L.ferr : (R*)L.err;
  R* B = fp32 ? (R*)L.fB : (R*)L.bufB;
  R* Cb = fp32 ? (R*)L.fC : (R*)L.bufC;
  const bool timed = c.timing;
  const int tk = l == 1 ? PT_K_RKC_STAGE_L1 : PT_K_RKC_STAGE_COARSE;
  const size_t lds = 3 * L.nelem * sizeof(R);
  const int ppt = small_ppt(g.n, g.m, c.P.ivis == 0);
  if constexpr (sizeof(R) == 8) {
    if (c.P.RKC1 == 0) {
      // RK1: per iteration F = stream_space(S0) into B, then the update
      const int dir = c.cyc[p + 1] > c.cyc[p] ? 0 : 1;
      dim3 b16(64, 4), gi((g.n - 1 + 63) / 64, (g.m - 1 + 3) / 4);
      for (int it = 1; it <= c.P.Nmulti; it++) {
        const int ntemp = c.P.Nmulti - it + 1;
        const double temp = std::cos((double)(2 * ntemp - 1) * 3.1415926535897932 / (double)(2 * c.P.Nmulti));
        StageArgs<R> a = stage_base<R>(c, l, false);
        a.Sx = err;
        if (timed) timer_begin(c, tk, (int64_t)(g.n - 1) * (g.m - 1));
        launch_residual<R>(c, grd, blk, a, B, c.red);
        hipLaunchKernelGGL(k_rk1_update, gi, b16, 0, c.stream, (double*)err, (const double*)B, L.dts, L.ratio, g.n,
                           g.m, g.pitch, temp, dir, (double)a.sz);
        if (timed) timer_end(c, tk);
      }
      return smoothed_residual<R>(c, l, grd, blk, fp32, l1_sum);
    }
  }
  if constexpr (sizeof(R) == 4) {
    // FP32 isoviscous levels too large for one workgroup: strip sweeps
    const StripCfg sc = strip_cfg(g);
    if (l >= 2 && c.P.ivis == 0 && ppt == 0 && sc.cw > 0 && !l1_sum) {
      const RkcCoef co = rkc_coef(rk);
      for (int it = 1; it <= c.P.Nmulti; it++) {
        int ntemp = c.P.Nmulti - it + 1;
        double temp = (c.P.Nmulti > 1) ? (c.P.c_2 - c.P.c_1) * (double)(ntemp - 1) / (double)(1 - c.P.Nmulti) : 0.0;
        PT_TRY(strip_sweep(c, l, sc, temp + c.P.c_2, co));
      }
      StageArgs<R> a = stage_base<R>(c, l, fp32);
      a.Sx = (R*)L.ferr;
      launch_residual<R>(c, grd, blk, a, (R*)L.fB, c.red);
      (void)p;
      return PT_OK;
    }
  }
  if (l >= 2 && ppt > 0 && lds <= SMALL_LDS_BYTES && !l1_sum) {
    // whole sweep + residual in one workgroup, buffers in LDS
    const RkcCoef co = rkc_coef(rk);
    R* res = fp32 ? (R*)L.fB : (R*)L.res;
    for (int it = 1; it <= c.P.Nmulti; it++) {
      int ntemp = c.P.Nmulti - it + 1;
      double temp = (c.P.Nmulti > 1) ? (c.P.c_2 - c.P.c_1) * (double)(ntemp - 1) / (double)(1 - c.P.Nmulti) : 0.0;
      double cc = temp + c.P.c_2;
      StageArgs<R> a = stage_base<R>(c, l, fp32);
      // the residual buffer aliases fB in FP32 mode: the fused kernel keeps its
      // RKC buffers in LDS, so fB is free
      if constexpr (sizeof(R) == 4) {
        const int rb = cb_rows(g.n, g.m);
        if (c.P.ivis == 0 && rb > 0) {
          PT_TRY(launch_cb(c, rb, lds, a, err, res, s, cc, co));
          continue;
        }
      }
      if (c.P.ivis == 0) PT_TRY((launch_small_v<R, true>(c, ppt, lds, a, err, res, s, cc, co)));
      else PT_TRY((launch_small_v<R, false>(c, ppt, lds, a, err, res, s, cc, co)));
    }
    (void)p;
    return PT_OK;
  }
  for (int it = 1; it <= c.P.Nmulti; it++) {
    int ntemp = c.P.Nmulti - it + 1;
    double temp = (c.P.Nmulti > 1) ? (c.P.c_2 - c.P.c_1) * (double)(ntemp - 1) / (double)(1 - c.P.Nmulti) : 0.0;
    double cc = temp + c.P.c_2;
    StageArgs<R> a = stage_base<R>(c, l, fp32);
    // stage 1: S1 = S0 + mewt1*c*dt*F(S0) -> B
    a.Sx = err; a.S0 = err; a.S2 = err; a.Y = B; a.stage1 = 1; a.cf = (R)(rk.mewt[1] * cc);
    if (timed) timer_begin(c, tk, (int64_t)(g.n - 1) * (g.m - 1));
    launch_stage<R>(c, grd, blk, a);
    if (timed) timer_end(c, tk);
    // rotating buffers: S1 in `cur`, S2 in `prev`; Y goes to the S2 buffer
    // except at stage 2 (S2 aliases S0) and at the last stage (Y -> err)
    R* cur = B;
    R* prev = err;   // S2 = S0 interior at stage 2
    R* spare = Cb;
    for (int j = 2; j <= s; j++) {
      a.stage1 = 0;
      a.Sx = cur; a.S0 = err; a.S2 = prev;
      a.c0 = (R)rk.cc[j]; a.c1 = (R)rk.mew[j]; a.c2 = (R)rk.vv[j]; a.cf = (R)(rk.mewt[j] * cc);
      R* out;
      if (j == s) out = err;
      else if (prev == err) out = spare;
      else out = prev;
      a.Y = out;
      if (timed) timer_begin(c, tk, (int64_t)(g.n - 1) * (g.m - 1));
      launch_stage<R>(c, grd, blk, a);
      if (timed) timer_end(c, tk);
      if (j < s) {
        R* np = cur;
        cur = out;
        prev = np;
        spare = (out == spare) ? nullptr : spare;
      }
    }
  }
  (void)p;
  return smoothed_residual<R>(c, l, grd, blk, fp32, l1_sum);
}

// smoother_vis (PS/smoother.f90:454-580) on the field set `vs` of level l
static int smoother_vis(Context& c, int l, VisSet& vs, int interpolate) {
  Level& L = c.L[l];
  const LevelGeom& g = L.g;
  const Tables& t = c.T;
  auto bcs = [&]() {
    int rows = (g.m + 7) * 3;
    hipLaunchKernelGGL(k_vis_bcs_side, dim3((rows + 255) / 256), dim3(256), 0, c.stream, vs.vis, g.n, g.m, g.pitch);
    int cols = (g.n + 1) * 3;
    hipLaunchKernelGGL(k_vis_bcs_tb, dim3((cols + 255) / 256), dim3(256), 0, c.stream, vs.vis, g.n, g.m, g.pitch);
  };
  bcs();
  int option, Niter;
  double lb, la;
  if (interpolate == 0) { option = 1; Niter = 3; lb = -3.0; la = -2.0; }
  else { option = 0; Niter = c.P.vis_iter; lb = -2.0; la = -1.0; }
  int Nvis_store = 0, nn_store = 1, n = 0;
  bool trigger = false;
  double ratio_min = 1.0e100;
  dim3 b16(64, 4);
  dim3 gnodes = grid2(g.n + 1, g.m + 1, b16);
  dim3 gint = grid2(g.n - 1, g.m - 1, b16);
  const int ivis0 = c.P.ivis == 0;
  // dt_stream for the level's own RKC1 stage count (the mixed path's coarse
  // levels may use fewer stages than nstageS)
  const LevelRkc rk = level_rkc(c, l, c.P.precision == PT_PREC_MIXED);
  double ae = -(rk.stability / 2.0), be = (double)rk.s;
  double ctemp = c.P.courant_stream * (rk.stability / (ae * ae));
  if (c.P.RKC1 == 0) {
    // RK1 (PS/iterative_solver.f90:1173-1186): dt = -2 courant eR/(eR^2+eI^2),
    // k_test_eig's RKC1 form with unit ellipse axes (eR/1 and eI/1 are exact)
    ae = 1.0; be = 1.0;
    ctemp = -(c.P.courant_stream * 2.0);
  }
  for (;;) {
    for (int nn = nn_store; nn <= Niter; nn++) {
      hipLaunchKernelGGL(k_vis_grad, gnodes, b16, 0, c.stream, vs, g.n, g.m, g.pitch, Nvis_store >= 1 ? 1 : 0, ivis0,
                         g.dr, g.ds, g.dr2, g.ds2, g.drds, t.dt_vis[l]);
      if (interpolate == 0 && (Nvis_store >= 1 || ivis0)) {
        hipLaunchKernelGGL(k_test_eig, gint, b16, 0, c.stream, vs, L.ratio, L.dts, g.n, g.m, g.pitch, t.SR1[l],
                           t.SR2[l], t.SR4[l], t.SR5[l], t.SR6[l], t.SI1[l], t.SI2[l], t.SI4[l], t.SI5[l], t.SI6[l],
                           ae, be, ctemp, c.red);
        int nb = gint.x * gint.y;
        hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, c.stream, c.red, nb, (int)RED_MIN, c.red_out);
        PT_HIP(hipMemcpyAsync(c.h_red, c.red_out, sizeof(double), hipMemcpyDeviceToHost, c.stream));
        PT_HIP(hipStreamSynchronize(c.stream));
        // the reference's loop starts from 1e100 and keeps it where every
        // ratio is +Inf (eigI = 0, isoviscous): PS/smoother.f90:501-506
        ratio_min = std::min(1.0e100, c.h_red[0]);
      }
      if (option == 1) {
        if ((Nvis_store >= 1 && ratio_min > c.P.eig_ratio_min) || ivis0 || n > 20) {
          c.eig_fine = ratio_min;   // module variables eig_fine, n_eta: last call wins
          c.n_eta = n;
          trigger = true;
          break;
        }
      } else if (interpolate >= 1 && (Nvis_store >= 1 || ivis0)) {
        trigger = true;
        break;
      }
      int ntemp = Niter - nn + 1;
      double cc = 2.0 / (-lb - la + (lb - la) * std::cos((double)(2 * ntemp - 1) * 3.1415926535897932 / (double)(2 * Niter)));
      cc = 0.99 * cc;
      hipLaunchKernelGGL(k_vis_update, gnodes, b16, 0, c.stream, vs.vis, vs.vsmooth, g.n, g.m, g.pitch, cc);
      bcs();
    }
    if (trigger) break;
    Nvis_store++;
    nn_store = 1;
    if (option == 1) { Niter = 6; lb = -2.0; la = -1.0; }
    n++;
  }
  return PT_OK;
}

// compute_viscosity (Newtonian, static=1): PS/iterative_solver.f90:560-634.
// The reference saves and restores the finer level's fields around the
// restriction; here the finer level is smoothed in its scratch set instead.
static int compute_viscosity(Context& c, int l) {
  Level& L = c.L[l];
  const LevelGeom& g = L.g;
  dim3 b16(64, 4);
  size_t bytes = sizeof(double) * L.nelem;
  if (l == 1) {
    const double lvT = std::log(c.P.visT), lvC = std::log(c.P.visC), lvP = std::log(c.P.visP);
    hipLaunchKernelGGL(k_visf, grid2(g.n + 1, g.m + 1, b16), b16, 0, c.stream, c.Tvis, c.Cvis, L.vis_static, g.n, g.m,
                       g.pitch, g.n + 11, c.P.ivis, lvT, lvC, lvP, c.T.ds, c.P.RaT != 0.0 ? 1 : 0);
    PT_HIP(hipMemcpyAsync(L.v.vis, L.vis_static, bytes, hipMemcpyDeviceToDevice, c.stream));
    return smoother_vis(c, 1, L.v, 0);
  }
  Level& F = c.L[l - 1];
  size_t fbytes = sizeof(double) * F.nelem;
  PT_HIP(hipMemcpyAsync(F.vs.vis, F.vis_static, fbytes, hipMemcpyDeviceToDevice, c.stream));
  PT_TRY(smoother_vis(c, l - 1, F.vs, 1));
  hipLaunchKernelGGL((k_restrict<double, double>), grid2(g.n - 1, g.m - 1, b16), b16, 0, c.stream, F.vs.vis,
                     F.g.pitch, L.v.vis, g.n, g.m, g.pitch);
  int tmax = std::max(g.n + 1, g.m - 1);
  hipLaunchKernelGGL(k_vis_bcs_init_boundary, dim3((tmax + 255) / 256), dim3(256), 0, c.stream, L.v.vis, F.vs.vis,
                     g.n, g.m, g.pitch, F.g.n, F.g.m, F.g.pitch);
  int ts = (g.m + 1) * 3;
  hipLaunchKernelGGL(k_vis_bcs_init_side, dim3((ts + 255) / 256), dim3(256), 0, c.stream, L.v.vis, g.n, g.m, g.pitch);
  int tt = (g.n + 7) * 3;
  hipLaunchKernelGGL(k_vis_bcs_init_tb, dim3((tt + 255) / 256), dim3(256), 0, c.stream, L.v.vis, g.n, g.m, g.pitch);
  PT_HIP(hipMemcpyAsync(L.vis_static, L.v.vis, bytes, hipMemcpyDeviceToDevice, c.stream));
  return smoother_vis(c, l, L.v, 0);
}

int launch_equilibrate(Context& c, double* ab, double* row, int n, int kl, int ku) {
  hipLaunchKernelGGL(k_coarse_equilibrate, dim3((n + 127) / 128), dim3(128), 0, c.stream, ab, row, n, kl, ku);
  return PT_OK;
}

static int coarse_factor(Context& c) {
  const Tables& t = c.T;
  Level& G = c.L[t.grids];
  CoarseArgs a;
  a.vis = G.v.vis; a.vxm = G.v.vxxmzz; a.vxz = G.v.vxz; a.vx = G.v.vx; a.vz = G.v.vz;
  a.nxc = G.g.n; a.nzc = G.g.m; a.p = G.g.pitch; a.kl = t.kl; a.ku = t.ku; a.ngc = t.ngc;
  a.ivis0 = c.P.ivis == 0; a.Vbc = c.P.Vbc;
  std::memcpy(a.SM1, t.SM1, sizeof a.SM1); std::memcpy(a.SM2, t.SM2, sizeof a.SM2);
  std::memcpy(a.SM4, t.SM4, sizeof a.SM4); std::memcpy(a.SM5, t.SM5, sizeof a.SM5);
  std::memcpy(a.SM6, t.SM6, sizeof a.SM6);
  size_t mbytes = sizeof(double) * (size_t)(2 * t.kl + t.ku + 1) * t.ngc;
  PT_HIP(hipMemsetAsync(c.Mc, 0, mbytes, c.stream));
  hipLaunchKernelGGL(k_coarse_assemble, dim3((t.ngc + 127) / 128), dim3(128), 0, c.stream, a, c.Mc, c.RowB);
  hipLaunchKernelGGL(k_coarse_equilibrate, dim3((t.ngc + 127) / 128), dim3(128), 0, c.stream, c.Mc, c.RowB, t.ngc,
                     t.kl, t.ku);
  hipLaunchKernelGGL(k_band_lu, dim3(1), dim3(64), 0, c.stream, c.Mc, c.ipiv, t.ngc, t.kl, t.ku, c.d_info);
  int32_t info = 0;
  PT_HIP(hipMemcpyAsync(&info, c.d_info, sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
  PT_HIP(hipStreamSynchronize(c.stream));
  if (info != 0) {
    set_error("coarse-grid band LU: zero pivot at column " + std::to_string(info));
    return PT_ERR_SINGULAR_COARSE;
  }
  hipLaunchKernelGGL(k_coarse_inverse, dim3(c.nci), dim3(64), sizeof(double) * t.ngc, c.stream, c.Mc, c.ipiv,
                     c.RowB, c.Gc, t.ngc, t.kl, t.ku, G.g.n, G.g.m);
  return PT_OK;
}

int compute_rhs1(Context& c) {
  Level& L = c.L[1];
  dim3 b16(64, 4);
  const double* D1 = c.T.D1;
  hipLaunchKernelGGL(k_rhs1, grid2(c.T.nx - 1, c.T.nz - 1, b16), b16, 0, c.stream, c.Tbuoy, c.Cbuoy, L.rhs, c.T.nx,
                     c.T.nz, L.g.pitch, c.T.nx + 11, c.T.dr, c.P.RaT, c.P.RaC, D1[1], D1[2], D1[3], D1[4], D1[5]);
  return PT_OK;
}

// FP64 copies of the FP32 mirrors (mixed precision): coefficients per solve
__global__ void k_to_f32(const double* a, float* b, long n) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = (float)a[t];
}
__global__ void k_add_f32_to_f64(double* psi, const float* e, int n, int m, int p, double sz) {
  int i = 1 + blockIdx.x * blockDim.x + threadIdx.x, k = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i > n - 1 || k > m - 1) return;
  double y = psi[IDX(i, k, p)] + (double)e[IDX(i, k, p)];
  store_with_ghosts<double>(psi, i, k, n, m, p, y, sz);
}

static int to_f32(Context& c, const double* a, float* b, size_t n) {
  hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, a, b, (long)n);
  return PT_OK;
}

__global__ void k_fill_f32(float* v, long n) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) v[t] = 0.0f;
}

// schedule entry p of the cycle (PS/iterative_solver.f90:143-340): the
// restriction into its level when the schedule descends, the visit (RKC1
// smoothing + residual, or the coarse solve), and the prolongation when it
// ascends next.  Kernel launches only (so a captured cycle holds kernel nodes
// and nothing else): a level-1 visit leaves the residual's L1 norm in
// device memory (red_out), which level1_residual copies to the host.
static int visit(Context& c, int p, bool mixed) {
  const Tables& t = c.T;
  const int* multi = c.cyc;
  dim3 b16(64, 4);
  const double sz = c.P.Vbc == 0 ? -1.0 : 1.0;
  const int l = multi[p];
  Level& L = c.L[l];
  if (multi[p - 1] < multi[p] && l >= 2 && mixed && c.coarse_zero && l < t.grids)
    hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)((L.nelem + 255) / 256)), dim3(256), 0, c.stream, L.ferr,
                       (long)L.nelem);
  if (multi[p - 1] < multi[p] && l >= 2) {
    // compute_RHS(l) = restrict(residual(l-1))
    Level& F = c.L[l - 1];
    if (!mixed)
      hipLaunchKernelGGL((k_restrict<double, double>), grid2(L.g.n - 1, L.g.m - 1, b16), b16, 0, c.stream, F.res,
                         F.g.pitch, L.rhs, L.g.n, L.g.m, L.g.pitch);
    else if (l == 2)
      hipLaunchKernelGGL((k_restrict<double, float>), grid2(L.g.n - 1, L.g.m - 1, b16), b16, 0, c.stream, F.res,
                         F.g.pitch, L.frhs, L.g.n, L.g.m, L.g.pitch);
    else
      hipLaunchKernelGGL((k_restrict<float, float>), grid2(L.g.n - 1, L.g.m - 1, b16), b16, 0, c.stream, F.fB,
                         F.g.pitch, L.frhs, L.g.n, L.g.m, L.g.pitch);
  }
  if (l < t.grids) {
    if (!mixed) {
      PT_TRY(iterate_stream_t<double>(c, l, p, false, l == 1));
    } else if (l == 1) {
      // FP64 defect -> FP32 RKC1 smoothing of the correction -> FP64 update
      StageArgs<double> a = stage_base<double>(c, 1, false);
      a.Sx = L.err;
      dim3 blk(TX, TZ), grd((L.g.n - 1 + TX - 1) / TX, (L.g.m - 1 + TZ * PZ - 1) / (TZ * PZ));
      launch_residual<double>(c, grd, blk, a, c.dres64, c.red);
      PT_TRY(to_f32(c, c.dres64, L.frhs, L.nelem));
      hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)((L.nelem + 255) / 256)), dim3(256), 0, c.stream, L.ferr,
                         (long)L.nelem);
      PT_TRY(iterate_stream_t<float>(c, 1, p, true, false));
      hipLaunchKernelGGL(k_add_f32_to_f64, grid2(L.g.n - 1, L.g.m - 1, b16), b16, 0, c.stream, L.err, L.ferr,
                         L.g.n, L.g.m, L.g.pitch, sz);
      launch_residual<double>(c, grd, blk, a, L.res, c.red);
      int nb = grd.x * grd.y;
      hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, c.stream, c.red, nb, (int)RED_SUM_ABS, c.red_out);
    } else if (l >= c.presmooth_from && c.coarse_zero && multi[p - 1] < multi[p] && multi[p + 1] > multi[p]) {
      // no smoothing on the way down: a level entered from the finer one
      // and left for the coarser one starts its correction from zero, so
      // the residual restricted next is the level's right-hand side (the
      // turning visits of F- and W-cycles, entered from a coarser level,
      // carry a prolonged correction and are smoothed)
      StageArgs<float> a = stage_base<float>(c, l, true);
      a.Sx = L.ferr;
      dim3 blk(TX, TZ), grd((L.g.n - 1 + TX - 1) / TX, (L.g.m - 1 + TZ * PZ - 1) / (TZ * PZ));
      launch_residual<float>(c, grd, blk, a, L.fB, c.red);
    } else {
      PT_TRY(iterate_stream_t<float>(c, l, p, true, false));
    }
  } else {
    // Compute_coarse_vector + DGBTRS + Stream_grid_coarse + enforceBCs_stream
    // as one dense matvec with the operator built from the LU factors
    dim3 gb((c.nci + 255) / 256);
    size_t sh = sizeof(double) * c.nci;
    if (!mixed)
      hipLaunchKernelGGL((k_coarse_apply<double, double>), gb, dim3(256), sh, c.stream, c.Gc, L.rhs, L.err, L.g.n,
                         L.g.m, L.g.pitch, sz);
    else
      PT_TRY(launch_coarse_apply_wg(c, L.frhs, L.ferr, L.g.n, L.g.m, L.g.pitch, (float)sz));
  }
  if (l > 1 && multi[p + 1] < multi[p]) {
    Level& F = c.L[l - 1];
    dim3 gp = grid2(F.g.n - 1, F.g.m - 1, b16);
    if (!mixed)
      hipLaunchKernelGGL((k_prolong<double, double>), gp, b16, 0, c.stream, L.err, L.g.pitch, F.err, F.g.n, F.g.m,
                         F.g.pitch, sz);
    else if (l - 1 == 1)
      hipLaunchKernelGGL((k_prolong<float, double>), gp, b16, 0, c.stream, L.ferr, L.g.pitch, F.err, F.g.n,
                         F.g.m, F.g.pitch, sz);
    else
      hipLaunchKernelGGL((k_prolong<float, float>), gp, b16, 0, c.stream, L.ferr, L.g.pitch, F.ferr, F.g.n,
                         F.g.m, F.g.pitch, (float)sz);
  }
  return PT_OK;
}

// FP64 residual RHS - L psi of a level-1 iterate (compute_residual's operator
// with the level's current viscosity set) into res, and sum |res| on the host
int level1_residual_of(Context& c, const double* psi, double* res, double* l1) {
  Level& L = c.L[1];
  StageArgs<double> a = stage_base<double>(c, 1, false);
  a.Sx = psi;
  dim3 blk(TX, TZ), grd((L.g.n - 1 + TX - 1) / TX, (L.g.m - 1 + TZ * PZ - 1) / (TZ * PZ));
  launch_residual<double>(c, grd, blk, a, res, c.red);
  const int nb = grd.x * grd.y;
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, c.stream, c.red, nb, (int)RED_SUM_ABS, c.red_out);
  PT_HIP(hipMemcpyAsync(c.h_red, c.red_out, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  PT_HIP(hipStreamSynchronize(c.stream));
  *l1 = c.h_red[0];
  return PT_OK;
}

// residual_mag(1) of the level-1 visit just enqueued, and the abnormal-
// residual stop (PS/iterative_solver.f90:438-447)
static int level1_residual(Context& c, double* resmag1) {
  PT_HIP(hipMemcpyAsync(c.h_red, c.red_out, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  PT_HIP(hipStreamSynchronize(c.stream));
  const LevelGeom& g = c.L[1].g;
  *resmag1 = c.h_red[0] / (double)(g.n - 2) / (double)(g.m - 2);
  if (*resmag1 > 1.0e10 * c.res_ref1) {
    set_error("residual " + std::to_string(*resmag1) + " > 1e10 * res_ref");
    return PT_ERR_HIGH_RESIDUAL;
  }
  return PT_OK;
}

// The level buffers whose roles rotate: a strip sweep swaps its level's
// error buffer with a work buffer, so a cycle may start from one of a few
// arrangements, and a replay must leave the pointers as the eager launches
// would have.
static void buffer_roles(Context& c, std::vector<void*>& out) {
  out.clear();
  for (int l = 1; l <= c.T.grids; l++) {
    Level& L = c.L[l];
    out.insert(out.end(), {(void*)L.ferr, (void*)L.fB, (void*)L.fC, (void*)L.fD, (void*)L.fE});
  }
}
static void restore_roles(Context& c, const std::vector<void*>& v) {
  size_t j = 0;
  for (int l = 1; l <= c.T.grids; l++) {
    Level& L = c.L[l];
    float** f[5] = {&L.ferr, &L.fB, &L.fC, &L.fD, &L.fE};
    for (float** q : f) *q = (float*)v[j++];
  }
}

// visits 2..nsteps of a cycle and visit 1 of the next.  After one eager run
// the sequence is captured as a HIP graph (kernel nodes only) per starting
// arrangement of the rotating buffers (at most MAX_CYCLE_GRAPHS) and
// replayed; the graphs' kernel arguments (buffers, level constants) do not
// change between cycles or solves of a context.  PT_GRAPH=0 at pt_create,
// per-kernel timing, a failed capture or more arrangements keep the launches
// eager; the results are the same either way.
static constexpr size_t MAX_CYCLE_GRAPHS = 4;
static int rest_and_next_level1(Context& c, bool mixed) {
  auto enqueue = [&]() -> int {
    for (int p = 2; p <= c.cyc_n; p++) PT_TRY(visit(c, p, mixed));
    return visit(c, 1, mixed);
  };
  if (!c.use_graph || c.timing || c.graph_state < 0) return enqueue();
  if (c.graph_state == 0) {
    c.graph_state = 1;   // the first run is eager (sets the kernels' attributes)
    return enqueue();
  }
  c.graph_state = 2;
  std::vector<void*> before;
  buffer_roles(c, before);
  for (const Context::CycleGraph& cg : c.cycle_graphs) {
    if (cg.before != before) continue;
    PT_HIP(hipGraphLaunch(cg.exec, c.stream));
    restore_roles(c, cg.after);
    return PT_OK;
  }
  if (c.cycle_graphs.size() >= MAX_CYCLE_GRAPHS) {
    c.graph_state = -1;
    return enqueue();
  }
  // capture: the launches are recorded, not run, while the host-side pointer
  // swaps take effect; the replay below then runs them
  Context::CycleGraph cg;
  cg.before = before;
  PT_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
  const int rc = enqueue();
  hipError_t e = hipStreamEndCapture(c.stream, &cg.graph);
  if (rc != PT_OK || e != hipSuccess) {
    if (cg.graph) (void)hipGraphDestroy(cg.graph);
    restore_roles(c, before);
    c.graph_state = -1;
    if (rc != PT_OK) return rc;
    return enqueue();
  }
  buffer_roles(c, cg.after);
  e = hipGraphInstantiate(&cg.exec, cg.graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    // nothing of the capture ran: back to the starting arrangement, eager from now on
    (void)hipGraphDestroy(cg.graph);
    restore_roles(c, before);
    c.graph_state = -1;
    return enqueue();
  }
  c.cycle_graphs.push_back(std::move(cg));
  PT_HIP(hipGraphLaunch(c.cycle_graphs.back().exec, c.stream));
  return PT_OK;
}

// multigrid: PS/iterative_solver.f90:124-341
static int multigrid_body(Context& c);
int multigrid(Context& c) {
  if (c.timing) timer_begin(c, PT_K_SOLVE);
  const int rc = multigrid_body(c);
  if (c.timing) timer_end(c, PT_K_SOLVE);
  return rc;
}

static int multigrid_body(Context& c) {
  const Tables& t = c.T;
  const bool mixed = c.P.precision == PT_PREC_MIXED;
  Level& L1 = c.L[1];
  size_t b1 = sizeof(double) * L1.nelem;
  PT_HIP(hipMemcpyAsync(L1.err, c.SF, b1, hipMemcpyDeviceToDevice, c.stream));
  PT_TRY(compute_rhs1(c));
  double rr;
  PT_TRY(reduce_to_host(c, L1.rhs, t.nx - 1, t.nz - 1, L1.g.pitch, 4, RED_SUM_ABS, &rr));
  c.res_ref1 = rr / (double)(t.nx - 2) / (double)(t.nz - 2);
  for (int l = 1; l <= t.grids; l++) PT_TRY(compute_viscosity(c, l));
  if (c.first_solve || c.P.ivis > 0) {
    PT_TRY(coarse_factor(c));
    c.first_solve = false;
  }
  if (mixed) {
    for (int l = 1; l <= t.grids; l++) {
      Level& L = c.L[l];
      PT_TRY(to_f32(c, L.v.vis, L.fvis, L.nelem)); PT_TRY(to_f32(c, L.v.vxxmzz, L.fvxm, L.nelem));
      PT_TRY(to_f32(c, L.v.vxz, L.fvxz, L.nelem)); PT_TRY(to_f32(c, L.v.vx, L.fvx, L.nelem));
      PT_TRY(to_f32(c, L.v.vz, L.fvz, L.nelem)); PT_TRY(to_f32(c, L.dts, L.fdts, L.nelem));
      // coarse corrections of the mixed path start each solve from zero
      if (l >= 2)
        hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)((L.nelem + 255) / 256)), dim3(256), 0, c.stream, L.ferr,
                           (long)L.nelem);
    }
  }
  if (c.guess_net) PT_TRY(learned_before_solve(c));
  int q = 1;
  double resmag1 = 0.0;
  // cycle q: visit 1 (level 1, ends with its residual's L1 norm on the host),
  // then visits 2..nsteps; the stop test follows every level-1 visit
  PT_TRY(visit(c, 1, mixed));
  int rc = level1_residual(c, &resmag1);
  if (rc != PT_OK) { c.last_q = q; return rc; }
  while (!((resmag1 / c.res_ref1) <= c.P.tolerance)) {
    if (q + 1 > c.P.max_cycles) {
      for (int p = 2; p <= c.cyc_n; p++) PT_TRY(visit(c, p, mixed));
      c.cycles++;
      set_error("q > max_cycles = " + std::to_string(c.P.max_cycles));
      return PT_ERR_MAX_CYCLES;
    }
    PT_TRY(rest_and_next_level1(c, mixed));
    c.cycles++;
    q++;
    rc = level1_residual(c, &resmag1);
    if (rc != PT_OK) { c.last_q = q; return rc; }   // the cycle the reference stops at
  }
  c.cycles++;
  if (c.guess_net) PT_TRY(learned_after_solve(c));
  PT_HIP(hipMemcpyAsync(c.SF, L1.err, b1, hipMemcpyDeviceToDevice, c.stream));
  c.last_q = q;
  c.residual1 = resmag1;
  c.solves++;
  return PT_OK;
}

int stream_bcs(Context& c, double* S) {
  dim3 b16(64, 4);
  hipLaunchKernelGGL(k_stream_bcs, grid2(c.T.nx + 7, c.T.nz + 7, b16), b16, 0, c.stream, S, c.T.nx, c.T.nz,
                     c.T.nx + 7, c.P.Vbc == 0 ? -1.0 : 1.0);
  return PT_OK;
}

int compute_velocities(Context& c) {
  dim3 b16(64, 4);
  const double* D1 = c.T.D1;
  int p = c.T.nx + 7;
  hipLaunchKernelGGL(k_velocities, grid2(c.T.nx + 1, c.T.nz + 1, b16), b16, 0, c.stream, c.SF, c.u, c.w, c.T.nx,
                     c.T.nz, p, c.T.dr, c.T.ds, D1[1], D1[2], D1[3], D1[4], D1[5]);
  hipLaunchKernelGGL(k_velocity_bcs, grid2(c.T.nx + 7, c.T.nz + 7, b16), b16, 0, c.stream, c.u, c.w, c.T.nx, c.T.nz,
                     p, c.P.Vbc);
  return PT_OK;
}

// ---------------------------------------------------------------------------
// allocation
// ---------------------------------------------------------------------------
static int dalloc(void** p, size_t bytes) {
  PT_HIP(hipMalloc(p, bytes));
  PT_HIP(hipMemset(*p, 0, bytes));
  return PT_OK;
}
#define DALLOC(ptr, bytes) PT_TRY(dalloc((void**)&(ptr), (bytes)))

int solver_alloc(Context& c) {
  const Tables& t = c.T;
  for (int l = 1; l <= t.grids; l++) {
    Level& L = c.L[l];
    LevelGeom& g = L.g;
    g.n = t.nxg[l]; g.m = t.nzg[l]; g.pitch = g.n + 7;
    g.dr = t.drg[l]; g.ds = t.dsg[l]; g.dr2 = t.dr2g[l]; g.ds2 = t.ds2g[l]; g.dr3 = t.dr3g[l]; g.ds3 = t.ds3g[l];
    g.dr4 = t.dr4g[l]; g.ds4 = t.ds4g[l]; g.drds = t.drdsg[l]; g.drds2 = t.drds2g[l]; g.dr2ds = t.dr2dsg[l];
    g.dr2ds2 = t.dr2ds2g[l];
    L.nelem = (size_t)(g.n + 7) * (g.m + 7);
    size_t b = sizeof(double) * L.nelem, f = sizeof(float) * L.nelem;
    DALLOC(L.err, b); DALLOC(L.bufB, b); DALLOC(L.bufC, b); DALLOC(L.rhs, b); DALLOC(L.res, b);
    DALLOC(L.dts, b); DALLOC(L.ratio, b); DALLOC(L.vis_static, b);
    VisSet* sets[2] = {&L.v, &L.vs};
    for (VisSet* s : sets) {
      DALLOC(s->vis, b); DALLOC(s->vx, b); DALLOC(s->vz, b); DALLOC(s->vxx, b); DALLOC(s->vzz, b);
      DALLOC(s->vxz, b); DALLOC(s->vxxpzz, b); DALLOC(s->vxxmzz, b); DALLOC(s->vsmooth, b);
    }
    if (c.P.precision == PT_PREC_MIXED) {
      DALLOC(L.ferr, f); DALLOC(L.fB, f); DALLOC(L.fC, f); DALLOC(L.fD, f); DALLOC(L.fE, f); DALLOC(L.frhs, f);
      DALLOC(L.fdts, f); DALLOC(L.fvis, f); DALLOC(L.fvxm, f); DALLOC(L.fvxz, f); DALLOC(L.fvx, f); DALLOC(L.fvz, f);
    }
  }
  if (c.P.precision == PT_PREC_MIXED) DALLOC(c.dres64, sizeof(double) * c.L[1].nelem);
  DALLOC(c.d_info, sizeof(int32_t));
  if (c.P.iterate == 1) {   // the coarsest-grid operator of the multigrid path
    size_t mb = sizeof(double) * (size_t)(2 * t.kl + t.ku + 1) * t.ngc;
    if (t.ngc > 4096) {
      set_error("coarsest grid has more than 4096 unknowns");
      return PT_ERR_UNSUPPORTED;
    }
    DALLOC(c.Mc, mb); DALLOC(c.RowB, sizeof(double) * t.ngc); DALLOC(c.Bc, sizeof(double) * t.ngc);
    c.nci = (t.nxg[t.grids] - 1) * (t.nzg[t.grids] - 1);
    DALLOC(c.Gc, sizeof(double) * (size_t)c.nci * c.nci);
    DALLOC(c.ipiv, sizeof(int32_t) * t.ngc);
  }
  c.red_cap = 1 << 20;
  DALLOC(c.red, sizeof(double) * c.red_cap);
  DALLOC(c.red_out, sizeof(double) * 16);
  PT_HIP(hipHostMalloc((void**)&c.h_red, sizeof(double) * 16, hipHostMallocDefault));
  // stencil constants
  for (int q = 0; q < 5; q++) { c.st.D1[q] = t.D1[q + 1]; c.st.D2[q] = t.D2[q + 1]; }
  for (int q = 0; q < 7; q++) { c.st.D3[q] = t.D3[q]; c.st.D4[q] = t.D4[q]; }
  for (int b = 0; b < 5; b++)
    for (int q = 0; q < 5; q++) {
      c.st.Drs[b][q] = t.Drs[b + 1][q + 1]; c.st.Drss[b][q] = t.Drss[b + 1][q + 1];
      c.st.Drrs[b][q] = t.Drrs[b + 1][q + 1]; c.st.Drrss[b][q] = t.Drrss[b + 1][q + 1];
    }
  return PT_OK;
}

int solver_free(Context& c) {
  for (int l = 1; l <= c.T.grids; l++) {
    Level& L = c.L[l];
    void* ps[] = {L.err, L.bufB, L.bufC, L.rhs, L.res, L.dts, L.ratio, L.vis_static, L.v.vis, L.v.vx, L.v.vz,
                  L.v.vxx, L.v.vzz, L.v.vxz, L.v.vxxpzz, L.v.vxxmzz, L.v.vsmooth, L.vs.vis, L.vs.vx, L.vs.vz,
                  L.vs.vxx, L.vs.vzz, L.vs.vxz, L.vs.vxxpzz, L.vs.vxxmzz, L.vs.vsmooth, L.ferr, L.fB, L.fC, L.fD, L.fE,
                  L.frhs, L.fdts, L.fvis, L.fvxm, L.fvxz, L.fvx, L.fvz};
    for (void* p : ps)
      if (p) (void)hipFree(p);
  }
  void* ps[] = {c.Mc, c.RowB, c.Bc, c.ipiv, c.d_info, c.red, c.red_out, c.dres64, c.Gc, c.lg_res, c.lg_res2, c.lg_psi, c.stats_rows,
                c.strain_buf};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  if (c.h_red) (void)hipHostFree(c.h_red);
  for (Context::CycleGraph& cg : c.cycle_graphs) {
    if (cg.exec) (void)hipGraphExecDestroy(cg.exec);
    if (cg.graph) (void)hipGraphDestroy(cg.graph);
  }
  c.cycle_graphs.clear();
  return PT_OK;
}

}  // namespace pt
