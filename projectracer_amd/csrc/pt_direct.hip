// Direct stream-function solver (iterate=0): PS/momentum.f90:172-653.
//
// build_SF_matrix assembles the banded operator of the whole grid (one row per
// node, Database 1 or 2 ordering, LDAB = 2 kl + ku + 1), scales its rows by
// DGBEQU's factors and factors it with partial pivoting (DGBTF2's column
// order, so the factors match the unblocked LAPACK routine bit for bit);
// solve_SF_equation forms the buoyancy vector, runs DGBTRS and scatters the
// solution into psi with its ghost images (Stream_grid).
//
// Device layout: AB column-major, A(p,q) at AB[(q-1)*LDAB + kl+ku+p-q]; the
// factorization is one workgroup of 1024 threads walking the columns; the
// solve is one workgroup whose band window lives in an LDS ring.
#include "pt_internal.hpp"

#include <algorithm>
#include <cstring>

namespace pt {

// ---------------------------------------------------------------------------
// assembly
// ---------------------------------------------------------------------------
struct SfArgs {
  double D1[7], D2[7], D3[7], D4[7];                        // offsets -3..3
  double Drs[7][7], Drss[7][7], Drrs[7][7], Drrss[7][7];    // [k+3][i+3] = D?(i,k)
  double dr, ds, dr2, ds2, dr3, ds3, dr4, ds4, drds, drds2, dr2ds, dr2ds2;
  int nx, nz, Vbc, Database, kl, ku, n, pitch;              // pitch of the vis array (nx+7)
};

// pf / indices (PS/momentum.f90:221-253)
__device__ __forceinline__ int sf_pf(const SfArgs& a, int ip, int kp) {
  return a.Database == 1 ? (a.nx + 1) * kp + ip + 1 : (a.nz + 1) * ip + kp + 1;
}
__device__ __forceinline__ void sf_indices(const SfArgs& a, int p, int& ip, int& kp) {
  if (a.Database == 1) {
    const int N = a.nx + 1;
    ip = p - N * ((p - 1) / N) - 1; kp = (p - 1) / N;
  } else {
    const int N = a.nz + 1;
    kp = p - N * ((p - 1) / N) - 1; ip = (p - 1) / N;
  }
}

struct VisTerms {
  double vis, visx, visz, visxx, viszz, visxz;
};

// Matrix_element (PS/momentum.f90:507-606) for the stencil point (i,k) of row
// (ip,kp), with the uniform-grid metrics xr = zs = 1 and their derivatives 0
// kept in the expressions so that signs and roundings follow the reference;
// `m` is the band entry, updated in place
__device__ __forceinline__ void sf_element(const SfArgs& a, int i, int k, int ip, int kp, int rsym, int ssym,
                                           const VisTerms& v, double& m) {
  const double xr = 1.0, xrr = 0.0, xrrr = 0.0, xrrrr = 0.0, zs = 1.0, zss = 0.0, zsss = 0.0, zssss = 0.0;
  double Ss = 0.0, Sss = 0.0, Ssss = 0.0, Sssss = 0.0, Sr = 0.0, Srr = 0.0, Srrr = 0.0, Srrrr = 0.0;
  if (i == 0) {
    Ss = a.D1[k + 3] / a.ds; Sss = a.D2[k + 3] / a.ds2; Ssss = a.D3[k + 3] / a.ds3; Sssss = a.D4[k + 3] / a.ds4;
  }
  if (ssym == 1 && a.Vbc == 0) { Ss = -Ss; Sss = -Sss; Ssss = -Ssss; Sssss = -Sssss; }
  if (k == 0) {
    Sr = a.D1[i + 3] / a.dr; Srr = a.D2[i + 3] / a.dr2; Srrr = a.D3[i + 3] / a.dr3; Srrrr = a.D4[i + 3] / a.dr4;
  }
  if (rsym == 1) { Sr = -Sr; Srr = -Srr; Srrr = -Srrr; Srrrr = -Srrrr; }
  const double Sx = Sr / xr;
  const double Sxx = (Srr - Sx * xrr) / (xr * xr);
  const double Sxxx = (Srrr - 3.0 * Sxx * xr * xrr - Sx * xrrr) / (xr * xr * xr);
  const double Sxxxx =
      (Srrrr - 6.0 * Sxxx * xrr * (xr * xr) - (3.0 * (xrr * xrr) + 4.0 * xr * xrrr) * Sxx - Sx * xrrrr) /
      ((xr * xr) * (xr * xr));
  const double Sz = Ss / zs;
  const double Szz = (Sss - Sz * zss) / (zs * zs);
  const double Szzz = (Ssss - 3.0 * Szz * zs * zss - Sz * zsss) / (zs * zs * zs);
  const double Szzzz =
      (Sssss - 6.0 * Szzz * zss * (zs * zs) - (3.0 * (zss * zss) + 4.0 * zs * zsss) * Szz - Sz * zssss) /
      ((zs * zs) * (zs * zs));
  double Srs = a.Drs[k + 3][i + 3] / a.drds, Srss = a.Drss[k + 3][i + 3] / a.drds2;
  double Srrs = a.Drrs[k + 3][i + 3] / a.dr2ds, Srrss = a.Drrss[k + 3][i + 3] / a.dr2ds2;
  const bool flip = (a.Vbc == 0) ? ((rsym == 1) != (ssym == 1)) : (rsym == 1);
  if (flip) { Srs = -Srs; Srss = -Srss; Srrs = -Srrs; Srrss = -Srrss; }
  const double Sxz = Srs / xr / zs;
  const double Sxxz = (Srrs - Sxz * zs * xrr) / zs / (xr * xr);
  const double Sxzz = (Srss - Sxz * xr * zss) / xr / (zs * zs);
  const double Sxxzz = (Srrss - Sxzz * xrr * (zs * zs) - Sxxz * zss * (xr * xr) - Sxz * xrr * zss) / ((xr * zs) * (xr * zs));
  const bool wall_x = (ip == 0 || ip == a.nx), wall_z = (kp == 0 || kp == a.nz);
  if (wall_x) {
    m = m + Sxx;
  } else if (wall_z) {
    if (a.Vbc == 0) m = m + Szz;
  } else {
    const double term1 = (v.visxx - v.viszz) * (Sxx - Szz);
    const double term2 = v.vis * (Sxxxx + Szzzz);
    const double term3 = 2.0 * v.vis * Sxxzz;
    const double term4 = 4.0 * v.visxz * Sxz;
    const double term5 = 2.0 * v.visx * (Sxxx + Sxzz);
    const double term6 = 2.0 * v.visz * (Szzz + Sxxz);
    m = m + term1 + term2 + term3 + term4 + term5 + term6;
  }
}

// one stencil point folded into the domain (PS/momentum.f90:383-474)
__device__ __forceinline__ void sf_point(const SfArgs& a, int i, int k, int ip, int kp, int p, const VisTerms& v,
                                         double* ab) {
  int rsym = 0, ssym = 0, iq = ip + i, kq = kp + k;
  if (iq < 0) { iq = -iq; rsym = 1; }
  else if (iq > a.nx) { iq = a.nx - (iq - a.nx); rsym = 1; }
  if (kq < 0) { kq = -kq; ssym = 1; }
  else if (kq > a.nz) { kq = a.nz - (kq - a.nz); ssym = 1; }
  const int q = sf_pf(a, iq, kq);
  const int ld = 2 * a.kl + a.ku + 1, kk1 = a.kl + a.ku + 1;
  double& m = ab[(long)(q - 1) * ld + (kk1 + p - q - 1)];
  sf_element(a, i, k, ip, kp, rsym, ssym, v, m);
}

// the rows of build_SF_matrix (PS/momentum.f90:362-478) into a zeroed AB; one
// thread per row, which owns every entry (p, q) of its row
__global__ void k_sf_assemble(SfArgs a, const double* vis, double* ab) {
  const int p = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (p > a.n) return;
  int ip, kp;
  sf_indices(a, p, ip, kp);
  const int ld = 2 * a.kl + a.ku + 1, kk1 = a.kl + a.ku + 1;
  const bool corner = (ip == 0 || ip == a.nx) && (kp == 0 || kp == a.nz);
  if (corner || (a.Vbc == 1 && (kp == 0 || kp == a.nz))) {
    ab[(long)(p - 1) * ld + (kk1 - 1)] = 1.0;
    return;
  }
  VisTerms v = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (ip > 0 && ip < a.nx && kp > 0 && kp < a.nz) {
    // second-order viscosity derivatives, D1_so = (-1/2, 0, 1/2), D2_so =
    // (1, -2, 1), Drs_so = D1_so x D1_so (PS/momentum.f90:369-381)
    const double d1[3] = {-0.5, 0.0, 0.5}, d2[3] = {1.0, -2.0, 1.0};
    const int P = a.pitch;
    auto V = [&](int i, int k) { return vis[(long)(k + H3) * P + (i + H3)]; };
    double visr = 0.0, visrr = 0.0, viss = 0.0, visss = 0.0, visrs = 0.0;
    for (int j = -1; j <= 1; j++) visr = visr + d1[j + 1] * V(ip + j, kp);
    for (int j = -1; j <= 1; j++) visrr = visrr + d2[j + 1] * V(ip + j, kp);
    for (int j = -1; j <= 1; j++) viss = viss + d1[j + 1] * V(ip, kp + j);
    for (int j = -1; j <= 1; j++) visss = visss + d2[j + 1] * V(ip, kp + j);
    for (int b = -1; b <= 1; b++)
      for (int q = -1; q <= 1; q++) visrs = visrs + (d1[q + 1] * d1[b + 1]) * V(ip + q, kp + b);
    visr = visr / a.dr; visrr = visrr / a.dr2; viss = viss / a.ds; visss = visss / a.ds2; visrs = visrs / a.drds;
    v.vis = V(ip, kp);
    v.visx = visr / 1.0;
    v.visz = viss / 1.0;
    v.visxx = (visrr - v.visx * 0.0) / 1.0;
    v.viszz = (visss - v.visz * 0.0) / 1.0;
    v.visxz = visrs / 1.0 / 1.0;
  }
  for (int i = -3; i <= 3; i += 6) sf_point(a, i, 0, ip, kp, p, v, ab);
  for (int k = -3; k <= 3; k += 6) sf_point(a, 0, k, ip, kp, p, v, ab);
  for (int i = -2; i <= 2; i++)
    for (int k = -2; k <= 2; k++) sf_point(a, i, k, ip, kp, p, v, ab);
}

// ---------------------------------------------------------------------------
// DGBTF2 (band LU with partial pivoting, LAPACK's unblocked column order) in
// panels of nb columns.
//   k_lu_panel: one workgroup factors columns p0 .. p0+nb-1 held in LDS
//     (rows p0 .. p0+nb-1+kl): per column the first-maximum pivot search, the
//     row interchange, the scaling and the rank-1 update, each restricted to
//     the panel's columns.  It records ju after every step (the last column
//     DGBTF2's step touches; -1 for a zero pivot, which DGBTF2 skips).
//   k_lu_trailing: one wave per column beyond the panel replays the panel's
//     steps on that column in order -- the interchange, then
//     x = x + L(t,j) (-A(j,c)) -- while the column is within the step's ju.
// Every element sees the operations of DGBTF2 in the same order, so the
// factors are DGBTF2's bit for bit.  AB's fill rows start zero (memset) and
// are first written by the step that DGBTF2 zeroes them at, so that zeroing
// is not repeated.
// ---------------------------------------------------------------------------
constexpr int LU_PANEL_THREADS = 1024;   // upper bound; launched with roundup64(kl+1) threads
constexpr int LU_LDS_LIMIT = 160 * 1024 - 1024;   // dynamic LDS per workgroup (gfx950: 160 KB)
constexpr int LU_TRAIL_WAVES = 8;   // columns per trailing workgroup

__global__ void __launch_bounds__(LU_PANEL_THREADS) k_lu_panel(double* ab, int32_t* ipiv, int32_t* ju_rec,
                                                               int32_t* state, int n, int kl, int ku, int p0,
                                                               int nb) {
  extern __shared__ double P[];   // P[(c-p0)*W + (r-p0)], W = kl + nb
  __shared__ double s_best[LU_PANEL_THREADS / 64];
  __shared__ int s_idx[LU_PANEL_THREADS / 64];
  __shared__ int s_jp;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, NT = blockDim.x, nwv = NT >> 6;
  const long ld = 2 * kl + ku + 1;
  const int kv = ku + kl, W = kl + nb, pe = min(p0 + nb - 1, n), np = pe - p0 + 1;
  // A(r, c) = AB(kv+1+r-c, c): rows above c+kl are stored for c >= p0
  for (int e = tid; e < np * W; e += NT) {
    const int cc = e / W, r = e - cc * W, col = p0 + cc, row = p0 + r;
    P[e] = (row <= min(n, col + kl)) ? ab[(long)(col - 1) * ld + (kv + row - col)] : 0.0;
  }
  // ju and the first zero pivot carried between panels
  int ju = (p0 == 1) ? 1 : state[0];
  int inf = (p0 == 1) ? 0 : state[1];
  __syncthreads();
  auto A = [&](int r, int col) -> double& { return P[(col - p0) * W + (r - p0)]; };
  for (int j = p0; j <= pe; j++) {
    const int km = min(kl, n - j);
    double best = -1.0;
    int bi = 0x7fffffff;
    for (int t = tid; t <= km; t += NT) {
      const double x = fabs(A(j + t, j));
      if (x > best) { best = x; bi = t; }
    }
    for (int off = 32; off > 0; off >>= 1) {
      const double ob = __shfl_down(best, off);
      const int oi = __shfl_down(bi, off);
      if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
    }
    if (lane == 0) { s_best[wave] = best; s_idx[wave] = bi; }
    __syncthreads();
    if (tid == 0) {
      double b = s_best[0];
      int x = s_idx[0];
      for (int w = 1; w < nwv; w++)
        if (s_best[w] > b || (s_best[w] == b && s_idx[w] < x)) { b = s_best[w]; x = s_idx[w]; }
      s_jp = x + 1;
      ipiv[j - 1] = x + j;
    }
    __syncthreads();
    const int jp = s_jp;
    if (A(j + jp - 1, j) != 0.0) {
      ju = max(ju, min(j + ku + jp - 1, n));
      const int ce = min(ju, pe);
      if (jp != 1)
        for (int col = j + tid; col <= ce; col += NT) {
          const double tmp = A(j + jp - 1, col);
          A(j + jp - 1, col) = A(j, col);
          A(j, col) = tmp;
        }
      __syncthreads();
      if (km > 0) {
        // thread t owns row j+1+t of the panel (km <= kl < blockDim): the
        // scaled multiplier stays in a register for the rank-1 update
        const double rcp = 1.0 / A(j, j);
        const int t = 1 + tid;
        double lt = 0.0;
        if (t <= km) {
          lt = rcp * A(j + t, j);
          A(j + t, j) = lt;
        }
#pragma unroll 8
        for (int col = j + 1; col <= ce; col++) {
          const double temp = -1.0 * A(j, col);
          if (t <= km) A(j + t, col) = A(j + t, col) + lt * temp;
        }
      }
      if (tid == 0) ju_rec[j - 1] = ju;
    } else {
      if (inf == 0) inf = j;
      if (tid == 0) ju_rec[j - 1] = -1;
    }
    __syncthreads();
  }
  for (int e = tid; e < np * W; e += NT) {
    const int cc = e / W, r = e - cc * W, col = p0 + cc, row = p0 + r;
    if (row <= min(n, col + kl)) ab[(long)(col - 1) * ld + (kv + row - col)] = P[e];
  }
  if (tid == 0) { state[0] = ju; state[1] = inf; }
}

// wave-level ordering of LDS accesses between lanes of one wave
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ void __launch_bounds__(LU_TRAIL_WAVES * 64) k_lu_trailing(double* ab, const int32_t* ipiv,
                                                                     const int32_t* ju_rec, int n, int kl, int ku,
                                                                     int p0, int nb) {
  extern __shared__ double sh[];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long ld = 2 * kl + ku + 1;
  const int kv = ku + kl, W = kl + nb, pe = min(p0 + nb - 1, n), np = pe - p0 + 1;
  double* Ls = sh;                                   // Ls[s*kl + t-1] = L(p0+s+t, p0+s)
  double* Cb = sh + (size_t)nb * kl + (size_t)wave * W;   // this wave's column, rows p0 .. p0+W-1
  for (int e = tid; e < np * kl; e += LU_TRAIL_WAVES * 64) {
    const int sc = e / kl, t = 1 + (e - sc * kl), j = p0 + sc;
    Ls[e] = (t <= min(kl, n - j)) ? ab[(long)(j - 1) * ld + (kv + t)] : 0.0;
  }
  __syncthreads();
  const int c = pe + 1 + blockIdx.x * LU_TRAIL_WAVES + wave;
  if (c > min(n, pe + kv)) return;
  for (int r = lane; r < W; r += 64) {
    const int row = p0 + r;
    Cb[r] = (row >= c - kv && row <= min(n, c + kl)) ? ab[(long)(c - 1) * ld + (kv + row - c)] : 0.0;
  }
  wave_lds_sync();
  for (int sc = 0; sc < np; sc++) {
    const int j = p0 + sc, jut = ju_rec[j - 1];
    if (jut < 0 || c > jut) continue;
    const int l = ipiv[j - 1];
    if (l != j) {
      if (lane == 0) {
        const double tmp = Cb[l - p0];
        Cb[l - p0] = Cb[sc];
        Cb[sc] = tmp;
      }
      wave_lds_sync();
    }
    const double temp = -1.0 * Cb[sc];
    const int km = min(kl, n - j);
    for (int t = 1 + lane; t <= km; t += 64) Cb[sc + t] = Cb[sc + t] + Ls[sc * kl + t - 1] * temp;
    wave_lds_sync();
  }
  for (int r = lane; r < W; r += 64) {
    const int row = p0 + r;
    if (row >= c - kv && row <= min(n, c + kl)) ab[(long)(c - 1) * ld + (kv + row - c)] = Cb[r];
  }
}

// ---------------------------------------------------------------------------
// DGBTRS('N', one right-hand side), blocked.  The forward and the backward
// substitution each run as one workgroup; b's band window lives in an LDS
// ring (slot = (position-1) mod RING) refilled from global memory in chunks,
// and finished positions are written back in place.  The column steps are
// taken NBH (one wavefront) at a time:
//   head: one wave runs the block's steps on the block's NBH rows, one row
//     per lane, with the step's value broadcast by v_readlane and the
//     block's triangle of L (or U) staged in LDS: no barrier per step;
//   tail: every thread then applies the block's steps to one of the kl (or
//     kl+ku) rows below (above) the head, in step order.
// Each element sees the same operations in the same order as DGBTRS's
// column-by-column loops, so the results are identical.  Forward blocks
// end early where a pivot row would leave the head (the block's first step
// may pivot anywhere: its swap is made in the ring while every row is
// current); the pivots of this operator are rare and adjacent.
//   forward (L with the interchanges): b(j) <-> b(ipiv(j)), then
//     b(j+1 .. j+kl) += L(., j) (-b(j));
//   backward (DTBSV, upper, bandwidth kl+ku): x(j) = b(j)/U(j,j) unless b(j)
//     is zero, then b(j-1 .. j-kl-ku) -= x(j) U(., j).
// ---------------------------------------------------------------------------
constexpr int SOLVE_MAX_THREADS = 1024;
constexpr int RING = 4096;
constexpr int CHUNK = 512;
constexpr int NBH = 64;   // head rows per block: one wavefront

__device__ __forceinline__ int ring_slot(int pos) { return (pos - 1) & (RING - 1); }

// the value of v in lane `lane` (uniform), as a uniform double
__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const long long x = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(x & 0xffffffffLL), lane);
  const int hi = __builtin_amdgcn_readlane((int)(x >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}

// forward: L y = P b, in place on b
template <int E>   // tail rows per thread
__global__ void __launch_bounds__(SOLVE_MAX_THREADS) k_band_forward(const double* __restrict__ ab,
                                                                    const int32_t* __restrict__ ipiv, double* b,
                                                                    int n, int kl, int ku) {
  __shared__ double R[RING];
  __shared__ double Lh[NBH][NBH];   // Lh[s][r] = L(j0+r, j0+s)
  __shared__ double T[NBH];         // -b(j) of the block's steps
  const int tid = threadIdx.x, NT = blockDim.x, lane = tid & 63;
  const long ld = 2 * kl + ku + 1;
  const int kd = ku + kl + 1;   // L(j+t, j) = AB(kd+t, j), t = 1..kl (1-based)
  if (kl == 0 || n < 2) return;
  auto Lg = [&](int r, int j) { return ab[(long)(j - 1) * ld + (kd - 1 + (r - j))]; };
  int hi = min(n, NBH + kl + CHUNK);
  for (int pos = 1 + tid; pos <= hi; pos += NT) R[ring_slot(pos)] = b[pos - 1];
  __syncthreads();
  int j0 = 1;
  while (j0 <= n - 1) {
    if (j0 + NBH - 1 + kl > hi && hi < n) {
      // rows up to j0+NBH-1+kl; the slots reused belong to rows below j0
      const int top = min(n, j0 + NBH - 1 + kl + CHUNK);
      for (int pos = hi + 1 + tid; pos <= top; pos += NT) R[ring_slot(pos)] = b[pos - 1];
      hi = top;
      __syncthreads();
    }
    // the block: steps j0 .. j1, the pivots of steps j0+1 .. j1 inside the head
    const int jp = j0 + lane;
    const bool out = lane >= 1 && jp <= n - 1 && ipiv[jp - 1] > j0 + NBH - 1;
    const unsigned long long m = __ballot(out);
    const int f = m ? __ffsll((long long)m) - 1 : NBH;
    const int j1 = min(min(j0 + f - 1, j0 + NBH - 1), n - 1);
    const int nb = j1 - j0 + 1;
    if (tid == 0) {
      const int l0 = ipiv[j0 - 1];
      if (l0 != j0) {
        const double t0 = R[ring_slot(j0)];
        R[ring_slot(j0)] = R[ring_slot(l0)];
        R[ring_slot(l0)] = t0;
      }
    }
    {
      // wave w stages columns w, w+nw, ...: up to 16 loads in flight per lane,
      // from clamped (valid) band entries, then the masked stores
      const int nw = NT >> 6, wv = tid >> 6;
      for (int c0 = 0; c0 < NBH; c0 += 16 * nw) {
        double tmp[16];
#pragma unroll
        for (int q = 0; q < 16; q++) {
          const int j = min(j0 + min(c0 + q * nw + wv, nb - 1), n - 1);
          tmp[q] = Lg(min(max(j0 + lane, j + 1), j + kl), j);
        }
#pragma unroll
        for (int q = 0; q < 16; q++) {
          const int sc = c0 + q * nw + wv, j = j0 + sc, row = j0 + lane;
          if (sc < NBH) Lh[sc][lane] = (sc < nb && lane > sc && row <= min(n, j + kl)) ? tmp[q] : 0.0;
        }
      }
    }
    __syncthreads();
    if (tid < 64) {
      double v = (j0 + lane <= n) ? R[ring_slot(j0 + lane)] : 0.0;
      // the block's pivot rows, one per lane, and each step's multiplier kept
      // in the step's lane: no memory operation inside the step loop
      const int pl = ipiv[min(j0 + lane, n - 1) - 1];
      double tv = 0.0;
      // a full block with no interchange after its first step and a band at
      // least a head wide (the common case): each step is one lane read and
      // one masked multiply-add
      const bool swaps = __ballot(lane >= 1 && lane < nb && pl != j0 + lane) != 0;
      const bool fast = nb == NBH && !swaps && kl >= NBH && j0 + NBH - 1 <= n;
      if (fast) {
#pragma unroll
        for (int half = 0; half < 2; half++) {
          double lc[NBH / 2];
#pragma unroll
          for (int q = 0; q < NBH / 2; q++) lc[q] = Lh[half * (NBH / 2) + q][lane];
#pragma unroll
          for (int q = 0; q < NBH / 2; q++) {
            const int sc = half * (NBH / 2) + q;
            const double temp = -1.0 * readlane_f64(v, sc);
            if (lane == sc) tv = temp;
            if (lane > sc) v = v + lc[q] * temp;
          }
        }
      }
      // the steps in two halves, each with its column of the triangle in
      // registers and constant lane indices
#pragma unroll
      for (int half = 0; half < 2 && !fast; half++) {
        double lc[NBH / 2];
#pragma unroll
        for (int q = 0; q < NBH / 2; q++) lc[q] = Lh[half * (NBH / 2) + q][lane];
#pragma unroll
        for (int q = 0; q < NBH / 2; q++) {
          const int sc = half * (NBH / 2) + q;
          if (sc < nb) {
            const int j = j0 + sc;
            if (sc > 0) {
              const int l = __builtin_amdgcn_readlane(pl, sc);
              if (l != j) {
                const double a = readlane_f64(v, sc), c = readlane_f64(v, l - j0);
                if (lane == sc) v = c;
                if (lane == l - j0) v = a;
              }
            }
            const double temp = -1.0 * readlane_f64(v, sc);
            if (lane == sc) tv = temp;
            if (lane > sc && j0 + lane <= min(n, j + kl)) v = v + lc[q] * temp;
          }
        }
      }
      T[lane] = tv;
      if (lane < nb) b[j0 + lane - 1] = v;                    // y(j0 .. j1): final
      else if (j0 + lane <= n) R[ring_slot(j0 + lane)] = v;   // head rows of the next block
    }
    __syncthreads();
    // tail rows j0+NBH .. j1+kl
#pragma unroll
    for (int e = 0; e < E; e++) {
      const int t = tid + e * NT, r = j0 + NBH + t;
      const bool act = t < kl && r <= min(n, j1 + kl);
      double v = act ? R[ring_slot(r)] : 0.0;
#pragma unroll
      for (int half = 0; half < 2; half++) {
        // unconditional loads from clamped (valid) band entries, all in flight at once
        double lt[NBH / 2];
#pragma unroll
        for (int q = 0; q < NBH / 2; q++) {
          const int j = min(j0 + half * (NBH / 2) + q, j1);
          lt[q] = Lg(min(max(r, j + 1), j + kl), j);
        }
#pragma unroll
        for (int q = 0; q < NBH / 2; q++) {
          const int sc = half * (NBH / 2) + q, j = j0 + sc;
          if (sc < nb && r <= j + kl) v = v + lt[q] * T[sc];
        }
      }
      if (act) R[ring_slot(r)] = v;
    }
    __syncthreads();
    j0 = j1 + 1;
  }
  if (tid == 0) b[n - 1] = R[ring_slot(n)];
}

// backward: U x = y (DTBSV, upper, non-unit, bandwidth kl+ku), in place on b
template <int E>   // tail rows per thread
__global__ void __launch_bounds__(SOLVE_MAX_THREADS) k_band_backward(const double* __restrict__ ab, double* b,
                                                                     int n, int kl, int ku) {
  __shared__ double R[RING];
  __shared__ double Uh[NBH][NBH];   // Uh[s][r] = U(j0+r, j0+s), r < s
  __shared__ double X[NBH];         // x(j) of the block's steps
  __shared__ int F[NBH];            // step taken (b(j) != 0)
  const int tid = threadIdx.x, NT = blockDim.x, lane = tid & 63;
  const long ld = 2 * kl + ku + 1;
  const int kk = kl + ku;   // U(i, j) = AB(kk+1+i-j, j)
  auto Ug = [&](int i, int j) { return ab[(long)(j - 1) * ld + (kk + i - j)]; };
  int lo = max(1, n - NBH - kk - CHUNK + 1);
  for (int pos = lo + tid; pos <= n; pos += NT) R[ring_slot(pos)] = b[pos - 1];
  __syncthreads();
  int j1 = n;
  while (j1 >= 1) {
    const int j0 = max(1, j1 - NBH + 1), nb = j1 - j0 + 1;
    if (j0 - kk < lo && lo > 1) {
      const int bot = max(1, j0 - kk - CHUNK);
      for (int pos = bot + tid; pos < lo; pos += NT) R[ring_slot(pos)] = b[pos - 1];
      lo = bot;
      __syncthreads();
    }
    {
      // as in the forward pass; r == sc is the diagonal U(j,j)
      const int nw = NT >> 6, wv = tid >> 6;
      for (int c0 = 0; c0 < NBH; c0 += 16 * nw) {
        double tmp[16];
#pragma unroll
        for (int q = 0; q < 16; q++) {
          const int j = j0 + min(c0 + q * nw + wv, nb - 1);
          tmp[q] = Ug(max(min(j0 + lane, j), j - kk), j);
        }
#pragma unroll
        for (int q = 0; q < 16; q++) {
          const int sc = c0 + q * nw + wv, j = j0 + sc, row = j0 + lane;
          if (sc < NBH) Uh[sc][lane] = (sc < nb && lane <= sc && row >= j - kk) ? tmp[q] : 0.0;
        }
      }
    }
    __syncthreads();
    if (tid < 64) {
      double v = (lane < nb) ? R[ring_slot(j0 + lane)] : 0.0;
      double xv = 0.0;   // x(j) and whether step j was taken, in the step's lane
      int fv = 0;
      // the band reaches every head row when kl+ku >= NBH
      const bool wide = kk >= NBH;
#pragma unroll
      for (int half = 1; half >= 0; half--) {
        double uc[NBH / 2];
#pragma unroll
        for (int q = 0; q < NBH / 2; q++) uc[q] = Uh[half * (NBH / 2) + q][lane];
#pragma unroll
        for (int q = NBH / 2 - 1; q >= 0; q--) {
          const int sc = half * (NBH / 2) + q;
          if (sc < nb) {
            const int j = j0 + sc;
            const double bj = readlane_f64(v, sc);
            if (bj != 0.0) {
              const double x = bj / readlane_f64(uc[q], sc);   // U(j,j)
              if (lane == sc) v = x;
              if (lane < sc && (wide || j0 + lane >= j - kk)) v = v - x * uc[q];
              if (lane == sc) { xv = x; fv = 1; }
            }
          }
        }
      }
      X[lane] = xv;
      F[lane] = fv;
      if (lane < nb) b[j0 + lane - 1] = v;   // x(j0 .. j1): final
    }
    __syncthreads();
    // tail rows j0-1 down to j0-kk
#pragma unroll
    for (int e = 0; e < E; e++) {
      const int t = tid + e * NT, i = j0 - 1 - t;
      const bool act = t < kk && i >= 1;
      double v = act ? R[ring_slot(i)] : 0.0;
#pragma unroll
      for (int half = 1; half >= 0; half--) {
        double ut[NBH / 2];
#pragma unroll
        for (int q = 0; q < NBH / 2; q++) {
          const int j = min(j0 + half * (NBH / 2) + q, j1);
          ut[q] = Ug(min(max(i, j - kk), j - 1 < 1 ? j : j - 1), j);
        }
#pragma unroll
        for (int q = NBH / 2 - 1; q >= 0; q--) {
          const int sc = half * (NBH / 2) + q, j = j0 + sc;
          if (sc < nb && F[sc] && i >= j - kk) v = v - X[sc] * ut[q];
        }
      }
      if (act) R[ring_slot(i)] = v;
    }
    __syncthreads();
    j1 = j0 - 1;
  }
}

// ---------------------------------------------------------------------------
// buoyancy (PS/momentum.f90:255-285): Bvec(p) = (RaT Tx - RaC Cx) RowB(p) at
// interior nodes, 0 elsewhere
// ---------------------------------------------------------------------------
__global__ void k_sf_buoyancy(SfArgs a, const double* Tb, const double* Cb, const double* row, double RaT, double RaC,
                              double* B) {
  const int p = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (p > a.n) return;
  int ip, kp;
  sf_indices(a, p, ip, kp);
  double v = 0.0;
  if (ip > 0 && ip < a.nx && kp > 0 && kp < a.nz) {
    const int p3 = a.nx + 7, p5 = a.nx + 11;
    double tr = 0.0, cr = 0.0;
    for (int j = -2; j <= 2; j++) tr = tr + a.D1[j + 3] * Tb[(long)(kp + H3) * p3 + (ip + j + H3)];
    for (int j = -2; j <= 2; j++) cr = cr + a.D1[j + 3] * Cb[(long)(kp + H5) * p5 + (ip + j + H5)];
    const double Tx = (tr / a.dr) / 1.0, Cx = (cr / a.dr) / 1.0;
    const double rc = 0.0 + RaC * Cx;
    v = (RaT * Tx - rc) * row[p - 1];
  }
  B[p - 1] = v;
}

// Stream_grid (PS/momentum.f90:287-319): psi = Bvec at the nodes; ghosts are
// exact (anti)symmetric images, written here in closed form per node
__global__ void k_sf_scatter(SfArgs a, const double* B, double* SF) {
  const int i = -H3 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int k = -H3 + (int)(blockIdx.y * blockDim.y + threadIdx.y);
  if (i > a.nx + H3 || k > a.nz + H3) return;
  int iq = i, kq = k;
  double sgn = 1.0;
  if (iq < 0) { iq = -iq; sgn = -sgn; }
  else if (iq > a.nx) { iq = 2 * a.nx - iq; sgn = -sgn; }
  if (kq < 0) { kq = -kq; if (a.Vbc == 0) sgn = -sgn; }
  else if (kq > a.nz) { kq = 2 * a.nz - kq; if (a.Vbc == 0) sgn = -sgn; }
  const double v = B[sf_pf(a, iq, kq) - 1];
  SF[(long)(k + H3) * (a.nx + 7) + (i + H3)] = sgn < 0.0 ? -v : v;
}

__global__ void k_fill_f64(double* v, long n, double x) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) v[t] = x;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static SfArgs sf_args(const Context& c) {
  SfArgs a;
  const Tables& t = c.T;
  std::memcpy(a.D1, t.D1, sizeof a.D1); std::memcpy(a.D2, t.D2, sizeof a.D2);
  std::memcpy(a.D3, t.D3, sizeof a.D3); std::memcpy(a.D4, t.D4, sizeof a.D4);
  std::memcpy(a.Drs, t.Drs, sizeof a.Drs); std::memcpy(a.Drss, t.Drss, sizeof a.Drss);
  std::memcpy(a.Drrs, t.Drrs, sizeof a.Drrs); std::memcpy(a.Drrss, t.Drrss, sizeof a.Drrss);
  a.dr = t.dr; a.ds = t.ds; a.dr2 = t.dr2; a.ds2 = t.ds2; a.dr3 = t.dr3; a.ds3 = t.ds3; a.dr4 = t.dr4; a.ds4 = t.ds4;
  a.drds = t.drdsg[1]; a.drds2 = t.drds2g[1]; a.dr2ds = t.dr2dsg[1]; a.dr2ds2 = t.dr2ds2g[1];
  a.nx = t.nx; a.nz = t.nz; a.Vbc = c.P.Vbc; a.Database = c.P.Database;
  a.kl = t.dkl; a.ku = t.dku; a.n = t.dn; a.pitch = t.nx + 7;
  return a;
}

int direct_alloc(Context& c) {
  const Tables& t = c.T;
  if (t.dkl + t.dku + NBH + CHUNK >= RING) {   // the solves' LDS window
    set_error("direct solver: band width " + std::to_string(t.dkl + t.dku) + " exceeds " +
              std::to_string(RING - NBH - CHUNK - 1));
    return PT_ERR_UNSUPPORTED;
  }
  const size_t ld = (size_t)(2 * t.dkl + t.dku + 1);
  if (hipMalloc((void**)&c.dAB, sizeof(double) * ld * (size_t)t.dn) != hipSuccess ||
      hipMalloc((void**)&c.dRow, sizeof(double) * (size_t)t.dn) != hipSuccess ||
      hipMalloc((void**)&c.dB, sizeof(double) * (size_t)t.dn) != hipSuccess ||
      hipMalloc((void**)&c.dipiv, sizeof(int32_t) * (size_t)t.dn) != hipSuccess ||
      hipMalloc((void**)&c.dju, sizeof(int32_t) * (size_t)t.dn) != hipSuccess ||
      hipMalloc((void**)&c.d_lustate, sizeof(int32_t) * 2) != hipSuccess) {
    set_error("direct solver: out of device memory for the band matrix");
    return PT_ERR_HIP;
  }
  return PT_OK;
}

int direct_free(Context& c) {
  void* ps[] = {c.dAB, c.dRow, c.dB, c.dipiv, c.dju, c.d_lustate};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  c.dAB = nullptr; c.dRow = nullptr; c.dB = nullptr; c.dipiv = nullptr; c.dju = nullptr; c.d_lustate = nullptr;
  return PT_OK;
}

// build_SF_matrix with tstep >= 0 (PS/momentum.f90:321-488): vis = visf at the
// nodes (isoviscous: 1), assembly, DGBEQU row scaling, band LU
int build_sf_matrix(Context& c) {
  const Tables& t = c.T;
  if (c.P.iterate != 0) {
    set_error("build_SF_matrix needs a context created with iterate=0");
    return PT_ERR_INVALID;
  }
  const SfArgs a = sf_args(c);
  Level& L1 = c.L[1];
  hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)((L1.nelem + 255) / 256)), dim3(256), 0, c.stream, L1.v.vis,
                     (long)L1.nelem, 1.0);
  const size_t ld = (size_t)(2 * t.dkl + t.dku + 1);
  PT_HIP(hipMemsetAsync(c.dAB, 0, sizeof(double) * ld * (size_t)t.dn, c.stream));
  hipLaunchKernelGGL(k_sf_assemble, dim3((t.dn + 127) / 128), dim3(128), 0, c.stream, a, L1.v.vis, c.dAB);
  PT_TRY(launch_equilibrate(c, c.dAB, c.dRow, t.dn, t.dkl, t.dku));
  // panels of nb columns: the panel in LDS (W = kl+nb rows), the trailing
  // workgroups hold the panel's L columns and one column per wave
  int nb = 32;
  auto panel_lds = [&](int b) { return sizeof(double) * (size_t)(t.dkl + b) * b; };
  auto trail_lds = [&](int b) { return sizeof(double) * ((size_t)b * t.dkl + (size_t)LU_TRAIL_WAVES * (t.dkl + b)); };
  while (nb > 4 && (panel_lds(nb) > LU_LDS_LIMIT || trail_lds(nb) > LU_LDS_LIMIT)) nb /= 2;
  if (panel_lds(nb) > LU_LDS_LIMIT || trail_lds(nb) > LU_LDS_LIMIT) {
    set_error("direct solver: lower band width too large for the LU panels");
    return PT_ERR_UNSUPPORTED;
  }
  // set once per process (a thread-safe static initialisation)
  static const hipError_t attr_rc = [] {
    const hipError_t e = hipFuncSetAttribute((const void*)k_lu_panel, hipFuncAttributeMaxDynamicSharedMemorySize, LU_LDS_LIMIT);
    return e != hipSuccess ? e : hipFuncSetAttribute((const void*)k_lu_trailing, hipFuncAttributeMaxDynamicSharedMemorySize,
                               LU_LDS_LIMIT);
  }();
  PT_HIP(attr_rc);
  const int kv = t.dkl + t.dku;
  // one thread per subdiagonal row of the pivot column (the panel keeps each
  // row's multiplier in a register)
  const int panel_threads = ((t.dkl + 1 + 63) / 64) * 64;
  if (panel_threads > LU_PANEL_THREADS) {
    set_error("direct solver: lower band width above " + std::to_string(LU_PANEL_THREADS - 1));
    return PT_ERR_UNSUPPORTED;
  }
  for (int p0 = 1; p0 <= t.dn; p0 += nb) {
    const int pe = std::min(p0 + nb - 1, t.dn);
    hipLaunchKernelGGL(k_lu_panel, dim3(1), dim3(panel_threads), panel_lds(nb), c.stream, c.dAB, c.dipiv, c.dju,
                       c.d_lustate, t.dn, t.dkl, t.dku, p0, nb);
    const int ncols = std::min(t.dn, pe + kv) - pe;
    if (ncols > 0)
      hipLaunchKernelGGL(k_lu_trailing, dim3((ncols + LU_TRAIL_WAVES - 1) / LU_TRAIL_WAVES),
                         dim3(LU_TRAIL_WAVES * 64), trail_lds(nb), c.stream, c.dAB, c.dipiv, c.dju, t.dn, t.dkl,
                         t.dku, p0, nb);
  }
  int32_t info = 0;
  PT_HIP(hipMemcpyAsync(&info, c.d_lustate + 1, sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
  PT_HIP(hipStreamSynchronize(c.stream));
  if (info != 0) {
    set_error("direct solver band LU: zero pivot at column " + std::to_string(info));
    return PT_ERR_SINGULAR_COARSE;
  }
  c.direct_built = true;
  return PT_OK;
}

// solve_SF_equation (PS/momentum.f90:621-653)
int solve_sf_equation(Context& c) {
  const Tables& t = c.T;
  if (!c.direct_built) {
    set_error("solve_SF_equation before build_SF_matrix");
    return PT_ERR_INVALID;
  }
  const SfArgs a = sf_args(c);
  hipLaunchKernelGGL(k_sf_buoyancy, dim3((t.dn + 255) / 256), dim3(256), 0, c.stream, a, c.Tbuoy, c.Cbuoy, c.dRow,
                     c.P.RaT, c.P.RaC, c.dB);
  // one workgroup per triangular solve, sized to its band: 64-thread waves,
  // at most 1024 threads, E band rows per thread
  auto shape = [](int band, int& nt, int& e) {
    nt = std::min(SOLVE_MAX_THREADS, ((band + 63) / 64) * 64);
    e = (band + nt - 1) / nt;
  };
  int ntf, ef, ntb, eb;
  shape(t.dkl, ntf, ef);
  shape(t.dkl + t.dku, ntb, eb);
  if (ef > 4 || eb > 4) {
    set_error("direct solver: band too wide for the solve kernels");
    return PT_ERR_UNSUPPORTED;
  }
  auto fwd = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(1), dim3(ntf), 0, c.stream, c.dAB, c.dipiv, c.dB, t.dn, t.dkl, t.dku);
  };
  auto bwd = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(1), dim3(ntb), 0, c.stream, c.dAB, c.dB, t.dn, t.dkl, t.dku);
  };
  if (ef == 1) fwd(k_band_forward<1>);
  else if (ef == 2) fwd(k_band_forward<2>);
  else fwd(k_band_forward<4>);
  if (eb == 1) bwd(k_band_backward<1>);
  else if (eb == 2) bwd(k_band_backward<2>);
  else bwd(k_band_backward<4>);
  dim3 b16(64, 4), g((t.nx + 7 + 63) / 64, (t.nz + 7 + 3) / 4);
  hipLaunchKernelGGL(k_sf_scatter, g, b16, 0, c.stream, a, c.dB, c.SF);
  c.solves++;
  return PT_OK;
}

// update_velocities' iterate=0 branch (PS/momentum.f90:175-179) with the time
// loop's build at tstep=0 (PS/main.f90:247-251): the matrix is built on the
// first solve and, for ivis >= 1, on every solve
int solve_direct(Context& c) {
  if (!c.direct_built || c.P.ivis >= 1) PT_TRY(build_sf_matrix(c));
  return solve_sf_equation(c);
}

}  // namespace pt
