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
// DGBTF2 with one workgroup: pivot search as a block-wide first-maximum
// reduction, then the row swap, the column scaling and the rank-1 update of
// the (ju-j) x km window spread over all threads
// ---------------------------------------------------------------------------
constexpr int LU_THREADS = 1024;

__global__ void __launch_bounds__(LU_THREADS) k_band_lu_wg(double* ab, int32_t* ipiv, int n, int kl, int ku,
                                                           int32_t* info) {
  __shared__ double s_best[LU_THREADS / 64];
  __shared__ int s_idx[LU_THREADS / 64];
  __shared__ int s_jp;
  const int ld = 2 * kl + ku + 1, kv = ku + kl;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
#define AB(r, q) ab[(long)((q) - 1) * ld + ((r) - 1)]
  // zero the fill-in rows of the first columns (DGBTF2's set-up)
  for (int j = ku + 2 + tid; j <= min(kv, n); j += LU_THREADS)
    for (int i = kv - j + 2; i <= kl; i++) AB(i, j) = 0.0;
  __syncthreads();
  int ju = 1, inf = 0;
  for (int j = 1; j <= n; j++) {
    if (j + kv <= n)
      for (int i = 1 + tid; i <= kl; i += LU_THREADS) AB(i, j + kv) = 0.0;
    const int km = min(kl, n - j);
    // IDAMAX over AB(kv+1 .. kv+1+km, j): the first index of the largest |.|
    double best = -1.0;
    int bi = 0x7fffffff;
    for (int t = tid; t <= km; t += LU_THREADS) {
      const double x = fabs(AB(kv + 1 + t, j));
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
      for (int w = 1; w < LU_THREADS / 64; w++)
        if (s_best[w] > b || (s_best[w] == b && s_idx[w] < x)) { b = s_best[w]; x = s_idx[w]; }
      s_jp = x + 1;
      ipiv[j - 1] = x + j;
    }
    __syncthreads();
    const int jp = s_jp;
    const double piv = AB(kv + jp, j);
    if (piv != 0.0) {
      ju = max(ju, min(j + ku + jp - 1, n));
      if (jp != 1)
        for (int t = tid; t <= ju - j; t += LU_THREADS) {
          double* pa = &ab[(long)(j - 1 + t) * ld + (kv + jp - 1 - t)];
          double* pb = &ab[(long)(j - 1 + t) * ld + (kv - t)];
          const double tmp = *pa;
          *pa = *pb;
          *pb = tmp;
        }
      __syncthreads();
      if (km > 0) {
        const double rcp = 1.0 / AB(kv + 1, j);
        for (int t = 1 + tid; t <= km; t += LU_THREADS) AB(kv + 1 + t, j) = rcp * AB(kv + 1 + t, j);
        __syncthreads();
        // DGER: A(t, jj) += x(t) * (-y(jj)); the pivot row y is not updated
        const int total = (ju - j) * km;
        for (int idx = tid; idx < total; idx += LU_THREADS) {
          const int jj = 1 + idx / km, t = 1 + idx % km;
          const double temp = -1.0 * ab[(long)(j - 1 + jj) * ld + (kv - jj)];
          double& x = ab[(long)(j - 1 + jj) * ld + (kv - jj + t)];
          x = x + AB(kv + 1 + t, j) * temp;
        }
      }
    } else if (inf == 0) {
      inf = j;
    }
    __syncthreads();
  }
  if (tid == 0) *info = inf;
#undef AB
}

// ---------------------------------------------------------------------------
// DGBTRS('N', one right-hand side): the forward and the backward substitution
// each run as one workgroup sized to its band (one band row per thread up to
// 1024 threads), so no wave of the block idles through the steps.  The band
// window of b (and, forward, the pivot indices) lives in LDS rings (slot =
// (position-1) mod RING), refilled from global memory in chunks; finished
// positions are written back in place.  The band columns stream from HBM
// through a register queue DEPTH steps ahead of the step that uses them, with
// unconditional loads from clamped addresses so that no branch separates the
// queue slots, and every LDS read of a step is issued before any is waited
// on.  One barrier per column step:
//   forward (L with the interchanges): the pivot swap of step j is read from
//     two broadcast slots filled during step j-1 (the values at positions j
//     and ipiv(j) after that step), so no thread reads a ring entry that
//     another thread writes in the same phase;
//   backward (DTBSV, upper, bandwidth kl+ku): x(j) = b(j)/U(j,j) unless b(j)
//     is zero, then b(j-1 .. j-kl-ku) -= x(j) U(., j).
// Thread t owns the band rows t, t+blockDim.x, ... (E of them).
// ---------------------------------------------------------------------------
constexpr int SOLVE_MAX_THREADS = 1024;
constexpr int RING = 4096;
constexpr int CHUNK = 512;
constexpr int DEPTH = 8;

// Workgroup barrier that orders LDS only: the per-step handoffs of the solve
// go through LDS, and a full __syncthreads() would also wait for the band
// columns still in flight from HBM.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ int ring_slot(int pos) { return (pos - 1) & (RING - 1); }

// forward: L y = P b, in place on b
template <int E>
__global__ void __launch_bounds__(SOLVE_MAX_THREADS) k_band_forward(const double* __restrict__ ab,
                                                                    const int32_t* __restrict__ ipiv, double* b,
                                                                    int n, int kl, int ku) {
  __shared__ double R[RING];
  __shared__ int32_t IP[RING];
  __shared__ double bx[2], by[2];
  const int tid = threadIdx.x, NT = blockDim.x;
  const long ld = 2 * kl + ku + 1;
  const int kd = ku + kl + 1;
  if (kl == 0 || n < 2) return;
  int hi = min(n, kl + CHUNK);
  for (int pos = 1 + tid; pos <= hi; pos += NT) {
    R[ring_slot(pos)] = b[pos - 1];
    IP[ring_slot(pos)] = ipiv[pos - 1];
  }
  __syncthreads();
  // L(t, j) = AB(kd+1+t, j), t = 0..lm-1 (0-based rows kd+t of column j);
  // entries past the band or the last column are loaded but never used
  double q[DEPTH][E];
  auto fetch_l = [&](double* dst, int j) {
    const long col = (long)(min(j, n - 1) - 1) * ld + kd;
#pragma unroll
    for (int e = 0; e < E; e++) dst[e] = ab[col + min(tid + e * NT, kl - 1)];
  };
#pragma unroll
  for (int u = 0; u < DEPTH; u++) fetch_l(q[u], 1 + u);
  if (tid == 0) {
    bx[1] = R[ring_slot(IP[ring_slot(1)])];
    by[1] = R[ring_slot(1)];
  }
  __syncthreads();
  for (int j0 = 1; j0 <= n - 1; j0 += DEPTH) {
    int pv[DEPTH + 1];   // pivot rows of this block's steps (and the next one's first)
#pragma unroll
    for (int u = 0; u < DEPTH; u++) {
      const int j = j0 + u;
      if (j > n - 1) break;
      if (j + kl + 1 > hi && hi < n) {
        // refill (up to j+kl+1, which the next pivot may name): positions
        // hi+1 .. hi+CHUNK reuse the slots of finished positions
        const int top = min(n, hi + CHUNK);
        for (int pos = hi + 1 + tid; pos <= top; pos += NT) {
          R[ring_slot(pos)] = b[pos - 1];
          IP[ring_slot(pos)] = ipiv[pos - 1];
        }
        hi = top;
        lds_barrier();
      }
      if (u == 0) {
        // positions up to j0+kl+1 (kl >= DEPTH) are in the ring
#pragma unroll
        for (int v = 0; v <= DEPTH; v++) pv[v] = IP[ring_slot(min(j0 + v, n - 1))];
      }
      double Lv[E];
#pragma unroll
      for (int e = 0; e < E; e++) Lv[e] = q[u][e];
      fetch_l(q[u], j + DEPTH);
      const int lm = min(kl, n - j);
      const int l = pv[u];
      const int ln = (j + 1 <= n - 1) ? pv[u + 1] : 0;   // next step's pivot row
      // the ring slots of positions past the update range are valid and never
      // written in this phase, so all reads are issued at once
      double rv[E];
#pragma unroll
      for (int e = 0; e < E; e++) rv[e] = R[ring_slot(j + 1 + tid + e * NT)];
      const double old_j = by[j & 1], old_l = bx[j & 1];
      const double bj = (l != j) ? old_l : old_j;   // b(j) after the interchange
      const double temp = -1.0 * bj;
      if (tid == 0) b[j - 1] = bj;
#pragma unroll
      for (int e = 0; e < E; e++) {
        const int t = tid + e * NT;
        if (t < lm) {
          const int pos = j + 1 + t;
          double v = (pos == l && l != j) ? old_j : rv[e];
          v = v + Lv[e] * temp;
          R[ring_slot(pos)] = v;
          if (pos == j + 1) by[(j + 1) & 1] = v;
          if (pos == ln) bx[(j + 1) & 1] = v;
        }
      }
      // the next pivot row lies outside this step's update range: its value is unchanged
      if (tid == 0 && ln > j + lm) bx[(j + 1) & 1] = R[ring_slot(ln)];
      lds_barrier();
    }
  }
  if (tid == 0) b[n - 1] = by[n & 1];
}

// backward: U x = y (DTBSV, upper, non-unit, bandwidth kl+ku), in place on b
template <int E>
__global__ void __launch_bounds__(SOLVE_MAX_THREADS) k_band_backward(const double* __restrict__ ab, double* b,
                                                                     int n, int kl, int ku) {
  __shared__ double R[RING];
  const int tid = threadIdx.x, NT = blockDim.x;
  const long ld = 2 * kl + ku + 1;
  const int kk = kl + ku;
  int lo = max(1, n - kk - CHUNK + 1);
  for (int pos = lo + tid; pos <= n; pos += NT) R[ring_slot(pos)] = b[pos - 1];
  // U(j-1-t, j) = AB(kk-t, j) (0-based row kk-1-t), diagonal AB(kk+1, j)
  double q[DEPTH][E], qd[DEPTH];
  auto fetch_u = [&](double* dst, double& dg, int j) {
    const long col = (long)(max(j, 1) - 1) * ld;
#pragma unroll
    for (int e = 0; e < E; e++) dst[e] = ab[col + (kk - 1 - min(tid + e * NT, kk - 1))];
    dg = ab[col + kk];
  };
#pragma unroll
  for (int u = 0; u < DEPTH; u++) fetch_u(q[u], qd[u], n - u);
  __syncthreads();
  for (int j0 = n; j0 >= 1; j0 -= DEPTH) {
#pragma unroll
    for (int u = 0; u < DEPTH; u++) {
      const int j = j0 - u;
      if (j < 1) break;
      if (j - kk < lo && lo > 1) {
        const int bot = max(1, lo - CHUNK);
        for (int pos = bot + tid; pos < lo; pos += NT) R[ring_slot(pos)] = b[pos - 1];
        lo = bot;
        lds_barrier();
      }
      double Uv[E];
#pragma unroll
      for (int e = 0; e < E; e++) Uv[e] = q[u][e];
      const double dg = qd[u];
      fetch_u(q[u], qd[u], j - DEPTH);
      double rv[E];   // issued together with the read of b(j)
#pragma unroll
      for (int e = 0; e < E; e++) rv[e] = R[ring_slot(j - 1 - tid - e * NT)];
      const double bj = R[ring_slot(j)];
      if (bj != 0.0) {
        const double x = bj / dg;
        if (tid == 0) b[j - 1] = x;
        const int m = min(kk, j - 1);
#pragma unroll
        for (int e = 0; e < E; e++) {
          const int t = tid + e * NT;
          if (t < m) R[ring_slot(j - 1 - t)] = rv[e] - x * Uv[e];
        }
      } else if (tid == 0) {
        b[j - 1] = bj;
      }
      lds_barrier();
    }
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
  if (t.dkl < DEPTH) {   // the solve reads a block's pivots from positions up to j0+kl+1
    set_error("direct solver: lower band width below " + std::to_string(DEPTH));
    return PT_ERR_UNSUPPORTED;
  }
  if (t.dkl + t.dku + 1 > RING - CHUNK) {
    set_error("direct solver: band width " + std::to_string(t.dkl + t.dku) + " exceeds " +
              std::to_string(RING - CHUNK - 1));
    return PT_ERR_UNSUPPORTED;
  }
  const size_t ld = (size_t)(2 * t.dkl + t.dku + 1);
  if (hipMalloc((void**)&c.dAB, sizeof(double) * ld * (size_t)t.dn) != hipSuccess ||
      hipMalloc((void**)&c.dRow, sizeof(double) * (size_t)t.dn) != hipSuccess ||
      hipMalloc((void**)&c.dB, sizeof(double) * (size_t)t.dn) != hipSuccess ||
      hipMalloc((void**)&c.dipiv, sizeof(int32_t) * (size_t)t.dn) != hipSuccess) {
    set_error("direct solver: out of device memory for the band matrix");
    return PT_ERR_HIP;
  }
  return PT_OK;
}

int direct_free(Context& c) {
  void* ps[] = {c.dAB, c.dRow, c.dB, c.dipiv};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  c.dAB = nullptr; c.dRow = nullptr; c.dB = nullptr; c.dipiv = nullptr;
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
  hipLaunchKernelGGL(k_band_lu_wg, dim3(1), dim3(LU_THREADS), 0, c.stream, c.dAB, c.dipiv, t.dn, t.dkl, t.dku,
                     c.d_info);
  int32_t info = 0;
  PT_HIP(hipMemcpyAsync(&info, c.d_info, sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
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
