// pt_equalize.hip -- tracer redistribution on the device:
// equalize_tracers_original (PS/composition.f90:84-263) for discrete tracer
// types (comp=1, ntype=1, shape_function=1, iC_continuous=0), called at the
// start of energy_time (PS/energy.f90:198-204).
//
// What the reference does in one call:
//  1. nexcess(i,k) = nempty - tpc at nodes with tracers whose neighbourhood
//     (i-buffer:i+buffer, k-buffer:k+buffer) holds one tracer type only
//     (count_type(tt,..) == nempty(..) there, halos mirrored, :107-137),
//     else 0; nfree = the sum of the positive excesses.  With tpc_max =
//     tpc_min = tpc the early return (:150-155) is taken only when nfree = 0.
//  2. donors (:163-190): scanning the tracers in ID order, a node with excess
//     gives up its first tracers near the walls (s < dist_s, s > 1-dist_s,
//     r < dist_r, r > aspect-dist_r), then its first other tracers, until its
//     excess is used up; IDsave lists the wall donors in ID order, then the
//     others in ID order.
//  3. lottery (:193-238): while some node lacks tracers (nexcess < 0), the
//     nodes at the lowest nexcess, listed i outer / k inner, receive one
//     round of ncells draws from mzran stream 0: a node q = floor(ncells*u)+1
//     and a position uniform in its dual cell, clamped off the walls; the
//     next donor moves there and takes the node's majority type.  Stops when
//     every nexcess >= 0 (then nfree = the moves made) or the donors run out.
//  4. moved tracers take T interpolated from Tratio (:242-250); all tracers
//     are projected again without the filters (:257-259).
//
// On the device: (1) one thread per node over the per-node counts; (2) one
// wavefront per node with excess ranks its tracers by ID within their class
// (the storage order is the cell order, so a node's tracers are one slot
// range), then an ID-order compaction writes IDsave; (3) one workgroup runs
// the lottery: per round the minimum and the node list by a workgroup
// reduction and ordered compactions over the nodes that lack tracers, the
// draws and moves by all lanes, each lane jumping the mzran stream to its
// own run of draws (the stream's two parts are a linear recurrence mod M
// and an affine map mod 2^32, so a jump is a product of precomputed
// powers); (4) the tracer pipeline's interpolation table and projection.
#include <array>
#include <climits>
#include <string>
#include "pt_internal.hpp"

namespace pt {

namespace {

constexpr int EQ_BLOCK = 1024;    // IDs per block of the IDsave compaction
constexpr int EQ_LOT = 1024;      // lanes of the lottery workgroup

__global__ void k_eq_fill(int32_t* a, int64_t n, int32_t v) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) a[j] = v;
}

// cell ranges [cstart, cend) of the cell-ordered storage (empty cells 0, 0)
__global__ void k_eq_bounds(const int32_t* okey, int64_t n, int ncell, int32_t* cstart, int32_t* cend) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int c = okey[p];
  if (c < 0 || c >= ncell) return;
  if (p == 0 || okey[p - 1] != c) cstart[c] = (int32_t)p;
  if (p == n - 1 || okey[p + 1] != c) cend[c] = (int32_t)(p + 1);
}

// count_type at the nodes: tracers and type-1 tracers of each dual cell
__global__ void k_eq_cells(const int32_t* cstart, const int32_t* cend, const double* Ct, int ncell, int32_t* cnt,
                           int32_t* cnt1) {
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= ncell) return;
  const int a = cstart[node], b = cend[node];
  int n1 = 0;
  for (int j = a; j < b; j++) n1 += Ct[j] != 0.0;
  cnt[node] = b - a;
  cnt1[node] = n1;
}

// an index mirrored into 0..n as the halo copies of count_type do it
// (count(-i) = count(i), count(n+i) = count(n-i), for i = 1..buffer in turn).
// An index that mirrors past the other edge needs buffer > n, and then the
// window already holds every index 0..n of that direction itself, so the
// test's outcome does not depend on where such an index lands.
__device__ __forceinline__ int reflect(int x, int n) {
  while (x < 0 || x > n) x = x < 0 ? -x : 2 * n - x;
  return x;
}

__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
  for (int o = 32; o; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(v, o);
    if (lane >= o) v += t;
  }
  return v;
}

// nexcess (PS/composition.f90:124-137, shape_function=1): a node with
// tracers whose (2 buffer + 1)^2 neighbourhood holds one type only.  Halo
// nodes outside both index ranges hold no counts (count_type there stays 0,
// equal to nempty for every type).  Written node-major (nex) and i-major
// (nexT, the lottery's order); nfree accumulates the positive parts.
__global__ void __launch_bounds__(256) k_eq_excess(const int32_t* __restrict__ cnt, const int32_t* __restrict__ cnt1,
                                                   int nx, int nz, int b, int tpc, int32_t* nex, int32_t* nexT,
                                                   int32_t* nfree) {
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  const int ncell = (nx + 1) * (nz + 1);
  int e = 0;
  if (node < ncell) {
    const int k = node / (nx + 1), i = node - k * (nx + 1);
    const int n = cnt[node];
    if (n > 0) {
      bool mixed0 = false, mixed1 = false;   // some node holds type 1 / type 0
      for (int kk = k - b; kk <= k + b && !(mixed0 && mixed1); kk++) {
        const bool kin = kk >= 0 && kk <= nz;
        const int kr = kin ? kk : reflect(kk, nz);
        for (int ii = i - b; ii <= i + b; ii++) {
          const bool iin = ii >= 0 && ii <= nx;
          if (!iin && !kin) continue;
          const int q = kr * (nx + 1) + (iin ? ii : reflect(ii, nx));
          const int c1 = cnt1[q];
          mixed0 |= c1 != 0;
          mixed1 |= c1 != cnt[q];
          if (mixed0 && mixed1) break;
        }
      }
      if (!mixed0 || !mixed1) e = n - tpc;
    }
    nex[node] = e;
    nexT[i * (nz + 1) + k] = e;
  }
  const int w = wave_sum(e > 0 ? e : 0);
  if ((threadIdx.x & 63) == 0 && w) atomicAdd(nfree, w);
}

// Donors (PS/composition.f90:163-190).  One wavefront per node with excess
// e: its m tracers are the slot range [cstart, cstart+m); each takes its
// rank by ID among the node's tracers of its class (0 near the walls, 1
// else); the first min(e, n_wall) wall tracers and the first e - that many
// others are donors.  cls[id] = 1 (wall donor) or 2 (other donor), slot[id]
// its storage slot.
__global__ void __launch_bounds__(256) k_eq_donors(const int32_t* __restrict__ cstart, const int32_t* __restrict__ cnt,
                                                   const int32_t* __restrict__ nex, int ncell,
                                                   const double* __restrict__ r, const double* __restrict__ s,
                                                   const int32_t* __restrict__ tid, double lo_r, double hi_r,
                                                   double lo_s, double hi_s, int32_t* cls, int32_t* slot) {
  const int lane = threadIdx.x & 63;
  const int node = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (node >= ncell) return;   // the whole wavefront
  const int e = nex[node];
  if (e <= 0) return;
  const int start = cstart[node], m = cnt[node];
  auto wall = [&](int j) {
    const double rv = r[j], sv = s[j];
    return sv < lo_s || sv > hi_s || rv < lo_r || rv > hi_r;
  };
  int n_wall = 0;
  for (int u0 = 0; u0 < m; u0 += 64) {
    const int u = u0 + lane;
    n_wall += __popcll(__ballot(u < m && wall(start + u)));
  }
  const int take_wall = e < n_wall ? e : n_wall, take_other = e - take_wall;
  for (int t0 = 0; t0 < m; t0 += 64) {
    const int t = t0 + lane;
    const bool live = t < m;
    const int my_id = live ? tid[start + t] : INT_MAX;
    const int my_cl = live ? (wall(start + t) ? 0 : 1) : 2;
    int rank = 0;
    for (int u0 = 0; u0 < m; u0 += 64) {
      const int u = u0 + lane;
      const int uid = u < m ? tid[start + u] : INT_MAX;
      const int ucl = u < m ? (wall(start + u) ? 0 : 1) : 3;
      for (int q = 0; q < 64; q++) {
        const int oid = __shfl(uid, q), ocl = __shfl(ucl, q);
        rank += (ocl == my_cl) & (oid < my_id);
      }
    }
    if (live) {
      const int d = my_cl == 0 ? (rank < take_wall ? 1 : 0) : (rank < take_other ? 2 : 0);
      if (d) {
        cls[my_id] = d;
        slot[my_id] = start + t;
      }
    }
  }
}

// IDsave compaction in ID order: per block of EQ_BLOCK IDs the numbers of
// wall donors and other donors
__global__ void __launch_bounds__(EQ_BLOCK) k_eq_bcount(const int32_t* cls, int64_t n, int32_t* bsum) {
  __shared__ int w1[EQ_BLOCK / 64], w2[EQ_BLOCK / 64];
  const int64_t id = (int64_t)blockIdx.x * EQ_BLOCK + threadIdx.x;
  const int d = id < n ? cls[id] : 0;
  const int c1 = __popcll(__ballot(d == 1)), c2 = __popcll(__ballot(d == 2));
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { w1[wv] = c1; w2[wv] = c2; }
  __syncthreads();
  if (threadIdx.x == 0) {
    int a = 0, b = 0;
    for (int q = 0; q < EQ_BLOCK / 64; q++) { a += w1[q]; b += w2[q]; }
    bsum[2 * blockIdx.x] = a;
    bsum[2 * blockIdx.x + 1] = b;
  }
}
// exclusive prefix sums of the block counts (in place), totals into tot[0..1]
__global__ void __launch_bounds__(1024) k_eq_bscan(int32_t* bsum, int nb, int32_t* tot) {
  __shared__ int wa[16], wb[16];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int per = (nb + 1023) / 1024, b0 = t * per, b1 = b0 + per < nb ? b0 + per : nb;
  int sa = 0, sb = 0;
  for (int b = b0; b < b1; b++) { sa += bsum[2 * b]; sb += bsum[2 * b + 1]; }
  const int ia = wave_incl_scan(sa), ib = wave_incl_scan(sb);
  if (lane == 63) { wa[wv] = ia; wb[wv] = ib; }
  __syncthreads();
  int oa = 0, ob = 0, ta = 0, tb = 0;
  for (int q = 0; q < 16; q++) {
    if (q < wv) { oa += wa[q]; ob += wb[q]; }
    ta += wa[q]; tb += wb[q];
  }
  int ra = oa + ia - sa, rb = ob + ib - sb;
  for (int b = b0; b < b1; b++) {
    const int ca = bsum[2 * b], cb = bsum[2 * b + 1];
    bsum[2 * b] = ra; bsum[2 * b + 1] = rb;
    ra += ca; rb += cb;
  }
  if (t == 0) { tot[0] = ta; tot[1] = tb; }
}
// IDsave: wall donors in ID order, then the other donors in ID order
__global__ void __launch_bounds__(EQ_BLOCK) k_eq_place(const int32_t* cls, const int32_t* slot, int64_t n,
                                                       const int32_t* bsum, const int32_t* tot, int32_t* save) {
  __shared__ int w1[EQ_BLOCK / 64], w2[EQ_BLOCK / 64];
  const int64_t id = (int64_t)blockIdx.x * EQ_BLOCK + threadIdx.x;
  const int d = id < n ? cls[id] : 0;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const unsigned long long m1 = __ballot(d == 1), m2 = __ballot(d == 2);
  const unsigned long long lt = (1ull << lane) - 1ull;
  if (lane == 0) { w1[wv] = __popcll(m1); w2[wv] = __popcll(m2); }
  __syncthreads();
  if (!d) return;
  int o1 = 0, o2 = 0;
  for (int q = 0; q < wv; q++) { o1 += w1[q]; o2 += w2[q]; }
  const int pos = d == 1 ? bsum[2 * blockIdx.x] + o1 + __popcll(m1 & lt)
                         : tot[0] + bsum[2 * blockIdx.x + 1] + o2 + __popcll(m2 & lt);
  save[pos] = slot[id];
}

// mzran (PS/composition.f90:7-20) on one stream's state
constexpr uint32_t MZ_M = 2147483579u;   // 2^31 - 69
// x mod M for x < 2^64: 2^31 = 69 (mod M), so two folds hi*69 + lo bring x
// below 2M
__device__ __forceinline__ uint32_t mz_reduce(uint64_t x) {
  x = (x >> 31) * 69u + (x & 0x7FFFFFFFu);
  x = (x >> 31) * 69u + (x & 0x7FFFFFFFu);
  return (uint32_t)(x >= MZ_M ? x - MZ_M : x);
}
struct Mz {
  int32_t i, j, k, n;
  __device__ double next() {
    int32_t ran = (int32_t)((uint32_t)i - (uint32_t)k);
    if (ran < 0) ran = ran + (int32_t)MZ_M;
    i = j; j = k; k = ran;
    n = (int32_t)(69069u * (uint32_t)n + 1013904243u);
    ran = (int32_t)((uint32_t)ran + (uint32_t)n);
    return 0.5 + 0.2328306e-9 * (double)ran;
  }
  // Advance by s draws at once.  With i, j, k in [0, M) the first part of
  // the stream is the linear recurrence x(d) = x(d-3) - x(d-1) mod M on the
  // state (i, j, k) = (x(d-3), x(d-2), x(d-1)), matrix [[0,1,0],[0,0,1],
  // [1,0,M-1]]; the second part is the affine map n -> 69069 n + 1013904243
  // mod 2^32.  jump holds both maps raised to 2^b, b = 0..31: 9 entries per
  // matrix (row-major), then (multiplier, increment) pairs.
  __device__ void advance(uint64_t s, const uint32_t* __restrict__ jump) {
    uint32_t v0 = (uint32_t)i, v1 = (uint32_t)j, v2 = (uint32_t)k, w = (uint32_t)n;
    for (int b = 0; s; b++, s >>= 1) {
      if (!(s & 1)) continue;
      const uint32_t* P = jump + 9 * b;
      const uint32_t a0 = mz_reduce((uint64_t)P[0] * v0 + (uint64_t)P[1] * v1 + (uint64_t)P[2] * v2);
      const uint32_t a1 = mz_reduce((uint64_t)P[3] * v0 + (uint64_t)P[4] * v1 + (uint64_t)P[5] * v2);
      const uint32_t a2 = mz_reduce((uint64_t)P[6] * v0 + (uint64_t)P[7] * v1 + (uint64_t)P[8] * v2);
      v0 = a0; v1 = a1; v2 = a2;
      const uint32_t* L = jump + 9 * 32 + 2 * b;
      w = L[0] * w + L[1];
    }
    i = (int32_t)v0; j = (int32_t)v1; k = (int32_t)v2; n = (int32_t)w;
  }
};

struct Lottery {
  int32_t *nexT, *actA, *actB, *listA, *listB;   // nexcess in i-major order (o = i (nz+1) + k), o lists
  const int32_t *save, *cnt, *cnt1;
  double *r, *s, *Ct;
  int32_t* dev;   // [0] donors, [1] moves made (out), [4..7] mzran state (in/out)
  const uint32_t* jump;
  int nx, nz;
  double dr, ds, aspect, dist_r, dist_s;
};

__device__ __forceinline__ int load_nex(int32_t* p) {
  // the moves update nexcess with device-scope atomics; read it at the same scope
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

constexpr int EQ_PT = 8;   // consecutive list entries per lane in one compaction pass

// exclusive prefix sums over the workgroup's lanes of two counts, and their totals
__device__ __forceinline__ void block_scan2(int a, int b, int* s_a, int* s_b, int* ea, int* eb, int* ta, int* tb) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ia = wave_incl_scan(a), ib = wave_incl_scan(b);
  if (lane == 63) { s_a[wv] = ia; s_b[wv] = ib; }
  __syncthreads();
  int oa = 0, ob = 0, sa = 0, sb = 0;
  for (int q = 0; q < EQ_LOT / 64; q++) {
    if (q < wv) { oa += s_a[q]; ob += s_b[q]; }
    sa += s_a[q]; sb += s_b[q];
  }
  __syncthreads();
  *ea = oa + ia - a; *eb = ob + ib - b; *ta = sa; *tb = sb;
}

// The lottery (PS/composition.f90:193-238) in one workgroup.  Only nodes
// that lack tracers at the start can be at the minimum, and a node leaves
// for good once it reaches 0, so a new minimum is found on a shrinking list
// of those nodes (kept in the reference's i-outer order), which also gives
// istore, the nodes at it; while the minimum stays, the next round's istore
// is the last one's nodes still at it.  Then the round's moves, lane t
// taking a contiguous run of them: it advances a copy of the round's mzran
// state to its first draw (three per move) and draws its moves in the
// reference's order.  Every loop bound is the same in all lanes; each round
// moves at least one tracer, so the rounds end after at most `donors`.
__global__ void __launch_bounds__(EQ_LOT) k_eq_lottery(Lottery a) {
  __shared__ int s_w[EQ_LOT / 64], s_w2[EQ_LOT / 64];
  __shared__ int s_min;
  __shared__ int32_t s_state[4], s_next[4];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int nnode = (a.nx + 1) * (a.nz + 1);
  const int64_t donors = a.dev[0];
  if (t < 4) s_state[t] = a.dev[4 + t];
  // the nodes that lack tracers, i outer and k inner; each pass, lane t
  // takes EQ_PT consecutive entries
  int32_t *act = a.actA, *act2 = a.actB;
  int nact = 0;
  for (int base = 0; base < nnode; base += EQ_LOT * EQ_PT) {
    const int o0 = base + t * EQ_PT;
    int c = 0;
    for (int u = 0; u < EQ_PT; u++) c += o0 + u < nnode && a.nexT[o0 + u] < 0;
    int e, e2, tot, tot2;
    block_scan2(c, 0, s_w, s_w2, &e, &e2, &tot, &tot2);
    for (int u = 0; u < EQ_PT; u++)
      if (o0 + u < nnode && a.nexT[o0 + u] < 0) act[nact + e++] = o0 + u;
    nact += tot;
  }
  __syncthreads();
  int64_t p = 0;   // moves made
  int32_t *list = a.listA, *list2 = a.listB;
  int level = 0, ncells = 0;   // the minimum and its nodes (istore), once found
  while (p < donors) {
    // While the minimum stays at `level` no node can join it (nexcess only
    // grows): the nodes still at it are the last list's, in the same order.
    if (ncells > 0) {
      int nk = 0;
      for (int base = 0; base < ncells; base += EQ_LOT * EQ_PT) {
        const int j0 = base + t * EQ_PT;
        int o[EQ_PT], c1 = 0;
#pragma unroll
        for (int u = 0; u < EQ_PT; u++) {
          o[u] = j0 + u < ncells ? list[j0 + u] : -1;
          if (o[u] >= 0 && load_nex(a.nexT + o[u]) != level) o[u] = -1;
          c1 += o[u] >= 0;
        }
        int e1, e2, tot1, tot2;
        block_scan2(c1, 0, s_w, s_w2, &e1, &e2, &tot1, &tot2);
#pragma unroll
        for (int u = 0; u < EQ_PT; u++)
          if (o[u] >= 0) list2[nk + e1++] = o[u];
        nk += tot1;
      }
      int32_t* sw = list; list = list2; list2 = sw;
      ncells = nk;
    }
    if (ncells == 0) {
      // a new minimum: scan the nodes still lacking
      int mn = INT_MAX;
      for (int j = t; j < nact; j += EQ_LOT) mn = min(mn, load_nex(a.nexT + act[j]));
#pragma unroll
      for (int o = 32; o; o >>= 1) mn = min(mn, __shfl_xor(mn, o));
      if (lane == 0) s_w[wv] = mn;
      __syncthreads();
      if (t == 0) {
        int m = s_w[0];
        for (int q = 1; q < EQ_LOT / 64; q++) m = min(m, s_w[q]);
        s_min = m;
      }
      __syncthreads();
      mn = s_min;
      if (mn >= 0) break;
      level = mn;
      // istore (the nodes at the minimum) and the nodes still lacking
      int nkeep = 0;
      for (int base = 0; base < nact; base += EQ_LOT * EQ_PT) {
        const int j0 = base + t * EQ_PT;
        int o[EQ_PT], v[EQ_PT], c1 = 0, c2 = 0;
#pragma unroll
        for (int u = 0; u < EQ_PT; u++) {
          const bool live = j0 + u < nact;
          o[u] = live ? act[j0 + u] : 0;
          v[u] = live ? load_nex(a.nexT + o[u]) : 0;
          c1 += live && v[u] == mn;
          c2 += live && v[u] < 0;
        }
        int e1, e2, tot1, tot2;
        block_scan2(c1, c2, s_w, s_w2, &e1, &e2, &tot1, &tot2);
#pragma unroll
        for (int u = 0; u < EQ_PT; u++) {
          const bool live = j0 + u < nact;
          if (live && v[u] == mn) list[ncells + e1++] = o[u];
          if (live && v[u] < 0) act2[nkeep + e2++] = o[u];
        }
        ncells += tot1;
        nkeep += tot2;
      }
      int32_t* sw = act; act = act2; act2 = sw;
      nact = nkeep;
    }
    if (t < 4) s_next[t] = s_state[t];
    __syncthreads();
    // this round's moves: ncells draws, or fewer when the donors run out
    const int m = (int)(donors - p < (int64_t)ncells ? donors - p : (int64_t)ncells);
    const int per = (m + EQ_LOT - 1) / EQ_LOT;
    const int j0 = t * per < m ? t * per : m, j1 = j0 + per < m ? j0 + per : m;
    if (j0 < j1) {
      Mz mz{s_state[0], s_state[1], s_state[2], s_state[3]};
      mz.advance(3ull * (uint64_t)j0, a.jump);
      for (int j = j0; j < j1; j++) {
        const double u = mz.next();
        const int o = list[(int)floor((double)ncells * u)];
        const double ur = mz.next(), us = mz.next();
        const int i = o / (a.nz + 1), k = o - i * (a.nz + 1);
        const double rc = a.dr * (double)i, sc = a.ds * (double)k;
        const double pr = (ur - 0.5) * a.dr, ps = (us - 0.5) * a.ds;
        double rn = rc + pr, sn = sc + ps;
        if (rn <= 0.0) rn = a.dist_r;
        if (rn >= a.aspect) rn = a.aspect - a.dist_r;
        if (sn <= 0.0) sn = a.dist_s;
        if (sn >= 1.0) sn = 1.0 - a.dist_s;
        const int q = a.save[p + j];
        a.r[q] = rn;
        a.s[q] = sn;
        // maxloc(count_type(0:1,i,k)) - 1: type 0 on a tie
        const int node = k * (a.nx + 1) + i;
        const int n1 = a.cnt1[node];
        a.Ct[q] = n1 > a.cnt[node] - n1 ? 1.0 : 0.0;
        atomicAdd(a.nexT + o, 1);
      }
      if (j1 == m) {
        s_next[0] = mz.i; s_next[1] = mz.j; s_next[2] = mz.k; s_next[3] = mz.n;
      }
    }
    __syncthreads();
    if (t < 4) s_state[t] = s_next[t];
    __syncthreads();
    p += m;
  }
  if (t == 0) {
    a.dev[1] = (int32_t)p;
    a.dev[4] = s_state[0]; a.dev[5] = s_state[1]; a.dev[6] = s_state[2]; a.dev[7] = s_state[3];
  }
}

unsigned blocks(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// mzran's two maps raised to 2^b, b = 0..31 (the layout Mz::advance reads)
constexpr int EQ_JUMP = 32 * 9 + 32 * 2;
void mz_jump_table(uint32_t* tab) {
  const uint64_t M = MZ_M;
  uint64_t P[9] = {0, 1, 0, 0, 0, 1, 1, 0, M - 1};
  uint32_t A = 69069u, C = 1013904243u;
  for (int b = 0; b < 32; b++) {
    for (int q = 0; q < 9; q++) tab[9 * b + q] = (uint32_t)P[q];
    tab[9 * 32 + 2 * b] = A;
    tab[9 * 32 + 2 * b + 1] = C;
    uint64_t Q[9];
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) {
        uint64_t acc = 0;
        for (int k = 0; k < 3; k++) acc = (acc + P[3 * r + k] * P[3 * k + c] % M) % M;
        Q[3 * r + c] = acc;
      }
    for (int q = 0; q < 9; q++) P[q] = Q[q];
    C = A * C + C;
    A = A * A;
  }
}

}  // namespace

void equalize_free(Context& c) {
  for (int32_t* p : {c.eq_node, c.eq_id, c.eq_save, c.eq_dev, c.eq_bsum})
    if (p) (void)hipFree(p);
  c.eq_node = c.eq_id = c.eq_save = c.eq_dev = c.eq_bsum = nullptr;
}

static int equalize_body(Context& c, int32_t* cnt, int32_t* cnt1, int32_t* nex, int32_t* list, int32_t* cls,
                         int32_t* slot, int nb);

int equalize_tracers(Context& c) {
  const int nx = c.T.nx, nz = c.T.nz, ncell = (nx + 1) * (nz + 1);
  const int64_t n = c.T.ntr;
  const int nb = (int)blocks(n, EQ_BLOCK);
  if (!c.eq_node) {
    PT_HIP(hipMalloc((void**)&c.eq_node, sizeof(int32_t) * 8 * (size_t)ncell));
    PT_HIP(hipMalloc((void**)&c.eq_id, sizeof(int32_t) * 2 * (size_t)n));
    PT_HIP(hipMalloc((void**)&c.eq_save, sizeof(int32_t) * (size_t)n));
    PT_HIP(hipMalloc((void**)&c.eq_dev, sizeof(int32_t) * (8 + EQ_JUMP)));
    PT_HIP(hipMalloc((void**)&c.eq_bsum, sizeof(int32_t) * 2 * (size_t)nb));
    // one table for the process, built once (thread-safe static
    // initialisation: ensemble cases call this from several host threads)
    static const std::array<uint32_t, EQ_JUMP> jump = [] {
      std::array<uint32_t, EQ_JUMP> tab{};
      mz_jump_table(tab.data());
      return tab;
    }();
    PT_HIP(hipMemcpyAsync(c.eq_dev + 8, jump.data(), sizeof(uint32_t) * EQ_JUMP, hipMemcpyHostToDevice, c.stream));
  }
  int32_t *cnt = c.eq_node, *cnt1 = cnt + ncell, *nex = cnt1 + ncell, *list = nex + ncell;
  int32_t *cls = c.eq_id, *slot = cls + n;
  if (c.timing) timer_begin(c, PT_K_EQUALIZE);
  const int rc = equalize_body(c, cnt, cnt1, nex, list, cls, slot, nb);
  if (c.timing) timer_end(c, PT_K_EQUALIZE);
  return rc;
}

static int equalize_body(Context& c, int32_t* cnt, int32_t* cnt1, int32_t* nex, int32_t* list, int32_t* cls,
                         int32_t* slot, int nb) {
  const int nx = c.T.nx, nz = c.T.nz, ncell = (nx + 1) * (nz + 1);
  const int64_t n = c.T.ntr;
  PT_TRY(ensure_cell_order(c));
  // per-node counts from the cell-ordered storage (c.cstart / c.cend are
  // free between projections)
  hipLaunchKernelGGL(k_eq_fill, dim3(blocks(ncell, 256)), dim3(256), 0, c.stream, c.cstart, (int64_t)ncell, 0);
  hipLaunchKernelGGL(k_eq_fill, dim3(blocks(ncell, 256)), dim3(256), 0, c.stream, c.cend, (int64_t)ncell, 0);
  hipLaunchKernelGGL(k_eq_bounds, dim3(blocks(n, 256)), dim3(256), 0, c.stream, c.okey, n, ncell, c.cstart, c.cend);
  hipLaunchKernelGGL(k_eq_cells, dim3(blocks(ncell, 256)), dim3(256), 0, c.stream, c.cstart, c.cend, c.Ct, ncell, cnt,
                     cnt1);
  hipLaunchKernelGGL(k_eq_fill, dim3(1), dim3(64), 0, c.stream, c.eq_dev, (int64_t)8, 0);
  int32_t* nexT = list + ncell;
  hipLaunchKernelGGL(k_eq_excess, dim3(blocks(ncell, 256)), dim3(256), 0, c.stream, cnt, cnt1, nx, nz, c.P.buffer,
                     c.P.tpc, nex, nexT, c.eq_dev);
  int32_t h[8];
  PT_HIP(hipMemcpyAsync(h, c.eq_dev, sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
  PT_HIP(hipStreamSynchronize(c.stream));
  const int nfree = h[0];
  if (nfree == 0) return PT_OK;   // PS/composition.f90:150-155
  c.nequalize++;
  // donors and IDsave
  const double dist_r = c.P.dist_factor * c.T.dr, dist_s = c.P.dist_factor * c.T.ds;   // PS/basics.f90:276-277
  hipLaunchKernelGGL(k_eq_fill, dim3(blocks(n, 256)), dim3(256), 0, c.stream, cls, n, 0);
  hipLaunchKernelGGL(k_eq_donors, dim3(blocks(ncell, 4)), dim3(256), 0, c.stream, c.cstart, cnt, nex, ncell, c.r, c.s,
                     c.tid, dist_r, c.T.aspect - dist_r, dist_s, 1.0 - dist_s, cls, slot);
  hipLaunchKernelGGL(k_eq_bcount, dim3(nb), dim3(EQ_BLOCK), 0, c.stream, cls, n, c.eq_bsum);
  hipLaunchKernelGGL(k_eq_bscan, dim3(1), dim3(1024), 0, c.stream, c.eq_bsum, nb, c.eq_dev + 2);
  hipLaunchKernelGGL(k_eq_place, dim3(nb), dim3(EQ_BLOCK), 0, c.stream, cls, slot, n, c.eq_bsum, c.eq_dev + 2,
                     c.eq_save);
  // the lottery, from the context's mzran stream 0
  PT_HIP(hipMemcpyAsync(c.eq_dev + 4, c.rng, sizeof(int32_t) * 4, hipMemcpyHostToDevice, c.stream));
  Lottery L;
  L.nexT = nexT; L.actA = nexT + ncell; L.actB = nexT + 2 * ncell; L.listA = list; L.listB = nexT + 3 * ncell;
  L.save = c.eq_save; L.cnt = cnt; L.cnt1 = cnt1;
  L.jump = reinterpret_cast<const uint32_t*>(c.eq_dev + 8);
  L.r = c.r; L.s = c.s; L.Ct = c.Ct; L.dev = c.eq_dev;
  L.nx = nx; L.nz = nz; L.dr = c.T.dr; L.ds = c.T.ds; L.aspect = c.T.aspect; L.dist_r = dist_r; L.dist_s = dist_s;
  hipLaunchKernelGGL(k_eq_lottery, dim3(1), dim3(EQ_LOT), 0, c.stream, L);
  PT_HIP(hipMemcpyAsync(h, c.eq_dev, sizeof(int32_t) * 8, hipMemcpyDeviceToHost, c.stream));
  PT_HIP(hipStreamSynchronize(c.stream));
  if (h[2] + h[3] != nfree) {
    set_error("equalize: " + std::to_string(h[2] + h[3]) + " donors for nfree = " + std::to_string(nfree));
    return PT_ERR_INVALID;
  }
  const int moved = h[1];
  for (int q = 0; q < 4; q++) c.rng[q] = h[4 + q];
  if (moved > 0) PT_TRY(interpolate_T_at(c, c.eq_save, moved));
  PT_TRY(equalize_projection(c, moved > 0));
  c.eq_moved_last = moved;
  c.eq_moved_total += moved;
  return PT_OK;
}

}  // namespace pt
