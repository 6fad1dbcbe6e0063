// pt_tracers.hip -- Lagrangian tracer pipeline on gfx950.
//
//   k_bin            nint binning of tracer positions to node dual cells + domain
//                    check (PS/composition.f90:1002-1011)
//   k_cs_count/scan/fill  stable counting sort of (cell, slot) pairs that reads
//                    each cell's tracers from the storage order's neighbouring
//                    cells (the tracers move less than a cell per stage);
//                    (rocPRIM) stable radix sort when the storage order is not
//                    that of a previous projection or a tracer moved farther
//   k_cell_bounds    run bounds per cell after the radix sort (nempty = run length)
//   k_cell_reduce    per-cell sums of T*w, C*w, w for the 3x3 neighbour nodes
//                    (shape_func, PS/composition.f90:896-921), fixed order
//   k_node_combine   node = sum of its 9 cells' partials in a fixed order;
//                    Tratio/Cnew = ratio where sum w > 0 (PS/composition.f90:942-1070)
//   k_reorder        keeps the tracer SoA in cell order (ID map for host copies)
//   k_sym_* / k_t_*  enforceBCs / extendT / enforceBCs_comp
//                    (PS/energy.f90:558-630, PS/composition.f90:1279-1292)
//   k_smoothness / k_filter
//                    smoother_time / smoother_time_T Richardson passes
//                    (PS/smoother.f90:582-672, 75-139)
//   k_conduct_*      load_conductivity_array (PS/energy.f90:59-86)
//   k_source         tracer_space_uniform (PS/energy.f90:135-151)
//   k_itab           compute_derivatives for u, w and the source into one
//                    interleaved interpolation table (PS/composition.f90:1457-1478)
//   k_rk1..k_rk4     interpolate (x3) + the RK4 stage updates of energy_time
//                    (PS/composition.f90:1359-1393, PS/energy.f90:258-537)
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include "pt_internal.hpp"

namespace pt {

#define I3(i, k, p) ((long)((k) + 3) * (p) + ((i) + 3))
#define I5(i, k, p) ((long)((k) + 5) * (p) + ((i) + 5))

static dim3 grid2(int ni, int nk, dim3 b) { return dim3((ni + b.x - 1) / b.x, (nk + b.y - 1) / b.y); }

// nint(x) for the reference's NINT: round half away from zero
__device__ __forceinline__ int nint(double x) { return (int)round(x); }

// ---------------------------------------------------------------------------
// binning, stable sort by cell, per-cell sums, node combine
// ---------------------------------------------------------------------------
// key = dual cell of nint(r/dr), nint(s/ds); positions outside the domain
// (or NaN) raise flag 0 and get the sentinel key ncell (sorted last).  With
// the storage order's cells (okey), a tracer whose cell is more than one
// cell from its slot's in either direction raises flag 1.
__global__ void k_bin(const double* r, const double* s, int64_t n, int nx, int nz, double dr, double ds,
                      double aspect, int32_t* key, int32_t* flags, const int32_t* okey) {
  // one flag store per wavefront that raises it: an under-resolved flow can
  // move half the tracers farther than a cell, and same-address atomics
  // from every one of them serialize
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = j < n;
  const int ncell = (nx + 1) * (nz + 1);
  bool outside = false, far = false;
  if (live) {
    const double rv = r[j], sv = s[j];
    if (!(rv >= 0.0 && rv <= aspect && sv >= 0.0 && sv <= 1.0)) {
      outside = true;
      key[j] = ncell;
    } else {
      const int i = nint(rv / dr), k = nint(sv / ds);
      key[j] = k * (nx + 1) + i;
      if (okey) {
        const int o = okey[j], oi = o % (nx + 1), ok = o / (nx + 1);
        far = abs(oi - i) > 1 || abs(ok - k) > 1;
      }
    }
  }
  const unsigned long long bo = __ballot(outside), bf = __ballot(far);
  if ((threadIdx.x & 63) == 0) {
    if (bo) atomicOr(&flags[0], 1);
    if (bf) atomicOr(&flags[1], 1);
  }
}

// ---------------------------------------------------------------------------
// Stable counting sort of the (cell, slot) pairs when the tracers are stored
// in the cell order of an earlier projection and none moved more than one
// cell from its slot's cell.  A block owns CS_CELLS consecutive cells of one
// row; every tracer now in them is stored in the slot ranges of the cells of
// the three rows around, one column beyond the block's on each side.  Those
// three windows are read in ascending slot order, so each cell's tracers are
// found in ascending slot order: the result (perm, skey, cstart, cend) is the
// one the stable radix sort gives, element for element.
// ---------------------------------------------------------------------------
constexpr int CS_CELLS = 64;
constexpr int CS_THREADS = 256;

struct CsBlock {
  int cfirst, ncb, nwin, lo[3], hi[3];
};

// The block's cells and its windows: for each of the rows around (up to
// three), the slot range [lo, hi) of the cells ia..ib in the storage order,
// from the first non-empty cell's start to the last one's end (empty cells
// may hold any bounds).  One thread per (row, cell); ends with a barrier.
__device__ void cs_block(const int32_t* ostart, const int32_t* oend, int nx, int nz, CsBlock& b) {
  const int tid = threadIdx.x;
  const int k = blockIdx.y, i0 = blockIdx.x * CS_CELLS;
  const int ka = max(k - 1, 0), kb = min(k + 1, nz);
  const int ia = max(i0 - 1, 0), ib = min(i0 + min(CS_CELLS, nx + 1 - i0), nx);
  const int ncol = ib - ia + 1;   // at most CS_CELLS + 2
  if (tid == 0) {
    b.cfirst = k * (nx + 1) + i0;
    b.ncb = min(CS_CELLS, nx + 1 - i0);
    b.nwin = kb - ka + 1;
    for (int w = 0; w < 3; w++) {
      b.lo[w] = INT32_MAX;
      b.hi[w] = 0;
    }
  }
  __syncthreads();
  if (tid < 3 * ncol) {
    const int w = tid / ncol, i = ia + tid % ncol, kk = ka + w;
    if (kk <= kb) {
      const int c = kk * (nx + 1) + i;
      const int s0 = ostart[c], s1 = oend[c];
      if (s1 > s0) {
        atomicMin(&b.lo[w], s0);
        atomicMax(&b.hi[w], s1);
      }
    }
  }
  __syncthreads();
  if (tid < 3 && b.lo[tid] == INT32_MAX) {
    b.lo[tid] = 0;
    b.hi[tid] = 0;
  }
  __syncthreads();
}

// per-cell counts of the block's cells, and the block's total
__global__ void __launch_bounds__(CS_THREADS) k_cs_count(const int32_t* __restrict__ key,
                                                          const int32_t* __restrict__ ostart,
                                                          const int32_t* __restrict__ oend, int nx, int nz,
                                                          int32_t* __restrict__ ccount, int32_t* __restrict__ bsum) {
  __shared__ CsBlock sb;
  __shared__ int lcount[CS_CELLS];
  const int tid = threadIdx.x;
  if (tid < CS_CELLS) lcount[tid] = 0;
  cs_block(ostart, oend, nx, nz, sb);
  const int cfirst = sb.cfirst, ncb = sb.ncb;
  for (int w = 0; w < sb.nwin; w++)
    for (int q = sb.lo[w] + tid; q < sb.hi[w]; q += CS_THREADS) {
      const int lc = key[q] - cfirst;
      if (lc >= 0 && lc < ncb) atomicAdd(&lcount[lc], 1);
    }
  __syncthreads();
  if (tid < CS_CELLS) {
    int v = tid < ncb ? lcount[tid] : 0;
    if (tid < ncb) ccount[cfirst + tid] = v;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (tid == 0) bsum[blockIdx.y * gridDim.x + blockIdx.x] = v;
  }
}

// exclusive prefix sum of the block totals (one workgroup); flags[2] is
// raised when the total is not n (a tracer outside every window)
__global__ void __launch_bounds__(1024) k_cs_scan(int32_t* bsum, int nblk, int64_t n, int32_t* flags) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int per = (nblk + 1023) / 1024;
  const int a = min(t * per, nblk), b = min(a + per, nblk);
  int loc = 0;
  for (int u = a; u < b; u++) loc += bsum[u];
  part[t] = loc;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = part[t] - loc;   // exclusive
  for (int u = a; u < b; u++) {
    const int v = bsum[u];
    bsum[u] = run;
    run += v;
  }
  if (t == 1023 && (int64_t)part[1023] != n) atomicOr(&flags[2], 1);
}

// cell bounds, then the sorted (cell, slot) pairs: each window is read in
// chunks of CS_THREADS slots; within a wave a tracer's rank among the lanes
// before it with the same cell comes from ballots, across waves from the
// per-wave counts, across chunks from the running count per cell
__global__ void __launch_bounds__(CS_THREADS) k_cs_fill(const int32_t* __restrict__ key,
                                                         const int32_t* __restrict__ ostart,
                                                         const int32_t* __restrict__ oend, int nx, int nz,
                                                         const int32_t* __restrict__ ccount,
                                                         const int32_t* __restrict__ bsum, int32_t* __restrict__ cstart,
                                                         int32_t* __restrict__ cend, int32_t* __restrict__ perm,
                                                         int32_t* __restrict__ skey) {
  __shared__ CsBlock sb;
  __shared__ int lstart[CS_CELLS], running[CS_CELLS], cnt[CS_THREADS / 64][CS_CELLS];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  cs_block(ostart, oend, nx, nz, sb);
  const int cfirst = sb.cfirst, ncb = sb.ncb;
  if (tid < CS_CELLS) {
    // the block's cell starts: its total's offset plus a wave prefix sum
    const int v = tid < ncb ? ccount[cfirst + tid] : 0;
    int incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int u = __shfl_up(incl, o, 64);
      if (lane >= o) incl += u;
    }
    const int st = bsum[blockIdx.y * gridDim.x + blockIdx.x] + incl - v;
    lstart[tid] = st;
    running[tid] = 0;
    if (tid < ncb) {
      cstart[cfirst + tid] = st;
      cend[cfirst + tid] = st + v;
    }
  }
  __syncthreads();
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int w = 0; w < sb.nwin; w++) {
    const int lo = sb.lo[w], hi = sb.hi[w];
    for (int base = lo; base < hi; base += CS_THREADS) {
      cnt[wv][lane] = 0;
      const int q = base + tid;
      int lc = -1;
      if (q < hi) {
        const int d = key[q] - cfirst;
        if (d >= 0 && d < ncb) lc = d;
      }
      __syncthreads();
      int rank = 0;
      unsigned long long pending = __ballot(lc >= 0);
      while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const int lv = __shfl(lc, leader, 64);
        const unsigned long long m = __ballot(lc == lv);
        if (lc == lv) rank = __popcll(m & below);
        if (lane == leader) cnt[wv][lv] = __popcll(m);
        pending &= ~m;
      }
      __syncthreads();
      if (lc >= 0) {
        int pos = lstart[lc] + running[lc] + rank;
        for (int u = 0; u < wv; u++) pos += cnt[u][lc];
        perm[pos] = q;
        skey[pos] = cfirst + lc;
      }
      __syncthreads();
      if (tid < CS_CELLS) {
        int add = 0;
        for (int u = 0; u < CS_THREADS / 64; u++) add += cnt[u][tid];
        running[tid] += add;
      }
      __syncthreads();
    }
  }
}

// segment bounds of each cell in the sorted key array (empty cells: 0,0)
__global__ void k_cell_bounds(const int32_t* skey, int64_t n, int ncell, int32_t* cstart, int32_t* cend) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int c = skey[p];
  if (c >= ncell) return;
  if (p == 0 || skey[p - 1] != c) cstart[c] = (int32_t)p;
  if (p == n - 1 || skey[p + 1] != c) cend[c] = (int32_t)(p + 1);
}

// Per-cell partial sums of the projection.  A block owns PJ_CELLS
// consecutive cells; their tracers are one contiguous run of the
// (cell, slot)-sorted order, processed in chunks of PJ_CHUNK:
//  phase 1, one thread per tracer: coalesced loads through the permutation
//    (close to the identity, since the tracer arrays are kept in cell
//    order), the 3+3 hat factors max(0,1-|r-ri|/dr), max(0,1-|s-sk|/ds) of
//    shape_func (PS/composition.f90:896-921) for the nodes around the
//    tracer's cell, staged in LDS with T and C;
//  phase 2, one thread per (cell, neighbour node): w = wx*wz, and T*w, C*w,
//    w summed over the cell's tracers in ascending slot order (across
//    chunks too, so each sum is one sequential pass).
// Fixed summation order, no atomics: the result is deterministic.
constexpr int PJ_CELLS = 28;     // 28 cells x 9 neighbour nodes = 252 phase-2 threads
constexpr int PJ_THREADS = 256;
constexpr int PJ_CHUNK = 512;
static_assert(PJ_CHUNK % PJ_THREADS == 0, "a chunk is a whole number of tracers per thread");
__global__ void __launch_bounds__(PJ_THREADS) k_cell_reduce(const double* __restrict__ r, const double* __restrict__ s,
                                                            const double* __restrict__ Tt, const double* __restrict__ Ct,
                                                            const int32_t* __restrict__ perm,
                                                            const int32_t* __restrict__ skey, int64_t n,
                                                            const int32_t* __restrict__ cstart,
                                                            const int32_t* __restrict__ cend,
                                                            double* __restrict__ cellsum, int nx, int nz, double dr,
                                                            double ds) {
  // hat-factor rows padded by one double: an owner thread's row na (nb) is
  // then two LDS banks from its neighbours', not in the same bank
  __shared__ double hx[3][PJ_CHUNK + 1], hz[3][PJ_CHUNK + 1], sT[PJ_CHUNK], sC[PJ_CHUNK];
  __shared__ int32_t range[2];
  const int ncell = (nx + 1) * (nz + 1);
  const int cfirst = blockIdx.x * PJ_CELLS;
  const int clast = min(cfirst + PJ_CELLS, ncell);   // exclusive
  const int tid = threadIdx.x;
  // the block's tracers: from the first non-empty cell's segment start to
  // the last one's end (consecutive cells are consecutive in the sorted order)
  if (tid == 0) {
    range[0] = INT32_MAX;
    range[1] = 0;
  }
  __syncthreads();
  if (tid < clast - cfirst) {
    const int32_t b0 = cstart[cfirst + tid], b1 = cend[cfirst + tid];
    if (b1 > b0) {
      atomicMin(&range[0], b0);
      atomicMax(&range[1], b1);
    }
  }
  // phase-2 role: cell c, neighbour node (a, b) of the 3x3 around it
  const int c = cfirst + tid / 9, ab = tid % 9, na = ab / 3, nb = ab % 3;
  const bool owner = tid < 9 * PJ_CELLS && c < clast;
  int64_t b0 = 0, b1 = 0;
  if (owner) {
    b0 = cstart[c];
    b1 = cend[c];
  }
  double txs = 0.0, cxs = 0.0, ws = 0.0;
  __syncthreads();
  const int64_t t0 = range[0] == INT32_MAX ? 0 : range[0], t1 = range[0] == INT32_MAX ? 0 : range[1];
  for (int64_t base = t0; base < t1; base += PJ_CHUNK) {
    const int cnt = (int)min((int64_t)PJ_CHUNK, t1 - base);
    // the thread's PJ_Q tracers of the chunk: slot and permutation loads
    // (clamped into the chunk) for all of them, then their r, s, T, C, so
    // the chunk waits for two memory round trips
    constexpr int PJ_Q = PJ_CHUNK / PJ_THREADS;
    int32_t jq[PJ_Q], kq[PJ_Q];
#pragma unroll
    for (int u = 0; u < PJ_Q; u++) {
      const int q = min(tid + u * PJ_THREADS, cnt - 1);
      jq[u] = perm[base + q];
      kq[u] = skey[base + q];
    }
    double rq[PJ_Q], sq[PJ_Q], tq[PJ_Q], cq[PJ_Q];
#pragma unroll
    for (int u = 0; u < PJ_Q; u++) {
      rq[u] = r[jq[u]];
      sq[u] = s[jq[u]];
      tq[u] = Tt[jq[u]];
      cq[u] = Ct[jq[u]];
    }
#pragma unroll
    for (int u = 0; u < PJ_Q; u++) {
      const int q = tid + u * PJ_THREADS;
      if (q >= cnt) continue;
      const int key = kq[u];
      const int i = key % (nx + 1), k = key / (nx + 1);
#pragma unroll
      for (int e = 0; e < 3; e++) {
        const double tx = 1.0 - fabs(rq[u] - (double)(i + e - 1) * dr) / dr;
        hx[e][q] = tx < 0.0 ? 0.0 : tx;
        const double tz = 1.0 - fabs(sq[u] - (double)(k + e - 1) * ds) / ds;
        hz[e][q] = tz < 0.0 ? 0.0 : tz;
      }
      sT[q] = tq[u];
      sC[q] = cq[u];
    }
    __syncthreads();
    if (owner) {
      const int q0 = (int)(max(b0, base) - base), q1 = (int)(min(b1, base + cnt) - base);
      int q = q0;
      // four tracers' LDS operands in flight; the sums stay in slot order
      for (; q + 4 <= q1; q += 4) {
        double w[4], tv[4], cv[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          w[u] = hx[na][q + u] * hz[nb][q + u];
          tv[u] = sT[q + u];
          cv[u] = sC[q + u];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
          txs = txs + tv[u] * w[u];
          cxs = cxs + cv[u] * w[u];
          ws = ws + w[u];
        }
      }
      for (; q < q1; q++) {
        const double w = hx[na][q] * hz[nb][q];
        txs = txs + sT[q] * w;
        cxs = cxs + sC[q] * w;
        ws = ws + w;
      }
    }
    __syncthreads();
  }
  if (owner) {
    double* out = cellsum + (long)c * 27 + ab * 3;
    out[0] = txs;
    out[1] = cxs;
    out[2] = ws;
  }
}

// node (i,k) sums the partials of the 9 cells around it in a fixed order;
// Tratio, Cnew = ratio where the weight sum is positive, nempty = own count
__global__ void k_node_combine(const double* __restrict__ cellsum, const int32_t* cstart, const int32_t* cend,
                               double* Tratio, double* Cnew, int32_t* nempty, int nx, int nz) {
  int i = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y * blockDim.y + threadIdx.y;
  if (i > nx || k > nz) return;
  double txs = 0.0, cxs = 0.0, ssum = 0.0;
  for (int a = -1; a <= 1; a++) {
    int ic = i + a;
    if (ic < 0 || ic > nx) continue;
    for (int b = -1; b <= 1; b++) {
      int kc = k + b;
      if (kc < 0 || kc > nz) continue;
      long c = (long)kc * (nx + 1) + ic;
      int slot = ((-a) + 1) * 3 + ((-b) + 1);   // this node seen from cell (ic,kc)
      txs = txs + cellsum[c * 27 + slot * 3 + 0];
      cxs = cxs + cellsum[c * 27 + slot * 3 + 1];
      ssum = ssum + cellsum[c * 27 + slot * 3 + 2];
    }
  }
  long own = (long)k * (nx + 1) + i;
  nempty[own] = cend[own] - cstart[own];
  if (ssum > 0.0) {
    Tratio[I3(i, k, nx + 7)] = txs / ssum;
    Cnew[I5(i, k, nx + 11)] = cxs / ssum;
  }
}

// ---------------------------------------------------------------------------
// boundary conditions
// ---------------------------------------------------------------------------
// side ghosts (symmetric), rows 0..nz, for a field with halo h and pitch p
__global__ void k_sym_side(double* f, int nx, int nz, int h, int p) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (nz + 1) * h) return;
  int k = t / h, i = 1 + t % h;
  long row = (long)(k + h) * p + h;
  f[row - i] = f[row + i];
  f[row + nx + i] = f[row + nx - i];
}
// enforceBCs_comp top/bottom: symmetric rows for all columns -h..nx+h
__global__ void k_sym_tb(double* f, int nx, int nz, int h, int p) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (nx + 1 + 2 * h) * h) return;
  int i = -h + t / h, k = 1 + t % h;
  f[(long)(h - k) * p + (i + h)] = f[(long)(h + k) * p + (i + h)];
  f[(long)(h + nz + k) * p + (i + h)] = f[(long)(h + nz - k) * p + (i + h)];
}
// T wall rows: ghost = 2*Twall + (2/D2(0))*heating*ds2 - mirror (isothermal) or
// mirror (insulating bottom), then the wall rows themselves (enforceBCs/extendT)
__global__ void k_t_tb(double* f, const double* Cvis, int nx, int nz, int h, int p, int p5, int Tbot, double Tb,
                       double Tt, double two_over_d20, double ds2, double H, double Htype) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (nx + 1 + 2 * h) * h) return;
  int i = -h + t / h, k = 1 + t % h;
  long ci = i + 5;
  double cb = Cvis[(long)(0 + 5) * p5 + ci], ct = Cvis[(long)(nz + 5) * p5 + ci];
  double hb = Htype * cb + H * (1.0 - (0.0 + cb));
  double ht = Htype * ct + H * (1.0 - (0.0 + ct));
  double in_b = f[(long)(h + k) * p + (i + h)];
  double in_t = f[(long)(h + nz - k) * p + (i + h)];
  if (Tbot == 0) f[(long)(h - k) * p + (i + h)] = 2.0 * Tb + two_over_d20 * hb * ds2 - in_b;
  else f[(long)(h - k) * p + (i + h)] = in_b;
  f[(long)(h + nz + k) * p + (i + h)] = 2.0 * Tt + two_over_d20 * ht * ds2 - in_t;
}
__global__ void k_t_walls(double* f, int nx, int nz, int h, int p, int Tbot, double Tb, double Tt) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nx + 1 + 2 * h) return;
  int i = -h + t;
  if (Tbot == 0) f[(long)h * p + (i + h)] = Tb;
  f[(long)(h + nz) * p + (i + h)] = Tt;
}
__global__ void k_copy_nodes(const double* a, int pa, int ha, double* b, int pb, int hb, int nx, int nz) {
  int i = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y * blockDim.y + threadIdx.y;
  if (i > nx || k > nz) return;
  b[(long)(k + hb) * pb + (i + hb)] = a[(long)(k + ha) * pa + (i + ha)];
}

static void launch1(Context& c, void (*kern)(double*, int, int, int, int), double* f, int nthreads, int nx, int nz,
                    int h, int p) {
  hipLaunchKernelGGL(kern, dim3((nthreads + 255) / 256), dim3(256), 0, c.stream, f, nx, nz, h, p);
}

// enforceBCs(T) on a halo-3 field
static void t_bcs(Context& c, double* f) {
  int nx = c.T.nx, nz = c.T.nz;
  launch1(c, k_sym_side, f, (nz + 1) * 3, nx, nz, 3, nx + 7);
  int n = (nx + 7) * 3;
  hipLaunchKernelGGL(k_t_tb, dim3((n + 255) / 256), dim3(256), 0, c.stream, f, c.Cvis, nx, nz, 3, nx + 7, nx + 11,
                     c.P.Tbot, c.Tbot_iso, c.Ttop_iso, 2.0 / c.T.D2[3], c.T.ds2, c.P.H, c.P.Htype);
  hipLaunchKernelGGL(k_t_walls, dim3((nx + 7 + 255) / 256), dim3(256), 0, c.stream, f, nx, nz, 3, nx + 7, c.P.Tbot,
                     c.Tbot_iso, c.Ttop_iso);
}
// enforceBCs_comp on a halo-5 field
static void c_bcs(Context& c, double* f) {
  int nx = c.T.nx, nz = c.T.nz;
  launch1(c, k_sym_side, f, (nz + 1) * 5, nx, nz, 5, nx + 11);
  launch1(c, k_sym_tb, f, (nx + 11) * 5, nx, nz, 5, nx + 11);
}
// extendT: Textend(0:nx,0:nz) = Tratio, halo-5 wall ghosts
static void extend_t(Context& c) {
  int nx = c.T.nx, nz = c.T.nz;
  dim3 b16(64, 4);
  hipLaunchKernelGGL(k_copy_nodes, grid2(nx + 1, nz + 1, b16), b16, 0, c.stream, c.Tratio, nx + 7, 3, c.Textend,
                     nx + 11, 5, nx, nz);
  launch1(c, k_sym_side, c.Textend, (nz + 1) * 5, nx, nz, 5, nx + 11);
  int n = (nx + 11) * 5;
  hipLaunchKernelGGL(k_t_tb, dim3((n + 255) / 256), dim3(256), 0, c.stream, c.Textend, c.Cvis, nx, nz, 5, nx + 11,
                     nx + 11, c.P.Tbot, c.Tbot_iso, c.Ttop_iso, 2.0 / c.T.D2[3], c.T.ds2, c.P.H, c.P.Htype);
  hipLaunchKernelGGL(k_t_walls, dim3((nx + 11 + 255) / 256), dim3(256), 0, c.stream, c.Textend, nx, nz, 5, nx + 11,
                     c.P.Tbot, c.Tbot_iso, c.Ttop_iso);
}

// ---------------------------------------------------------------------------
// filters
// ---------------------------------------------------------------------------
// smoothness (PS/smoother.f90:75-95): block maxima over i=0..nx-1, k=1..nz-1
__global__ void k_smoothness(const double* f, int nx, int nz, int h, int p, double* partial) {
  __shared__ double red[256];
  int tid = threadIdx.y * blockDim.x + threadIdx.x;
  int i = blockIdx.x * blockDim.x + threadIdx.x, k = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  double d = 0.0;
  if (i <= nx - 1 && k <= nz - 1) {
    auto F = [&](int a, int b) { return f[(long)(b + h) * p + (a + h)]; };
    double t = F(i, k);
    double a1 = fabs(F(i, k - 1) - t), a2 = fabs(F(i + 1, k - 1) - t), a3 = fabs(F(i + 1, k) - t),
           a4 = fabs(F(i, k + 1) - t), a5 = fabs(F(i + 1, k + 1) - t);
    d = fmax(fmax(fmax(fmax(a1, a2), a3), a4), a5);
  }
  red[tid] = d;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (tid < st) red[tid] = fmax(red[tid], red[tid + st]);
    __syncthreads();
  }
  if (tid == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

// one Richardson pass: out = T0 + (c*dts)*lap(T0) (clamped to [0,1] for C)
__global__ void k_filter(const double* T0, double* out, int nx, int nz, int h, int p, double cdts, int clamp01,
                         double dr, double ds, double dr2, double ds2, double d10, double d11, double d12, double d13,
                         double d14, double d20, double d21, double d22, double d23, double d24) {
  int i = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y * blockDim.y + threadIdx.y;
  if (i > nx || k > nz) return;
  const double D1[5] = {d10, d11, d12, d13, d14}, D2[5] = {d20, d21, d22, d23, d24};
  auto F = [&](int a, int b) { return T0[(long)(b + h) * p + (a + h)]; };
  double sr = 0.0, ss = 0.0, srr = 0.0, sss = 0.0;
  for (int q = -2; q <= 2; q++) sr = sr + D1[q + 2] * F(i + q, k);
  for (int q = -2; q <= 2; q++) ss = ss + D1[q + 2] * F(i, k + q);
  for (int q = -2; q <= 2; q++) srr = srr + D2[q + 2] * F(i + q, k);
  for (int q = -2; q <= 2; q++) sss = sss + D2[q + 2] * F(i, k + q);
  double Tr = sr / dr, Ts = ss / ds, Trr = srr / dr2, Tss = sss / ds2;
  double Tx = Tr / 1.0, Tz = Ts / 1.0;
  double Txx = (Trr - Tx * 0.0) / 1.0, Tzz = (Tss - Tz * 0.0) / 1.0;
  double temp = Txx + Tzz;
  double v = F(i, k) + cdts * temp;
  if (clamp01) {
    if (v < 0.0) v = 0.0;
    if (v > 1.0) v = 1.0;
  }
  out[(long)(k + h) * p + (i + h)] = v;
}

static int smoothness_host(Context& c, const double* f, int h, int p, double* out) {
  dim3 b16(64, 4);
  dim3 g = grid2(c.T.nx, c.T.nz - 1, b16);
  hipLaunchKernelGGL(k_smoothness, g, b16, 0, c.stream, f, c.T.nx, c.T.nz, h, p, c.red);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, c.stream, c.red, (int)(g.x * g.y), (int)RED_MAX, c.red_out);
  PT_HIP(hipMemcpyAsync(c.h_red, c.red_out, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  PT_HIP(hipStreamSynchronize(c.stream));
  *out = c.h_red[0];
  return PT_OK;
}

static void filter_pass(Context& c, const double* T0, double* out, int h, double cdts, int clamp01) {
  dim3 b16(64, 4);
  const double* D1 = c.T.D1;
  const double* D2 = c.T.D2;
  hipLaunchKernelGGL(k_filter, grid2(c.T.nx + 1, c.T.nz + 1, b16), b16, 0, c.stream, T0, out, c.T.nx, c.T.nz, h,
                     c.T.nx + 1 + 2 * h, cdts, clamp01, c.T.dr, c.T.ds, c.T.dr2, c.T.ds2, D1[1], D1[2], D1[3], D1[4],
                     D1[5], D2[1], D2[2], D2[3], D2[4], D2[5]);
}

// ---------------------------------------------------------------------------
// Filter loops driven on the device.  The reference checks the field's
// smoothness before every pass and stops at the first check below the
// threshold (PS/smoother.f90:582-672).  Here one kernel per pass does both
// halves of a loop iteration on its input field: the check (a partial
// maximum per check workgroup, stored in the pass's slot of a ring) and the
// pass itself, written to the other buffer.  The pass is speculative: every
// workgroup of the next pass kernel reduces the partials (max is order-free
// and exact) and, when the check is below the threshold, the kernel records
// the pass count and raises the `done` flag instead of going on, so the
// field the loop returns is the input of the pass whose check succeeded, as
// in the reference.  No atomics: 256 same-address atomics per pass cost more
// than the pass itself.  The host enqueues pass kernels in batches without
// waiting; once `done` is set every later kernel returns at once.
// Passes alternate between two buffers; each writes its field's ghost
// images itself (the reference's enforceBCs / enforceBCs_comp after the pass).
// ---------------------------------------------------------------------------
// At most this many workgroups of a pass kernel take a share of the check,
// striding over its points, each leaving one partial maximum.
constexpr int SC_BLOCKS = 256;
constexpr int FSLOTS = 64;   // ring of check slots, SC_BLOCKS partial maxima (FP64 bit patterns) each

__global__ void k_filter_init(int32_t* ctl, unsigned long long* sm) {
  const int t = threadIdx.x;
  if (t < 4) ctl[t] = 0;
  for (int u = t; u < FSLOTS * SC_BLOCKS; u += blockDim.x) sm[u] = 0ull;
}

struct FilterArgs {
  int nx, nz, h, p;
  double dr, ds, dr2, ds2;
  double D1[5], D2[5];
  int clamp01;          // C: clamp to [0,1], symmetric ghosts (enforceBCs_comp)
  // T (clamp01 == 0): enforceBCs with the isothermal / insulating walls
  int Tbot;
  double Tb, Tt, two_over_d20, H, Htype;
  const double* Cvis;   // halo 5, for the heating term of the T wall ghosts
  int check_h, check_p; // the check's view of the field: halo and row pitch
};


// Pass j of the loop, in -> out, on a (64, 4) block grid over the nodes
// 0..nx x 0..nz.  First the decision on the previous kernel's check (the
// check before pass j-1, in slot prev): below the threshold, the loop ends
// with j-2 passes made.  Otherwise: the check before pass j (the max
// 5-neighbour jump of `in` viewed with halo check_h and pitch check_p,
// smoothness in PS/smoother.f90:75-95, over i = 0..nx-1, k = 1..nz-1) into
// slot cur, and the pass (one Richardson pass, smoother_space at every node,
// then the ghost images) into out.
__global__ void __launch_bounds__(256) k_filter_pass(const double* in, double* out, FilterArgs a, int32_t* ctl,
                                                     unsigned long long* sm, int prev, int cur, int pass,
                                                     double cdts, double delta) {
  const int bid = blockIdx.y * gridDim.x + blockIdx.x;
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  const int nx = a.nx, nz = a.nz, h = a.h, p = a.p;
  // XCD-aware tiles: workgroups are dispatched round-robin over the 8 XCDs
  // (bid % 8); each XCD gets a contiguous band of tiles, the same every
  // pass, so a tile's field and most of its halo stay in that XCD's L2
  const int nblk = (int)(gridDim.x * gridDim.y);
  const int q8 = nblk / 8, r8 = nblk % 8, xcd = bid % 8;
  const int lb = xcd * q8 + min(xcd, r8) + bid / 8;
  const int tx = lb % (int)gridDim.x, ty = lb / (int)gridDim.x;
  const int i = tx * blockDim.x + threadIdx.x, k = ty * blockDim.y + threadIdx.y;
  const bool node = i <= nx && k <= nz;
  const int ncheck = min(nblk, SC_BLOCKS);
  // The control word, the previous check and every load of `in` this
  // thread needs are issued together, before the decision: the loads do
  // not depend on it, and a kernel after the loop's end only reads.
  const int32_t done = ctl[0];
  // the previous check: one partial maximum per check workgroup
  const double prev_part = pass > 1 && tid < ncheck ? __longlong_as_double((long long)sm[(long)prev * SC_BLOCKS + tid])
                                                      : 0.0;
  double y = 0.0;
  if (node) {
    auto F = [&](int u, int v) { return in[(long)(v + h) * p + (u + h)]; };
    double sr = 0.0, ss = 0.0, srr = 0.0, sss = 0.0;
    for (int q = -2; q <= 2; q++) sr = sr + a.D1[q + 2] * F(i + q, k);
    for (int q = -2; q <= 2; q++) ss = ss + a.D1[q + 2] * F(i, k + q);
    for (int q = -2; q <= 2; q++) srr = srr + a.D2[q + 2] * F(i + q, k);
    for (int q = -2; q <= 2; q++) sss = sss + a.D2[q + 2] * F(i, k + q);
    double Tr = sr / a.dr, Ts = ss / a.ds, Trr = srr / a.dr2, Tss = sss / a.ds2;
    double Tx = Tr / 1.0, Tz = Ts / 1.0;
    double Txx = (Trr - Tx * 0.0) / 1.0, Tzz = (Tss - Tz * 0.0) / 1.0;
    double temp = Txx + Tzz;
    y = F(i, k) + cdts * temp;
  }
  double d = 0.0;
  if (bid < ncheck) {
    const long npts = (long)nx * (nz - 1);
    auto G = [&](int u, int v) { return in[(long)(v + a.check_h) * a.check_p + (u + a.check_h)]; };
    // with a full set of check workgroups, the 32 on each XCD stride over
    // that XCD's eighth of the points, near its band of tiles
    long e0 = (long)bid * 256 + tid, e_end = npts, estride = (long)ncheck * 256;
    if (ncheck == SC_BLOCKS) {
      const int x = bid % 8, jx = bid / 8;
      e0 = npts * x / 8 + (long)jx * 256 + tid;
      e_end = npts * (x + 1) / 8;
      estride = (long)(SC_BLOCKS / 8) * 256;
    }
    for (long e = e0; e < e_end; e += estride) {
      const int kk = 1 + (int)(e / nx), ii = (int)(e - (long)(kk - 1) * nx);
      const double t = G(ii, kk);
      const double a1 = fabs(G(ii, kk - 1) - t), a2 = fabs(G(ii + 1, kk - 1) - t), a3 = fabs(G(ii + 1, kk) - t),
                   a4 = fabs(G(ii, kk + 1) - t), a5 = fabs(G(ii + 1, kk + 1) - t);
      d = fmax(d, fmax(fmax(fmax(fmax(a1, a2), a3), a4), a5));
    }
  }
  if (done) return;
  // both block maxima at once: the previous check's (the decision) and this
  // pass's check (this block's partial)
  __shared__ double red[2][4];
  double pm = prev_part;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    pm = fmax(pm, __shfl_xor(pm, o, 64));
    d = fmax(d, __shfl_xor(d, o, 64));
  }
  if ((tid & 63) == 0) {
    red[0][tid >> 6] = pm;
    red[1][tid >> 6] = d;
  }
  __syncthreads();
  if (pass > 1 && fmax(fmax(red[0][0], red[0][1]), fmax(red[0][2], red[0][3])) < delta) {
    if (bid == 0 && tid == 0) {
      ctl[1] = pass - 2;   // passes made before the successful check
      ctl[0] = 1;
    }
    return;
  }
  if (bid < ncheck && tid == 0)
    sm[(long)cur * SC_BLOCKS + bid] =
        (unsigned long long)__double_as_longlong(fmax(fmax(red[1][0], red[1][1]), fmax(red[1][2], red[1][3])));
  if (!node) return;
  auto W = [&](int u, int v, double val) { out[(long)(v + h) * p + (u + h)] = val; };
  // the columns whose ghost images equal this node's column: itself, and
  // its mirror beyond a side wall when it lies within h of that wall
  int cols[2] = {i, i};
  int ncol = 1;
  if (i >= 1 && i <= h) cols[ncol++] = -i;
  else if (i >= nx - h && i <= nx - 1) cols[ncol++] = 2 * nx - i;
  if (a.clamp01) {
    if (y < 0.0) y = 0.0;
    if (y > 1.0) y = 1.0;
    for (int c = 0; c < ncol; c++) {
      W(cols[c], k, y);
      if (k >= 1 && k <= h) W(cols[c], -k, y);
      if (k >= nz - h && k <= nz - 1) W(cols[c], 2 * nz - k, y);
    }
    return;
  }
  // T: the wall rows hold the isothermal values (the insulating bottom keeps
  // its filtered value); ghost rows are 2 T_wall + (2/D2(0)) H ds^2 - mirror
  // (isothermal) or the mirror (insulating bottom)
  if (k == 0 && a.Tbot == 0) y = a.Tb;
  if (k == nz) y = a.Tt;
  for (int c = 0; c < ncol; c++) {
    const int col = cols[c];
    W(col, k, y);
    if (k >= 1 && k <= h) {
      const double cb = a.Cvis[(long)5 * (nx + 11) + (col + 5)];
      const double hb = a.Htype * cb + a.H * (1.0 - (0.0 + cb));
      W(col, -k, a.Tbot == 0 ? 2.0 * a.Tb + a.two_over_d20 * hb * a.ds2 - y : y);
    }
    if (k >= nz - h && k <= nz - 1) {
      const double ct = a.Cvis[(long)(nz + 5) * (nx + 11) + (col + 5)];
      const double ht = a.Htype * ct + a.H * (1.0 - (0.0 + ct));
      W(col, 2 * nz - k, 2.0 * a.Tt + a.two_over_d20 * ht * a.ds2 - y);
    }
  }
}

// one device-driven filter loop on x0 (x1: the second buffer).  cdts_of(j)
// is pass j's c*dts; on return x0 holds the filtered field with its ghosts
// and *passes the number of passes made before the successful check
static int filter_loop(Context& c, double* x0, double* x1, const FilterArgs& fa, double delta,
                       double (*cdts_of)(const Context&, int), int* passes, int* hint) {
  dim3 b16(64, 4);
  dim3 gflt = grid2(c.T.nx + 1, c.T.nz + 1, b16);
  unsigned long long* sm = c.d_fsm;
  if (c.timing) timer_begin(c, PT_K_FILTER);
  hipLaunchKernelGGL(k_filter_init, dim3(1), dim3(256), 0, c.stream, c.d_fctl, sm);
  const int cap = 100000;
  // the first batch covers the previous call's pass count (consecutive
  // projections need similar counts) and the kernel that decides on its
  // last check; later batches double
  int j = 0, batch = std::min(std::max(*hint + 3, 8), 1024);
  for (;;) {
    for (int b = 0; b < batch; b++) {
      j++;
      const double* in = (j & 1) ? x0 : x1;
      double* out = (j & 1) ? x1 : x0;
      hipLaunchKernelGGL(k_filter_pass, gflt, b16, 0, c.stream, in, out, fa, c.d_fctl, sm, (j - 1) % FSLOTS,
                         j % FSLOTS, j, cdts_of(c, j), delta);
    }
    PT_HIP(hipMemcpyAsync(c.h_fctl, c.d_fctl, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
    PT_HIP(hipStreamSynchronize(c.stream));
    if (c.h_fctl[0]) break;
    if (j > cap) {
      set_error("filter did not reach its smoothness threshold in 100000 passes");
      return PT_ERR_INVALID;
    }
    batch = std::min(2 * batch, 256);
  }
  if (c.timing) timer_end(c, PT_K_FILTER);
  *passes = c.h_fctl[1];
  *hint = *passes;
  if (*passes & 1) {
    // the loop ended on the second buffer
    const size_t bytes = sizeof(double) * (size_t)(c.T.nx + 1 + 2 * fa.h) * (c.T.nz + 1 + 2 * fa.h);
    PT_HIP(hipMemcpyAsync(x0, x1, bytes, hipMemcpyDeviceToDevice, c.stream));
  }
  return PT_OK;
}

static FilterArgs filter_args(const Context& c, int h, int clamp01) {
  FilterArgs a;
  a.nx = c.T.nx; a.nz = c.T.nz; a.h = h; a.p = c.T.nx + 1 + 2 * h;
  a.dr = c.T.dr; a.ds = c.T.ds; a.dr2 = c.T.dr2; a.ds2 = c.T.ds2;
  for (int q = 0; q < 5; q++) { a.D1[q] = c.T.D1[q + 1]; a.D2[q] = c.T.D2[q + 1]; }
  a.clamp01 = clamp01;
  a.Tbot = c.P.Tbot; a.Tb = c.Tbot_iso; a.Tt = c.Ttop_iso; a.two_over_d20 = 2.0 / c.T.D2[3];
  a.H = c.P.H; a.Htype = c.P.Htype; a.Cvis = c.Cvis;
  a.check_h = 3; a.check_p = c.T.nx + 7;
  return a;
}

// C passes: the first ten with the Chebyshev-ordered steps that damp the
// unresolved frequencies first, then c = 1 (PS/smoother.f90:596-617)
static double cdts_comp(const Context& c, int j) {
  double cc = 1.0;
  if (j <= 10) {
    const int Niter = 10, ntemp = Niter - j + 1;
    const double lb = -5.0, la = -1.0;
    cc = 2.0 / (-lb - la + (lb - la) * std::cos((double)(2 * ntemp - 1) * 3.1415926535897932 / (double)(2 * Niter)));
  }
  return cc * (c.T.dts * c.P.courantC);
}
static double cdts_temp(const Context& c, int) { return c.T.dts * c.P.courantT; }

static int smoother_time_host(Context& c);

// smoother_time (C): PS/smoother.f90:582-636
static int smoother_time(Context& c) {
  size_t b5 = sizeof(double) * (size_t)(c.T.nx + 11) * (c.T.nz + 11);
  PT_HIP(hipMemcpyAsync(c.Cvis, c.Cnew, b5, hipMemcpyDeviceToDevice, c.stream));
  PT_HIP(hipMemcpyAsync(c.Cbuoy, c.Cnew, b5, hipMemcpyDeviceToDevice, c.stream));
  if (!c.filter_on_device) return smoother_time_host(c);
  // the reference passes the halo-5 C array to smoothness, whose dummy has
  // the halo-3 shape (PS/smoother.f90:79, :602): sequence association makes
  // it read the C storage with the halo-3 row length
  int P = 0;
  PT_TRY(filter_loop(c, c.Cvis, c.TxS, filter_args(c, 5, 1), c.T.delta_C, cdts_comp, &P, &c.hint_comp));
  // n_comp counts the outer iterations after the first ten-pass one
  c.n_comp = P < 10 ? 0 : P - 9;
  return PT_OK;
}

// smoother_time with a host round trip per check, the loop as the reference
// writes it (kept as the cross-check of the device-driven loop)
static int smoother_time_host(Context& c) {
  size_t b5 = sizeof(double) * (size_t)(c.T.nx + 11) * (c.T.nz + 11);
  int Niter = 10, nn = 0;
  double lb = -5.0, la = -1.0;
  bool trigger = false;
  for (;;) {
    if (nn > 100000) {
      set_error("smoother_time: C filter did not reach delta_C");
      return PT_ERR_INVALID;
    }
    for (int n = 1; n <= Niter; n++) {
      // the reference passes the halo-5 C array to smoothness, whose dummy has
      // the halo-3 shape (PS/smoother.f90:79, :602): sequence association
      // makes it read the C storage with the halo-3 row length
      double diff;
      PT_TRY(smoothness_host(c, c.Cvis, 3, c.T.nx + 7, &diff));
      if (diff < c.T.delta_C) { trigger = true; break; }
      double cc;
      if (nn == 0) {
        int ntemp = Niter - n + 1;
        cc = 2.0 / (-lb - la + (lb - la) * std::cos((double)(2 * ntemp - 1) * 3.1415926535897932 / (double)(2 * Niter)));
      } else {
        cc = 1.0;
      }
      // T0 = Cvis (copy), Cvis = filter(T0); c*dts*temp = (c*dts)*temp
      PT_HIP(hipMemcpyAsync(c.TxS, c.Cvis, b5, hipMemcpyDeviceToDevice, c.stream));
      filter_pass(c, c.TxS, c.Cvis, 5, cc * (c.T.dts * c.P.courantC), 1);
      c_bcs(c, c.Cvis);
    }
    if (trigger) break;
    Niter = 1;
    nn++;
  }
  c.n_comp = nn;
  return PT_OK;
}

// smoother_time_T: PS/smoother.f90:638-672 (Tin = Tratio)
static int smoother_time_T(Context& c) {
  size_t b3 = sizeof(double) * (size_t)(c.T.nx + 7) * (c.T.nz + 7);
  PT_HIP(hipMemcpyAsync(c.Tvis, c.Tratio, b3, hipMemcpyDeviceToDevice, c.stream));
  PT_HIP(hipMemcpyAsync(c.Tbuoy, c.Tratio, b3, hipMemcpyDeviceToDevice, c.stream));
  if (c.P.RaT == 0.0 || c.P.ivis == 0) return PT_OK;
  if (c.filter_on_device) {
    int P = 0;
    PT_TRY(filter_loop(c, c.Tvis, c.CxS, filter_args(c, 3, 0), c.T.delta_T, cdts_temp, &P, &c.hint_temp));
    c.n_temp = P;
    return PT_OK;
  }
  int n = 0;
  for (;;) {
    if (n > 100000) {
      set_error("smoother_time_T: T filter did not reach delta_T");
      return PT_ERR_INVALID;
    }
    double diff;
    PT_TRY(smoothness_host(c, c.Tvis, 3, c.T.nx + 7, &diff));
    if (diff < c.T.delta_T) { c.n_temp = n; break; }
    PT_HIP(hipMemcpyAsync(c.CxS, c.Tvis, b3, hipMemcpyDeviceToDevice, c.stream));
    filter_pass(c, c.CxS, c.Tvis, 3, c.T.dts * c.P.courantT, 0);
    t_bcs(c, c.Tvis);
    n++;
  }
  return PT_OK;
}

// load_conductivity_array: PS/energy.f90:59-86
__global__ void k_conduct_nodes(const double* Cvis, double* cond, int nx, int nz, double cf) {
  int i = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y * blockDim.y + threadIdx.y;
  if (i > nx || k > nz) return;
  cond[I5(i, k, nx + 11)] = 1.0 + Cvis[I5(i, k, nx + 11)] * (cf - 1.0);
}
__global__ void k_conduct_grad(const double* cond, double* cr, double* cs, int nx, int nz, double dr, double ds,
                               double d0, double d1, double d2, double d3, double d4) {
  int i = -3 + (int)(blockIdx.x * blockDim.x + threadIdx.x), k = -3 + (int)(blockIdx.y * blockDim.y + threadIdx.y);
  if (i > nx + 3 || k > nz + 3) return;
  const double D[5] = {d0, d1, d2, d3, d4};
  double sr = 0.0, ss = 0.0;
  for (int q = -2; q <= 2; q++) sr = sr + D[q + 2] * cond[I5(i + q, k, nx + 11)];
  for (int q = -2; q <= 2; q++) ss = ss + D[q + 2] * cond[I5(i, k + q, nx + 11)];
  cr[I3(i, k, nx + 7)] = sr / dr;
  cs[I3(i, k, nx + 7)] = ss / ds;
}
static void load_conductivity(Context& c) {
  int nx = c.T.nx, nz = c.T.nz;
  dim3 b16(64, 4);
  hipLaunchKernelGGL(k_conduct_nodes, grid2(nx + 1, nz + 1, b16), b16, 0, c.stream, c.Cvis, c.cond, nx, nz,
                     c.P.conduct_factor);
  launch1(c, k_sym_side, c.cond, (nz + 1) * 5, nx, nz, 5, nx + 11);
  launch1(c, k_sym_tb, c.cond, (nx + 11) * 5, nx, nz, 5, nx + 11);
  const double* D1 = c.T.D1;
  hipLaunchKernelGGL(k_conduct_grad, grid2(nx + 7, nz + 7, b16), b16, 0, c.stream, c.cond, c.cond_r, c.cond_s, nx, nz,
                     c.T.dr, c.T.ds, D1[1], D1[2], D1[3], D1[4], D1[5]);
}

// tracers_to_corners: PS/composition.f90:923-1087.  filters = false is the
// reference's smooth_vis=0 (no smoother_time / smoother_time_T).  With
// discrete types (iC_continuous=0, tracers_to_corners_OG :1090-1276 with
// shape_function=1) Ct holds Ttype as 0.0 / 1.0: the type-1 weight sum over
// the weight sum is count_type_real(1)/nempty_real, the same weighted
// average, and nempty counts the same nearest-node tracers.
int tracers_to_corners(Context& c, const double* r, const double* s, const double* Tt, bool filters) {
  const int nx = c.T.nx, nz = c.T.nz, ncell = (nx + 1) * (nz + 1);
  const int64_t n = c.T.ntr;
  size_t b3 = sizeof(double) * (size_t)(nx + 7) * (nz + 7);
  if (c.timing) timer_begin(c, PT_K_PROJECT);
  PT_HIP(hipMemcpyAsync(c.Tratio, c.T0, b3, hipMemcpyDeviceToDevice, c.stream));
  PT_HIP(hipMemsetAsync(c.d_flags, 0, sizeof(int32_t) * 4, c.stream));
  unsigned nbt = (unsigned)((n + 255) / 256);
  // The windowed sort needs the storage order of an earlier projection and
  // no tracer more than one cell from its slot's cell.  Under-resolved flows
  // move tracers several cells per stage; after such a projection the next
  // WINDOW_BACKOFF projections take the radix sort without asking.
  constexpr int WINDOW_BACKOFF = 8;
  const bool try_window = c.order_valid && !c.project_radix_only && c.window_skip == 0;
  if (c.window_skip > 0) c.window_skip--;
  hipLaunchKernelGGL(k_bin, dim3(nbt), dim3(256), 0, c.stream, r, s, n, nx, nz, c.T.dr, c.T.ds, c.T.aspect,
                     c.cellkey, c.d_flags, try_window ? c.okey : nullptr);
  bool windowed = false;
  if (try_window) {
    // the domain check first, as the reference's binning loop does
    // (PS/composition.f90:1002-1011); then the sort the order allows
    PT_HIP(hipMemcpyAsync(c.h_fctl, c.d_flags, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
    PT_HIP(hipStreamSynchronize(c.stream));
    if (c.h_fctl[0]) {
      if (c.timing) timer_end(c, PT_K_PROJECT);
      set_error("tracer position outside [0,aspect]x[0,1]");
      return PT_ERR_TRACER_OUTSIDE;
    }
    windowed = c.h_fctl[1] == 0;
    if (!windowed) c.window_skip = WINDOW_BACKOFF;
  }
  if (windowed) {
    const dim3 gcs((unsigned)((nx + 1 + CS_CELLS - 1) / CS_CELLS), (unsigned)(nz + 1));
    const int nblk = (int)(gcs.x * gcs.y);
    hipLaunchKernelGGL(k_cs_count, gcs, dim3(CS_THREADS), 0, c.stream, c.cellkey, c.ostart, c.oend, nx, nz, c.ccount,
                       c.bsum);
    hipLaunchKernelGGL(k_cs_scan, dim3(1), dim3(1024), 0, c.stream, c.bsum, nblk, n, c.d_flags);
    hipLaunchKernelGGL(k_cs_fill, gcs, dim3(CS_THREADS), 0, c.stream, c.cellkey, c.ostart, c.oend, nx, nz, c.ccount,
                       c.bsum, c.cstart, c.cend, c.perm, c.skey);
    c.proj_window++;
  } else {
    PT_HIP(hipMemsetAsync(c.cstart, 0, sizeof(int32_t) * ncell, c.stream));
    PT_HIP(hipMemsetAsync(c.cend, 0, sizeof(int32_t) * ncell, c.stream));
    PT_TRY(sort_pairs(c.sort_tmp, c.sort_bytes, c.cellkey, c.skey, c.iota, c.perm, n, c.sort_bits, c.stream));
    hipLaunchKernelGGL(k_cell_bounds, dim3(nbt), dim3(256), 0, c.stream, c.skey, n, ncell, c.cstart, c.cend);
    c.proj_fallback++;
  }
  hipLaunchKernelGGL(k_cell_reduce, dim3((ncell + PJ_CELLS - 1) / PJ_CELLS), dim3(PJ_THREADS), 0, c.stream, r, s, Tt,
                     c.Ct, c.perm, c.skey, n, c.cstart, c.cend, c.cellsum, nx, nz, c.T.dr, c.T.ds);
  dim3 b16(64, 4);
  hipLaunchKernelGGL(k_node_combine, grid2(nx + 1, nz + 1, b16), b16, 0, c.stream, c.cellsum, c.cstart, c.cend,
                     c.Tratio, c.Cnew, c.nempty, nx, nz);
  if (c.timing) timer_end(c, PT_K_PROJECT);
  PT_HIP(hipMemcpyAsync(c.h_fctl, c.d_flags, 3 * sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
  PT_HIP(hipStreamSynchronize(c.stream));
  if (c.h_fctl[0]) {
    set_error("tracer position outside [0,aspect]x[0,1]");
    return PT_ERR_TRACER_OUTSIDE;
  }
  if (c.h_fctl[2]) {
    set_error("projection: the windowed counting sort did not place every tracer");
    return PT_ERR_INVALID;
  }
  c_bcs(c, c.Cnew);
  if (filters) PT_TRY(smoother_time(c));
  t_bcs(c, c.Tratio);
  if (filters) PT_TRY(smoother_time_T(c));
  load_conductivity(c);
  return PT_OK;
}

// ---------------------------------------------------------------------------
// source term, derivative tables, RK stages
// ---------------------------------------------------------------------------
__global__ void k_source(const double* Tx_, const double* cond, const double* cr, const double* cs, const double* Cvis,
                         double* src, int nx, int nz, double dr, double ds, double dr2, double ds2, double H,
                         double Htype, double d10, double d11, double d12, double d13, double d14, double d20,
                         double d21, double d22, double d23, double d24) {
  int i = -3 + (int)(blockIdx.x * blockDim.x + threadIdx.x), k = -3 + (int)(blockIdx.y * blockDim.y + threadIdx.y);
  if (i > nx + 3 || k > nz + 3) return;
  const double D1[5] = {d10, d11, d12, d13, d14}, D2[5] = {d20, d21, d22, d23, d24};
  const int p5 = nx + 11;
  double sx = 0, sz = 0, sxx = 0, szz = 0;
  for (int q = -2; q <= 2; q++) sx = sx + D1[q + 2] * Tx_[I5(i + q, k, p5)];
  for (int q = -2; q <= 2; q++) sz = sz + D1[q + 2] * Tx_[I5(i, k + q, p5)];
  for (int q = -2; q <= 2; q++) sxx = sxx + D2[q + 2] * Tx_[I5(i + q, k, p5)];
  for (int q = -2; q <= 2; q++) szz = szz + D2[q + 2] * Tx_[I5(i, k + q, p5)];
  double Tx = sx / dr, Tz = sz / ds, Txx = sxx / dr2, Tzz = szz / ds2;
  double kx = cr[I3(i, k, nx + 7)], kz = cs[I3(i, k, nx + 7)];
  double cv = Cvis[I5(i, k, p5)];
  double h = Htype * cv + H * (1.0 - (0.0 + cv));
  src[I3(i, k, nx + 7)] = cond[I5(i, k, p5)] * (Txx + Tzz) + (kx * Tx + kz * Tz) + h;
}

// interpolation table: tab[((i*(nz+7)) + (k+3))*12 + f*4 + d], f in {u,w,src},
// d = {value, DERr(1), DERr(2), DERr(3)}; i in 0..nx, k in -3..nz+3
__global__ void k_itab(const double* u, const double* w, const double* src, double* tab, int nx, int nz,
                       const double* Dm, double p1, double p2, double p3) {
  int k = -3 + (int)(blockIdx.x * blockDim.x + threadIdx.x), i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i > nx || k > nz + 3) return;
  const double* fs[3] = {u, w, src};
  const double pw[3] = {p1, p2, p3};
  long base = ((long)i * (nz + 7) + (k + 3)) * 12;
  for (int f = 0; f < 3; f++) {
    const double* F = fs[f];
    tab[base + f * 4] = F[I3(i, k, nx + 7)];
    for (int jj = 1; jj <= 3; jj++) {
      double sacc = 0.0;
      for (int a = -3; a <= 3; a++) sacc = sacc + Dm[(jj - 1) * 7 + a + 3] * F[I3(i + a, k, nx + 7)];
      tab[base + f * 4 + jj] = sacc / pw[jj - 1];
    }
  }
}

// hr**3 as the correctly rounded cube (the reference evaluates hr**real(3,8)
// with libm pow); FMA-based double-double product, rounded once
__device__ __forceinline__ double cube_cr(double x) {
  double p = x * x;
  double e = __fma_rn(x, x, -p);
  double t = p * x;
  double te = __fma_rn(p, x, -t);
  return t + (te + e * x);
}

struct InterpConst {
  int nx, nz;
  double dr, ds, f3;             // factorial(3)
};
struct InterpConstTab {
  double dsp1, dsp2, dsp3;       // ds_power(1..3)
  double Dm[3][7];               // D(1..3, -3..3)
};

// Per-node coefficient table of the interpolation polynomial.  interpolate
// (PS/composition.f90:1359-1393) evaluates, around the tracer's node (i,k),
//   fR(jj) = T0(jj) + a1 T1(jj) + a2 T2(jj) + a3 T3(jj),   jj = k-3..k+3,
//   out    = fR(k) + b1 d1 + b2 d2 + b3 d3,  d_q = sum_jj D(q,jj) fR(jj) / ds^q,
// with T_p = DERr(p) (T_0 = f), a_p = hr^p/p!, b_q = hs^q/q!.  That is the
// bicubic  out = sum_q b_q sum_p a_p C_pq  with  C_p0 = T_p(k),
// C_pq = sum_jj D(q,jj) T_p(jj) / ds^q.  The C_pq depend on the node only, so
// they are formed once per stage here (16 per field) and a tracer evaluates
// 4+1 short sums instead of 7 rows of Taylor sums and three derivative sums
// with their divisions.  Same polynomial; the rounding differs from the
// reference's order at the level of the last bit.
// ctab[node*48 + f*16 + q*4 + p], node = k*(nx+1) + i, f in {u, w, source}.
// One thread per coefficient, in ctab order, so a wave's stores are one
// contiguous run; the 7-row sums keep the order jj = 0..6.
__global__ void k_ctab(const double* __restrict__ tab, double* __restrict__ ctab, int nx, int nz, InterpConstTab C) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long nnode = (long)(nx + 1) * (nz + 1);
  if (t >= nnode * 48) return;
  const int node = (int)(t / 48), e = (int)(t - (long)node * 48);
  const int k = node / (nx + 1), i = node - k * (nx + 1);
  const int f = e >> 4, q = (e >> 2) & 3, p = e & 3;
  const double* col = tab + ((long)i * (nz + 7) + k) * 12 + f * 4 + p;   // row kk = k-3 at offset 0
  double v;
  if (q == 0) {
    v = col[3 * 12];
  } else {
    const double* D = C.Dm[q - 1];
    double acc = 0.0;
#pragma unroll
    for (int jj = 0; jj < 7; jj++) acc = acc + D[jj] * col[jj * 12];
    v = acc / (q == 1 ? C.dsp1 : (q == 2 ? C.dsp2 : C.dsp3));
  }
  ctab[t] = v;
}

// ---- interpolation at the tracers ------------------------------------------
// A workgroup holds RK_THREADS consecutive tracers of the cell-ordered SoA.
// Sorted by nearest node at the step's start, they reach a small box of
// nodes (at 1024x256 and 50 tracers per node: 6-7 nodes at stage 1, 14-24
// at the drifted stage positions).  The workgroup reduces the box, copies
// its coefficients from ctab into LDS with contiguous loads, and every lane
// reads its node's 48 coefficients (384 B) there, so the vector-memory path
// carries the tracer streams and one copy of the box per workgroup.  A box
// larger than RK_BOX nodes (a workgroup that wraps a grid row, or positions
// spread far) reads ctab directly.  Same coefficients in the same order:
// the results do not depend on the path.
constexpr int RK_THREADS = 256;
constexpr int RK_BOX = 42;     // staged nodes: 42 * 24 double2 fit 4 loads per thread
constexpr int RK_NODE = 25;    // double2 per staged node: 24 + 1 pad spreads nodes over LDS banks
static_assert(RK_BOX * 24 <= 4 * RK_THREADS, "box copy is 4 double2 per thread");

struct Located {
  int i, k;                      // nearest node
  double a1, a2, a3, b1, b2, b3; // hr^p/p!, hs^q/q!
};
__device__ __forceinline__ Located locate(const InterpConst& C, double r, double s) {
  Located p;
  int i = (int)round(r / C.dr), k = (int)round(s / C.ds);
  // positions reaching this point passed the projection's domain check;
  // the clamp only guards the table reads
  i = i < 0 ? 0 : (i > C.nx ? C.nx : i);
  k = k < 0 ? 0 : (k > C.nz ? C.nz : k);
  // factorial(1) = 1 and factorial(2) = 2 exactly, so h/factorial(1) = h and
  // h^2/factorial(2) = h^2 * 0.5 are the reference's quotients bit for bit
  const double hr = r - (double)i * C.dr;
  p.a1 = hr; p.a2 = (hr * hr) * 0.5; p.a3 = cube_cr(hr) / C.f3;
  const double hs = s - (double)k * C.ds;
  p.b1 = hs; p.b2 = (hs * hs) * 0.5; p.b3 = cube_cr(hs) / C.f3;
  p.i = i; p.k = k;
  return p;
}
// the bicubic from one node's 24 double2 of C_pq (LDS or global pointer)
template <typename P>
__device__ __forceinline__ void eval3(P cf, const Located& p, double out[3]) {
#pragma unroll
  for (int f = 0; f < 3; f++) {
    double g[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const double2 lo = cf[f * 8 + q * 2], hi = cf[f * 8 + q * 2 + 1];
      double v = lo.x;
      v = v + lo.y * p.a1;
      v = v + hi.x * p.a2;
      v = v + hi.y * p.a3;
      g[q] = v;
    }
    double fi = g[0];
    fi = fi + g[1] * p.b1;
    fi = fi + g[2] * p.b2;
    fi = fi + g[3] * p.b3;
    out[f] = fi;
  }
}
// workgroup barrier that orders LDS only: global loads issued before it stay
// in flight
__device__ __forceinline__ void lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}
__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int o = 32; o; o >>= 1) v = min(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int o = 32; o; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}
// interpolate for the three fields at (r, s).  Every thread of the workgroup
// calls it (threads past the last tracer pass that tracer's position), since
// it holds two workgroup barriers.
__device__ __forceinline__ void interp3(const double* __restrict__ ctab, const InterpConst& C, double r, double s,
                                        double out[3]) {
  __shared__ double2 box[RK_BOX * RK_NODE];
  __shared__ int bb[RK_THREADS / 64][4];
  const Located p = locate(C, r, s);
  const int i0w = wave_min(p.i), i1w = wave_max(p.i), k0w = wave_min(p.k), k1w = wave_max(p.k);
  if ((threadIdx.x & 63) == 0) {
    const int wv = threadIdx.x >> 6;
    bb[wv][0] = i0w; bb[wv][1] = i1w; bb[wv][2] = k0w; bb[wv][3] = k1w;
  }
  lds_sync();
  int i0 = bb[0][0], i1 = bb[0][1], k0 = bb[0][2], k1 = bb[0][3];
#pragma unroll
  for (int q = 1; q < RK_THREADS / 64; q++) {
    i0 = min(i0, bb[q][0]); i1 = max(i1, bb[q][1]); k0 = min(k0, bb[q][2]); k1 = max(k1, bb[q][3]);
  }
  // the box [i0,i1] x [k0,k1] holds every thread's node; `staged` is the
  // same in all threads
  const int w = i1 - i0 + 1, h = k1 - k0 + 1;
  const bool staged = w * h <= RK_BOX;
  const double2* g2 = reinterpret_cast<const double2*>(ctab);
  if (staged) {
    // box rows are contiguous runs of w nodes in ctab; loads are clamped to
    // the box so they can all be issued before the first store
    const int total = w * h * 24, per = w * 24;
    double2 v[4];
    int dst[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int t = threadIdx.x + u * RK_THREADS;
      const int tc = t < total ? t : total - 1;
      const int row = tc / per, rem = tc - row * per;
      const int nd = rem / 24, q = rem - nd * 24;
      v[u] = g2[((long)(k0 + row) * (C.nx + 1) + i0) * 24 + rem];
      dst[u] = t < total ? (row * w + nd) * RK_NODE + q : -1;
    }
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (dst[u] >= 0) box[dst[u]] = v[u];
  }
  lds_sync();
  if (staged)
    eval3(box + ((p.k - k0) * w + (p.i - i0)) * RK_NODE, p, out);
  else
    eval3(g2 + ((long)p.k * (C.nx + 1) + p.i) * 24, p, out);
}

struct RkArgs {
  double *r, *s, *T;                 // r, s, Ttr (accumulators after stage 1)
  double *r0, *s0, *T0;              // step-start copies
  double *r1, *s1, *T1, *r2, *s2, *T2;
  const double* tab;                 // ctab (per-node polynomial coefficients)
  int64_t n;
  double dtstage;                    // stage-specific step
  InterpConst C;
};

// The tracer streams are read and written once per stage kernel:
// nontemporal loads and stores keep them from displacing the coefficient
// table in the caches.  Each kernel issues all of its stream loads before
// the interpolation's barriers; threads past the last tracer load that
// tracer (j clamped) and store nothing.
#define NT_LD(x) __builtin_nontemporal_load(&(x))
#define NT_ST(x, v) __builtin_nontemporal_store((v), &(x))

// stage 1: PS/energy.f90:258-301
__global__ void __launch_bounds__(RK_THREADS) k_rk1(RkArgs a) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, jc = j < a.n ? j : a.n - 1;
  const double rv = NT_LD(a.r[jc]), sv = NT_LD(a.s[jc]), Tv = NT_LD(a.T[jc]);
  double o[3];
  interp3(a.tab, a.C, rv, sv, o);
  if (j >= a.n) return;
  NT_ST(a.r0[j], rv); NT_ST(a.s0[j], sv); NT_ST(a.T0[j], Tv);
  NT_ST(a.r[j], o[0]); NT_ST(a.s[j], o[1]); NT_ST(a.T[j], o[2]);
  NT_ST(a.r1[j], rv + a.dtstage * o[0]);
  NT_ST(a.s1[j], sv + a.dtstage * o[1]);
  NT_ST(a.T1[j], Tv + a.dtstage * o[2]);
}
// stage 2: PS/energy.f90:350-402
__global__ void __launch_bounds__(RK_THREADS) k_rk2(RkArgs a) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, jc = j < a.n ? j : a.n - 1;
  const double rv = NT_LD(a.r[jc]), sv = NT_LD(a.s[jc]), Tv = NT_LD(a.T[jc]);
  const double r0 = NT_LD(a.r0[jc]), s0 = NT_LD(a.s0[jc]), T0 = NT_LD(a.T0[jc]);
  double o[3];
  interp3(a.tab, a.C, NT_LD(a.r1[jc]), NT_LD(a.s1[jc]), o);
  if (j >= a.n) return;
  NT_ST(a.r[j], rv + 2.0 * o[0]);
  NT_ST(a.s[j], sv + 2.0 * o[1]);
  NT_ST(a.T[j], Tv + 2.0 * o[2]);
  NT_ST(a.r2[j], r0 + a.dtstage * o[0]);
  NT_ST(a.s2[j], s0 + a.dtstage * o[1]);
  NT_ST(a.T2[j], T0 + a.dtstage * o[2]);
}
// stage 3: PS/energy.f90:450-473
__global__ void __launch_bounds__(RK_THREADS) k_rk3(RkArgs a) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, jc = j < a.n ? j : a.n - 1;
  const double rv = NT_LD(a.r[jc]), sv = NT_LD(a.s[jc]), Tv = NT_LD(a.T[jc]);
  const double r0 = NT_LD(a.r0[jc]), s0 = NT_LD(a.s0[jc]), T0 = NT_LD(a.T0[jc]);
  double o[3];
  interp3(a.tab, a.C, NT_LD(a.r2[jc]), NT_LD(a.s2[jc]), o);
  if (j >= a.n) return;
  NT_ST(a.r[j], rv + 2.0 * o[0]);
  NT_ST(a.r1[j], r0 + a.dtstage * o[0]);
  NT_ST(a.s[j], sv + 2.0 * o[1]);
  NT_ST(a.s1[j], s0 + a.dtstage * o[1]);
  NT_ST(a.T[j], Tv + 2.0 * o[2]);
  NT_ST(a.T1[j], T0 + a.dtstage * o[2]);
}
// stage 4: PS/energy.f90:517-537 (dtstage = dt/6)
__global__ void __launch_bounds__(RK_THREADS) k_rk4(RkArgs a) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, jc = j < a.n ? j : a.n - 1;
  const double rv = NT_LD(a.r[jc]), sv = NT_LD(a.s[jc]), Tv = NT_LD(a.T[jc]);
  const double r0 = NT_LD(a.r0[jc]), s0 = NT_LD(a.s0[jc]), T0 = NT_LD(a.T0[jc]);
  double o[3];
  interp3(a.tab, a.C, NT_LD(a.r1[jc]), NT_LD(a.s1[jc]), o);
  if (j >= a.n) return;
  NT_ST(a.r[j], r0 + a.dtstage * (rv + o[0]));
  NT_ST(a.s[j], s0 + a.dtstage * (sv + o[1]));
  NT_ST(a.T[j], T0 + a.dtstage * (Tv + o[2]));
}

static int source_and_tables(Context& c, bool velocity_too) {
  int nx = c.T.nx, nz = c.T.nz;
  dim3 b16(64, 4);
  (void)velocity_too;
  extend_t(c);
  const double *D1 = c.T.D1, *D2 = c.T.D2;
  hipLaunchKernelGGL(k_source, grid2(nx + 7, nz + 7, b16), b16, 0, c.stream, c.Textend, c.cond, c.cond_r, c.cond_s,
                     c.Cvis, c.src, nx, nz, c.T.dr, c.T.ds, c.T.dr2, c.T.ds2, c.P.H, c.P.Htype, D1[1], D1[2], D1[3],
                     D1[4], D1[5], D2[1], D2[2], D2[3], D2[4], D2[5]);
  return PT_OK;
}

static int build_tables(Context& c, const double* u, const double* w, const double* src) {
  int nx = c.T.nx, nz = c.T.nz;
  dim3 b(64, 4);
  hipLaunchKernelGGL(k_itab, grid2(nz + 7, nx + 1, b), b, 0, c.stream, u, w, src, c.itab, nx, nz,
                     (const double*)c.dtab, c.T.dr_power[1], c.T.dr_power[2], c.T.dr_power[3]);
  InterpConstTab C;
  C.dsp1 = c.T.ds_power[1]; C.dsp2 = c.T.ds_power[2]; C.dsp3 = c.T.ds_power[3];
  for (int m = 0; m < 3; m++)
    for (int q = 0; q < 7; q++) C.Dm[m][q] = c.T.D[m + 1][q];
  const long ncoef = (long)(nx + 1) * (nz + 1) * 48;
  hipLaunchKernelGGL(k_ctab, dim3((unsigned)((ncoef + 255) / 256)), dim3(256), 0, c.stream, c.itab, c.ctab, nx, nz, C);
  return PT_OK;
}
static int build_itab(Context& c) { return build_tables(c, c.u, c.w, c.src); }

static RkArgs rk_args(Context& c) {
  RkArgs a;
  a.r = c.r; a.s = c.s; a.T = c.Tt; a.r0 = c.r0; a.s0 = c.s0; a.T0 = c.T0t;
  a.r1 = c.r1; a.s1 = c.s1; a.T1 = c.T1t; a.r2 = c.r2; a.s2 = c.s2; a.T2 = c.T2t;
  a.tab = c.ctab; a.n = c.T.ntr;
  a.C.nx = c.T.nx; a.C.nz = c.T.nz;
  a.C.dr = c.T.dr; a.C.ds = c.T.ds; a.C.f3 = c.T.factorial[3];
  return a;
}

// stability_RK4 with tracers (CFL=0): PS/energy.f90:756-810
__global__ void k_eig_max(const double* cond, const double* cr, const double* cs, int nx, int nz, double Tr_max,
                          double Ts_max, double Trr_max, double Tss_max, double* partial) {
  __shared__ double ra[256], rd[256];
  int tid = threadIdx.y * blockDim.x + threadIdx.x;
  int i = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y * blockDim.y + threadIdx.y;
  double eA = 0.0, eD = 0.0;
  if (i <= nx && k <= nz) {
    double cd = cond[I5(i, k, nx + 11)];
    eA = fabs(-(cd * 0.0 / 1.0) + cr[I3(i, k, nx + 7)] / 1.0) * Tr_max +
         fabs(-(cd * 0.0 / 1.0) + cs[I3(i, k, nx + 7)] / 1.0) * Ts_max;
    eD = cd * (fabs(Trr_max / 1.0) + fabs(Tss_max / 1.0));
  }
  ra[tid] = eA; rd[tid] = eD;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (tid < st) { ra[tid] = fmax(ra[tid], ra[tid + st]); rd[tid] = fmax(rd[tid], rd[tid + st]); }
    __syncthreads();
  }
  if (tid == 0) {
    int b = blockIdx.y * gridDim.x + blockIdx.x;
    partial[2 * b] = ra[0];
    partial[2 * b + 1] = rd[0];
  }
}
__global__ void k_eig_final(const double* partial, int nb, double* out) {
  if (threadIdx.x == 0) {
    double a = 0.0, d = 0.0;
    for (int b = 0; b < nb; b++) { a = fmax(a, partial[2 * b]); d = fmax(d, partial[2 * b + 1]); }
    out[0] = a; out[1] = d;
  }
}
static int stability_rk4(Context& c) {
  const Tables& t = c.T;
  double Ts_max = t.Tr_base / t.ds, Tss_max = -t.Trr_base / t.ds2, Tr_max = t.Tr_base / t.dr,
         Trr_max = -t.Trr_base / t.dr2;
  dim3 b16(64, 4);
  dim3 g = grid2(t.nx + 1, t.nz + 1, b16);
  hipLaunchKernelGGL(k_eig_max, g, b16, 0, c.stream, c.cond, c.cond_r, c.cond_s, t.nx, t.nz, Tr_max, Ts_max, Trr_max,
                     Tss_max, c.red);
  hipLaunchKernelGGL(k_eig_final, dim3(1), dim3(64), 0, c.stream, c.red, (int)(g.x * g.y), c.red_out);
  PT_HIP(hipMemcpyAsync(c.h_red, c.red_out, 2 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  PT_HIP(hipStreamSynchronize(c.stream));
  double eA = c.h_red[0], eD = c.h_red[1];
  double dtA = 2.8 / eA, dtD = 2.8 / eD, dt;
  if (eA == 0.0) dt = dtD;
  else dt = 1.0 / (1.0 / dtA + 1.0 / dtD);
  c.dt = c.P.courant * dt;
  return PT_OK;
}

static int update_velocities(Context& c) {
  PT_TRY(solve_stream(c));
  return compute_velocities(c);
}

// `which` names the stage's own timer (PT_K_TRACER_RK2 for stage 2, else -1);
// every stage also counts towards PT_K_TRACER_STAGE
static void rk_launch(Context& c, void (*kern)(RkArgs), RkArgs a, int which = -1) {
  unsigned nb = (unsigned)((a.n + RK_THREADS - 1) / RK_THREADS);
  if (c.timing) timer_begin(c, PT_K_TRACER_STAGE);
  if (c.timing && which >= 0) timer_begin(c, which);
  hipLaunchKernelGGL(kern, dim3(nb), dim3(RK_THREADS), 0, c.stream, a);
  if (c.timing && which >= 0) timer_end(c, which);
  if (c.timing) timer_end(c, PT_K_TRACER_STAGE);
}

// energy_time: PS/energy.f90:181-556 (tracer=1, comp=1, torder=4, RK_update=1)
static int reorder_tracers(Context& c);

int energy_time(Context& c) {
  PT_TRY(stability_rk4(c));
  if (c.P.equalize == 1) PT_TRY(equalize_tracers(c));   // PS/energy.f90:198-204
  const double dt = c.dt;
  const int nx = c.T.nx, nz = c.T.nz;
  size_t b3 = sizeof(double) * (size_t)(nx + 7) * (nz + 7), b5 = sizeof(double) * (size_t)(nx + 11) * (nz + 11);
  PT_HIP(hipMemcpyAsync(c.C0, c.Cnew, b5, hipMemcpyDeviceToDevice, c.stream));
  PT_HIP(hipMemcpyAsync(c.T0, c.Tratio, b3, hipMemcpyDeviceToDevice, c.stream));
  RkArgs a = rk_args(c);
  // stage 1
  PT_TRY(source_and_tables(c, true));
  PT_TRY(build_itab(c));
  a.dtstage = 0.5 * dt;
  rk_launch(c, k_rk1, a);
  // stage 2
  PT_TRY(tracers_to_corners(c, c.r1, c.s1, c.T1t));
  PT_TRY(source_and_tables(c, false));
  PT_TRY(update_velocities(c));
  PT_TRY(build_itab(c));
  a.dtstage = 0.5 * dt;
  rk_launch(c, k_rk2, a, PT_K_TRACER_RK2);
  // stage 3
  PT_HIP(hipMemcpyAsync(c.Cnew, c.C0, b5, hipMemcpyDeviceToDevice, c.stream));
  PT_TRY(tracers_to_corners(c, c.r2, c.s2, c.T2t));
  PT_TRY(source_and_tables(c, false));
  PT_TRY(update_velocities(c));
  PT_TRY(build_itab(c));
  a.dtstage = dt;
  rk_launch(c, k_rk3, a);
  // stage 4
  PT_HIP(hipMemcpyAsync(c.Cnew, c.C0, b5, hipMemcpyDeviceToDevice, c.stream));
  PT_TRY(tracers_to_corners(c, c.r1, c.s1, c.T1t));
  PT_TRY(source_and_tables(c, false));
  PT_TRY(update_velocities(c));
  PT_TRY(build_itab(c));
  a.dtstage = dt / 6.0;
  rk_launch(c, k_rk4, a);
  // wrap up
  PT_HIP(hipMemcpyAsync(c.Cnew, c.C0, b5, hipMemcpyDeviceToDevice, c.stream));
  PT_TRY(tracers_to_corners(c, c.r, c.s, c.Tt));
  PT_TRY(reorder_tracers(c));
  c.t_initial = false;   // T = Tratio (PS/energy.f90:552)
  return PT_OK;
}

// initial state: PS/basics.f90:352-359, PS/composition.f90:848, PS/main.f90:228-229
__global__ void k_load_T(const double* Tn, double* T, int nx, int nz) {
  int i = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y * blockDim.y + threadIdx.y;
  if (i > nx || k > nz) return;
  T[I3(i, k, nx + 7)] = Tn[(long)k * (nx + 1) + i];
}

// Storage order of the tracers: after the projection of the step's final
// positions the SoA arrays are permuted into that projection's cell order
// (stable: ascending slot within a cell), so the next step's interpolation
// gathers and projection reads walk the grid cell by cell.
__global__ void k_reorder(const int32_t* __restrict__ perm, int64_t n, const double* __restrict__ r,
                          const double* __restrict__ s, const double* __restrict__ T, const double* __restrict__ C,
                          const int32_t* __restrict__ id, double* __restrict__ r2, double* __restrict__ s2,
                          double* __restrict__ T2, double* __restrict__ C2, int32_t* __restrict__ id2) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int32_t q = perm[j];
  r2[j] = r[q];
  s2[j] = s[q];
  T2[j] = T[q];
  C2[j] = C[q];
  id2[j] = id[q];
}
static int reorder_tracers(Context& c) {
  const int64_t n = c.T.ntr;
  hipLaunchKernelGGL(k_reorder, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, c.perm, n, c.r, c.s,
                     c.Tt, c.Ct, c.tid, c.r1, c.s1, c.T1t, c.r2, c.tid_tmp);
  std::swap(c.r, c.r1);
  std::swap(c.s, c.s1);
  std::swap(c.Tt, c.T1t);
  std::swap(c.Ct, c.r2);
  std::swap(c.tid, c.tid_tmp);
  // the new storage order is the last projection's cell order
  std::swap(c.okey, c.skey);
  std::swap(c.ostart, c.cstart);
  std::swap(c.oend, c.cend);
  c.order_valid = true;
  return PT_OK;
}
__global__ void k_to_id_order(const int32_t* __restrict__ id, int64_t n, const double* __restrict__ a,
                              double* __restrict__ out) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) out[id[j]] = a[j];
}
// the RK4 registers r1, s1, T1t, r2 are free between steps and hold the result
int tracers_in_id_order(Context& c, double** r, double** s, double** Tt, double** Ct) {
  const int64_t n = c.T.ntr;
  const dim3 g((unsigned)((n + 255) / 256)), b(256);
  if (r) { hipLaunchKernelGGL(k_to_id_order, g, b, 0, c.stream, c.tid, n, c.r, c.r1); *r = c.r1; }
  if (s) { hipLaunchKernelGGL(k_to_id_order, g, b, 0, c.stream, c.tid, n, c.s, c.s1); *s = c.s1; }
  if (Tt) { hipLaunchKernelGGL(k_to_id_order, g, b, 0, c.stream, c.tid, n, c.Tt, c.T1t); *Tt = c.T1t; }
  if (Ct) { hipLaunchKernelGGL(k_to_id_order, g, b, 0, c.stream, c.tid, n, c.Ct, c.r2); *Ct = c.r2; }
  return PT_OK;
}

// ---- helpers of the tracer redistribution (pt_equalize.hip) ----------------
// The storage order as the cell order of the current positions (ostart /
// oend per cell, okey per slot).  Every projection of a step's final
// positions leaves it so; otherwise bin, sort and reorder here.
int ensure_cell_order(Context& c) {
  if (c.order_valid) return PT_OK;
  const int nx = c.T.nx, nz = c.T.nz, ncell = (nx + 1) * (nz + 1);
  const int64_t n = c.T.ntr;
  const unsigned nbt = (unsigned)((n + 255) / 256);
  PT_HIP(hipMemsetAsync(c.d_flags, 0, sizeof(int32_t) * 4, c.stream));
  hipLaunchKernelGGL(k_bin, dim3(nbt), dim3(256), 0, c.stream, c.r, c.s, n, nx, nz, c.T.dr, c.T.ds, c.T.aspect,
                     c.cellkey, c.d_flags, nullptr);
  PT_HIP(hipMemsetAsync(c.cstart, 0, sizeof(int32_t) * ncell, c.stream));
  PT_HIP(hipMemsetAsync(c.cend, 0, sizeof(int32_t) * ncell, c.stream));
  PT_TRY(sort_pairs(c.sort_tmp, c.sort_bytes, c.cellkey, c.skey, c.iota, c.perm, n, c.sort_bits, c.stream));
  hipLaunchKernelGGL(k_cell_bounds, dim3(nbt), dim3(256), 0, c.stream, c.skey, n, ncell, c.cstart, c.cend);
  PT_HIP(hipMemcpyAsync(c.h_fctl, c.d_flags, sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
  PT_HIP(hipStreamSynchronize(c.stream));
  if (c.h_fctl[0]) {
    set_error("tracer position outside [0,aspect]x[0,1]");
    return PT_ERR_TRACER_OUTSIDE;
  }
  return reorder_tracers(c);
}

// Textend(-3:nx+3,-3:nz+3) into a halo-3 field (the reference's Tdumb)
__global__ void k_window3(const double* T5, double* T3, int nx, int nz) {
  int i = -3 + (int)(blockIdx.x * blockDim.x + threadIdx.x), k = -3 + (int)(blockIdx.y * blockDim.y + threadIdx.y);
  if (i > nx + 3 || k > nz + 3) return;
  T3[I3(i, k, nx + 7)] = T5[I5(i, k, nx + 11)];
}
// T at the listed tracer slots from the coefficient table of one field
__global__ void k_interp_slots(const double* __restrict__ ctab, InterpConst C, const int32_t* __restrict__ slots,
                               int64_t n, const double* __restrict__ r, const double* __restrict__ s,
                               double* __restrict__ T) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int32_t q = slots[j];
  const Located p = locate(C, r[q], s[q]);
  double o[3];
  eval3(reinterpret_cast<const double2*>(ctab) + ((long)p.k * (C.nx + 1) + p.i) * 24, p, o);
  T[q] = o[0];
}
// Ttr of redistributed tracers: extendT(Tratio), compute_derivatives of
// Textend(-3:nx+3,-3:nz+3) and interpolate (PS/composition.f90:174-177,
// 245-251).  src, itab and ctab are scratch here: stage 1 of the RK4
// rebuilds them.
int interpolate_T_at(Context& c, const int32_t* slots, int64_t n) {
  const int nx = c.T.nx, nz = c.T.nz;
  extend_t(c);
  dim3 b16(64, 4);
  hipLaunchKernelGGL(k_window3, grid2(nx + 7, nz + 7, b16), b16, 0, c.stream, c.Textend, c.src, nx, nz);
  PT_TRY(build_tables(c, c.src, c.src, c.src));
  InterpConst C;
  C.nx = nx; C.nz = nz; C.dr = c.T.dr; C.ds = c.T.ds; C.f3 = c.T.factorial[3];
  if (n > 0)
    hipLaunchKernelGGL(k_interp_slots, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, c.ctab, C, slots, n,
                       c.r, c.s, c.Tt);
  return PT_OK;
}
// The projection that ends the redistribution: tracers_to_corners(rtr, str,
// Ttr, T) with smooth_vis=0 (PS/composition.f90:257-259); nodes without
// tracers keep the module T (the initial field before the first
// energy_time, Tratio after it).  The storage order then follows the new
// positions.
int equalize_projection(Context& c, bool moved) {
  const size_t b3 = sizeof(double) * (size_t)(c.T.nx + 7) * (c.T.nz + 7);
  if (!c.t_initial) PT_HIP(hipMemcpyAsync(c.T0, c.Tratio, b3, hipMemcpyDeviceToDevice, c.stream));
  if (moved) c.order_valid = false;
  PT_TRY(tracers_to_corners(c, c.r, c.s, c.Tt, false));
  return reorder_tracers(c);
}

// tracers (reference ID order) onto the device, projected onto the nodes, in cell order
static int place_tracers(Context& c, const double* rtr, const double* str, const double* Ttr, const double* Ctr) {
  size_t tb = sizeof(double) * (size_t)c.T.ntr;
  c.order_valid = false;
  c.window_skip = 0;
  PT_HIP(hipMemcpyAsync(c.r, rtr, tb, hipMemcpyHostToDevice, c.stream));
  PT_HIP(hipMemcpyAsync(c.s, str, tb, hipMemcpyHostToDevice, c.stream));
  PT_HIP(hipMemcpyAsync(c.Tt, Ttr, tb, hipMemcpyHostToDevice, c.stream));
  PT_HIP(hipMemcpyAsync(c.Ct, Ctr, tb, hipMemcpyHostToDevice, c.stream));
  PT_HIP(hipMemcpyAsync(c.tid, c.iota, sizeof(int32_t) * (size_t)c.T.ntr, hipMemcpyDeviceToDevice, c.stream));
  PT_TRY(tracers_to_corners(c, c.r, c.s, c.Tt));
  return reorder_tracers(c);
}

static int clear_coarse_errors(Context& c) {
  for (int l = 2; l <= c.T.grids; l++) {
    PT_HIP(hipMemsetAsync(c.L[l].err, 0, sizeof(double) * c.L[l].nelem, c.stream));
    if (c.L[l].ferr) PT_HIP(hipMemsetAsync(c.L[l].ferr, 0, sizeof(float) * c.L[l].nelem, c.stream));
  }
  return PT_OK;
}

// restart=1 start (PS/grid.f90:55-103, PS/main.f90:204-209, :226-229): T,
// Cnew and SF as print_restart wrote them (whole arrays with halos), the
// tracers in ID order.  T and Cnew are the values kept at nodes without
// tracers; the module T becomes the re-projected field.  The walls take
// Tbot/Ttop: the reference leaves Tbot_iso/Ttop_iso at 0 on this path.
int restart_state(Context& c, const double* T3, const double* C5, const double* SF3, double Tbot, double Ttop,
                  const double* rtr, const double* str, const double* Ttr, const double* Ctr) {
  const int nx = c.T.nx, nz = c.T.nz;
  c.Tbot_iso = Tbot; c.Ttop_iso = Ttop; c.time = 0.0; c.first_solve = true;
  c.t_initial = true; c.mass_set = false;
  const size_t b3 = sizeof(double) * (size_t)(nx + 7) * (nz + 7), b5 = sizeof(double) * (size_t)(nx + 11) * (nz + 11);
  PT_HIP(hipMemcpyAsync(c.T0, T3, b3, hipMemcpyHostToDevice, c.stream));
  hipLaunchKernelGGL(k_t_walls, dim3((nx + 7 + 255) / 256), dim3(256), 0, c.stream, c.T0, nx, nz, 3, nx + 7, 0, Tbot,
                     Ttop);
  t_bcs(c, c.T0);
  PT_HIP(hipMemcpyAsync(c.Cnew, C5, b5, hipMemcpyHostToDevice, c.stream));
  c_bcs(c, c.Cnew);
  PT_TRY(place_tracers(c, rtr, str, Ttr, Ctr));
  PT_HIP(hipMemcpyAsync(c.T0, c.Tratio, b3, hipMemcpyDeviceToDevice, c.stream));
  PT_HIP(hipMemcpyAsync(c.SF, SF3, b3, hipMemcpyHostToDevice, c.stream));
  if (c.P.iterate == 1) PT_TRY(stream_bcs(c, c.SF));   // PS/grid.f90:68
  PT_TRY(clear_coarse_errors(c));
  PT_HIP(hipStreamSynchronize(c.stream));
  return PT_OK;
}

int load_state(Context& c, const double* T_nodes, double Tbot, double Ttop, const double* rtr, const double* str,
               const double* Ttr, const double* Ctr) {
  const int nx = c.T.nx, nz = c.T.nz;
  c.Tbot_iso = Tbot; c.Ttop_iso = Ttop; c.time = 0.0; c.first_solve = true;
  c.t_initial = true; c.mass_set = false;
  size_t nb = sizeof(double) * (size_t)(nx + 1) * (nz + 1);
  double* dTn = c.src;   // staging (src is rebuilt every stage)
  PT_HIP(hipMemcpyAsync(dTn, T_nodes, nb, hipMemcpyHostToDevice, c.stream));
  PT_HIP(hipMemsetAsync(c.T0, 0, sizeof(double) * (size_t)(nx + 7) * (nz + 7), c.stream));
  dim3 b16(64, 4);
  hipLaunchKernelGGL(k_load_T, grid2(nx + 1, nz + 1, b16), b16, 0, c.stream, dTn, c.T0, nx, nz);
  hipLaunchKernelGGL(k_t_walls, dim3((nx + 7 + 255) / 256), dim3(256), 0, c.stream, c.T0, nx, nz, 3, nx + 7, 0, Tbot,
                     Ttop);
  t_bcs(c, c.T0);
  PT_HIP(hipMemsetAsync(c.Cnew, 0, sizeof(double) * (size_t)(nx + 11) * (nz + 11), c.stream));
  PT_TRY(place_tracers(c, rtr, str, Ttr, Ctr));
  PT_HIP(hipMemsetAsync(c.SF, 0, sizeof(double) * (size_t)(nx + 7) * (nz + 7), c.stream));
  PT_TRY(clear_coarse_errors(c));
  PT_HIP(hipStreamSynchronize(c.stream));
  return PT_OK;
}

// ---------------------------------------------------------------------------
static int dalloc(void** p, size_t bytes) {
  PT_HIP(hipMalloc(p, bytes));
  PT_HIP(hipMemset(*p, 0, bytes));
  return PT_OK;
}
#define DALLOC(ptr, bytes) PT_TRY(dalloc((void**)&(ptr), (bytes)))

__global__ void k_iota(int32_t* v, int64_t n) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) v[j] = (int32_t)j;
}

int tracers_alloc(Context& c) {
  const int nx = c.T.nx, nz = c.T.nz, ncell = (nx + 1) * (nz + 1);
  size_t b3 = sizeof(double) * (size_t)(nx + 7) * (nz + 7), b5 = sizeof(double) * (size_t)(nx + 11) * (nz + 11);
  double** f3[] = {&c.SF, &c.Tratio, &c.T0, &c.Tvis, &c.Tbuoy, &c.u, &c.w, &c.src, &c.cond_r, &c.cond_s};
  for (double** p : f3) DALLOC(*p, b3);
  double** f5[] = {&c.Cnew, &c.C0, &c.Cvis, &c.Cbuoy, &c.cond, &c.Textend, &c.TxS, &c.CxS};
  for (double** p : f5) DALLOC(*p, b5);
  DALLOC(c.itab, sizeof(double) * (size_t)(nx + 1) * (nz + 7) * 12);
  DALLOC(c.ctab, sizeof(double) * (size_t)(nx + 1) * (nz + 1) * 48);
  // D(1..3,-3..3) table for k_itab
  DALLOC(c.dtab, sizeof(double) * 21);
  double Dm[21];
  for (int m = 0; m < 3; m++)
    for (int q = 0; q < 7; q++) Dm[m * 7 + q] = c.T.D[m + 1][q];
  PT_HIP(hipMemcpy(c.dtab, Dm, sizeof Dm, hipMemcpyHostToDevice));
  size_t tb = sizeof(double) * (size_t)c.T.ntr;
  double** tr[] = {&c.r, &c.s, &c.Tt, &c.Ct, &c.r0, &c.s0, &c.T0t, &c.r1, &c.s1, &c.T1t, &c.r2, &c.s2, &c.T2t};
  for (double** p : tr) DALLOC(*p, tb);
  size_t ib = sizeof(int32_t) * (size_t)c.T.ntr;
  DALLOC(c.cellkey, ib); DALLOC(c.skey, ib); DALLOC(c.iota, ib); DALLOC(c.perm, ib);
  DALLOC(c.tid, ib); DALLOC(c.tid_tmp, ib);
  DALLOC(c.cstart, sizeof(int32_t) * ncell); DALLOC(c.cend, sizeof(int32_t) * ncell);
  DALLOC(c.okey, ib); DALLOC(c.ostart, sizeof(int32_t) * ncell); DALLOC(c.oend, sizeof(int32_t) * ncell);
  DALLOC(c.ccount, sizeof(int32_t) * ncell);
  DALLOC(c.bsum, sizeof(int32_t) * (size_t)((nx + 1 + CS_CELLS - 1) / CS_CELLS) * (nz + 1));
  DALLOC(c.nempty, sizeof(int32_t) * ncell);
  DALLOC(c.cellsum, sizeof(double) * 27 * (size_t)ncell);
  DALLOC(c.d_flags, sizeof(int32_t) * 4);
  DALLOC(c.d_fctl, sizeof(int32_t) * 4);
  DALLOC(c.d_fsm, sizeof(unsigned long long) * FSLOTS * SC_BLOCKS);
  PT_HIP(hipHostMalloc((void**)&c.h_fctl, sizeof(int32_t) * 4, hipHostMallocDefault));
  hipLaunchKernelGGL(k_iota, dim3((unsigned)((c.T.ntr + 255) / 256)), dim3(256), 0, 0, c.iota, c.T.ntr);
  c.sort_bits = 1;
  while ((1 << c.sort_bits) <= ncell) c.sort_bits++;
  PT_TRY(sort_pairs_temp_bytes(c.T.ntr, c.sort_bits, &c.sort_bytes));
  DALLOC(c.sort_tmp, c.sort_bytes + 256);
  PT_HIP(hipDeviceSynchronize());
  return PT_OK;
}

int tracers_free(Context& c) {
  void* ps[] = {c.SF, c.Tratio, c.T0, c.Tvis, c.Tbuoy, c.u, c.w, c.src, c.cond_r, c.cond_s, c.Cnew, c.C0, c.Cvis,
                c.Cbuoy, c.cond, c.Textend, c.TxS, c.CxS, c.itab, c.ctab, c.dtab, c.r, c.s, c.Tt, c.Ct, c.r0, c.s0, c.T0t,
                c.r1, c.s1, c.T1t, c.r2, c.s2, c.T2t, c.cellkey, c.skey, c.iota, c.perm, c.tid, c.tid_tmp, c.cstart, c.cend,
                c.nempty, c.cellsum, c.d_flags, c.sort_tmp, c.d_fctl, c.d_fsm, c.okey, c.ostart, c.oend, c.ccount,
                c.bsum};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  if (c.h_fctl) (void)hipHostFree(c.h_fctl);
  equalize_free(c);
  return PT_OK;
}

}  // namespace pt
