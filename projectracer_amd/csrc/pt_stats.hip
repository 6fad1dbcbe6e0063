// pt_stats.hip -- per-step diagnostics of `stats` (PS/basics.f90:363-462) on
// the device: the stats.dat row (comp=1, tracer=1 format, :421-422) and the
// viscosity_smoother.dat values (:453-455).
//
//   k_stats_rows   one workgroup per grid row k: the x-integrals (integrate1Dx,
//                  trapezoid rule, :494-516) of Cnew, u^2+w^2, T and RaT/conduct/eta,
//                  and the row's minima/maxima of eta and of the tracer counts
//   k_stats_edges  the wall rows: Nu, Nu_bot integrands (D1 z-derivative of T),
//                  corner fluxes q1..q4 and the maximum strain-rate invariant on
//                  the top wall (compute_strain_rate_invariant_direct,
//                  PS/momentum.f90:73-93)
//   k_stats_final  the z-integrals (integrate1Dz, integrate1Dz_ent, :518-541)
//                  and the row of scalars
// Every reduction has a fixed order, so the result does not depend on timing.
#include <cfloat>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include "pt_internal.hpp"

namespace pt {

#define S3(i, k, p) ((long)((k) + 3) * (p) + ((i) + 3))
#define S5(i, k, p) ((long)((k) + 5) * (p) + ((i) + 5))

constexpr int ST_THREADS = 256;

// per-row outputs: rows[k*NROW + q]
enum { R_C = 0, R_V, R_T, R_R, R_VMIN, R_VMAX, R_NEMPTY0, R_NMIN, R_NMAX, NROW };

__device__ __forceinline__ void block_sum(double* sh, double v) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = ST_THREADS / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + s];
    __syncthreads();
  }
}
__device__ __forceinline__ void block_minmax(double* sh, double v, bool is_max) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = ST_THREADS / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      double o = sh[threadIdx.x + s];
      sh[threadIdx.x] = is_max ? fmax(sh[threadIdx.x], o) : fmin(sh[threadIdx.x], o);
    }
    __syncthreads();
  }
}

// x-integral term i of integrate1Dx on a uniform grid: 0.5*(f(i)+f(i+1))*dr
__global__ void __launch_bounds__(ST_THREADS) k_stats_rows(const double* Cnew, const double* u, const double* w,
                                                           const double* T, const double* cond, const double* vis,
                                                           const int32_t* nempty, int nx, int nz, double dr,
                                                           double RaT, double* rows) {
  __shared__ double sh[ST_THREADS];
  const int k = blockIdx.x;
  const int p3 = nx + 7, p5 = nx + 11;
  double sc = 0.0, sv = 0.0, st = 0.0, sr = 0.0;
  for (int i = threadIdx.x; i < nx; i += ST_THREADS) {
    const double c0 = Cnew[S5(i, k, p5)], c1 = Cnew[S5(i + 1, k, p5)];
    sc = sc + 0.5 * (c0 * 1.0 + c1 * 1.0) * dr;
    const double u0 = u[S3(i, k, p3)], w0 = w[S3(i, k, p3)], u1 = u[S3(i + 1, k, p3)], w1 = w[S3(i + 1, k, p3)];
    const double v0 = u0 * u0 + w0 * w0, v1 = u1 * u1 + w1 * w1;
    sv = sv + 0.5 * (v0 * 1.0 + v1 * 1.0) * dr;
    st = st + 0.5 * (T[S3(i, k, p3)] * 1.0 + T[S3(i + 1, k, p3)] * 1.0) * dr;
    const double r0 = RaT / cond[S5(i, k, p5)] / vis[S3(i, k, p3)];
    const double r1 = RaT / cond[S5(i + 1, k, p5)] / vis[S3(i + 1, k, p3)];
    sr = sr + 0.5 * (r0 * 1.0 + r1 * 1.0) * dr;
  }
  double* out = rows + (long)k * NROW;
  block_sum(sh, sc);
  if (threadIdx.x == 0) out[R_C] = sh[0];
  __syncthreads();
  block_sum(sh, sv);
  if (threadIdx.x == 0) out[R_V] = sh[0];
  __syncthreads();
  block_sum(sh, st);
  if (threadIdx.x == 0) out[R_T] = sh[0];
  __syncthreads();
  block_sum(sh, sr);
  if (threadIdx.x == 0) out[R_R] = sh[0];
  __syncthreads();
  // minima / maxima over the row's nodes 0..nx
  double vmin = DBL_MAX, vmax = -DBL_MAX, e0 = 0.0, nmin = DBL_MAX, nmax = -DBL_MAX;
  for (int i = threadIdx.x; i <= nx; i += ST_THREADS) {
    const double v = vis[S3(i, k, p3)];
    vmin = fmin(vmin, v);
    vmax = fmax(vmax, v);
    const double ne = (double)nempty[(long)k * (nx + 1) + i];
    e0 = e0 + (ne == 0.0 ? 1.0 : 0.0);
    nmin = fmin(nmin, ne);
    nmax = fmax(nmax, ne);
  }
  block_minmax(sh, vmin, false);
  if (threadIdx.x == 0) out[R_VMIN] = sh[0];
  __syncthreads();
  block_minmax(sh, vmax, true);
  if (threadIdx.x == 0) out[R_VMAX] = sh[0];
  __syncthreads();
  block_sum(sh, e0);
  if (threadIdx.x == 0) out[R_NEMPTY0] = sh[0];
  __syncthreads();
  block_minmax(sh, nmin, false);
  if (threadIdx.x == 0) out[R_NMIN] = sh[0];
  __syncthreads();
  block_minmax(sh, nmax, true);
  if (threadIdx.x == 0) out[R_NMAX] = sh[0];
}

struct EdgeArgs {
  double D1[5], D2[5], Drs[5][5];   // offsets -2..2; Drs[b+2][a+2] = D1(a) D1(b)
  double dr, ds, dr2, ds2, drds;
  const double *u, *w;              // velocities (halo 3)
  int from_velocity;                // 1: the direct path's strain from u, w
};

// compute_strain_rate_invariant_direct (PS/momentum.f90:73-93, uniform grid)
// at node (i, k) from psi
__device__ __forceinline__ double strain_at(const double* SF, int i, int k, int p3, const EdgeArgs& a) {
  double sxx = 0.0, szz = 0.0, sxz = 0.0;
  for (int j = -2; j <= 2; j++) sxx = sxx + a.D2[j + 2] * SF[S3(i + j, k, p3)];
  for (int j = -2; j <= 2; j++) szz = szz + a.D2[j + 2] * SF[S3(i, k + j, p3)];
  for (int b = -2; b <= 2; b++)
    for (int q = -2; q <= 2; q++) sxz = sxz + a.Drs[b + 2][q + 2] * SF[S3(i + q, k + b, p3)];
  const double Sxx = sxx / a.dr2, Szz = szz / a.ds2, Sxz = sxz / a.drds;
  const double d = Szz - Sxx;
  return sqrt(2.0 * (Sxz * Sxz) + 0.5 * (d * d));
}

// compute_strain_rate_invariant (PS/momentum.f90:1-23, uniform grid), the
// direct path's strain: from the velocities' first derivatives
__device__ __forceinline__ double strain_vel_at(int i, int k, int p3, const EdgeArgs& a) {
  double ur = 0.0, us = 0.0, wr = 0.0, ws = 0.0;
  for (int j = -2; j <= 2; j++) ur = ur + a.D1[j + 2] * a.u[S3(i + j, k, p3)];
  for (int j = -2; j <= 2; j++) us = us + a.D1[j + 2] * a.u[S3(i, k + j, p3)];
  for (int j = -2; j <= 2; j++) wr = wr + a.D1[j + 2] * a.w[S3(i + j, k, p3)];
  for (int j = -2; j <= 2; j++) ws = ws + a.D1[j + 2] * a.w[S3(i, k + j, p3)];
  const double ux = (ur / a.dr) / 1.0, uz = (us / a.ds) / 1.0, wx = (wr / a.dr) / 1.0, wz = (ws / a.ds) / 1.0;
  const double e = uz + wx;
  return sqrt(0.5 * (2.0 * (ux * ux + wz * wz) + e * e));
}
__device__ __forceinline__ double strain_node(const double* SF, int i, int k, int p3, const EdgeArgs& a) {
  return a.from_velocity ? strain_vel_at(i, k, p3, a) : strain_at(SF, i, k, p3, a);
}

// D1 . T(i, k-2..k+2) / ds / zs(k), zs = 1 on a uniform grid
__device__ __forceinline__ double ddz(const double* T, int i, int k, int p3, const EdgeArgs& a) {
  double s = 0.0;
  for (int j = -2; j <= 2; j++) s = s + a.D1[j + 2] * T[S3(i, k + j, p3)];
  return s / a.ds / 1.0;
}

// out: [0] x-integral of -conduct*Tz on the top row, [1] same on the bottom
// row, [2] x-integral of T on the bottom row, [3] max strain on the top row,
// [4..7] q1..q4
__global__ void __launch_bounds__(ST_THREADS) k_stats_edges(const double* T, const double* cond, const double* SF,
                                                            int nx, int nz, EdgeArgs a, double* out) {
  __shared__ double sh[ST_THREADS];
  const int p3 = nx + 7, p5 = nx + 11;
  double top = 0.0, bot = 0.0, tb = 0.0, smax = -DBL_MAX;
  for (int i = threadIdx.x; i < nx; i += ST_THREADS) {
    const double ft0 = -cond[S5(i, nz, p5)] * ddz(T, i, nz, p3, a);
    const double ft1 = -cond[S5(i + 1, nz, p5)] * ddz(T, i + 1, nz, p3, a);
    top = top + 0.5 * (ft0 * 1.0 + ft1 * 1.0) * a.dr;
    const double fb0 = -cond[S5(i, 0, p5)] * ddz(T, i, 0, p3, a);
    const double fb1 = -cond[S5(i + 1, 0, p5)] * ddz(T, i + 1, 0, p3, a);
    bot = bot + 0.5 * (fb0 * 1.0 + fb1 * 1.0) * a.dr;
    tb = tb + 0.5 * (T[S3(i, 0, p3)] * 1.0 + T[S3(i + 1, 0, p3)] * 1.0) * a.dr;
  }
  for (int i = threadIdx.x; i <= nx; i += ST_THREADS) smax = fmax(smax, strain_node(SF, i, nz, p3, a));
  block_sum(sh, top);
  if (threadIdx.x == 0) out[0] = sh[0];
  __syncthreads();
  block_sum(sh, bot);
  if (threadIdx.x == 0) out[1] = sh[0];
  __syncthreads();
  block_sum(sh, tb);
  if (threadIdx.x == 0) out[2] = sh[0];
  __syncthreads();
  block_minmax(sh, smax, true);
  if (threadIdx.x == 0) {
    out[3] = sh[0];
    out[4] = -cond[S5(0, nz, p5)] * ddz(T, 0, nz, p3, a);
    out[5] = -cond[S5(nx, nz, p5)] * ddz(T, nx, nz, p3, a);
    out[6] = -cond[S5(nx, 0, p5)] * ddz(T, nx, 0, p3, a);
    out[7] = -cond[S5(0, 0, p5)] * ddz(T, 0, 0, p3, a);
  }
}

// z-integrals over the rows (sequential, as integrate1Dz) and the extrema
// out: [0] int C, [1] int C from kent, [2] int |v|^2, [3] int T, [4] int RaT/k/eta,
//      [5] top-row int |v|^2, [6] vis min, [7] vis max, [8] empty, [9] min, [10] max
__global__ void k_stats_final(const double* rows, int nz, int kent, double ds, double* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double c = 0.0, ce = 0.0, v = 0.0, t = 0.0, r = 0.0;
  for (int k = 0; k <= nz - 1; k++) {
    const double* a = rows + (long)k * NROW;
    const double* b = rows + (long)(k + 1) * NROW;
    c = c + 0.5 * (a[R_C] * 1.0 + b[R_C] * 1.0) * ds;
    if (k >= kent) ce = ce + 0.5 * (a[R_C] * 1.0 + b[R_C] * 1.0) * ds;
    v = v + 0.5 * (a[R_V] * 1.0 + b[R_V] * 1.0) * ds;
    t = t + 0.5 * (a[R_T] * 1.0 + b[R_T] * 1.0) * ds;
    r = r + 0.5 * (a[R_R] * 1.0 + b[R_R] * 1.0) * ds;
  }
  double vmin = DBL_MAX, vmax = -DBL_MAX, e0 = 0.0, nmin = DBL_MAX, nmax = -DBL_MAX;
  for (int k = 0; k <= nz; k++) {
    const double* a = rows + (long)k * NROW;
    vmin = fmin(vmin, a[R_VMIN]);
    vmax = fmax(vmax, a[R_VMAX]);
    e0 = e0 + a[R_NEMPTY0];
    nmin = fmin(nmin, a[R_NMIN]);
    nmax = fmax(nmax, a[R_NMAX]);
  }
  out[0] = c; out[1] = ce; out[2] = v; out[3] = t; out[4] = r;
  out[5] = rows[(long)nz * NROW + R_V];
  out[6] = vmin; out[7] = vmax; out[8] = e0; out[9] = nmin; out[10] = nmax;
}

// strain(0:nx, 0:nz) of the current psi, row-major (k, i)
__global__ void k_strain_field(const double* SF, int nx, int nz, EdgeArgs a, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y * blockDim.y + threadIdx.y;
  if (i > nx || k > nz) return;
  out[(long)k * (nx + 1) + i] = strain_node(SF, i, k, nx + 7, a);
}

static EdgeArgs edge_args(Context& c) {
  EdgeArgs a;
  for (int q = 0; q < 5; q++) { a.D1[q] = c.st.D1[q]; a.D2[q] = c.st.D2[q]; }
  for (int b = 0; b < 5; b++)
    for (int q = 0; q < 5; q++) a.Drs[b][q] = c.T.Drs[b + 1][q + 1];
  a.dr = c.T.dr; a.ds = c.T.ds; a.dr2 = c.T.dr2; a.ds2 = c.T.ds2; a.drds = c.L[1].g.drds;
  a.u = c.u; a.w = c.w;
  a.from_velocity = c.P.iterate == 0;   // stats: PS/basics.f90:371-376
  return a;
}

int strain_field(Context& c, double* out) {
  dim3 b(64, 4), g((c.T.nx + 1 + 63) / 64, (c.T.nz + 1 + 3) / 4);
  hipLaunchKernelGGL(k_strain_field, g, b, 0, c.stream, c.SF, c.T.nx, c.T.nz, edge_args(c), out);
  return PT_OK;
}

// Scratch of the no-allocation variant (pt_stats_noalloc): the per-row
// partial integrals, the wall values and the final scalars of one call, for
// grids of up to STATS_NOALLOC_MAX_NZ rows, in a file-scope device array.
// Each device holds its own copy; its address there is resolved once per
// device at pt_create (stats_scratch_setup) and only that address reaches
// kernels.  Every call on a device shares the array: calls that overlap on
// one device race.  Every element a call reads is written by the same call
// first, so nothing needs clearing between calls.
constexpr int STATS_NOALLOC_MAX_NZ = 2047;
constexpr size_t STATS_SCRATCH_DOUBLES = (size_t)NROW * (STATS_NOALLOC_MAX_NZ + 1) + 32;
static __device__ __align__(256) double scratch[STATS_SCRATCH_DOUBLES];

static std::mutex g_scratch_mu;
static std::map<int, double*> g_scratch_ptr;   // device -> address of `scratch` there

int stats_scratch_setup(Context& c) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto it = g_scratch_ptr.find(c.device);
  if (it == g_scratch_ptr.end()) {
    void* ptr = nullptr;
    PT_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    it = g_scratch_ptr.emplace(c.device, static_cast<double*>(ptr)).first;
  }
  c.stats_scratch = it->second;
  return PT_OK;
}

static int stats_with(Context& c, pt_stats_row* row, double* rows);

int stats(Context& c, pt_stats_row* row) {
  if (!c.stats_rows) {
    PT_HIP(hipMalloc((void**)&c.stats_rows, sizeof(double) * (size_t)NROW * (c.T.nz + 1) + sizeof(double) * 32));
  }
  return stats_with(c, row, c.stats_rows);
}

int stats_noalloc(Context& c, pt_stats_row* row) {
  if (c.T.nz > STATS_NOALLOC_MAX_NZ) {
    set_error("pt_stats_noalloc: nz = " + std::to_string(c.T.nz) + " exceeds the scratch limit of " +
              std::to_string(STATS_NOALLOC_MAX_NZ) + " rows; use pt_stats");
    return PT_ERR_UNSUPPORTED;
  }
  if (!c.stats_scratch) {
    set_error("pt_stats_noalloc: scratch address not resolved at pt_create");
    return PT_ERR_HIP;
  }
  return stats_with(c, row, c.stats_scratch);
}

static int stats_with(Context& c, pt_stats_row* row, double* rows) {
  const int nx = c.T.nx, nz = c.T.nz;
  double* edge = rows + (size_t)NROW * (nz + 1);
  double* fin = edge + 8;
  const double* T = c.t_initial ? c.T0 : c.Tratio;   // the reference's module T
  hipLaunchKernelGGL(k_stats_rows, dim3(nz + 1), dim3(ST_THREADS), 0, c.stream, c.Cnew, c.u, c.w, T, c.cond,
                     c.L[1].v.vis, c.nempty, nx, nz, c.T.dr, c.P.RaT, rows);
  const EdgeArgs a = edge_args(c);
  hipLaunchKernelGGL(k_stats_edges, dim3(1), dim3(ST_THREADS), 0, c.stream, T, c.cond, c.SF, nx, nz, a, edge);
  // kent: last row at or below the entrainment height (entrainment_setup,
  // PS/composition.f90:765-774, uniform grid)
  int kent = 0;
  for (int k = 0; k <= nz; k++)
    if (c.T.ds * (double)k <= c.P.dent) kent = k;
  hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(64), 0, c.stream, rows, nz, kent, c.T.ds, fin);
  double h[19];
  PT_HIP(hipMemcpyAsync(h, edge, sizeof h, hipMemcpyDeviceToHost, c.stream));
  PT_HIP(hipStreamSynchronize(c.stream));
  const double* e = h;       // edge values
  const double* f = h + 8;   // final values
  const double aspect = c.T.aspect;
  const double Cavg = f[0] / aspect;
  if (!c.mass_set) {
    c.mass = Cavg;   // the reference's tstep=0 call (PS/basics.f90:381-389)
    c.mass_set = true;
  }
  std::memset(row, 0, sizeof *row);
  row->time = c.time;
  row->Tavg = f[3] / aspect;
  row->vrms = std::sqrt(f[2] / aspect);
  row->mobility = std::sqrt(f[5] / aspect) / row->vrms;
  row->RaT_avg = f[4] / aspect;
  row->Nu = e[0] / e[2];
  row->Nu_bot = e[1] / aspect;
  row->strain_max = e[3];
  row->q1 = e[4]; row->q2 = e[5]; row->q3 = e[6]; row->q4 = e[7];
  row->empty_cells = (int32_t)f[8];
  row->tracer_min = (int32_t)f[9];
  row->tracer_max = (int32_t)f[10];
  row->Cavg = Cavg;
  row->ent = f[1] / aspect / c.mass;
  row->vis_min = f[6];
  row->vis_max = f[7];
  row->n_eta = c.n_eta; row->n_temp = c.n_temp; row->n_comp = c.n_comp; row->eig_fine = c.eig_fine;
  return PT_OK;
}

}  // namespace pt
