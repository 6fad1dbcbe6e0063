// pt_setup.cpp -- init-time tables of the hot path (host, once per context).
//
// Each block restates the reference routine cited next to it, keeping the
// reference's operation order so the FP64 device path sees the same
// constants (stencil weights, RKC1 coefficients, eigenvalue bounds, local
// pseudo-time steps) as the Fortran program.
#include <cmath>
#include <cstring>
#include <cfloat>
#include <vector>
#include "pt_internal.hpp"

namespace pt {

static constexpr double PII = 3.1415926535897932;   // PS/basics.f90:32

// optimize_grid: PS/iterative_solver.f90:1-61
static void optimize_grid(int& nx, int& nz, int& grids, int n_points_min) {
  int n_points_max = std::min(nx, nz);
  int n_coarse_max = grids - 1;
  int n_max = n_points_max * (1 << n_coarse_max);
  int nxbig = (nz <= nx) ? 1 : 0;
  int n_in = nxbig ? nz : nx;
  int n_temp = n_max, n_coarse_temp = 0;
  for (int np = n_points_min; np <= n_points_max; np++)
    for (int nc = 1; nc <= n_coarse_max; nc++) {
      int n = np * (1 << nc);
      if (n >= n_in && n < n_temp) { n_temp = n; n_coarse_temp = nc; }
    }
  grids = n_coarse_temp + 1;
  if (nxbig == 0) nx = n_temp; else nz = n_temp;
  n_in = (nz <= nx) ? nx : nz;
  int npt = (int)std::lround((double)n_in / (double)(1 << n_coarse_temp));
  n_temp = npt * (1 << n_coarse_temp);
  if (nxbig == 0) nz = n_temp; else nx = n_temp;
}

// W-cycle schedules: PS/multigrid_cycle.f90:9-97
static int cycle_schedule(int grids, int* multi) {
  static const int w2[] = {1, 2};
  static const int w3[] = {1, 2, 3, 2, 3, 2};
  static const int w4[] = {1, 2, 3, 4, 3, 4, 3, 2, 3, 4, 3, 4, 3, 2};
  static const int w5[] = {1, 2, 3, 4, 5, 4, 5, 4, 3, 4, 5, 4, 3, 2, 3, 4, 5, 4, 3, 4, 5, 4, 5, 4, 3, 2};
  static const int w6[] = {1, 2, 3, 4, 5, 6, 5, 6, 5, 4, 5, 6, 5, 4, 3, 4, 5, 6, 5, 4, 3,
                           2, 3, 4, 5, 6, 5, 4, 3, 4, 5, 6, 5, 4, 5, 6, 5, 6, 5, 4, 3, 2};
  const int* m; int n;
  switch (grids) {
    case 2: m = w2; n = 2; break;
    case 3: m = w3; n = 6; break;
    case 4: m = w4; n = 14; break;
    case 5: m = w5; n = 26; break;
    case 6: m = w6; n = 42; break;
    default: return -1;
  }
  for (int p = 1; p <= n; p++) multi[p] = m[p - 1];
  multi[0] = multi[n];
  multi[n + 1] = multi[1];
  return n;
}

// The cycle the mixed path may follow instead: 'V' the standard V-cycle, 'F'
// the standard F-cycle (both listed, commented out, beside the W-cycle at
// PS/multigrid_cycle.f90:43-96), 'W' the W-cycle.  Where a kind has no
// distinct form for `grids` levels it is the W-cycle.
int cycle_schedule_kind(int grids, char kind, int* multi) {
  static const int f4[] = {1, 2, 3, 4, 3, 4, 3, 2, 3, 4, 3, 2};
  static const int f5[] = {1, 2, 3, 4, 5, 4, 5, 4, 3, 4, 5, 4, 3, 2, 3, 4, 5, 4, 3, 2};
  static const int f6[] = {1, 2, 3, 4, 5, 6, 5, 6, 5, 4, 5, 6, 5, 4, 3, 4, 5, 6, 5, 4, 3, 2, 3, 4, 5, 6, 5, 4, 3, 2};
  int n = 0;
  if (kind == 'V' && grids >= 3) {
    for (int l = 1; l <= grids; l++) multi[++n] = l;
    for (int l = grids - 1; l >= 2; l--) multi[++n] = l;
  } else if (kind == 'F' && grids >= 4) {
    const int* m = grids == 4 ? f4 : grids == 5 ? f5 : f6;
    n = grids == 4 ? 12 : grids == 5 ? 20 : 30;
    for (int p = 1; p <= n; p++) multi[p] = m[p - 1];
  } else {
    return cycle_schedule(grids, multi);
  }
  multi[0] = multi[n];
  multi[n + 1] = multi[1];
  return n;
}

// FD_coefficients: PS/main.f90:499-618 (Fornberg's recursion for nodes
// 0,1,-1,2,-2,3,-3 and derivative orders up to 4)
static void fd_tables(Tables& t) {
  const int order = 4, bigM = 4, bigN = 6;
  double a[8];
  double dl[6][7][7];   // dl[m+1][n][i]
  std::memset(dl, 0, sizeof dl);
  a[0] = 0.0;
  for (int i = 1, k = 1; i <= bigN; i += 2, k++) { a[i] = (double)k; a[i + 1] = -(double)k; }
  auto D = [&](int m, int n, int i) -> double& { return dl[m + 1][n][i]; };
  D(0, 0, 0) = 1.0;
  double c1 = 1.0;
  for (int n = 1; n <= bigN; n++) {
    double c2 = 1.0;
    int mmax = n < bigM ? n : bigM;
    for (int i = 0; i < n; i++) {
      double c3 = a[n] - a[i];
      c2 = c2 * c3;
      if (n <= bigM) D(n, n - 1, i) = 0.0;
      for (int m = 0; m <= mmax; m++) D(m, n, i) = ((a[n] - 0.0) * D(m, n - 1, i) - (double)m * D(m - 1, n - 1, i)) / c3;
    }
    for (int m = 0; m <= mmax; m++)
      D(m, n, n) = (c1 / c2) * ((double)m * D(m - 1, n - 1, n - 1) - (a[n - 1] - 0.0) * D(m, n - 1, n - 1));
    c1 = c2;
  }
  std::memset(t.D1, 0, sizeof t.D1); std::memset(t.D2, 0, sizeof t.D2);
  std::memset(t.D3, 0, sizeof t.D3); std::memset(t.D4, 0, sizeof t.D4); std::memset(t.D, 0, sizeof t.D);
  t.D1[3] = D(1, order, 0); t.D2[3] = D(2, order, 0);
  for (int i = 1, k = 1; i <= order; i += 2, k++) {
    t.D1[3 + k] = D(1, order, i); t.D1[3 - k] = D(1, order, i + 1);
    t.D2[3 + k] = D(2, order, i); t.D2[3 - k] = D(2, order, i + 1);
  }
  for (int m = 2; m <= 4; m += 2) {
    t.D[m - 1][3] = D(m - 1, order + m - 2, 0); t.D[m][3] = D(m, order + m - 2, 0);
    for (int i = 1, k = 1; i <= order + m - 2; i += 2, k++) {
      t.D[m - 1][3 + k] = D(m - 1, order + m - 2, i); t.D[m - 1][3 - k] = D(m - 1, order + m - 2, i + 1);
      t.D[m][3 + k] = D(m, order + m - 2, i); t.D[m][3 - k] = D(m, order + m - 2, i + 1);
    }
  }
  for (int m = 1; m <= 4; m += 2) {
    t.D[m][3] = 0.0;
    for (int i = 1; i <= 3; i++) { t.D[m][3 + i] = -t.D[m][3 - i]; t.D[m + 1][3 + i] = t.D[m + 1][3 - i]; }
  }
  t.D3[3] = D(3, order + 2, 0); t.D4[3] = D(4, order + 2, 0);
  for (int i = 1, k = 1; i <= order + 2; i += 2, k++) {
    t.D3[3 + k] = D(3, order + 2, i); t.D3[3 - k] = D(3, order + 2, i + 1);
    t.D4[3 + k] = D(4, order + 2, i); t.D4[3 - k] = D(4, order + 2, i + 1);
  }
  t.D1[3] = 0.0; t.D3[3] = 0.0;
  for (int i = 1; i <= 3; i++) {
    t.D1[3 + i] = -t.D1[3 - i]; t.D2[3 + i] = t.D2[3 - i];
    t.D3[3 + i] = -t.D3[3 - i]; t.D4[3 + i] = t.D4[3 - i];
  }
  std::memset(t.Drs, 0, sizeof t.Drs); std::memset(t.Drss, 0, sizeof t.Drss);
  std::memset(t.Drrs, 0, sizeof t.Drrs); std::memset(t.Drrss, 0, sizeof t.Drrss);
  for (int k = -2; k <= 2; k++)
    for (int i = -2; i <= 2; i++) {
      t.Drs[k + 3][i + 3] = t.D1[k + 3] * t.D1[i + 3];
      t.Drss[k + 3][i + 3] = t.D2[k + 3] * t.D1[i + 3];
      t.Drrs[k + 3][i + 3] = t.D1[k + 3] * t.D2[i + 3];
      t.Drrss[k + 3][i + 3] = t.D2[k + 3] * t.D2[i + 3];
    }
}

// Chebyshev_prep_stream: PS/Chebyshev.f90:20-34, 82-110, into the given arrays
static void rkc1_coefficients(int s, double damping, double* mewt, double* mew, double* vv, double* cc,
                              double* stability) {
  double T[64], Tx[64], b[64];
  double w0 = 1.0 + damping / (double)(s * s);
  T[0] = 1.0; T[1] = w0; Tx[0] = 0.0; Tx[1] = 1.0;
  for (int j = 2; j <= s; j++) {
    T[j] = 2.0 * w0 * T[j - 1] - T[j - 2];
    Tx[j] = 2.0 * T[j - 1] + 2.0 * w0 * Tx[j - 1] - Tx[j - 2];
  }
  double w1 = T[s] / Tx[s];
  for (int j = 0; j <= s; j++) b[j] = 1.0 / T[j];
  std::memset(mewt, 0, 64 * sizeof(double)); std::memset(mew, 0, 64 * sizeof(double));
  std::memset(vv, 0, 64 * sizeof(double)); std::memset(cc, 0, 64 * sizeof(double));
  mewt[1] = b[1] * w1;
  for (int j = 2; j <= s; j++) {
    mew[j] = 2.0 * b[j] * w0 / b[j - 1];
    vv[j] = -b[j] / b[j - 2];
    cc[j] = 1.0 - mew[j] - vv[j];
  }
  for (int j = 2; j <= s; j++) mewt[j] = 2.0 * b[j] * w1 / b[j - 1];
  *stability = -(2.0 * w0 / w1);
}
static void rkc1_tables(Tables& t, int s, double damping) {
  rkc1_coefficients(s, damping, t.mewt, t.mew, t.vv, t.cc, &t.RKC1_stability);
}
// the mixed path's coarse-level smoother: the same recurrence with s stages
void rkc1_coarse_tables(Tables& t, int s, double damping) {
  rkc1_coefficients(s, damping, t.mewt_c, t.mew_c, t.vv_c, t.cc_c, &t.RKC1_stability_c);
  t.s_coarse = s;
}

// derivative_eigenvalues: PS/energy.f90:632-754.  The Fourier samples are
// formed as the reference's default-kind (single precision) complex
// products cmplx(cos,sin)*cmplx(cos,sin); the per-offset samples are
// tabulated once so the 2049x2049 sweep is cheap.
static void eigen_bases(Tables& t) {
  const int na = 2049;
  const double da = 2.0 * PII / (double)(na - 1);
  std::vector<float> cr(na * 7), ci(na * 7);
  for (int a = 0; a < na; a++) {
    double kr = (double)a * da;
    for (int j = -3; j <= 3; j++) {
      cr[a * 7 + j + 3] = (float)std::cos(kr * (double)j);
      ci[a * 7 + j + 3] = (float)std::sin(kr * (double)j);
    }
  }
  const float one_re = (float)std::cos(0.0), one_im = (float)std::sin(0.0);
  auto prod = [](float ar, float ai, float br, float bi, double& re, double& im) {
    float pr = ar * br - ai * bi;
    float pi = ar * bi + ai * br;
    re = pr; im = pi;
  };
  double Tr_max = 0, Trr_max = 0, Trrr_max = 0, Trrrr_max = 0, Trs_max = 0, Trrs_max = 0, Trrss_max = 0,
         Tr_so = 0, Trr_so = 0;
  const double D1so[3] = {-0.5, 0.0, 0.5}, D2so[3] = {1.0, -2.0, 1.0};
  for (int a = 0; a < na; a++) {
    double r1 = 0, i1 = 0, r2 = 0, i2 = 0, r3 = 0, i3 = 0, r4 = 0, i4 = 0, rs1 = 0, is1 = 0, rs2 = 0, is2 = 0;
    for (int j = -2; j <= 2; j++) {
      double re, im; prod(cr[a * 7 + j + 3], ci[a * 7 + j + 3], one_re, one_im, re, im);
      r1 = r1 + t.D1[j + 3] * re; i1 = i1 + t.D1[j + 3] * im;
      r2 = r2 + t.D2[j + 3] * re; i2 = i2 + t.D2[j + 3] * im;
    }
    for (int j = -1; j <= 1; j++) {
      double re, im; prod(cr[a * 7 + j + 3], ci[a * 7 + j + 3], one_re, one_im, re, im);
      rs1 = rs1 + D1so[j + 1] * re; is1 = is1 + D1so[j + 1] * im;
      rs2 = rs2 + D2so[j + 1] * re; is2 = is2 + D2so[j + 1] * im;
    }
    for (int j = -3; j <= 3; j++) {
      double re, im; prod(cr[a * 7 + j + 3], ci[a * 7 + j + 3], one_re, one_im, re, im);
      r3 = r3 + t.D3[j + 3] * re; i3 = i3 + t.D3[j + 3] * im;
      r4 = r4 + t.D4[j + 3] * re; i4 = i4 + t.D4[j + 3] * im;
    }
    // division by T_test(0,0) = (1,0) is exact for every complex division algorithm
    (void)r1; (void)r3; (void)i2; (void)i4; (void)rs1; (void)is2;
    if (std::fabs(i1) > Tr_max) Tr_max = std::fabs(i1);
    if (std::fabs(r2) > Trr_max) Trr_max = std::fabs(r2);
    if (std::fabs(i3) > Trrr_max) Trrr_max = std::fabs(i3);
    if (std::fabs(r4) > Trrrr_max) Trrrr_max = std::fabs(r4);
    if (std::fabs(is1) > Tr_so) Tr_so = std::fabs(is1);
    if (std::fabs(rs2) > Trr_so) Trr_so = std::fabs(rs2);
    for (int a2 = 0; a2 < na; a2++) {
      double rs = 0, ims = 0, rrs = 0, irs = 0, rrss = 0, irss = 0;
      for (int j = -2; j <= 2; j++)
        for (int jj = -2; jj <= 2; jj++) {
          double re, im;
          prod(cr[a * 7 + j + 3], ci[a * 7 + j + 3], cr[a2 * 7 + jj + 3], ci[a2 * 7 + jj + 3], re, im);
          double d1 = t.Drs[jj + 3][j + 3], d2 = t.Drrs[jj + 3][j + 3], d3 = t.Drrss[jj + 3][j + 3];
          rs = rs + d1 * re; ims = ims + d1 * im;
          rrs = rrs + d2 * re; irs = irs + d2 * im;
          rrss = rrss + d3 * re; irss = irss + d3 * im;
        }
      (void)ims; (void)rrs; (void)irss;
      if (std::fabs(rs) > Trs_max) Trs_max = std::fabs(rs);
      if (std::fabs(irs) > Trrs_max) Trrs_max = std::fabs(irs);
      if (std::fabs(rrss) > Trrss_max) Trrss_max = std::fabs(rrss);
    }
  }
  t.Tr_base = Tr_max; t.Trr_base = Trr_max; t.Trrr_base = Trrr_max; t.Trrrr_base = Trrrr_max;
  t.Trs_base = Trs_max; t.Trrs_base = Trrs_max; t.Trrss_base = Trrss_max;
  t.Tr_base_so = Tr_so; t.Trr_base_so = Trr_so;
}

// eigenvalue_prep for a uniform grid: PS/iterative_solver.f90:1232-1362
// (metric factors xr=zs=1 and their derivatives 0 kept in the expressions)
static void eigen_tables(Tables& t) {
  const double xr = 1.0, xrr = 0.0, xrrr = 0.0, xrrrr = 0.0, zs = 1.0, zss = 0.0, zsss = 0.0, zssss = 0.0;
  for (int l = 1; l <= t.grids; l++) {
    double Ts = t.Tr_base / t.dsg[l], Tss = -t.Trr_base / t.ds2g[l], Tsss = t.Trrr_base / t.ds3g[l],
           Tssss = t.Trrrr_base / t.ds4g[l];
    double Tr = t.Tr_base / t.drg[l], Trr = -t.Trr_base / t.dr2g[l], Trrr = t.Trrr_base / t.dr3g[l],
           Trrrr = t.Trrrr_base / t.dr4g[l];
    double Trs = -t.Trs_base / t.drdsg[l], Trrss = t.Trrss_base / t.dr2ds2g[l], Trss = t.Trrs_base / t.drds2g[l],
           Trrs = t.Trrs_base / t.dr2dsg[l];
    {
      double Sr = 0, Ss = 0, Srr = Trr, Sss = Tss, Srrr = 0, Ssss = 0, Srrrr = Trrrr, Sssss = Tssss, Srs = Trs,
             Srrs = 0, Srss = 0, Srrss = Trrss;
      double Sx = Sr / xr, Sz = Ss / zs;
      double Sxx = (Srr - Sx * xrr) / std::pow(xr, 2.0);
      double Szz = (Sss - Sz * zss) / std::pow(zs, 2.0);
      double Sxxx = (Srrr - 3.0 * Sxx * xr * xrr - Sx * xrrr) / std::pow(xr, 3.0);
      double Szzz = (Ssss - 3.0 * Szz * zs * zss - Sz * zsss) / std::pow(zs, 3.0);
      double Sxxxx = (Srrrr - 6.0 * Sxxx * xrr * std::pow(xr, 2.0) - (3.0 * std::pow(xrr, 2.0) + 4.0 * xr * xrrr) * Sxx - Sx * xrrrr) / std::pow(xr, 4.0);
      double Szzzz = (Sssss - 6.0 * Szzz * zss * std::pow(zs, 2.0) - (3.0 * std::pow(zss, 2.0) + 4.0 * zs * zsss) * Szz - Sz * zssss) / std::pow(zs, 4.0);
      double Sxz = Srs / xr / zs;
      double Sxxz = (Srrs - Sxz * zs * xrr) / zs / std::pow(xr, 2.0);
      double Sxzz = (Srss - Sxz * xr * zss) / xr / std::pow(zs, 2.0);
      double Sxxzz = (Srrss - Sxzz * xrr * std::pow(zs, 2.0) - Sxxz * zss * std::pow(xr, 2.0) - Sxz * xrr * zss) / std::pow(xr * zs, 2.0);
      t.SR1[l] = -(Sxx - Szz);
      t.SR2[l] = -((Sxxxx + Szzzz) + 2.0 * Sxxzz);
      t.SR4[l] = -(4.0 * Sxz);
      t.SR5[l] = -(2.0 * (Sxxx + Sxzz));
      t.SR6[l] = -(2.0 * (Szzz + Sxxz));
    }
    {
      double Sr = Tr, Ss = Ts, Srr = 0, Sss = 0, Srrr = Trrr, Ssss = Tsss, Srrrr = 0, Sssss = 0, Srs = 0,
             Srrs = Trrs, Srss = Trss, Srrss = 0;
      double Sx = Sr / xr;
      double Sxx = (Srr + std::fabs(Sx * xrr)) / std::pow(xr, 2.0);
      double Sxxx = (Srrr + std::fabs(3.0 * Sxx * xr * xrr) + std::fabs(Sx * xrrr)) / std::fabs(std::pow(xr, 3.0));
      double Sxxxx = (Srrrr + std::fabs(6.0 * Sxxx * xrr * std::pow(xr, 2.0)) +
                      (std::fabs(3.0 * std::pow(xrr, 2.0) + 4.0 * xr * xrrr) * std::fabs(Sxx)) + std::fabs(Sx * xrrrr)) /
                     std::pow(xr, 4.0);
      double Sz = Ss / zs;
      double Szz = (Sss + std::fabs(Sz * zss)) / std::pow(zs, 2.0);
      double Szzz = (Ssss + std::fabs(3.0 * Szz * zs * zss) + std::fabs(Sz * zsss)) / std::fabs(std::pow(zs, 3.0));
      double Szzzz = (Sssss + std::fabs(6.0 * Szzz * zss * std::pow(zs, 2.0)) +
                      (std::fabs(3.0 * std::pow(zss, 2.0) + 4.0 * zs * zsss) * std::fabs(Szz)) + std::fabs(Sz * zssss)) /
                     std::pow(zs, 4.0);
      double Sxz = Srs / xr / zs;
      double Sxxz = (Srrs + std::fabs(Sxz * zs * xrr)) / std::fabs(zs) / std::pow(xr, 2.0);
      double Sxzz = (Srss + std::fabs(Sxz * xr * zss)) / std::fabs(xr) / std::pow(zs, 2.0);
      double Sxxzz = (Srrss + std::fabs(Sxzz * xrr) * std::pow(zs, 2.0) + std::fabs(Sxxz * zss) * std::pow(xr, 2.0) +
                      std::fabs(Sxz * xrr * zss)) / std::pow(xr * zs, 2.0);
      t.SI1[l] = (Sxx + Szz);
      t.SI2[l] = std::fabs(Sxxxx + Szzzz) + std::fabs(2.0 * Sxxzz);
      t.SI4[l] = 4.0 * Sxz;
      t.SI5[l] = 2.0 * (Sxxx + Sxzz);
      t.SI6[l] = 2.0 * (Szzz + Sxxz);
    }
  }
}

// stream_space_matrix_prep: PS/iterative_solver.f90:692-743
static void coarse_stencils(Tables& t) {
  int g = t.grids;
  for (int dk = -3; dk <= 3; dk++)
    for (int di = -3; di <= 3; di++) {
      double Sxx = 0, Sxxx = 0, Sxxxx = 0, Szz = 0, Szzz = 0, Szzzz = 0;
      if (dk == 0) { Sxx = t.D2[di + 3] / t.dr2g[g]; Sxxx = t.D3[di + 3] / t.dr3g[g]; Sxxxx = t.D4[di + 3] / t.dr4g[g]; }
      if (di == 0) { Szz = t.D2[dk + 3] / t.ds2g[g]; Szzz = t.D3[dk + 3] / t.ds3g[g]; Szzzz = t.D4[dk + 3] / t.ds4g[g]; }
      double Sxz = t.Drs[dk + 3][di + 3] / t.drdsg[g];
      double Sxzz = t.Drss[dk + 3][di + 3] / t.drds2g[g];
      double Sxxz = t.Drrs[dk + 3][di + 3] / t.dr2dsg[g];
      double Sxxzz = t.Drrss[dk + 3][di + 3] / t.dr2ds2g[g];
      t.SM1[dk + 3][di + 3] = Sxx - Szz;
      t.SM2[dk + 3][di + 3] = Sxxxx + Szzzz + 2.0 * Sxxzz;
      t.SM4[dk + 3][di + 3] = 4.0 * Sxz;
      t.SM5[dk + 3][di + 3] = 2.0 * (Sxxx + Sxzz);
      t.SM6[dk + 3][di + 3] = 2.0 * (Szzz + Sxxz);
    }
}

// Compute_Matrix_coarse with analysis=1 (band widths): PS/iterative_solver.f90:1723-1745
static void coarse_bandwidth(Tables& t) {
  static const int limit[7] = {0, 2, 2, 3, 2, 2, 0};
  int nxc = t.nxg[t.grids], nzc = t.nzg[t.grids], N = nzc + 1;
  t.ngc = (nxc + 1) * (nzc + 1);
  int ku = 0, kl = 0;
  for (int p = 1; p <= t.ngc; p++) {
    int kp = p - N * ((p - 1) / N) - 1, ip = (p - 1) / N;
    for (int j = -3; j <= 3; j++) {
      int iq = j + ip;
      if (iq > 0 && iq < nxc)
        for (int kq = kp - limit[j + 3]; kq <= kp + limit[j + 3]; kq++)
          if (kq > 0 && kq < nzc) {
            int d = (N * iq + kq + 1) - p;
            if (d > ku) ku = d;
            if (-d > kl) kl = -d;
          }
    }
  }
  t.kl = kl; t.ku = ku;
}

// build_SF_matrix with tstep=-1 (PS/momentum.f90:341-483, 608-619): the band
// widths of the whole-grid operator, min/max of q-p over the folded stencil
// points (i = +-3 at k = 0, k = +-3 at i = 0, the 5x5 block) of every row that
// is not an identity row (corners; the top and bottom rows when rigid)
static void direct_bandwidth(Tables& t, int Vbc, int Database) {
  const int nx = t.nx, nz = t.nz;
  auto pf = [&](int ip, int kp) { return Database == 1 ? (nx + 1) * kp + ip + 1 : (nz + 1) * ip + kp + 1; };
  int sub = 0, sup = 0;
  auto visit = [&](int ip, int kp, int i, int k) {
    int iq = ip + i, kq = kp + k;
    if (iq < 0) iq = -iq; else if (iq > nx) iq = 2 * nx - iq;
    if (kq < 0) kq = -kq; else if (kq > nz) kq = 2 * nz - kq;
    const int d = pf(iq, kq) - pf(ip, kp);
    if (d < sub) sub = d;
    else if (d > sup) sup = d;
  };
  for (int ip = 0; ip <= nx; ip++)
    for (int kp = 0; kp <= nz; kp++) {
      const bool corner = (ip == 0 || ip == nx) && (kp == 0 || kp == nz);
      if (corner || (Vbc == 1 && (kp == 0 || kp == nz))) continue;
      for (int i = -3; i <= 3; i += 6) visit(ip, kp, i, 0);
      for (int k = -3; k <= 3; k += 6) visit(ip, kp, 0, k);
      for (int i = -2; i <= 2; i++)
        for (int k = -2; k <= 2; k++) visit(ip, kp, i, k);
    }
  t.dn = (nx + 1) * (nz + 1);
  t.dkl = -sub;
  t.dku = sup;
}

int build_tables(const pt_params& P, Tables& t) {
  std::memset(&t, 0, sizeof t);
  t.nx = P.nx; t.nz = P.nz; t.grids = P.grids; t.aspect = P.aspect;
  // compute_smoothing_parameters: PS/smoother.f90:31-73
  if (P.ivis == 0 || P.RaT == 0.0) t.delta_T = 2.0;
  else t.delta_T = (P.visT != 1.0) ? std::log(P.delta_visT) / std::log(P.visT) : 2.0;
  t.delta_C = 2.0;
  if (P.conduct_factor != 1.0) t.delta_C = std::fabs(std::log(P.delta_visC) / std::log(1.0 / P.conduct_factor));
  if (P.ivis >= 2 && P.visC != 1.0) t.delta_C = std::fabs(std::log(P.delta_visC) / std::log(P.visC));
  // the direct path (iterate=0) keeps nx, nz (PS/main.f90:170); two levels are
  // kept for the shared level-1 buffers
  if (P.iterate != 0) optimize_grid(t.nx, t.nz, t.grids, P.mingrid);
  else t.grids = 2;
  t.nsteps = cycle_schedule(t.grids, t.multi);
  if (t.nsteps < 0) return PT_ERR_UNSUPPORTED;
  // compute_basics: PS/basics.f90:204-300
  t.dr = t.aspect / (double)t.nx; t.ds = 1.0 / (double)t.nz;
  t.dr2 = std::pow(t.dr, 2.0); t.ds2 = std::pow(t.ds, 2.0); t.dr3 = std::pow(t.dr, 3.0); t.ds3 = std::pow(t.ds, 3.0);
  t.dr4 = std::pow(t.dr, 4.0); t.ds4 = std::pow(t.ds, 4.0);
  t.factorial[0] = 1.0; t.dr_power[0] = 1.0; t.ds_power[0] = 1.0;
  for (int j = 1; j <= 4; j++) {
    t.factorial[j] = (double)j * t.factorial[j - 1];
    t.dr_power[j] = std::pow(t.dr, (double)j); t.ds_power[j] = std::pow(t.ds, (double)j);
  }
  t.ntr = (int64_t)(t.nx + 1) * (t.nz + 1) * P.tpc;
  for (int j = 1; j <= t.grids; j++) {
    t.nxg[j] = t.nx / (1 << (j - 1)); t.nzg[j] = t.nz / (1 << (j - 1));
    t.drg[j] = t.aspect / (double)t.nxg[j]; t.dsg[j] = 1.0 / (double)t.nzg[j];
    t.dr2g[j] = std::pow(t.drg[j], 2.0); t.ds2g[j] = std::pow(t.dsg[j], 2.0);
    t.dr3g[j] = std::pow(t.drg[j], 3.0); t.ds3g[j] = std::pow(t.dsg[j], 3.0);
    t.dr4g[j] = std::pow(t.drg[j], 4.0); t.ds4g[j] = std::pow(t.dsg[j], 4.0);
    t.drdsg[j] = t.drg[j] * t.dsg[j]; t.drds2g[j] = t.drg[j] * t.ds2g[j];
    t.dr2dsg[j] = t.dr2g[j] * t.dsg[j]; t.dr2ds2g[j] = t.dr2g[j] * t.ds2g[j];
  }
  coarse_bandwidth(t);
  if (P.iterate == 0) direct_bandwidth(t, P.Vbc, P.Database);
  rkc1_tables(t, P.nstageS, P.dampingS);
  fd_tables(t);
  eigen_bases(t);
  // stability_RK4_smoother: PS/smoother.f90:719-740 (eI = 0 on a uniform grid)
  {
    double Tss = -t.Trr_base / t.ds2, Trr = -t.Trr_base / t.dr2;
    double eR = std::fabs(Trr / std::pow(1.0, 2.0)) + std::fabs(Tss / std::pow(1.0, 2.0));
    t.dts = 2.0 * eR / (std::pow(eR, 2.0) + std::pow(0.0, 2.0));
  }
  eigen_tables(t);
  coarse_stencils(t);
  // stability_vis_smoother: PS/smoother.f90:761-782
  for (int l = 1; l <= t.grids; l++) {
    double Tss = -t.Trr_base_so / t.ds2g[l], Trr = -t.Trr_base_so / t.dr2g[l];
    double eR = std::fabs(Trr / std::pow(1.0, 2.0)) + std::fabs(Tss / std::pow(1.0, 2.0));
    t.dt_vis[l] = 2.0 * eR / (std::pow(eR, 2.0) + std::pow(0.0, 2.0));
  }
  return PT_OK;
}

}  // namespace pt
