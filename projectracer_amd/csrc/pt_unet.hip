// pt_unet.hip -- the learned initial guess for the stream-function solve: a
// four-level convolutional encoder-decoder (3x3 convolutions, channels
// 32-64-128-256, skip connections, about 2.1M parameters) evaluated and
// trained with hand-written gfx950 kernels.  No reference counterpart: the
// reference warm-starts every solve from the previous psi
// (PS/iterative_solver.f90:134); this network proposes a correction to that
// warm start, and the multigrid cycles (the plain solver) always run to the
// same stop test afterwards.
//
// Data layout: activations are pixel-major (N, H, W, Cs) with the channels of
// a pixel contiguous (Cs the buffer's channel stride, a multiple of 8, so a
// skip connection and the decoder input share one buffer as two channel
// slices).  MFMA operands are BF16; every product accumulates in FP32.
// Weights keep an FP32 master copy (AdamW) and a BF16 copy for the MFMA
// kernels, conv weights laid out (tap, Cout, Cin) with Cin contiguous.
//
// Kernels:
//   k_conv_fwd    3x3 convolution as an implicit GEMM (pixels x Cout, K = 9 Cin)
//                 with v_mfma_f32_16x16x32_bf16; bias, ReLU, BF16 or FP32 out
//   k_conv_wgrad  weight gradient as a GEMM (Cout x Cin per tap, K = pixels),
//                 split over pixel chunks into FP32 partials, then k_sum_parts
//                 adds the chunks in a fixed order
//   k_wflip       weights of the data-gradient convolution (taps flipped, Cin
//                 and Cout exchanged): the data gradient is k_conv_fwd on them
//   k_pool/k_pool_bwd, k_up/k_up_bwd, k_relu_bwd, k_head/k_head_bwd, k_loss,
//   k_adamw       the elementwise and reduction parts
#include <cmath>
#include <cstring>
#include <random>
#include <vector>
#include "pt_internal.hpp"

namespace pt {

typedef __attribute__((ext_vector_type(4))) float f4;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__host__ __device__ inline uint16_t f2bf(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }
__device__ inline uint4 pack8(const uint16_t* v) {
  uint4 r;
  r.x = (uint32_t)v[0] | ((uint32_t)v[1] << 16);
  r.y = (uint32_t)v[2] | ((uint32_t)v[3] << 16);
  r.z = (uint32_t)v[4] | ((uint32_t)v[5] << 16);
  r.w = (uint32_t)v[6] | ((uint32_t)v[7] << 16);
  return r;
}

// ---------------------------------------------------------------------------
// tensors
// ---------------------------------------------------------------------------
struct Act {            // BF16 activation slice: channels [off, off+C) of stride Cs
  uint16_t* p;
  int Cs, off, C;
};
struct Grad {           // FP32 gradient slice
  float* p;
  int Cs, off, C;
};

// ---------------------------------------------------------------------------
// 3x3 convolution, implicit GEMM on MFMA.  One wavefront computes a 32-pixel
// x 32-channel output tile as 2 x 2 16x16 MFMA tiles; its K loop runs over
// the 9 taps and the input channels in steps of 32 (each lane holds 8
// consecutive channels of one pixel or one output channel).  The A fragment
// (pixels x channels) reads the input pixel shifted by the tap, zero outside
// the image; the B fragment (channels x outputs) reads the weights.
// Output modes: 0 BF16 store, 1 FP32 store, 2 FP32 accumulate.
// ---------------------------------------------------------------------------
struct ConvArgs {
  const uint16_t* x; int xs, xo, Cin;
  const uint16_t* w;      // (9, Cout, Cin)
  const float* bias;      // Cout or null
  void* y; int ys, yo, Cout;
  int N, H, W;
  int relu, mode;
};

__global__ void __launch_bounds__(256) k_conv_fwd(ConvArgs a) {
  const int lane = threadIdx.x & 63;
  const long wave = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const long P = (long)a.N * a.H * a.W;
  const long npt = (P + 31) / 32;
  const int nct = (a.Cout + 31) / 32;
  if (wave >= npt * nct) return;
  const long pt0 = (wave % npt) * 32;
  const int ct0 = (int)(wave / npt) * 32;
  const int r16 = lane & 15, g = lane >> 4;   // row within a 16-tile, K group
  // the two pixels this lane feeds into the A fragments
  long px[2];
  int pn[2], py[2], pxx[2];
  bool pv[2];
  for (int t = 0; t < 2; t++) {
    px[t] = pt0 + t * 16 + r16;
    pv[t] = px[t] < P;
    long q = pv[t] ? px[t] : 0;
    pxx[t] = (int)(q % a.W);
    q /= a.W;
    py[t] = (int)(q % a.H);
    pn[t] = (int)(q / a.H);
  }
  f4 acc[2][2];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  const uint4 zero = {0u, 0u, 0u, 0u};
  for (int tap = 0; tap < 9; tap++) {
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    const uint16_t* xr[2];
    bool in[2];
    for (int t = 0; t < 2; t++) {
      const int yy = py[t] + dy, xx = pxx[t] + dx;
      in[t] = pv[t] && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W;
      const long pp = in[t] ? (((long)pn[t] * a.H + yy) * a.W + xx) : 0;
      xr[t] = a.x + pp * a.xs + a.xo;
    }
    const uint16_t* wt = a.w + (long)tap * a.Cout * a.Cin;
    for (int c0 = 0; c0 < a.Cin; c0 += 32) {
      const int ck = c0 + 8 * g;
      const bool kin = ck < a.Cin;
      uint4 av[2], bv[2];
      for (int t = 0; t < 2; t++)
        av[t] = (in[t] && kin) ? *reinterpret_cast<const uint4*>(xr[t] + ck) : zero;
      for (int u = 0; u < 2; u++) {
        const int co = ct0 + u * 16 + r16;
        bv[u] = (kin && co < a.Cout) ? *reinterpret_cast<const uint4*>(wt + (long)co * a.Cin + ck) : zero;
      }
      for (int t = 0; t < 2; t++)
        for (int u = 0; u < 2; u++)
          acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, av[t]),
                                                              __builtin_bit_cast(bf16x8, bv[u]), acc[t][u], 0, 0, 0);
    }
  }
  // C layout: lane holds output channel (col) r16 of the tile and pixels
  // (rows) 4g .. 4g+3
  for (int t = 0; t < 2; t++)
    for (int u = 0; u < 2; u++) {
      const int co = ct0 + u * 16 + r16;
      if (co >= a.Cout) continue;
      const float b = a.bias ? a.bias[co] : 0.f;
      for (int e = 0; e < 4; e++) {
        const long p = pt0 + t * 16 + 4 * g + e;
        if (p >= P) continue;
        float v = acc[t][u][e] + b;
        if (a.relu) v = fmaxf(v, 0.f);
        const long o = p * a.ys + a.yo + co;
        if (a.mode == 0) reinterpret_cast<uint16_t*>(a.y)[o] = f2bf(v);
        else if (a.mode == 1) reinterpret_cast<float*>(a.y)[o] = v;
        else reinterpret_cast<float*>(a.y)[o] += v;
      }
    }
}

// weights of the data-gradient convolution: wb(tap', ci, co) = w(8 - tap, co, ci)
__global__ void k_wflip(const uint16_t* w, uint16_t* wb, int Cout, int Cin) {
  const long n = 9L * Cout * Cin;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int co = (int)(t % Cout);
  const long r = t / Cout;
  const int ci = (int)(r % Cin);
  const int tp = (int)(r / Cin);
  wb[t] = w[((long)(8 - tp) * Cout + co) * Cin + ci];
}

// ---------------------------------------------------------------------------
// weight gradient: dW(tap, co, ci) = sum_p G(p, co) X(p + off(tap), ci).  A
// wavefront computes a 32 (co) x 32 (ci) tile of one tap over one chunk of
// KCH pixels; lane operands are 8 consecutive pixels of one channel.
// part[chunk][tap][co][ci] then k_sum_parts adds the chunks in order.
// ---------------------------------------------------------------------------
constexpr int KCH = 2048;
struct WgradArgs {
  const uint16_t* g; int gs, go, Cout;   // BF16 output gradient (ReLU applied)
  const uint16_t* x; int xs, xo, Cin;
  float* part;
  int N, H, W;
};
__global__ void __launch_bounds__(256) k_conv_wgrad(WgradArgs a) {
  const int lane = threadIdx.x & 63;
  const long wave = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const long P = (long)a.N * a.H * a.W;
  const int nchunk = (int)((P + KCH - 1) / KCH);
  const int nco = (a.Cout + 31) / 32, nci = (a.Cin + 31) / 32;
  const long total = (long)nchunk * 9 * nco * nci;
  if (wave >= total) return;
  long r = wave;
  const int cit = (int)(r % nci); r /= nci;
  const int cot = (int)(r % nco); r /= nco;
  const int tap = (int)(r % 9);
  const int chunk = (int)(r / 9);
  const int dy = tap / 3 - 1, dx = tap % 3 - 1;
  const int r16 = lane & 15, g = lane >> 4;
  f4 acc[2][2];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  const long p_end = P < (long)(chunk + 1) * KCH ? P : (long)(chunk + 1) * KCH;
  for (long p0 = (long)chunk * KCH; p0 < p_end; p0 += 32) {
    uint16_t av[2][8], bv[2][8];
    for (int e = 0; e < 8; e++) {
      const long p = p0 + 8 * g + e;
      const bool pv = p < p_end;
      long q = pv ? p : 0;
      const int xx = (int)(q % a.W);
      q /= a.W;
      const int yy = (int)(q % a.H);
      const int nn = (int)(q / a.H);
      const int sy = yy + dy, sx = xx + dx;
      const bool sin = pv && sy >= 0 && sy < a.H && sx >= 0 && sx < a.W;
      const long sp = sin ? (((long)nn * a.H + sy) * a.W + sx) : 0;
      for (int t = 0; t < 2; t++) {
        const int co = cot * 32 + t * 16 + r16;
        av[t][e] = (pv && co < a.Cout) ? a.g[p * a.gs + a.go + co] : (uint16_t)0;
      }
      for (int u = 0; u < 2; u++) {
        const int ci = cit * 32 + u * 16 + r16;
        bv[u][e] = (sin && ci < a.Cin) ? a.x[sp * a.xs + a.xo + ci] : (uint16_t)0;
      }
    }
    for (int t = 0; t < 2; t++)
      for (int u = 0; u < 2; u++) {
        const uint4 ua = pack8(av[t]), ub = pack8(bv[u]);
        acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, ua),
                                                            __builtin_bit_cast(bf16x8, ub), acc[t][u], 0, 0, 0);
      }
  }
  // C: lane holds ci (col) r16 of the tile, co (rows) 4g..4g+3
  for (int t = 0; t < 2; t++)
    for (int u = 0; u < 2; u++) {
      const int ci = cit * 32 + u * 16 + r16;
      if (ci >= a.Cin) continue;
      for (int e = 0; e < 4; e++) {
        const int co = cot * 32 + t * 16 + 4 * g + e;
        if (co >= a.Cout) continue;
        a.part[(((long)chunk * 9 + tap) * a.Cout + co) * a.Cin + ci] = acc[t][u][e];
      }
    }
}
__global__ void k_sum_parts(const float* part, int nchunk, long n, float* out) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  float s = 0.f;
  for (int c = 0; c < nchunk; c++) s += part[(long)c * n + t];
  out[t] = s;
}
// bias gradient: db(co) = sum_p G(p, co), one workgroup per channel, fixed order
__global__ void k_bias_grad(const uint16_t* g, int gs, int go, long P, float* db) {
  __shared__ float red[256];
  const int co = blockIdx.x;
  float s = 0.f;
  for (long p = threadIdx.x; p < P; p += 256) s += bf2f(g[p * gs + go + co]);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[co] = red[0];
}

// ReLU backward: G = dY * (y > 0), to BF16 for the MFMA kernels
__global__ void k_relu_bwd(const float* dy, int ds, int dof, const uint16_t* y, int ys, int yo, uint16_t* g, int C,
                           long P) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P * C) return;
  const long p = t / C;
  const int c = (int)(t % C);
  const float v = bf2f(y[p * ys + yo + c]) > 0.f ? dy[p * ds + dof + c] : 0.f;
  g[p * C + c] = f2bf(v);
}

// 2x2 average pooling and its gradient
__global__ void k_pool(const uint16_t* x, int xs, int xo, uint16_t* y, int C, int N, int H, int W) {
  const int Ho = H / 2, Wo = W / 2;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)N * Ho * Wo * C) return;
  const int c = (int)(t % C);
  long r = t / C;
  const int xo_ = (int)(r % Wo); r /= Wo;
  const int yo_ = (int)(r % Ho);
  const int n = (int)(r / Ho);
  float s = 0.f;
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) s += bf2f(x[(((long)n * H + 2 * yo_ + a) * W + 2 * xo_ + b) * xs + xo + c]);
  y[t] = f2bf(0.25f * s);
}
__global__ void k_pool_bwd(const float* dy, float* dx, int xs, int xo, int C, int N, int H, int W) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)N * H * W * C) return;
  const int c = (int)(t % C);
  long r = t / C;
  const int xx = (int)(r % W); r /= W;
  const int yy = (int)(r % H);
  const int n = (int)(r / H);
  const float v = dy[(((long)n * (H / 2) + yy / 2) * (W / 2) + xx / 2) * C + c];
  dx[(((long)n * H + yy) * W + xx) * xs + xo + c] += 0.25f * v;
}
// nearest-neighbour upsampling by 2 and its gradient (sum of the four)
__global__ void k_up(const uint16_t* x, uint16_t* y, int C, int N, int H, int W) {   // H, W: output
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)N * H * W * C) return;
  const int c = (int)(t % C);
  long r = t / C;
  const int xx = (int)(r % W); r /= W;
  const int yy = (int)(r % H);
  const int n = (int)(r / H);
  y[t] = x[(((long)n * (H / 2) + yy / 2) * (W / 2) + xx / 2) * C + c];
}
__global__ void k_up_bwd(const float* dy, float* dx, int C, int N, int H, int W) {   // H, W: input (coarse)
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)N * H * W * C) return;
  const int c = (int)(t % C);
  long r = t / C;
  const int xx = (int)(r % W); r /= W;
  const int yy = (int)(r % H);
  const int n = (int)(r / H);
  float s = 0.f;
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) s += dy[(((long)n * 2 * H + 2 * yy + a) * 2 * W + 2 * xx + b) * C + c];
  dx[t] = s;
}

// output head: 1x1 convolution of C channels to one (FP32)
__global__ void k_head(const uint16_t* x, int C, const float* w, const float* b, float* y, long P) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  float s = b[0];
  for (int c = 0; c < C; c++) s += w[c] * bf2f(x[p * C + c]);
  y[p] = s;
}
// its gradient: dx(p, c) = w(c) dy(p); dw, db by fixed-order reductions
__global__ void k_head_bwd_x(const float* dy, const float* w, float* dx, int C, long P) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P * C) return;
  dx[t] = w[t % C] * dy[t / C];
}
__global__ void k_head_bwd_w(const float* dy, const uint16_t* x, int C, long P, float* dw, float* db) {
  __shared__ float red[256];
  const int c = blockIdx.x;   // C blocks for dw, one more for db
  float s = 0.f;
  for (long p = threadIdx.x; p < P; p += 256) s += (c < C) ? dy[p] * bf2f(x[p * C + c]) : dy[p];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (c < C) dw[c] = red[0];
    else db[0] = red[0];
  }
}

// Loss: masked mean square of e = y - t plus lam times the masked mean
// square of the 5-point Laplacian of e on each image (zero outside it); the
// second term weights an error by its squared curvature, the way the
// stream-function operator amplifies a correction's error:
//   loss = (sum m e^2 + lam sum (m lap e)^2) / sum m
//   dy   = (2 m e + 2 lam lap(m lap e)) / sum m
__device__ inline float lap_at(const float* f, long p, int i, int k, int H, int W) {
  const float c = f[p];
  float s = -4.f * c;
  if (i > 0) s += f[p - 1];
  if (i < W - 1) s += f[p + 1];
  if (k > 0) s += f[p - W];
  if (k < H - 1) s += f[p + W];
  return s;
}
__global__ void k_err(const float* y, const float* tg, float* e, long P) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < P) e[p] = y[p] - tg[p];
}
__global__ void k_lap_err(const float* e, const float* m, int H, int W, long P, float* le) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const int i = (int)(p % W), k = (int)((p / W) % H);
  le[p] = m[p] * lap_at(e, p, i, k, H, W);
}
__global__ void k_loss(const float* e, const float* le, const float* m, long P, double* part) {
  __shared__ double r1[256], r2[256], r3[256];
  double s1 = 0.0, s2 = 0.0, s3 = 0.0;
  for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < P; p += (long)gridDim.x * 256) {
    const double d = (double)e[p];
    s1 += (double)m[p] * d * d;
    s2 += (double)m[p];
    s3 += (double)le[p] * (double)le[p];
  }
  r1[threadIdx.x] = s1;
  r2[threadIdx.x] = s2;
  r3[threadIdx.x] = s3;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      r1[threadIdx.x] += r1[threadIdx.x + st];
      r2[threadIdx.x] += r2[threadIdx.x + st];
      r3[threadIdx.x] += r3[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = r1[0];
    part[3 * blockIdx.x + 1] = r2[0];
    part[3 * blockIdx.x + 2] = r3[0];
  }
}
__global__ void k_loss_final(const double* part, int nb, double lam, double* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s1 = 0.0, s2 = 0.0, s3 = 0.0;
  for (int b = 0; b < nb; b++) {
    s1 += part[3 * b];
    s2 += part[3 * b + 1];
    s3 += part[3 * b + 2];
  }
  out[0] = (s1 + lam * s3) / fmax(s2, 1.0);
  out[1] = s2;
}
__global__ void k_loss_grad(const float* e, const float* le, const float* m, const double* lm, double lam, int H,
                            int W, long P, float* dy) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const int i = (int)(p % W), k = (int)((p / W) % H);
  const double g = 2.0 * (double)m[p] * (double)e[p] + 2.0 * lam * (double)lap_at(le, p, i, k, H, W);
  dy[p] = (float)(g / fmax(lm[1], 1.0));
}

// AdamW on the flat FP32 parameters, then their BF16 copy
__global__ void k_adamw(float* w, const float* g, float* m, float* v, uint16_t* wb, long n, float lr, float b1,
                        float b2, float eps, float wd, float c1, float c2) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float gt = g[t];
  const float mt = b1 * m[t] + (1.f - b1) * gt;
  const float vt = b2 * v[t] + (1.f - b2) * gt * gt;
  m[t] = mt;
  v[t] = vt;
  const float mh = mt / c1, vh = vt / c2;
  const float wt = w[t] - lr * (mh / (sqrtf(vh) + eps) + wd * w[t]);
  w[t] = wt;
  wb[t] = f2bf(wt);
}
__global__ void k_to_bf16(const float* a, uint16_t* b, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = f2bf(a[t]);
}
__global__ void k_fill0(float* a, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = 0.f;
}

// ---------------------------------------------------------------------------
// network description
// ---------------------------------------------------------------------------
constexpr int IN_CH = 8;    // input channels in memory (5 used, the rest zero)
constexpr int NUSED_IN = 5;
struct ConvL {
  int Cin, Cout;
  long w, b;   // offsets into the flat parameter array
};
// the 17 convolutions in order: enc0a, enc0b, enc1a, enc1b, enc2a, enc2b,
// enc3a, enc3b, up3, dec2a, dec2b, up2, dec1a, dec1b, up1, dec0a, dec0b
static const int kConvDims[17][2] = {{IN_CH, 32}, {32, 32},   {32, 64},   {64, 64},   {64, 128},  {128, 128},
                                     {128, 256},  {256, 256}, {256, 128}, {256, 128}, {128, 128}, {128, 64},
                                     {128, 64},   {64, 64},   {64, 32},   {64, 32},   {32, 32}};

struct Net {
  ConvL conv[17];
  long head_w, head_b, nparam;
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;   // flat FP32
  uint16_t* wb = nullptr;     // flat BF16 copy
  uint16_t* wflip = nullptr;  // scratch: flipped weights of one layer
  float* part = nullptr; size_t part_cap = 0;
  long step = 0;
  float out_scale = 1.f;      // target scale c0 (the dataset's rms target), fixed at the first training call
  bool out_scale_set = false;
  double lap_weight = 0.0;    // weight of the error's Laplacian term in the loss
  int mask_N = 0, mask_H = 0, mask_W = 0;   // shape n.msk holds the interior mask for
  // activation buffers for the current (N, H, W)
  int N = 0, H = 0, W = 0;
  std::vector<void*> bufs;
  uint16_t *x0, *e0a, *cat0, *p1, *e1a, *cat1, *p2, *e2a, *cat2, *p3, *e3a, *e3b, *u3, *d2a, *d2b, *u2, *d1a, *d1b,
      *u1, *d0a, *d0b, *gbf;
  float *y, *dy, *gA, *gB, *gcat0, *gcat1, *gcat2, *tgt, *msk;
  double* lpart = nullptr;
  double* lout = nullptr;
  double* h_lout = nullptr;
  // dataset (device resident): inputs BF16 (cap, H, W, IN_CH), targets FP32
  int dH = 0, dW = 0, dcap = 0, dcount = 0;
  uint16_t* dx = nullptr;
  float* dt = nullptr;
  hipStream_t stream = nullptr;
};

static long npix(const Net& n) { return (long)n.N * n.H * n.W; }

static int net_alloc_bufs(Net& n, int N, int H, int W) {
  if (n.N == N && n.H == H && n.W == W) return PT_OK;
  for (void* p : n.bufs) (void)hipFree(p);
  n.bufs.clear();
  n.N = N; n.H = H; n.W = W;
  auto abf = [&](uint16_t** p, long px, int C) -> int {
    PT_HIP(hipMalloc((void**)p, sizeof(uint16_t) * (size_t)px * C));
    PT_HIP(hipMemset(*p, 0, sizeof(uint16_t) * (size_t)px * C));
    n.bufs.push_back(*p);
    return PT_OK;
  };
  auto af = [&](float** p, long cnt) -> int {
    PT_HIP(hipMalloc((void**)p, sizeof(float) * (size_t)cnt));
    PT_HIP(hipMemset(*p, 0, sizeof(float) * (size_t)cnt));
    n.bufs.push_back(*p);
    return PT_OK;
  };
  const long P0 = (long)N * H * W, P1 = P0 / 4, P2 = P1 / 4, P3 = P2 / 4;
  PT_TRY(abf(&n.x0, P0, IN_CH)); PT_TRY(abf(&n.e0a, P0, 32)); PT_TRY(abf(&n.cat0, P0, 64));
  PT_TRY(abf(&n.p1, P1, 32)); PT_TRY(abf(&n.e1a, P1, 64)); PT_TRY(abf(&n.cat1, P1, 128));
  PT_TRY(abf(&n.p2, P2, 64)); PT_TRY(abf(&n.e2a, P2, 128)); PT_TRY(abf(&n.cat2, P2, 256));
  PT_TRY(abf(&n.p3, P3, 128)); PT_TRY(abf(&n.e3a, P3, 256)); PT_TRY(abf(&n.e3b, P3, 256));
  PT_TRY(abf(&n.u3, P2, 256)); PT_TRY(abf(&n.d2a, P2, 128)); PT_TRY(abf(&n.d2b, P2, 128));
  PT_TRY(abf(&n.u2, P1, 128)); PT_TRY(abf(&n.d1a, P1, 64)); PT_TRY(abf(&n.d1b, P1, 64));
  PT_TRY(abf(&n.u1, P0, 64)); PT_TRY(abf(&n.d0a, P0, 32)); PT_TRY(abf(&n.d0b, P0, 32));
  PT_TRY(abf(&n.gbf, P0, 64));   // BF16 output gradient of the layer in hand (largest: 64 ch at full size)
  PT_TRY(af(&n.y, P0)); PT_TRY(af(&n.dy, P0)); PT_TRY(af(&n.tgt, P0)); PT_TRY(af(&n.msk, P0));
  PT_TRY(af(&n.gA, P0 * 64)); PT_TRY(af(&n.gB, P0 * 64));
  PT_TRY(af(&n.gcat0, P0 * 64)); PT_TRY(af(&n.gcat1, P1 * 128)); PT_TRY(af(&n.gcat2, P2 * 256));
  return PT_OK;
}

static unsigned nblk(long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

// y = conv(x) with the layer's BF16 weights
static void conv_fwd(Net& n, int li, Act x, Act y, int N, int H, int W, bool relu) {
  const ConvL& L = n.conv[li];
  ConvArgs a;
  a.x = x.p; a.xs = x.Cs; a.xo = x.off; a.Cin = L.Cin;
  a.w = n.wb + L.w; a.bias = n.w + L.b;
  a.y = y.p; a.ys = y.Cs; a.yo = y.off; a.Cout = L.Cout;
  a.N = N; a.H = H; a.W = W; a.relu = relu ? 1 : 0; a.mode = 0;
  const long P = (long)N * H * W;
  const long waves = ((P + 31) / 32) * ((L.Cout + 31) / 32);
  hipLaunchKernelGGL(k_conv_fwd, dim3(nblk(waves, 4)), dim3(256), 0, n.stream, a);
}

// backward of y = relu(conv(x)): dY (FP32 slice) -> weight and bias grads,
// and dX (FP32, stored or accumulated; skipped when dx.p is null)
static int conv_bwd(Net& n, int li, Act x, Act y, Grad dyv, Grad dx, bool accumulate, int N, int H, int W) {
  const ConvL& L = n.conv[li];
  const long P = (long)N * H * W;
  hipLaunchKernelGGL(k_relu_bwd, dim3(nblk(P * L.Cout)), dim3(256), 0, n.stream, dyv.p, dyv.Cs, dyv.off, y.p, y.Cs,
                     y.off, n.gbf, L.Cout, P);
  // weight gradient
  const int nchunk = (int)((P + KCH - 1) / KCH);
  const size_t need = (size_t)nchunk * 9 * L.Cout * L.Cin;
  if (need > n.part_cap) {
    if (n.part) (void)hipFree(n.part);
    PT_HIP(hipMalloc((void**)&n.part, sizeof(float) * need));
    n.part_cap = need;
  }
  WgradArgs wa;
  wa.g = n.gbf; wa.gs = L.Cout; wa.go = 0; wa.Cout = L.Cout;
  wa.x = x.p; wa.xs = x.Cs; wa.xo = x.off; wa.Cin = L.Cin;
  wa.part = n.part; wa.N = N; wa.H = H; wa.W = W;
  const long waves = (long)nchunk * 9 * ((L.Cout + 31) / 32) * ((L.Cin + 31) / 32);
  hipLaunchKernelGGL(k_conv_wgrad, dim3(nblk(waves, 4)), dim3(256), 0, n.stream, wa);
  const long nw = 9L * L.Cout * L.Cin;
  hipLaunchKernelGGL(k_sum_parts, dim3(nblk(nw)), dim3(256), 0, n.stream, n.part, nchunk, nw, n.g + L.w);
  hipLaunchKernelGGL(k_bias_grad, dim3(L.Cout), dim3(256), 0, n.stream, n.gbf, L.Cout, 0, P, n.g + L.b);
  if (!dx.p) return PT_OK;
  // data gradient: the convolution of G with the flipped, transposed weights
  hipLaunchKernelGGL(k_wflip, dim3(nblk(nw)), dim3(256), 0, n.stream, n.wb + L.w, n.wflip, L.Cout, L.Cin);
  ConvArgs a;
  a.x = n.gbf; a.xs = L.Cout; a.xo = 0; a.Cin = L.Cout;
  a.w = n.wflip; a.bias = nullptr;
  a.y = dx.p; a.ys = dx.Cs; a.yo = dx.off; a.Cout = L.Cin;
  a.N = N; a.H = H; a.W = W; a.relu = 0; a.mode = accumulate ? 2 : 1;
  const long dwaves = ((P + 31) / 32) * ((L.Cin + 31) / 32);
  hipLaunchKernelGGL(k_conv_fwd, dim3(nblk(dwaves, 4)), dim3(256), 0, n.stream, a);
  return PT_OK;
}

static void pool(Net& n, Act x, uint16_t* y, int C, int N, int H, int W) {
  hipLaunchKernelGGL(k_pool, dim3(nblk((long)N * (H / 2) * (W / 2) * C)), dim3(256), 0, n.stream, x.p, x.Cs, x.off, y,
                     C, N, H, W);
}
static void up(Net& n, const uint16_t* x, uint16_t* y, int C, int N, int H, int W) {
  hipLaunchKernelGGL(k_up, dim3(nblk((long)N * H * W * C)), dim3(256), 0, n.stream, x, y, C, N, H, W);
}

// the forward pass on n.x0 (N, H, W): prediction in n.y
static void forward(Net& n) {
  const int N = n.N, H = n.H, W = n.W;
  const int H1 = H / 2, W1 = W / 2, H2 = H / 4, W2 = W / 4, H3 = H / 8, W3 = W / 8;
  conv_fwd(n, 0, {n.x0, IN_CH, 0, IN_CH}, {n.e0a, 32, 0, 32}, N, H, W, true);
  conv_fwd(n, 1, {n.e0a, 32, 0, 32}, {n.cat0, 64, 32, 32}, N, H, W, true);
  pool(n, {n.cat0, 64, 32, 32}, n.p1, 32, N, H, W);
  conv_fwd(n, 2, {n.p1, 32, 0, 32}, {n.e1a, 64, 0, 64}, N, H1, W1, true);
  conv_fwd(n, 3, {n.e1a, 64, 0, 64}, {n.cat1, 128, 64, 64}, N, H1, W1, true);
  pool(n, {n.cat1, 128, 64, 64}, n.p2, 64, N, H1, W1);
  conv_fwd(n, 4, {n.p2, 64, 0, 64}, {n.e2a, 128, 0, 128}, N, H2, W2, true);
  conv_fwd(n, 5, {n.e2a, 128, 0, 128}, {n.cat2, 256, 128, 128}, N, H2, W2, true);
  pool(n, {n.cat2, 256, 128, 128}, n.p3, 128, N, H2, W2);
  conv_fwd(n, 6, {n.p3, 128, 0, 128}, {n.e3a, 256, 0, 256}, N, H3, W3, true);
  conv_fwd(n, 7, {n.e3a, 256, 0, 256}, {n.e3b, 256, 0, 256}, N, H3, W3, true);
  up(n, n.e3b, n.u3, 256, N, H2, W2);
  conv_fwd(n, 8, {n.u3, 256, 0, 256}, {n.cat2, 256, 0, 128}, N, H2, W2, true);
  conv_fwd(n, 9, {n.cat2, 256, 0, 256}, {n.d2a, 128, 0, 128}, N, H2, W2, true);
  conv_fwd(n, 10, {n.d2a, 128, 0, 128}, {n.d2b, 128, 0, 128}, N, H2, W2, true);
  up(n, n.d2b, n.u2, 128, N, H1, W1);
  conv_fwd(n, 11, {n.u2, 128, 0, 128}, {n.cat1, 128, 0, 64}, N, H1, W1, true);
  conv_fwd(n, 12, {n.cat1, 128, 0, 128}, {n.d1a, 64, 0, 64}, N, H1, W1, true);
  conv_fwd(n, 13, {n.d1a, 64, 0, 64}, {n.d1b, 64, 0, 64}, N, H1, W1, true);
  up(n, n.d1b, n.u1, 64, N, H, W);
  conv_fwd(n, 14, {n.u1, 64, 0, 64}, {n.cat0, 64, 0, 32}, N, H, W, true);
  conv_fwd(n, 15, {n.cat0, 64, 0, 64}, {n.d0a, 32, 0, 32}, N, H, W, true);
  conv_fwd(n, 16, {n.d0a, 32, 0, 32}, {n.d0b, 32, 0, 32}, N, H, W, true);
  hipLaunchKernelGGL(k_head, dim3(nblk(npix(n))), dim3(256), 0, n.stream, n.d0b, 32, n.w + n.head_w, n.w + n.head_b,
                     n.y, npix(n));
}

// loss of n.y against n.tgt under n.msk, gradient into n.g; loss to host
static int loss_backward(Net& n, double* loss) {
  const int N = n.N, H = n.H, W = n.W;
  const int H1 = H / 2, W1 = W / 2, H2 = H / 4, W2 = W / 4, H3 = H / 8, W3 = W / 8;
  const long P0 = npix(n), P1 = P0 / 4, P2 = P1 / 4, P3 = P2 / 4;
  const int nb = 256;
  // e in gB, m lap e in gA (both free until the head's backward)
  hipLaunchKernelGGL(k_err, dim3(nblk(P0)), dim3(256), 0, n.stream, n.y, n.tgt, n.gB, P0);
  hipLaunchKernelGGL(k_lap_err, dim3(nblk(P0)), dim3(256), 0, n.stream, n.gB, n.msk, H, W, P0, n.gA);
  hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, n.stream, n.gB, n.gA, n.msk, P0, n.lpart);
  hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(64), 0, n.stream, n.lpart, nb, n.lap_weight, n.lout);
  hipLaunchKernelGGL(k_loss_grad, dim3(nblk(P0)), dim3(256), 0, n.stream, n.gB, n.gA, n.msk, n.lout, n.lap_weight, H,
                     W, P0, n.dy);
  hipLaunchKernelGGL(k_fill0, dim3(nblk(n.nparam)), dim3(256), 0, n.stream, n.g, n.nparam);
  hipLaunchKernelGGL(k_fill0, dim3(nblk(P0 * 64)), dim3(256), 0, n.stream, n.gcat0, P0 * 64);
  hipLaunchKernelGGL(k_fill0, dim3(nblk(P1 * 128)), dim3(256), 0, n.stream, n.gcat1, P1 * 128);
  hipLaunchKernelGGL(k_fill0, dim3(nblk(P2 * 256)), dim3(256), 0, n.stream, n.gcat2, P2 * 256);
  // head
  hipLaunchKernelGGL(k_head_bwd_w, dim3(33), dim3(256), 0, n.stream, n.dy, n.d0b, 32, P0, n.g + n.head_w,
                     n.g + n.head_b);
  hipLaunchKernelGGL(k_head_bwd_x, dim3(nblk(P0 * 32)), dim3(256), 0, n.stream, n.dy, n.w + n.head_w, n.gA, 32, P0);
  // level 0 decoder: gA = d(d0b)
  PT_TRY(conv_bwd(n, 16, {n.d0a, 32, 0, 32}, {n.d0b, 32, 0, 32}, {n.gA, 32, 0, 32}, {n.gB, 32, 0, 32}, false, N, H, W));
  PT_TRY(conv_bwd(n, 15, {n.cat0, 64, 0, 64}, {n.d0a, 32, 0, 32}, {n.gB, 32, 0, 32}, {n.gcat0, 64, 0, 64}, true, N, H,
                  W));
  // up1 conv: its output is cat0[0:32]; input u1 (64 ch)
  PT_TRY(conv_bwd(n, 14, {n.u1, 64, 0, 64}, {n.cat0, 64, 0, 32}, {n.gcat0, 64, 0, 32}, {n.gA, 64, 0, 64}, false, N, H,
                  W));
  hipLaunchKernelGGL(k_up_bwd, dim3(nblk(P1 * 64)), dim3(256), 0, n.stream, n.gA, n.gB, 64, N, H1, W1);   // d(d1b)
  PT_TRY(conv_bwd(n, 13, {n.d1a, 64, 0, 64}, {n.d1b, 64, 0, 64}, {n.gB, 64, 0, 64}, {n.gA, 64, 0, 64}, false, N, H1,
                  W1));
  PT_TRY(conv_bwd(n, 12, {n.cat1, 128, 0, 128}, {n.d1a, 64, 0, 64}, {n.gA, 64, 0, 64}, {n.gcat1, 128, 0, 128}, true, N,
                  H1, W1));
  PT_TRY(conv_bwd(n, 11, {n.u2, 128, 0, 128}, {n.cat1, 128, 0, 64}, {n.gcat1, 128, 0, 64}, {n.gA, 128, 0, 128}, false,
                  N, H1, W1));
  hipLaunchKernelGGL(k_up_bwd, dim3(nblk(P2 * 128)), dim3(256), 0, n.stream, n.gA, n.gB, 128, N, H2, W2);   // d(d2b)
  PT_TRY(conv_bwd(n, 10, {n.d2a, 128, 0, 128}, {n.d2b, 128, 0, 128}, {n.gB, 128, 0, 128}, {n.gA, 128, 0, 128}, false,
                  N, H2, W2));
  PT_TRY(conv_bwd(n, 9, {n.cat2, 256, 0, 256}, {n.d2a, 128, 0, 128}, {n.gA, 128, 0, 128}, {n.gcat2, 256, 0, 256},
                  true, N, H2, W2));
  PT_TRY(conv_bwd(n, 8, {n.u3, 256, 0, 256}, {n.cat2, 256, 0, 128}, {n.gcat2, 256, 0, 128}, {n.gA, 256, 0, 256}, false,
                  N, H2, W2));
  hipLaunchKernelGGL(k_up_bwd, dim3(nblk(P3 * 256)), dim3(256), 0, n.stream, n.gA, n.gB, 256, N, H3, W3);   // d(e3b)
  PT_TRY(conv_bwd(n, 7, {n.e3a, 256, 0, 256}, {n.e3b, 256, 0, 256}, {n.gB, 256, 0, 256}, {n.gA, 256, 0, 256}, false,
                  N, H3, W3));
  PT_TRY(conv_bwd(n, 6, {n.p3, 128, 0, 128}, {n.e3a, 256, 0, 256}, {n.gA, 256, 0, 256}, {n.gB, 128, 0, 128}, false,
                  N, H3, W3));   // d(p3)
  // pool3 backward into the skip slice of cat2 (accumulating onto the decoder's share)
  hipLaunchKernelGGL(k_pool_bwd, dim3(nblk(P2 * 128)), dim3(256), 0, n.stream, n.gB, n.gcat2, 256, 128, 128, N, H2,
                     W2);
  PT_TRY(conv_bwd(n, 5, {n.e2a, 128, 0, 128}, {n.cat2, 256, 128, 128}, {n.gcat2, 256, 128, 128}, {n.gA, 128, 0, 128},
                  false, N, H2, W2));
  PT_TRY(conv_bwd(n, 4, {n.p2, 64, 0, 64}, {n.e2a, 128, 0, 128}, {n.gA, 128, 0, 128}, {n.gB, 64, 0, 64}, false, N, H2,
                  W2));
  hipLaunchKernelGGL(k_pool_bwd, dim3(nblk(P1 * 64)), dim3(256), 0, n.stream, n.gB, n.gcat1, 128, 64, 64, N, H1, W1);
  PT_TRY(conv_bwd(n, 3, {n.e1a, 64, 0, 64}, {n.cat1, 128, 64, 64}, {n.gcat1, 128, 64, 64}, {n.gA, 64, 0, 64}, false, N,
                  H1, W1));
  PT_TRY(conv_bwd(n, 2, {n.p1, 32, 0, 32}, {n.e1a, 64, 0, 64}, {n.gA, 64, 0, 64}, {n.gB, 32, 0, 32}, false, N, H1, W1));
  hipLaunchKernelGGL(k_pool_bwd, dim3(nblk(P0 * 32)), dim3(256), 0, n.stream, n.gB, n.gcat0, 64, 32, 32, N, H, W);
  PT_TRY(conv_bwd(n, 1, {n.e0a, 32, 0, 32}, {n.cat0, 64, 32, 32}, {n.gcat0, 64, 32, 32}, {n.gA, 32, 0, 32}, false, N, H,
                  W));
  PT_TRY(conv_bwd(n, 0, {n.x0, IN_CH, 0, IN_CH}, {n.e0a, 32, 0, 32}, {n.gA, 32, 0, 32}, {nullptr, 0, 0, 0}, false, N,
                  H, W));
  if (loss) {
    PT_HIP(hipMemcpyAsync(n.h_lout, n.lout, 2 * sizeof(double), hipMemcpyDeviceToHost, n.stream));
    PT_HIP(hipStreamSynchronize(n.stream));
    *loss = n.h_lout[0];
  }
  return PT_OK;
}

static int adamw_step(Net& n, double lr, double wd) {
  n.step++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)n.step), c2 = 1.f - std::pow(b2, (float)n.step);
  hipLaunchKernelGGL(k_adamw, dim3(nblk(n.nparam)), dim3(256), 0, n.stream, n.w, n.g, n.m, n.v, n.wb, n.nparam,
                     (float)lr, b1, b2, 1e-8f, (float)wd, c1, c2);
  return PT_OK;
}

static int net_create(int seed, Net** out) {
  Net* n = new Net();
  *out = n;   // freed by the caller on failure
  long off = 0;
  for (int l = 0; l < 17; l++) {
    n->conv[l].Cin = kConvDims[l][0];
    n->conv[l].Cout = kConvDims[l][1];
    n->conv[l].w = off; off += 9L * kConvDims[l][0] * kConvDims[l][1];
    n->conv[l].b = off; off += kConvDims[l][1];
  }
  n->head_w = off; off += 32;
  n->head_b = off; off += 1;
  n->nparam = off;
  // He-normal initialisation of the conv weights (fan-in 9 Cin_used), zero
  // biases; the head starts at zero so the first guess is the warm start
  std::vector<float> h(off, 0.f);
  std::mt19937 rng((uint32_t)seed);
  for (int l = 0; l < 17; l++) {
    const int cin_used = l == 0 ? NUSED_IN : kConvDims[l][0];
    std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (9.f * cin_used)));
    for (long t = 0; t < 9L * kConvDims[l][1] * kConvDims[l][0]; t++) {
      const int ci = (int)(t % kConvDims[l][0]);
      h[n->conv[l].w + t] = (l == 0 && ci >= NUSED_IN) ? 0.f : nd(rng);
    }
  }
  std::normal_distribution<float> nh(0.f, 1e-3f);
  for (int c = 0; c < 32; c++) h[n->head_w + c] = nh(rng);
  auto fail = [&](hipError_t e) {
    set_error(std::string("net allocation: ") + hipGetErrorString(e));
    return PT_ERR_HIP;
  };
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking)) != hipSuccess) return fail(e);
  float** f[] = {&n->w, &n->g, &n->m, &n->v};
  for (float** p : f) {
    if ((e = hipMalloc((void**)p, sizeof(float) * off)) != hipSuccess) return fail(e);
    if ((e = hipMemset(*p, 0, sizeof(float) * off)) != hipSuccess) return fail(e);
  }
  if ((e = hipMalloc((void**)&n->wb, sizeof(uint16_t) * off)) != hipSuccess) return fail(e);
  if ((e = hipMalloc((void**)&n->wflip, sizeof(uint16_t) * 9 * 256 * 256)) != hipSuccess) return fail(e);
  if ((e = hipMalloc((void**)&n->lpart, sizeof(double) * 3 * 256)) != hipSuccess) return fail(e);
  if ((e = hipMalloc((void**)&n->lout, sizeof(double) * 2)) != hipSuccess) return fail(e);
  if ((e = hipHostMalloc((void**)&n->h_lout, sizeof(double) * 2, hipHostMallocDefault)) != hipSuccess) return fail(e);
  if ((e = hipMemcpy(n->w, h.data(), sizeof(float) * off, hipMemcpyHostToDevice)) != hipSuccess) return fail(e);
  hipLaunchKernelGGL(k_to_bf16, dim3(nblk(off)), dim3(256), 0, n->stream, n->w, n->wb, off);
  if ((e = hipStreamSynchronize(n->stream)) != hipSuccess) return fail(e);
  return PT_OK;
}

static void net_free(Net* n) {
  if (!n) return;
  if (n->stream) (void)hipStreamSynchronize(n->stream);
  for (void* p : n->bufs) (void)hipFree(p);
  void* ps[] = {n->w, n->g, n->m, n->v, n->wb, n->wflip, n->part, n->lpart, n->lout, n->dx, n->dt};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  if (n->h_lout) (void)hipHostFree(n->h_lout);
  if (n->stream) (void)hipStreamDestroy(n->stream);
  delete n;
}

// host (N, H, W, 5) FP32 inputs -> n.x0 (BF16, 8 channels)
static int upload_inputs(Net& n, const float* x) {
  const long P = npix(n);
  std::vector<uint16_t> h((size_t)P * IN_CH, 0);
  for (long p = 0; p < P; p++)
    for (int c = 0; c < NUSED_IN; c++) h[p * IN_CH + c] = f2bf(x[p * NUSED_IN + c]);
  PT_HIP(hipMemcpyAsync(n.x0, h.data(), sizeof(uint16_t) * h.size(), hipMemcpyHostToDevice, n.stream));
  PT_HIP(hipStreamSynchronize(n.stream));
  return PT_OK;
}

// ---------------------------------------------------------------------------
// the network inside a solve
// ---------------------------------------------------------------------------
#define G3(i, k, p) ((long)((k) + 3) * (p) + ((i) + 3))
#define G5(i, k, p) ((long)((k) + 5) * (p) + ((i) + 5))

// input image of one solve, pixel (i, k) = node (i, k), 0 <= i < nx, 0 <= k < nz:
// Tbuoy, Cbuoy, ln(eta), psi / rms(psi), residual / rms(residual) (zero on the walls)
__global__ void k_guess_inputs(const double* Tb, const double* Cb, const double* vis, const double* SF,
                               const double* res, double rpsi, double rres, int nx, int nz, uint16_t* x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= nx || k >= nz) return;
  const int p3 = nx + 7, p5 = nx + 11;
  uint16_t* o = x + ((long)k * nx + i) * IN_CH;
  o[0] = f2bf((float)Tb[G3(i, k, p3)]);
  o[1] = f2bf((float)Cb[G5(i, k, p5)]);
  o[2] = f2bf((float)log(vis[G3(i, k, p3)]));
  o[3] = f2bf((float)(SF[G3(i, k, p3)] * rpsi));
  o[4] = f2bf((i >= 1 && k >= 1) ? (float)(res[G3(i, k, p3)] * rres) : 0.f);
  for (int c = NUSED_IN; c < IN_CH; c++) o[c] = 0;
}

// psi_out(i, k) = SF(i, k) + scale * y(k, i) at the interior nodes, with the
// ghost images of enforceBCs_stream (antisymmetric side walls, sz-symmetric
// top and bottom); psi_out holds a copy of SF beforehand (walls and ghosts)
__global__ void k_guess_apply(const double* SF, const float* y, double scale, double* out, int nx, int nz, double sz) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x, k = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i > nx - 1 || k > nz - 1) return;
  const int p = nx + 7;
  const double v = SF[G3(i, k, p)] + scale * (double)y[(long)k * nx + i];
  out[G3(i, k, p)] = v;
  const bool lo_i = i <= 3, hi_i = i >= nx - 3, lo_k = k <= 3, hi_k = k >= nz - 3;
  if (lo_i) out[G3(-i, k, p)] = -v;
  if (hi_i) out[G3(2 * nx - i, k, p)] = -v;
  if (lo_k) {
    out[G3(i, -k, p)] = sz * v;
    if (lo_i) out[G3(-i, -k, p)] = -(sz * v);
    if (hi_i) out[G3(2 * nx - i, -k, p)] = -(sz * v);
  }
  if (hi_k) {
    out[G3(i, 2 * nz - k, p)] = sz * v;
    if (lo_i) out[G3(-i, 2 * nz - k, p)] = -(sz * v);
    if (hi_i) out[G3(2 * nx - i, 2 * nz - k, p)] = -(sz * v);
  }
}

// training target of one solve: (psi_new - psi_old) / rms(residual), zero on the walls
__global__ void k_guess_target(const double* psi, const double* SF, double inv_sr, int nx, int nz, float* t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= nx || k >= nz) return;
  const int p = nx + 7;
  t[(long)k * nx + i] = (i >= 1 && k >= 1) ? (float)((psi[G3(i, k, p)] - SF[G3(i, k, p)]) * inv_sr) : 0.f;
}

__global__ void k_interior_mask(float* m, int N, int H, int W) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)N * H * W) return;
  const int i = (int)(t % W), k = (int)((t / W) % H);
  m[t] = (i >= 1 && k >= 1) ? 1.f : 0.f;
}
__global__ void k_scale_copy(const float* a, float* b, float s, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = s * a[t];
}

// 3x3 binomial smoothing of an H x W image (zero outside)
__global__ void k_smooth121(const float* a, float* b, int H, int W) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= W || k >= H) return;
  float s = 0.f;
  for (int dk = -1; dk <= 1; dk++)
    for (int di = -1; di <= 1; di++) {
      const int ii = i + di, kk = k + dk;
      if (ii < 0 || ii >= W || kk < 0 || kk >= H) continue;
      s += (float)((2 - di * di) * (2 - dk * dk)) * a[(long)kk * W + ii];
    }
  b[(long)k * W + i] = s * (1.f / 16.f);
}
// sums over the interior nodes of r0 * (r0 - r1) and (r0 - r1)^2: with
// r1 = r0 - L d these are <r0, L d> and <L d, L d>
__global__ void k_dot2(const double* r0, const double* r1, int nx, int nz, double* part) {
  __shared__ double s1[256], s2[256];
  double a = 0.0, b = 0.0;
  const long tot = (long)(nx - 1) * (nz - 1);
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < tot; e += (long)gridDim.x * 256) {
    const int i = 1 + (int)(e % (nx - 1)), k = 1 + (int)(e / (nx - 1));
    const long g = G3(i, k, nx + 7);
    const double ld = r0[g] - r1[g];
    a += r0[g] * ld;
    b += ld * ld;
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      s1[threadIdx.x] += s1[threadIdx.x + st];
      s2[threadIdx.x] += s2[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s1[0];
    part[2 * blockIdx.x + 1] = s2[0];
  }
}
// the same two sums over 8 x 8 node tiles' sums of the residuals (their
// smooth part): one thread per tile
__global__ void k_dot2_tiles(const double* r0, const double* r1, int nx, int nz, double* part) {
  __shared__ double s1[256], s2[256];
  const int tx = nx / 8, tz = nz / 8;
  double a = 0.0, b = 0.0;
  for (int t = blockIdx.x * 256 + threadIdx.x; t < tx * tz; t += gridDim.x * 256) {
    const int i0 = (t % tx) * 8, k0 = (t / tx) * 8;
    double u = 0.0, v = 0.0;
    for (int k = k0; k < k0 + 8; k++)
      for (int i = i0; i < i0 + 8; i++) {
        if (i < 1 || k < 1) continue;
        const long g = G3(i, k, nx + 7);
        u += r0[g];
        v += r0[g] - r1[g];
      }
    a += u * v;
    b += v * v;
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      s1[threadIdx.x] += s1[threadIdx.x + st];
      s2[threadIdx.x] += s2[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s1[0];
    part[2 * blockIdx.x + 1] = s2[0];
  }
}
__global__ void k_sum2_final(const double* part, int nb, double* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double a = 0.0, b = 0.0;
  for (int q = 0; q < nb; q++) {
    a += part[2 * q];
    b += part[2 * q + 1];
  }
  out[0] = a;
  out[1] = b;
}

int learned_before_solve(Context& c) {
  Net& n = *static_cast<Net*>(c.guess_net);
  c.lg_slot = -1;
  if (!c.guess_on && !c.capture) return PT_OK;
  const int nx = c.T.nx, nz = c.T.nz;
  if (nx % 8 || nz % 8) {
    set_error("learned guess: nx and nz must be multiples of 8");
    return PT_ERR_UNSUPPORTED;
  }
  Level& L1 = c.L[1];
  const size_t b1 = sizeof(double) * L1.nelem;
  if (!c.lg_res) {
    PT_HIP(hipMalloc((void**)&c.lg_res, b1));
    PT_HIP(hipMalloc((void**)&c.lg_res2, b1));
    PT_HIP(hipMalloc((void**)&c.lg_psi, b1));
    PT_HIP(hipMemset(c.lg_res, 0, b1));
    PT_HIP(hipMemset(c.lg_res2, 0, b1));
  }
  // the warm start's residual, its rms and psi's rms
  double l1w = 0.0, ssr = 0.0, ssp = 0.0;
  PT_TRY(level1_residual_of(c, L1.err, c.lg_res, &l1w));
  PT_TRY(reduce_to_host(c, c.lg_res, nx - 1, nz - 1, nx + 7, 4, RED_SUM_SQ, &ssr));
  PT_TRY(reduce_to_host(c, c.SF, nx - 1, nz - 1, nx + 7, 4, RED_SUM_SQ, &ssp));
  const double npts = (double)(nx - 1) * (double)(nz - 1);
  const double sr = std::sqrt(ssr / npts), sp = std::sqrt(ssp / npts);
  if (!(sr > 0.0)) return PT_OK;   // nothing to correct
  c.lg_sr = sr;
  c.lg_l1_warm = l1w;
  hipStream_t saved = n.stream;
  n.stream = c.stream;   // the network's kernels join the context's stream for this solve
  int rc = net_alloc_bufs(n, 1, nz, nx);
  if (rc != PT_OK) { n.stream = saved; return rc; }
  dim3 b16(64, 4), g((nx + 63) / 64, (nz + 3) / 4);
  hipLaunchKernelGGL(k_guess_inputs, g, b16, 0, c.stream, c.Tbuoy, c.Cbuoy, L1.v.vis, c.SF, c.lg_res,
                     sp > 0.0 ? 1.0 / sp : 0.0, 1.0 / sr, nx, nz, n.x0);
  const long img = (long)nx * nz;
  if (c.capture && n.dcap > 0 && n.dH == nz && n.dW == nx && n.dcount < n.dcap) {
    c.lg_slot = n.dcount++;
    PT_HIP(hipMemcpyAsync(n.dx + (long)c.lg_slot * img * IN_CH, n.x0, sizeof(uint16_t) * img * IN_CH,
                          hipMemcpyDeviceToDevice, c.stream));
  }
  if (c.guess_on) {
    forward(n);
    // the proposed correction d, smoothed (its high frequencies are what L
    // amplifies most and what the smoother removes cheapest)
    float* ya = n.y;
    float* yb = n.dy;
    for (int q = 0; q < c.guess_smooth; q++) {
      hipLaunchKernelGGL(k_smooth121, g, b16, 0, c.stream, ya, yb, nz, nx);
      std::swap(ya, yb);
    }
    const double sz = c.P.Vbc == 0 ? -1.0 : 1.0;
    const double scale = sr * (double)n.out_scale;
    dim3 gi((nx - 1 + 63) / 64, (nz - 1 + 3) / 4);
    // psi_w + d, its residual r1 = r0 - L d, and the step alpha along d that
    // minimises |r0 - alpha L d|_2
    PT_HIP(hipMemcpyAsync(c.lg_psi, c.SF, b1, hipMemcpyDeviceToDevice, c.stream));
    hipLaunchKernelGGL(k_guess_apply, gi, b16, 0, c.stream, c.SF, ya, scale, c.lg_psi, nx, nz, sz);
    double l1d = 0.0;
    rc = level1_residual_of(c, c.lg_psi, c.lg_res2, &l1d);
    if (rc != PT_OK) { n.stream = saved; return rc; }
    if (c.guess_judge == 1)
      hipLaunchKernelGGL(k_dot2_tiles, dim3(256), dim3(256), 0, c.stream, c.lg_res, c.lg_res2, nx, nz, n.lpart);
    else
      hipLaunchKernelGGL(k_dot2, dim3(256), dim3(256), 0, c.stream, c.lg_res, c.lg_res2, nx, nz, n.lpart);
    hipLaunchKernelGGL(k_sum2_final, dim3(1), dim3(64), 0, c.stream, n.lpart, 256, n.lout);
    PT_HIP(hipMemcpyAsync(n.h_lout, n.lout, 2 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    PT_HIP(hipStreamSynchronize(c.stream));
    const double alpha = c.guess_judge == 2 ? 1.0 : (n.h_lout[1] > 0.0 ? n.h_lout[0] / n.h_lout[1] : 0.0);
    c.lg_alpha = alpha;
    PT_HIP(hipMemcpyAsync(c.lg_psi, c.SF, b1, hipMemcpyDeviceToDevice, c.stream));
    hipLaunchKernelGGL(k_guess_apply, gi, b16, 0, c.stream, c.SF, ya, scale * alpha, c.lg_psi, nx, nz, sz);
    double l1g = 0.0;
    rc = level1_residual_of(c, c.lg_psi, c.lg_res2, &l1g);
    if (rc != PT_OK) { n.stream = saved; return rc; }
    c.guess_tried++;
    c.lg_l1_guess = l1g;
    // the plain solver judges: the proposal replaces the warm start only when
    // its FP64 L1 residual is lower
    const bool keep = c.guess_judge == 2 ? (l1g < 1e3 * l1w)
                      : c.guess_judge == 1 ? (alpha > 0.0 && l1g < 2.0 * l1w) : (l1g < l1w);
    if (keep) {
      PT_HIP(hipMemcpyAsync(L1.err, c.lg_psi, b1, hipMemcpyDeviceToDevice, c.stream));
      c.guess_used++;
    }
  }
  n.stream = saved;
  return PT_OK;
}

int learned_after_solve(Context& c) {
  if (c.lg_slot < 0) return PT_OK;
  Net& n = *static_cast<Net*>(c.guess_net);
  const int nx = c.T.nx, nz = c.T.nz;
  dim3 b16(64, 4), g((nx + 63) / 64, (nz + 3) / 4);
  hipLaunchKernelGGL(k_guess_target, g, b16, 0, c.stream, c.L[1].err, c.SF, 1.0 / c.lg_sr, nx, nz,
                     n.dt + (long)c.lg_slot * nx * nz);
  c.lg_slot = -1;
  return PT_OK;
}

// AdamW on random batches of the device dataset
static int net_train(Net& n, int iters, int batch, double lr, double wd, uint64_t seed, double* loss) {
  if (n.dcount < 1) {
    set_error("training: the dataset is empty");
    return PT_ERR_INVALID;
  }
  const long img = (long)n.dH * n.dW;
  // the stored targets' interior rms fixes the output scale at the first call
  if (!n.out_scale_set) {
    std::vector<float> h((size_t)n.dcount * img);
    PT_HIP(hipMemcpyAsync(h.data(), n.dt, sizeof(float) * h.size(), hipMemcpyDeviceToHost, n.stream));
    PT_HIP(hipStreamSynchronize(n.stream));
    double s = 0.0;
    long cnt = 0;
    for (int q = 0; q < n.dcount; q++) {
      double sq = 0.0;
      long cq = 0;
      for (long pix = 0; pix < img; pix++) {
        if (pix % n.dW == 0 || pix / n.dW == 0) continue;
        const double v = h[(size_t)q * img + pix];
        sq += v * v;
        cq++;
      }
      s += sq;
      cnt += cq;
    }
    n.out_scale = cnt > 0 && s > 0.0 ? (float)std::sqrt(s / cnt) : 1.f;
    n.out_scale_set = true;
  }
  const int N = std::min(batch, n.dcount);
  PT_TRY(net_alloc_bufs(n, N, n.dH, n.dW));
  if (n.mask_N != N || n.mask_H != n.dH || n.mask_W != n.dW) {
    hipLaunchKernelGGL(k_interior_mask, dim3(nblk((long)N * img)), dim3(256), 0, n.stream, n.msk, N, n.dH, n.dW);
    n.mask_N = N; n.mask_H = n.dH; n.mask_W = n.dW;
  }
  std::mt19937_64 rng(seed);
  std::vector<int> idx(n.dcount);
  for (int it = 0; it < iters; it++) {
    for (int q = 0; q < n.dcount; q++) idx[q] = q;
    for (int b = 0; b < N; b++) {   // a batch without repetition
      std::uniform_int_distribution<int> ud(b, n.dcount - 1);
      std::swap(idx[b], idx[ud(rng)]);
      PT_HIP(hipMemcpyAsync(n.x0 + (long)b * img * IN_CH, n.dx + (long)idx[b] * img * IN_CH,
                            sizeof(uint16_t) * img * IN_CH, hipMemcpyDeviceToDevice, n.stream));
      hipLaunchKernelGGL(k_scale_copy, dim3(nblk(img)), dim3(256), 0, n.stream, n.dt + (long)idx[b] * img,
                         n.tgt + (long)b * img, 1.f / n.out_scale, img);
    }
    forward(n);
    PT_TRY(loss_backward(n, it == iters - 1 ? loss : nullptr));
    PT_TRY(adamw_step(n, lr, wd));
  }
  PT_HIP(hipStreamSynchronize(n.stream));
  return PT_OK;
}

}  // namespace pt

using namespace pt;

extern "C" {

int pt_net_dataset(void* net, int32_t H, int32_t W, int32_t capacity) {
  Net* n = static_cast<Net*>(net);
  if (!n || capacity < 1 || H < 8 || W < 8 || H % 8 || W % 8) {
    set_error("pt_net_dataset: capacity >= 1 and H, W multiples of 8 required");
    return PT_ERR_INVALID;
  }
  if (n->dx) (void)hipFree(n->dx);
  if (n->dt) (void)hipFree(n->dt);
  n->dx = nullptr; n->dt = nullptr;
  const long img = (long)H * W;
  PT_HIP(hipMalloc((void**)&n->dx, sizeof(uint16_t) * img * IN_CH * capacity));
  PT_HIP(hipMalloc((void**)&n->dt, sizeof(float) * img * capacity));
  n->dH = H; n->dW = W; n->dcap = capacity; n->dcount = 0;
  return PT_OK;
}

int pt_net_dataset_size(void* net, int32_t* count) {
  Net* n = static_cast<Net*>(net);
  if (!n || !count) return PT_ERR_INVALID;
  *count = n->dcount;
  return PT_OK;
}

int pt_net_train(void* net, int32_t iters, int32_t batch, double lr, double weight_decay, uint64_t seed,
                 double* loss) {
  Net* n = static_cast<Net*>(net);
  if (!n || iters < 1 || batch < 1 || !loss) return PT_ERR_INVALID;
  return net_train(*n, iters, batch, lr, weight_decay, seed, loss);
}

int pt_net_dataset_get(void* net, int32_t slot, float* x, float* t) {
  Net* n = static_cast<Net*>(net);
  if (!n || !x || !t || slot < 0 || slot >= n->dcount) return PT_ERR_INVALID;
  const long img = (long)n->dH * n->dW;
  std::vector<uint16_t> hx((size_t)img * IN_CH);
  PT_HIP(hipMemcpyAsync(hx.data(), n->dx + (long)slot * img * IN_CH, sizeof(uint16_t) * hx.size(),
                        hipMemcpyDeviceToHost, n->stream));
  PT_HIP(hipMemcpyAsync(t, n->dt + (long)slot * img, sizeof(float) * img, hipMemcpyDeviceToHost, n->stream));
  PT_HIP(hipStreamSynchronize(n->stream));
  for (long p = 0; p < img; p++)
    for (int c = 0; c < NUSED_IN; c++) x[p * NUSED_IN + c] = bf2f(hx[p * IN_CH + c]);
  return PT_OK;
}

int pt_net_set_loss(void* net, double lap_weight) {
  Net* n = static_cast<Net*>(net);
  if (!n || !(lap_weight >= 0.0)) return PT_ERR_INVALID;
  n->lap_weight = lap_weight;
  return PT_OK;
}

int pt_net_out_scale(void* net, double* scale) {
  Net* n = static_cast<Net*>(net);
  if (!n || !scale) return PT_ERR_INVALID;
  *scale = n->out_scale;
  return PT_OK;
}

int pt_net_create(int32_t seed, void** net) {
  if (!net) { set_error("null argument"); return PT_ERR_INVALID; }
  *net = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    set_error("no HIP device visible");
    return PT_ERR_NO_DEVICE;
  }
  Net* n = nullptr;
  const int rc = net_create(seed, &n);
  if (rc != PT_OK) { net_free(n); return rc; }
  *net = n;
  return PT_OK;
}

int pt_net_destroy(void* net) {
  net_free(static_cast<Net*>(net));
  return PT_OK;
}

int pt_net_num_params(void* net, int64_t* count) {
  Net* n = static_cast<Net*>(net);
  if (!n || !count) return PT_ERR_INVALID;
  *count = n->nparam;
  return PT_OK;
}

int pt_net_get_params(void* net, float* host) {
  Net* n = static_cast<Net*>(net);
  if (!n || !host) return PT_ERR_INVALID;
  PT_HIP(hipMemcpyAsync(host, n->w, sizeof(float) * n->nparam, hipMemcpyDeviceToHost, n->stream));
  PT_HIP(hipStreamSynchronize(n->stream));
  return PT_OK;
}

int pt_net_set_params(void* net, const float* host) {
  Net* n = static_cast<Net*>(net);
  if (!n || !host) return PT_ERR_INVALID;
  PT_HIP(hipMemcpyAsync(n->w, host, sizeof(float) * n->nparam, hipMemcpyHostToDevice, n->stream));
  hipLaunchKernelGGL(k_to_bf16, dim3(nblk(n->nparam)), dim3(256), 0, n->stream, n->w, n->wb, n->nparam);
  PT_HIP(hipStreamSynchronize(n->stream));
  return PT_OK;
}

static int check_shape(int N, int H, int W) {
  if (N < 1 || H < 8 || W < 8 || H % 8 || W % 8) {
    set_error("network input: N >= 1 and H, W multiples of 8 (>= 8) required");
    return PT_ERR_INVALID;
  }
  return PT_OK;
}

int pt_net_forward(void* net, int32_t N, int32_t H, int32_t W, const float* x, float* y) {
  Net* n = static_cast<Net*>(net);
  if (!n || !x || !y) return PT_ERR_INVALID;
  PT_TRY(check_shape(N, H, W));
  PT_TRY(net_alloc_bufs(*n, N, H, W));
  PT_TRY(upload_inputs(*n, x));
  forward(*n);
  PT_HIP(hipMemcpyAsync(y, n->y, sizeof(float) * npix(*n), hipMemcpyDeviceToHost, n->stream));
  PT_HIP(hipStreamSynchronize(n->stream));
  return PT_OK;
}

int pt_net_loss_grad(void* net, int32_t N, int32_t H, int32_t W, const float* x, const float* t, const float* mask,
                     double* loss, float* grad) {
  Net* n = static_cast<Net*>(net);
  if (!n || !x || !t || !mask || !loss || !grad) return PT_ERR_INVALID;
  PT_TRY(check_shape(N, H, W));
  PT_TRY(net_alloc_bufs(*n, N, H, W));
  PT_TRY(upload_inputs(*n, x));
  const long P = npix(*n);
  PT_HIP(hipMemcpyAsync(n->tgt, t, sizeof(float) * P, hipMemcpyHostToDevice, n->stream));
  PT_HIP(hipMemcpyAsync(n->msk, mask, sizeof(float) * P, hipMemcpyHostToDevice, n->stream));
  forward(*n);
  PT_TRY(loss_backward(*n, loss));
  PT_HIP(hipMemcpyAsync(grad, n->g, sizeof(float) * n->nparam, hipMemcpyDeviceToHost, n->stream));
  PT_HIP(hipStreamSynchronize(n->stream));
  return PT_OK;
}

int pt_net_adamw_step(void* net, double lr, double wd) {
  Net* n = static_cast<Net*>(net);
  if (!n) return PT_ERR_INVALID;
  PT_TRY(adamw_step(*n, lr, wd));
  PT_HIP(hipStreamSynchronize(n->stream));
  return PT_OK;
}

}  // extern "C"
