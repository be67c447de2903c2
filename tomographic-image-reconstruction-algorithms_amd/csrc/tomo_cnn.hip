// tomo_cnn.hip -- the learned preconditioner: a 5-layer 3x3 convolutional
// network (3 -> 32 -> 32 -> 32 -> 32 -> 1 channels, ReLU after layers 1-4)
// that maps (field, SIRT update direction, sensor coverage) to a per-pixel
// field update; its backward pass; Adam training on device-resident samples
// taken from plain SIRT trajectories; and the learned field update of an
// engine context.  The reference has no learned component, so nothing here
// has a reference counterpart.
//
// Arithmetic: activations, weights and back-propagated gradients in bfloat16,
// products on the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16) with FP32
// accumulation; biases, the output layer, the loss and Adam state in FP32;
// cross-wave sums of the weight gradient and the loss in FP64 in a fixed
// order (deterministic).
//
// Layouts (pixel-major, channels fastest, "NHWC"):
//   input      bf16 [B][H][W][8]    channels: field, direction, coverage, 0...
//   hidden     bf16 [B][H][W][32]
//   output     f32  [B][H][W]
//   weights of layer l for the forward product, bf16 [Cout16][Kpad]
//     with k = tap * Cin8 + c (tap = 3 ky + kx), zero beyond the real ranges;
//   weights for the backward data product, bf16 [Cin of l][Kpad'] with
//     k' = tap' * Cout8 + o holding W[o][c][8 - tap'] (mirrored taps).
//   Parameters (host ABI, FP32): per layer W[o][c][ky][kx] then b[o].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "tomo.h"
#include "tomo_ctx.h"
#include "tomo_kernels.h"

using namespace tomo;
using namespace tomo_internal;

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kLayers = 5;
constexpr int kCin[kLayers] = {3, 32, 32, 32, 32};
constexpr int kCout[kLayers] = {32, 32, 32, 32, 1};
// storage channel counts: input and the 1-channel gradient padded to 8
constexpr int kCinS[kLayers] = {8, 32, 32, 32, 32};
constexpr int kCoutS[kLayers] = {32, 32, 32, 32, 8};

constexpr int param_offset(int l) {
  int off = 0;
  for (int i = 0; i < l; ++i) off += kCout[i] * kCin[i] * 9 + kCout[i];
  return off;
}
constexpr int kParams = param_offset(kLayers);  // 28929

constexpr int kpad(int cin_s) { return ((9 * cin_s + 31) / 32) * 32; }

// Input normalisation: field, plain SIRT step and output update in units of
// kFieldUnit, coverage by its maximum.
constexpr float kFieldUnit = 0.1f;

__device__ __forceinline__ __bf16 to_bf16(float v) { return (__bf16)v; }

// --------------------------------------------------------------- conv kernel
// out[p][m] = sum_{tap, c} Wf[m][tap * CIN + c] * in[p + d(tap)][c] as the
// product D = A * B with A = weights (M = output channels), B = input
// patches (N = 16 consecutive pixels of a row), K = (tap, c) in blocks of 32.
// mfma_f32_16x16x32_bf16 operand layouts (lane l, element e):
//   A[i][k]: i = l % 16, k = 8 (l / 16) + e;  B[k][j]: j = l % 16, k = 8 (l / 16) + e;
//   D[i][j]: j = l % 16, i = 4 (l / 16) + r  (r = accumulator register).
enum ConvEpilogue {
  kEpiBiasRelu = 0,  // + bias, ReLU, bf16 [p][COUT]
  kEpiBiasOut = 1,   // + bias, channel 0 only, f32 [p]
  kEpiMask = 2,      // * (gate[p][m] > 0), bf16 [p][COUT]  (backward data)
};

struct ConvArgs {
  const __bf16 *in;    // [B][H][W][CIN]
  const __bf16 *w;     // [COUT16][KPAD]
  const float *bias;   // [COUT] (forward epilogues)
  const __bf16 *gate;  // [B][H][W][COUT] (mask epilogue)
  void *out;
  int B, H, W;
  int groups;          // B * H * W / 16
};

template <int CIN, int COUT, int EPI>
__global__ __launch_bounds__(256) void conv3x3_kernel(ConvArgs a) {
  constexpr int KP = kpad(CIN);
  constexpr int KB = KP / 32;
  constexpr int MB = (COUT + 15) / 16;
  const int lane = threadIdx.x & 63;
  const int q = lane >> 4, col = lane & 15;
  const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int nwaves = gridDim.x * (blockDim.x >> 6);

  bf16x8 A[MB][KB];
#pragma unroll
  for (int mb = 0; mb < MB; ++mb)
#pragma unroll
    for (int kb = 0; kb < KB; ++kb)
      A[mb][kb] = *reinterpret_cast<const bf16x8 *>(a.w + (size_t)(mb * 16 + col) * KP + kb * 32 + 8 * q);

  const int gpr = a.W >> 4;  // groups per row
  for (int g = wave; g < a.groups; g += nwaves) {
    const int x0 = (g % gpr) * 16;
    const int rowi = g / gpr;  // b * H + y
    const int y = rowi % a.H;
    const int px = x0 + col;
    f32x4 acc[MB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) acc[mb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
      const int k0 = kb * 32 + 8 * q;
      const int tap = k0 / CIN, c0 = k0 % CIN;
      bf16x8 Bv;
#pragma unroll
      for (int e = 0; e < 8; ++e) Bv[e] = to_bf16(0.f);
      if (tap < 9) {
        const int yy = y + tap / 3 - 1, xx = px + tap % 3 - 1;
        if (yy >= 0 && yy < a.H && xx >= 0 && xx < a.W)
          Bv = *reinterpret_cast<const bf16x8 *>(a.in + ((size_t)(rowi + (yy - y)) * a.W + xx) * CIN + c0);
      }
#pragma unroll
      for (int mb = 0; mb < MB; ++mb)
        acc[mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[mb][kb], Bv, acc[mb], 0, 0, 0);
    }
    const size_t pix = (size_t)rowi * a.W + px;
    if (EPI == kEpiBiasOut) {
      if (q == 0) reinterpret_cast<float *>(a.out)[pix] = acc[0][0] + a.bias[0];
    } else {
#pragma unroll
      for (int mb = 0; mb < MB; ++mb) {
        const int m0 = mb * 16 + 4 * q;
        if (m0 >= COUT) continue;
        bf16x4 v;
        if (EPI == kEpiBiasRelu) {
#pragma unroll
          for (int r = 0; r < 4; ++r) v[r] = to_bf16(fmaxf(acc[mb][r] + a.bias[m0 + r], 0.f));
        } else {
          const bf16x4 gt = *reinterpret_cast<const bf16x4 *>(a.gate + pix * COUT + m0);
#pragma unroll
          for (int r = 0; r < 4; ++r) v[r] = to_bf16((float)gt[r] > 0.f ? acc[mb][r] : 0.f);
        }
        *reinterpret_cast<bf16x4 *>(reinterpret_cast<__bf16 *>(a.out) + pix * COUT + m0) = v;
      }
    }
  }
}

// ------------------------------------------------------ weight-gradient kernel
// dW[o][tap * CIN + c] = sum_p d[p][o] * in[p + d(tap)][c] and db[o] =
// sum_p d[p][o] as D = A * B with A = d^T (M = o), B = patches (N = (tap, c)
// plus one column of ones for the bias), K = 32 consecutive pixels of a row.
// Each wave stages its 32 pixels (gradient transposed, input rows y-1..y+1
// channel-major) in LDS, accumulates over its chunks in registers and writes
// one FP32 partial; a second kernel sums the partials in FP64 in wave order.
constexpr int kWgWaves = 4;

template <int CIN, int COUT>
struct WgradShape {
  static constexpr int NREAL = 9 * CIN;              // (tap, c) columns
  static constexpr int NB = (NREAL + 1 + 15) / 16;  // + bias column
  static constexpr int MB = (COUT + 15) / 16;
  static constexpr int NCOL = NB * 16;
  static constexpr int MROW = MB * 16;
  static constexpr int XS = 36;  // staged pixels per input row (34 used)
};

struct WgradArgs {
  const __bf16 *in;  // [B][H][W][CIN]
  const __bf16 *d;   // [B][H][W][COUT]
  float *part;       // [nwaves][MROW][NCOL]
  int B, H, W;
  int chunks;        // B * H * W / 32
};

template <int CIN, int COUT>
__global__ __launch_bounds__(256) void wgrad_kernel(WgradArgs a) {
  using S = WgradShape<CIN, COUT>;
  __shared__ __bf16 sd[kWgWaves][S::MROW][32];
  __shared__ __bf16 sx[kWgWaves][CIN][3][S::XS];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int q = lane >> 4, col = lane & 15;
  const int wave = blockIdx.x * kWgWaves + wv;
  const int nwaves = gridDim.x * kWgWaves;
  const int cpr = a.W >> 5;  // chunks per row

  f32x4 acc[S::MB][S::NB];
#pragma unroll
  for (int mb = 0; mb < S::MB; ++mb)
#pragma unroll
    for (int nb = 0; nb < S::NB; ++nb) acc[mb][nb] = f32x4{0.f, 0.f, 0.f, 0.f};

  // every wave of the block runs the same number of rounds (barriers)
  const int rounds = (a.chunks + nwaves - 1) / nwaves;
  for (int it = 0; it < rounds; ++it) {
    const int chunk = wave + it * nwaves;
    const bool valid = chunk < a.chunks;
    if (valid) {
      const int x0 = (chunk % cpr) * 32;
      const int rowi = chunk / cpr;
      const int y = rowi % a.H;
      // gradient, transposed: sd[o][p]
      for (int idx = lane; idx < 32 * (COUT / 8); idx += 64) {
        const int p = idx / (COUT / 8), cb = idx % (COUT / 8);
        const bf16x8 v = *reinterpret_cast<const bf16x8 *>(a.d + ((size_t)rowi * a.W + x0 + p) * COUT + cb * 8);
#pragma unroll
        for (int e = 0; e < 8; ++e) sd[wv][cb * 8 + e][p] = v[e];
      }
      for (int idx = lane; idx < (S::MROW - COUT) * 32; idx += 64)
        sd[wv][COUT + idx / 32][idx % 32] = to_bf16(0.f);
      // input rows y-1..y+1, pixels x0-1..x0+32, channel-major: sx[c][r][p+1]
      for (int idx = lane; idx < 3 * 34 * (CIN / 8); idx += 64) {
        const int cb = idx % (CIN / 8), rp = idx / (CIN / 8);
        const int r = rp / 34, pp = rp % 34;
        const int yy = y + r - 1, xx = x0 + pp - 1;
        bf16x8 v;
        if (yy >= 0 && yy < a.H && xx >= 0 && xx < a.W) {
          v = *reinterpret_cast<const bf16x8 *>(a.in + ((size_t)(rowi + r - 1) * a.W + xx) * CIN + cb * 8);
        } else {
#pragma unroll
          for (int e = 0; e < 8; ++e) v[e] = to_bf16(0.f);
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) sx[wv][cb * 8 + e][r][pp] = v[e];
      }
    }
    __syncthreads();
    if (valid) {
      bf16x8 Av[S::MB];
#pragma unroll
      for (int mb = 0; mb < S::MB; ++mb) Av[mb] = *reinterpret_cast<const bf16x8 *>(&sd[wv][mb * 16 + col][8 * q]);
#pragma unroll
      for (int nb = 0; nb < S::NB; ++nb) {
        const int n = nb * 16 + col;
        bf16x8 Bv;
        if (n < S::NREAL) {
          const int tap = n / CIN, c = n % CIN;
          const int r = tap / 3, dx = tap % 3 - 1;
#pragma unroll
          for (int e = 0; e < 8; ++e) Bv[e] = sx[wv][c][r][8 * q + e + dx + 1];
        } else {
          const float one = n == S::NREAL ? 1.f : 0.f;
#pragma unroll
          for (int e = 0; e < 8; ++e) Bv[e] = to_bf16(one);
        }
#pragma unroll
        for (int mb = 0; mb < S::MB; ++mb)
          acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Av[mb], Bv, acc[mb][nb], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  float *dst = a.part + (size_t)wave * S::MROW * S::NCOL;
#pragma unroll
  for (int mb = 0; mb < S::MB; ++mb)
#pragma unroll
    for (int nb = 0; nb < S::NB; ++nb)
#pragma unroll
      for (int r = 0; r < 4; ++r) dst[(size_t)(mb * 16 + 4 * q + r) * S::NCOL + nb * 16 + col] = acc[mb][nb][r];
}

// grad (canonical layout of layer l) = sum over waves of the partials, FP64,
// in a fixed order: kWredGroups consecutive runs of waves are summed in
// parallel (pass 1, into red[group][i]), then the runs in order (pass 2)
constexpr int kWredGroups = 32;

template <int CIN, int COUT>
__device__ __forceinline__ size_t wgrad_part_index(int i, int cin_real, int cout_real) {
  using S = WgradShape<CIN, COUT>;
  const int nw = cout_real * cin_real * 9;
  if (i < nw) {
    const int o = i / (cin_real * 9), rem = i % (cin_real * 9);
    const int c = rem / 9, tap = rem % 9;
    return (size_t)o * S::NCOL + tap * CIN + c;
  }
  return (size_t)(i - nw) * S::NCOL + S::NREAL;
}

template <int CIN, int COUT>
__global__ void wgrad_reduce_kernel(const float *part, int nwaves, int cin_real, int cout_real, double *red) {
  using S = WgradShape<CIN, COUT>;
  const int total = cout_real * cin_real * 9 + cout_real;
  const int i = blockIdx.x * blockDim.x + threadIdx.x, g = blockIdx.y;
  if (i >= total) return;
  const size_t at = wgrad_part_index<CIN, COUT>(i, cin_real, cout_real);
  const int run = (nwaves + kWredGroups - 1) / kWredGroups;
  const int w1 = min(nwaves, (g + 1) * run);
  double s = 0.0;
  for (int w = g * run; w < w1; ++w) s += (double)part[(size_t)w * S::MROW * S::NCOL + at];
  red[(size_t)g * total + i] = s;
}

__global__ void wgrad_finish_kernel(const double *red, int total, float *grad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  double s = 0.0;
  for (int g = 0; g < kWredGroups; ++g) s += red[(size_t)g * total + i];
  grad[i] = (float)s;
}

// ------------------------------------------------------------ small kernels
// x[p][0..2] (f32) -> bf16 [p][8]
__global__ void pack_input_kernel(const float *x3, __bf16 *x8, size_t npix) {
  const size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (p >= npix) return;
  bf16x8 v;
#pragma unroll
  for (int e = 0; e < 8; ++e) v[e] = to_bf16(e < 3 ? x3[p * 3 + e] : 0.f);
  *reinterpret_cast<bf16x8 *>(x8 + p * 8) = v;
}

// dy (f32 [p]) -> bf16 [p][8] with channel 0 set
__global__ void pack_dy_kernel(const float *dy, __bf16 *d8, size_t npix) {
  const size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (p >= npix) return;
  bf16x8 v;
#pragma unroll
  for (int e = 0; e < 8; ++e) v[e] = to_bf16(e == 0 ? dy[p] : 0.f);
  *reinterpret_cast<bf16x8 *>(d8 + p * 8) = v;
}

// masked MSE: d8[p][0] = m (y - t) / count (bf16), loss partials (FP64)
__global__ void loss_kernel(const float *y, const float *t, __bf16 *d8, int B, int H, int W, int mx0, int mx1,
                            int my0, int my1, float count, double *part) {
  __shared__ double red[256];
  const size_t npix = (size_t)B * H * W;
  double s = 0.0;
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < npix; p += (size_t)gridDim.x * blockDim.x) {
    const int x = (int)(p % W), yy = (int)((p / W) % H);
    float g = 0.f;
    if (x >= mx0 && x < mx1 && yy >= my0 && yy < my1) {
      const float r = y[p] - t[p];
      s += (double)r * (double)r;
      g = r / count;
    }
    bf16x8 v;
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = to_bf16(e == 0 ? g : 0.f);
    *reinterpret_cast<bf16x8 *>(d8 + p * 8) = v;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void sum_kernel(const double *part, int n, double scale, double *out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += part[i];
    *out = s * scale;
  }
}

// canonical FP32 parameters -> forward and backward-data weight layouts
struct PackArgs {
  const float *params;
  __bf16 *wf[kLayers];
  __bf16 *wb[kLayers];
  float *bias[kLayers];
};

__global__ void pack_weights_kernel(PackArgs a) {
  const int l = blockIdx.y;
  const int cin = kCin[l], cout = kCout[l], cin_s = kCinS[l], cout_s = kCoutS[l];
  const int kpf = kpad(cin_s), kpb = kpad(cout_s);
  const int mrow_f = ((cout + 15) / 16) * 16;
  const float *W = a.params + param_offset(l);
  const int nf = mrow_f * kpf;
  const int nb = cin * kpb;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nf + nb + cout; i += gridDim.x * blockDim.x) {
    if (i < nf) {
      const int m = i / kpf, k = i % kpf;
      const int tap = k / cin_s, c = k % cin_s;
      float v = 0.f;
      if (m < cout && tap < 9 && c < cin) v = W[(m * cin + c) * 9 + tap];
      a.wf[l][i] = to_bf16(v);
    } else if (i < nf + nb) {
      const int j = i - nf;
      const int c = j / kpb, k = j % kpb;
      const int tap = k / cout_s, o = k % cout_s;
      float v = 0.f;
      if (tap < 9 && o < cout) v = W[(o * cin + c) * 9 + (8 - tap)];
      a.wb[l][j] = to_bf16(v);
    } else {
      const int o = i - nf - nb;
      a.bias[l][o] = W[cout * cin * 9 + o];
    }
  }
}

__global__ void adam_kernel(float *p, const float *g, float *m, float *v, int n, float lr, float b1, float b2,
                            float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---------------------------------------------------- engine-facing kernels
// direction image d = mask * (sum_s G_s) / (1 + f) / S (FP64 sum over shots)
__global__ void direction_kernel(const float *f, const float *G, int S, int nx, int ny, Mask m, float *dir) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nx * ny) return;
  const int x = i % nx, y = i / nx;
  float d = 0.f;
  if (x >= m.x0 && x < m.x1 && y >= m.y0 && y < m.y1) {
    double s = 0.0;
    for (int g = 0; g < S; ++g) s += (double)G[(size_t)g * nx * ny + i];
    d = (float)(s / (1.0 + (double)f[i]) / (double)S);
  }
  dir[i] = d;
}

// coverage image summed over shots (FP64)
__global__ void coverage_sum_kernel(const float *cov, int S, long long n2, float *out) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n2) return;
  double s = 0.0;
  for (int g = 0; g < S; ++g) s += (double)cov[(size_t)g * n2 + i];
  out[i] = (float)s;
}

// stats[0] = sum_mask dir^2, stats[1] = max cov, stats[2] = mask cells
__global__ void input_stats_kernel(const float *dir, const float *cov, int nx, int ny, Mask m, double *stats) {
  __shared__ double s2[256];
  __shared__ float mx[256];
  double s = 0.0;
  float c = 0.f;
  for (int i = threadIdx.x; i < nx * ny; i += blockDim.x) {
    const int x = i % nx, y = i / nx;
    if (x >= m.x0 && x < m.x1 && y >= m.y0 && y < m.y1) s += (double)dir[i] * (double)dir[i];
    c = fmaxf(c, cov ? cov[i] : 0.f);
  }
  s2[threadIdx.x] = s;
  mx[threadIdx.x] = c;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      s2[threadIdx.x] += s2[threadIdx.x + o];
      mx[threadIdx.x] = fmaxf(mx[threadIdx.x], mx[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    stats[0] = s2[0];
    stats[1] = (double)mx[0];
    stats[2] = (double)(m.x1 - m.x0) * (double)(m.y1 - m.y0);
  }
}

// network input [p][8] bf16 from field, plain SIRT step (scale * direction)
// and coverage, field and step in units of kFieldUnit, coverage over its
// maximum; optional target (truth - field) / unit
__global__ void make_input_kernel(const float *f, const float *dir, float scale, const float *cov,
                                  const double *stats, int n2, __bf16 *x8, const float *fo, float *target) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n2) return;
  const float inv_c = stats[1] > 0.0 ? (float)(1.0 / stats[1]) : 0.f;
  bf16x8 v;
  const float c3[3] = {f[i] / kFieldUnit, scale * dir[i] / kFieldUnit, cov ? cov[i] * inv_c : 0.f};
#pragma unroll
  for (int e = 0; e < 8; ++e) v[e] = to_bf16(e < 3 ? c3[e] : 0.f);
  *reinterpret_cast<bf16x8 *>(x8 + (size_t)i * 8) = v;
  if (target) target[i] = (fo[i] - f[i]) / kFieldUnit;
}

// f += mask * (alpha * unit * y + max(1 - alpha, 0) * scale * dir): the
// learned step blended with the plain step (dir is already masked)
__global__ void learned_update_kernel(float *f, const float *y, const float *dir, int nx, int ny, Mask m,
                                      float alpha, float scale) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nx * ny) return;
  const int x = i % nx, yy = i / nx;
  if (x >= m.x0 && x < m.x1 && yy >= m.y0 && yy < m.y1)
    f[i] += alpha * kFieldUnit * y[i] + fmaxf(1.f - alpha, 0.f) * scale * dir[i];
}

// target[s][p] = (field[p] - f_s[p]) / unit for samples s of a block
__global__ void retarget_kernel(const float *field, const float *fs, float *target, int count, size_t n2) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n2 * (size_t)count) return;
  target[i] = (field[i % n2] - fs[i]) / kFieldUnit;
}

__global__ void copy_f32_kernel(const float *src, float *dst, size_t n) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// dst = src + scale * dir (masked direction already)
__global__ void axpy_kernel(const float *src, const float *dir, float scale, float *dst, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i] + scale * dir[i];
}

unsigned blocks_for(size_t n, unsigned t = 256) { return (unsigned)((n + t - 1) / t); }

}  // namespace

// ====================================================================== network object
struct tomo_cnn {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // parameters and Adam state (FP32, canonical layout)
  float *d_params = nullptr, *d_grad = nullptr, *d_m = nullptr, *d_v = nullptr;
  int adam_t = 0;
  __bf16 *d_wf[kLayers] = {}, *d_wb[kLayers] = {};
  float *d_bias[kLayers] = {};
  bool packed = false;
  // activations for a batch of npix pixels
  size_t cap_pix = 0;
  __bf16 *d_x8 = nullptr;               // input [p][8]
  __bf16 *d_act[kLayers - 1] = {};      // hidden outputs [p][32]
  float *d_y = nullptr;                 // output [p]
  __bf16 *d_g8 = nullptr;               // output-layer gradient [p][8]
  __bf16 *d_ga = nullptr, *d_gb = nullptr;  // hidden gradients [p][32] (ping-pong)
  float *d_part = nullptr;              // weight-gradient partials
  double *d_wred = nullptr;             // their sums over runs of waves
  size_t part_n = 0;
  double *d_red = nullptr;              // reductions (loss partials, stats)
  float *d_tmp3 = nullptr;              // host-staged input [p][3] / target / dy
  size_t tmp3_n = 0;
  // training samples (all of one grid size)
  int sH = 0, sW = 0, nsamples = 0, cap_samples = 0;
  size_t cap_n2 = 0;  // cells per sample the storage was sized for
  Mask smask{0, 0, 0, 0};
  __bf16 *d_sx = nullptr;   // [N][H][W][8]
  float *d_st = nullptr;    // [N][H][W]  targets
  float *d_sf = nullptr;    // [N][H][W]  field of each sample (for new targets)
  float *d_batch_t = nullptr;
  // engine-side scratch for inference / sampling
  float *d_dir = nullptr, *d_cov = nullptr, *d_fprev = nullptr, *d_dprev = nullptr;
  long long eng_n2 = 0;
  // learned iteration state: the context it runs on and the current learned
  // step factor (<= the caller's alpha); d_fprev holds the plain fallback
  const tomo_ctx *ls_ctx = nullptr;
  float ls_alpha = 0.f;
  bool sampling = false;  // learned iterations add a training sample per state
};

namespace {

int cfail(tomo_cnn *n, int code, const std::string &msg) {
  if (n) n->err = msg;
  return code;
}

#define CNNCHK(net, call)                                                                     \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess) return cfail((net), TOMO_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
int nalloc(tomo_cnn *n, T **p, size_t count) {
  if (*p) {
    hipFree(*p);
    *p = nullptr;
  }
  if (count == 0) return TOMO_OK;
  if (hipMalloc((void **)p, count * sizeof(T)) != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return cfail(n, TOMO_ERR_ALLOC, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed");
  }
  return TOMO_OK;
}

template <int CIN, int COUT>
size_t wgrad_part_elems(int nwaves) {
  using S = WgradShape<CIN, COUT>;
  return (size_t)nwaves * S::MROW * S::NCOL;
}

// waves of a weight-gradient launch: one per 32-pixel chunk up to 1024,
// rounded up to whole blocks
int wgrad_waves(size_t npix) {
  const size_t w = std::min<size_t>(std::max<size_t>(npix / 32, 1), 1024);
  return (int)((w + kWgWaves - 1) / kWgWaves * kWgWaves);
}

int ensure_batch(tomo_cnn *n, size_t npix) {
  if (n->cap_pix >= npix) return TOMO_OK;
  int rc;
  if ((rc = nalloc(n, &n->d_x8, npix * 8)) || (rc = nalloc(n, &n->d_y, npix)) ||
      (rc = nalloc(n, &n->d_g8, npix * 8)) || (rc = nalloc(n, &n->d_ga, npix * 32)) ||
      (rc = nalloc(n, &n->d_gb, npix * 32)) || (rc = nalloc(n, &n->d_batch_t, npix)))
    return rc;
  for (int l = 0; l < kLayers - 1; ++l)
    if ((rc = nalloc(n, &n->d_act[l], npix * 32))) return rc;
  const int nw = wgrad_waves(npix);
  const size_t part = std::max(wgrad_part_elems<32, 32>(nw), wgrad_part_elems<8, 32>(nw));
  if ((rc = nalloc(n, &n->d_part, std::max(part, wgrad_part_elems<32, 8>(nw))))) return rc;
  if (!n->d_wred && (rc = nalloc(n, &n->d_wred, (size_t)kWredGroups * (32 * 32 * 9 + 32)))) return rc;
  n->part_n = part;
  n->cap_pix = npix;
  return TOMO_OK;
}

int ensure_tmp(tomo_cnn *n, size_t count) {
  if (n->tmp3_n >= count) return TOMO_OK;
  int rc = nalloc(n, &n->d_tmp3, count);
  if (!rc) n->tmp3_n = count;
  return rc;
}

int pack(tomo_cnn *n) {
  if (n->packed) return TOMO_OK;
  PackArgs a;
  a.params = n->d_params;
  for (int l = 0; l < kLayers; ++l) {
    a.wf[l] = n->d_wf[l];
    a.wb[l] = n->d_wb[l];
    a.bias[l] = n->d_bias[l];
  }
  hipLaunchKernelGGL(pack_weights_kernel, dim3(48, kLayers), dim3(256), 0, n->stream, a);
  CNNCHK(n, hipGetLastError());
  n->packed = true;
  return TOMO_OK;
}

int conv_grid(size_t groups) { return (int)std::min<size_t>((groups + 3) / 4, 4096); }

template <int CIN, int COUT, int EPI>
int launch_conv(tomo_cnn *n, const __bf16 *in, const __bf16 *w, const float *bias, const __bf16 *gate, void *out,
                int B, int H, int W) {
  ConvArgs a{in, w, bias, gate, out, B, H, W, (int)((size_t)B * H * W / 16)};
  hipLaunchKernelGGL((conv3x3_kernel<CIN, COUT, EPI>), dim3(conv_grid(a.groups)), dim3(256), 0, n->stream, a);
  CNNCHK(n, hipGetLastError());
  return TOMO_OK;
}

// forward over x8 -> act[0..3], y
int forward_net(tomo_cnn *n, int B, int H, int W) {
  int rc;
  if ((rc = pack(n))) return rc;
  if ((rc = launch_conv<8, 32, kEpiBiasRelu>(n, n->d_x8, n->d_wf[0], n->d_bias[0], nullptr, n->d_act[0], B, H, W)))
    return rc;
  for (int l = 1; l < 4; ++l)
    if ((rc = launch_conv<32, 32, kEpiBiasRelu>(n, n->d_act[l - 1], n->d_wf[l], n->d_bias[l], nullptr, n->d_act[l],
                                                 B, H, W)))
      return rc;
  return launch_conv<32, 16, kEpiBiasOut>(n, n->d_act[3], n->d_wf[4], n->d_bias[4], nullptr, n->d_y, B, H, W);
}

template <int CIN, int COUT>
int launch_wgrad(tomo_cnn *n, const __bf16 *in, const __bf16 *d, int layer, int B, int H, int W) {
  const size_t npix = (size_t)B * H * W;
  const int nw = wgrad_waves(npix);
  WgradArgs a{in, d, n->d_part, B, H, W, (int)(npix / 32)};
  hipLaunchKernelGGL((wgrad_kernel<CIN, COUT>), dim3(nw / kWgWaves), dim3(64 * kWgWaves), 0, n->stream, a);
  CNNCHK(n, hipGetLastError());
  const int total = kCout[layer] * kCin[layer] * 9 + kCout[layer];
  hipLaunchKernelGGL((wgrad_reduce_kernel<CIN, COUT>), dim3(blocks_for(total), kWredGroups), dim3(256), 0,
                     n->stream, (const float *)n->d_part, nw, kCin[layer], kCout[layer], n->d_wred);
  CNNCHK(n, hipGetLastError());
  hipLaunchKernelGGL(wgrad_finish_kernel, dim3(blocks_for(total)), dim3(256), 0, n->stream,
                     (const double *)n->d_wred, total, n->d_grad + param_offset(layer));
  CNNCHK(n, hipGetLastError());
  return TOMO_OK;
}

// backward from d_g8 (output-layer gradient) -> d_grad (canonical)
int backward_net(tomo_cnn *n, int B, int H, int W) {
  int rc;
  // layer 5: weights from act[3] and g8; data gradient into ga (gated by act[3])
  if ((rc = launch_wgrad<32, 8>(n, n->d_act[3], n->d_g8, 4, B, H, W))) return rc;
  if ((rc = launch_conv<8, 32, kEpiMask>(n, n->d_g8, n->d_wb[4], nullptr, n->d_act[3], n->d_ga, B, H, W)))
    return rc;
  __bf16 *cur = n->d_ga, *nxt = n->d_gb;
  for (int l = 3; l >= 1; --l) {
    if ((rc = launch_wgrad<32, 32>(n, n->d_act[l - 1], cur, l, B, H, W))) return rc;
    if ((rc = launch_conv<32, 32, kEpiMask>(n, cur, n->d_wb[l], nullptr, n->d_act[l - 1], nxt, B, H, W))) return rc;
    std::swap(cur, nxt);
  }
  return launch_wgrad<8, 32>(n, n->d_x8, cur, 0, B, H, W);
}

int check_shape(tomo_cnn *n, int B, int H, int W) {
  if (B <= 0 || H <= 0 || W <= 0) return cfail(n, TOMO_ERR_ARG, "empty batch");
  if (W % 32 != 0) return cfail(n, TOMO_ERR_ARG, "image width must be a multiple of 32");
  if ((size_t)B * H * W > ((size_t)1 << 31)) return cfail(n, TOMO_ERR_ARG, "batch too large");
  return TOMO_OK;
}

int upload_x3(tomo_cnn *n, const float *x, size_t npix) {
  int rc;
  if ((rc = ensure_tmp(n, npix * 3))) return rc;
  CNNCHK(n, hipMemcpyAsync(n->d_tmp3, x, sizeof(float) * npix * 3, hipMemcpyHostToDevice, n->stream));
  hipLaunchKernelGGL(pack_input_kernel, dim3(blocks_for(npix)), dim3(256), 0, n->stream, (const float *)n->d_tmp3,
                     n->d_x8, npix);
  CNNCHK(n, hipGetLastError());
  return TOMO_OK;
}

int adam_step(tomo_cnn *n, double lr) {
  n->adam_t += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)n->adam_t), c2 = 1.f - std::pow(b2, (float)n->adam_t);
  hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(kParams)), dim3(256), 0, n->stream, n->d_params,
                     (const float *)n->d_grad, n->d_m, n->d_v, kParams, (float)lr, b1, b2, 1e-8f, c1, c2);
  CNNCHK(n, hipGetLastError());
  n->packed = false;
  return TOMO_OK;
}

// loss of the batch in d_x8 / d_batch_t (device), gradient in d_grad
int loss_and_grad(tomo_cnn *n, int B, int H, int W, Mask m, double *loss_dev) {
  int rc;
  if ((rc = forward_net(n, B, H, W))) return rc;
  const int nb = 256;
  const float count = (float)((double)B * (m.x1 - m.x0) * (m.y1 - m.y0));
  hipLaunchKernelGGL(loss_kernel, dim3(nb), dim3(256), 0, n->stream, (const float *)n->d_y,
                     (const float *)n->d_batch_t, n->d_g8, B, H, W, m.x0, m.x1, m.y0, m.y1, count, n->d_red);
  CNNCHK(n, hipGetLastError());
  hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(64), 0, n->stream, (const double *)n->d_red, nb,
                     0.5 / (double)count, loss_dev);
  CNNCHK(n, hipGetLastError());
  return backward_net(n, B, H, W);
}

int ensure_engine_scratch(tomo_cnn *n, long long n2) {
  if (n->eng_n2 >= n2) return TOMO_OK;
  int rc;
  if ((rc = nalloc(n, &n->d_dir, (size_t)n2)) || (rc = nalloc(n, &n->d_cov, (size_t)n2)) ||
      (rc = nalloc(n, &n->d_fprev, (size_t)n2)) || (rc = nalloc(n, &n->d_dprev, (size_t)n2)))
    return rc;
  n->eng_n2 = n2;
  return TOMO_OK;
}

// sample storage for `cap` samples of n2 cells (the stored ones are kept;
// on failure the old storage stays and nothing new is held)
int reserve_samples(tomo_cnn *n, int cap, size_t n2) {
  int rc;
  __bf16 *sx = nullptr;
  float *st = nullptr, *sf = nullptr;
  if ((rc = nalloc(n, &sx, (size_t)cap * n2 * 8)) || (rc = nalloc(n, &st, (size_t)cap * n2)) ||
      (rc = nalloc(n, &sf, (size_t)cap * n2))) {
    if (sx) (void)hipFree(sx);
    if (st) (void)hipFree(st);
    if (sf) (void)hipFree(sf);
    return rc;
  }
  if (n->nsamples) {
    CNNCHK(n, hipMemcpy(sx, n->d_sx, sizeof(__bf16) * n->nsamples * n2 * 8, hipMemcpyDeviceToDevice));
    CNNCHK(n, hipMemcpy(st, n->d_st, sizeof(float) * n->nsamples * n2, hipMemcpyDeviceToDevice));
    CNNCHK(n, hipMemcpy(sf, n->d_sf, sizeof(float) * n->nsamples * n2, hipMemcpyDeviceToDevice));
  }
  if (n->d_sx) (void)hipFree(n->d_sx);
  if (n->d_st) (void)hipFree(n->d_st);
  if (n->d_sf) (void)hipFree(n->d_sf);
  n->d_sx = sx;
  n->d_st = st;
  n->d_sf = sf;
  n->cap_samples = cap;
  n->cap_n2 = n2;
  return TOMO_OK;
}

// network input of phantom p of the context's current state into d_x8 (its
// field, its own shots' plain step and coverage) and, when `target` is
// non-null, the target from its truth; needs gradients of every shot
int engine_input(tomo_cnn *n, tomo_ctx *c, int p, float scale, float *target) {
  int rc;
  if ((rc = ensure_engine_scratch(n, c->n2))) return rc;
  if ((rc = ensure_batch(n, (size_t)c->n2))) return rc;
  const Mask m = ctx_mask(c);
  const int G = c->nshots / c->nph;
  const float *f = c->d_f + (size_t)p * c->n2, *fo = c->d_fo + (size_t)p * c->n2;
  const float *grad = c->d_grad + (size_t)p * G * c->n2;
  // the engine stream orders these after the gradient pass; the network
  // stream waits on a sync below
  hipLaunchKernelGGL(direction_kernel, dim3(blocks_for(c->n2)), dim3(256), 0, c->stream, f, grad, G, c->nx, c->ny,
                     m, n->d_dir);
  HIPCHK(c, hipGetLastError());
  const float *cov = nullptr;
  if (c->coverage && c->d_cov) {
    hipLaunchKernelGGL(coverage_sum_kernel, dim3(blocks_for(c->n2)), dim3(256), 0, c->stream,
                       (const float *)(c->d_cov + (size_t)p * G * c->n2), G, c->n2, n->d_cov);
    HIPCHK(c, hipGetLastError());
    cov = n->d_cov;
  }
  hipLaunchKernelGGL(input_stats_kernel, dim3(1), dim3(256), 0, c->stream, (const float *)n->d_dir, cov, c->nx,
                     c->ny, m, n->d_red);
  HIPCHK(c, hipGetLastError());
  hipLaunchKernelGGL(make_input_kernel, dim3(blocks_for(c->n2)), dim3(256), 0, c->stream, f,
                     (const float *)n->d_dir, scale, cov, (const double *)n->d_red, (int)c->n2, n->d_x8, fo, target);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TOMO_OK;
}

}  // namespace

// ====================================================================== ABI
extern "C" {

int tomo_cnn_num_params(void) { return kParams; }

int tomo_cnn_create(int device, tomo_cnn **out) {
  if (!out) return TOMO_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    return TOMO_ERR_NODEV;
  }
  if (device < 0) device = 0;
  if (device >= ndev) return TOMO_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return TOMO_ERR_HIP;
  tomo_cnn *n = new tomo_cnn;
  n->device = device;
  int rc = TOMO_OK;
  if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess) rc = TOMO_ERR_HIP;
  if (!rc) rc = nalloc(n, &n->d_params, kParams);
  if (!rc) rc = nalloc(n, &n->d_grad, kParams);
  if (!rc) rc = nalloc(n, &n->d_m, kParams);
  if (!rc) rc = nalloc(n, &n->d_v, kParams);
  if (!rc) rc = nalloc(n, &n->d_red, 1024);
  for (int l = 0; l < kLayers && !rc; ++l) {
    const int mrow_f = ((kCout[l] + 15) / 16) * 16;
    rc = nalloc(n, &n->d_wf[l], (size_t)mrow_f * kpad(kCinS[l]));
    if (!rc) rc = nalloc(n, &n->d_wb[l], (size_t)kCin[l] * kpad(kCoutS[l]));
    if (!rc) rc = nalloc(n, &n->d_bias[l], kCout[l]);
  }
  if (rc) {
    tomo_cnn_destroy(n);
    return rc;
  }
  *out = n;
  return tomo_cnn_init(n, 1);
}

void tomo_cnn_destroy(tomo_cnn *n) {
  if (!n) return;
  (void)hipSetDevice(n->device);
  if (n->stream) (void)hipStreamSynchronize(n->stream);
  auto fr = [](auto *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
  };
  fr(n->d_params); fr(n->d_grad); fr(n->d_m); fr(n->d_v); fr(n->d_red);
  for (int l = 0; l < kLayers; ++l) { fr(n->d_wf[l]); fr(n->d_wb[l]); fr(n->d_bias[l]); }
  fr(n->d_x8); fr(n->d_y); fr(n->d_g8); fr(n->d_ga); fr(n->d_gb); fr(n->d_part); fr(n->d_wred); fr(n->d_tmp3);
  fr(n->d_batch_t);
  for (int l = 0; l < kLayers - 1; ++l) fr(n->d_act[l]);
  fr(n->d_sx); fr(n->d_st); fr(n->d_sf);
  fr(n->d_dir); fr(n->d_cov); fr(n->d_fprev); fr(n->d_dprev);
  if (n->stream) (void)hipStreamDestroy(n->stream);
  delete n;
}

const char *tomo_cnn_last_error(const tomo_cnn *n) { return n ? n->err.c_str() : ""; }

int tomo_cnn_init(tomo_cnn *n, uint64_t seed) {
  if (!n) return TOMO_ERR_ARG;
  // He-normal weights (std sqrt(2 / (9 cin))), zero biases; the output layer
  // starts at a tenth of that so the first updates are small
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  std::vector<float> p(kParams, 0.f);
  for (int l = 0; l < kLayers; ++l) {
    const double sd = std::sqrt(2.0 / (9.0 * kCin[l])) * (l == kLayers - 1 ? 0.1 : 1.0);
    float *w = p.data() + param_offset(l);
    for (int i = 0; i < kCout[l] * kCin[l] * 9; ++i) w[i] = (float)(sd * nd(rng));
  }
  return tomo_cnn_set_params(n, p.data());
}

int tomo_cnn_set_params(tomo_cnn *n, const float *params) {
  if (!n || !params) return TOMO_ERR_ARG;
  CNNCHK(n, hipSetDevice(n->device));
  CNNCHK(n, hipMemcpyAsync(n->d_params, params, sizeof(float) * kParams, hipMemcpyHostToDevice, n->stream));
  CNNCHK(n, hipMemsetAsync(n->d_m, 0, sizeof(float) * kParams, n->stream));
  CNNCHK(n, hipMemsetAsync(n->d_v, 0, sizeof(float) * kParams, n->stream));
  n->adam_t = 0;
  n->packed = false;
  CNNCHK(n, hipStreamSynchronize(n->stream));
  return TOMO_OK;
}

int tomo_cnn_get_params(tomo_cnn *n, float *params) {
  if (!n || !params) return TOMO_ERR_ARG;
  CNNCHK(n, hipSetDevice(n->device));
  CNNCHK(n, hipMemcpyAsync(params, n->d_params, sizeof(float) * kParams, hipMemcpyDeviceToHost, n->stream));
  CNNCHK(n, hipStreamSynchronize(n->stream));
  return TOMO_OK;
}

int tomo_cnn_forward(tomo_cnn *n, int B, int H, int W, const float *x, float *y) {
  if (!n || !x || !y) return TOMO_ERR_ARG;
  int rc;
  if ((rc = check_shape(n, B, H, W))) return rc;
  CNNCHK(n, hipSetDevice(n->device));
  const size_t npix = (size_t)B * H * W;
  if ((rc = ensure_batch(n, npix)) || (rc = upload_x3(n, x, npix)) || (rc = forward_net(n, B, H, W))) return rc;
  CNNCHK(n, hipMemcpyAsync(y, n->d_y, sizeof(float) * npix, hipMemcpyDeviceToHost, n->stream));
  CNNCHK(n, hipStreamSynchronize(n->stream));
  return TOMO_OK;
}

int tomo_cnn_backward(tomo_cnn *n, int B, int H, int W, const float *x, const float *dy, float *grad) {
  if (!n || !x || !dy || !grad) return TOMO_ERR_ARG;
  int rc;
  if ((rc = check_shape(n, B, H, W))) return rc;
  CNNCHK(n, hipSetDevice(n->device));
  const size_t npix = (size_t)B * H * W;
  if ((rc = ensure_batch(n, npix)) || (rc = upload_x3(n, x, npix)) || (rc = forward_net(n, B, H, W))) return rc;
  // the staging buffer is free again once the input is packed (same stream)
  CNNCHK(n, hipMemcpyAsync(n->d_tmp3, dy, sizeof(float) * npix, hipMemcpyHostToDevice, n->stream));
  hipLaunchKernelGGL(pack_dy_kernel, dim3(blocks_for(npix)), dim3(256), 0, n->stream, (const float *)n->d_tmp3,
                     n->d_g8, npix);
  CNNCHK(n, hipGetLastError());
  if ((rc = backward_net(n, B, H, W))) return rc;
  CNNCHK(n, hipMemcpyAsync(grad, n->d_grad, sizeof(float) * kParams, hipMemcpyDeviceToHost, n->stream));
  CNNCHK(n, hipStreamSynchronize(n->stream));
  return TOMO_OK;
}

int tomo_cnn_loss(tomo_cnn *n, int B, int H, int W, const float *x, const float *target, const int *mask4,
                  double *loss, float *grad) {
  if (!n || !x || !target || !mask4 || !loss) return TOMO_ERR_ARG;
  int rc;
  if ((rc = check_shape(n, B, H, W))) return rc;
  const Mask m{mask4[0], mask4[1], mask4[2], mask4[3]};
  if (m.x0 < 0 || m.x1 > W || m.y0 < 0 || m.y1 > H || m.x0 >= m.x1 || m.y0 >= m.y1)
    return cfail(n, TOMO_ERR_ARG, "mask outside the image or empty");
  CNNCHK(n, hipSetDevice(n->device));
  const size_t npix = (size_t)B * H * W;
  if ((rc = ensure_batch(n, npix)) || (rc = upload_x3(n, x, npix))) return rc;
  CNNCHK(n, hipMemcpyAsync(n->d_batch_t, target, sizeof(float) * npix, hipMemcpyHostToDevice, n->stream));
  if ((rc = loss_and_grad(n, B, H, W, m, n->d_red + 512))) return rc;
  CNNCHK(n, hipMemcpyAsync(loss, n->d_red + 512, sizeof(double), hipMemcpyDeviceToHost, n->stream));
  if (grad) CNNCHK(n, hipMemcpyAsync(grad, n->d_grad, sizeof(float) * kParams, hipMemcpyDeviceToHost, n->stream));
  CNNCHK(n, hipStreamSynchronize(n->stream));
  return TOMO_OK;
}

int tomo_cnn_clear_samples(tomo_cnn *n) {
  if (!n) return TOMO_ERR_ARG;
  n->nsamples = 0;
  return TOMO_OK;
}

int tomo_cnn_num_samples(const tomo_cnn *n) { return n ? n->nsamples : 0; }

int tomo_cnn_add_sample_phantom(tomo_cnn *n, tomo_ctx *c, int p, float scale) {
  if (!n || !c) return TOMO_ERR_ARG;
  int rc = check_ctx(c);
  if (rc) return rc;
  if (p < 0 || p >= c->nph) return cfail(n, TOMO_ERR_ARG, "phantom index out of range");
  if ((rc = need_shots(c)) || (rc = need_all_gradients(c))) return cfail(n, rc, c->err);
  if (!c->have_truth) return cfail(n, TOMO_ERR_STATE, "the context has no ground truth");
  if (c->nx % 32 != 0) return cfail(n, TOMO_ERR_ARG, "grid width must be a multiple of 32");
  const Mask m = ctx_mask(c);
  if (n->nsamples == 0) {
    n->sH = c->ny;
    n->sW = c->nx;
    n->smask = m;
  } else if (n->sH != c->ny || n->sW != c->nx || n->smask.x0 != m.x0 || n->smask.x1 != m.x1 ||
             n->smask.y0 != m.y0 || n->smask.y1 != m.y1) {
    return cfail(n, TOMO_ERR_ARG, "samples must share one grid size and mask");
  }
  const size_t n2 = (size_t)c->n2;
  // storage reserved for another grid size holds no sample yet: resize it
  if (n->nsamples == 0 && n->cap_samples > 0 && n->cap_n2 != n2 && (rc = reserve_samples(n, n->cap_samples, n2)))
    return rc;
  if (n->nsamples == n->cap_samples && (rc = reserve_samples(n, std::max(16, n->cap_samples * 2), n2))) return rc;
  if ((rc = engine_input(n, c, p, scale, n->d_st + (size_t)n->nsamples * n2))) return rc;
  CNNCHK(n, hipMemcpyAsync(n->d_sx + (size_t)n->nsamples * n2 * 8, n->d_x8, sizeof(__bf16) * n2 * 8,
                           hipMemcpyDeviceToDevice, c->stream));
  CNNCHK(n, hipMemcpyAsync(n->d_sf + (size_t)n->nsamples * n2, c->d_f + (size_t)p * n2, sizeof(float) * n2,
                           hipMemcpyDeviceToDevice, c->stream));
  CNNCHK(n, hipStreamSynchronize(c->stream));
  n->nsamples += 1;
  return TOMO_OK;
}

int tomo_cnn_add_sample(tomo_cnn *n, tomo_ctx *c, float scale) {
  if (!n || !c) return TOMO_ERR_ARG;
  if (c->nph != 1) return cfail(n, TOMO_ERR_STATE, "tomo_cnn_add_sample takes a context of one phantom");
  return tomo_cnn_add_sample_phantom(n, c, 0, scale);
}

int tomo_cnn_reserve_samples(tomo_cnn *n, int count, int ny, int nx) {
  if (!n || count < 0 || ny <= 0 || nx <= 0) return TOMO_ERR_ARG;
  if (n->nsamples > 0 && (n->sH != ny || n->sW != nx)) return cfail(n, TOMO_ERR_ARG, "samples must share one grid size");
  const size_t n2 = (size_t)ny * nx;
  if (count <= n->cap_samples && n2 == n->cap_n2) return TOMO_OK;
  CNNCHK(n, hipSetDevice(n->device));
  return reserve_samples(n, std::max(count, n->nsamples), n2);
}

int tomo_cnn_set_targets(tomo_cnn *n, int first, int count, const float *field) {
  if (!n || !field || first < 0 || count < 0 || first + count > n->nsamples) return TOMO_ERR_ARG;
  if (count == 0) return TOMO_OK;
  CNNCHK(n, hipSetDevice(n->device));
  const size_t n2 = (size_t)n->sH * n->sW;
  int rc;
  if ((rc = ensure_tmp(n, n2))) return rc;
  CNNCHK(n, hipMemcpyAsync(n->d_tmp3, field, sizeof(float) * n2, hipMemcpyHostToDevice, n->stream));
  hipLaunchKernelGGL(retarget_kernel, dim3(blocks_for(n2 * count)), dim3(256), 0, n->stream, (const float *)n->d_tmp3,
                     (const float *)(n->d_sf + (size_t)first * n2), n->d_st + (size_t)first * n2, count, n2);
  CNNCHK(n, hipGetLastError());
  CNNCHK(n, hipStreamSynchronize(n->stream));
  return TOMO_OK;
}

int tomo_cnn_train(tomo_cnn *n, int steps, int batch, double lr, uint64_t seed, double *losses) {
  if (!n || steps < 0 || batch <= 0 || !(lr > 0.0)) return TOMO_ERR_ARG;
  if (n->nsamples <= 0) return cfail(n, TOMO_ERR_STATE, "no training samples");
  CNNCHK(n, hipSetDevice(n->device));
  const int B = std::min(batch, n->nsamples);
  const int H = n->sH, W = n->sW;
  const size_t n2 = (size_t)H * W;
  int rc;
  if ((rc = ensure_batch(n, (size_t)B * n2))) return rc;
  std::mt19937_64 rng(seed);
  std::vector<int> order(n->nsamples);
  for (int i = 0; i < n->nsamples; ++i) order[i] = i;
  size_t pos = order.size();
  std::vector<double> lossv(steps, 0.0);
  for (int s = 0; s < steps; ++s) {
    for (int b = 0; b < B; ++b) {
      if (pos >= order.size()) {
        std::shuffle(order.begin(), order.end(), rng);
        pos = 0;
      }
      const int k = order[pos++];
      CNNCHK(n, hipMemcpyAsync(n->d_x8 + (size_t)b * n2 * 8, n->d_sx + (size_t)k * n2 * 8,
                               sizeof(__bf16) * n2 * 8, hipMemcpyDeviceToDevice, n->stream));
      CNNCHK(n, hipMemcpyAsync(n->d_batch_t + (size_t)b * n2, n->d_st + (size_t)k * n2, sizeof(float) * n2,
                               hipMemcpyDeviceToDevice, n->stream));
    }
    if ((rc = loss_and_grad(n, B, H, W, n->smask, n->d_red + 512 + (s % 256)))) return rc;
    if ((rc = adam_step(n, lr))) return rc;
    if (s % 256 == 255 || s == steps - 1) {
      const int cnt = s % 256 + 1;
      CNNCHK(n, hipMemcpyAsync(lossv.data() + (s - cnt + 1), n->d_red + 512, sizeof(double) * cnt,
                               hipMemcpyDeviceToHost, n->stream));
      CNNCHK(n, hipStreamSynchronize(n->stream));
    }
  }
  if (losses) std::copy(lossv.begin(), lossv.end(), losses);
  return TOMO_OK;
}

// Learned field update of the context: f += mask * (alpha * unit * CNN(input)
// + max(1 - alpha, 0) * plain step), with the input from the current field,
// the plain SIRT step (scale times the uniform direction of the stored
// gradients) and the coverage.
int tomo_learned_update(tomo_ctx *c, tomo_cnn *n, float alpha, float scale) {
  if (!n || !c) return TOMO_ERR_ARG;
  int rc = check_ctx(c);
  if (rc) return rc;
  if (c->nph != 1) return fail(c, TOMO_ERR_STATE, "tomo_learned_update takes a context of one phantom");
  if ((rc = need_shots(c)) || (rc = need_all_gradients(c))) return rc;
  if (c->nx % 32 != 0) return fail(c, TOMO_ERR_ARG, "grid width must be a multiple of 32");
  if ((rc = engine_input(n, c, 0, scale, nullptr))) return fail(c, rc, n->err);
  if ((rc = forward_net(n, 1, c->ny, c->nx))) return fail(c, rc, n->err);
  CNNCHK(n, hipStreamSynchronize(n->stream));
  hipLaunchKernelGGL(learned_update_kernel, dim3(blocks_for(c->n2)), dim3(256), 0, c->stream, c->d_f,
                     (const float *)n->d_y, (const float *)n->d_dir, c->nx, c->ny, ctx_mask(c), alpha, scale);
  HIPCHK(c, hipGetLastError());
  if ((rc = upload_coefficients(c))) return rc;
  std::fill(c->grad_valid.begin(), c->grad_valid.end(), 0);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TOMO_OK;
}

int tomo_cnn_set_sampling(tomo_cnn *n, int on) {
  if (!n) return TOMO_ERR_ARG;
  n->sampling = on != 0;
  return TOMO_OK;
}

int tomo_learned_reset(tomo_cnn *n) {
  if (!n) return TOMO_ERR_ARG;
  n->ls_ctx = nullptr;
  n->ls_alpha = 0.f;
  return TOMO_OK;
}

int tomo_learned_iteration(tomo_ctx *c, tomo_cnn *n, float alpha, float scale, int weighted, double *misfit,
                           double *eps, int *accepted) {
  if (!n || !c) return TOMO_ERR_ARG;
  int rc = check_ctx(c);
  if (rc) return rc;
  if (c->nph != 1) return fail(c, TOMO_ERR_STATE, "tomo_learned_iteration takes a context of one phantom");
  if ((rc = need_shots(c))) return rc;
  if (c->nx % 32 != 0) return fail(c, TOMO_ERR_ARG, "grid width must be a multiple of 32");
  if (!(alpha > 0.f)) return fail(c, TOMO_ERR_ARG, "alpha must be positive");
  if (n->ls_ctx != c) {
    n->ls_ctx = c;
    n->ls_alpha = alpha;
  }
  n->ls_alpha = std::min(n->ls_alpha, alpha);
  double m0 = 0.0;
  if ((rc = tomo_gradient(c, 0, c->nshots, &m0))) return rc;
  if (n->sampling && (rc = tomo_cnn_add_sample(n, c, scale))) return fail(c, rc, n->err);
  // the plain step from the current field: the fallback
  if ((rc = ensure_engine_scratch(n, c->n2))) return fail(c, rc, n->err);
  hipLaunchKernelGGL(copy_f32_kernel, dim3(blocks_for(c->n2)), dim3(256), 0, c->stream, (const float *)c->d_f,
                     n->d_fprev, (size_t)c->n2);
  HIPCHK(c, hipGetLastError());
  const int M = c->nshots;
  if (weighted) {
    const int pb = weighted_part_blocks((int)c->n2);
    if ((rc = ensure_scratch(c, (size_t)c->n2 + (size_t)M * pb + (size_t)M + 1))) return rc;
    const float *pw = nullptr;
    if ((rc = pixel_weights(c, 0, M, &pw))) return rc;  // before the scratch below is taken
    double *mean = c->d_scratch, *part = mean + c->n2, *w = part + (size_t)M * pb;
    HIPCHK(c, launch_weighted(n->d_fprev, c->d_grad, M, c->nx, c->ny, ctx_mask(c), (double)scale, mean, part, pb, w,
                              c->stream, pw));
  } else {
    const float *pw = nullptr;
    if ((rc = pixel_weights(c, 0, M, &pw))) return rc;
    HIPCHK(c, launch_update_uniform(n->d_fprev, c->d_grad, M, c->nx, c->ny, ctx_mask(c), (double)scale, (double)M,
                                    c->stream, pw));
  }
  // the learned candidate, checked by the data misfit before it is kept
  if ((rc = tomo_learned_update(c, n, n->ls_alpha, scale))) return rc;
  double m1 = 0.0;
  if ((rc = tomo_misfit(c, &m1))) return rc;
  int acc = m1 < m0;
  if (acc) {
    n->ls_alpha = std::min(alpha, n->ls_alpha * 1.5f);
  } else {
    // the misfit rose (or is NaN): keep the plain step instead
    n->ls_alpha *= 0.5f;
    hipLaunchKernelGGL(copy_f32_kernel, dim3(blocks_for(c->n2)), dim3(256), 0, c->stream,
                       (const float *)n->d_fprev, c->d_f, (size_t)c->n2);
    HIPCHK(c, hipGetLastError());
    if ((rc = upload_coefficients(c))) return rc;
  }
  std::fill(c->grad_valid.begin(), c->grad_valid.end(), 0);
  if (misfit) *misfit = m0;
  if (accepted) *accepted = acc;
  if (eps && c->have_truth) return tomo_error(c, eps);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TOMO_OK;
}

}  // extern "C"
