// cnn_oracle.cpp -- CPU oracle of the learned preconditioner network.
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/ (and nothing in the product
// path).  The network has no counterpart in the reference; this file states
// it independently of the GPU code so that the HIP kernels can be checked
// against it:
//
//   5 convolution layers, 3x3 taps, zero padding ("same"), stride 1,
//   channels 3 -> 32 -> 32 -> 32 -> 32 -> 1, ReLU after layers 1-4;
//   y[p, o] = b[o] + sum_{c, ky, kx} W[o][c][ky][kx] * x[p + (ky - 1, kx - 1), c]
//   (cross-correlation).  Parameters are stored layer by layer as W (OIHW)
//   then b; 28929 values in all.
//
// Two arithmetic modes:
//   mode 0: everything in FP64 (pinned by finite differences of the loss);
//   mode 1: the GPU's rounding points -- inputs, weights, hidden activations
//           and the back-propagated pre-activation gradients rounded to
//           bfloat16 (round to nearest even) after an FP32 rounding of the
//           FP64 sum; sums themselves exact.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

// Planted mistake for the finite-difference test (ORC_MUTATION 5): the
// backward data pass does not mirror the taps.
#ifndef ORC_MUTATION
#define ORC_MUTATION 0
#endif

namespace {

constexpr int kLayers = 5;
constexpr int kCin[kLayers] = {3, 32, 32, 32, 32};
constexpr int kCout[kLayers] = {32, 32, 32, 32, 1};

int layer_offset(int l) {
  int off = 0;
  for (int i = 0; i < l; ++i) off += kCout[i] * kCin[i] * 9 + kCout[i];
  return off;
}

float f32_round(double v) { return (float)v; }

float bf16_round(float v) {
  uint32_t u;
  std::memcpy(&u, &v, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return v;  // inf / nan unchanged
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  float r;
  std::memcpy(&r, &u, 4);
  return r;
}

double rnd(double v, int mode) { return mode ? (double)bf16_round(f32_round(v)) : v; }

struct Net {
  int B, H, W, mode;
  std::vector<double> w[kLayers], b[kLayers];  // w[l][((o * cin + c) * 9 + t)]
  std::vector<double> act[kLayers + 1];         // act[0] input, act[l] output of layer l ([B][H][W][C])
  Net(const float *params, int B_, int H_, int W_, int mode_) : B(B_), H(H_), W(W_), mode(mode_) {
    for (int l = 0; l < kLayers; ++l) {
      const float *p = params + layer_offset(l);
      const int nw = kCout[l] * kCin[l] * 9;
      w[l].resize(nw);
      b[l].resize(kCout[l]);
      for (int i = 0; i < nw; ++i) w[l][i] = mode ? (double)bf16_round(p[i]) : (double)p[i];
      for (int o = 0; o < kCout[l]; ++o) b[l][o] = (double)p[nw + o];  // biases stay FP32
    }
  }
  size_t npix() const { return (size_t)B * H * W; }

  // pre-activation of layer l at every pixel (double, exact sums)
  void conv(int l, const std::vector<double> &in, std::vector<double> &pre) const {
    const int ci = kCin[l], co = kCout[l];
    pre.assign(npix() * co, 0.0);
#pragma omp parallel for collapse(2) schedule(static)
    for (int bb = 0; bb < B; ++bb)
      for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
          double *out = &pre[(((size_t)bb * H + y) * W + x) * co];
          for (int o = 0; o < co; ++o) {
            double s = b[l][o];
            for (int t = 0; t < 9; ++t) {
              const int yy = y + t / 3 - 1, xx = x + t % 3 - 1;
              if (yy < 0 || yy >= H || xx < 0 || xx >= W) continue;
              const double *src = &in[(((size_t)bb * H + yy) * W + xx) * ci];
              for (int c = 0; c < ci; ++c) s += w[l][((size_t)o * ci + c) * 9 + t] * src[c];
            }
            out[o] = s;
          }
        }
  }

  void forward(const float *x) {
    act[0].resize(npix() * 3);
    for (size_t i = 0; i < act[0].size(); ++i) act[0][i] = mode ? (double)bf16_round(x[i]) : (double)x[i];
    std::vector<double> pre;
    for (int l = 0; l < kLayers; ++l) {
      conv(l, act[l], pre);
      if (l < kLayers - 1) {
        for (auto &v : pre) v = v > 0.0 ? rnd(v, mode) : 0.0;
      } else if (mode) {
        for (auto &v : pre) v = (double)f32_round(v);
      }
      act[l + 1] = pre;
    }
  }
  const std::vector<double> &out() const { return act[kLayers]; }

  // gradient of sum(dy * y) with respect to the parameters (after forward)
  void backward(const double *dy, double *grad) {
    std::vector<double> d(npix());  // pre-activation gradient of the current layer
    for (size_t i = 0; i < npix(); ++i) d[i] = rnd(dy[i], mode);
    for (int l = kLayers - 1; l >= 0; --l) {
      const int ci = kCin[l], co = kCout[l];
      double *gw = grad + layer_offset(l), *gb = gw + (size_t)co * ci * 9;
      const std::vector<double> &in = act[l];
      // weights and bias
#pragma omp parallel for schedule(static)
      for (int o = 0; o < co; ++o) {
        double sb = 0.0;
        for (size_t p = 0; p < npix(); ++p) sb += d[p * co + o];
        gb[o] = sb;
        for (int c = 0; c < ci; ++c)
          for (int t = 0; t < 9; ++t) {
            double s = 0.0;
            for (int bb = 0; bb < B; ++bb)
              for (int y = 0; y < H; ++y) {
                const int yy = y + t / 3 - 1;
                if (yy < 0 || yy >= H) continue;
                for (int x = 0; x < W; ++x) {
                  const int xx = x + t % 3 - 1;
                  if (xx < 0 || xx >= W) continue;
                  s += d[(((size_t)bb * H + y) * W + x) * co + o] * in[(((size_t)bb * H + yy) * W + xx) * ci + c];
                }
              }
            gw[((size_t)o * ci + c) * 9 + t] = s;
          }
      }
      if (l == 0) break;
      // data: gradient with respect to the layer input, then ReLU of layer l-1
      std::vector<double> dn(npix() * ci, 0.0);
#pragma omp parallel for collapse(2) schedule(static)
      for (int bb = 0; bb < B; ++bb)
        for (int y = 0; y < H; ++y)
          for (int x = 0; x < W; ++x)
            for (int c = 0; c < ci; ++c) {
              double s = 0.0;
              for (int t = 0; t < 9; ++t) {
#if ORC_MUTATION == 5
                const int yy = y + t / 3 - 1, xx = x + t % 3 - 1;
#else
                const int yy = y - (t / 3 - 1), xx = x - (t % 3 - 1);
#endif
                if (yy < 0 || yy >= H || xx < 0 || xx >= W) continue;
                const double *src = &d[(((size_t)bb * H + yy) * W + xx) * co];
                for (int o = 0; o < co; ++o) s += w[l][((size_t)o * ci + c) * 9 + t] * src[o];
              }
              const size_t i = (((size_t)bb * H + y) * W + x) * ci + c;
              dn[i] = act[l][i] > 0.0 ? rnd(s, mode) : 0.0;
            }
      d.swap(dn);
    }
  }
};

}  // namespace

extern "C" {

int orc_cnn_num_params(void) { return layer_offset(kLayers); }

// y[B][H][W] from x[B][H][W][3]
void orc_cnn_forward(const float *params, int B, int H, int W, const float *x, float *y, int mode) {
  Net n(params, B, H, W, mode);
  n.forward(x);
  for (size_t i = 0; i < n.npix(); ++i) y[i] = (float)n.out()[i];
}

// gradient of sum(dy * y) with respect to the parameters
void orc_cnn_backward(const float *params, int B, int H, int W, const float *x, const float *dy, double *grad,
                      int mode) {
  Net n(params, B, H, W, mode);
  n.forward(x);
  std::vector<double> d(dy, dy + n.npix());
  n.backward(d.data(), grad);
}

// masked mean-square loss L = sum_{b, p in mask} (y - t)^2 / (2 count),
// count = B * |mask|, mask = [x0, x1) x [y0, y1); and its gradient (optional).
// Mode 1 forms dL/dy = (y - t) / count in FP32, as the GPU does.
double orc_cnn_loss(const float *params, int B, int H, int W, const float *x, const float *target, const int *mask4,
                    double *grad, int mode) {
  Net n(params, B, H, W, mode);
  n.forward(x);
  const double count = (double)B * (mask4[1] - mask4[0]) * (mask4[3] - mask4[2]);
  std::vector<double> dy(n.npix(), 0.0);
  double loss = 0.0;
  for (int bb = 0; bb < B; ++bb)
    for (int yy = mask4[2]; yy < mask4[3]; ++yy)
      for (int xx = mask4[0]; xx < mask4[1]; ++xx) {
        const size_t i = ((size_t)bb * H + yy) * W + xx;
        const double r = n.out()[i] - (double)target[i];
        loss += r * r;
        dy[i] = mode ? (double)(((float)n.out()[i] - target[i]) / (float)count) : r / count;
      }
  loss /= 2.0 * count;
  if (grad) n.backward(dy.data(), grad);
  return loss;
}

}  // extern "C"
