// tomo_host.h -- host-side helpers shared by the engine and the C ABI
// (internal).
#pragma once

#include <cmath>
#include <string>
#include <vector>

#include "tomo.h"

namespace tomo_host {

// last tomo_create() failure (no context exists to hold it)
void set_create_error(const std::string &msg);
const char *create_error();

// sponge damping g(d) = gamma ((W - d) / W)^2 for edge distance d < W
inline double sponge_gamma(const tomo_config &c, int d) {
  if (c.sponge_width <= 0 || d >= c.sponge_width) return 0.0;
  const double x = double(c.sponge_width - d) / double(c.sponge_width);
  return c.sponge_gamma * x * x;
}

// Source pulse at step k (0 outside the window), evaluated in T arithmetic.
// Reference pulse: cos(wc t) exp(-t^2 / (2 tau^2)), t = k dt - t_shift,
// active while k + 1 <= src_last_level (sirt-mimo.cu:553-562).
template <typename T>
T pulse(const tomo_config &c, int k) {
  if (c.src_last_level > 0 && k + 1 > c.src_last_level) return T(0);
  if (c.wavelet == TOMO_WAVELET_REF_GAUSS_COS) {
    const T t = T(k) * T(c.dt) - T(c.wl_shift);
    const T tau = T(c.wl_tau);
    return std::cos(T(c.wl_omega) * t) * std::exp(-(t * t) / (T(2) * tau * tau));
  }
  const T t = T(k) * T(c.dt) - T(c.ricker_t0);
  const T a = T(3.14159265358979323846) * T(c.ricker_fp) * t;
  const T x = a * a;
  return (T(1) - T(2) * x) * std::exp(-x);
}

// Receiver cells and anchors of the four sensor lines (order bottom, right,
// top, left; sirt-mimo.cu:629-647, 724-740).
void line_receivers(const tomo_config &c, std::vector<int> &cell, std::vector<int> &anchor);

int traces_to_sides(const tomo_config &c, int nshots, const float *traces, float *bottom, float *right,
                    float *top, float *left);
int sides_to_traces(const tomo_config &c, int nshots, const float *bottom, const float *right,
                    const float *top, const float *left, float *traces);

}  // namespace tomo_host
