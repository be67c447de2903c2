// tomo_host.cpp -- host half of the C ABI: presets, transducer geometry,
// phantoms, measurement noise, the reference text layout and the
// relaxation schedules.  No device code.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "tomo.h"
#include "tomo_host.h"

// Built with -ffp-contract=off: every product is rounded on its own, which
// the phantom's disk tests need (sums of squares compared against r^2).

namespace tomo_host {

namespace {
std::mutex g_err_mu;
std::string g_create_err;

// SplitMix64: a fixed, platform-independent generator for the seeded phantoms
// and noise.
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  double uniform() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
  double normal() {  // Box-Muller
    double u1 = uniform();
    while (u1 <= 0.0) u1 = uniform();
    const double u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * 3.14159265358979323846 * u2);
  }
};
}  // namespace

void set_create_error(const std::string &msg) {
  std::lock_guard<std::mutex> g(g_err_mu);
  g_create_err = msg;
}

const char *create_error() {
  std::lock_guard<std::mutex> g(g_err_mu);
  return g_create_err.c_str();
}

void line_receivers(const tomo_config &c, std::vector<int> &cell, std::vector<int> &anchor) {
  cell.clear();
  anchor.clear();
  const int nx = c.nx, lo = c.line_lo, hi = c.line_hi;
  for (int side = 0; side < 4; ++side)
    for (int i = c.line_first; i <= c.line_last; ++i) {
      int x, y, ax, ay;
      switch (side) {
        case 0: x = i; y = hi; ax = i; ay = hi - 1; break;   // bottom row
        case 1: x = hi; y = i; ax = hi - 1; ay = i; break;   // right column
        case 2: x = i; y = lo; ax = i; ay = lo + 1; break;   // top row
        default: x = lo; y = i; ax = lo + 1; ay = i; break;  // left column
      }
      cell.push_back(x + nx * y);
      anchor.push_back(ax + nx * ay);
    }
}

int traces_to_sides(const tomo_config &c, int S, const float *tr, float *bottom, float *right, float *top,
                    float *left) {
  const int L = c.line_last - c.line_first + 1, nr = 4 * L, nt = c.nt;
  float *side[4] = {bottom, right, top, left};
  const int len[4] = {c.nx, c.ny, c.nx, c.ny};
  for (int q = 0; q < 4; ++q) std::memset(side[q], 0, sizeof(float) * (size_t)S * nt * len[q]);
  for (int s = 0; s < S; ++s)
    for (int k = 0; k < nt; ++k)
      for (int q = 0; q < 4; ++q)
        for (int i = 0; i < L; ++i)
          side[q][((size_t)s * nt + k) * len[q] + c.line_first + i] = tr[((size_t)s * nt + k) * nr + q * L + i];
  return TOMO_OK;
}

int sides_to_traces(const tomo_config &c, int S, const float *bottom, const float *right, const float *top,
                    const float *left, float *tr) {
  const int L = c.line_last - c.line_first + 1, nr = 4 * L, nt = c.nt;
  const float *side[4] = {bottom, right, top, left};
  const int len[4] = {c.nx, c.ny, c.nx, c.ny};
  for (int s = 0; s < S; ++s)
    for (int k = 0; k < nt; ++k)
      for (int q = 0; q < 4; ++q)
        for (int i = 0; i < L; ++i)
          tr[((size_t)s * nt + k) * nr + q * L + i] = side[q][((size_t)s * nt + k) * len[q] + c.line_first + i];
  return TOMO_OK;
}

}  // namespace tomo_host

using namespace tomo_host;

extern "C" {

void tomo_config_reference(tomo_config *c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  // constants of sirt-mimo.cu:17-35 (float literals, widened exactly)
  c->nx = 192;
  c->ny = 192;
  c->nt = 401;
  c->wavelet = TOMO_WAVELET_REF_GAUSS_COS;
  c->h = (double)0.001f;
  c->dt = (double)3.3333e-07f;
  c->c0 = 1500.0;                          // sirt-mimo.cu:525
  c->wl_omega = (double)7.8540e+05f;
  c->wl_tau = (double)4.0000e-06f;
  c->wl_shift = (double)8.1573e-06f;
  c->src_last_level = 24;                  // "k + 1 <= 24", sirt-mimo.cu:554
  c->sponge_width = 0;
  c->sponge_gamma = 0.0;
  c->recv_mode = TOMO_RECV_LINES;
  c->record_first = 2;                     // "k > 1", sirt-mimo.cu:629
  c->line_lo = 20;
  c->line_hi = 180;
  c->line_first = 21;
  c->line_last = 179;                      // "i > 20 && i < 180", sirt-mimo.cu:629
  c->recv_gain = (double)0.001f * 1000.0;  // "H * 1000.f", sirt-mimo.cu:727
  c->mask_x0 = 21;
  c->mask_x1 = 180;                        // sirt-mimo.cu:870
  c->mask_y0 = 21;
  c->mask_y1 = 180;
  c->checkpoint_every = 0;
  c->shots_per_pass = 0;
}

void tomo_config_unit(tomo_config *c, int n, int nt) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->nx = n;
  c->ny = n;
  c->nt = nt;
  c->wavelet = TOMO_WAVELET_RICKER;
  c->h = 1.0;
  c->dt = 0.4;
  c->c0 = 1.0;
  c->ricker_fp = 0.1;
  c->ricker_t0 = 1.5 / 0.1;
  c->src_last_level = 0;
  c->sponge_width = 20;
  c->sponge_gamma = 0.12;
  c->recv_mode = TOMO_RECV_POINTS;
  c->record_first = 2;
  c->recv_gain = 1.0;
  c->mask_x0 = n / 4;
  c->mask_x1 = n - n / 4;
  c->mask_y0 = n / 4;
  c->mask_y1 = n - n / 4;
  c->checkpoint_every = 0;
  c->shots_per_pass = 0;
}

const char *tomo_strerror(int status) {
  switch (status) {
    case TOMO_OK: return "ok";
    case TOMO_ERR_ARG: return "bad argument";
    case TOMO_ERR_HIP: return "HIP runtime error";
    case TOMO_ERR_ALLOC: return "allocation failed";
    case TOMO_ERR_STATE: return "call made before its prerequisites";
    case TOMO_ERR_IO: return "file missing or malformed";
    case TOMO_ERR_NODEV: return "no usable GPU";
    default: return "unknown status";
  }
}

int tomo_abi_version(void) { return TOMO_ABI_VERSION; }

// Position_Transducers (sirt-mimo.cu:438-466): square ring [21,181]^2, 1-based
int tomo_transducers_reference(int *ii, int *jj, int ns) {
  if (!ii || !jj || ns <= 0 || ns > 640) return TOMO_ERR_ARG;
  for (int p = 0; p < ns; ++p) {
    const int side = p / 160, q = p - 160 * side + 1;  // q = 1..160 along the side
    switch (side) {
      case 0: ii[p] = 21 + q; jj[p] = 181; break;
      case 1: ii[p] = 181; jj[p] = 181 - q; break;
      case 2: ii[p] = 181 - q; jj[p] = 21; break;
      default: ii[p] = 21; jj[p] = 21 + q; break;
    }
  }
  return TOMO_OK;
}

int tomo_transducers_ring(int nx, int ny, int count, double radius, int *ii, int *jj) {
  if (!ii || !jj || count <= 0 || nx < 8 || ny < 8 || !(radius > 0)) return TOMO_ERR_ARG;
  const double cx = 0.5 * (nx - 1), cy = 0.5 * (ny - 1);
  for (int p = 0; p < count; ++p) {
    const double th = 2.0 * 3.14159265358979323846 * p / count;
    const int x = (int)std::lround(cx + radius * std::cos(th));
    const int y = (int)std::lround(cy + radius * std::sin(th));
    if (x < 2 || x > nx - 3 || y < 2 || y > ny - 3) return TOMO_ERR_ARG;
    for (int q = 0; q < p; ++q)
      if (ii[q] == x && jj[q] == y) return TOMO_ERR_ARG;  // ring too small for the count
    ii[p] = x;
    jj[p] = y;
  }
  return TOMO_OK;
}

// field_setup (sirt-mimo.cu:470-509; generate-field.cu:117-172) in float
// arithmetic; powf(x, 2) taken as the correctly rounded x * x
int tomo_phantom_reference(float *fo, int nx, int ny, double h) {
  if (!fo || nx <= 0 || ny <= 0 || !(h > 0)) return TOMO_ERR_ARG;
  const float hf = (float)h;
  const float rc = 0.015f, rp = 0.005f, sc = 0.03f, sp = 0.05f;
  const float ox = rc * std::cos(-30 * (3.14159265f / 180));
  const float oy = rc * std::sin(30 * (3.14159265f / 180));
  const float rc2 = rc * rc, rp2 = rp * rp;
  for (int j = 0; j < ny; ++j) {
    const float y = -0.1f + (float)j * hf;
    for (int i = 0; i < nx; ++i) {
      const float x = -0.1f + (float)i * hf;
      float v = 0.f;
      if (x * x + y * y <= rc2) v = sc;
      const float a1 = x - ox, b1 = y - oy, a2 = x + ox, b3 = y + rc;
      if (a1 * a1 + b1 * b1 <= rp2) v = sp;
      if (a2 * a2 + b1 * b1 <= rp2) v = sp;
      if (x * x + b3 * b3 <= rp2) v = sp;
      fo[i + nx * j] = v;
    }
  }
  return TOMO_OK;
}

int tomo_phantom_inclusions(float *fo, int nx, int ny, uint64_t seed, int count, double rmin, double rmax,
                            double cmin, double cmax, int x0, int x1, int y0, int y1) {
  if (!fo || nx <= 0 || ny <= 0 || count < 0 || rmin <= 0 || rmax < rmin || cmin < 0 || cmax < cmin ||
      x0 < 0 || y0 < 0 || x1 > nx || y1 > ny || x0 >= x1 || y0 >= y1)
    return TOMO_ERR_ARG;
  Rng rng(seed);
  for (int n = 0; n < count; ++n) {
    const double r = rmin + (rmax - rmin) * rng.uniform();
    const double sign = rng.uniform() < 0.5 ? -1.0 : 1.0;
    const double con = sign * (cmin + (cmax - cmin) * rng.uniform());
    const double wx = std::max(0.0, (x1 - x0) - 2 * r), wy = std::max(0.0, (y1 - y0) - 2 * r);
    const double cx = x0 + std::min(r, 0.5 * (x1 - x0)) + wx * rng.uniform();
    const double cy = y0 + std::min(r, 0.5 * (y1 - y0)) + wy * rng.uniform();
    const float val = (float)((1.0 + con) * (1.0 + con) - 1.0);
    for (int j = std::max(0, (int)std::floor(cy - r)); j <= std::min(ny - 1, (int)std::ceil(cy + r)); ++j)
      for (int i = std::max(0, (int)std::floor(cx - r)); i <= std::min(nx - 1, (int)std::ceil(cx + r)); ++i) {
        const double dx = i - cx, dy = j - cy;
        if (dx * dx + dy * dy <= r * r) fo[i + nx * j] = val;
      }
  }
  return TOMO_OK;
}

int tomo_phantom_smooth(float *fo, int nx, int ny, uint64_t seed, double sigma, double amp, int x0, int x1,
                        int y0, int y1) {
  // a filter wider than the grid (sigma above its larger side) is refused:
  // its taps would outnumber the cells they act on
  if (!fo || nx <= 0 || ny <= 0 || !(sigma > 0) || sigma > std::max(nx, ny) || amp < 0 || x0 < 0 || y0 < 0 ||
      x1 > nx || y1 > ny || x0 >= x1 || y0 >= y1)
    return TOMO_ERR_ARG;
  Rng rng(seed);
  const size_t n2 = (size_t)nx * ny;
  std::vector<double> a(n2), b(n2);
  for (size_t i = 0; i < n2; ++i) a[i] = rng.normal();
  const int R = (int)std::ceil(4 * sigma);
  std::vector<double> k(2 * R + 1);
  double ks = 0;
  for (int d = -R; d <= R; ++d) ks += (k[d + R] = std::exp(-0.5 * d * d / (sigma * sigma)));
  for (auto &v : k) v /= ks;
  for (int j = 0; j < ny; ++j)  // along x, zero outside the grid
    for (int i = 0; i < nx; ++i) {
      double s = 0;
      for (int d = -R; d <= R; ++d) {
        const int q = i + d;
        if (q >= 0 && q < nx) s += k[d + R] * a[q + (size_t)nx * j];
      }
      b[i + (size_t)nx * j] = s;
    }
  for (int j = 0; j < ny; ++j)  // along y
    for (int i = 0; i < nx; ++i) {
      double s = 0;
      for (int d = -R; d <= R; ++d) {
        const int q = j + d;
        if (q >= 0 && q < ny) s += k[d + R] * b[i + (size_t)nx * q];
      }
      a[i + (size_t)nx * j] = s;
    }
  // taper to zero over sigma cells at the region edge, then scale the peak to amp
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      double t = 0;
      if (i >= x0 && i < x1 && j >= y0 && j < y1) {
        const double e = std::min(std::min(i - x0, x1 - 1 - i), std::min(j - y0, y1 - 1 - j));
        t = e >= sigma ? 1.0 : 0.5 * (1.0 - std::cos(3.14159265358979323846 * e / sigma));
      }
      b[i + (size_t)nx * j] = a[i + (size_t)nx * j] * t;
    }
  double peak = 0;
  for (size_t i = 0; i < n2; ++i) peak = std::max(peak, std::fabs(b[i]));
  if (!(peak > 0)) peak = 1;
  for (size_t i = 0; i < n2; ++i) {
    const double c = amp * b[i] / peak;
    fo[i] = (float)((1.0 + c) * (1.0 + c) - 1.0);
  }
  return TOMO_OK;
}

int tomo_add_noise(float *tr, int S, int nt, int nr, double rel, uint64_t seed) {
  if (!tr || S <= 0 || nt <= 0 || nr <= 0 || rel < 0) return TOMO_ERR_ARG;
  Rng rng(seed);
  for (int s = 0; s < S; ++s)
    for (int r = 0; r < nr; ++r) {
      double e = 0;
      for (int k = 0; k < nt; ++k) {
        const double v = tr[((size_t)s * nt + k) * nr + r];
        e += v * v;
      }
      const double sd = rel * std::sqrt(e / nt);
      for (int k = 0; k < nt; ++k) tr[((size_t)s * nt + k) * nr + r] += (float)(sd * rng.normal());
    }
  return TOMO_OK;
}

// relaxation schedules (sirt-mimo-io.cu:355; sirt-mimo-norm-alt.cu:472-477)
double tomo_schedule(int kind, int k, const double *p) {
  if (!p) return std::numeric_limits<double>::quiet_NaN();
  if (kind == TOMO_SCHED_POWER) return p[0] * (p[1] + p[2] / std::pow(double(k + 1), p[3]));
  const int it = k + 1;
  if (it % 2 != 0) return p[0] + p[1] / std::pow(double(it), p[2]);
  return p[3] - p[4] / std::pow(double(it), p[5]);
}

// mimo-io.cuh:22-45 (blank lines after each k and l block) or, with
// trailing_blank_lines == 0, the IO_Files layout of sirt-mimo.cu:389-425
int tomo_write_text(const char *path, const float *a, int nx, int ny, int nz, int nw, int precision,
                    int trailing_blank_lines) {
  if (!path || !a || nx <= 0 || ny <= 0 || nz <= 0 || nw <= 0) return TOMO_ERR_ARG;
  std::ofstream o(path);
  if (!o) return TOMO_ERR_IO;
  if (trailing_blank_lines) {
    o.unsetf(std::ios_base::floatfield);
    o.precision(precision > 0 ? precision : std::numeric_limits<float>::max_digits10);
  } else {
    o.precision(precision > 0 ? precision : 6);
  }
  for (int l = 0; l < nw; ++l) {
    for (int k = 0; k < nz; ++k) {
      for (int j = 0; j < ny; ++j) {
        const float *row = a + (((size_t)l * nz + k) * ny + j) * nx;
        for (int i = 0; i < nx; ++i) o << row[i] << " ";
        o << "\n";
      }
      if (trailing_blank_lines) o << "\n";
    }
    if (trailing_blank_lines) o << "\n";
  }
  return o.good() ? TOMO_OK : TOMO_ERR_IO;
}

int tomo_read_text(const char *path, float *a, int nx, int ny, int nz, int nw) {
  if (!path || !a || nx <= 0 || ny <= 0 || nz <= 0 || nw <= 0) return TOMO_ERR_ARG;
  std::ifstream in(path);
  if (!in) return TOMO_ERR_IO;
  const size_t n = (size_t)nx * ny * nz * nw;
  for (size_t i = 0; i < n; ++i)
    if (!(in >> a[i])) return TOMO_ERR_IO;
  return TOMO_OK;
}

}  // extern "C"
