// tomo_oracle.cpp -- CPU restatement of the reference's wave-tomography
// update rules.  TEST INFRASTRUCTURE ONLY: used by tests/, by the smoke
// check and by bench.py's CPU-baseline leg as the checker / baseline, never
// by the product library.
//
// Written from the reference sources read as text
// (crogers719/Tomographic-Image-Reconstruction-Algorithms; file:line below).
// Templated on the real type: T = double is the correctness oracle,
// T = float (same code) is the timed CPU baseline.  OpenMP over shots.
//
// Parity status of each function (what pins it) is listed in the header
// comment of oracle/oracle.py.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/tomo.h"

// Planted mistakes for the tests that show a check fails when the rule is
// wrong (tests/test_oracle.py builds variants with -DORC_MUTATION=n):
//  1 correlation pairs z(k+1) with Lu(k)      2 additive receiver injection
//  3 source window one level longer           4 adjoint ignores f at neighbours
//  5 Ricker with 2 pi fp in place of pi fp     6 sponge profile (d / W)^2, inverted
//  7 residual of level k+1 injected into z(k)
#ifndef ORC_MUTATION
#define ORC_MUTATION 0
#endif

namespace {

// ---------------------------------------------------------------- presets
// Constants of sirt-mimo.cu:17-35 (float literals, widened exactly).
void preset_reference(tomo_config *c) {
  std::memset(c, 0, sizeof(*c));
  c->nx = 192;
  c->ny = 192;
  c->nt = 401;
  c->wavelet = TOMO_WAVELET_REF_GAUSS_COS;
  c->h = (double)0.001f;
  c->dt = (double)3.3333e-07f;
  c->c0 = 1500.0;
  c->wl_omega = (double)7.8540e+05f;
  c->wl_tau = (double)4.0000e-06f;
  c->wl_shift = (double)8.1573e-06f;
  c->src_last_level = 24;                       // "k + 1 <= 24", sirt-mimo.cu:554
  c->sponge_width = 0;
  c->recv_mode = TOMO_RECV_LINES;
  c->record_first = 2;                          // "k > 1", sirt-mimo.cu:629
  c->line_lo = 20;
  c->line_hi = 180;                             // rows/cols 20, 180 (sirt-mimo.cu:631-647)
  c->line_first = 21;
  c->line_last = 179;                           // "i > 20 && i < 180"
  c->recv_gain = (double)0.001f * 1000.0;       // "H * 1000.f", sirt-mimo.cu:727
  c->mask_x0 = 21; c->mask_x1 = 180;            // sirt-mimo.cu:870
  c->mask_y0 = 21; c->mask_y1 = 180;
  c->checkpoint_every = 0;
  c->shots_per_pass = 0;
}

void preset_unit(tomo_config *c, int n, int nt) {
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
  c->mask_x0 = n / 4; c->mask_x1 = n - n / 4;
  c->mask_y0 = n / 4; c->mask_y1 = n - n / 4;
  c->checkpoint_every = 0;
  c->shots_per_pass = 0;
}

// ------------------------------------------------------------------ model
template <typename T>
struct Model {
  int nx, ny, nt, n2;
  bool sponge;
  std::vector<T> r;      // v dt / h, v = c0 sqrt(1 + f)      (sirt-mimo.cu:525-526)
  std::vector<T> onepf;  // 1 + f
  std::vector<T> a, ab;  // sponge factors 1/(1+g), (1-g)/(1+g); 1 outside
  std::vector<T> amp;    // source pulse per step k (0 when inactive)
  T c0, dt, h, gain;
};

// Sponge damping profile g(d) = gamma * ((W - d) / W)^2 for cell distance
// d < W from the nearest edge.  No reference counterpart.
inline double sponge_g(const tomo_config &c, int d) {
  if (c.sponge_width <= 0 || d >= c.sponge_width) return 0.0;
#if ORC_MUTATION == 6
  double x = double(d) / double(c.sponge_width);
#else
  double x = double(c.sponge_width - d) / double(c.sponge_width);
#endif
  return c.sponge_gamma * x * x;
}

template <typename T>
T pulse_at(const tomo_config &c, int k) {
  if (c.src_last_level > 0 && k + 1 > c.src_last_level + (ORC_MUTATION == 3 ? 1 : 0)) return T(0);
  if (c.wavelet == TOMO_WAVELET_REF_GAUSS_COS) {
    // t = k * DT - TT; cosf(OMEGAC * t) * expf(-(t*t) / (2 TAO^2))  (sirt-mimo.cu:555-561)
    T t = T(k) * T(c.dt) - T(c.wl_shift);
    T tau = T(c.wl_tau);
    return std::cos(T(c.wl_omega) * t) * std::exp(-(t * t) / (T(2) * tau * tau));
  }
  // Ricker
  T t = T(k) * T(c.dt) - T(c.ricker_t0);
  T pf = T(ORC_MUTATION == 5 ? 2 * M_PI : M_PI) * T(c.ricker_fp) * t;
  T x = pf * pf;
  return (T(1) - T(2) * x) * std::exp(-x);
}

template <typename T>
void build_model(const tomo_config &c, const T *f, Model<T> &m) {
  m.nx = c.nx; m.ny = c.ny; m.nt = c.nt; m.n2 = c.nx * c.ny;
  m.c0 = T(c.c0); m.dt = T(c.dt); m.h = T(c.h); m.gain = T(c.recv_gain);
  m.sponge = c.sponge_width > 0;
  m.r.resize(m.n2); m.onepf.resize(m.n2); m.a.assign(m.n2, T(1)); m.ab.assign(m.n2, T(1));
  for (int j = 0; j < c.ny; ++j)
    for (int i = 0; i < c.nx; ++i) {
      int id = i + c.nx * j;
      T v = m.c0 * std::sqrt(T(1) + f[id]);
      m.r[id] = v * m.dt / m.h;
      m.onepf[id] = T(1) + f[id];
      if (m.sponge) {
        int d = std::min(std::min(i, j), std::min(c.nx - 1 - i, c.ny - 1 - j));
        double g = sponge_g(c, d);
        m.a[id] = T(1.0 / (1.0 + g));
        m.ab[id] = T((1.0 - g) / (1.0 + g));
      }
    }
  m.amp.resize(c.nt);
  for (int k = 0; k < c.nt; ++k) m.amp[k] = pulse_at<T>(c, k);
}

// ------------------------------------------------------------ forward step
// Level k+1 from levels k, k-1 for one shot (propagation, sirt-mimo.cu:511-593).
// bbox = 0-based inclusive source rectangle {x0, x1, y0, y1}.
template <typename T>
void forward_step(const Model<T> &m, const int *bbox, const T *um, const T *u0, T *up, int k) {
  const int nx = m.nx, ny = m.ny;
  const T amp = m.amp[k];
  for (int j = 0; j < ny; ++j) {
    for (int i = 0; i < nx; ++i) {
      const int id = i + nx * j;
      const T r = m.r[id];
      const T s = T(2) - T(4) * r * r;
      T val;
      if (i != 0 && i != nx - 1 && j != 0 && j != ny - 1) {
        val = r * r * (u0[id + 1] + u0[id - 1] + u0[id - nx] + u0[id + nx]) + s * u0[id];
        val = m.a[id] * val - m.ab[id] * um[id];
        if (amp != T(0) && i >= bbox[0] && i <= bbox[1] && j >= bbox[2] && j <= bbox[3]) {
          T v = m.c0 * std::sqrt(m.onepf[id]);
          val += m.a[id] * (v * v * m.dt * m.dt * amp);
        }
      } else {
        // one-way edge rule; corners take the diagonal neighbour (sirt-mimo.cu:566-587)
        const bool top = (j == 0), bottom = (j == ny - 1), left = (i == 0), right = (i == nx - 1);
        const int ja = top ? j + 1 : bottom ? j - 1 : j;
        const int jb = top ? j + 2 : bottom ? j - 2 : j;
        const int ia = left ? i + 1 : right ? i - 1 : i;
        const int ib = left ? i + 2 : right ? i - 2 : i;
        const int ida = ia + nx * ja, idb = ib + nx * jb;
        val = (T(2) - T(2) * r - r * r) * u0[id] + T(2) * r * (T(1) + r) * u0[ida] -
              r * r * u0[idb] + (T(2) * r - T(1)) * um[id] - T(2) * r * um[ida];
      }
      up[id] = val;
    }
  }
}

// corner u = mean of the two edge neighbours (propagation_at_corners, sirt-mimo.cu:595-615)
template <typename T>
void corners_mean(T *u, int nx, int ny) {
  u[0] = T(0.5) * (u[nx] + u[1]);
  u[nx - 1] = T(0.5) * (u[nx - 2] + u[nx - 1 + nx]);
  u[nx * (ny - 1)] = T(0.5) * (u[nx * (ny - 2)] + u[1 + nx * (ny - 1)]);
  u[nx - 1 + nx * (ny - 1)] = T(0.5) * (u[nx - 2 + nx * (ny - 1)] + u[nx - 1 + nx * (ny - 2)]);
}

// Lu = clamped 5-point Laplacian / h^2, corners = mean of edge neighbours
// (laplace + laplace_corners, sirt-mimo.cu:776-831).  u must be corner-averaged.
template <typename T>
void laplacian(const T *u, T *L, int nx, int ny, T h) {
  const T h2 = h * h;
  for (int j = 0; j < ny; ++j) {
    const int ja = j > 0 ? j - 1 : j, jb = j < ny - 1 ? j + 1 : j;
    for (int i = 0; i < nx; ++i) {
      const int ia = i > 0 ? i - 1 : i, ib = i < nx - 1 ? i + 1 : i;
      L[i + nx * j] = (u[i + nx * ja] + u[i + nx * jb] + u[ia + nx * j] + u[ib + nx * j] -
                       T(4) * u[i + nx * j]) / h2;
    }
  }
  L[0] = (L[1] + L[nx]) / T(2);
  L[nx - 1] = (L[nx - 2] + L[nx - 1 + nx]) / T(2);
  L[nx * (ny - 1)] = (L[1 + nx * (ny - 1)] + L[nx * (ny - 2)]) / T(2);
  L[nx - 1 + nx * (ny - 1)] = (L[nx - 2 + nx * (ny - 1)] + L[nx - 1 + nx * (ny - 2)]) / T(2);
}

// ------------------------------------------------------------ adjoint step
// z(k) from z(k+1), z(k+2) (backpropagation1, sirt-mimo.cu:686-710), then
// residual injection, Neumann edge copy and corner mean in that order
// (backpropagation2, sirt-mimo.cu:712-774; the corner mean reads the edge
// values after the copy).  rr_k = residual row of level k, float[nr].
template <typename T>
void adjoint_step(const Model<T> &m, const T *z1, const T *z2, T *z0, int nr, const int *rcell,
                  const int *ranchor, const T *rr_k, std::vector<T> &tmp) {
  const int nx = m.nx, ny = m.ny;
  const T K = m.c0 * m.c0 * (m.dt * m.dt);
  const T h2 = m.h * m.h;
  for (int j = 1; j < ny - 1; ++j)
    for (int i = 1; i < nx - 1; ++i) {
      const int id = i + nx * j;
      T val;
      if (!m.sponge) {
        const T *q = m.onepf.data();
#if ORC_MUTATION == 4
        val = K * q[id] * (z1[id - nx] + z1[id + nx] + z1[id - 1] + z1[id + 1] - T(4) * z1[id]) / h2 +
              T(2) * z1[id] - z2[id];
#else
        val = K * (q[id - nx] * z1[id - nx] + q[id + nx] * z1[id + nx] + q[id - 1] * z1[id - 1] +
                   q[id + 1] * z1[id + 1] - T(4) * q[id] * z1[id]) / h2 +
              T(2) * z1[id] - z2[id];
#endif
      } else {
        // transpose of the damped forward operator (general form)
        auto w = [&](int c) { return m.a[c] * m.r[c] * m.r[c]; };
        val = w(id - nx) * z1[id - nx] + w(id + nx) * z1[id + nx] + w(id - 1) * z1[id - 1] +
              w(id + 1) * z1[id + 1] + m.a[id] * (T(2) - T(4) * m.r[id] * m.r[id]) * z1[id] -
              m.ab[id] * z2[id];
      }
      z0[id] = val;
    }
  // receivers: value = interior value at the anchor + gain * residual
  tmp.resize(nr);
#if ORC_MUTATION == 2
  for (int r = 0; r < nr; ++r) tmp[r] = z0[rcell[r]] + rr_k[r] * m.gain;
#else
  for (int r = 0; r < nr; ++r) tmp[r] = z0[ranchor[r]] + rr_k[r] * m.gain;
#endif
  for (int r = 0; r < nr; ++r) z0[rcell[r]] = tmp[r];
  // Neumann edges (i = 1 .. N-2)
  for (int i = 1; i < nx - 1; ++i) {
    z0[i] = z0[i + nx];
    z0[i + nx * (ny - 1)] = z0[i + nx * (ny - 2)];
  }
  for (int j = 1; j < ny - 1; ++j) {
    z0[nx * j] = z0[1 + nx * j];
    z0[nx - 1 + nx * j] = z0[nx - 2 + nx * j];
  }
  z0[0] = (z0[1] + z0[nx]) / T(2);
  z0[nx - 1] = (z0[nx - 2] + z0[nx - 1 + nx]) / T(2);
  z0[nx * (ny - 1)] = (z0[1 + nx * (ny - 1)] + z0[nx * (ny - 2)]) / T(2);
  z0[nx - 1 + nx * (ny - 1)] = (z0[nx - 2 + nx * (ny - 1)] + z0[nx - 1 + nx * (ny - 2)]) / T(2);
}

// Residual level injected into z(k): level k itself (backpropagation2 reads
// the residual of the level it writes, sirt-mimo.cu:724-740).
inline int injected_level(int k) { return ORC_MUTATION == 7 ? k + 1 : k; }

// ---------------------------------------------------------- one shot, all
// Forward of one shot: traces[nt][nr] (initial_signal, sirt-mimo.cu:617-649),
// optional history hist[nt][n2] with corners averaged.
template <typename T>
void shot_forward(const tomo_config &c, const Model<T> &m, const int *bbox, int nr,
                  const int *rcell, T *traces, T *hist) {
  const int n2 = m.n2, nt = m.nt;
  std::vector<T> ring(3 * (size_t)n2, T(0));
  auto lvl = [&](int k) -> T * { return hist ? hist + (size_t)k * n2 : ring.data() + (size_t)(k % 3) * n2; };
  std::fill(lvl(0), lvl(0) + n2, T(0));
  std::fill(lvl(1), lvl(1) + n2, T(0));
  if (traces) std::fill(traces, traces + (size_t)nt * nr, T(0));
  for (int k = 1; k < nt - 1; ++k) {
    forward_step(m, bbox, lvl(k - 1), lvl(k), lvl(k + 1), k);
    if (traces && k + 1 >= c.record_first)
      for (int r = 0; r < nr; ++r) traces[(size_t)(k + 1) * nr + r] = lvl(k + 1)[rcell[r]];
  }
  if (hist)
    for (int k = 0; k < nt; ++k) corners_mean(hist + (size_t)k * n2, m.nx, m.ny);
}

// Gradient of one shot: G[n2] = sum_k z(k) Lu(k) (update_differential,
// sirt-mimo-io.cu:750-772 -- no 1/(1+f) inside), residual rr = meas - u at the
// receivers (difference_signal, sirt-mimo.cu:651-684), misfit = sum rr^2 in FP64.
// Forward levels are rebuilt segment by segment from checkpoint pairs every C
// levels (C = checkpoint_every, or one segment when 0).
template <typename T>
void shot_gradient(const tomo_config &c, const Model<T> &m, const int *bbox, int nr,
                   const int *rcell, const int *ranchor, const T *meas, T *G, double *misfit,
                   T *coverage) {
  const int n2 = m.n2, nt = m.nt;
  const int C = (c.checkpoint_every > 0 && c.checkpoint_every < nt) ? c.checkpoint_every : nt;
  const int nseg = (nt + C - 1) / C;
  std::vector<T> ring(3 * (size_t)n2, T(0));
  std::vector<T> ckpt((size_t)nseg * 2 * n2, T(0));
  std::vector<T> rr((size_t)nt * nr, T(0));
  // pass 1: forward with residual, keep level pairs (mC, mC+1)
  auto lvl = [&](int k) -> T * {
    int q = k % C;
    if (q < 2) return ckpt.data() + ((size_t)(k / C) * 2 + q) * n2;
    return ring.data() + (size_t)(k % 3) * n2;
  };
  std::fill(lvl(0), lvl(0) + n2, T(0));
  std::fill(lvl(1), lvl(1) + n2, T(0));
  double mis = 0.0;
  for (int k = 1; k < nt - 1; ++k) {
    forward_step(m, bbox, lvl(k - 1), lvl(k), lvl(k + 1), k);
    if (k + 1 >= c.record_first)
      for (int r = 0; r < nr; ++r) {
        T d = meas[(size_t)(k + 1) * nr + r] - lvl(k + 1)[rcell[r]];
        rr[(size_t)(k + 1) * nr + r] = d;
        mis += double(d) * double(d);
      }
  }
  if (misfit) *misfit = mis;
  std::fill(G, G + n2, T(0));
  if (coverage) std::fill(coverage, coverage + n2, T(0));
  // pass 2: segments in reverse; recompute the segment's levels, then adjoint
  std::vector<T> seg((size_t)C * n2);
  std::vector<T> zr(3 * (size_t)n2, T(0));
  std::vector<T> L(n2), tmp;
  auto zl = [&](int k) -> T * { return zr.data() + (size_t)(k % 3) * n2; };
  for (int sidx = nseg - 1; sidx >= 0; --sidx) {
    const int k0 = sidx * C;
    const int cnt = std::min(C, nt - k0);
    std::memcpy(seg.data(), ckpt.data() + (size_t)sidx * 2 * n2, sizeof(T) * n2);
    if (cnt > 1) std::memcpy(seg.data() + n2, ckpt.data() + ((size_t)sidx * 2 + 1) * n2, sizeof(T) * n2);
    for (int q = 2; q < cnt; ++q) {
      const int k = k0 + q - 1;  // step producing level k0 + q
      forward_step(m, bbox, seg.data() + (size_t)(q - 2) * n2, seg.data() + (size_t)(q - 1) * n2,
                   seg.data() + (size_t)q * n2, k);
    }
    for (int q = cnt - 1; q >= 0; --q) {
      const int k = k0 + q;
      if (k < 1 || k > nt - 2) continue;
      T *u = seg.data() + (size_t)q * n2;
      corners_mean(u, m.nx, m.ny);
      adjoint_step(m, zl(k + 1), zl(k + 2), zl(k), nr, rcell, ranchor,
                   rr.data() + (size_t)injected_level(k) * nr, tmp);
      laplacian(u, L.data(), m.nx, m.ny, m.h);
      const T *z = zl(ORC_MUTATION == 1 ? k + 1 : k);
      for (int id = 0; id < n2; ++id) G[id] += z[id] * L[id];
      if (coverage)
        for (int id = 0; id < n2; ++id) coverage[id] += u[id] * u[id];
    }
  }
}

// ---------------------------------------------------------------- updates
inline bool in_mask(const tomo_config &c, int i, int j) {
  return i >= c.mask_x0 && i < c.mask_x1 && j >= c.mask_y0 && j < c.mask_y1;
}

}  // namespace

// =================================================================== C ABI
extern "C" {

void orc_config_reference(tomo_config *c) { preset_reference(c); }
void orc_config_unit(tomo_config *c, int n, int nt) { preset_unit(c, n, nt); }

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

// Phantom of field_setup (sirt-mimo.cu:470-509), float arithmetic.  powf(x, 2)
// is taken as the correctly rounded square x * x.
void orc_phantom_reference(float *fo, int nx, int ny, double h) {
  const float hf = (float)h;
  const float rc = 0.015f, rp = 0.005f, sc = 0.03f, sp = 0.05f;
  const float ca = std::cos(-30 * (3.14159265f / 180));
  const float sa = std::sin(30 * (3.14159265f / 180));
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      float x = -0.1f + i * hf, y = -0.1f + j * hf;
      float value = 0.f;
      if (x * x + y * y <= rc * rc) value = sc;
      float dx = x - rc * ca, dy = y - rc * sa;
      if (dx * dx + dy * dy <= rp * rp) value = sp;
      dx = x + rc * ca;
      if (dx * dx + dy * dy <= rp * rp) value = sp;
      dy = y + rc;
      if (x * x + dy * dy <= rp * rp) value = sp;
      fo[i + nx * j] = value;
    }
}

// Position_Transducers (sirt-mimo.cu:438-466), 1-based.
void orc_transducers_reference(int *ii, int *jj, int num) {
  for (int p = 0; p < num; ++p) {
    if (p < 160) { ii[p] = 21 + (p + 1); jj[p] = 181; }
    else if (p < 320) { ii[p] = 181; jj[p] = 181 - ((p + 1) - 160); }
    else if (p < 480) { ii[p] = 181 - ((p + 1) - 320); jj[p] = 21; }
    else { ii[p] = 21; jj[p] = 21 + ((p + 1) - 480); }
  }
}

// Group bounding boxes (sirt-mimo.cu:543-554): group g uses transducers
// g*G and g*G+G-1 (1-based), sorted per axis; the 1-based test
// "j + 1 >= jp1 && j + 1 <= jp2" is the 0-based box [jp1-1, jp2-1].
int orc_groups(const int *ii, const int *jj, int ns, int G, int *bbox) {
  if (G <= 0) return -1;
  int ng = ns / G;
  for (int g = 0; g < ng; ++g) {
    int p = g * G;
    int ip1 = ii[p], ip2 = ii[p + G - 1], jp1 = jj[p], jp2 = jj[p + G - 1];
    if (ip1 > ip2) std::swap(ip1, ip2);
    if (jp1 > jp2) std::swap(jp1, jp2);
    bbox[4 * g + 0] = ip1 - 1; bbox[4 * g + 1] = ip2 - 1;
    bbox[4 * g + 2] = jp1 - 1; bbox[4 * g + 3] = jp2 - 1;
  }
  return ng;
}

// Sensor lines (sirt-mimo.cu:629-647, 724-740): order bottom (y = hi),
// right (x = hi), top (y = lo), left (x = lo); anchor = inward neighbour.
int orc_receivers_lines(const tomo_config *c, int *rcell, int *ranchor) {
  const int nx = c->nx, lo = c->line_lo, hi = c->line_hi;
  const int L = c->line_last - c->line_first + 1;
  int r = 0;
  for (int side = 0; side < 4; ++side)
    for (int i = c->line_first; i <= c->line_last; ++i, ++r) {
      int x, y, ax, ay;
      switch (side) {
        case 0: x = i; y = hi; ax = i; ay = hi - 1; break;
        case 1: x = hi; y = i; ax = hi - 1; ay = i; break;
        case 2: x = i; y = lo; ax = i; ay = lo + 1; break;
        default: x = lo; y = i; ax = lo + 1; ay = i; break;
      }
      if (rcell) rcell[r] = x + nx * y;
      if (ranchor) ranchor[r] = ax + nx * ay;
    }
  return 4 * L;
}

}  // extern "C"

// ------------------------------------------------------- typed entry points
namespace {

template <typename T>
int forward_all(const tomo_config *c, const T *f, int nshots, const int *bbox, int nr,
                const int *rcell, T *traces, T *hist) {
  Model<T> m;
  build_model(*c, f, m);
  const size_t tr = (size_t)c->nt * nr, hs = (size_t)c->nt * m.n2;
#pragma omp parallel for schedule(dynamic, 1)
  for (int s = 0; s < nshots; ++s)
    shot_forward(*c, m, bbox + 4 * s, nr, rcell, traces ? traces + s * tr : nullptr,
                 hist ? hist + s * hs : nullptr);
  return 0;
}

template <typename T>
int gradient_all(const tomo_config *c, const T *f, int nshots, const int *bbox, int nr,
                 const int *rcell, const int *ranchor, const T *meas, T *G, double *misfit,
                 T *coverage) {
  Model<T> m;
  build_model(*c, f, m);
  const size_t tr = (size_t)c->nt * nr;
  std::vector<T> cov;
  if (coverage) cov.assign((size_t)nshots * m.n2, T(0));
#pragma omp parallel for schedule(dynamic, 1)
  for (int s = 0; s < nshots; ++s)
    shot_gradient(*c, m, bbox + 4 * s, nr, rcell, ranchor, meas + s * tr, G + (size_t)s * m.n2,
                  misfit ? misfit + s : nullptr, coverage ? cov.data() + (size_t)s * m.n2 : nullptr);
  if (coverage) {
    std::fill(coverage, coverage + m.n2, T(0));
    for (int s = 0; s < nshots; ++s)
      for (int id = 0; id < m.n2; ++id) coverage[id] += cov[(size_t)s * m.n2 + id];
  }
  return 0;
}

// uniform update (sirt-mimo.cu:858-877 with mode SUM; sirt-mimo-io.cu:774-796 with MEAN)
template <typename T>
void update_uniform(const tomo_config *c, T *f, int nshots, const T *G, double scale, int mode) {
  const int nx = c->nx, ny = c->ny, n2 = nx * ny;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      if (!in_mask(*c, i, j)) continue;
      const int id = i + nx * j;
      double s = 0.0;
      for (int g = 0; g < nshots; ++g) s += double(G[(size_t)g * n2 + id]);
      double d = s / (1.0 + double(f[id]));
      if (mode == TOMO_UPDATE_MEAN) d /= nshots;
      f[id] = T(double(f[id]) + scale * d);
    }
}

// non-uniform weighted update (sirt-mimo-norm.cu:56-118, 873-892)
template <typename T>
void update_weighted(const tomo_config *c, T *f, int M, const T *G, double scale, double *wout) {
  const int nx = c->nx, ny = c->ny, n2 = nx * ny;
  std::vector<double> mean(n2, 0.0), d2(M, 0.0), w(M);
  for (int id = 0; id < n2; ++id) {
    double s = 0.0;
    for (int m = 0; m < M; ++m) s += double(G[(size_t)m * n2 + id]) / (1.0 + double(f[id]));
    mean[id] = s / M;
  }
  for (int m = 0; m < M; ++m) {
    double acc = 0.0;
    for (int id = 0; id < n2; ++id) {
      double v = double(G[(size_t)m * n2 + id]) / (1.0 + double(f[id])) - mean[id];
      acc += v * v;
    }
    d2[m] = acc;
  }
  bool degenerate = false;
  double W = 0.0;
  for (int m = 0; m < M; ++m) {
    if (!(d2[m] > 0.0)) degenerate = true;
    w[m] = 1.0 / std::sqrt(d2[m]);
    W += w[m];
  }
  if (degenerate || !std::isfinite(W))
    for (int m = 0; m < M; ++m) w[m] = 1.0 / M;
  else
    for (int m = 0; m < M; ++m) w[m] /= W;
  if (wout)
    for (int m = 0; m < M; ++m) wout[m] = w[m];
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      if (!in_mask(*c, i, j)) continue;
      const int id = i + nx * j;
      double s = 0.0;
      for (int m = 0; m < M; ++m) s += w[m] * double(G[(size_t)m * n2 + id]) / (1.0 + double(f[id]));
      f[id] = T(double(f[id]) + scale * s);
    }
}

// Weighted average of intermediate fields (sirt-mimo-avg-weight.cu:80-203,
// 530-544), step by step as the reference runs it:
//   fg_g = f + scale mask G_g / (1 + f)           (update_intermediate_fields)
//   favg = sum_g fg_g / M                           (sum_ / average_intermediate_fields)
//   w_g  = || fg_g - favg ||_2 over the full grid   (get_normalized_weights)
//   f   += sum_g fg_g / w_g                         (update_weighted_average_field, update_field)
// The whole intermediate fields, f included, are added.  A zero w_g (one
// group, or identical groups) makes f non-finite, as in the reference.
// wout (may be NULL) receives w_g and, in wout[M], their sum (get_total_weight).
template <typename T>
void update_avg_weight(const tomo_config *c, T *f, int M, const T *G, double scale, double *wout) {
  const int nx = c->nx, ny = c->ny, n2 = nx * ny;
  auto fg = [&](int m, int id) {
    const int i = id % nx, j = id / nx;
    const double df = double(G[(size_t)m * n2 + id]) / (1.0 + double(f[id]));
    return double(f[id]) + (in_mask(*c, i, j) ? scale * df : 0.0);
  };
  std::vector<double> favg(n2, 0.0), w(M, 0.0);
  for (int id = 0; id < n2; ++id) {
    double s = 0.0;
    for (int m = 0; m < M; ++m) s += fg(m, id);
    favg[id] = s / M;
  }
  double total = 0.0;
  for (int m = 0; m < M; ++m) {
    double acc = 0.0;
    for (int id = 0; id < n2; ++id) {
      const double v = fg(m, id) - favg[id];
      acc += v * v;
    }
    w[m] = std::sqrt(acc);
    total += w[m];
  }
  if (wout) {
    for (int m = 0; m < M; ++m) wout[m] = w[m];
    wout[M] = total;
  }
  std::vector<double> add(n2, 0.0);
  for (int id = 0; id < n2; ++id)
    for (int m = 0; m < M; ++m) add[id] += fg(m, id) / w[m];
  for (int id = 0; id < n2; ++id) f[id] = T(double(f[id]) + add[id]);
}

// ART step (art-mimo-io.cu:381-400)
template <typename T>
void update_group(const tomo_config *c, T *f, const T *Gg, double scale) {
  for (int j = 0; j < c->ny; ++j)
    for (int i = 0; i < c->nx; ++i) {
      if (!in_mask(*c, i, j)) continue;
      const int id = i + c->nx * j;
      f[id] = T(double(f[id]) + scale * double(Gg[id]) / (1.0 + double(f[id])));
    }
}

template <typename T>
double error_pct(const T *f, const float *fo, int n2) {
  double a = 0.0, b = 0.0;
  for (int id = 0; id < n2; ++id) {
    double d = double(f[id]) - double(fo[id]);
    a += d * d;
    b += double(fo[id]) * double(fo[id]);
  }
  return 100.0 * std::sqrt(a) / std::sqrt(b);
}

}  // namespace

#define ORC_TYPED(SUF, T)                                                                          \
  extern "C" int orc_forward_##SUF(const tomo_config *c, const T *f, int nshots, const int *bbox,  \
                                   int nr, const int *rcell, T *traces, T *hist) {                  \
    return forward_all<T>(c, f, nshots, bbox, nr, rcell, traces, hist);                            \
  }                                                                                                \
  extern "C" int orc_gradient_##SUF(const tomo_config *c, const T *f, int nshots, const int *bbox, \
                                    int nr, const int *rcell, const int *ranchor, const T *meas,   \
                                    T *G, double *misfit, T *coverage) {                           \
    return gradient_all<T>(c, f, nshots, bbox, nr, rcell, ranchor, meas, G, misfit, coverage);     \
  }                                                                                                \
  extern "C" void orc_update_uniform_##SUF(const tomo_config *c, T *f, int nshots, const T *G,     \
                                           double scale, int mode) {                               \
    update_uniform<T>(c, f, nshots, G, scale, mode);                                               \
  }                                                                                                \
  extern "C" void orc_update_weighted_##SUF(const tomo_config *c, T *f, int nshots, const T *G,    \
                                            double scale, double *w) {                             \
    update_weighted<T>(c, f, nshots, G, scale, w);                                                 \
  }                                                                                                \
  extern "C" void orc_update_avg_weight_##SUF(const tomo_config *c, T *f, int nshots, const T *G,  \
                                              double scale, double *w) {                           \
    update_avg_weight<T>(c, f, nshots, G, scale, w);                                               \
  }                                                                                                \
  extern "C" void orc_update_group_##SUF(const tomo_config *c, T *f, const T *Gg, double scale) {  \
    update_group<T>(c, f, Gg, scale);                                                              \
  }                                                                                                \
  extern "C" double orc_error_##SUF(const T *f, const float *fo, int n2) {                         \
    return error_pct<T>(f, fo, n2);                                                                \
  }                                                                                                \
  extern "C" void orc_pulse_##SUF(const tomo_config *c, T *amp) {                                  \
    for (int k = 0; k < c->nt; ++k) amp[k] = pulse_at<T>(*c, k);                                   \
  }

ORC_TYPED(f64, double)

// Operator probes for the dot-product and Laplacian tests (FP64).
// y = interior forward operator applied to x (one leapfrog step with
// u(k-1) = 0 and no source); edge cells of y hold the edge rule.
extern "C" void orc_apply_forward_op_f64(const tomo_config *c, const double *f, const double *x,
                                         double *y) {
  Model<double> m;
  build_model(*c, f, m);
  std::vector<double> zero(m.n2, 0.0);
  const int nobox[4] = {1, 0, 1, 0};
  forward_step(m, nobox, zero.data(), x, y, c->nt);  // k = nt: past any source window
}
// x = adjoint step applied to y with z(k+2) = 0 and no receivers.
extern "C" void orc_apply_adjoint_op_f64(const tomo_config *c, const double *f, const double *y,
                                         double *x) {
  Model<double> m;
  build_model(*c, f, m);
  std::vector<double> zero(m.n2, 0.0), tmp;
  std::fill(x, x + m.n2, 0.0);
  adjoint_step<double>(m, y, zero.data(), x, 0, nullptr, nullptr, nullptr, tmp);
}
// The adjoint sweep of shot_gradient alone (levels nt-2 .. 1), driven by
// given residual traces rr[nt][nr]: out[k] = z(k) at `cell` (0 for levels
// the sweep does not form).  Probe for the adjoint-state dot-product test
// <traces, rr> = sum_k s(k) z(k+1)(source cell).
extern "C" void orc_adjoint_probe_f64(const tomo_config *c, const double *f, int nr, const int *rcell,
                                      const int *ranchor, const double *rr, int cell, double *out) {
  Model<double> m;
  build_model(*c, f, m);
  const int n2 = m.n2, nt = m.nt;
  std::vector<double> zr(3 * (size_t)n2, 0.0), tmp;
  auto zl = [&](int k) -> double * { return zr.data() + (size_t)(k % 3) * n2; };
  std::fill(out, out + nt, 0.0);
  for (int k = nt - 2; k >= 1; --k) {
    adjoint_step(m, zl(k + 1), zl(k + 2), zl(k), nr, rcell, ranchor, rr + (size_t)injected_level(k) * nr, tmp);
    out[k] = zl(k)[cell];
  }
}
extern "C" void orc_laplacian_f64(const tomo_config *c, const double *u, double *L) {
  laplacian(u, L, c->nx, c->ny, double(c->h));
}
ORC_TYPED(f32, float)

// relaxation schedules (sirt-mimo-io.cu:355; sirt-mimo-norm-alt.cu:472-477)
extern "C" double orc_schedule(int kind, int k, const double *p) {
  if (kind == TOMO_SCHED_POWER) return p[0] * (p[1] + p[2] / std::pow(double(k + 1), p[3]));
  int it = k + 1;
  if (it % 2 != 0) return p[0] + p[1] / std::pow(double(it), p[2]);
  return p[3] - p[4] / std::pow(double(it), p[5]);
}
