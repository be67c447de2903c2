// tomo_engine.hip -- context, memory plan and the engine half of the C ABI
// (include/tomo.h).  Replaces the per-executable Ultrasonic_Tomography()
// host loops of the reference (e.g. sirt-mimo-io.cu:97-417,
// sirt-mimo-norm-mem.cu:419-498, art-mimo-io.cu:631-718).
//
// Memory plan for a pass of P shots (all float, [level][P][ny][nx]):
//   full history (checkpoint_every == 0 or >= nt): nt levels, one forward;
//   checkpointing (C): level pairs (mC, mC+1) of every segment m, plus a
//     C-level segment buffer; the backward sweep recomputes each segment
//     from its pair, newest segment first, then runs the adjoint over it.
//   adjoint: a ring of 3 levels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "tomo.h"
#include "tomo_host.h"
#include "tomo_kernels.h"
#include "tomo_ctx.h"

using namespace tomo;
using namespace tomo_internal;

namespace {

int full_history(const tomo_ctx *c) {
  const int C = c->cfg.checkpoint_every;
  return C <= 0 || C >= c->nt;
}

int pass_shots(const tomo_ctx *c) {
  const int P = c->cfg.shots_per_pass;
  return (P > 0 && P < c->nshots) ? P : std::max(c->nshots, 1);
}

// Stencil kernel shape: the vectorised kernels (4 or 2 cells per lane) when
// the row length allows, with the requested fields overriding the defaults.
StencilShape stencil_shape(const tomo_ctx *c) {
  StencilShape s;
  s.vec = c->shape_req.vec;
  // 4-wide lanes when a 256-column tile is filled (or nearly), else 2-wide
  if (s.vec == 0) s.vec = (c->nx % 4 == 0 && (c->nx % 256 == 0 || c->nx > 256)) ? 4 : (c->nx % 2 == 0 ? 2 : 1);
  if (c->nx % s.vec != 0 || c->nx < 4 * s.vec) s.vec = 1;
  // vectorised kernels stream a strip of rows per lane; taller strips cut
  // the halo re-read (2/rows) but leave fewer strips to fill the GPU.
  // Measured at 512^2 x 128 shots: forward 32 rows with 2 rows of loads in
  // flight, adjoint 16 rows with 1 (its deeper queues cost occupancy).
  s.fwd_rows = c->shape_req.fwd_rows ? c->shape_req.fwd_rows
                                     : (s.vec > 1 ? (c->ny >= 512 ? 32 : c->ny >= 256 ? 16 : 8) : 4);
  // 64-row strips exist only for the three-step adjoint; the other adjoint
  // kernels take 32 then
  s.adj_rows = c->shape_req.adj_rows ? std::min(c->shape_req.adj_rows, 32) : (s.vec > 1 ? (c->ny >= 256 ? 16 : 8) : 4);
  // prefetch depth: 2 rows ahead for the one-step forward; 1 for the
  // two-step forward, whose registers already hold two levels (measured at
  // 512^2 x 128 shots: 129.2 us per launch at depth 1, 133.3 at depth 2)
  const bool fwd2 = c->time_block >= 2 && forward2_supported(c->nx, c->cfg.sponge_width);
  s.fwd_depth = c->shape_req.fwd_depth ? c->shape_req.fwd_depth : (fwd2 ? 1 : 2);
  s.adj_depth = c->shape_req.adj_depth ? c->shape_req.adj_depth : 1;
  s.rows = s.fwd_rows;
  s.depth = s.fwd_depth;
  s.ring = c->d_ring;
  s.nring = c->nring;
  return s;
}

// strip height of the blocked forward launches over passes of np shots: the
// shape's, or 8 rows when the strips would fill less than a quarter of the
// wave slots (a launch over a few shots, e.g. an ART step's single shot,
// runs on the latency of each strip's row-by-row walk, which shorter strips
// cut; fuller launches keep the taller strips and their smaller halo share)
int forward_rows(const tomo_ctx *c, int np) {
  const int r = stencil_shape(c).fwd_rows;
  if (c->shape_req.fwd_rows || r <= 8 || c->ny % 8 != 0) return r;
  const long long slots = 8LL * c->num_cus;  // two waves per SIMD
  return 4LL * np * ((c->ny + r - 1) / r) < slots ? 8 : r;
}

// strip height of the two-step adjoint: the forward's rule (its halo is two
// rows of z(k+1) per strip, its register window does not grow with it);
// measured at 512^2 x 128 shots: 32 rows 5 % faster than 16
int adjoint2_rows(const tomo_ctx *c) {
  if (c->shape_req.adj_rows) return std::min(c->shape_req.adj_rows, 32);
  return c->ny >= 512 ? 32 : c->ny >= 256 ? 16 : 8;
}

// read-ahead of the three-step adjoint (1 or 2 rows; tomo_set_shape's
// adjoint depth)
int adjoint3_depth(const tomo_ctx *c) { return c->shape_req.adj_depth == 2 ? 2 : 1; }

// strip height of the three-step adjoint (16, 32 or 64 rows) for passes of
// np shots.  64-row strips halve the strips' halo rows and start-up against
// 32-row ones, so they are taken when the pass still has a wave for every
// slot (one wave per SIMD at nx = 512, two at nx = 256); measured at 512^2 x
// 128 shots: 525.3 against 517.0 Gcell/s per iteration
int adjoint3_rows(const tomo_ctx *c, int np) {
  const int r = c->shape_req.adj_rows;
  if (r == 16 || r == 32 || r == 64) return r;
  const long long slots = 4LL * c->num_cus * (c->nx == 512 ? 1 : 2);
  if ((long long)np * ((c->ny + 63) / 64) >= slots) return 64;
  return c->ny >= 512 ? 32 : 16;
}

bool depth_valid(int d) { return d == 0 || d == 1 || d == 2 || d == 4; }

bool shape_valid(int vec, int fr, int ar) {
  if (vec != 0 && vec != 1 && vec != 2 && vec != 4) return false;
  if (vec == 1) return (fr == 0 || fr == 4 || fr == 8 || fr == 16) && (ar == 0 || ar == 4 || ar == 8);
  auto ok = [](int r) { return r == 0 || r == 4 || r == 8 || r == 16 || r == 32; };
  return ok(fr) && (ok(ar) || ar == 64);
}

hipEvent_t take_event(tomo_ctx *c) {
  if (!c->ev_free.empty()) {
    hipEvent_t e = c->ev_free.back();
    c->ev_free.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Inside a graph capture a span's events become event-record nodes of the
// graph (hipEventRecordExternal): every replay records them, so the stencil
// loops of a replayed pass are timed like eager ones.  The events then belong
// to the graph entry.
void span_begin(tomo_ctx *c, TimedSpan &s, int kind) {
  s.kind = kind;
  s.launches = 0;
  s.cells = 0;
  s.owned = false;
  s.a = s.b = nullptr;
  if (c->capturing) {
    if (!c->cap_events_ok) return;
    if (hipEventCreate(&s.a) != hipSuccess || hipEventCreate(&s.b) != hipSuccess ||
        hipEventRecordWithFlags(s.a, c->stream, hipEventRecordExternal) != hipSuccess) {
      (void)hipGetLastError();
      c->cap_events_ok = false;
      if (s.a) hipEventDestroy(s.a);
      if (s.b) hipEventDestroy(s.b);
      s.a = s.b = nullptr;
    }
    s.owned = true;
    return;
  }
  s.a = take_event(c);
  s.b = take_event(c);
  if (s.a) hipEventRecord(s.a, c->stream);
}

void span_end(tomo_ctx *c, TimedSpan &s) {
  if (c->capturing) {
    c->cap_launches[s.kind] += s.launches;
    c->cap_cells[s.kind] += s.cells;
    if (s.a && s.b) {
      if (hipEventRecordWithFlags(s.b, c->stream, hipEventRecordExternal) == hipSuccess) {
        c->cap_spans.push_back(s);
        return;
      }
      (void)hipGetLastError();
      c->cap_events_ok = false;
    }
    if (s.a) hipEventDestroy(s.a);
    if (s.b) hipEventDestroy(s.b);
    return;
  }
  if (s.b) hipEventRecord(s.b, c->stream);
  c->spans.push_back(s);
}

void destroy_spans(std::vector<TimedSpan> &v) {
  for (auto &s : v) {
    if (s.a) hipEventDestroy(s.a);
    if (s.b) hipEventDestroy(s.b);
  }
  v.clear();
}

// after a stream sync: fold recorded spans into the totals
void spans_collect(tomo_ctx *c) {
  for (auto &s : c->spans) {
    float ms = 0.f;
    if (s.a && s.b && hipEventElapsedTime(&ms, s.a, s.b) == hipSuccess) {
      c->t_ms[s.kind] += ms;
      c->t_launch[s.kind] += s.launches;
      c->t_cells[s.kind] += s.cells;
    }
    if (s.owned) continue;
    if (s.a) c->ev_free.push_back(s.a);
    if (s.b) c->ev_free.push_back(s.b);
  }
  c->spans.clear();
}

// Spans of a replayed graph wait in c->spans until the call's closing sync
// reads them; a call that failed part-way leaves them there.  They hold the
// graph's own events, so they are read before any graph's events go.
void collect_graph_spans(tomo_ctx *c) {
  for (const auto &s : c->spans)
    if (s.owned) {
      (void)hipStreamSynchronize(c->stream);
      spans_collect(c);
      return;
    }
}

void graphs_clear(tomo_ctx *c) {
  collect_graph_spans(c);
  for (auto &g : c->graphs) {
    if (g.exec) hipGraphExecDestroy(g.exec);
    destroy_spans(g.spans);
  }
  c->graphs.clear();
}

}  // namespace

namespace tomo_internal {

int upload_coefficients(tomo_ctx *c) {
  const double c0dt_h = c->cfg.c0 * c->cfg.dt / c->cfg.h;
  for (int p = 0; p < c->nph; ++p)
    HIPCHK(c, launch_coefficients<float>(c->d_f + (size_t)p * c->n2, c->d_pw + (size_t)p * c->n2, c->nx, c->ny,
                                         c0dt_h, c->cfg.sponge_width, c->d_spa, c->stream));
  c->pw64_valid = false;
  return TOMO_OK;
}

}  // namespace tomo_internal

namespace {


int ensure_pw64(tomo_ctx *c) {
  if (c->pw64_valid) return TOMO_OK;
  if (!c->d_pw64) {
    int rc = dalloc(c, &c->d_pw64, (size_t)c->nph * c->n2);
    if (rc) return rc;
  }
  const double c0dt_h = c->cfg.c0 * c->cfg.dt / c->cfg.h;
  for (int p = 0; p < c->nph; ++p)
    HIPCHK(c, launch_coefficients<double>(c->d_f + (size_t)p * c->n2, c->d_pw64 + (size_t)p * c->n2, c->nx, c->ny,
                                          c0dt_h, c->cfg.sponge_width, c->d_spa64, c->stream));
  c->pw64_valid = true;
  return TOMO_OK;
}

int validate_config(const tomo_config &g, std::string &why) {
  if (g.nx < 8 || g.ny < 8) return why = "grid must be at least 8 x 8", TOMO_ERR_ARG;
  if ((long long)g.nx * g.ny > (1LL << 30)) return why = "grid too large", TOMO_ERR_ARG;
  if (g.nt < 3) return why = "nt must be >= 3", TOMO_ERR_ARG;
  if (!(g.h > 0) || !(g.dt > 0) || !(g.c0 > 0)) return why = "h, dt, c0 must be positive", TOMO_ERR_ARG;
  if (g.wavelet != TOMO_WAVELET_REF_GAUSS_COS && g.wavelet != TOMO_WAVELET_RICKER)
    return why = "unknown wavelet", TOMO_ERR_ARG;
  if (g.sponge_width < 0 || 2 * g.sponge_width >= std::min(g.nx, g.ny) || g.sponge_width > 256)
    return why = "sponge width out of range", TOMO_ERR_ARG;
  if (g.recv_mode != TOMO_RECV_LINES && g.recv_mode != TOMO_RECV_POINTS)
    return why = "unknown receiver mode", TOMO_ERR_ARG;
  if (g.recv_mode == TOMO_RECV_LINES) {
    const int n = std::min(g.nx, g.ny);
    if (g.line_lo < 2 || g.line_hi > n - 3 || g.line_lo + 2 > g.line_hi)
      return why = "sensor lines must lie >= 2 cells inside the edge", TOMO_ERR_ARG;
    if (g.line_first < 2 || g.line_last > n - 3 || g.line_first > g.line_last ||
        (g.line_first <= g.line_lo && g.line_last >= g.line_lo) ||
        (g.line_first <= g.line_hi && g.line_last >= g.line_hi))
      return why = "sensor span must avoid the line rows", TOMO_ERR_ARG;
  }
  if (g.mask_x0 < 0 || g.mask_y0 < 0 || g.mask_x1 > g.nx || g.mask_y1 > g.ny || g.mask_x0 > g.mask_x1 ||
      g.mask_y0 > g.mask_y1)
    return why = "mask outside the grid", TOMO_ERR_ARG;
  if (g.checkpoint_every > 0 && g.checkpoint_every < 3) return why = "checkpoint_every must be 0 or >= 3", TOMO_ERR_ARG;
  if (g.shots_per_pass < 0) return why = "shots_per_pass must be >= 0", TOMO_ERR_ARG;
  if (g.record_first < 0) return why = "record_first must be >= 0", TOMO_ERR_ARG;
  return TOMO_OK;
}

int upload_receivers(tomo_ctx *c) {
  std::vector<int> map((size_t)c->n2, -1);
  for (size_t r = 0; r < c->rcell.size(); ++r) map[c->rcell[r]] = (int)r;
  HIPCHK(c, hipMemcpyAsync(c->d_cellmap, map.data(), sizeof(int) * map.size(), hipMemcpyHostToDevice, c->stream));
  // cells of the edge/receiver pass of the vectorised kernels: the top and
  // bottom rows (the side edges of the other rows are done in the main
  // kernels), then the receivers
  std::vector<int2> ring;
  for (int x = 0; x < c->nx; ++x) ring.push_back(make_int2(x, -1));
  for (int x = 0; x < c->nx; ++x) ring.push_back(make_int2((c->ny - 1) * c->nx + x, -1));
  std::vector<unsigned char> rowflag(c->ny, 0);
  for (size_t r = 0; r < c->rcell.size(); ++r) {
    ring.push_back(make_int2(c->rcell[r], (int)r));
    rowflag[c->rcell[r] / c->nx] = 1;
  }
  int rc = dalloc(c, &c->d_ring, ring.size());
  if (rc) return rc;
  if ((rc = dalloc(c, &c->d_rowflag, (size_t)c->ny))) return rc;
  c->nring = (int)ring.size();
  HIPCHK(c, hipMemcpyAsync(c->d_ring, ring.data(), sizeof(int2) * ring.size(), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_rowflag, rowflag.data(), c->ny, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->recvseg_ok = false;
  // host source of an asynchronous copy: lives until the synchronisation below
  std::vector<int> seg;
  if (adjoint2_supported(c->nx, c->cfg.sponge_width)) {
    const int vec = c->nx / 64;
    seg.assign((size_t)c->ny * 64, -1);
    bool ok = true;
    for (size_t r = 0; r < c->rcell.size(); ++r) {
      const int x = c->rcell[r] % c->nx, y = c->rcell[r] / c->nx;
      int &d = seg[(size_t)y * 64 + x / vec];
      if (d >= 0) ok = false;
      d = (int)r * vec + x % vec;
    }
    if ((rc = dalloc(c, &c->d_recvseg, seg.size()))) return rc;
    HIPCHK(c, hipMemcpyAsync(c->d_recvseg, seg.data(), sizeof(int) * seg.size(), hipMemcpyHostToDevice, c->stream));
    c->recvseg_ok = ok;
  }
  if ((rc = dalloc(c, &c->d_anchor, std::max<size_t>(c->ranchor.size(), 1)))) return rc;
  if (!c->ranchor.empty())
    HIPCHK(c, hipMemcpyAsync(c->d_anchor, c->ranchor.data(), sizeof(int) * c->ranchor.size(),
                             hipMemcpyHostToDevice, c->stream));
  c->nr = (int)c->rcell.size();
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TOMO_OK;
}

// (re)allocate per-shot buffers after the shot or receiver count changed
int alloc_shot_buffers(tomo_ctx *c) {
  const size_t S = (size_t)std::max(c->nshots, 1), tr = (size_t)c->nt * std::max(c->nr, 1);
  int rc;
  if ((rc = dalloc(c, &c->d_meas, S * tr))) return rc;
  if ((rc = dalloc(c, &c->d_rr, S * tr))) return rc;
  if ((rc = dalloc(c, &c->d_grad, S * (size_t)c->n2))) return rc;
  if ((rc = dalloc(c, &c->d_misfit, S))) return rc;
  if (c->coverage && (rc = dalloc(c, &c->d_cov, S * (size_t)c->n2))) return rc;
  HIPCHK(c, launch_fill_f32(c->d_rr, S * tr, 0.f, c->stream));
  HIPCHK(c, launch_fill_f32(c->d_grad, S * (size_t)c->n2, 0.f, c->stream));
  c->have_meas = false;
  c->grad_valid.assign(S, 0);
  return TOMO_OK;
}

}  // namespace

namespace tomo_internal {

int ensure_scratch(tomo_ctx *c, size_t n) {
  if (c->scratch_n >= n) return TOMO_OK;
  int rc = dalloc(c, &c->d_scratch, n);
  if (rc) return rc;
  c->scratch_n = n;
  return TOMO_OK;
}

}  // namespace tomo_internal

namespace {


// wavefield storage for passes of P shots
int ensure_wavefields(tomo_ctx *c, int P, bool need_history) {
  const size_t lvl = (size_t)P * c->n2;
  size_t seg_levels, ckpt_levels;
  if (!need_history) {
    seg_levels = 5;  // rolling levels of a forward pass (ring_levels: at most 5)
    ckpt_levels = 0;
  } else if (full_history(c)) {
    seg_levels = (size_t)c->nt;
    ckpt_levels = 0;
  } else {
    const int C = c->cfg.checkpoint_every;
    seg_levels = (size_t)C;
    ckpt_levels = 2 * (size_t)((c->nt + C - 1) / C);
  }
  // adjoint levels: a ring of 5 (three-step launches write z(k-2) while
  // other strips still read z(k+2) on their halo rows)
  const size_t z_levels = need_history ? kAdjRing : 0;
  // the storage grows, never shrinks: a shared workspace keeps the largest
  // need of its contexts, and each context views it
  tomo_workspace own{};
  tomo_workspace *w = c->ws;
  if (!w) {
    own.d_ckpt = c->d_ckpt; own.d_seg = c->d_seg; own.d_z = c->d_z;
    own.ckpt_n = c->ckpt_n; own.seg_n = c->seg_n; own.z_n = c->z_n;
    w = &own;
  }
  int rc = TOMO_OK;
  if (w->seg_n < seg_levels * lvl) {
    if ((rc = dalloc(c, &w->d_seg, seg_levels * lvl)) == TOMO_OK) w->seg_n = seg_levels * lvl;
    else w->seg_n = 0;
  }
  if (!rc && w->ckpt_n < ckpt_levels * lvl) {
    if ((rc = dalloc(c, &w->d_ckpt, ckpt_levels * lvl)) == TOMO_OK) w->ckpt_n = ckpt_levels * lvl;
    else w->ckpt_n = 0;
  }
  if (!rc && w->z_n < z_levels * lvl) {
    if ((rc = dalloc(c, &w->d_z, z_levels * lvl)) == TOMO_OK) w->z_n = z_levels * lvl;
    else w->z_n = 0;
  }
  c->d_ckpt = w->d_ckpt; c->d_seg = w->d_seg; c->d_z = w->d_z;
  c->ckpt_n = w->ckpt_n; c->seg_n = w->seg_n; c->z_n = w->z_n;
  return rc;
}

// two forward steps per launch in the gradient passes: the sponge preset,
// the vectorised kernels, full-row waves, and room for a 4-level ring
bool two_step(const tomo_ctx *c) {
  if (c->time_block < 2 || !forward2_supported(c->nx, c->cfg.sponge_width)) return false;
  if (stencil_shape(c).vec < 2) return false;
  return full_history(c) || c->cfg.checkpoint_every >= 4;
}

// three forward steps per launch: the two-step conditions, strips that tile
// the rows exactly, and checkpoint segments of at least 5 levels
bool three_step(const tomo_ctx *c) {
  if (c->time_block < 3 || !two_step(c)) return false;
  if (!forward3_supported(c->nx, c->ny, c->cfg.sponge_width, stencil_shape(c).fwd_rows)) return false;
  return full_history(c) || c->cfg.checkpoint_every >= 5;
}

// rolling levels of a forward pass: a launch of m steps writes levels k+1 ..
// k+m while other strips still read levels k-1 and k on their halo rows, so
// the ring holds m + 2 levels
int ring_levels(const tomo_ctx *c) { return three_step(c) ? 5 : two_step(c) ? 4 : 3; }

// two adjoint steps per launch: with two or three forward steps per launch,
// point receivers with additive injection and at most one receiver per lane
// segment (coverage, when on, is summed per checkpoint segment by its own
// kernel)
bool two_step_adjoint(const tomo_ctx *c) {
  return c->adj_block >= 2 && two_step(c) && adjoint2_supported(c->nx, c->cfg.sponge_width) && c->recvseg_ok &&
         c->cfg.recv_mode == TOMO_RECV_POINTS;
}

// three adjoint steps per launch: the two-step conditions and a grid the
// three-step kernel covers
bool three_step_adjoint(const tomo_ctx *c) {
  return c->adj_block >= 3 && two_step_adjoint(c) && adjoint3_supported(c->nx, c->cfg.sponge_width);
}

// location of forward level k during the first (recording) pass (the
// rolling levels use a ring of ring_levels(c))
float *pass1_level(tomo_ctx *c, int P, int k) {
  const size_t lvl = (size_t)P * c->n2;
  if (full_history(c)) return c->d_seg + (size_t)k * lvl;
  const int C = c->cfg.checkpoint_every;
  const int q = k % C;
  if (q < 2) return c->d_ckpt + ((size_t)(k / C) * 2 + q) * lvl;
  return c->d_seg + (size_t)(k % ring_levels(c)) * lvl;
}

// location of level k while segment m is being replayed
float *segment_level(tomo_ctx *c, int P, int m, int k) {
  const size_t lvl = (size_t)P * c->n2;
  if (full_history(c)) return c->d_seg + (size_t)k * lvl;
  const int C = c->cfg.checkpoint_every;
  const int q = k - m * C;
  if (q < 2) return c->d_ckpt + ((size_t)m * 2 + q) * lvl;
  return c->d_seg + (size_t)q * lvl;
}

template <typename T>
FwdArgs<T> fwd_args(tomo_ctx *c, int shot0, const T *pw, const T *spa, const T *spab) {
  FwdArgs<T> a{};
  a.sstride = c->n2;
  a.pw = pw;
  a.f = c->d_f;
  a.cellmap = c->d_cellmap;
  a.bbox = c->d_bbox;
  a.nx = c->nx;
  a.ny = c->ny;
  a.nt = c->nt;
  a.nr = c->nr;
  a.shot0 = shot0;
  a.c0dt_h = T(c->cfg.c0 * c->cfg.dt / c->cfg.h);
  a.sponge_w = c->cfg.sponge_width;
  a.sp_a = spa;
  a.sp_ab = spab;
  a.spp = c->nph > 1 ? c->groups : 0;
  a.pstride = c->n2;
  return a;
}

template <typename T>
T pulse_h2(const tomo_ctx *c, int k);
template <>
float pulse_h2<float>(const tomo_ctx *c, int k) {
  return c->amp32[k] * (float)(c->cfg.h * c->cfg.h);
}
template <>
double pulse_h2<double>(const tomo_ctx *c, int k) {
  return c->amp64[k] * (c->cfg.h * c->cfg.h);
}

// One forward step k (levels k-1, k -> k+1) for P shots from shot0.
template <typename T>
int forward_step(tomo_ctx *c, FwdArgs<T> a, const T *um, const T *u0, T *up, int k, int P, bool record,
                 const float *meas, float *rec) {
  a.um = um;
  a.u0 = u0;
  a.up = up;
  a.level = k + 1;
  a.record = record && (k + 1 >= c->cfg.record_first) && rec != nullptr && c->nr > 0;
  a.meas = meas;
  a.rec = rec;
  a.amp_h2 = pulse_h2<T>(c, k);
  StencilShape sh = stencil_shape(c);
  if (!std::is_same<T, float>::value) sh.vec = 1;  // the FP64 solver runs the scalar kernels
  if (sh.vec == 1 && sh.fwd_rows < 4) sh.fwd_rows = 4;
  sh.rows = sh.fwd_rows;
  HIPCHK(c, launch_forward_step<T>(a, P, sh, c->stream));
  return TOMO_OK;
}

// Forward steps k and k+1 (levels k-1, k -> k+1, k+2) in one launch, edge
// rows included; the receiver records of both levels in the same launch when
// every receiver has its own lane segment (else one small pass after).
int forward_step2(tomo_ctx *c, FwdArgs<float> a, const float *um, const float *u0, float *up1, float *up2, int k,
                  int P, bool record, const float *meas, float *rec) {
  const bool rec_on = record && rec != nullptr && c->nr > 0;
  Fwd2Args b{};
  b.um = um;
  b.u0 = u0;
  b.up1 = up1;
  b.up2 = up2;
  b.sstride = a.sstride;
  b.pw = a.pw;
  b.f = a.f;
  b.spp = a.spp;
  b.pstride = a.pstride;
  b.bbox = a.bbox;
  b.nx = c->nx;
  b.ny = c->ny;
  b.shot0 = a.shot0;
  b.amp1 = pulse_h2<float>(c, k);
  b.amp2 = pulse_h2<float>(c, k + 1);
  b.c0dt_h = a.c0dt_h;
  b.sponge_w = a.sponge_w;
  b.sp_a = a.sp_a;
  b.sp_ab = a.sp_ab;
  b.rec1 = rec_on && k + 1 >= c->cfg.record_first;
  b.rec2 = rec_on && k + 2 >= c->cfg.record_first;
  const bool in_kernel = (b.rec1 || b.rec2) && c->recvseg_ok;
  if (in_kernel) {
    b.recv_seg = c->d_recvseg;
    b.row_has_recv = c->d_rowflag;
    b.meas = meas;
    b.rec = rec;
    b.nt = c->nt;
    b.nr = c->nr;
    b.level1 = k + 1;
  }
  const StencilShape sh = stencil_shape(c);
  HIPCHK(c, launch_forward2(b, P, forward_rows(c, P), sh.fwd_depth, c->stream));
  if ((b.rec1 || b.rec2) && !in_kernel) {  // records over the receivers
    Fwd2EdgeArgs e{};
    e.u0 = u0;
    e.up1 = up1;
    e.up2 = up2;
    e.sstride = a.sstride;
    e.f = a.f;
    e.spp = a.spp;
    e.pstride = a.pstride;
    e.ring = c->d_ring + 2 * c->nx;
    e.nring = c->nr;
    e.nx = c->nx;
    e.ny = c->ny;
    e.nt = c->nt;
    e.nr = c->nr;
    e.shot0 = a.shot0;
    e.level1 = k + 1;
    e.rec1 = b.rec1;
    e.rec2 = b.rec2;
    e.meas = meas;
    e.rec = rec;
    e.c0dt_h = a.c0dt_h;
    HIPCHK(c, launch_forward2_edges(e, P, c->stream));
  }
  return TOMO_OK;
}

// Forward steps k, k+1, k+2 (levels k-1, k -> k+1, k+2, k+3) in one launch,
// edge rows included; level k+1 is stored when store1 is set.  Receiver
// records in the same launch when every receiver has its own lane segment,
// else (level k+1 then stored) in small passes after.
int forward_step3(tomo_ctx *c, FwdArgs<float> a, const float *um, const float *u0, float *up1, float *up2,
                  float *up3, int k, int P, bool record, const float *meas, float *rec, bool store1) {
  const bool rec_on = record && rec != nullptr && c->nr > 0;
  Fwd3Args t{};
  Fwd2Args &b = t.b;
  b.um = um;
  b.u0 = u0;
  b.up1 = up1;
  b.up2 = up2;
  b.sstride = a.sstride;
  b.pw = a.pw;
  b.f = a.f;
  b.spp = a.spp;
  b.pstride = a.pstride;
  b.bbox = a.bbox;
  b.nx = c->nx;
  b.ny = c->ny;
  b.shot0 = a.shot0;
  b.amp1 = pulse_h2<float>(c, k);
  b.amp2 = pulse_h2<float>(c, k + 1);
  t.amp3 = pulse_h2<float>(c, k + 2);
  b.c0dt_h = a.c0dt_h;
  b.sponge_w = a.sponge_w;
  b.sp_a = a.sp_a;
  b.sp_ab = a.sp_ab;
  t.up3 = up3;
  b.rec1 = rec_on && k + 1 >= c->cfg.record_first;
  b.rec2 = rec_on && k + 2 >= c->cfg.record_first;
  t.rec3 = rec_on && k + 3 >= c->cfg.record_first;
  const bool any = b.rec1 || b.rec2 || t.rec3;
  const bool in_kernel = any && c->recvseg_ok;
  t.store1 = store1 || (any && !in_kernel);
  if (in_kernel) {
    b.recv_seg = c->d_recvseg;
    b.row_has_recv = c->d_rowflag;
    b.meas = meas;
    b.rec = rec;
    b.nt = c->nt;
    b.nr = c->nr;
    b.level1 = k + 1;
  }
  HIPCHK(c, launch_forward3(t, P, forward_rows(c, P), c->stream));
  if (any && !in_kernel) {  // records over the receivers: levels k+1, k+2, then k+3
    Fwd2EdgeArgs e{};
    e.sstride = a.sstride;
    e.f = a.f;
    e.spp = a.spp;
    e.pstride = a.pstride;
    e.ring = c->d_ring + 2 * c->nx;
    e.nring = c->nr;
    e.nx = c->nx;
    e.ny = c->ny;
    e.nt = c->nt;
    e.nr = c->nr;
    e.shot0 = a.shot0;
    e.meas = meas;
    e.rec = rec;
    e.c0dt_h = a.c0dt_h;
    e.u0 = u0;
    e.up1 = up1;
    e.up2 = up2;
    e.level1 = k + 1;
    e.rec1 = b.rec1;
    e.rec2 = b.rec2;
    HIPCHK(c, launch_forward2_edges(e, P, c->stream));
    e.up1 = up3;
    e.up2 = up3;
    e.level1 = k + 3;
    e.rec1 = t.rec3;
    e.rec2 = 0;
    HIPCHK(c, launch_forward2_edges(e, P, c->stream));
  }
  return TOMO_OK;
}

// traces of shots [g0, g0+ng) into d_rr (as u values, or the residuals
// meas - u when meas is given), float or double solver
template <typename T>
int forward_traces(tomo_ctx *c, int g0, int ng, T *ring, size_t ring_lvl, const T *pw, const T *spa,
                   const T *spab, const float *meas = nullptr) {
  const int P = std::min(pass_shots(c), ng);
  const size_t tr = (size_t)c->nt * c->nr;
  for (int p0 = g0; p0 < g0 + ng; p0 += P) {
    const int np = std::min(P, g0 + ng - p0);
    HIPCHK(c, launch_fill_f32(c->d_rr + (size_t)p0 * tr, (size_t)np * tr, 0.f, c->stream));
    // FP32 with m steps per launch: a ring of m + 2 levels (levels k+1 .. k+m
    // are written while other strips still read levels k-1, k on their halo
    // rows), as ring_levels
    bool two = false, three = false;
    if constexpr (std::is_same<T, float>::value) {
      two = two_step(c);
      three = three_step(c);
    }
    const int R = three ? 5 : two ? 4 : 3;
    auto L = [&](int k) { return ring + (size_t)(k % R) * ring_lvl; };
    if (std::is_same<T, float>::value) {
      HIPCHK(c, launch_fill_f32((float *)L(0), (size_t)np * c->n2, 0.f, c->stream));
      HIPCHK(c, launch_fill_f32((float *)L(1), (size_t)np * c->n2, 0.f, c->stream));
    } else {
      HIPCHK(c, launch_fill_f64((double *)L(0), (size_t)np * c->n2, 0.0, c->stream));
      HIPCHK(c, launch_fill_f64((double *)L(1), (size_t)np * c->n2, 0.0, c->stream));
    }
    FwdArgs<T> a = fwd_args<T>(c, p0, pw, spa, spab);
    TimedSpan s;
    span_begin(c, s, 0);
    for (int k = 1; k < c->nt - 1;) {
      int rc;
      if constexpr (std::is_same<T, float>::value) {
        if (three && k + 2 < c->nt - 1) {
          rc = forward_step3(c, a, L(k - 1), L(k), L(k + 1), L(k + 2), L(k + 3), k, np, true, meas, c->d_rr,
                             false);
          if (rc) return rc;
          k += 3;
          continue;
        }
        if (two && k + 1 < c->nt - 1) {
          rc = forward_step2(c, a, L(k - 1), L(k), L(k + 1), L(k + 2), k, np, true, meas, c->d_rr);
          if (rc) return rc;
          k += 2;
          continue;
        }
      }
      rc = forward_step<T>(c, a, L(k - 1), L(k), L(k + 1), k, np, true, meas, c->d_rr);
      if (rc) return rc;
      ++k;
    }
    s.launches = c->nt - 2;
    s.cells = (double)(c->nt - 2) * np * c->n2;
    span_end(c, s);
  }
  return TOMO_OK;
}

AdjArgs adj_args(tomo_ctx *c, int shot0) {
  AdjArgs a{};
  a.sstride = c->n2;
  a.pw = c->d_pw;
  a.cellmap = c->d_cellmap;
  a.anchor = c->d_anchor;
  a.rr = c->d_rr;
  a.acc = c->d_grad;
  a.cov = c->coverage ? c->d_cov : nullptr;
  a.nx = c->nx;
  a.ny = c->ny;
  a.nt = c->nt;
  a.nr = c->nr;
  a.shot0 = shot0;
  a.gain = (float)c->cfg.recv_gain;
  a.inv_h2 = (float)(1.0 / (c->cfg.h * c->cfg.h));
  a.sponge_w = c->cfg.sponge_width;
  a.sp_a = c->d_spa;
  a.sp_ab = c->d_spab;
  a.row_has_recv = c->d_rowflag;
  a.spp = c->nph > 1 ? c->groups : 0;
  a.pstride = c->n2;
  return a;
}

Adj2Args adj2_args(tomo_ctx *c, int shot0) {
  Adj2Args a{};
  a.sstride = c->n2;
  a.pw = c->d_pw;
  a.recv_seg = c->d_recvseg;
  a.rr = c->d_rr;
  a.acc = c->d_grad;
  a.nx = c->nx;
  a.ny = c->ny;
  a.nt = c->nt;
  a.nr = c->nr;
  a.shot0 = shot0;
  a.gain = (float)c->cfg.recv_gain;
  a.inv_h2 = (float)(1.0 / (c->cfg.h * c->cfg.h));
  a.row_has_recv = c->d_rowflag;
  a.sponge_w = c->cfg.sponge_width;
  a.sp_a = c->d_spa;
  a.sp_ab = c->d_spab;
  a.spp = c->nph > 1 ? c->groups : 0;
  a.pstride = c->n2;
  return a;
}

// Gradient of shots [p0, p0+np) in one pass (np <= P).
int gradient_pass(tomo_ctx *c, int p0, int np, int P) {
  const size_t tr = (size_t)c->nt * c->nr;
  const size_t lvl = (size_t)P * c->n2;
  const int nt = c->nt;
  HIPCHK(c, launch_fill_f32(c->d_rr + (size_t)p0 * tr, (size_t)np * tr, 0.f, c->stream));
  // pass 1: forward with the residual meas - u recorded at the receivers
  HIPCHK(c, launch_fill_f32(pass1_level(c, P, 0), (size_t)np * c->n2, 0.f, c->stream));
  HIPCHK(c, launch_fill_f32(pass1_level(c, P, 1), (size_t)np * c->n2, 0.f, c->stream));
  FwdArgs<float> fa = fwd_args<float>(c, p0, c->d_pw, c->d_spa, c->d_spab);
  {
    TimedSpan s;
    span_begin(c, s, 0);
    const bool two = two_step(c), three = three_step(c);
    const int C = full_history(c) ? 0 : c->cfg.checkpoint_every;
    for (int k = 1; k < nt - 1;) {
      int rc;
      if (three && k + 2 < nt - 1) {
        // level k+1 stays in registers unless the history keeps it
        const bool store1 = C == 0 || (k + 1) % C < 2;
        rc = forward_step3(c, fa, pass1_level(c, P, k - 1), pass1_level(c, P, k),
                           pass1_level(c, P, k + 1), pass1_level(c, P, k + 2),
                           pass1_level(c, P, k + 3), k, np, true, c->d_meas, c->d_rr, store1);
        k += 3;
      } else if (two && k + 1 < nt - 1) {
        rc = forward_step2(c, fa, pass1_level(c, P, k - 1), pass1_level(c, P, k),
                           pass1_level(c, P, k + 1), pass1_level(c, P, k + 2), k, np, true, c->d_meas,
                           c->d_rr);
        k += 2;
      } else {
        rc = forward_step<float>(c, fa, pass1_level(c, P, k - 1), pass1_level(c, P, k),
                                 pass1_level(c, P, k + 1), k, np, true, c->d_meas, c->d_rr);
        k += 1;
      }
      if (rc) return rc;
    }
    s.launches = nt - 2;
    s.cells = (double)(nt - 2) * np * c->n2;
    span_end(c, s);
  }
  HIPCHK(c, launch_misfit(c->d_rr + (size_t)p0 * tr, (long long)tr, np, c->d_misfit + p0, c->stream));
  // pass 2: adjoint sweep, newest segment first
  HIPCHK(c, launch_fill_f32(c->d_grad + (size_t)p0 * c->n2, (size_t)np * c->n2, 0.f, c->stream));
  if (c->coverage) HIPCHK(c, launch_fill_f32(c->d_cov + (size_t)p0 * c->n2, (size_t)np * c->n2, 0.f, c->stream));
  auto Z = [&](int k) { return c->d_z + (size_t)(k % kAdjRing) * lvl; };
  HIPCHK(c, launch_fill_f32(Z(nt - 1), (size_t)np * c->n2, 0.f, c->stream));
  HIPCHK(c, launch_fill_f32(Z(nt), (size_t)np * c->n2, 0.f, c->stream));
  const bool full = full_history(c);
  const int C = full ? nt : c->cfg.checkpoint_every;
  const int nseg = (nt + C - 1) / C;
  AdjArgs aa = adj_args(c, p0);
  StencilShape ash = stencil_shape(c);
  ash.rows = ash.adj_rows;
  ash.depth = ash.adj_depth;
  const bool two_adj = two_step_adjoint(c), three_adj = three_step_adjoint(c);
  // with two-step adjoint launches the coverage is summed per segment by its
  // own kernel, so the one-step tails of those passes carry none
  const bool seg_cov = two_adj && c->coverage;
  if (seg_cov) aa.cov = nullptr;
  for (int m = nseg - 1; m >= 0; --m) {
    const int k0 = m * C;
    const int cnt = std::min(C, nt - k0);
    if (!full && cnt > 2) {
      TimedSpan s;
      span_begin(c, s, 0);
      const bool two = two_step(c), three = three_step(c);
      for (int q = 2; q < cnt;) {
        const int k = k0 + q - 1;
        int rc;
        if (three && q + 2 < cnt) {
          rc = forward_step3(c, fa, segment_level(c, P, m, k - 1), segment_level(c, P, m, k),
                             segment_level(c, P, m, k + 1), segment_level(c, P, m, k + 2),
                             segment_level(c, P, m, k + 3), k, np, false, nullptr, nullptr, true);
          q += 3;
        } else if (two && q + 1 < cnt) {
          rc = forward_step2(c, fa, segment_level(c, P, m, k - 1), segment_level(c, P, m, k),
                             segment_level(c, P, m, k + 1), segment_level(c, P, m, k + 2), k, np, false, nullptr,
                             nullptr);
          q += 2;
        } else {
          rc = forward_step<float>(c, fa, segment_level(c, P, m, k - 1), segment_level(c, P, m, k),
                                   segment_level(c, P, m, k + 1), k, np, false, nullptr, nullptr);
          q += 1;
        }
        if (rc) return rc;
      }
      s.launches = cnt - 2;
      s.cells = (double)(cnt - 2) * np * c->n2;
      span_end(c, s);
    }
    TimedSpan s;
    span_begin(c, s, 1);
    int launches = 0;
    for (int q = cnt - 1; q >= 0;) {
      const int k = k0 + q;
      if (k < 1 || k > nt - 2) {
        --q;
        continue;
      }
      if (three_adj && q >= 2 && k - 2 >= 1) {  // levels k, k-1 and k-2 of this segment
        Adj3Args t{};
        t.b = adj2_args(c, p0);
        t.b.z1 = Z(k + 1);
        t.b.z2 = Z(k + 2);
        t.b.zm = Z(k - 1);
        t.zm2 = Z(k - 2);
        t.b.u0 = segment_level(c, P, m, k);
        t.b.um = segment_level(c, P, m, k - 1);
        t.um2 = segment_level(c, P, m, k - 2);
        t.b.level = k;
        HIPCHK(c, launch_adjoint3(t, np, adjoint3_rows(c, np), adjoint3_depth(c), c->stream));
        launches += 3;
        q -= 3;
        continue;
      }
      if (two_adj && q >= 1 && k - 1 >= 1) {  // levels k and k-1 of this segment
        Adj2Args b = adj2_args(c, p0);
        b.z1 = Z(k + 1);
        b.z2 = Z(k + 2);
        b.z0 = Z(k);
        b.zm = Z(k - 1);
        b.u0 = segment_level(c, P, m, k);
        b.um = segment_level(c, P, m, k - 1);
        b.level = k;
        HIPCHK(c, launch_adjoint2(b, np, adjoint2_rows(c), c->stream));
        launches += 2;
        q -= 2;
        continue;
      }
      aa.z1 = Z(k + 1);
      aa.z2 = Z(k + 2);
      aa.z0 = Z(k);
      aa.u = segment_level(c, P, m, k);
      aa.level = k;
      HIPCHK(c, launch_adjoint_step(aa, np, ash, c->stream));
      ++launches;
      --q;
    }
    s.launches = launches;
    s.cells = (double)launches * np * c->n2;
    span_end(c, s);
    if (seg_cov) {  // this segment's levels, newest first, as the one-step sweep adds them
      CovSegArgs cv{};
      cv.lev01[0] = segment_level(c, P, m, k0);
      cv.lev01[1] = cnt > 1 ? segment_level(c, P, m, k0 + 1) : cv.lev01[0];
      cv.seg = c->d_seg;
      cv.lstride = (long long)lvl;
      cv.sstride = (long long)c->n2;
      cv.cov = c->d_cov;
      cv.nx = c->nx;
      cv.ny = c->ny;
      cv.n2 = (int)c->n2;
      cv.nshots = np;
      cv.shot0 = p0;
      cv.k0 = k0;
      cv.klo = std::max(k0, 1);
      cv.khi = std::min(k0 + cnt - 1, nt - 2);
      HIPCHK(c, launch_coverage_segment(cv, c->stream));
    }
  }
  return TOMO_OK;
}

// every pointer and parameter a captured gradient pass depends on
std::vector<uintptr_t> graph_signature(const tomo_ctx *c) {
  const StencilShape sh = stencil_shape(c);
  const uintptr_t v[] = {(uintptr_t)c->d_ckpt, (uintptr_t)c->d_seg, (uintptr_t)c->d_z, (uintptr_t)c->d_rr,
                         (uintptr_t)c->d_meas, (uintptr_t)c->d_pw, (uintptr_t)c->d_f, (uintptr_t)c->d_grad,
                         (uintptr_t)c->d_cov, (uintptr_t)c->d_misfit, (uintptr_t)c->d_ring, (uintptr_t)c->nring,
                         (uintptr_t)c->d_bbox, (uintptr_t)c->d_cellmap, (uintptr_t)c->d_anchor,
                         (uintptr_t)c->d_rowflag, (uintptr_t)c->d_spa, (uintptr_t)c->d_spab,
                         (uintptr_t)c->coverage, (uintptr_t)c->nshots, (uintptr_t)c->nr, (uintptr_t)sh.vec,
                         (uintptr_t)sh.fwd_rows, (uintptr_t)sh.adj_rows, (uintptr_t)sh.fwd_depth,
                         (uintptr_t)sh.adj_depth, (uintptr_t)c->cfg.checkpoint_every, (uintptr_t)two_step(c), (uintptr_t)three_step(c),
                         (uintptr_t)two_step_adjoint(c), (uintptr_t)three_step_adjoint(c),
                         (uintptr_t)c->shape_req.adj_rows,
                         (uintptr_t)c->d_recvseg, (uintptr_t)c->nph};
  return std::vector<uintptr_t>(v, v + sizeof(v) / sizeof(v[0]));
}

// gradient pass through a captured graph: captured on first use (and again
// whenever a pointer or parameter it baked in changed), then replayed
int gradient_pass_graph(tomo_ctx *c, int p0, int np, int P) {
  const std::vector<uintptr_t> sig = graph_signature(c);
  GraphEntry *ge = nullptr;
  for (auto &g : c->graphs)
    if (g.p0 == p0 && g.np == np && g.P == P) ge = &g;
  if (ge && ge->sig != sig) {
    collect_graph_spans(c);
    hipGraphExecDestroy(ge->exec);
    ge->exec = nullptr;
    destroy_spans(ge->spans);
  }
  if (!ge || !ge->exec) {
    hipGraph_t graph = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    c->cap_launches[0] = c->cap_launches[1] = c->cap_cells[0] = c->cap_cells[1] = 0;
    c->cap_spans.clear();
    c->cap_events_ok = true;
    int rc = gradient_pass(c, p0, np, P);
    c->capturing = false;
    const hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (rc || e != hipSuccess) {
      if (graph) hipGraphDestroy(graph);
      destroy_spans(c->cap_spans);
      (void)hipGetLastError();
      return rc ? rc : fail(c, TOMO_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    }
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (ei != hipSuccess) {
      destroy_spans(c->cap_spans);
      return fail(c, TOMO_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    }
    if (!ge) {
      c->graphs.emplace_back();
      ge = &c->graphs.back();
      ge->p0 = p0;
      ge->np = np;
      ge->P = P;
    }
    ge->exec = exec;
    ge->sig = sig;
    for (int k = 0; k < 2; ++k) {
      ge->launches[k] = c->cap_launches[k];
      ge->cells[k] = c->cap_cells[k];
    }
    // the loops' timings need every span of the pass; a partial set is dropped
    ge->spans.swap(c->cap_spans);
    if (!c->cap_events_ok) destroy_spans(ge->spans);
    c->cap_spans.clear();
  }
  // the graph's own event pairs hold one replay's times: the spans of an
  // earlier replay that nobody has read yet are read before they are overwritten
  if (!ge->spans.empty()) {
    bool pending = false;
    for (const auto &s : c->spans)
      if (s.owned && s.a == ge->spans.front().a) pending = true;
    if (pending) {
      HIPCHK(c, hipStreamSynchronize(c->stream));
      spans_collect(c);
    }
  }
  TimedSpan s;
  span_begin(c, s, 2);
  HIPCHK(c, hipGraphLaunch(ge->exec, c->stream));
  s.launches = 1;  // replays; the launches and cells are the loops' own spans
  if (ge->spans.empty()) s.cells = ge->cells[0] + ge->cells[1];  // no inner timings
  span_end(c, s);
  for (const auto &g : ge->spans) c->spans.push_back(g);
  return TOMO_OK;
}

}  // namespace

namespace tomo_internal {

int check_ctx(tomo_ctx *c) {
  if (!c) return TOMO_ERR_ARG;
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return fail(c, TOMO_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
  return TOMO_OK;
}

// the buffers, geometry and shapes a launch program built for c depends on
std::vector<uintptr_t> context_signature(const tomo_ctx *c) { return graph_signature(c); }

}  // namespace tomo_internal

namespace tomo_internal {

int need_shots(tomo_ctx *c) {
  if (c->nshots <= 0) return fail(c, TOMO_ERR_STATE, "no source groups: call tomo_set_groups first");
  if (c->nr <= 0) return fail(c, TOMO_ERR_STATE, "no receivers");
  return TOMO_OK;
}

}  // namespace tomo_internal

namespace {



}  // namespace

namespace tomo_internal {

int need_all_gradients(tomo_ctx *c) {
  for (int g = 0; g < c->nshots; ++g)
    if (!c->grad_valid[g]) return fail(c, TOMO_ERR_STATE, "gradients of all shots are required");
  return TOMO_OK;
}

int pixel_weights(tomo_ctx *c, int g0, int ng, const float **w) {
  *w = nullptr;
  if (c->pixw_mode == 0) return TOMO_OK;
  if (!c->coverage || !c->d_cov) return fail(c, TOMO_ERR_STATE, "coverage weights need tomo_enable_coverage");
  int rc;
  if (!c->d_pixw && (rc = dalloc(c, &c->d_pixw, (size_t)c->n2))) return rc;
  if ((rc = ensure_scratch(c, 2))) return rc;
  HIPCHK(c, launch_pixel_weights(c->d_cov + (size_t)g0 * c->n2, ng, c->nx, c->ny, ctx_mask(c), c->pixw_eps,
                                 c->d_scratch, c->d_pixw, c->stream));
  *w = c->d_pixw;
  return TOMO_OK;
}

}  // namespace tomo_internal

namespace tomo_internal {

// entry points of the batched stepping (tomo_batch.hip)
StencilShape context_shape(const tomo_ctx *c) { return stencil_shape(c); }
float *level_pass1(tomo_ctx *c, int P, int k) { return pass1_level(c, P, k); }
float *level_segment(tomo_ctx *c, int P, int m, int k) { return segment_level(c, P, m, k); }
FwdArgs<float> forward_args(tomo_ctx *c, int shot0) {
  FwdArgs<float> a = fwd_args<float>(c, shot0, c->d_pw, c->d_spa, c->d_spab);
  a.ring = c->d_ring;
  a.nring = c->nring;
  return a;
}
AdjArgs adjoint_args(tomo_ctx *c, int shot0) {
  AdjArgs a = adj_args(c, shot0);
  a.ring = c->d_ring;
  a.nring = c->nring;
  return a;
}
float pulse_h2_f32(const tomo_ctx *c, int k) { return pulse_h2<float>(c, k); }
int wavefields(tomo_ctx *c, int P) { return ensure_wavefields(c, P, true); }
int shots_per_pass(const tomo_ctx *c) { return pass_shots(c); }

}  // namespace tomo_internal

namespace {

// the shots of every phantom: the group boxes repeated per phantom, the
// per-shot buffers sized for nph * groups shots
int stack_shots(tomo_ctx *c) {
  const int ng = c->groups;
  std::vector<int4> all((size_t)ng * c->nph);
  for (int p = 0; p < c->nph; ++p) std::copy(c->bbox.begin(), c->bbox.end(), all.begin() + (size_t)p * ng);
  // no shots until every per-shot buffer is in place: a failure below
  // leaves none to run (tomo_set_groups again)
  c->nshots = 0;
  int rc;
  if ((rc = dalloc(c, &c->d_bbox, all.size()))) return rc;
  HIPCHK(c, hipMemcpy(c->d_bbox, all.data(), sizeof(int4) * all.size(), hipMemcpyHostToDevice));
  c->nshots = (int)all.size();
  if ((rc = alloc_shot_buffers(c))) {
    c->nshots = 0;
    return rc;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TOMO_OK;
}

int one_phantom(tomo_ctx *c, const char *what) {
  if (c->nph == 1) return TOMO_OK;
  return fail(c, TOMO_ERR_STATE, std::string(what) + " takes a context of one phantom (tomo_set_phantoms)");
}

}  // namespace

// ====================================================================== ABI
extern "C" {

int tomo_create(const tomo_config *cfg, int device, tomo_ctx **out) {
  if (!cfg || !out) return TOMO_ERR_ARG;
  *out = nullptr;
  std::string why;
  int rc = validate_config(*cfg, why);
  if (rc) {
    tomo_host::set_create_error(why);
    return rc;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    tomo_host::set_create_error("no HIP device");
    return TOMO_ERR_NODEV;
  }
  if (device < 0) {
    if (hipGetDevice(&device) != hipSuccess) device = 0;
  }
  if (device >= ndev) {
    tomo_host::set_create_error("device index out of range");
    return TOMO_ERR_ARG;
  }
  tomo_ctx *c = new tomo_ctx();
  c->cfg = *cfg;
  c->device = device;
  c->nx = cfg->nx;
  c->ny = cfg->ny;
  c->nt = cfg->nt;
  c->n2 = (long long)cfg->nx * cfg->ny;
  auto bail = [&](int code) {
    tomo_host::set_create_error(c->err);
    tomo_destroy(c);
    return code;
  };
  if (hipSetDevice(device) != hipSuccess) return bail(fail(c, TOMO_ERR_HIP, "hipSetDevice failed"));
  if (hipDeviceGetAttribute(&c->num_cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
      c->num_cus <= 0)
    return bail(fail(c, TOMO_ERR_HIP, "hipDeviceGetAttribute(multiprocessor count) failed"));
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(c, TOMO_ERR_HIP, "hipStreamCreate failed"));
  if ((rc = dalloc(c, &c->d_f, c->n2)) || (rc = dalloc(c, &c->d_fo, c->n2)) || (rc = dalloc(c, &c->d_pw, c->n2)) ||
      (rc = dalloc(c, &c->d_cellmap, c->n2)) || (rc = dalloc(c, &c->d_bbox, 1)))
    return bail(rc);
  // sponge tables (also allocated, 1 entry, when there is no sponge)
  const int W = std::max(cfg->sponge_width, 1);
  std::vector<double> a64(W, 1.0), ab64(W, 1.0);
  std::vector<float> a32(W, 1.f), ab32(W, 1.f);
  for (int d = 0; d < cfg->sponge_width; ++d) {
    const double g = tomo_host::sponge_gamma(*cfg, d);
    a64[d] = 1.0 / (1.0 + g);
    ab64[d] = (1.0 - g) / (1.0 + g);
    a32[d] = (float)a64[d];
    ab32[d] = (float)ab64[d];
  }
  if ((rc = dalloc(c, &c->d_spa, W)) || (rc = dalloc(c, &c->d_spab, W)) || (rc = dalloc(c, &c->d_spa64, W)) ||
      (rc = dalloc(c, &c->d_spab64, W)))
    return bail(rc);
  if (hipMemcpy(c->d_spa, a32.data(), 4 * W, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->d_spab, ab32.data(), 4 * W, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->d_spa64, a64.data(), 8 * W, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->d_spab64, ab64.data(), 8 * W, hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(c, TOMO_ERR_HIP, "upload of sponge tables failed"));
  // pulse per step, in float and in double arithmetic
  c->amp32.resize(c->nt);
  c->amp64.resize(c->nt);
  for (int k = 0; k < c->nt; ++k) {
    c->amp32[k] = tomo_host::pulse<float>(*cfg, k);
    c->amp64[k] = tomo_host::pulse<double>(*cfg, k);
  }
  // f = 0 (sirt-mimo.cu:238), receivers of the line layout
  if (launch_fill_f32(c->d_f, c->n2, 0.f, c->stream) != hipSuccess ||
      launch_fill_f32(c->d_fo, c->n2, 0.f, c->stream) != hipSuccess)
    return bail(fail(c, TOMO_ERR_HIP, "initial fill failed"));
  if ((rc = upload_coefficients(c))) return bail(rc);
  if (cfg->recv_mode == TOMO_RECV_LINES) {
    tomo_host::line_receivers(*cfg, c->rcell, c->ranchor);
  }
  if ((rc = upload_receivers(c))) return bail(rc);
  *out = c;
  return TOMO_OK;
}

void tomo_destroy(tomo_ctx *c) {
  if (!c) return;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  for (auto &s : c->spans) {
    if (s.owned) continue;  // destroyed with their graph
    if (s.a) hipEventDestroy(s.a);
    if (s.b) hipEventDestroy(s.b);
  }
  c->spans.clear();
  for (auto e : c->ev_free) hipEventDestroy(e);
  graphs_clear(c);
  dfree(c->d_f); dfree(c->d_fo); dfree(c->d_pw); dfree(c->d_pw64);
  dfree(c->d_cellmap); dfree(c->d_anchor); dfree(c->d_bbox); dfree(c->d_ring); dfree(c->d_rowflag);
  dfree(c->d_recvseg);
  dfree(c->d_spa); dfree(c->d_spab); dfree(c->d_spa64); dfree(c->d_spab64);
  dfree(c->d_meas); dfree(c->d_rr); dfree(c->d_grad); dfree(c->d_cov); dfree(c->d_misfit); dfree(c->d_pixw);
  if (c->ws) {  // the workspace's storage outlives the context
    c->ws->users--;
    c->d_ckpt = c->d_seg = c->d_z = nullptr;
  }
  dfree(c->d_scratch); dfree(c->d_ckpt); dfree(c->d_seg); dfree(c->d_z); dfree(c->d_ring64);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

const char *tomo_last_error(const tomo_ctx *c) {
  if (!c) return tomo_host::create_error();
  return c->err.c_str();
}

int tomo_set_transducers(tomo_ctx *c, const int *ii, const int *jj, int ns, int one_based) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!ii || !jj || ns <= 0) return fail(c, TOMO_ERR_ARG, "bad transducer arrays");
  const int o = one_based ? 1 : 0;
  c->tx.resize(ns);
  c->ty.resize(ns);
  for (int p = 0; p < ns; ++p) {
    const int x = ii[p] - o, y = jj[p] - o;
    if (x < 1 || x > c->nx - 2 || y < 1 || y > c->ny - 2)
      return fail(c, TOMO_ERR_ARG, "transducer " + std::to_string(p) + " not in the grid interior");
    c->tx[p] = x;
    c->ty[p] = y;
  }
  // the groups of the previous transducers are gone (tomo_set_groups again)
  c->nshots = 0;
  c->groups = 0;
  c->bbox.clear();
  return TOMO_OK;
}

int tomo_set_groups(tomo_ctx *c, int G) {
  int rc = check_ctx(c);
  if (rc) return rc;
  const int ns = (int)c->tx.size();
  if (ns == 0) return fail(c, TOMO_ERR_STATE, "no transducers");
  if (G <= 0 || G > ns) return fail(c, TOMO_ERR_ARG, "group size out of range");
  const int ng = ns / G;
  c->bbox.resize(ng);
  for (int g = 0; g < ng; ++g) {
    const int p = g * G, q = g * G + G - 1;
    // filled bounding box of the first and last transducer (minmax, sirt-mimo.cu:543-551)
    c->bbox[g] = make_int4(std::min(c->tx[p], c->tx[q]), std::max(c->tx[p], c->tx[q]), std::min(c->ty[p], c->ty[q]),
                           std::max(c->ty[p], c->ty[q]));
  }
  c->groups = ng;
  if ((rc = stack_shots(c))) return rc;
  return c->nshots;
}

int tomo_set_phantoms(tomo_ctx *c, int count) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (count < 1 || count > 65536) return fail(c, TOMO_ERR_ARG, "phantom count out of range");
  if ((double)count * c->n2 > 4e9) return fail(c, TOMO_ERR_ARG, "phantom images exceed 4e9 cells");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  // the new images first: on an allocation failure the context keeps its own
  const size_t n = (size_t)count * c->n2;
  float *f = nullptr, *fo = nullptr, *pw = nullptr;
  if ((rc = dalloc(c, &f, n)) || (rc = dalloc(c, &fo, n)) || (rc = dalloc(c, &pw, n))) {
    dfree(f);
    dfree(fo);
    dfree(pw);
    return rc;
  }
  graphs_clear(c);
  dfree(c->d_f);
  dfree(c->d_fo);
  dfree(c->d_pw);
  c->d_f = f;
  c->d_fo = fo;
  c->d_pw = pw;
  c->nph = count;
  dfree(c->d_pw64);
  HIPCHK(c, launch_fill_f32(c->d_f, n, 0.f, c->stream));
  HIPCHK(c, launch_fill_f32(c->d_fo, n, 0.f, c->stream));
  c->have_truth = false;
  if ((rc = upload_coefficients(c))) return rc;
  if (c->groups > 0 && !c->bbox.empty() && (rc = stack_shots(c))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TOMO_OK;
}

int tomo_num_phantoms(const tomo_ctx *c) { return c ? c->nph : TOMO_ERR_ARG; }

int tomo_set_receivers(tomo_ctx *c, const int *ri, const int *rj, int nr, int one_based) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (c->cfg.recv_mode != TOMO_RECV_POINTS) return fail(c, TOMO_ERR_STATE, "receivers are fixed by the line layout");
  if (!ri || !rj || nr <= 0) return fail(c, TOMO_ERR_ARG, "bad receiver arrays");
  const int o = one_based ? 1 : 0;
  std::vector<int> cells(nr);
  std::vector<char> seen((size_t)c->n2, 0);
  for (int r = 0; r < nr; ++r) {
    const int x = ri[r] - o, y = rj[r] - o;
    if (x < 2 || x > c->nx - 3 || y < 2 || y > c->ny - 3)
      return fail(c, TOMO_ERR_ARG, "receiver " + std::to_string(r) + " must lie >= 2 cells inside the edge");
    const int id = x + c->nx * y;
    if (seen[id]) return fail(c, TOMO_ERR_ARG, "duplicate receiver cell");
    seen[id] = 1;
    cells[r] = id;
  }
  c->rcell = cells;
  c->ranchor = cells;  // additive injection: anchor = the cell itself
  if ((rc = upload_receivers(c))) return rc;
  if (c->nshots > 0 && (rc = alloc_shot_buffers(c))) return rc;
  return TOMO_OK;
}

int tomo_num_shots(const tomo_ctx *c) { return c ? c->nshots : TOMO_ERR_ARG; }
int tomo_num_receivers(const tomo_ctx *c) { return c ? c->nr : TOMO_ERR_ARG; }

int tomo_get_receivers(const tomo_ctx *c, int *ri, int *rj) {
  if (!c || !ri || !rj) return TOMO_ERR_ARG;
  for (int r = 0; r < c->nr; ++r) {
    ri[r] = c->rcell[r] % c->nx;
    rj[r] = c->rcell[r] / c->nx;
  }
  return TOMO_OK;
}

int tomo_set_field(tomo_ctx *c, const float *f) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!f) return fail(c, TOMO_ERR_ARG, "null field");
  HIPCHK(c, hipMemcpyAsync(c->d_f, f, sizeof(float) * c->n2 * c->nph, hipMemcpyHostToDevice, c->stream));
  if ((rc = upload_coefficients(c))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TOMO_OK;
}

int tomo_get_field(tomo_ctx *c, float *f) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!f) return fail(c, TOMO_ERR_ARG, "null field");
  HIPCHK(c, hipMemcpyAsync(f, c->d_f, sizeof(float) * c->n2 * c->nph, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TOMO_OK;
}

int tomo_set_truth(tomo_ctx *c, const float *fo) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!fo) return fail(c, TOMO_ERR_ARG, "null truth");
  HIPCHK(c, hipMemcpyAsync(c->d_fo, fo, sizeof(float) * c->n2 * c->nph, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->have_truth = true;
  return TOMO_OK;
}

int tomo_forward(tomo_ctx *c, int g0, int ng, int fp64, float *traces) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if ((rc = need_shots(c))) return rc;
  if (g0 < 0 || ng <= 0 || g0 + ng > c->nshots) return fail(c, TOMO_ERR_ARG, "shot range out of bounds");
  const int P = std::min(pass_shots(c), ng);
  const size_t lvl = (size_t)P * c->n2;
  if (fp64) {
    if ((rc = ensure_pw64(c))) return rc;
    if (c->ring64_n < 3 * lvl) {
      if ((rc = dalloc(c, &c->d_ring64, 3 * lvl))) return rc;
      c->ring64_n = 3 * lvl;
    }
    rc = forward_traces<double>(c, g0, ng, c->d_ring64, lvl, c->d_pw64, c->d_spa64, c->d_spab64);
  } else {
    if ((rc = ensure_wavefields(c, P, false))) return rc;
    rc = forward_traces<float>(c, g0, ng, c->d_seg, lvl, c->d_pw, c->d_spa, c->d_spab);
  }
  if (rc) return rc;
  // traces now live in d_rr: residuals of those shots must be recomputed
  for (int g = g0; g < g0 + ng; ++g) c->grad_valid[g] = 0;
  if (traces) {
    const size_t tr = (size_t)c->nt * c->nr;
    HIPCHK(c, hipMemcpyAsync(traces, c->d_rr + (size_t)g0 * tr, sizeof(float) * tr * ng, hipMemcpyDeviceToHost,
                             c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  spans_collect(c);
  return TOMO_OK;
}

int tomo_misfit(tomo_ctx *c, double *misfit) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if ((rc = need_shots(c))) return rc;
  if (!misfit) return fail(c, TOMO_ERR_ARG, "null output");
  if (!c->have_meas) return fail(c, TOMO_ERR_STATE, "no measured traces: call tomo_set_measured first");
  const int P = pass_shots(c);
  const size_t lvl = (size_t)P * c->n2;
  if ((rc = ensure_wavefields(c, P, false))) return rc;
  if ((rc = forward_traces<float>(c, 0, c->nshots, c->d_seg, lvl, c->d_pw, c->d_spa, c->d_spab, c->d_meas)))
    return rc;
  std::fill(c->grad_valid.begin(), c->grad_valid.end(), 0);  // d_rr now holds these residuals
  const size_t tr = (size_t)c->nt * c->nr;
  HIPCHK(c, launch_misfit(c->d_rr, (long long)tr, c->nshots, c->d_misfit, c->stream));
  std::vector<double> mis(c->nshots);
  HIPCHK(c, hipMemcpyAsync(mis.data(), c->d_misfit, sizeof(double) * c->nshots, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  spans_collect(c);
  double sum = 0.0;
  for (double v : mis) sum += v;
  *misfit = sum;
  return TOMO_OK;
}

int tomo_set_measured(tomo_ctx *c, const float *traces) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if ((rc = need_shots(c))) return rc;
  if (!traces) return fail(c, TOMO_ERR_ARG, "null traces");
  const size_t n = (size_t)c->nshots * c->nt * c->nr;
  HIPCHK(c, hipMemcpyAsync(c->d_meas, traces, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->have_meas = true;
  return TOMO_OK;
}

int tomo_gradient(tomo_ctx *c, int g0, int ng, double *misfit) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if ((rc = need_shots(c))) return rc;
  if (!c->have_meas) return fail(c, TOMO_ERR_STATE, "no measured traces: call tomo_set_measured first");
  if (g0 < 0 || ng <= 0 || g0 + ng > c->nshots) return fail(c, TOMO_ERR_ARG, "shot range out of bounds");
  const int P = std::min(pass_shots(c), ng);
  if ((rc = ensure_wavefields(c, P, true))) return rc;
  for (int p0 = g0; p0 < g0 + ng; p0 += P) {
    const int np = std::min(P, g0 + ng - p0);
    if ((rc = c->use_graphs ? gradient_pass_graph(c, p0, np, P) : gradient_pass(c, p0, np, P))) return rc;
  }
  std::vector<double> mis(ng);
  HIPCHK(c, hipMemcpyAsync(mis.data(), c->d_misfit + g0, sizeof(double) * ng, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  spans_collect(c);
  for (int g = g0; g < g0 + ng; ++g) c->grad_valid[g] = 1;
  if (misfit) {
    double s = 0.0;
    for (double v : mis) s += v;
    *misfit = s;
  }
  return TOMO_OK;
}

int tomo_get_gradient(tomo_ctx *c, int g, float *out) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!out || g < 0 || g >= c->nshots) return fail(c, TOMO_ERR_ARG, "bad shot index");
  if (!c->grad_valid[g]) return fail(c, TOMO_ERR_STATE, "gradient of this shot not computed");
  HIPCHK(c, hipMemcpyAsync(out, c->d_grad + (size_t)g * c->n2, sizeof(float) * c->n2, hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TOMO_OK;
}

int tomo_enable_coverage(tomo_ctx *c, int on) {
  int rc = check_ctx(c);
  if (rc) return rc;
  c->coverage = on != 0;
  if (c->coverage && !c->d_cov && c->nshots > 0) {
    if ((rc = dalloc(c, &c->d_cov, (size_t)c->nshots * c->n2))) return rc;
    HIPCHK(c, launch_fill_f32(c->d_cov, (size_t)c->nshots * c->n2, 0.f, c->stream));
  }
  return TOMO_OK;
}

int tomo_get_coverage(tomo_ctx *c, float *out) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!out) return fail(c, TOMO_ERR_ARG, "null output");
  if (!c->coverage || !c->d_cov) return fail(c, TOMO_ERR_STATE, "coverage not enabled");
  std::vector<float> all((size_t)c->nshots * c->n2);
  HIPCHK(c, hipMemcpyAsync(all.data(), c->d_cov, sizeof(float) * all.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const int G = c->nshots / c->nph;
  for (int p = 0; p < c->nph; ++p)
    for (long long id = 0; id < c->n2; ++id) {
      double s = 0.0;
      for (int g = p * G; g < (p + 1) * G; ++g) s += all[(size_t)g * c->n2 + id];
      out[(size_t)p * c->n2 + id] = (float)s;
    }
  return TOMO_OK;
}


namespace {

// the uniform update of every phantom, phantom p at relaxation scales[p]
// (scales == nullptr: `scale` for all)
int update_uniform_phantoms(tomo_ctx *c, float scale, const float *scales, int mode) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if ((rc = need_shots(c)) || (rc = need_all_gradients(c))) return rc;
  // each phantom from its own shots' gradients
  const int G = c->nshots / c->nph;
  const double div = mode == TOMO_UPDATE_MEAN ? (double)G : 1.0;
  for (int p = 0; p < c->nph; ++p) {
    const float *pw = nullptr;
    if ((rc = pixel_weights(c, p * G, G, &pw))) return rc;
    HIPCHK(c, launch_update_uniform(c->d_f + (size_t)p * c->n2, c->d_grad + (size_t)p * G * c->n2, G, c->nx, c->ny,
                                    ctx_mask(c), (double)(scales ? scales[p] : scale), div, c->stream, pw));
  }
  if ((rc = upload_coefficients(c))) return rc;
  std::fill(c->grad_valid.begin(), c->grad_valid.end(), 0);  // the field changed
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TOMO_OK;
}

// the weighted update of every phantom, phantom p at relaxation scales[p]
// (scales == nullptr: `scale` for all)
int update_weighted_phantoms(tomo_ctx *c, float scale, const float *scales, double *weights_out) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if ((rc = need_shots(c)) || (rc = need_all_gradients(c))) return rc;
  // each phantom from its own M shots, one after another on the stream
  // (they share the scratch)
  const int M = c->nshots / c->nph;
  const int pb = weighted_part_blocks((int)c->n2);
  const size_t need = (size_t)c->n2 + (size_t)M * pb + (size_t)M + 1;
  if ((rc = ensure_scratch(c, need + 2))) return rc;
  // the weights' own scratch: a separate slot past the weighted-update scratch
  double *mean = c->d_scratch, *part = mean + c->n2, *w = part + (size_t)M * pb;
  if (c->pixw_mode) {
    if (!c->coverage || !c->d_cov) return fail(c, TOMO_ERR_STATE, "coverage weights need tomo_enable_coverage");
    if (!c->d_pixw && (rc = dalloc(c, &c->d_pixw, (size_t)c->n2))) return rc;
  }
  for (int p = 0; p < c->nph; ++p) {
    const float *pw = nullptr;
    if (c->pixw_mode) {
      HIPCHK(c, launch_pixel_weights(c->d_cov + (size_t)p * M * c->n2, M, c->nx, c->ny, ctx_mask(c), c->pixw_eps,
                                     c->d_scratch + need, c->d_pixw, c->stream));
      pw = c->d_pixw;
    }
    HIPCHK(c, launch_weighted(c->d_f + (size_t)p * c->n2, c->d_grad + (size_t)p * M * c->n2, M, c->nx, c->ny,
                              ctx_mask(c), (double)(scales ? scales[p] : scale), mean, part, pb, w, c->stream, pw));
    if (weights_out)
      HIPCHK(c, hipMemcpyAsync(weights_out + (size_t)p * M, w, sizeof(double) * M, hipMemcpyDeviceToHost,
                               c->stream));
  }
  if ((rc = upload_coefficients(c))) return rc;
  std::fill(c->grad_valid.begin(), c->grad_valid.end(), 0);  // the field changed
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TOMO_OK;
}

}  // namespace

int tomo_update_uniform(tomo_ctx *c, float scale, int mode) { return update_uniform_phantoms(c, scale, nullptr, mode); }

int tomo_update_uniform_each(tomo_ctx *c, const float *scales, int mode) {
  if (!scales) return fail(c, TOMO_ERR_ARG, "null scales");
  return update_uniform_phantoms(c, 0.f, scales, mode);
}

int tomo_update_weighted(tomo_ctx *c, float scale, double *weights_out) {
  return update_weighted_phantoms(c, scale, nullptr, weights_out);
}

int tomo_update_weighted_each(tomo_ctx *c, const float *scales, double *weights_out) {
  if (!scales) return fail(c, TOMO_ERR_ARG, "null scales");
  return update_weighted_phantoms(c, 0.f, scales, weights_out);
}

int tomo_update_avg_weight(tomo_ctx *c, float scale, double *weights_out) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if ((rc = one_phantom(c, "tomo_update_avg_weight"))) return rc;
  if ((rc = need_shots(c)) || (rc = need_all_gradients(c))) return rc;
  const int M = c->nshots;
  const int pb = weighted_part_blocks((int)c->n2);
  const size_t need = (size_t)c->n2 + (size_t)M * pb + (size_t)M + 1;
  if ((rc = ensure_scratch(c, need))) return rc;
  double *mean = c->d_scratch, *part = mean + c->n2, *w = part + (size_t)M * pb;
  HIPCHK(c, launch_avg_weight(c->d_f, c->d_grad, M, c->nx, c->ny, ctx_mask(c), (double)scale, mean, part, pb, w,
                              c->stream));
  if ((rc = upload_coefficients(c))) return rc;
  std::fill(c->grad_valid.begin(), c->grad_valid.end(), 0);  // the field changed
  if (weights_out)
    HIPCHK(c, hipMemcpyAsync(weights_out, w, sizeof(double) * (M + 1), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TOMO_OK;
}

int tomo_update_group(tomo_ctx *c, int g, float scale) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if ((rc = one_phantom(c, "tomo_update_group"))) return rc;
  if ((rc = need_shots(c))) return rc;
  if (g < 0 || g >= c->nshots) return fail(c, TOMO_ERR_ARG, "bad shot index");
  if (!c->grad_valid[g]) return fail(c, TOMO_ERR_STATE, "gradient of this shot not computed");
  const float *pw = nullptr;
  if ((rc = pixel_weights(c, g, 1, &pw))) return rc;
  HIPCHK(c, launch_update_group(c->d_f, c->d_grad + (size_t)g * c->n2, c->nx, c->ny, ctx_mask(c), (double)scale,
                                c->stream, pw));
  if ((rc = upload_coefficients(c))) return rc;
  // the field changed: every stored gradient is stale
  std::fill(c->grad_valid.begin(), c->grad_valid.end(), 0);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TOMO_OK;
}

int tomo_error(tomo_ctx *c, double *eps) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!eps) return fail(c, TOMO_ERR_ARG, "null output");
  if (!c->have_truth) return fail(c, TOMO_ERR_STATE, "no ground truth: call tomo_set_truth first");
  // one value per phantom
  if ((rc = ensure_scratch(c, 2 * (size_t)c->nph))) return rc;
  for (int p = 0; p < c->nph; ++p)
    HIPCHK(c, launch_error(c->d_f + (size_t)p * c->n2, c->d_fo + (size_t)p * c->n2, (int)c->n2,
                           c->d_scratch + 2 * p, c->stream));
  std::vector<double> s(2 * (size_t)c->nph);
  HIPCHK(c, hipMemcpyAsync(s.data(), c->d_scratch, sizeof(double) * s.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int p = 0; p < c->nph; ++p) eps[p] = 100.0 * std::sqrt(s[2 * p]) / std::sqrt(s[2 * p + 1]);
  return TOMO_OK;
}

int tomo_sirt_iteration(tomo_ctx *c, float scale, int weighted, int mode, double *misfit, double *eps) {
  int rc = tomo_gradient(c, 0, c ? c->nshots : 0, misfit);
  if (rc) return rc;
  rc = weighted ? tomo_update_weighted(c, scale, nullptr) : tomo_update_uniform(c, scale, mode);
  if (rc) return rc;
  if (eps && c->have_truth) return tomo_error(c, eps);
  return TOMO_OK;
}

int tomo_art_iteration(tomo_ctx *c, float scale, double *misfit, double *eps) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if ((rc = one_phantom(c, "tomo_art_iteration"))) return rc;
  if ((rc = need_shots(c))) return rc;
  double total = 0.0;
  for (int g = 0; g < c->nshots; ++g) {
    double m = 0.0;
    if ((rc = tomo_gradient(c, g, 1, &m))) return rc;
    total += m;
    if ((rc = tomo_update_group(c, g, scale))) return rc;
  }
  if (misfit) *misfit = total;
  if (eps && c->have_truth) return tomo_error(c, eps);
  return TOMO_OK;
}

int tomo_traces_to_sides(const tomo_ctx *c, const float *traces, float *bottom, float *right, float *top,
                         float *left) {
  if (!c || !traces || !bottom || !right || !top || !left) return TOMO_ERR_ARG;
  if (c->cfg.recv_mode != TOMO_RECV_LINES) return TOMO_ERR_STATE;
  return tomo_host::traces_to_sides(c->cfg, c->nshots, traces, bottom, right, top, left);
}

int tomo_sides_to_traces(const tomo_ctx *c, const float *bottom, const float *right, const float *top,
                         const float *left, float *traces) {
  if (!c || !traces || !bottom || !right || !top || !left) return TOMO_ERR_ARG;
  if (c->cfg.recv_mode != TOMO_RECV_LINES) return TOMO_ERR_STATE;
  return tomo_host::sides_to_traces(c->cfg, c->nshots, bottom, right, top, left, traces);
}

int tomo_timing(tomo_ctx *c, double *out6) {
  if (!c || !out6) return TOMO_ERR_ARG;
  out6[0] = c->t_ms[0];
  out6[1] = c->t_launch[0];
  out6[2] = c->t_ms[1];
  out6[3] = c->t_launch[1];
  out6[4] = c->t_cells[0];
  out6[5] = c->t_cells[1];
  return TOMO_OK;
}

int tomo_timing_reset(tomo_ctx *c) {
  if (!c) return TOMO_ERR_ARG;
  for (int i = 0; i < 3; ++i) c->t_ms[i] = c->t_launch[i] = c->t_cells[i] = 0;
  return TOMO_OK;
}

int tomo_line_search(tomo_ctx *c, int weighted, double s_lo, double s_step, int count, double *best_scale,
                     double *best_eps) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if ((rc = one_phantom(c, "tomo_line_search"))) return rc;
  if ((rc = need_shots(c)) || (rc = need_all_gradients(c))) return rc;
  if (!c->have_truth) return fail(c, TOMO_ERR_STATE, "no ground truth: call tomo_set_truth first");
  if (count <= 0 || !(s_step > 0.0) || !best_scale) return fail(c, TOMO_ERR_ARG, "bad scale grid");
  const int M = c->nshots;
  const int pb = weighted_part_blocks((int)c->n2);
  // scratch: [weighted mean n2][partials M pb][weights M][pad][direction n2][sums 4][pixel stats 2]
  const size_t off_w = (size_t)c->n2 + (size_t)M * pb;
  const size_t off_d = off_w + (size_t)M + 1;
  const size_t off_s = off_d + (size_t)c->n2;
  if ((rc = ensure_scratch(c, off_s + 6))) return rc;
  const float *pw = nullptr;
  if (c->pixw_mode) {
    if (!c->coverage || !c->d_cov) return fail(c, TOMO_ERR_STATE, "coverage weights need tomo_enable_coverage");
    if (!c->d_pixw && (rc = dalloc(c, &c->d_pixw, (size_t)c->n2))) return rc;
    HIPCHK(c, launch_pixel_weights(c->d_cov, M, c->nx, c->ny, ctx_mask(c), c->pixw_eps, c->d_scratch + off_s + 4,
                                   c->d_pixw, c->stream));
    pw = c->d_pixw;
  }
  const double *w = nullptr;
  if (weighted) {
    HIPCHK(c, launch_weights_only(c->d_f, c->d_grad, M, c->nx, c->ny, c->d_scratch, c->d_scratch + c->n2, pb,
                                  c->d_scratch + off_w, c->stream));
    w = c->d_scratch + off_w;
  }
  HIPCHK(c, launch_error_quadratic(c->d_f, c->d_fo, c->d_grad, M, c->nx, c->ny, ctx_mask(c), (double)M, w, pw,
                                   c->d_scratch + off_d, c->d_scratch + off_s, c->stream));
  double q[4];
  HIPCHK(c, hipMemcpyAsync(q, c->d_scratch + off_s, sizeof(q), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  // eps(s)^2 is the quadratic a + 2 b s + c s^2 over the scale grid
  auto err2 = [&](int k) {
    const double s = s_lo + s_step * k;
    return q[0] + 2.0 * q[1] * s + q[2] * s * s;
  };
  int kbest = 0;
  if (q[2] > 0.0) {
    const double ks = (-q[1] / q[2] - s_lo) / s_step;
    const int k0 = (int)std::max(0.0, std::min((double)(count - 1), std::floor(ks)));
    kbest = k0;
    for (int k = std::max(0, k0 - 1); k <= std::min(count - 1, k0 + 2); ++k)
      if (err2(k) < err2(kbest)) kbest = k;
    if (err2(0) <= err2(kbest)) kbest = 0;  // the first grid point wins ties
  }
  *best_scale = s_lo + s_step * kbest;
  if (best_eps) *best_eps = 100.0 * std::sqrt(std::max(err2(kbest), 0.0)) / std::sqrt(q[3]);
  return TOMO_OK;
}

int tomo_set_time_blocking(tomo_ctx *c, int steps) {
  if (!c) return TOMO_ERR_ARG;
  if (steps < 1 || steps > 3) return fail(c, TOMO_ERR_ARG, "steps per launch must be 1, 2 or 3");
  c->time_block = steps;
  return TOMO_OK;
}

int tomo_get_time_blocking(const tomo_ctx *c) {
  return c ? (three_step(c) ? 3 : two_step(c) ? 2 : 1) : TOMO_ERR_ARG;
}
int tomo_get_adjoint_blocking(const tomo_ctx *c) {
  return c ? (three_step_adjoint(c) ? 3 : two_step_adjoint(c) ? 2 : 1) : TOMO_ERR_ARG;
}

int tomo_set_adjoint_blocking(tomo_ctx *c, int steps) {
  if (!c) return TOMO_ERR_ARG;
  if (steps < 1 || steps > 3) return fail(c, TOMO_ERR_ARG, "adjoint steps per launch must be 1, 2 or 3");
  c->adj_block = steps;
  return TOMO_OK;
}

int tomo_workspace_create(int device, tomo_workspace **out) {
  if (!out) return TOMO_ERR_ARG;
  *out = nullptr;
  int dev = device;
  if (dev < 0 && hipGetDevice(&dev) != hipSuccess) {
    tomo_host::set_create_error("tomo_workspace_create: no usable GPU");
    return TOMO_ERR_NODEV;
  }
  tomo_workspace *w = new tomo_workspace;
  w->device = dev;
  *out = w;
  return TOMO_OK;
}

int tomo_workspace_destroy(tomo_workspace *w) {
  if (!w) return TOMO_ERR_ARG;
  if (w->users > 0) {
    tomo_host::set_create_error("tomo_workspace_destroy: contexts still use the workspace");
    return TOMO_ERR_STATE;
  }
  (void)hipSetDevice(w->device);
  dfree(w->d_ckpt);
  dfree(w->d_seg);
  dfree(w->d_z);
  delete w;
  return TOMO_OK;
}

int tomo_set_workspace(tomo_ctx *c, tomo_workspace *w) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (w == c->ws) return TOMO_OK;
  if (w && w->device != c->device) return fail(c, TOMO_ERR_ARG, "workspace and context are on different devices");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->ws) {
    c->ws->users--;
    c->d_ckpt = c->d_seg = c->d_z = nullptr;  // the old workspace keeps its storage
  } else {
    dfree(c->d_ckpt);  // the context's own storage
    dfree(c->d_seg);
    dfree(c->d_z);
  }
  c->ckpt_n = c->seg_n = c->z_n = 0;
  c->ws = w;
  if (w) w->users++;
  return TOMO_OK;
}

int tomo_set_pixel_weights(tomo_ctx *c, int mode, double eps) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (mode != 0 && mode != 1) return fail(c, TOMO_ERR_ARG, "pixel weight mode must be 0 or 1");
  if (!(eps > 0.0)) return fail(c, TOMO_ERR_ARG, "eps must be positive");
  c->pixw_mode = mode;
  c->pixw_eps = eps;
  return TOMO_OK;
}

int tomo_timing_graphs(tomo_ctx *c, double *out3) {
  if (!c || !out3) return TOMO_ERR_ARG;
  out3[0] = c->t_ms[2];
  out3[1] = c->t_launch[2];
  out3[2] = c->t_cells[2];
  return TOMO_OK;
}

int tomo_set_graphs(tomo_ctx *c, int on) {
  int rc = check_ctx(c);
  if (rc) return rc;
  c->use_graphs = on != 0;
  if (!c->use_graphs) graphs_clear(c);
  return TOMO_OK;
}

}  // extern "C"

namespace {

// Shapes chosen by the autotuner, per size class (grid, shots per pass,
// sponge, coverage), shared by every context of the process.
struct TuneEntry {
  int nx, ny, P, sponge, cov;
  StencilShape shape;
  double fwd_ms, adj_ms;
};
std::mutex g_tune_mu;
std::vector<TuneEntry> g_tune;

// average ms of `reps` launches of fn on the context stream (min of 2 trials)
template <typename F>
int time_launches(tomo_ctx *c, int reps, F fn, double *ms) {
  hipEvent_t a = take_event(c), b = take_event(c);
  if (!a || !b) return fail(c, TOMO_ERR_HIP, "hipEventCreate failed");
  double best = 1e30;
  int rc = TOMO_OK;
  for (int trial = 0; trial < 3 && !rc; ++trial) {
    HIPCHK(c, hipEventRecord(a, c->stream));
    for (int r = 0; r < reps && !rc; ++r) rc = fn();
    HIPCHK(c, hipEventRecord(b, c->stream));
    HIPCHK(c, hipEventSynchronize(b));
    float t = 0.f;
    HIPCHK(c, hipEventElapsedTime(&t, a, b));
    if (trial > 0) best = std::min(best, (double)t / reps);  // trial 0 warms up
  }
  c->ev_free.push_back(a);
  c->ev_free.push_back(b);
  *ms = best;
  return rc;
}

}  // namespace

extern "C" {

int tomo_autotune(tomo_ctx *c, int force, int *out5, double *ms2) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if ((rc = need_shots(c))) return rc;
  const int P = pass_shots(c);
  const int sponge = c->cfg.sponge_width > 0;
  const int cov = c->coverage ? 1 : 0;
  auto finish = [&](const TuneEntry &t) {
    c->shape_req = t.shape;
    c->shape_req.rows = c->shape_req.depth = 0;
    c->shape_req.ring = nullptr;
    c->shape_req.nring = 0;
    if (out5) {
      out5[0] = t.shape.vec;
      out5[1] = t.shape.fwd_rows;
      out5[2] = t.shape.adj_rows;
      out5[3] = t.shape.fwd_depth;
      out5[4] = t.shape.adj_depth;
    }
    if (ms2) {
      ms2[0] = t.fwd_ms;
      ms2[1] = t.adj_ms;
    }
    return TOMO_OK;
  };
  if (!force) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (const auto &t : g_tune)
      if (t.nx == c->nx && t.ny == c->ny && t.P == P && t.sponge == sponge && t.cov == cov) return finish(t);
  }
  if ((rc = ensure_wavefields(c, P, true))) return rc;
  const size_t lvl = (size_t)P * c->n2;
  float *Z0 = c->d_z, *Z1 = c->d_z + lvl, *Z2 = c->d_z + 2 * lvl;
  const StencilShape keep = c->shape_req;
  const int mid = c->nt / 2;
  FwdArgs<float> fa = fwd_args<float>(c, 0, c->d_pw, c->d_spa, c->d_spab);
  AdjArgs aa = adj_args(c, 0);
  aa.z1 = Z1;
  aa.z2 = Z2;
  aa.z0 = Z0;
  aa.u = pass1_level(c, P, 0);
  aa.level = mid;
  const int reps = std::max(4, std::min(32, (int)(4.0e8 / ((double)lvl + 1.0))));
  TuneEntry best{c->nx, c->ny, P, sponge, cov, StencilShape{}, 1e30, 1e30};
  double best_cost = 1e30;
  const int fw = full_history(c) ? 1 : 2;  // forward launches per adjoint launch in a gradient pass
  for (int vec : {4, 2}) {
    if (c->nx % vec != 0 || c->nx < 4 * vec) continue;
    double bf = 1e30, ba = 1e30;
    int brf = 0, bdf = 0, bra = 0, bda = 0;
    for (int rows : {8, 16, 32}) {
      if (rows > c->ny) continue;
      for (int depth : {1, 2, 4}) {
        c->shape_req = StencilShape{};
        c->shape_req.vec = vec;
        c->shape_req.fwd_rows = rows;
        c->shape_req.fwd_depth = depth;
        c->shape_req.adj_rows = rows;
        c->shape_req.adj_depth = depth;
        double tf = 0.0, ta = 0.0;
        rc = time_launches(
            c, reps, [&]() -> int { return forward_step<float>(c, fa, Z0, Z1, Z2, mid, P, false, nullptr, nullptr); },
            &tf);
        if (!rc && depth <= 2) {
          StencilShape sh = stencil_shape(c);
          sh.rows = sh.adj_rows;
          sh.depth = sh.adj_depth;
          rc = time_launches(c, reps, [&]() -> int {
            HIPCHK(c, launch_adjoint_step(aa, P, sh, c->stream));
            return TOMO_OK;
          }, &ta);
        } else {
          ta = 1e30;
        }
        if (rc) {
          c->shape_req = keep;
          return rc;
        }
        if (tf < bf) { bf = tf; brf = rows; bdf = depth; }
        if (ta < ba) { ba = ta; bra = rows; bda = depth; }
      }
    }
    const double cost = fw * bf + ba;
    if (brf && bra && cost < best_cost) {
      best_cost = cost;
      best.shape = StencilShape{};
      best.shape.vec = vec;
      best.shape.fwd_rows = brf;
      best.shape.fwd_depth = bdf;
      best.shape.adj_rows = bra;
      best.shape.adj_depth = bda;
      best.fwd_ms = bf;
      best.adj_ms = ba;
    }
  }
  // with two or three steps per launch the one-step kernels only run the
  // tails: the strip heights are then those of the blocked kernels (same
  // lane width)
  if (best.shape.vec != 0) {
    c->shape_req = best.shape;
    if (two_step(c)) {
      float *Z3 = c->d_z + 3 * lvl;
      const bool adj2 = two_step_adjoint(c);
      double bf = 1e30, ba = 1e30;
      int brf = 0, bra = 0, bdf = 1;
      for (int rows : {8, 16, 32}) {
        if (rows > c->ny) continue;
        c->shape_req.fwd_rows = rows;
        c->shape_req.adj_rows = rows;
        double tf = 1e30, ta = 1e30;
        int df = 1;
        for (int depth : {1, 2}) {
          c->shape_req.fwd_depth = depth;
          double t = 0.0;
          rc = time_launches(c, reps, [&]() -> int {
            return forward_step2(c, fa, Z0, Z1, Z2, Z3, mid, P, false, nullptr, nullptr);
          }, &t);
          if (rc) break;
          if (t < tf) { tf = t; df = depth; }
        }
        // where three steps per launch apply at this strip height, those
        // kernels run all but the tails: they are timed instead (a recording
        // launch, level k+1 dropped, and a replay launch storing all three)
        if (!rc && three_step(c)) {
          float *U3 = pass1_level(c, P, 2);
          double t = 0.0;
          rc = time_launches(c, reps, [&]() -> int {
            int r = forward_step3(c, fa, Z0, Z1, Z2, Z3, U3, mid, P, false, nullptr, nullptr, false);
            return r ? r : forward_step3(c, fa, Z0, Z1, Z2, Z3, U3, mid, P, false, nullptr, nullptr, true);
          }, &t);
          tf = t / 3;  // two launches of three levels, per two levels as below
        }
        if (!rc && adj2) {
          Adj2Args b = adj2_args(c, 0);
          b.z1 = Z1;
          b.z2 = Z2;
          b.z0 = Z0;
          b.zm = Z3;
          b.u0 = pass1_level(c, P, 0);
          b.um = pass1_level(c, P, 1);
          b.level = mid;
          rc = time_launches(c, reps, [&]() -> int {
            HIPCHK(c, launch_adjoint2(b, P, adjoint2_rows(c), c->stream));
            return TOMO_OK;
          }, &ta);
        }
        if (rc) {
          c->shape_req = keep;
          return rc;
        }
        if (tf < bf) { bf = tf; brf = rows; bdf = df; }
        if (ta < ba) { ba = ta; bra = rows; }
      }
      best.shape.fwd_rows = brf;
      best.shape.fwd_depth = bdf;
      best.fwd_ms = bf / 2;  // per level
      if (adj2) {
        best.shape.adj_rows = bra;
        best.adj_ms = ba / 2;
      }
      // where three adjoint steps per launch apply, that kernel runs all but
      // the tails: its strip heights are timed instead
      c->shape_req = best.shape;
      if (three_step_adjoint(c)) {
        Adj3Args t{};
        t.b = adj2_args(c, 0);
        t.b.z1 = Z1;
        t.b.z2 = Z2;
        t.b.zm = c->d_z + 3 * lvl;
        t.zm2 = c->d_z + 4 * lvl;
        t.b.u0 = pass1_level(c, P, 0);
        t.b.um = pass1_level(c, P, 1);
        t.um2 = pass1_level(c, P, 2);
        t.b.level = mid;
        // strip height and read-ahead depth of that kernel (the one-step
        // sweep's depth does not carry over to it)
        double b3 = 1e30;
        int br3 = 0, bd3 = 1;
        for (int rows : {16, 32, 64}) {
          if (rows > c->ny) continue;
          for (int depth : {1, 2}) {
            if (rows == 64 && depth == 2) continue;  // 64-row strips read one row ahead only
            double t3 = 0.0;
            rc = time_launches(c, reps, [&]() -> int {
              HIPCHK(c, launch_adjoint3(t, P, rows, depth, c->stream));
              return TOMO_OK;
            }, &t3);
            if (rc) {
              c->shape_req = keep;
              return rc;
            }
            if (t3 < b3) { b3 = t3; br3 = rows; bd3 = depth; }
          }
        }
        if (br3) {
          best.shape.adj_rows = br3;
          best.shape.adj_depth = bd3;
          best.adj_ms = b3 / 3;
        }
      }
    }
  }
  c->shape_req = keep;
  // the timing launches wrote the gradient accumulators
  std::fill(c->grad_valid.begin(), c->grad_valid.end(), 0);
  if (best.shape.vec == 0) return TOMO_OK;  // no vectorised shape fits: keep the scalar kernels
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    bool found = false;
    for (auto &t : g_tune)
      if (t.nx == best.nx && t.ny == best.ny && t.P == best.P && t.sponge == best.sponge && t.cov == best.cov) {
        t = best;
        found = true;
      }
    if (!found) g_tune.push_back(best);
  }
  return finish(best);
}

int tomo_set_shape(tomo_ctx *c, int vec, int fwd_rows, int adj_rows, int fwd_depth, int adj_depth) {
  if (!c) return TOMO_ERR_ARG;
  if (!shape_valid(vec, fwd_rows, adj_rows) || !depth_valid(fwd_depth) || !depth_valid(adj_depth))
    return fail(c, TOMO_ERR_ARG, "unsupported stencil shape");
  if (vec > 1 && (c->nx % vec != 0 || c->nx < 4 * vec))
    return fail(c, TOMO_ERR_ARG, "nx is not a multiple of the vector width (or below 4 widths)");
  c->shape_req = StencilShape{};
  c->shape_req.vec = vec;
  c->shape_req.fwd_rows = fwd_rows;
  c->shape_req.adj_rows = adj_rows;
  c->shape_req.fwd_depth = fwd_depth;
  c->shape_req.adj_depth = adj_depth;
  return TOMO_OK;
}

int tomo_get_shape(const tomo_ctx *c, int *out5) {
  if (!c || !out5) return TOMO_ERR_ARG;
  const StencilShape s = stencil_shape(c);
  out5[0] = s.vec;
  out5[1] = s.fwd_rows;
  // the three-step adjoint's strips where it runs (for passes of the most
  // shots a pass holds)
  out5[2] = three_step_adjoint(c) ? adjoint3_rows(c, pass_shots(c)) : s.adj_rows;
  out5[3] = s.fwd_depth;
  out5[4] = s.adj_depth;
  return TOMO_OK;
}

}  // extern "C"
